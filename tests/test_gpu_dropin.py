"""GPU parity of the reference-named drop-ins (update_v_PML,
update_s_elastic_PML, update_s_visc_PML) called with the reference's
Numerical-Recipes float** arrays, against the oracle's single sweeps."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from common import ROOT, model, oracle, product

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a)
    sys.stdout.flush()


def _setup(nx, ny, fdorder, L, free_surf, fw, seed=5):
    from pyoracle import make_cfg
    cfg = make_cfg(nx, ny, fdorder=fdorder, fw=fw, free_surf=free_surf, L=L, nt=10, dt=1e-4, dh=0.25,
                   damping=300.0, fpml=50.0, fl=[6.25, 50.0, 12.0, 30.0][:L])
    o = oracle("f32")
    vs, rho, q = model(nx, ny, 0.25, seed=seed, qs=20.0 if L else None)
    P = o.model_setup(cfg, vs, rho, q)
    hc = o.holberg(fdorder, 1)
    pml = o.pml(cfg)
    rng = np.random.default_rng(seed)
    sh = o.padded_shape(cfg)
    nd = fdorder // 2 + 1
    inner = (slice(nd, nd + ny), slice(nd, nd + nx))
    F = {}
    for k, s in (("vy", 1e-3), ("sxy", 1.0), ("syz", 1.0)):
        a = np.zeros(sh, np.float32)
        a[inner] = (rng.standard_normal((ny, nx)) * s).astype(np.float32)
        F[k] = a
    F["psi_a"] = (rng.standard_normal(2 * fw * nx) * 1e-2).astype(np.float32)   # [2fw][nx]
    F["psi_b"] = (rng.standard_normal(2 * fw * ny) * 1e-2).astype(np.float32)   # [ny][2fw]
    F["r"] = (rng.standard_normal((max(L, 1),) + sh) * 1e-2).astype(np.float32)
    F["q"] = (rng.standard_normal((max(L, 1),) + sh) * 1e-2).astype(np.float32)
    return cfg, o, P, hc, pml, F, nd


def _nr(dn, a, nd, nx, ny):
    return dn.NRMatrix(-nd + 1, ny + nd, -nd + 1, nx + nd, a.copy())


def _vec(dn, a, fw):
    return dn.NRVector(1, 2 * fw, a[1:2 * fw + 1].copy())


@pytest.mark.parametrize("fdorder,free_surf,sw", [(4, 1, 0), (8, 0, 0), (12, 0, 1), (2, 1, 1)])
def test_update_v_PML(fdorder, free_surf, sw):
    nx, ny, fw = 60, 34, 8
    say("update_v_PML fdorder", fdorder, "free_surf", free_surf, "sw", sw)
    cfg, o, P, hc, pml, F, nd = _setup(nx, ny, fdorder, 0, free_surf, fw)
    dn = product()
    dn.set_globals(NX=nx, NY=ny, FDORDER=fdorder, FW=fw, FREE_SURF=free_surf, L=0, DT=1e-4, DH=0.25, NT=10)
    vy, vyp1 = _nr(dn, F["vy"], nd, nx, ny), _nr(dn, np.zeros_like(F["vy"]), nd, nx, ny)
    syz, sxy, rho = _nr(dn, F["syz"], nd, nx, ny), _nr(dn, F["sxy"], nd, nx, ny), _nr(dn, P["rho"].astype(np.float32), nd, nx, ny)
    # the drop-in updates the CPML memory in place: hand it copies
    psy = dn.NRMatrix(1, 2 * fw, 1, nx, F["psi_a"].reshape(2 * fw, nx).copy())
    psx = dn.NRMatrix(1, ny, 1, 2 * fw, F["psi_b"].reshape(ny, 2 * fw).copy())
    hcv = dn.NRVector(0, 6, hc.astype(np.float32))
    vecs = [_vec(dn, pml[k].astype(np.float32), fw) for k in
            ("K_x", "a_x", "b_x", "K_x_half", "a_x_half", "b_x_half", "K_y", "a_y", "b_y", "K_y_half", "a_y_half",
             "b_y_half")]
    # one source (type 3) at (i=20, j=5) with signal value 0.25 at nt=3
    srcpos = dn.NRMatrix(1, 8, 1, 1, np.array([[20], [5], [0], [0], [50], [1], [0], [3]], np.float32))
    signals = dn.NRMatrix(1, 1, 1, 10, np.zeros((1, 10), np.float32))
    signals.a[0, 2] = 0.25
    L = dn.lib()
    L.update_v_PML.restype = None
    L.update_v_PML(1, nx, 1, ny, 3, vy.ptr, vyp1.ptr, None, None, syz.ptr, sxy.ptr, rho.ptr, srcpos.ptr,
                   signals.ptr, 1, None, hcv.ptr, 0, sw, *[v.ptr for v in vecs], psy.ptr, psx.ptr)
    # oracle
    rv = F["vy"].copy(); rvp = np.zeros_like(rv); ra = F["psi_a"].copy(); rb = F["psi_b"].copy()
    o.update_v(cfg, sw, rv, rvp, F["syz"], F["sxy"], P["rho"], hc, pml, ra, rb)
    if sw == 0:
        rv[5 + nd - 1, 20 + nd - 1] += np.float32(0.25)
    else:
        # adjoint: the drop-in injects at srcpos with QUELLTYPB = 2 for every listed point
        rv[5 + nd - 1, 20 + nd - 1] += np.float32(0.25)
    inner = (slice(nd, nd + ny), slice(nd, nd + nx))
    for name, got, ref in (("vy", vy.a[inner], rv[inner]), ("vyp1", vyp1.a[inner], rvp[inner]),
                           ("psi_syy_y", psy.a.ravel(), ra), ("psi_sxy_x", psx.a.ravel(), rb)):
        d = np.max(np.abs(got.astype(np.float64) - ref))
        say("  %s max |diff| %.3e" % (name, d))
        assert np.array_equal(got, ref.astype(np.float32)), name


@pytest.mark.parametrize("fdorder,free_surf,L", [(4, 0, 0), (8, 1, 0), (8, 0, 2), (12, 0, 4), (6, 1, 1)])
def test_update_s_PML(fdorder, free_surf, L):
    nx, ny, fw = 60, 34, 8
    say("update_s fdorder", fdorder, "free_surf", free_surf, "L", L)
    cfg, o, P, hc, pml, F, nd = _setup(nx, ny, fdorder, L, free_surf, fw)
    dn = product()
    dn.set_globals(NX=nx, NY=ny, FDORDER=fdorder, FW=fw, FREE_SURF=free_surf, L=L, DT=1e-4, DH=0.25, NT=10)
    vy = _nr(dn, F["vy"], nd, nx, ny)
    syz, sxy = _nr(dn, F["syz"], nd, nx, ny), _nr(dn, F["sxy"], nd, nx, ny)
    uy, uyx = _nr(dn, np.zeros_like(F["vy"]), nd, nx, ny), _nr(dn, np.zeros_like(F["vy"]), nd, nx, ny)
    psy = dn.NRMatrix(1, 2 * fw, 1, nx, F["psi_a"].reshape(2 * fw, nx).copy())
    psx = dn.NRMatrix(1, ny, 1, 2 * fw, F["psi_b"].reshape(ny, 2 * fw).copy())
    hcv = dn.NRVector(0, 6, hc.astype(np.float32))
    vecs = [_vec(dn, pml[k].astype(np.float32), fw) for k in
            ("K_x", "a_x", "b_x", "K_x_half", "a_x_half", "b_x_half", "K_y", "a_y", "b_y", "K_y_half", "a_y_half",
             "b_y_half")]
    Lib = dn.lib()
    ruy = np.zeros_like(F["vy"]); ruyx = np.zeros_like(F["vy"])
    rsxy = F["sxy"].copy(); rsyz = F["syz"].copy(); ra = F["psi_a"].copy(); rb = F["psi_b"].copy()
    if L == 0:
        uip, ujp = _nr(dn, P["uip"].astype(np.float32), nd, nx, ny), _nr(dn, P["ujp"].astype(np.float32), nd, nx, ny)
        Lib.update_s_elastic_PML.restype = None
        Lib.update_s_elastic_PML(1, nx, 1, ny, vy.ptr, uy.ptr, uyx.ptr, syz.ptr, sxy.ptr, None, uip.ptr, ujp.ptr,
                                 None, None, hcv.ptr, 0, *[v.ptr for v in vecs], psy.ptr, psx.ptr)
        o.update_s_elastic(cfg, F["vy"], ruy, ruyx, rsyz, rsxy, P["uip"], P["ujp"], hc, pml, ra, rb)
    else:
        def t3(a):  # [L][j][i] padded -> f3tensor [j][i][1..L]
            return dn.NRTensor3(-nd + 1, ny + nd, -nd + 1, nx + nd, 1, L, np.moveaxis(a[:L], 0, -1).copy())
        r, q = t3(F["r"]), t3(F["q"])
        dip, dd = t3(P["dip"].astype(np.float32)), t3(P["d"].astype(np.float32))
        fipjp, f = _nr(dn, P["fipjp"].astype(np.float32), nd, nx, ny), _nr(dn, P["f"].astype(np.float32), nd, nx, ny)
        vv = [dn.NRVector(1, L, P[k][:L].astype(np.float32)) for k in ("bip", "bjm", "cip", "cjm")]
        Lib.update_s_visc_PML.restype = None
        Lib.update_s_visc_PML(1, nx, 1, ny, vy.ptr, uy.ptr, uyx.ptr, syz.ptr, sxy.ptr, None, None, None, hcv.ptr, 0,
                              r.ptr, None, q.ptr, fipjp.ptr, f.ptr, None, vv[0].ptr, vv[1].ptr, vv[2].ptr, vv[3].ptr,
                              dd.ptr, None, dip.ptr, *[v.ptr for v in vecs], psy.ptr, psx.ptr)
        rr = F["r"].copy(); rq = F["q"].copy()
        o.update_s_visc(cfg, F["vy"], ruy, ruyx, rsyz, rsxy, rr, rq, P["fipjp"], P["f"], P["dip"], P["d"], P["bip"],
                        P["bjm"], P["cip"], P["cjm"], hc, pml, ra, rb)
        inner3 = (slice(nd, nd + ny), slice(nd, nd + nx))
        assert np.array_equal(np.moveaxis(r.a, -1, 0)[(slice(None),) + inner3], rr[:L][(slice(None),) + inner3])
        assert np.array_equal(np.moveaxis(q.a, -1, 0)[(slice(None),) + inner3], rq[:L][(slice(None),) + inner3])
    inner = (slice(nd, nd + ny), slice(nd, nd + nx))
    for name, got, ref in (("sxy", sxy.a[inner], rsxy[inner]), ("syz", syz.a[inner], rsyz[inner]),
                           ("uy", uy.a[inner], ruy[inner]), ("uyx", uyx.a[inner], ruyx[inner]),
                           ("psi_vyy", psy.a.ravel(), ra), ("psi_vyx", psx.a.ravel(), rb)):
        d = np.max(np.abs(got.astype(np.float64) - ref))
        say("  %s max |diff| %.3e" % (name, d))
        assert np.array_equal(got, ref.astype(np.float32)), name


# --------------------------------------------------------------------------
# the drop-ins driven from C (tests/dropin_caller.c) like denise.c's loop
# --------------------------------------------------------------------------
CALLER = os.path.join(ROOT, "tests", "dropin_caller")


def _caller_input(path, nx, ny, fdorder, fw, free_surf, L, nt, mode, seed=3):
    from pyoracle import make_cfg
    o = oracle("f32")
    fl = [6.25, 50.0, 12.0, 30.0][:L]
    cfg = make_cfg(nx, ny, fdorder=fdorder, fw=fw, free_surf=free_surf, L=L, nt=nt, dt=1e-4, dh=0.25,
                   damping=300.0, fpml=40.0, fl=fl)
    vs, rho, q = model(nx, ny, 0.25, seed=seed, qs=20.0 if L else None)
    P = o.model_setup(cfg, vs, rho, q)
    nd = fdorder // 2 + 1
    I = (slice(nd, nd + ny), slice(nd, nd + nx))
    hc = o.holberg(fdorder, 1).astype(np.float32)
    pml = o.pml(cfg)
    sig = o.wavelet(1, nt, 1e-4, 0.0, 40.0, 1.0).astype(np.float32)
    rec_i = np.arange(fw + 3, nx - fw - 2, 7, dtype=np.int32)
    rec_j = np.full_like(rec_i, 4)
    src = (nx // 3, 5)
    with open(path, "wb") as f:
        np.array([nx, ny, fdorder, fw, free_surf, L, nt, len(rec_i), mode, src[0], src[1]], np.int32).tofile(f)
        np.concatenate([[1e-4, 0.25], hc]).astype(np.float32).tofile(f)
        for k in ("K_x", "a_x", "b_x", "K_x_half", "a_x_half", "b_x_half", "K_y", "a_y", "b_y", "K_y_half",
                  "a_y_half", "b_y_half"):
            pml[k][:2 * fw + 1].astype(np.float32).tofile(f)
        planes = [P["rho"]] + ([P["uip"], P["ujp"]] if L == 0 else [P["fipjp"], P["f"]] + [P["dip"][l] for l in range(L)]
                                + [P["d"][l] for l in range(L)])
        for a in planes:
            np.ascontiguousarray(a[I], np.float32).tofile(f)
        if L:
            for k in ("bip", "bjm", "cip", "cjm"):
                P[k][:L].astype(np.float32).tofile(f)
        sig.tofile(f)
        rec_i.tofile(f)
        rec_j.tofile(f)
    seis = o.forward(cfg, vs, rho, q, np.array([src[0]], np.int32), np.array([src[1]], np.int32), sig[None],
                     rec_i, rec_j)[0]
    return seis.astype(np.float32), len(rec_i)


def _run_caller(tmp, nx, ny, nt, ntr, mode):
    import subprocess
    out = os.path.join(tmp, "out%d.bin" % mode)
    r = subprocess.run([CALLER, os.path.join(tmp, "in%d.bin" % mode), out], capture_output=True, text=True,
                       timeout=300)
    say("  " + r.stdout.strip() + (" | " + r.stderr.strip() if r.stderr.strip() else ""))
    assert r.returncode == 0, r.stderr
    a = np.fromfile(out, np.float32)
    ms = float(r.stdout.split(";")[-1].split()[0])
    return a[:ntr * nt].reshape(ntr, nt), a[ntr * nt:].reshape(ny, nx), ms


@pytest.mark.parametrize("L,fdorder,free_surf", [(0, 4, 1), (2, 8, 0)])
def test_c_caller_all_modes_bit_equal_to_oracle(tmp_path, L, fdorder, free_surf):
    """denise.c's forward loop in C through the drop-ins: copy, resident and
    resident with deferred seismograms give seismograms bit-equal to the
    oracle's forward and the same final wavefield."""
    assert os.path.exists(CALLER), "tests/dropin_caller not built (python -c 'import __graft_entry__ as g; g.build()')"
    nx, ny, nt, fw = 90, 50, 300, 10
    say("C caller L", L, "FDORDER", fdorder, "FREE_SURF", free_surf)
    finals = []
    for mode in (0, 1, 2):
        seis, ntr = _caller_input(os.path.join(tmp_path, "in%d.bin" % mode), nx, ny, fdorder, fw, free_surf, L, nt,
                                  mode)
        got, vy, _ = _run_caller(str(tmp_path), nx, ny, nt, ntr, mode)
        assert np.array_equal(got, seis), "mode %d" % mode
        finals.append(vy)
    assert np.array_equal(finals[0], finals[1]) and np.array_equal(finals[0], finals[2])


def test_c_caller_resident_speed(tmp_path):
    """The same loop at 1200 x 300 (visco L = 2, FDORDER 8): step time of the
    three modes, the seismograms of the resident modes bit-equal to the
    oracle's; resident steps run without the per-call host copies.  The
    step time includes the final download of the wavefields, spread over
    the 1000 steps."""
    assert os.path.exists(CALLER)
    nx, ny, nt, fw = 1200, 300, 1000, 20
    ms = {}
    for mode in (0, 1, 2):
        seis, ntr = _caller_input(os.path.join(tmp_path, "in%d.bin" % mode), nx, ny, 8, fw, 1, 2, nt, mode)
        got, _, ms[mode] = _run_caller(str(tmp_path), nx, ny, nt, ntr, mode)
        assert np.array_equal(got, seis), "mode %d" % mode
    say("  ms per step: copy %.3f, resident %.3f, deferred %.3f" % (ms[0], ms[1], ms[2]))
    assert ms[1] < ms[0] and ms[2] < ms[0]
