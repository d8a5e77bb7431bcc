"""BOUNDARY 1 (periodic left and right edges, exchange_v.c:99-161 and
exchange_s.c:95-157 on one domain, no x-CPML) on the GPU: the fused step's
seismograms and gradients against the FP32 oracle (bit for bit), translation
invariance of the fused step itself (model, sources and receivers rolled
across the wrap and across strip boundaries), and the reference-named
drop-ins against the oracle's sweeps with the reference's halo exchange."""
import sys

import numpy as np
import pytest

from common import model, oracle, product, rel_l2, ricker

pytestmark = pytest.mark.gpu

DT, DH = 1e-4, 0.25


def say(*a):
    print(*a)
    sys.stdout.flush()


def _case(nx, ny, fdorder, L, free_surf, fw, nt, fc, nshots=3, seed=11, vs=None):
    dn = product()
    qs = 20.0 if L else None
    v, rho, q = model(nx, ny, DH, seed=seed, qs=qs)
    vs = v if vs is None else vs
    fl = [6.25, 50.0][:L]
    ctx = dn.Context(nx, ny, fdorder=fdorder, maxrelerror=1, fw=fw, free_surf=free_surf, L=L, fl=fl, dt=DT, dh=DH,
                     damping=300.0, fpml=fc, dtinv=2, boundary=1)
    rec_i = np.unique(np.concatenate([np.arange(1, nx + 1, 7), [2, nx - 1, nx]])).astype(np.int32)
    rec_j = np.full_like(rec_i, 2)
    src_i = np.array([1, nx // 2 + 3, nx][:nshots], np.int32)
    src_j = np.full(nshots, 3, np.int32)
    sig = np.stack([ricker(nt, DT, fc, tshift=0.001 * s) for s in range(nshots)])
    ctx.set_receivers(rec_i, rec_j)
    ctx.set_shots(src_i, src_j, sig)
    ctx.set_model(vs, rho, q)
    from pyoracle import make_cfg
    cfg = make_cfg(nx, ny, fdorder=fdorder, maxrelerror=1, fw=fw, free_surf=free_surf, L=L, nt=nt, dtinv=2, dt=DT,
                   dh=DH, damping=300.0, fpml=fc, fl=fl, boundary=1)
    return ctx, cfg, dict(vs=vs, rho=rho, qs=q, rec_i=rec_i, rec_j=rec_j, src_i=src_i, src_j=src_j, sig=sig)


# name, nx, ny, fdorder, L, free_surf, fw, nt
CASES = [
    ("el_fd4_fs_fw10", 150, 40, 4, 0, 1, 10, 500),
    ("visc_fd8_nofs_fw10", 150, 44, 8, 2, 0, 10, 500),
    ("visc_fd8_fs_fw0", 130, 36, 8, 2, 1, 0, 400),
    ("el_fd12_fs_fw8", 97, 36, 12, 0, 1, 8, 300),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_periodic_seismograms_and_gradient_match_oracle(case):
    name, nx, ny, fdorder, L, fs, fw, nt = case
    say("case", name)
    ctx, cfg, d = _case(nx, ny, fdorder, L, fs, fw, nt, 60.0)
    o32, o64 = oracle("f32"), oracle("f64")
    seis = ctx.forward()
    for s in range(len(d["src_i"])):
        ref, _, _ = o32.forward(cfg, d["vs"], d["rho"], d["qs"], d["src_i"][s:s + 1], d["src_j"][s:s + 1], d["sig"][s],
                                d["rec_i"], d["rec_j"])
        r64, _, _ = o64.forward(cfg, d["vs"], d["rho"], d["qs"], d["src_i"][s:s + 1], d["src_j"][s:s + 1], d["sig"][s],
                                d["rec_i"], d["rec_j"])
        say("  shot %d (source column %d): seismograms vs oracle f32 %.3e, vs f64 %.3e" % (
            s, d["src_i"][s], rel_l2(seis[s], ref), rel_l2(seis[s], r64)))
        assert np.max(np.abs(ref)) > 0 and np.array_equal(seis[s], ref.astype(np.float32))
    # gradient: observed data from the model 3 % faster
    vt = (d["vs"] * np.float32(1.03)).astype(np.float32)
    obs = np.stack([o32.forward(cfg, vt, d["rho"], d["qs"], d["src_i"][s:s + 1], d["src_j"][s:s + 1], d["sig"][s],
                                d["rec_i"], d["rec_j"])[0] for s in range(len(d["src_i"]))]).astype(np.float32)
    ctx.set_observed(obs)
    l2, en = ctx.gradient()
    gvs, grho, _, _ = ctx.get_gradient()
    r32 = o32.gradient(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"], d["rec_j"], obs)
    r64 = o64.gradient(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"], d["rec_j"], obs)
    say("  gradient vs oracle f32 %.3e / %.3e, vs f64 %.3e; l2 %.9e oracle %.9e" % (
        rel_l2(gvs, r32["grad_vs"]), rel_l2(grho, r32["grad_rho"]), rel_l2(gvs, r64["grad_vs"]), l2, r32["l2"]))
    assert np.array_equal(gvs, r32["grad_vs"].astype(np.float32))
    assert np.array_equal(grho, r32["grad_rho"].astype(np.float32))
    ctx.close()


@pytest.mark.parametrize("shift", [1, 57, 120])
def test_fused_step_is_translation_invariant(shift):
    """Rolling model, sources and receivers by `shift` columns on a periodic
    grid leaves the fused step's seismograms unchanged bit for bit (the
    wrap, the strips of 57 owned columns and the halo lanes all move)."""
    nx, ny, nt = 150, 40, 400
    say("case roll by", shift)
    dn = product()
    v, rho, q = model(nx, ny, DH, seed=4, qs=20.0)
    roll = lambda a: np.ascontiguousarray(np.roll(a, shift, axis=1))
    mv = lambda i: ((np.asarray(i) - 1 + shift) % nx + 1).astype(np.int32)
    rec_i = np.array([1, 5, 60, 100, 149, 150], np.int32)
    src_i = np.array([2, 140], np.int32)
    sig = np.stack([ricker(nt, DT, 60.0), ricker(nt, DT, 60.0, tshift=0.002)])
    out = []
    for k in range(2):
        ctx = dn.Context(nx, ny, fdorder=8, maxrelerror=1, fw=10, free_surf=1, L=2, fl=[6.25, 50.0], dt=DT, dh=DH,
                         damping=300.0, fpml=60.0, boundary=1)
        if k == 0:
            ctx.set_model(v, rho, q)
            ctx.set_receivers(rec_i, np.full(len(rec_i), 2, np.int32))
            ctx.set_shots(src_i, np.full(2, 3, np.int32), sig)
        else:
            ctx.set_model(roll(v), roll(rho), roll(q))
            ctx.set_receivers(mv(rec_i), np.full(len(rec_i), 2, np.int32))
            ctx.set_shots(mv(src_i), np.full(2, 3, np.int32), sig)
        out.append(ctx.forward())
        ctx.close()
    assert np.max(np.abs(out[0])) > 0
    assert np.array_equal(out[0], out[1])


def _nr(dn, a, nd, nx, ny):
    return dn.NRMatrix(-nd + 1, ny + nd, -nd + 1, nx + nd, a.copy())


@pytest.mark.parametrize("fdorder,L", [(4, 0), (8, 2)])
def test_periodic_dropins_match_oracle_sweeps(fdorder, L):
    """update_v_PML and update_s_*_PML with BOUNDARY 1 against the oracle's
    sweeps after the reference's exchange (orc_wrap_x); the host arrays are
    passed with stale halos, the library wraps by itself."""
    nx, ny, fw = 70, 30, 8
    say("drop-ins, BOUNDARY 1, FDORDER", fdorder, "L", L)
    from pyoracle import make_cfg
    cfg = make_cfg(nx, ny, fdorder=fdorder, fw=fw, free_surf=1, L=L, nt=10, dt=DT, dh=DH, damping=300.0,
                   fpml=50.0, fl=[6.25, 50.0][:L], boundary=1)
    o = oracle("f32")
    vs, rho, q = model(nx, ny, DH, seed=9, qs=20.0 if L else None)
    P = o.model_setup(cfg, vs, rho, q)
    hc = o.holberg(fdorder, 1)
    pml = o.pml(cfg)
    rng = np.random.default_rng(9)
    sh = o.padded_shape(cfg)
    nd = fdorder // 2 + 1
    inner = (slice(nd, nd + ny), slice(nd, nd + nx))
    F = {}
    for k, s in (("vy", 1e-3), ("sxy", 1.0), ("syz", 1.0)):
        a = np.zeros(sh, np.float32)
        a[inner] = (rng.standard_normal((ny, nx)) * s).astype(np.float32)
        F[k] = a
    dn = product()
    dn.set_globals(NX=nx, NY=ny, FDORDER=fdorder, FW=fw, FREE_SURF=1, L=L, DT=DT, DH=DH, NT=10, BOUNDARY=1)
    Lib = dn.lib()
    hcv = dn.NRVector(0, 6, hc.astype(np.float32))
    vecs = [dn.NRVector(1, 2 * fw, pml[k][1:2 * fw + 1].astype(np.float32).copy()) for k in
            ("K_x", "a_x", "b_x", "K_x_half", "a_x_half", "b_x_half", "K_y", "a_y", "b_y", "K_y_half", "a_y_half",
             "b_y_half")]
    psa = np.zeros(2 * fw * nx, np.float32)
    psb = np.zeros(2 * fw * ny, np.float32)
    # velocity
    vy, vyp1 = _nr(dn, F["vy"], nd, nx, ny), _nr(dn, np.zeros_like(F["vy"]), nd, nx, ny)
    syz, sxy, rh = _nr(dn, F["syz"], nd, nx, ny), _nr(dn, F["sxy"], nd, nx, ny), _nr(dn, P["rho"].astype(np.float32), nd, nx, ny)
    psy = dn.NRMatrix(1, 2 * fw, 1, nx, psa.reshape(2 * fw, nx).copy())
    psx = dn.NRMatrix(1, ny, 1, 2 * fw, psb.reshape(ny, 2 * fw).copy())
    srcpos = dn.NRMatrix(1, 8, 1, 1, np.array([[1], [5], [0], [0], [50], [1], [0], [3]], np.float32))
    signals = dn.NRMatrix(1, 1, 1, 10, np.zeros((1, 10), np.float32))
    Lib.update_v_PML.restype = None
    Lib.update_v_PML(1, nx, 1, ny, 3, vy.ptr, vyp1.ptr, None, None, syz.ptr, sxy.ptr, rh.ptr, srcpos.ptr,
                     signals.ptr, 1, None, hcv.ptr, 0, 0, *[v.ptr for v in vecs], psy.ptr, psx.ptr)
    rsxy = F["sxy"].copy()
    o.wrap_x(cfg, rsxy)
    rv, rvp = F["vy"].copy(), np.zeros_like(F["vy"])
    o.update_v(cfg, 0, rv, rvp, F["syz"], rsxy, P["rho"], hc, pml, psa.copy(), psb.copy())
    assert np.array_equal(vy.a[inner], rv[inner]) and np.array_equal(vyp1.a[inner], rvp[inner])
    # stress
    rvy = F["vy"].copy()
    o.wrap_x(cfg, rvy)
    uy, uyx = _nr(dn, np.zeros_like(F["vy"]), nd, nx, ny), _nr(dn, np.zeros_like(F["vy"]), nd, nx, ny)
    vy2 = _nr(dn, F["vy"], nd, nx, ny)
    syz2, sxy2 = _nr(dn, F["syz"], nd, nx, ny), _nr(dn, F["sxy"], nd, nx, ny)
    psy2 = dn.NRMatrix(1, 2 * fw, 1, nx, psa.reshape(2 * fw, nx).copy())
    psx2 = dn.NRMatrix(1, ny, 1, 2 * fw, psb.reshape(ny, 2 * fw).copy())
    ruy, ruyx = np.zeros_like(F["vy"]), np.zeros_like(F["vy"])
    rsxy2, rsyz2 = F["sxy"].copy(), F["syz"].copy()
    if L == 0:
        uip, ujp = _nr(dn, P["uip"].astype(np.float32), nd, nx, ny), _nr(dn, P["ujp"].astype(np.float32), nd, nx, ny)
        Lib.update_s_elastic_PML.restype = None
        Lib.update_s_elastic_PML(1, nx, 1, ny, vy2.ptr, uy.ptr, uyx.ptr, syz2.ptr, sxy2.ptr, None, uip.ptr, ujp.ptr,
                                 None, None, hcv.ptr, 0, *[v.ptr for v in vecs], psy2.ptr, psx2.ptr)
        o.update_s_elastic(cfg, rvy, ruy, ruyx, rsyz2, rsxy2, P["uip"], P["ujp"], hc, pml, psa.copy(), psb.copy())
    else:
        def t3(a):
            return dn.NRTensor3(-nd + 1, ny + nd, -nd + 1, nx + nd, 1, L, np.moveaxis(a[:L], 0, -1).copy())
        z = np.zeros((L,) + sh, np.float32)
        r, qq = t3(z), t3(z)
        dip, dd = t3(P["dip"].astype(np.float32)), t3(P["d"].astype(np.float32))
        fipjp, f = _nr(dn, P["fipjp"].astype(np.float32), nd, nx, ny), _nr(dn, P["f"].astype(np.float32), nd, nx, ny)
        vv = [dn.NRVector(1, L, P[k][:L].astype(np.float32)) for k in ("bip", "bjm", "cip", "cjm")]
        Lib.update_s_visc_PML.restype = None
        Lib.update_s_visc_PML(1, nx, 1, ny, vy2.ptr, uy.ptr, uyx.ptr, syz2.ptr, sxy2.ptr, None, None, None, hcv.ptr, 0,
                              r.ptr, None, qq.ptr, fipjp.ptr, f.ptr, None, vv[0].ptr, vv[1].ptr, vv[2].ptr, vv[3].ptr,
                              dd.ptr, None, dip.ptr, *[v.ptr for v in vecs], psy2.ptr, psx2.ptr)
        o.update_s_visc(cfg, rvy, ruy, ruyx, rsyz2, rsxy2, z.copy(), z.copy(), P["fipjp"], P["f"], P["dip"], P["d"],
                        P["bip"], P["bjm"], P["cip"], P["cjm"], hc, pml, psa.copy(), psb.copy())
    for nm, got, ref in (("sxy", sxy2.a[inner], rsxy2[inner]), ("syz", syz2.a[inner], rsyz2[inner]),
                         ("uy", uy.a[inner], ruy[inner]), ("uyx", uyx.a[inner], ruyx[inner])):
        say("  %s max |diff| %.3e" % (nm, np.max(np.abs(got.astype(np.float64) - ref))))
        assert np.array_equal(got, ref.astype(np.float32)), nm
    dn.set_globals(NX=nx, NY=ny, FDORDER=fdorder, FW=fw, FREE_SURF=1, L=L, DT=DT, DH=DH, NT=10)
