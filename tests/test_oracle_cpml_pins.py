"""Pins of the oracle's CPML (oracle/sh_oracle.c: orc_pml_pro and the CPML
branches of orc_update_v / orc_update_s_elastic / orc_update_s_visc) that do
not rest on reading its code.  CPU only.

What pins what:
  profiles      every array entry of d, K, alpha, b and a recomputed in
                float64 from the grid coordinate of the point it belongs to
                (full grid x = (i-1) DH, half grid x + DH/2, distance into the
                layer measured from its inner edge) and the published forms:
                d = d0 (s/L)^N with d0 = -(N+1) V ln(R) / (2L), R = 0.001
                (INRIA RR-6020 / Komatitsch & Martin 2007), K = 1 + (Kmax-1)
                (s/L)^N, alpha = pi FPML (1 - s/L) + 0.1 pi FPML (Festa &
                Vilotte), b = exp(-(d/K + alpha) DT), a = d (b-1) / (K (d +
                K alpha)) where |int(d)| > 0 (the reference's integer abs()).
                Reference: PML_pro.c:103-297.
  recursion     psi(n) = b psi(n-1) + a u with u = 1 switched on at n = 1
                equals the convolution of u with the CPML memory kernel
                -(d/K^2) exp(-(d/K + alpha) t) sampled at n DT, for every n:
                fixes b's exponent and a together (the steady-state identity
                alone holds for any alpha and b).
  reflection    one source in a homogeneous medium on a 100 x 100 grid with
                20-point CPML on all four sides (FREE_SURF 0), against the same
                source on a grid large enough that no edge arrival enters the
                record: the seismogram residual stays below a bound that a
                full/half-grid profile swap, a missing psi term in either
                kernel, a missing 1/K (K_MAX_PML = 2), a mirrored right strip,
                a halved a or a doubled decay each exceed (each mistake is
                planted in the profiles the oracle's sweeps read).  Elastic
                and viscoelastic, FDORDER 4 and 8.
                Reference: update_v_PML.c:142-179, update_s_elastic_PML.c:
                116-150, update_s_visc_PML.c:61-93.
  visco top     with taus = 0 and no free surface the viscoelastic stress
                update equals the elastic one everywhere except the top FW
                rows, and there it equals the elastic update fed the full-grid
                y profile on its top strip: the reference's visco kernel reads
                b_y / a_y / K_y on the top strip and the half-grid profile on
                the bottom one (update_s_visc_PML.c:77-93).
"""
import numpy as np
import pytest

from common import oracle, rel_l2

O32 = oracle("f32")
O64 = oracle("f64")


def cfg(**kw):
    from pyoracle import make_cfg
    return make_cfg(**kw)


# --------------------------------------------------------------------------
# profiles from the grid coordinates
# --------------------------------------------------------------------------
def brute_force_profiles(n, fw, dh, dt, damping, fpml, npower=2.0, kmax=1.0):
    """d, K, alpha, a, b of the full and half grid for one axis of n points,
    indexed like the oracle's arrays: 1..fw for the strip at index 1,
    fw+1..2fw for the strip at index n (entry 0 unused)."""
    dh, dt, damping, fpml, npower, kmax = (float(np.float32(v)) for v in (dh, dt, damping, fpml, npower, kmax))
    Lw = fw * dh
    d0 = -(npower + 1.0) * damping * np.log(0.001) / (2.0 * Lw)
    amax = np.pi * fpml
    gidx = np.concatenate([np.arange(1, fw + 1), np.arange(n - fw + 1, n + 1)])   # global point of each entry
    low = np.arange(2 * fw) < fw
    out = {}
    for half, sfx in ((0.0, ""), (0.5, "_half")):
        pos = (gidx - 1) * dh + half * dh
        s = np.where(low, Lw - pos, pos - ((n - 1) * dh - Lw))                       # distance into the layer
        inside = s >= 0.0
        sn = np.where(inside, s / Lw, 0.0)
        d = np.where(inside, d0 * sn ** npower, 0.0)
        K = np.where(inside, 1.0 + (kmax - 1.0) * sn ** npower, 1.0)
        al = np.where(inside, amax * (1.0 - sn) + 0.1 * amax, 0.0)
        al = np.maximum(al, 0.0)
        b = np.exp(-(d / K + al) * dt)
        a = np.where(np.abs(np.trunc(d)) > 0, d * (b - 1.0) / (K * (d + K * al)), 0.0)
        for q, v in (("d", d), ("K", K), ("alpha", al), ("a", a), ("b", b)):
            full = np.zeros(2 * fw + 1)
            full[1:] = v
            out[q + sfx] = full
    return out


PROFILE_CFGS = [dict(nx=400, ny=120, fw=20, dh=0.25, dt=1e-4, damping=300.0, fpml=40.0, npower=2.0, kmax=1.0),
                dict(nx=1200, ny=300, fw=20, dh=0.2, dt=1e-4, damping=350.0, fpml=25.0, npower=2.0, kmax=2.0),
                dict(nx=96, ny=64, fw=10, dh=25.0, dt=1e-2, damping=1000.0, fpml=6.5, npower=3.0, kmax=1.5),
                dict(nx=50, ny=70, fw=7, dh=1.0, dt=2e-4, damping=3000.0, fpml=15.0, npower=2.0, kmax=1.0)]


@pytest.mark.parametrize("pc", PROFILE_CFGS)
def test_profiles_from_grid_coordinates(pc):
    c = cfg(nx=pc["nx"], ny=pc["ny"], fw=pc["fw"], dh=pc["dh"], dt=pc["dt"], damping=pc["damping"], fpml=pc["fpml"],
            npower=pc["npower"], k_max_pml=pc["kmax"], fdorder=4)
    fw = pc["fw"]
    for o, rtol in ((O64, 1e-12), (O32, 1e-4)):     # float32 position arithmetic
        P = o.pml(c)
        for ax, n in (("x", pc["nx"]), ("y", pc["ny"])):
            ref = brute_force_profiles(n, fw, pc["dh"], pc["dt"], pc["damping"], pc["fpml"], pc["npower"], pc["kmax"])
            for q in ("d", "K", "alpha", "b", "a"):
                for sfx in ("", "_half"):
                    got = P["%s_%s%s" % (q, ax, sfx)][1:2 * fw + 1]
                    exp = ref[q + sfx][1:]
                    np.testing.assert_allclose(got, exp, rtol=rtol, atol=rtol * np.max(np.abs(exp)) * 1e-3,
                                               err_msg="%s_%s%s" % (q, ax, sfx))


@pytest.mark.parametrize("pc", PROFILE_CFGS[:3])
def test_psi_recursion_is_the_cpml_convolution(pc):
    c = cfg(nx=pc["nx"], ny=pc["ny"], fw=pc["fw"], dh=pc["dh"], dt=pc["dt"], damping=pc["damping"], fpml=pc["fpml"],
            npower=pc["npower"], k_max_pml=pc["kmax"], fdorder=4)
    fw = pc["fw"]
    dt = float(np.float32(pc["dt"]))
    P = O64.pml(c)
    ref = brute_force_profiles(pc["nx"], fw, pc["dh"], pc["dt"], pc["damping"], pc["fpml"], pc["npower"], pc["kmax"])
    for sfx in ("", "_half"):
        a, b = P["a_x" + sfx][1:2 * fw + 1], P["b_x" + sfx][1:2 * fw + 1]
        d, K, al = (ref[q + sfx][1:] for q in ("d", "K", "alpha"))
        use = np.trunc(d) != 0
        psi = np.zeros(2 * fw)
        for n in range(1, 301):
            psi = b * psi + a * 1.0
            if n in (1, 2, 7, 40, 300):
                # integral over [0, n DT] of -(d/K^2) exp(-(d/K + alpha) t)
                rate = d / K + al
                exact = -(d / (K * K)) * (1.0 - np.exp(-rate * n * dt)) / rate
                np.testing.assert_allclose(psi[use], exact[use], rtol=1e-10)
        assert np.all(psi[~use] == 0.0)


# --------------------------------------------------------------------------
# reflection of the absorbing frame
# --------------------------------------------------------------------------
N_SMALL, FW, EXTRA = 100, 20, 50
DH, DT, NT, FC, V0 = 0.5, 5e-4, 800, 20.0, 200.0
REFLECTION_BOUND = 2.5e-3
_POS = np.array([22, 30, 40, 50, 60, 70, 79])
REC_I = np.concatenate([_POS, _POS, _POS, np.full(7, 22), np.full(7, 79)]).astype(np.int32)
REC_J = np.concatenate([np.full(7, 22), np.full(7, 50), np.full(7, 79), _POS, _POS]).astype(np.int32)


def _homogeneous(n, L):
    vs = np.full((n, n), V0, np.float32)
    rho = np.full((n, n), 1800.0, np.float32)
    return vs, rho, (np.full((n, n), 20.0, np.float32) if L else None)


def sweep_forward(o, c, vs, rho, qs, si, sj, sig, ri, rj, pml):
    """One shot through the oracle's single sweeps (update_v, point force,
    stress update) with the CPML profiles given, recording vy."""
    nd = c.fdorder // 2 + 1
    P = o.model_setup(c, vs, rho, qs)
    hc = o.holberg(c.fdorder, c.maxrelerror)
    sh = o.padded_shape(c)
    vy, vyp1, sxy, syz, uy, uyx = (np.zeros(sh, o.dt) for _ in range(6))
    Lm = max(c.L, 1)
    r = np.zeros((Lm,) + sh, o.dt)
    q = np.zeros_like(r)
    psi_v = [np.zeros(2 * c.fw * c.nx, o.dt), np.zeros(2 * c.fw * c.ny, o.dt)]
    psi_s = [np.zeros(2 * c.fw * c.nx, o.dt), np.zeros(2 * c.fw * c.ny, o.dt)]
    seis = np.zeros((len(ri), c.nt), o.dt)
    for n in range(c.nt):
        o.update_v(c, 0, vy, vyp1, syz, sxy, P["rho"], hc, pml, psi_v[0], psi_v[1])
        vy[sj + nd - 1, si + nd - 1] += sig[n]
        if c.L > 0:
            o.update_s_visc(c, vy, uy, uyx, syz, sxy, r, q, P["fipjp"], P["f"], P["dip"], P["d"], P["bip"], P["bjm"],
                            P["cip"], P["cjm"], hc, pml, psi_s[0], psi_s[1])
        else:
            o.update_s_elastic(c, vy, uy, uyx, syz, sxy, P["uip"], P["ujp"], hc, pml, psi_s[0], psi_s[1])
        seis[:, n] = vy[rj + nd - 1, ri + nd - 1]
    return seis


def planted_mistakes(P, kmax):
    """CPML mistakes a restatement could plausibly make, as the profile
    arrays its sweeps would then effectively read."""
    out = {}
    Q = dict(P)
    for q in ("d", "K", "alpha", "a", "b"):
        for ax in ("x", "y"):
            Q["%s_%s" % (q, ax)], Q["%s_%s_half" % (q, ax)] = P["%s_%s_half" % (q, ax)], P["%s_%s" % (q, ax)]
    out["full/half-grid profiles swapped"] = Q
    for key, name in (("a_x", "velocity x"), ("a_y", "velocity y"), ("a_x_half", "stress x"),
                      ("a_y_half", "stress y (bottom)")):
        Q = dict(P)
        Q[key] = np.zeros_like(P[key])
        out["psi term missing in %s" % name] = Q
    if kmax > 1.0:
        Q = dict(P)
        for k in ("K_x", "K_y", "K_x_half", "K_y_half"):
            Q[k] = np.ones_like(P[k])
        out["1/K missing"] = Q
    Q = dict(P)
    for k in ("K_x", "a_x", "b_x"):
        v = P[k].copy()
        v[FW + 1:2 * FW + 1] = v[FW + 1:2 * FW + 1][::-1].copy()
        Q[k] = v
    out["right strip mirrored"] = Q
    Q = dict(P)
    for k in ("a_x", "a_x_half", "a_y", "a_y_half"):
        Q[k] = 0.5 * P[k]
    out["a halved"] = Q
    Q = dict(P)
    for k in ("b_x", "b_x_half", "b_y", "b_y_half"):
        Q[k] = P[k] * P[k]
    out["decay per step doubled"] = Q
    return out


@pytest.mark.parametrize("L,fdorder,kmax", [(0, 4, 1.0), (0, 4, 2.0), (2, 4, 1.0), (2, 8, 2.0)])
def test_cpml_reflection_and_planted_mistakes(L, fdorder, kmax):
    fl = [FC / 4, 2 * FC] if L else []
    cs = cfg(nx=N_SMALL, ny=N_SMALL, fdorder=fdorder, fw=FW, free_surf=0, L=L, nt=NT, dt=DT, dh=DH, damping=V0,
             fpml=FC, fl=fl, k_max_pml=kmax)
    nb = N_SMALL + 2 * EXTRA
    cb = cfg(nx=nb, ny=nb, fdorder=fdorder, fw=0, free_surf=0, L=L, nt=NT, dt=DT, dh=DH, damping=V0, fpml=FC, fl=fl)
    sig = O32.wavelet(1, NT, DT, 0.0, FC, 1.0).astype(np.float32)
    # the big grid: the nearest edge is 100 cells from the source, its first
    # arrival at any receiver comes after 0.43 s > NT DT = 0.4 s
    vs, rho, qs = _homogeneous(nb, L)
    big, _, _ = O32.forward(cb, vs, rho, qs, np.array([50 + EXTRA], np.int32), np.array([50 + EXTRA], np.int32),
                            sig[None], REC_I + EXTRA, REC_J + EXTRA)
    vs, rho, qs = _homogeneous(N_SMALL, L)
    src = (np.array([50], np.int32), np.array([50], np.int32))
    small, _, _ = O32.forward(cs, vs, rho, qs, src[0], src[1], sig[None], REC_I, REC_J)
    P = O32.pml(cs)
    loop = sweep_forward(O32, cs, vs, rho, qs, np.int32(50), np.int32(50), sig, REC_I, REC_J, P)
    assert np.array_equal(loop, small)           # the sweep loop is the oracle's own shot
    res = rel_l2(small, big)
    print("L=%d FDORDER=%d K_MAX=%.0f: CPML residual %.3e (bound %.1e)" % (L, fdorder, kmax, res, REFLECTION_BOUND))
    assert res < REFLECTION_BOUND
    for name, Q in planted_mistakes(P, kmax).items():
        bad = rel_l2(sweep_forward(O32, cs, vs, rho, qs, np.int32(50), np.int32(50), sig, REC_I, REC_J, Q), big)
        print("  planted: %-36s residual %.3e" % (name, bad))
        assert bad > REFLECTION_BOUND, name


# --------------------------------------------------------------------------
# visco top strip: full-grid profile (Q2)
# --------------------------------------------------------------------------
def test_visco_top_strip_reads_full_grid_profile():
    nx, ny, fw = 40, 36, 6
    cv = cfg(nx=nx, ny=ny, fdorder=8, fw=fw, free_surf=0, L=2, dt=1e-4, dh=0.25, fl=[6.25, 50.0], k_max_pml=2.0)
    ce = cfg(nx=nx, ny=ny, fdorder=8, fw=fw, free_surf=0, L=0, dt=1e-4, dh=0.25, k_max_pml=2.0)
    rng = np.random.default_rng(11)
    vs = (150.0 + 50.0 * rng.random((ny, nx))).astype(np.float32)
    rho = np.full((ny, nx), 1900.0, np.float32)
    qs = np.full((ny, nx), np.inf, np.float32)
    o = O64
    Pv = o.model_setup(cv, vs, rho, qs)
    Pe = o.model_setup(ce, vs, rho)
    hc = o.holberg(8, 1)
    pml = o.pml(cv)
    sh = o.padded_shape(cv)
    nd = 5
    inner = (slice(nd, nd + ny), slice(nd, nd + nx))
    vy = np.zeros(sh)
    vy[inner] = rng.standard_normal((ny, nx))
    z = lambda: np.zeros(sh)

    def visco():
        sxy, syz = z(), z()
        r = np.zeros((2,) + sh)
        psi = [np.full(2 * fw * nx, 0.01), np.full(2 * fw * ny, 0.02)]
        o.update_s_visc(cv, vy, z(), z(), syz, sxy, r, r.copy(), Pv["fipjp"], Pv["f"], Pv["dip"], Pv["d"], Pv["bip"],
                        Pv["bjm"], Pv["cip"], Pv["cjm"], hc, pml, psi[0], psi[1])
        return sxy[inner], syz[inner]

    def elastic(prof):
        sxy, syz = z(), z()
        # the elastic kernel's psi carries the factor DT of its dhi = DT/DH
        dt = float(np.float32(1e-4))
        psi = [np.full(2 * fw * nx, 0.01 * dt), np.full(2 * fw * ny, 0.02 * dt)]
        o.update_s_elastic(ce, vy, z(), z(), syz, sxy, Pe["uip"], Pe["ujp"], hc, prof, psi[0], psi[1])
        return sxy[inner], syz[inner]

    v_sxy, v_syz = visco()
    e_sxy, e_syz = elastic(pml)
    # same sxy everywhere, same syz below the top strip
    assert rel_l2(v_sxy, e_sxy) < 1e-12
    assert rel_l2(v_syz[fw:], e_syz[fw:]) < 1e-12
    # differs on the top strip ...
    assert rel_l2(v_syz[:fw], e_syz[:fw]) > 1e-3
    # ... where it is the elastic update with the full-grid y profile there
    top = dict(pml)
    for q in ("K", "a", "b"):
        arr = pml["%s_y_half" % q].copy()
        arr[1:fw + 1] = pml["%s_y" % q][1:fw + 1]
        top["%s_y_half" % q] = arr
    t_sxy, t_syz = elastic(top)
    assert rel_l2(v_syz, t_syz) < 1e-12
    assert rel_l2(v_sxy, t_sxy) < 1e-12
