"""Pins of the CPU oracle (oracle/sh_oracle.c) by what the mathematics and the
reference's own third-party code fix, and parity of the product's host-side
setup with the oracle.  CPU only.

What pins what:
  holberg          Taylor rows = solution of the staggered-grid moment
                   equations; Holberg rows = max phase-velocity error E over
                   the band [0, 2 pi / ppw] (their defining property).
  pml_pro          INRIA d0 reflection identity, power-law profile shape,
                   CPML steady-state identity 1/K + a/(1-b) = alpha/(d+K alpha),
                   and the integer-abs quirk (a = 0 where 0 < d < 1).
  butterworth      bit-equal to libcseife built from the reference sources
                   (oracle/_ref) and equal to scipy's bilinear transform of
                   the analog Butterworth sections (no prewarp) + lfilter.
  wavelet          DENISE Ricker == textbook Ricker with peak frequency
                   fc*sqrt(2)/1.5 delayed by 1.5/fc + tshift.
  update_v/update_s  FP64 leapfrog energy conservation (skew-adjoint
                   operators with the zero frame) for FDORDER 2..10; FDORDER 12
                   breaks it by the hc[5] typo (quirk present); travel time of
                   a pulse in a homogeneous medium.
  update_s_visc    taus = 0 reduces to the elastic update; memory variables
                   decay by ((1-eta/2)/(1+eta/2))^n, the Crank-Nicolson
                   approximation of exp(-t/tau_sigma).
  residual/energy  time reversal identity and closed-form energy of a
                   constant trace.
  eprecond1/means  EPRECOND 0 / EPRECOND 1 ratio fields: (mean vs / mean rho)^2
                   between the parameters, the 0.005 max floor, invariance
                   under a joint scaling of signals and data (square root),
                   per-shot weights (two shots = sum of one-shot gradients).
  lbfgs/pcg/opt    secant condition and textbook two-loop agreement;
                   A-conjugacy of PR directions with exact line search;
                   vertex of an exact parabola; relative step rule.
  gauss_filt_var   scipy.ndimage.correlate (mode "nearest") with the
                   normalised closed-form Gaussian, per velocity (uniform and
                   two-layer models, hand-computed half widths); constant
                   gradient unchanged; WD_DAMP = 0 copies.
  taper_grad       closed forms of the vertical/horizontal windows and of the
                   erf bowl around a source in float64; zeroed cells.
  time_window      exp(-GAMMA (t - pick)^2) in float64 for t >= pick.
  stf (numpy)      Wiener deconvolution identity: observed = modelled * h
                   with EPS_STF = 0 returns signal * h; NORMALIZE helpers on
                   hand-computed traces.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

from common import ROOT, model, oracle, product, rel_l2

O32 = oracle("f32")
O64 = oracle("f64")


def cfg(**kw):
    from pyoracle import make_cfg
    return make_cfg(**kw)


# --------------------------------------------------------------------------
# holbergcoeff
# --------------------------------------------------------------------------
@pytest.mark.parametrize("fdorder", [2, 4, 6, 8, 10, 12])
def test_taylor_coefficients_solve_moment_equations(fdorder):
    k = fdorder // 2
    A = np.array([[(2 * m - 1) ** (2 * n - 1) for m in range(1, k + 1)] for n in range(1, k + 1)], float)
    b = np.zeros(k)
    b[0] = 1.0
    c = np.linalg.solve(A, b)
    hc = O64.holberg(fdorder, 0)
    np.testing.assert_allclose(hc[1:k + 1], c, rtol=2e-7)
    assert np.all(hc[k + 1:] == 0)


@pytest.mark.parametrize("maxrelerror,E", [(1, 0.001), (2, 0.005), (3, 0.01), (4, 0.03)])
@pytest.mark.parametrize("fdorder", [2, 4, 6, 8, 10, 12])
def test_holberg_rows_meet_phase_velocity_error(fdorder, maxrelerror, E):
    hc = O64.holberg(fdorder, maxrelerror)
    k = fdorder // 2
    kap = np.linspace(1e-6, 2 * np.pi / hc[0], 20000)
    m = np.arange(1, k + 1)[:, None]
    ph = np.sum(hc[1:k + 1, None] * 2 * np.sin((2 * m - 1) * kap / 2), axis=0) / kap
    err = np.max(np.abs(ph - 1))
    assert 0.9 * E <= err <= 1.06 * E


def test_product_holberg_equals_oracle():
    dn = product()
    for me in range(5):
        for fd in (2, 4, 6, 8, 10, 12):
            assert np.array_equal(dn.holbergcoeff(fd, me), O32.holberg(fd, me))
    with pytest.raises(Exception):
        dn.holbergcoeff(7, 1)


# --------------------------------------------------------------------------
# PML_pro
# --------------------------------------------------------------------------
PML_CFGS = [dict(nx=400, ny=120, fw=20, dh=0.25, dt=1e-4, damping=300.0, fpml=40.0),
            dict(nx=1200, ny=300, fw=20, dh=0.2, dt=1e-4, damping=350.0, fpml=25.0),
            dict(nx=64, ny=64, fw=10, dh=25.0, dt=1e-2, damping=1000.0, fpml=6.5)]


@pytest.mark.parametrize("pc", PML_CFGS)
def test_pml_profiles_physics(pc):
    c = cfg(nx=pc["nx"], ny=pc["ny"], fw=pc["fw"], dh=pc["dh"], dt=pc["dt"], damping=pc["damping"],
            fpml=pc["fpml"], fdorder=4)
    P = O64.pml(c)
    fw = pc["fw"]
    Lw = fw * np.float32(pc["dh"])
    N = 2.0
    # left full-grid: distance into the layer at point i is L - dh*(i-1)
    d = P["d_x"][1:fw + 1]
    x = Lw - pc["dh"] * np.arange(fw)
    d0 = d[0] / (x[0] / Lw) ** N
    np.testing.assert_allclose(d, d0 * (x / Lw) ** N, rtol=1e-5)
    # INRIA identity: theoretical reflection of the continuous profile is 0.001
    np.testing.assert_allclose(np.exp(-2.0 * d0 * Lw / ((N + 1) * pc["damping"])), 0.001, rtol=1e-5)
    for side in ("x", "x_half", "y", "y_half"):
        d, K, al, a, b = (P["%s_%s" % (q, side)][1:2 * fw + 1] for q in ("d", "K", "alpha", "a", "b"))
        set_ = a != 0
        # steady-state response of the recursive convolution equals the CPML stretch at omega = 0
        lhs = 1.0 / K[set_] + a[set_] / (1.0 - b[set_])
        rhs = al[set_] / (d[set_] + K[set_] * al[set_])
        np.testing.assert_allclose(lhs, rhs, rtol=1e-8)
        # Q1: integer abs() -> no a where 0 < d < 1
        small = (d > 0) & (d < 1)
        assert np.all(a[small] == 0)
        assert np.all(a[d >= 1] != 0)
        assert np.all((b > 0) & (b <= 1))
    # the quirk is exercised by these profiles (innermost half-grid points)
    assert any(np.any((P[k] > 0) & (P[k] < 1)) for k in ("d_x", "d_x_half", "d_y", "d_y_half"))


@pytest.mark.parametrize("pc", PML_CFGS)
def test_product_pml_equals_oracle(pc):
    dn = product()
    got = dn.pml_pro(pc["nx"], pc["ny"], pc["fw"], pc["dh"], pc["dt"], pc["damping"], pc["fpml"])
    c = cfg(nx=pc["nx"], ny=pc["ny"], fw=pc["fw"], dh=pc["dh"], dt=pc["dt"], damping=pc["damping"],
            fpml=pc["fpml"], fdorder=4)
    ref = O32.pml(c)
    for k, v in got.items():
        assert np.array_equal(v[1:], ref[k][1:2 * pc["fw"] + 1]), k


# --------------------------------------------------------------------------
# Butterworth / timedomain_filt
# --------------------------------------------------------------------------
def _cseife_ref():
    p = os.path.join(ROOT, "oracle", "_ref", "libcseife_ref.so")
    if not os.path.exists(p):
        pytest.skip("oracle/_ref/libcseife_ref.so not built (reference tree absent)")
    L = C.CDLL(p)
    for f in ("seife_lpb", "seife_hpb"):
        getattr(L, f).argtypes = [C.POINTER(C.c_double), C.c_int, C.c_double, C.c_double, C.c_int]
    return L


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("method", [1, 2])
def test_butterworth_equals_libcseife(order, method):
    L = _cseife_ref()
    rng = np.random.default_rng(order * 10 + method)
    x = rng.standard_normal(3000)
    y_ref = x.copy()
    f = L.seife_lpb if method == 1 else L.seife_hpb
    f(y_ref.ctypes.data_as(C.POINTER(C.c_double)), len(y_ref), 1e-4, 1.0 / 30.0, order)
    y = O64.butterworth(x, 1e-4, 1.0 / 30.0, order, method)
    assert np.array_equal(y, y_ref)


@pytest.mark.parametrize("order", [2, 4, 5])
def test_butterworth_is_bilinear_analog_butterworth(order):
    from scipy import signal
    dt, fc = 1e-4, 60.0
    w0 = 2 * np.pi * fc
    rng = np.random.default_rng(order)
    x = rng.standard_normal(2000)
    y = x.copy()
    m = order // 2
    if order % 2:
        b, a = signal.bilinear([w0], [1.0, w0], fs=1.0 / dt)
        y = signal.lfilter(b, a, y)
    for j in range(1, m + 1):
        h = np.sin(np.pi / 2 * (2 * j - 1) / order)
        b, a = signal.bilinear([w0 ** 2], [1.0, 2 * h * w0, w0 ** 2], fs=1.0 / dt)
        y = signal.lfilter(b, a, y)
    got = O64.butterworth(x, dt, 1.0 / fc, order, 1)
    np.testing.assert_allclose(got, y, rtol=1e-9, atol=1e-12 * np.max(np.abs(y)))


@pytest.mark.parametrize("method,order", [(1, 4), (2, 4), (1, 3)])
def test_timedomain_filt_offset_quirk_and_product(method, order):
    rng = np.random.default_rng(7)
    data = rng.standard_normal((5, 600)).astype(np.float32)
    got = O32.timedomain_filt(data, 1e-4, 40.0, order, method)
    # Q8: the filter sees [0, s1 .. s_{ns-1}], the last sample is untouched
    for t in range(5):
        buf = np.concatenate([[0.0], data[t, :-1].astype(np.float64)])
        y = O64.butterworth(buf, float(np.float32(1e-4)), 1.0 / float(np.float32(40.0)), order, method)
        np.testing.assert_array_equal(got[t, :-1], y[1:].astype(np.float32))
        assert got[t, -1] == data[t, -1]
    prod = product().timedomain_filt(data, 1e-4, 40.0, order, method)
    assert np.array_equal(prod, got)


# --------------------------------------------------------------------------
# wavelet
# --------------------------------------------------------------------------
def test_denise_ricker_is_textbook_ricker():
    dt, fc, ts = 1e-4, 40.0, 0.002
    nt = 1000
    w = O64.wavelet(1, nt, dt, ts, fc, 1.0)
    t = np.arange(1, nt + 1) * np.float32(dt)
    fp = fc * np.sqrt(2.0) / 1.5
    tt = t - 1.5 / fc - ts
    ref = (1 - 2 * (np.pi * fp * tt) ** 2) * np.exp(-(np.pi * fp * tt) ** 2)
    np.testing.assert_allclose(w, ref, atol=2e-6)
    assert abs(t[np.argmax(w)] - (1.5 / fc + ts)) <= dt


def test_spike_and_product_wavelets():
    dn = product()
    s = O32.wavelet(6, 100, 1e-3, 0.0105, 10.0, 2.0)
    assert np.count_nonzero(s) == 1 and s[int(np.floor(0.0105 / np.float32(1e-3) + 0.5))] == 2.0
    for qa in (1, 2, 4, 5, 6):
        for fc, ts in ((25.0, 0.0), (40.0, 0.003)):
            a = dn.wavelet(qa, 700, 1e-4, ts, fc, 1.5)
            b = O32.wavelet(qa, 700, 1e-4, ts, fc, 1.5)
            assert np.array_equal(a, b), qa


# --------------------------------------------------------------------------
# stencils
# --------------------------------------------------------------------------
def _random_state(c, o, seed):
    nd = c.fdorder // 2 + 1
    sh = o.padded_shape(c)
    inner = (slice(nd, nd + c.ny), slice(nd, nd + c.nx))
    rng = np.random.default_rng(seed)
    F = {}
    for k, s in (("vy", 1.0), ("sxy", 1e3), ("syz", 1e3)):
        a = np.zeros(sh, o.dt)
        a[inner] = rng.standard_normal((c.ny, c.nx)) * s
        F[k] = a
    return F, inner


@pytest.mark.parametrize("fdorder", [2, 4, 6, 8, 10, 12])
def test_leapfrog_energy_conservation(fdorder):
    nx, ny = 40, 30
    c = cfg(nx=nx, ny=ny, fdorder=fdorder, fw=0, free_surf=1, L=0, dt=1e-4, dh=0.25)
    vs, rho, _ = model(nx, ny, 0.25, seed=2)
    P = O64.model_setup(c, vs, rho)
    hc = O64.holberg(fdorder, 1)
    pml = O64.pml(c)
    F, I = _random_state(c, O64, 4)
    psi = np.zeros(8)
    vyp1 = np.zeros_like(F["vy"])
    uy = np.zeros_like(F["vy"])
    uyx = np.zeros_like(F["vy"])
    rhoI, uip, ujp = P["rho"][I], P["uip"][I], P["ujp"][I]
    E = []
    es_prev = None
    for n in range(150):
        vold = F["vy"][I].copy()
        O64.update_v(c, 0, F["vy"], vyp1, F["syz"], F["sxy"], P["rho"], hc, pml, psi, psi)
        if es_prev is not None:
            E.append(0.5 * np.sum(rhoI * vold * F["vy"][I]) + es_prev)
        O64.update_s_elastic(c, F["vy"], uy, uyx, F["syz"], F["sxy"], P["uip"], P["ujp"], hc, pml, psi, psi)
        es_prev = 0.5 * np.sum(F["sxy"][I] ** 2 / uip + F["syz"][I] ** 2 / ujp)
    E = np.array(E)
    drift = np.max(np.abs(E - E[0])) / E[0]
    if fdorder < 12:
        assert drift < 1e-12
    else:
        assert drift > 1e-6  # Q6: hc[5] used twice in update_v_PML.c:469 breaks adjointness


def test_travel_time_homogeneous():
    """Peak of a Ricker pulse arrives after distance/vs (+-2%) in a homogeneous medium."""
    nx, ny, dh, vs0 = 220, 220, 0.5, 200.0
    c = cfg(nx=nx, ny=ny, fdorder=8, fw=20, free_surf=0, L=0, nt=2200, dt=1e-4, dh=dh, damping=200.0, fpml=30.0)
    vs = np.full((ny, nx), vs0, np.float32)
    rho = np.full((ny, nx), 2000.0, np.float32)
    fc = 30.0
    sig = O32.wavelet(1, c.nt, 1e-4, 0.0, fc, 1.0)
    ri = np.array([110 + 20, 110 + 50], np.int32)
    rj = np.array([110, 110], np.int32)
    seis, _, _ = O64.forward(c, vs, rho, None, np.array([110], np.int32), np.array([110], np.int32), sig, ri, rj)
    t = np.arange(1, c.nt + 1) * 1e-4
    # far-field 2-D line source: peak delayed by r/c (plus a fixed phase lag common to both receivers)
    tp = [t[np.argmax(np.abs(seis[k]))] for k in range(2)]
    dt_meas = tp[1] - tp[0]
    assert abs(dt_meas - 30 * dh / vs0) <= 0.02 * 30 * dh / vs0


def test_visco_with_zero_taus_equals_elastic():
    nx, ny = 40, 30
    cv = cfg(nx=nx, ny=ny, fdorder=8, fw=6, free_surf=1, L=2, dt=1e-4, dh=0.25, fl=[6.25, 50.0])
    ce = cfg(nx=nx, ny=ny, fdorder=8, fw=6, free_surf=1, L=0, dt=1e-4, dh=0.25)
    vs, rho, _ = model(nx, ny, 0.25, seed=2)
    qs = np.full((ny, nx), np.inf, np.float32)
    for o, tol in ((O64, 1e-12), (O32, 2e-6)):
        Pv = o.model_setup(cv, vs, rho, qs)
        Pe = o.model_setup(ce, vs, rho)
        hc = o.holberg(8, 1)
        pml = o.pml(cv)
        Fv, I = _random_state(cv, o, 9)
        Fe = {k: v.copy() for k, v in Fv.items()}
        z = lambda: np.zeros_like(Fv["vy"])
        r = np.zeros((2,) + Fv["vy"].shape, o.dt)
        q = np.zeros_like(r)
        pa, pb = np.zeros(12 * nx, o.dt), np.zeros(12 * ny, o.dt)
        pa2, pb2 = pa.copy(), pb.copy()
        o.update_s_visc(cv, Fv["vy"], z(), z(), Fv["syz"], Fv["sxy"], r, q, Pv["fipjp"], Pv["f"], Pv["dip"], Pv["d"],
                        Pv["bip"], Pv["bjm"], Pv["cip"], Pv["cjm"], hc, pml, pa, pb)
        o.update_s_elastic(ce, Fe["vy"], z(), z(), Fe["syz"], Fe["sxy"], Pe["uip"], Pe["ujp"], hc, pml, pa2, pb2)
        # visco uses dhi = 1/DH with fipjp = mu*DT; elastic dhi = DT/DH with mu
        assert rel_l2(Fv["sxy"][I], Fe["sxy"][I]) < tol
        assert rel_l2(Fv["syz"][I], Fe["syz"][I]) < tol
        assert np.all(r == 0) and np.all(q == 0)


def test_memory_variable_decay():
    nx, ny = 20, 16
    fl = [6.25, 50.0]
    c = cfg(nx=nx, ny=ny, fdorder=4, fw=0, free_surf=1, L=2, dt=1e-4, dh=0.25, fl=fl)
    vs, rho, _ = model(nx, ny, 0.25, seed=2)
    P = O64.model_setup(c, vs, rho, np.full((ny, nx), 20.0, np.float32))
    hc = O64.holberg(4, 1)
    pml = O64.pml(c)
    sh = O64.padded_shape(c)
    r = np.ones((2,) + sh)
    q = np.ones((2,) + sh) * 2.0
    z = lambda: np.zeros(sh)
    psi = np.zeros(4)
    vy, sxy, syz = z(), z(), z()
    n = 50
    for _ in range(n):
        O64.update_s_visc(c, vy, z(), z(), syz, sxy, r, q, P["fipjp"], P["f"], P["dip"], P["d"], P["bip"], P["bjm"],
                          P["cip"], P["cjm"], hc, pml, psi, psi)
    for l, f in enumerate(fl):
        eta = np.float64(np.float32(1e-4)) * 2 * np.pi * f
        g = (1 - eta / 2) / (1 + eta / 2)
        np.testing.assert_allclose(r[l, 5, 5], g ** n, rtol=1e-10)
        np.testing.assert_allclose(q[l, 5, 5], 2 * g ** n, rtol=1e-10)
        # Crank-Nicolson: per-step factor matches exp(-eta) to O(eta^3)
        assert abs(g - np.exp(-eta)) <= eta ** 3 / 12 * 1.01


# --------------------------------------------------------------------------
# residual / energy
# --------------------------------------------------------------------------
def test_residual_reversal_and_energy_closed_form():
    ntr, ns, dt = 3, 50, 1e-3
    rng = np.random.default_rng(3)
    syn = rng.standard_normal((ntr, ns))
    obs = np.ones((ntr, ns))
    diff = np.zeros((ntr, ns))
    l2 = C.c_double(0.0)
    en = C.c_double(0.0)
    f = O64.fn("orc_residual")
    rp = C.POINTER(C.c_double)
    obs_c = obs.copy()
    f(ntr, ns, C.c_double(dt), syn.ctypes.data_as(rp), obs_c.ctypes.data_as(rp), diff.ctypes.data_as(rp),
      C.byref(l2), C.byref(en))
    # Q7: the first observed sample is replaced by the synthetic one
    exp = syn - obs
    exp[:, 0] = 0.0
    np.testing.assert_array_equal(diff, exp[:, ::-1])
    np.testing.assert_allclose(l2.value, np.sum(exp ** 2), rtol=1e-14)
    # energy of the time-integrated observed trace: obs' = [syn0, 1, 1, ...]
    cum = np.cumsum(np.concatenate([syn[:, :1], np.ones((ntr, ns - 1))], axis=1), axis=1)
    np.testing.assert_allclose(en.value, np.sum(cum ** 2) * dt * dt, rtol=1e-12)


# --------------------------------------------------------------------------
# optimiser pieces
# --------------------------------------------------------------------------
def test_calc_opt_step_parabola_vertex():
    a, b, cc = 2.0, -3.0, 1.5  # vertex at 1.0
    eps = np.array([0.0, 0.8, 1.6])
    L2t = a + b * eps + cc * eps ** 2
    assert abs(O64.calc_opt_step(L2t, eps) - 1.0) < 1e-12
    # vertex beyond the bracket is clamped to epst[3]
    eps2 = np.array([0.0, 0.2, 0.4])
    assert O64.calc_opt_step(a + b * eps2 + cc * eps2 ** 2, eps2) == 0.4


def _textbook_two_loop(S, Y, g):
    q = g.copy()
    al = []
    for s, y in zip(S[::-1], Y[::-1]):
        a = (s @ q) / (y @ s)
        al.append(a)
        q -= a * y
    r = (S[-1] @ Y[-1]) / (Y[-1] @ Y[-1]) * q
    for (s, y), a in zip(zip(S, Y), al[::-1]):
        b = (y @ r) / (y @ s)
        r += s * (a - b)
    return r


def test_lbfgs_two_loop():
    rng = np.random.default_rng(0)
    n, m = 30, 5
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)
    S = [rng.standard_normal(n) for _ in range(m)]
    Y = [A @ s for s in S]
    nslots = 8  # slots m+1..nslots empty (rho = 0), as before the ring wraps
    ys = np.zeros((nslots, n))
    ss = np.zeros((nslots, n))
    ys[:m] = Y
    ss[:m] = S
    g = rng.standard_normal(n)
    r = O64.lbfgs(ys, ss, g, cur=m)
    np.testing.assert_allclose(r, _textbook_two_loop(S, Y, g), rtol=1e-10)
    # secant condition of the newest pair
    np.testing.assert_allclose(O64.lbfgs(ys, ss, Y[-1], cur=m), S[-1], rtol=1e-9)


def test_pcg_polak_ribiere_conjugacy():
    rng = np.random.default_rng(1)
    n = 12
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)
    bvec = rng.standard_normal(n)
    x = np.zeros(n)
    gold = np.zeros((1, n))
    d = np.zeros((1, n))
    dirs = []
    for _ in range(4):
        g = A @ x - bvec
        gnew = -g[None, :]          # descent direction as stored by the reference
        d = O64.pcg(gnew, gold, d)
        p = d[0]
        alpha = -(g @ p) / (p @ A @ p)  # exact line search
        x = x + alpha * p
        dirs.append(p.copy())
        gold = gnew
    for i in range(len(dirs)):
        for j in range(i):
            assert abs(dirs[i] @ A @ dirs[j]) < 1e-8 * np.linalg.norm(A @ dirs[i]) * np.linalg.norm(dirs[j])


def test_calc_mat_change_rule():
    f = O64.fn("orc_calc_mat_change")
    f.restype = C.c_double
    rp = C.POINTER(C.c_double)
    nx, ny = 4, 3
    u = np.full((ny, nx), 200.0)
    rho = np.full((ny, nx), 1800.0)
    gu = np.zeros((ny, nx))
    grho = np.zeros((ny, nx))
    gu[1, 2] = -4.0
    gu[0, 0] = 2.0
    grho[2, 3] = 10.0
    un = np.zeros_like(u)
    rn = np.zeros_like(u)
    er = C.c_double()
    eps_u = f(nx, ny, gu.ctypes.data_as(rp), grho.ctypes.data_as(rp), u.ctypes.data_as(rp), rho.ctypes.data_as(rp),
              un.ctypes.data_as(rp), rn.ctypes.data_as(rp), C.c_double(0.01), 5, 0, 0, C.c_double(0.5),
              C.c_double(0.0), C.c_double(205.0), C.c_double(0.0), C.c_double(5000.0), C.byref(er))
    assert abs(eps_u - 0.01 * 200.0 / 4.0) < 1e-12
    assert abs(er.value - 0.01 * 1800.0 / 10.0 * 0.5) < 1e-12
    # +2.0 cell: 200 + 0.5*4 = 202 within the limit; -4 cell: 200 + 2 = 202 -> kept; check revert above 205
    assert un[1, 2] == 202.0 and un[0, 0] == 199.0
    f(nx, ny, gu.ctypes.data_as(rp), grho.ctypes.data_as(rp), u.ctypes.data_as(rp), rho.ctypes.data_as(rp),
      un.ctypes.data_as(rp), rn.ctypes.data_as(rp), C.c_double(0.02), 5, 0, 0, C.c_double(0.5), C.c_double(0.0),
      C.c_double(203.0), C.c_double(0.0), C.c_double(5000.0), C.byref(er))
    assert un[1, 2] == 200.0  # 204 > VSUPPERLIM 203: reverted, not clamped (Q17)


def test_dtstab_homogeneous_closed_form():
    c = cfg(nx=30, ny=20, fdorder=8, maxrelerror=1, fw=5, dt=1e-4, dh=0.25)
    vs = np.full((20, 30), 300.0, np.float32)
    rho = np.full((20, 30), 1800.0, np.float32)
    hc = O64.holberg(8, 1)
    ref = 0.25 / (np.sqrt(2) * np.sum(np.abs(hc[1:])) * 300.0)
    np.testing.assert_allclose(O64.dtstab(c, vs, rho, None, 0.025), ref, rtol=1e-12)


# --------------------------------------------------------------------------
# product host pieces that need no device
# --------------------------------------------------------------------------
def test_relaxation_constants_equal_oracle():
    dn = product()
    fl = [6.25, 50.0]
    c = cfg(nx=10, ny=8, fdorder=4, fw=2, L=2, dt=1e-4, dh=0.2, fl=fl)
    vs, rho, _ = model(10, 8, 0.2, seed=1)
    P = O32.model_setup(c, vs, rho, np.full((8, 10), 20.0, np.float32))
    got = dn.relaxation_constants(fl, 1e-4)
    for k in ("bip", "bjm", "cip", "cjm"):
        assert np.array_equal(got[k], P[k][:2]), k


def test_library_exports_every_declared_symbol():
    dn = product()
    hdr = open(os.path.join(ROOT, "include", "denise_sh.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b([A-Za-z_]\w*)\s*\(", hdr)) - {"if", "sizeof"}
    assert names, "no declarations parsed"
    L = C.CDLL(dn.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(dn.EXPORTS) == names


def test_binding_declares_the_argument_types_of_every_call():
    """Every C-ABI function the Python binding calls has its argument types
    declared somewhere in the binding: without them ctypes passes a context
    handle as a C int and cuts the pointer to 32 bits."""
    src = open(os.path.join(ROOT, "denise-sh_amd", "__init__.py")).read()
    called = set(re.findall(r"(?:lib\(\)|\bL)\.(dn_\w+)\(", src))
    declared = set(re.findall(r"\.(dn_\w+)\.argtypes\s*=", src))
    assert called, "no calls parsed"
    assert not sorted(called - declared), sorted(called - declared)


def test_seismo_ssg_and_psource_host_semantics():
    dn = product()
    nx, ny, nt = 8, 6, 10
    dn.set_globals(NX=nx, NY=ny, FDORDER=4, FW=2, FREE_SURF=1, L=0, DT=1e-3, DH=1.0, NT=nt)
    vy = dn.NRMatrix(-2, ny + 3, -2, nx + 3, np.arange((ny + 6) * (nx + 6), dtype=np.float32).reshape(ny + 6, nx + 6))
    rec = dn.NRIMatrix(1, 3, 1, 2, np.array([[2, 5], [3, 1], [1, 2]], np.int32))
    sec = dn.NRMatrix(1, 2, 1, nt)
    L = dn.lib()
    L.seismo_ssg.restype = None
    L.seismo_ssg(4, 2, rec.ptr, sec.ptr, vy.ptr, None, None, None)
    assert sec[1, 4] == vy[3, 2] and sec[2, 4] == vy[1, 5]
    sxx = dn.NRMatrix(-2, ny + 3, -2, nx + 3)
    syy = dn.NRMatrix(-2, ny + 3, -2, nx + 3)
    src = dn.NRMatrix(1, 8, 1, 1, np.array([[3], [2], [0], [0], [10], [1], [0], [1]], np.float32))
    sig = dn.NRMatrix(1, 1, 1, nt, np.arange(1, nt + 1, dtype=np.float32)[None, :] ** 2)
    L.psource.restype = None
    L.psource(5, sxx.ptr, syy.ptr, src.ptr, sig.ptr, 1)
    amp = np.float32((np.float32(36.0) - np.float32(16.0)) / np.float32(1e-3))
    assert sxx[2, 3] == amp and syy[2, 3] == amp


@pytest.mark.parametrize("L", [0, 2])
def test_checkfd_equals_oracle_and_flags(L):
    dn = product()
    fl = [6.25, 50.0][:L]
    c = cfg(nx=40, ny=30, fdorder=8, maxrelerror=1, fw=10, dt=1e-4, dh=0.2, L=L, fl=fl)
    vs, rho, _ = model(40, 30, 0.2, seed=3)
    qs = np.full((30, 40), 20.0, np.float32) if L else None
    rc, r = dn.checkfd(vs, rho, qs, 8, 1, fl, 1e-4, 0.2, 0.04, 10)
    assert rc == 0 and not r["dt_unstable"] and not r["fw_too_large"]
    assert np.float32(r["dtstab"]) == np.float32(O32.dtstab(c, vs, rho, qs, 0.04))
    if L == 0:
        assert r["cmax"] == vs.max() and r["cmin"] == vs.min()
    else:
        # homogeneous closed form of the relaxed/unrelaxed velocities (GSLS, tau = 2/Q)
        rc, h = dn.checkfd(np.full((4, 5), 300.0), np.full((4, 5), 1800.0), np.full((4, 5), 20.0), 8, 1, fl,
                           1e-4, 0.2, 0.04, 1)
        ws = 2 * np.pi / 0.04
        tl = 1 / (2 * np.pi * np.array(fl))
        sumu = np.sum(ws ** 2 * tl ** 2 * 0.1 / (1 + ws ** 2 * tl ** 2))
        np.testing.assert_allclose(h["cmin"], 300 / np.sqrt(1 + sumu), rtol=1e-5)
        np.testing.assert_allclose(h["cmax"], 300 * np.sqrt((1 + L * 0.1) / (1 + sumu)), rtol=1e-5)
        assert h["fw_small"] == 1
    # DT above the limit and a frame wider than the grid abort (err() in the reference)
    rc, r2 = dn.checkfd(vs, rho, qs, 8, 1, fl, 1.01 * r["dtstab"], 0.2, 0.04, 10)
    assert rc == 1 and r2["dt_unstable"]
    rc, r3 = dn.checkfd(vs, rho, qs, 8, 1, fl, 1e-4, 0.2, 0.04, 31)
    assert rc == 1 and r3["fw_too_large"]
    hc = dn.holbergcoeff(8, 1)
    np.testing.assert_allclose(r["dhstab"], r["cmin"] / (hc[0] * 2 / 0.04), rtol=1e-6)


# --------------------------------------------------------------------------
# adjoint gradient: directional derivative of the misfit
# --------------------------------------------------------------------------
def _misfit(o, c, vs, rho, qs, si, sj, sig, ri, rj, obs):
    seis, _, _ = o.forward(c, vs, rho, qs, si, sj, sig, ri, rj)
    d = seis[:, 1:] - obs[:, 1:].astype(seis.dtype)   # calc_res.c: obs[1] := syn[1]
    return float(np.sum(d.astype(np.float64) ** 2))


@pytest.mark.parametrize("L", [0, 2])
def test_gradient_is_misfit_derivative(L):
    """Directional derivatives of the misfit sum_(t>1) (syn - obs)^2 by
    central differences against the oracle's gradient with EPRECOND = 0, on
    smooth bumps in a homogeneous medium.  What the adjoint-state method
    fixes for this velocity-stress scheme with raw residual injection:

      density at fixed mu:   -DT * sum_t vyp1_adj * vyp1_fwd = c * dJ/drho
      shear modulus:         -DT * sum_t (uyx sxy + uy syz)  = c * DTF * mu * dJ/dmu
      vs (chain rule 2 rho vs): the same ratio as mu, since d/d ln vs = 2 d/d ln mu

    with one positive constant c for both parameters and DTF = DT for the
    elastic update (its uy, uyx carry the DT factor) and 1 for the
    viscoelastic one.  A wrong sign, a dropped correlation term, a
    mis-paired snapshot, a transposed field or a wrong chain-rule factor
    breaks one of the ratios below."""
    nx, ny, nt, dh, dt = 70, 44, 700, 0.25, 1e-4
    fl = [6.25, 50.0][:L]
    c = cfg(nx=nx, ny=ny, fdorder=4, maxrelerror=1, fw=10, free_surf=1, L=L, nt=nt, dtinv=1, dt=dt, dh=dh,
            damping=300.0, fpml=30.0, fl=fl, eprecond=0)
    vs = np.full((ny, nx), 200.0, np.float32)
    rho = np.full((ny, nx), 1800.0, np.float32)
    qs = np.full((ny, nx), 20.0, np.float32) if L else None
    t = np.arange(1, nt + 1) * dt
    tau = np.pi * (t - 1.5 / 30.0) / (1.5 / 30.0)
    sig = ((1 - 4 * tau ** 2) * np.exp(-2 * tau ** 2)).astype(np.float32)[None, :]
    si, sj = np.array([35], np.int32), np.array([3], np.int32)
    ri = np.arange(14, 57, 3, dtype=np.int32)
    rj = np.full_like(ri, 2)
    vt = vs.copy()
    vt[14:24, 25:45] *= np.float32(1.06)
    obs, _, _ = O64.forward(c, vt, rho, qs, si, sj, sig[0], ri, rj)
    obs = obs.astype(np.float32)
    g = O64.gradient(c, vs, rho, qs, si, sj, sig, ri, rj, obs[None])
    J = lambda v, r: _misfit(O64, c, v, r, qs, si, sj, sig[0], ri, rj, obs)
    jj, ii = np.mgrid[1:ny + 1, 1:nx + 1]
    mu = 1800.0 * 200.0 ** 2
    dtf = dt if L == 0 else 1.0
    e = 2e-3
    rows = []
    for cj, ci in [(10, 24), (16, 25), (24, 44), (8, 50), (30, 35)]:
        d = np.exp(-((jj - cj) ** 2 + (ii - ci) ** 2) / (2 * 5.0 ** 2))
        fd_mu = (J(vs * np.sqrt(1 + e * d), rho) - J(vs * np.sqrt(1 - e * d), rho)) / (2 * e)       # d/d ln mu
        fd_rho = (J(vs / np.sqrt(1 + e * d), rho * (1 + e * d)) -
                  J(vs / np.sqrt(1 - e * d), rho * (1 - e * d))) / (2 * e)                          # d/d ln rho | mu
        fd_vs = (J(vs * (1 + e * d), rho) - J(vs * (1 - e * d), rho)) / (2 * e)                     # d/d ln vs | rho
        k_mu = float(np.sum(-dt * g["raw_u"] * mu * d))
        k_rho = float(np.sum(-dt * g["raw_rho"] * 1800.0 * d))
        k_vs = float(np.sum(g["grad_vs"] * vs * d))
        rows.append((k_rho / fd_rho, k_mu / fd_mu / (dtf * mu), k_vs / fd_vs / (dtf * mu)))
    r = np.array(rows)
    print("ratios [rho|mu, mu/(DTF mu), vs/(DTF mu)]:\n", r)
    cst = np.median(r[:, 0])
    assert cst > 0
    assert np.all(np.abs(r[:, 0] / cst - 1) < 0.08)
    assert np.all((r[:, 1] / cst > 0.75) & (r[:, 1] / cst < 1.12))
    np.testing.assert_allclose(r[:, 2], r[:, 1], rtol=0.02)


def test_eprecond1_scaling_and_normalisation_means():
    """eprecond1.c:15-70 and denise.c:746-819, 1671-1694 by what they fix,
    from the oracle's EPRECOND 0 and 1 gradients of the same shots:

      per shot  g1 = g0 / (We * C^2),  We = sqrt(Ws Wr) + 0.005 max sqrt(Ws Wr),
                C_vs, C_rho = the means of the vs and rho models;

    so for one shot g0_u / g1_u = We C_vs^2 and g0_rho / g1_rho = We C_rho^2:
    (a) their ratio is (mean vs / mean rho)^2 in every cell (one We for both
        parameters, the means of the right models, not swapped);
    (b) We = g0_u / (g1_u C_vs^2) is nowhere below the floor
        0.005 / 1.005 * max We (the regularisation is relative to the maximum);
    (c) source signals and observed data scaled together by a = 2 (exact in
        floating point) scale every field energy and the correlation by a^2:
        the EPRECOND 1 gradient stays the same only if We is of degree one in
        Ws Wr (the square root);
    (d) the two-shot gradient is the sum of the one-shot gradients (We is
        taken per shot, not from energies summed over shots)."""
    nx, ny, nt, dh, dt = 60, 40, 500, 0.25, 1e-4
    kw = dict(nx=nx, ny=ny, fdorder=4, maxrelerror=1, fw=10, free_surf=1, L=0, nt=nt, dtinv=2, dt=dt, dh=dh,
              damping=300.0, fpml=30.0)
    c0, c1 = cfg(eprecond=0, **kw), cfg(eprecond=1, **kw)
    rng = np.random.default_rng(7)
    jj, ii = np.mgrid[0:ny, 0:nx]
    vs = (180.0 + 60.0 * jj / ny + 8.0 * rng.standard_normal((ny, nx))).astype(np.float32)
    rho = (1500.0 + 400.0 * ii / nx + 30.0 * rng.standard_normal((ny, nx))).astype(np.float32)
    t = np.arange(1, nt + 1) * dt
    tau = np.pi * (t - 1.5 / 30.0) / (1.5 / 30.0)
    w = ((1 - 4 * tau ** 2) * np.exp(-2 * tau ** 2)).astype(np.float32)
    si, sj = np.array([20, 41], np.int32), np.array([3, 3], np.int32)
    ri = np.arange(12, 50, 3, dtype=np.int32)
    rj = np.full_like(ri, 2)
    vt = vs.copy()
    vt[12:22, 22:40] *= np.float32(1.07)
    obs = np.stack([O64.forward(c0, vt, rho, None, si[k:k + 1], sj[k:k + 1], w, ri, rj)[0] for k in range(2)])
    obs = obs.astype(np.float32)
    sig = np.stack([w, w])

    def grad(c, k, a=1.0):
        s = slice(k, k + 1) if k is not None else slice(0, 2)
        return O64.gradient(c, vs, rho, None, si[s], sj[s], (a * sig[s]).astype(np.float32), ri, rj,
                            (a * obs[s]).astype(np.float32))

    g0, g1 = grad(c0, 0), grad(c1, 0)
    cvs, crho = float(np.mean(vs.astype(np.float64))), float(np.mean(rho.astype(np.float64)))
    m = (np.abs(g0["raw_u"]) > 1e-6 * np.abs(g0["raw_u"]).max()) & \
        (np.abs(g0["raw_rho"]) > 1e-6 * np.abs(g0["raw_rho"]).max())
    ru = g0["raw_u"][m] / g1["raw_u"][m]
    rr = g0["raw_rho"][m] / g1["raw_rho"][m]
    print("cells %d, C_vs %.6g, C_rho %.6g" % (m.sum(), cvs, crho))
    np.testing.assert_allclose(ru / rr, (cvs / crho) ** 2, rtol=1e-9)                    # (a)
    We = ru / cvs ** 2
    floor = 0.005 / 1.005 * We.max()
    print("min We / floor = %.6f" % (We.min() / floor))
    assert We.min() >= floor * (1 - 1e-9)                                                 # (b)
    g1a = grad(c1, 0, a=2.0)
    g0a = grad(c0, 0, a=2.0)
    assert rel_l2(g0a["raw_u"], 4.0 * g0["raw_u"]) < 1e-13
    for key in ("raw_u", "raw_rho", "grad_vs", "grad_rho"):                               # (c)
        assert rel_l2(g1a[key], g1[key]) < 1e-13, key
    g12 = grad(c1, None)                                                                  # (d)
    g1b = grad(c1, 1)
    for key in ("raw_u", "raw_rho", "grad_vs", "grad_rho"):
        assert rel_l2(g12[key], g1[key] + g1b[key]) < 1e-12, key


def test_eprecond3_receiver_approximation():
    """EPRECOND 3 (eprecond1.c:34-50): per shot g3 = g0 / (We3 C^2) with
    We3 = sqrt(Ws) (asinh((xmax - x)/y) - asinh((xmin - x)/y)) + 0.005 max,
    the Plessix-Mulder (2004) approximation of the receiver-side energy for
    receivers along the whole surface (xmin = DH, xmax = NX DH, x = i DH,
    y = j DH).  Ws from the oracle's own forward, the rest in numpy: the
    ratio of the EPRECOND 0 and 3 gradients must be We3 C^2 in every cell;
    receiver energy Wr plays no part (a different receiver line leaves g3 /
    g0 unchanged)."""
    nx, ny, nt, dh, dt = 60, 40, 500, 0.25, 1e-4
    kw = dict(nx=nx, ny=ny, fdorder=4, maxrelerror=1, fw=10, free_surf=1, L=0, nt=nt, dtinv=2, dt=dt, dh=dh,
              damping=300.0, fpml=30.0)
    c0, c3 = cfg(eprecond=0, **kw), cfg(eprecond=3, **kw)
    rng = np.random.default_rng(11)
    jj, ii = np.mgrid[0:ny, 0:nx]
    vs = (180.0 + 60.0 * jj / ny + 8.0 * rng.standard_normal((ny, nx))).astype(np.float32)
    rho = (1500.0 + 400.0 * ii / nx + 30.0 * rng.standard_normal((ny, nx))).astype(np.float32)
    t = np.arange(1, nt + 1) * dt
    tau = np.pi * (t - 1.5 / 30.0) / (1.5 / 30.0)
    w = ((1 - 4 * tau ** 2) * np.exp(-2 * tau ** 2)).astype(np.float32)
    si, sj = np.array([25], np.int32), np.array([3], np.int32)
    vt = vs.copy()
    vt[12:22, 22:40] *= np.float32(1.07)
    Ws = None
    ratios = []
    for ri in (np.arange(12, 50, 3, dtype=np.int32), np.arange(5, 30, 2, dtype=np.int32)):
        rj = np.full_like(ri, 2)
        obs = O64.forward(c0, vt, rho, None, si, sj, w, ri, rj)[0][None].astype(np.float32)
        g0 = O64.gradient(c0, vs, rho, None, si, sj, w[None], ri, rj, obs)
        g3 = O64.gradient(c3, vs, rho, None, si, sj, w[None], ri, rj, obs)
        if Ws is None:
            nd = 3
            Ws = O64.forward(c0, vs, rho, None, si, sj, w, ri, rj, want_snaps=True)[2][nd:nd + ny, nd:nd + nx]
        m = np.abs(g0["raw_u"]) > 1e-6 * np.abs(g0["raw_u"]).max()
        ratios.append((g0["raw_u"][m] / g3["raw_u"][m], m))
    f = np.float32
    x = (np.arange(1, nx + 1, dtype=np.float32) * f(dh))[None, :]
    y = (np.arange(1, ny + 1, dtype=np.float32) * f(dh))[:, None]
    xmin, xmax = f(dh), f(nx) * f(dh)
    w3 = np.sqrt(Ws) * (np.arcsinh(((xmax - x) / y).astype(np.float64)) - np.arcsinh(((xmin - x) / y).astype(np.float64)))
    we3 = w3 + 0.005 * w3.max()
    cvs = float(np.mean(vs.astype(np.float64)))
    for r, m in ratios:
        np.testing.assert_allclose(r / cvs ** 2, we3[m], rtol=1e-9)
    assert we3.max() / we3.min() > 10.0     # the weight varies: the check sees the geometry


def test_pseudo_hessian_is_sum_of_squared_snapshots():
    """hessian.c: H_mu = sum over stored steps 1..NTDTINV-1 of uy^2 + (uyx/2)^2,
    H_rho = sum vyp1^2, running on across shots; hessian_out.c chain rule.
    Checked against numpy sums of the forward snapshots of orc_forward."""
    nx, ny, nt, dtinv = 60, 40, 300, 3
    c = cfg(nx=nx, ny=ny, fdorder=6, maxrelerror=1, fw=8, free_surf=1, L=2, nt=nt, dtinv=dtinv, dt=1e-4, dh=0.25,
            damping=300.0, fpml=60.0, fl=[6.25, 50.0])
    vs, rho, qs = model(nx, ny, 0.25, seed=7, qs=20.0)
    sig = np.zeros((2, nt), np.float32)
    sig[:, 0] = 1.0                       # QUELLART 6 spike ...
    sig = O64.timedomain_filt(sig, 1e-4, 80.0, 4, 1).astype(np.float32)   # ... low-passed at FC_HESSIAN
    si = np.array([20, 40], np.int32)
    sj = np.array([3, 3], np.int32)
    hv, hr = O64.pseudo_hessian(c, vs, rho, qs, si, sj, sig)
    hu = np.zeros((ny, nx))
    hrr = np.zeros((ny, nx))
    last = np.zeros((ny, nx))
    for s in range(2):
        _, snaps, _ = O64.forward(c, vs, rho, qs, si[s:s + 1], sj[s:s + 1], sig[s], np.zeros(0, np.int32),
                                  np.zeros(0, np.int32), want_snaps=True)
        n = snaps.shape[0]
        hu += np.sum(snaps[:n - 1, 1] ** 2 + (0.5 * snaps[:n - 1, 2]) ** 2, axis=0)
        hrr += np.sum(snaps[:n - 1, 0] ** 2, axis=0)
        last += snaps[n - 1, 1] ** 2
    v, r = vs.astype(np.float64), rho.astype(np.float64)
    np.testing.assert_allclose(hr, hrr + v ** 4 * hu, rtol=1e-11)
    np.testing.assert_allclose(hv, 4 * r * r * v * v * hu, rtol=1e-11)
    assert np.max(last) > 0          # the left-out last stored step carries energy (quirk matters)
    h32v, h32r = O32.pseudo_hessian(c, vs, rho, qs, si, sj, sig)
    assert rel_l2(h32v, hv) < 1e-5 and rel_l2(h32r, hr) < 1e-5


# ---------------------------------------------------------------------------
# gradient post-processing (gauss_filt_var.c, taper_grad.c)
# ---------------------------------------------------------------------------
def _gauss_kernel(hx, hz):
    """Normalised Gaussian with half widths hx (x) and hz (y), sigma = h // 2."""
    sx, sz = 2.0 * (hx // 2) ** 2, 2.0 * (hz // 2) ** 2
    jj = np.arange(-hz, hz + 1, dtype=np.float64)[:, None]
    ii = np.arange(-hx, hx + 1, dtype=np.float64)[None, :]
    k = np.exp(-ii * ii / sx - jj * jj / sz)
    return k / k.sum()


def test_gauss_filt_var_uniform_is_normalised_gaussian_correlation():
    from scipy.ndimage import correlate
    rng = np.random.default_rng(11)
    ny, nx = 37, 53
    g = rng.standard_normal((ny, nx))
    # vs 400, FC_END 20: lam 20 m; WD_DAMP 0.25 -> 5 m / 0.25 m = 20 -> 21 cells,
    # hz = 10; WD_DAMP1 0.15 -> 12 -> 13 cells, hx = 6
    vel = np.full((ny, nx), 400.0, np.float32)
    out = O64.gauss_filt_var(g, vel, 0.25, 0.25, 0.15, 20.0)
    ref = correlate(g, _gauss_kernel(6, 10), mode="nearest")
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()
    # a transposed kernel (hx <-> hz) is far off
    wrong = correlate(g, _gauss_kernel(10, 6), mode="nearest")
    assert np.abs(out - wrong).max() > 1e-2
    out32 = O32.gauss_filt_var(g.astype(np.float32), vel, 0.25, 0.25, 0.15, 20.0)
    assert rel_l2(out32, ref) < 1e-6


def test_gauss_filt_var_minimum_size_and_two_layers():
    from scipy.ndimage import correlate
    rng = np.random.default_rng(12)
    ny, nx = 30, 41
    g = rng.standard_normal((ny, nx))
    vel = np.full((ny, nx), 400.0, np.float32)
    vel[ny // 2:, :] = 100.0
    # top: hx 6, hz 10 as above; bottom (vs 100): lam 5 m -> 5 and 3 cells,
    # odd, raised to 6 -> hx = hz = 3
    out = O64.gauss_filt_var(g, vel, 0.25, 0.25, 0.15, 20.0)
    top = correlate(g, _gauss_kernel(6, 10), mode="nearest")
    bot = correlate(g, _gauss_kernel(3, 3), mode="nearest")
    assert np.abs(out[:ny // 2] - top[:ny // 2]).max() < 1e-12
    assert np.abs(out[ny // 2:] - bot[ny // 2:]).max() < 1e-12


def test_gauss_filt_var_constant_and_copy():
    vel = (150.0 + 100.0 * np.random.default_rng(3).random((20, 25))).astype(np.float32)
    c = np.full((20, 25), 2.5)
    assert np.abs(O64.gauss_filt_var(c, vel, 0.25, 0.3, 0.2, 30.0) - 2.5).max() < 1e-14
    g = np.random.default_rng(4).standard_normal((20, 25)).astype(np.float32)
    # WD_DAMP 0: FILT_SIZE_Z = 0 -> copy
    assert np.array_equal(O32.gauss_filt_var(g, vel, 0.25, 0.0, 0.2, 30.0), g)


def _win(h, tl):
    return np.exp(-0.5 * (3.0 * (h - tl) / (tl / 2.0)) ** 2)


def test_taper_grad_vertical_window_closed_form():
    ny, nx = 12, 60
    g = np.ones((ny, nx), np.float32)
    gt = (5, 15, 40, 50)
    src_i, src_j = [30], [4]
    rec_i, rec_j = [20, 22], [2, 2]
    out, tap = O32.taper_grad(g, 1, 0.25, gradt=gt, src_i=src_i, src_j=src_j, rec_i=rec_i, rec_j=rec_j)
    exp = np.zeros(nx)
    for i in range(1, nx + 1):
        if 5 < i < 50:
            h = i - 5
            w = 1.0
            if h <= 10:
                w = _win(h, 10)
            if h >= 45 - 10:
                w = _win(h - 35 + 10, 10)
            exp[i - 1] = w
    assert np.abs(tap - exp[None, :]).max() < 1e-6
    # equal flanks: the window is symmetric about the middle of (GRADT1, GRADT4),
    # peaks at 1 and vanishes outside
    assert np.array_equal(tap[0, 5:49], tap[0, 5:49][::-1])
    assert tap.max() == 1.0 and not tap[:, :5].any() and not tap[:, 49:].any()
    want = np.broadcast_to(exp, (ny, nx)).copy()
    want[3, 29] = 0.0
    want[1, 19] = want[1, 21] = 0.0
    assert np.abs(out - want).max() < 1e-6
    assert out[3, 29] == 0.0 and out[1, 19] == 0.0


def test_taper_grad_horizontal_window_closed_form():
    ny, nx = 40, 7
    g = np.full((ny, nx), 2.0, np.float32)
    out, tap = O32.taper_grad(g, 2, 0.5, gradt=(10, 20, 0, 0))
    col = np.zeros(ny)
    for j in range(1, ny + 1):
        if 10 <= j <= 20:
            col[j - 1] = _win(j - 10 + 1, 10) if j - 10 + 1 <= 10 else 1.0
        elif j > 20:
            col[j - 1] = j * 0.5
    assert np.abs(tap - col[:, None]).max() < 1e-5
    assert np.abs(out - 2.0 * col[:, None]).max() < 1e-5


def test_taper_grad_source_bowl_closed_form():
    from math import erf
    ny, nx, dh, R, fs = 41, 45, 0.5, 4.0, 1
    si, sj = 23, 21
    g = np.ones((ny, nx), np.float32)
    out, tap = O32.taper_grad(g, 3, dh, srtshape=1, srtradius=R, filtsize=fs, src_i=[si], src_j=[sj])
    ng = int(np.ceil(2 * R / dh))
    maxrad = np.sqrt(2 * R * R)
    m = np.zeros((ng, ng))
    for j in range(1, ng + 1):
        for i in range(1, ng + 1):
            x, y = (i - ng / 2 - 0.5) * dh, (j - ng / 2 - 0.5) * dh
            m[j - 1, i - 1] = erf(np.hypot(x, y) / maxrad)
    m -= m.min()
    edge = max(m[0, ng // 2 - 1], m[ng // 2 - 1, 0], m[ng // 2 - 1, ng - 1], m[ng - 1, ng // 2 - 1])
    m = np.minimum(m / edge, 1.0)
    ms = np.ones((ny, nx))
    c = int(np.ceil(ng / 2))
    for iy in range(1, ng + 1):
        for ix in range(1, ng + 1):
            xx, yy = si + ix - c, sj + iy - c
            if 1 <= xx <= nx and 1 <= yy <= ny:
                ms[yy - 1, xx - 1] = min(ms[yy - 1, xx - 1], m[iy - 1, ix - 1])
    ms = (ms - ms.min()) / (ms - ms.min()).max()
    ms[sj - 1 - fs:sj + fs, si - 1 - fs:si + fs] = 0.0
    assert np.abs(tap - ms).max() < 1e-6
    assert np.abs(out - ms).max() < 1e-6
    assert tap.max() == 1.0 and out[sj - 1, si - 1] == 0.0


# ---------------------------------------------------------------------------
# data conditioning: time_window.c, stf.c
# ---------------------------------------------------------------------------
def test_time_window_closed_form_and_product_host_code():
    rng = np.random.default_rng(41)
    ntr, ns, dt, gamma = 5, 300, 1e-3, 2000.0
    d = rng.standard_normal((ntr, ns)).astype(np.float32)
    t = (np.arange(1, ns + 1) * np.float32(dt)).astype(np.float64)
    out2 = O32.time_window(d, dt, 2, gamma)
    want2 = d * np.exp(-gamma * (t - np.float32(dt)) ** 2)[None, :]
    assert np.abs(out2 - want2).max() < 1e-6 * np.abs(d).max()
    assert np.array_equal(out2[:, 0], d[:, 0])        # t = DT: taper 1
    picks = np.array([0.01, 0.05, 0.1, 0.2, 0.25], np.float32)
    out1 = O32.time_window(d, dt, 1, gamma, twlength_plus=0.02, picks=picks)
    for i in range(ntr):
        p = float(np.float32(picks[i] + np.float32(0.02)))
        after = t >= p
        assert np.array_equal(out1[i, ~after], d[i, ~after])
        assert np.abs(out1[i, after] - d[i, after] * np.exp(-gamma * (t[after] - p) ** 2)).max() < 1e-6 * np.abs(d).max()
    dn = product()
    assert np.array_equal(dn.time_window(d, dt, 2, gamma), out2)
    assert np.array_equal(dn.time_window(d, dt, 1, gamma, 0.02, picks), out1)


def test_stf_oracle_recovers_the_convolution_filter():
    import stf_oracle
    rng = np.random.default_rng(42)
    ntr, ns = 7, 200
    h = np.array([0.0, 0.5, 1.0, -0.3, 0.1])
    syn = rng.standard_normal((ntr, ns))
    syn[:, ns - len(h):] = 0.0                       # the convolution fits the record
    obs = np.stack([np.convolve(syn[i], h)[:ns] for i in range(ntr)])
    sig = rng.standard_normal(ns)
    sig[ns - len(h):] = 0.0
    est = stf_oracle.stf(obs, syn, sig, eps_stf=0.0)
    want = np.convolve(sig, h)[:ns]
    assert np.abs(est - want).max() < 1e-9 * np.abs(want).max()
    # a correlation (conjugate on the wrong factor) would not match
    wrong = np.correlate(np.concatenate([sig, np.zeros(len(h))]), h, "full")[len(h) - 1:len(h) - 1 + ns]
    assert np.abs(est - wrong).max() > 1e-2
    # muted traces drop out: muting half of them keeps the identity
    mute = np.arange(ntr) % 2 == 0
    assert np.abs(stf_oracle.stf(obs, syn, sig, mute=mute, eps_stf=0.0) - want).max() < 1e-9 * np.abs(want).max()


def test_stf_normalisation_helpers():
    import stf_oracle
    x = np.array([[10.0, 1.0, -2.0, 0.5], [3.0, 0.0, 0.0, 0.0]], np.float32)
    y = stf_oracle.normalize_max(x)
    assert np.array_equal(y[0], np.array([10.0, 0.5, -1.0, 0.25], np.float32))   # sample 1 excluded
    assert np.array_equal(y[1], x[1])                                            # zero max -> 1
    assert np.array_equal(stf_oracle.trace_max(x), np.array([2.0, 1.0], np.float32))
