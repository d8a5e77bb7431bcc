"""Pins of the oracle's BOUNDARY 1 path (periodic left and right edges:
exchange_v.c:99-161 and exchange_s.c:95-157 on one domain, the x-CPML
switched off by the (!BOUNDARY) guards of update_v_PML.c:63,71 and
update_s_*_PML.c).  CPU only.

  energy       with FW = 0 the periodic x direction keeps the leapfrog
               scheme's discrete energy exactly (FP64, drift < 1e-12); an
               exchange shifted by one column, planted in the harness,
               breaks it.
  translation  on a periodic grid, rolling model, sources and receivers by
               s columns leaves the seismograms unchanged bit for bit (every
               cell sees the same neighbours); without BOUNDARY 1 the same
               roll changes them.
"""
import numpy as np
import pytest

from common import model, oracle

O32 = oracle("f32")
O64 = oracle("f64")


def cfg(**kw):
    from pyoracle import make_cfg
    return make_cfg(**kw)


def _state(c, seed):
    nd = c.fdorder // 2 + 1
    sh = O64.padded_shape(c)
    inner = (slice(nd, nd + c.ny), slice(nd, nd + c.nx))
    rng = np.random.default_rng(seed)
    F = {}
    for k, s in (("vy", 1.0), ("sxy", 1e3), ("syz", 1e3)):
        a = np.zeros(sh, np.float64)
        a[inner] = rng.standard_normal((c.ny, c.nx)) * s
        F[k] = a
    return F, inner


def _wrap_shifted(c, a):
    """A planted mistake: the exchange with every halo column taken one
    column too far inside."""
    nd = c.fdorder // 2 + 1
    k = c.fdorder // 2
    for l in range(1, k + 1):
        a[:, nd - 1 + c.nx + l] = a[:, nd - 1 + l + 1]
        a[:, nd - 1 + 1 - l] = a[:, nd - 1 + c.nx - l]


@pytest.mark.parametrize("fdorder", [2, 4, 8])
@pytest.mark.parametrize("planted", [False, True])
def test_periodic_leapfrog_energy_conservation(fdorder, planted):
    nx, ny = 40, 30
    c = cfg(nx=nx, ny=ny, fdorder=fdorder, fw=0, free_surf=1, L=0, dt=1e-4, dh=0.25, boundary=1)
    vs, rho, _ = model(nx, ny, 0.25, seed=2)
    P = O64.model_setup(c, vs, rho)
    hc = O64.holberg(fdorder, 1)
    pml = O64.pml(c)
    F, I = _state(c, 4)
    wrap = (lambda a: _wrap_shifted(c, a)) if planted else (lambda a: O64.wrap_x(c, a))
    wrap(F["vy"])
    wrap(F["sxy"])
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
        wrap(F["vy"])
        if es_prev is not None:
            E.append(0.5 * np.sum(rhoI * vold * F["vy"][I]) + es_prev)
        O64.update_s_elastic(c, F["vy"], uy, uyx, F["syz"], F["sxy"], P["uip"], P["ujp"], hc, pml, psi, psi)
        wrap(F["sxy"])
        es_prev = 0.5 * np.sum(F["sxy"][I] ** 2 / uip + F["syz"][I] ** 2 / ujp)
    E = np.array(E)
    drift = np.max(np.abs(E - E[0])) / E[0]
    if planted:
        assert drift > 1e-6
    else:
        assert drift < 1e-12


def _forward(o, c, vs, rho, qs, si, sj, sig, ri, rj):
    return o.forward(c, vs, rho, qs, si, sj, sig, ri, rj)[0]


@pytest.mark.parametrize("L", [0, 2])
@pytest.mark.parametrize("kind", ["f32", "f64"])
def test_periodic_translation_invariance(L, kind):
    o = O32 if kind == "f32" else O64
    nx, ny, nt, shift = 96, 40, 500, 37
    vs, rho, qs = model(nx, ny, 0.25, seed=7, qs=20.0 if L else None)
    base = dict(nx=nx, ny=ny, fdorder=8, fw=10, free_surf=1, L=L, nt=nt, dt=1e-4, dh=0.25, fl=[10.0, 80.0][:L])
    sig = O32.wavelet(1, nt, 1e-4, 0.0, 60.0, 1.0)[None]
    si, sj = np.array([3], np.int32), np.array([4], np.int32)
    ri = np.array([1, 2, 30, 70, 95, 96], np.int32)
    rj = np.full(len(ri), 3, np.int32)
    roll = lambda a: np.ascontiguousarray(np.roll(a, shift, axis=1))
    mv = lambda i: ((i - 1 + shift) % nx + 1).astype(np.int32)
    out = {}
    for boundary in (1, 0):
        c = cfg(boundary=boundary, **base)
        a = _forward(o, c, vs, rho, qs, si, sj, sig, ri, rj)
        b = _forward(o, c, roll(vs), roll(rho), roll(qs) if L else None, mv(si), sj, sig, mv(ri), rj)
        assert np.all(np.isfinite(a)) and np.max(np.abs(a)) > 0
        out[boundary] = (a, b)
    assert np.array_equal(out[1][0], out[1][1])
    assert not np.array_equal(out[0][0], out[0][1])
