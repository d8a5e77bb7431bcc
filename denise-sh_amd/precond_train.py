"""Training of the learned diagonal-Hessian preconditioner (not in the
reference).

Training pairs come from randomised models of the FWI model family: a vs
gradient with depth plus seeded Gaussian anomalies, rho gradient, Q.  For
each pair the context computes the misfit gradient at a starting model (the
smooth part of the true model), which yields the network's input maps
(dn_precond_features: source and receiver energy maps and vs), and the
label: the diagonal of the vs Hessian estimated with finite-difference
Hessian-vector probes (dn_hessian_probe), smoothed over 3x3 cells and taken
as log10 relative to its maximum (floored at 1e-4).

Everything runs on the GPU through the C-ABI (libdenise_sh.so); this module
only orchestrates.
"""
import time

import numpy as np


def random_model(nx, ny, dh, seed, vmin=100.0, vmax=350.0, nanom=6, amp=0.12):
    """vs gradient vmin -> vmax with depth plus seeded Gaussian anomalies
    (sigma 2..6 m, +-amp); rho 1700 -> 2000; Q = 20."""
    rng = np.random.default_rng(seed)
    j = np.arange(ny, dtype=np.float64)[:, None]
    i = np.arange(nx, dtype=np.float64)[None, :]
    v1 = vmin + (vmax - vmin) * j / max(ny - 1, 1) + 0.0 * i
    pert = np.zeros((ny, nx))
    for _ in range(nanom):
        cj, ci = rng.uniform(0.15 * ny, ny), rng.uniform(0, nx)
        sg = rng.uniform(2.0, 6.0) / dh
        pert += rng.uniform(-amp, amp) * np.exp(-((j - cj) ** 2 + (i - ci) ** 2) / (2 * sg * sg))
    vs_true = (v1 * (1.0 + pert)).astype(np.float32)
    vs_start = v1.astype(np.float32)
    rho = (1700.0 + 300.0 * j / max(ny - 1, 1) + 0.0 * i).astype(np.float32)
    qs = np.full((ny, nx), 20.0, np.float32)
    return vs_true, vs_start, rho, qs


def random_checkerboard(nx, ny, dh, seed):
    """The checkerboard model family with seeded parameters: vs rising
    linearly with depth from vmin (85..115 m/s) to vmax (320..380 m/s) times
    1 +- a checkerboard (cells 5..11 m, random horizontal phase, amplitude
    5..12 %) as the true model, the plain gradient as the starting model;
    rho 1700 -> 2000; Q = 20."""
    rng = np.random.default_rng(seed)
    j = np.arange(ny, dtype=np.float64)[:, None]
    i = np.arange(nx, dtype=np.float64)[None, :]
    vmin, vmax = rng.uniform(85.0, 115.0), rng.uniform(320.0, 380.0)
    v1 = vmin + (vmax - vmin) * j / max(ny - 1, 1) + 0.0 * i
    cell = rng.uniform(5.0, 11.0) / dh
    cb = np.where(((np.floor(i / cell + rng.uniform(0.0, 2.0)) + np.floor(j / cell)) % 2) == 0, 1.0, -1.0)
    vs_true = (v1 * (1.0 + rng.uniform(0.05, 0.12) * cb)).astype(np.float32)
    rho = (1700.0 + 300.0 * j / max(ny - 1, 1) + 0.0 * i).astype(np.float32)
    return vs_true, v1.astype(np.float32), rho, np.full((ny, nx), 20.0, np.float32)


def pairs_on_context(ctx, models, nprobes=8, eps=1e-2, seed0=100, log=None, geometry=None):
    """Training pairs on a context that already holds receivers and shots
    (EPRECOND 0 for the labels): for each (vs_true, vs_start, rho, qs) the
    observed data from vs_true, the gradient at vs_start, the input maps and
    the probe label.  geometry(k), if given, sets the shots of pair k on the
    context first (acquisitions that vary over the training set).  Returns
    X [n][3][H8][W8], T [n][H8][W8]."""
    n = len(models)
    X = T = None
    for k, (vt, v0, rho, qs) in enumerate(models):
        if geometry is not None:
            geometry(k)
        ctx.set_model(vt, rho, qs)
        obs = ctx.forward()
        ctx.set_observed(obs)
        ctx.set_model(v0, rho, qs)
        ctx.gradient()
        x = ctx.precond_features()
        if X is None:
            h8, w8 = x.shape[1], x.shape[2]
            X = np.zeros((n, 3, h8, w8), np.float32)
            T = np.zeros((n, h8, w8), np.float32)
        X[k] = x
        T[k] = label_from_diag(ctx.hessian_probe(nprobes=nprobes, eps=eps, seed=seed0 + k), X.shape[2], X.shape[3])
        if log:
            log("pair %d of %d" % (k + 1, n))
    return X, T


def smooth3(a):
    p = np.pad(a, 1, mode="edge")
    return sum(p[dj:dj + a.shape[0], di:di + a.shape[1]] for dj in range(3) for di in range(3)) / 9.0


def label_from_diag(diag, h8, w8):
    d = smooth3(np.abs(diag.astype(np.float64)))
    m = d.max()
    t = np.log10(np.maximum(d / m, 1e-4)) if m > 0 else np.zeros_like(d)
    out = np.zeros((h8, w8), np.float32)
    ny, nx = d.shape
    out[:ny, :nx] = t
    out[ny:, :nx] = t[ny - 1:ny, :]
    out[:, nx:] = out[:, nx - 1:nx]
    return out


class Survey:
    """One acquisition geometry on one context: surface sources and receivers."""

    def __init__(self, dn, nx, ny, dh=0.25, nshots=8, nt=900, fc=40.0, dt=1e-4, fdorder=8, fw=12, L=2, dtinv=2,
                 eprecond=1):
        self.dn, self.nx, self.ny, self.nt, self.dt = dn, nx, ny, nt, dt
        self.ctx = dn.Context(nx, ny, fdorder=fdorder, maxrelerror=1, fw=fw, free_surf=1, L=L,
                              fl=[fc / 4.0, 2.0 * fc][:L], dt=dt, dh=dh, damping=350.0, fpml=fc, dtinv=dtinv,
                              eprecond=eprecond)
        rec_i = np.arange(fw + 2, nx - fw - 1, 2, dtype=np.int32)
        self.ctx.set_receivers(rec_i, np.full_like(rec_i, 2))
        src_i = np.linspace(fw + 4, nx - fw - 4, nshots).astype(np.int32)
        sig = np.tile(dn.wavelet(1, nt, dt, 0.0, fc, 1.0), (nshots, 1))
        self.ctx.set_shots(src_i, np.full(nshots, 3, np.int32), sig)
        self.nshots = nshots

    def setup(self, vs_true, vs_start, rho, qs):
        """Observed data from the true model; the context is left at the
        starting model (the normalisation means C_vs, C_rho are taken at the
        first gradient or at dn_fwi_start, from the starting model)."""
        self.ctx.set_model(vs_true, rho, qs)
        obs = self.ctx.forward()
        self.ctx.set_observed(obs)
        self.ctx.set_model(vs_start, rho, qs)
        return obs


def make_pairs(dn, npairs, nx, ny, seed0=100, nprobes=8, eps=1e-2, **survey):
    """Training pairs (X [n][3][H8][W8], T [n][H8][W8]) and the seconds spent."""
    t0 = time.perf_counter()
    survey.setdefault("eprecond", 0)      # labels: Hessian of the unscaled misfit
    sv = Survey(dn, nx, ny, **survey)
    h8, w8 = (ny + 7) // 8 * 8, (nx + 7) // 8 * 8
    X = np.zeros((npairs, 3, h8, w8), np.float32)
    T = np.zeros((npairs, h8, w8), np.float32)
    for k in range(npairs):
        vt, v0, rho, qs = random_model(nx, ny, 0.25, seed0 + k)
        sv.setup(vt, v0, rho, qs)
        sv.ctx.gradient()
        X[k] = sv.ctx.precond_features()
        diag = sv.ctx.hessian_probe(nprobes=nprobes, eps=eps, seed=seed0 + k)
        T[k] = label_from_diag(diag, h8, w8)
    sv.ctx.close()
    return X, T, time.perf_counter() - t0


def train(net, X, T, steps=300, lr=2e-3, batch=8, seed=0):
    """Adam on random minibatches; returns the loss history."""
    rng = np.random.default_rng(seed)
    hist = []
    n = X.shape[0]
    for s in range(steps):
        idx = rng.choice(n, size=min(batch, n), replace=False)
        hist.append(net.train_step(X[idx], T[idx], lr))
    return hist


def run_fwi(sv, vs_true, vs_start, rho, qs, iters, net=None, judge_every=3, revert=True, **opts):
    """FWI of vs (rho kept) on one survey; returns per-iteration records with
    the elapsed seconds, and the final model."""
    sv.setup(vs_true, vs_start, rho, qs)
    sv.ctx.set_learned(net, judge_every, revert=revert)
    o = dict(testshot_start=1, testshot_end=sv.nshots, testshot_incr=2, vs_lower=50.0, vs_upper=800.0,
             rho_lower=500.0, rho_upper=4000.0, inv_rho_iter=1000000, eps_scale=0.01, stepmax=4)
    o.update(opts)
    sv.ctx.fwi_start(**o)
    t0 = time.perf_counter()
    recs = []
    for _ in range(iters):
        r = sv.ctx.fwi_iteration()
        r["seconds"] = time.perf_counter() - t0
        recs.append(r)
    vs, _ = sv.ctx.fwi_model()
    return recs, vs
