"""The four workloads at their full sizes on the GPU.

* Elastic SH, 400 x 120, FDORDER 4 Holberg, 5000 steps, one shot: forward
  seismograms and the gradient against the FP32 oracle (bars 1e-5 and 1e-4
  relative L2), the FP64 oracle's distance printed beside them.
* Viscoelastic SH, 1200 x 300, FDORDER 8, L 2, 10 000 steps (the measured
  workload): one of its 40 shots through forward and gradient against the
  FP32 oracle, the FP64 distance printed.
* Batched gradient, 2000 x 400: eight shots with record lengths between 0.4
  and 1.2 s and wavelets low-passed at 15, 30 and 60 Hz (Butterworth order 4)
  in one batch give the seismograms, misfit and gradient of the same shots
  advanced one at a time, bit for bit; the shortest of them against the FP32
  and FP64 oracles.
* Reduced-precision snapshots, 4000 x 800, 12 000 steps, every fourth step
  stored: BF16 and FP16 storage against FP32 storage, gradient error printed
  and bounded; on a cropped window of the same model with a 6000-step
  record, FP32, BF16 and FP16 storage against the FP32 and FP64 oracles.
The oracle runs on the host cores (OpenMP), tens of seconds in all.
"""
import os
import sys

import numpy as np
import pytest

from common import ROOT, model, oracle, product, rel_l2, ricker

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
GRAD_TOL = 1e-4


def say(*a):
    print(*a)
    sys.stdout.flush()


def _checkerboard(nx, ny, dh, seed_shift=0):
    """vs 100 -> 350 m/s with depth times a +-10 % checkerboard of 8 m cells
    (true) and the plain gradient (start); rho 1700 -> 2000; Q 20."""
    j = np.arange(ny, dtype=np.float64)[:, None]
    i = np.arange(nx, dtype=np.float64)[None, :]
    vs1d = 100.0 + 250.0 * j / (ny - 1) + 0.0 * i
    cell = 8.0 / dh
    cb = np.where(((np.floor(i / cell + seed_shift) + np.floor(j / cell)) % 2) == 0, 1.0, -1.0)
    vs_true = (vs1d * (1.0 + 0.1 * cb)).astype(np.float32)
    rho = (1700.0 + 300.0 * j / (ny - 1) + 0.0 * i).astype(np.float32)
    return vs_true, vs1d.astype(np.float32), rho, np.full((ny, nx), 20.0, np.float32)


def _report_grad(tag, g, r32, r64):
    say("  %s rel L2 vs oracle f32 %.3e, vs f64 %.3e (oracle f32 vs f64 %.3e)" % (
        tag, rel_l2(g, r32), rel_l2(g, r64), rel_l2(r32, r64)))
    return rel_l2(g, r32)


def test_elastic_400x120_one_shot():
    from pyoracle import make_cfg
    dn = product()
    nx, ny, dh, dt, nt, fc = 400, 120, 0.25, 1e-4, 5000, 40.0
    vs_true, rho, _ = model(nx, ny, dh, seed=11)                    # linear 120 -> 300 m/s, Gaussian anomalies
    vs0, _, _ = model(nx, ny, dh, seed=11, nanom=0)                 # the plain gradient
    src_i = np.array([int(np.floor(10.0 / dh + 0.5))], np.int32)
    src_j = np.array([int(np.floor(0.5 / dh + 0.5))], np.int32)
    rec_x = 6.0 + 0.5 * np.arange(96)
    rec_i = np.floor(rec_x / dh + 0.5).astype(np.int32)
    rec_j = np.full(96, int(np.floor(0.5 / dh + 0.5)), np.int32)
    sig = ricker(nt, dt, fc)[None]
    kw = dict(fdorder=4, maxrelerror=1, fw=20, free_surf=1, L=0, dt=dt, dh=dh, damping=300.0, fpml=fc, dtinv=1)
    ctx = dn.Context(nx, ny, **kw)
    cfg = make_cfg(nx, ny, nt=nt, fl=[], **kw)
    ctx.set_receivers(rec_i, rec_j)
    ctx.set_shots(src_i, src_j, sig)
    o32, o64 = oracle("f32"), oracle("f64")
    say("case elastic 400x120 FD4, 5000 steps: oracle forward of the true model")
    obs, _, _ = o32.forward(cfg, vs_true, rho, None, src_i, src_j, sig[0], rec_i, rec_j)
    obs64, _, _ = o64.forward(cfg, vs_true, rho, None, src_i, src_j, sig[0], rec_i, rec_j)
    ctx.set_model(vs_true, rho)
    seis = ctx.forward()[0]
    e32, e64 = rel_l2(seis, obs), rel_l2(seis, obs64)
    say("  seismograms rel L2 vs oracle f32 %.3e, vs f64 %.3e (oracle f32 vs f64 %.3e)" % (
        e32, e64, rel_l2(obs, obs64)))
    assert np.max(np.abs(obs)) > 0 and e32 <= SEIS_TOL
    obs = np.ascontiguousarray(obs[None], np.float32)
    ctx.set_model(vs0, rho)
    ctx.set_observed(obs)
    l2, en = ctx.gradient()
    gvs, grho, _, _ = ctx.get_gradient()
    say("  oracle gradients (f32, f64)")
    r32 = o32.gradient(cfg, vs0, rho, None, src_i, src_j, sig, rec_i, rec_j, obs)
    r64 = o64.gradient(cfg, vs0, rho, None, src_i, src_j, sig, rec_i, rec_j, obs)
    ev = _report_grad("grad_vs ", gvs, r32["grad_vs"], r64["grad_vs"])
    er = _report_grad("grad_rho", grho, r32["grad_rho"], r64["grad_rho"])
    say("  l2 %.9e oracle %.9e; energy %.9e oracle %.9e" % (l2, r32["l2"], en, r32["energy"]))
    assert np.max(np.abs(r32["grad_vs"])) > 0
    assert ev <= GRAD_TOL and er <= GRAD_TOL
    assert abs(l2 - r32["l2"]) <= 1e-9 * abs(r32["l2"])
    assert abs(en - r32["energy"]) <= 1e-9 * abs(r32["energy"])
    ctx.close()


def test_visco_1200x300_one_shot_full_record():
    from pyoracle import make_cfg
    sys.path.insert(0, ROOT)
    import bench
    dn = product()
    W = bench.workload("c2", 10000)
    nx, ny, nt = W["nx"], W["ny"], W["nt"]
    k = 20                                                            # a shot in the middle of the line
    src_i, src_j = W["src_i"][k:k + 1], W["src_j"][k:k + 1]
    sig = bench.ricker(dn, nt, W["dt"], W["fc"])[None]
    kw = dict(fdorder=W["fdorder"], maxrelerror=1, fw=W["fw"], free_surf=1, L=W["L"], dt=W["dt"], dh=W["dh"],
              damping=350.0, fpml=W["fc"], dtinv=W["dtinv"])
    ctx = dn.Context(nx, ny, fl=W["fl"], **kw)
    cfg = make_cfg(nx, ny, nt=nt, fl=W["fl"], **kw)
    ctx.set_receivers(W["rec_i"], W["rec_j"])
    ctx.set_shots(src_i, src_j, sig)
    o32, o64 = oracle("f32"), oracle("f64")
    say("case visco 1200x300 FD8 L2, 10000 steps, shot %d: oracle forward of the true model" % k)
    obs, _, _ = o32.forward(cfg, W["vs_true"], W["rho"], W["qs"], src_i, src_j, sig[0], W["rec_i"], W["rec_j"])
    say("  oracle forward f64")
    obs64, _, _ = o64.forward(cfg, W["vs_true"], W["rho"], W["qs"], src_i, src_j, sig[0], W["rec_i"], W["rec_j"])
    ctx.set_model(W["vs_true"], W["rho"], W["qs"])
    seis = ctx.forward()[0]
    e32, e64 = rel_l2(seis, obs), rel_l2(seis, obs64)
    say("  seismograms rel L2 vs oracle f32 %.3e, vs f64 %.3e (oracle f32 vs f64 %.3e)" % (
        e32, e64, rel_l2(obs, obs64)))
    assert e64 <= 1e-3
    assert np.max(np.abs(obs)) > 0 and e32 <= SEIS_TOL
    obs = np.ascontiguousarray(obs[None], np.float32)
    ctx.set_model(W["vs_start"], W["rho"], W["qs"])
    ctx.set_observed(obs)
    l2, en = ctx.gradient()
    gvs, grho, _, _ = ctx.get_gradient()
    say("  oracle gradient f32")
    r32 = o32.gradient(cfg, W["vs_start"], W["rho"], W["qs"], src_i, src_j, sig, W["rec_i"], W["rec_j"], obs)
    say("  oracle gradient f64")
    r64 = o64.gradient(cfg, W["vs_start"], W["rho"], W["qs"], src_i, src_j, sig, W["rec_i"], W["rec_j"], obs)
    ev = _report_grad("grad_vs ", gvs, r32["grad_vs"], r64["grad_vs"])
    er = _report_grad("grad_rho", grho, r32["grad_rho"], r64["grad_rho"])
    say("  l2 %.9e oracle %.9e; energy %.9e oracle %.9e" % (l2, r32["l2"], en, r32["energy"]))
    assert np.max(np.abs(r32["grad_vs"])) > 0
    assert ev <= GRAD_TOL and er <= GRAD_TOL
    assert abs(l2 - r32["l2"]) <= 1e-9 * abs(r32["l2"])
    assert abs(en - r32["energy"]) <= 1e-9 * abs(r32["energy"])
    ctx.close()


def test_batched_uneven_stages_2000x400_equals_one_at_a_time():
    dn = product()
    nx, ny, dh, dt, fc = 2000, 400, 0.2, 1e-4, 25.0
    vs_true, vs0, rho, qs = _checkerboard(nx, ny, dh)
    rng = np.random.default_rng(7)
    ns = 8
    nts = np.sort(rng.integers(4000, 12001, ns)).astype(np.int32)[rng.permutation(ns)]
    ntm = int(nts.max())
    stages = [15.0, 30.0, 60.0]
    sig = np.zeros((ns, ntm), np.float32)
    base = dn.wavelet(1, ntm, dt, 0.0, fc, 1.0)
    for s in range(ns):
        sig[s] = dn.timedomain_filt(base[None], dt, stages[s % 3], 4, 1)[0]
        sig[s, nts[s]:] = 0.0
    src_i = np.linspace(20, nx - 21, ns).astype(np.int32)
    src_j = np.full(ns, 3, np.int32)
    rec_i = np.arange(25, nx - 25, 8, dtype=np.int32)
    rec_j = np.full_like(rec_i, 3)
    ctx = dn.Context(nx, ny, fdorder=8, maxrelerror=1, fw=20, free_surf=1, L=2, fl=[6.25, 50.0], dt=dt, dh=dh,
                     damping=350.0, fpml=fc, dtinv=4)
    ctx.set_receivers(rec_i, rec_j)
    ctx.set_shots(src_i, src_j, sig, nts)
    say("case batched 2000x400: %d shots, steps %s, low-pass stages %s" % (ns, [int(v) for v in nts], stages))
    ctx.set_model(vs_true, rho, qs)
    obs = ctx.forward()
    ctx.set_model(vs0, rho, qs)
    ctx.set_observed(obs)
    out = {}
    for batch in (0, 1):
        ctx.set_batch(batch)
        seis = ctx.forward()
        l2, en = ctx.gradient()
        gvs, grho, _, _ = ctx.get_gradient()
        out[batch] = (seis, gvs, grho, l2, en)
        say("  %s: l2 %.9e energy %.9e" % ("one batch" if batch == 0 else "one shot at a time", l2, en))
    a, b = out[0], out[1]
    for s in range(ns):
        assert np.all(a[0][s][:, nts[s]:] == 0.0)
    assert np.max(np.abs(a[1])) > 0
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[3] == b[3] and a[4] == b[4]
    ctx.close()
    # the shortest shot of the batch against the oracles, alone
    from pyoracle import make_cfg
    k = int(np.argmin(nts))
    ntk = int(nts[k])
    kw = dict(fdorder=8, maxrelerror=1, fw=20, free_surf=1, L=2, dt=dt, dh=dh, damping=350.0, fpml=fc, dtinv=4)
    c1 = dn.Context(nx, ny, fl=[6.25, 50.0], **kw)
    c1.set_receivers(rec_i, rec_j)
    c1.set_shots(src_i[k:k + 1], src_j[k:k + 1], sig[k:k + 1, :ntk])
    c1.set_model(vs0, rho, qs)
    ob = np.ascontiguousarray(obs[k:k + 1, :, :ntk], np.float32)
    c1.set_observed(ob)
    c1.gradient()
    gvs, grho, _, _ = c1.get_gradient()
    cfg = make_cfg(nx, ny, nt=ntk, fl=[6.25, 50.0], **kw)
    say("  shot %d (%d steps, %g Hz stage) against the oracles f32 and f64" % (k, ntk, stages[k % 3]))
    o32, o64 = oracle("f32"), oracle("f64")
    r32 = o32.gradient(cfg, vs0, rho, qs, src_i[k:k + 1], src_j[k:k + 1], sig[k:k + 1, :ntk], rec_i, rec_j, ob)
    r64 = o64.gradient(cfg, vs0, rho, qs, src_i[k:k + 1], src_j[k:k + 1], sig[k:k + 1, :ntk], rec_i, rec_j, ob)
    ev = _report_grad("grad_vs ", gvs, r32["grad_vs"], r64["grad_vs"])
    er = _report_grad("grad_rho", grho, r32["grad_rho"], r64["grad_rho"])
    assert ev <= GRAD_TOL and er <= GRAD_TOL
    c1.close()


def test_reduced_precision_snapshots_4000x800():
    sys.path.insert(0, ROOT)
    import bench
    dn = product()
    nx, ny, dh, dt, nt, fc = 4000, 800, 0.1, 5e-5, 12000, 60.0
    vs_true, vs0, rho, qs = bench.layered_lenses(nx, ny, dh, seed=5)
    ns = 2
    src_i = np.array([1000, 3000], np.int32)
    src_j = np.full(ns, 5, np.int32)
    rec_i = np.arange(30, nx - 30, 10, dtype=np.int32)
    rec_j = np.full_like(rec_i, 5)
    sig = np.tile(dn.wavelet(1, nt, dt, 0.0, fc, 1.0), (ns, 1))
    kw = dict(fdorder=8, maxrelerror=1, fw=20, free_surf=1, L=2, fl=[6.25, 50.0], dt=dt, dh=dh, damping=350.0,
              fpml=fc, dtinv=4)
    say("case reduced-precision snapshots 4000x800, %d shots x %d steps, every 4th stored" % (ns, nt))
    grads = {}
    obs = None
    for prec in (0, 1, 2):
        ctx = dn.Context(nx, ny, snap_precision=prec, **kw)
        ctx.set_receivers(rec_i, rec_j)
        ctx.set_shots(src_i, src_j, sig)
        if obs is None:
            ctx.set_model(vs_true, rho, qs)
            obs = ctx.forward()
            assert np.all(np.isfinite(obs)) and np.max(np.abs(obs)) > 0
        ctx.set_model(vs0, rho, qs)
        ctx.set_observed(obs)
        l2, en = ctx.gradient()
        gvs, grho, _, _ = ctx.get_gradient()
        grads[prec] = (gvs, grho, l2)
        say("  %s snapshots: l2 %.9e" % (["fp32", "bf16", "fp16"][prec], l2))
        ctx.close()
    g0 = grads[0]
    assert np.all(np.isfinite(g0[0])) and np.max(np.abs(g0[0])) > 0
    for prec, tol in ((1, 1e-2), (2, 2e-3)):
        ev, er = rel_l2(grads[prec][0], g0[0]), rel_l2(grads[prec][1], g0[1])
        say("  %s vs fp32 storage: grad_vs rel L2 %.3e, grad_rho rel L2 %.3e" % (["", "bf16", "fp16"][prec], ev, er))
        assert ev <= tol and er <= tol
        assert grads[prec][2] == g0[2]          # the misfit comes from the forward pass, not the snapshots


def test_4000x800_window_against_fp64_oracle():
    """The reduced-precision workload's grid spacing, time step and model on
    a cropped window (600 x 300 cells: x 319-379 m, the top 30 m, over four
    lenses, the shallowest 2.3 m deep) with a 6000-step record, one shot,
    every fourth step stored, FP32 / BF16 / FP16 snapshot storage.  Against
    the FP32 oracle: FP32 storage within the gradient bar.  Against the FP64
    oracle (the observed data modelled in FP64 as well): two residuals.
    (a) The workload's own: layered start model, the lenses the only
        difference.  The FP32 oracle's gradient is 6e-7 from the FP64
        one; the product's FP32 path within 1e-5, BF16 and FP16 storage
        within 1e-2.
    (b) The start model 5 % slower than the true one everywhere: a residual
        spread over the whole record, where the FP32 oracle's gradient is
        1.1e-3 from the FP64 one (cancellation in the correlation sum);
        the product's FP32 path within 2e-3, BF16 and FP16 within 1e-2."""
    sys.path.insert(0, ROOT)
    import bench
    from pyoracle import make_cfg
    dn = product()
    dh, dt, nt, fc = 0.1, 5e-5, 6000, 60.0
    vt, v0, rh, q = bench.layered_lenses(4000, 800, dh, seed=5)
    win = (slice(0, 300), slice(3190, 3790))
    vs_true, vs_layer, rho, qs = (np.ascontiguousarray(a[win]) for a in (vt, v0, rh, q))
    ny, nx = vs_true.shape
    src_i, src_j = np.array([300], np.int32), np.array([5], np.int32)
    rec_i = np.arange(30, nx - 30, 4, dtype=np.int32)
    rec_j = np.full_like(rec_i, 5)
    sig = dn.wavelet(1, nt, dt, 0.0, fc, 1.0)[None].astype(np.float32)
    kw = dict(fdorder=8, maxrelerror=1, fw=20, free_surf=1, L=2, dt=dt, dh=dh, damping=350.0, fpml=fc, dtinv=4)
    cfg = make_cfg(nx, ny, nt=nt, fl=[6.25, 50.0], **kw)
    o32, o64 = oracle("f32"), oracle("f64")
    say("case 4000x800 model window %dx%d, 1 shot x %d steps: observed data (oracle f32, f64)" % (nx, ny, nt))
    obs = np.ascontiguousarray(o32.forward(cfg, vs_true, rho, qs, src_i, src_j, sig[0], rec_i, rec_j)[0][None],
                               np.float32)
    obs64 = o64.forward(cfg, vs_true, rho, qs, src_i, src_j, sig[0], rec_i, rec_j)[0][None]
    for name, vs0 in (("a: layered start model", vs_layer), ("b: start model 5% slow", (0.95 * vs_true).astype(np.float32))):
        say("  %s: oracle gradients f32, f64" % name)
        r32 = o32.gradient(cfg, vs0, rho, qs, src_i, src_j, sig, rec_i, rec_j, obs)
        r64 = o64.gradient(cfg, vs0, rho, qs, src_i, src_j, sig, rec_i, rec_j, obs64)
        say("    residual l2 %.3e (fp32) %.3e (fp64); data energy %.3e" % (r32["l2"], r64["l2"], r64["energy"]))
        assert r32["l2"] > 0 and np.all(np.isfinite(r32["grad_vs"])) and np.max(np.abs(r32["grad_vs"])) > 0
        for prec in (0, 1, 2):
            ctx = dn.Context(nx, ny, fl=[6.25, 50.0], snap_precision=prec, **kw)
            ctx.set_receivers(rec_i, rec_j)
            ctx.set_shots(src_i, src_j, sig)
            ctx.set_model(vs0, rho, qs)
            ctx.set_observed(obs)
            ctx.gradient()
            gvs, grho, _, _ = ctx.get_gradient()
            tag = ["fp32", "bf16", "fp16"][prec]
            ev = _report_grad("%s grad_vs " % tag, gvs, r32["grad_vs"], r64["grad_vs"])
            er = _report_grad("%s grad_rho" % tag, grho, r32["grad_rho"], r64["grad_rho"])
            if prec == 0:
                assert ev <= GRAD_TOL and er <= GRAD_TOL
            bound = (1e-5 if name.startswith("a") else 2e-3) if prec == 0 else 1e-2
            assert rel_l2(gvs, r64["grad_vs"]) <= bound and rel_l2(grho, r64["grad_rho"]) <= bound
            ctx.close()
