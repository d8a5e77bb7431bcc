"""Data conditioning and source-time-function inversion on the GPU against
the oracle: trace kill and offset mute in the residual, misfit and energy
(calc_res.c:37-113, calc_energy.c:23-49, with the energy's transposed
reading of the kill list), and stf.c's Wiener deconvolution (tolerance 1e-5
relative: FFT rounding differs from numpy's)."""
import sys

import numpy as np
import pytest

from common import oracle, product, rel_l2, ricker
from test_gpu_parity import DH, DT, _obs_for, setup_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a)
    sys.stdout.flush()


def masks(kill, src_i, src_j, rec_i, rec_j, mute, offsetc):
    """Per (shot, receiver) masks as calc_res / calc_energy derive them:
    float32 offsets from srcpos (here grid index * DH) and recpos * DH."""
    ns, nr = len(src_i), len(rec_i)
    nsrc, ntr = kill.shape
    flat = kill.ravel()
    skip_res = np.zeros((ns, nr), np.uint8)
    skip_en = np.zeros((ns, nr), np.uint8)
    f = np.float32
    for s in range(ns):
        for r in range(nr):
            if flat[s * ntr + r] == 1:
                skip_res[s, r] = 1
            if flat[r * nsrc + s] == 1:
                skip_en[s, r] = 1
            if mute:
                xs, ys = f(src_i[s]) * f(DH), f(src_j[s]) * f(DH)
                xr, yr = f(rec_i[r]) * f(DH), f(rec_j[r]) * f(DH)
                off = f(np.sqrt(np.float64(f((xs - xr) * (xs - xr)) + f((ys - yr) * (ys - yr)))))
                if (mute == 1 and off >= f(offsetc)) or (mute == 2 and off <= f(offsetc)):
                    skip_res[s, r] = 1
    return skip_res, skip_en


@pytest.mark.parametrize("mute", [1, 2])
def test_gradient_with_trace_kill_and_offset_mute(mute):
    say("case offset mute", mute)
    nx, ny, nt, nshots = 100, 45, 300, 3
    ctx, cfg, d = setup_case(nx, ny, 4, 2, 1, 10, nt, 60.0, nshots=nshots)
    obs = _obs_for(ctx, cfg, d, nx, ny, nshots)
    ctx.set_observed(obs)
    ntr = len(d["rec_i"])
    rng = np.random.default_rng(50 + mute)
    # the kill list of the file: nsrc x ntr, with a few 1s (and a 2, which is not 1)
    kill = (rng.random((nshots, ntr)) < 0.2).astype(np.int32)
    kill[0, 0] = 2
    offsetc = 6.0
    ctx.set_datacond(trkill=1, kill=kill, offset_mute=mute, offsetc=offsetc)
    l2, en = ctx.gradient()
    gvs, grho, _, _ = ctx.get_gradient()
    sr, se = masks(kill, d["src_i"], d["src_j"], d["rec_i"], d["rec_j"], mute, offsetc)
    say("  traces left out of the residual %d, of the energy %d (of %d)" % (sr.sum(), se.sum(), sr.size))
    assert 0 < sr.sum() < sr.size and se.sum() > 0
    o32 = oracle("f32")
    cfg.skip_res = sr.ctypes.data
    cfg.skip_en = se.ctypes.data
    r = o32.gradient(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"], d["rec_j"], obs)
    cfg.skip_res = cfg.skip_en = None
    r0 = o32.gradient(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"], d["rec_j"], obs)
    say("  grad_vs rel L2 %.2e (unconditioned differs by %.2e); l2 %.9e / %.9e; energy %.9e / %.9e" % (
        rel_l2(gvs, r["grad_vs"]), rel_l2(r0["grad_vs"], r["grad_vs"]), l2, r["l2"], en, r["energy"]))
    assert rel_l2(gvs, r["grad_vs"]) <= 1e-6 and rel_l2(grho, r["grad_rho"]) <= 1e-6
    assert rel_l2(r0["grad_vs"], r["grad_vs"]) > 1e-3
    assert abs(l2 - r["l2"]) <= 1e-9 * abs(r["l2"])
    assert abs(en - r["energy"]) <= 1e-9 * abs(r["energy"])
    # switching conditioning off restores the plain gradient
    ctx.set_datacond()
    ctx.gradient()
    g0, _, _, _ = ctx.get_gradient()
    assert rel_l2(g0, r0["grad_vs"]) <= 1e-6
    ctx.close()


@pytest.mark.parametrize("normalize", [0, 1, 2])
def test_stf_matches_oracle(normalize):
    import stf_oracle
    dn = product()
    rng = np.random.default_rng(60 + normalize)
    ntr, ns = 24, 700
    syn = rng.standard_normal((ntr, ns)).astype(np.float32)
    obs = (0.7 * syn + 0.3 * rng.standard_normal((ntr, ns))).astype(np.float32)
    sig = ricker(ns, DT, 40.0)
    mute = (np.arange(ntr) % 5 == 0).astype(np.uint8)
    got = dn.stf(obs, syn, sig, mute=mute, eps_stf=0.05, normalize=normalize)
    want = stf_oracle.stf(obs, syn, sig, mute=mute, eps_stf=0.05, normalize=normalize)
    e = rel_l2(got, want)
    say("  normalize %d: rel L2 vs oracle %.2e" % (normalize, e))
    assert e < 1e-5


def test_stf_shot_recovers_the_true_signal():
    """Observed data modelled with a 30 Hz Ricker, the context's signal a
    40 Hz Ricker: the estimate follows the 30 Hz wavelet (correlation 0.91
    measured; the regularised deconvolution loses the band edges) and equals
    the oracle's estimate from the same traces."""
    import stf_oracle
    nx, ny, nt, nshots = 120, 50, 600, 2
    ctx, cfg, d = setup_case(nx, ny, 4, 0, 1, 10, nt, 40.0, nshots=nshots)
    true = np.stack([ricker(nt, DT, 30.0) for _ in range(nshots)])
    ctx.set_signals(true)
    obs = ctx.forward()
    ctx.set_signals(d["sig"])
    ctx.set_observed(obs)
    syn = ctx.forward()
    est = ctx.stf_shot(1, eps_stf=1e-3)
    want = stf_oracle.stf(obs[1], syn[1], d["sig"][1], eps_stf=1e-3)
    cc = float(np.dot(est, true[1]) / np.linalg.norm(est) / np.linalg.norm(true[1]))
    say("  rel L2 vs oracle %.2e; correlation with the true signal %.4f" % (rel_l2(est, want), cc))
    assert rel_l2(est, want) < 1e-5
    assert cc > 0.85
    ctx.close()
