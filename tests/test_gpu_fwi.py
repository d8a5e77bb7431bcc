"""The device FWI loop (dn_fwi_iteration: gradient, L-BFGS, line search,
model update, stage rule) against the oracle's restatement of denise.c's
iteration (oracle/fwi_oracle.py) on the same inputs, with and without the
gradient post-processing of denise.c:1818-1847, with L-BFGS and with PCG.

The first iteration is bit-identical.  From the second on, the product's
FP64 L-BFGS dot products and the reference's (and the oracle's) float ones
give directions that differ by rounding, which the trial steps of the line
search amplify: misfits agree to 1e-4 relative, model updates to 1e-4
relative L2, and the line-search decisions (step1, step3) are identical."""
import os
import sys

import numpy as np
import pytest

from common import ROOT, model, oracle, product, rel_l2

sys.path.insert(0, os.path.join(ROOT, "oracle"))
from fwi_oracle import FwiOracle  # noqa: E402
from pyoracle import make_cfg  # noqa: E402

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a)
    sys.stdout.flush()


CASES = [
    ("viscoelastic L-BFGS vs+rho", dict(L=2, opts=dict(inv_rho_iter=0))),
    ("viscoelastic PCG vs+rho", dict(L=2, opts=dict(inv_rho_iter=0, grad_method=1))),
    ("elastic L-BFGS with smoothing and taper",
     dict(L=0, opts=dict(inv_rho_iter=1000, spatfilter=4, wd_damp=0.25, wd_damp1=0.25, fc_end=40.0,
                         taper_vert=1, gradt=(3, 10, 86, 96)))),
]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_fwi_iterations_match_oracle(case):
    dn = product()
    O = oracle("f32")
    name, c = case
    nx, ny, nt, dt, dh, fc = 100, 40, 900, 1e-4, 0.25, 40.0
    L = c["L"]
    fl = [10.0, 60.0][:L]
    vs, rho, _ = model(nx, ny, dh, seed=31, vs0=120.0, vs1=280.0, rho0=1700.0, rho1=1900.0)
    qs = np.full((ny, nx), 20.0, np.float32) if L else None
    j = np.arange(ny)[:, None]
    i = np.arange(nx)[None, :]
    bump = np.exp(-((i - 50.0) ** 2 + (j - 10.0) ** 2) / (2 * 5.0 ** 2))
    vt = (vs * (1.0 + 0.1 * bump)).astype(np.float32)
    rec_i = np.arange(12, nx - 12, 3, dtype=np.int32)
    rec_j = np.full_like(rec_i, 2)
    src_i = np.array([20, 50, 80], np.int32)
    src_j = np.full(3, 3, np.int32)
    sig = np.tile(dn.wavelet(1, nt, dt, 0.0, fc, 1.0), (3, 1)).astype(np.float32)
    cfg = make_cfg(nx, ny, fdorder=4, fw=12, free_surf=1, L=L, nt=nt, dtinv=1, dt=dt, dh=dh, damping=300.0,
                   fpml=fc, fl=fl)
    obs = np.stack([O.forward(cfg, vt, rho, qs, src_i[k:k + 1], src_j[k:k + 1], sig[k:k + 1], rec_i, rec_j)[0]
                    for k in range(3)]).astype(np.float32)
    opts = dict(testshot_start=1, testshot_end=3, testshot_incr=2, vs_lower=50.0, vs_upper=800.0,
                rho_lower=500.0, rho_upper=4000.0, eps_scale=0.01, stepmax=4)
    opts.update(c["opts"])
    say("case", name)
    ref = FwiOracle(O, cfg, vs, rho, qs, src_i, src_j, sig, rec_i, rec_j, obs, **opts)
    ctx = dn.Context(nx, ny, fdorder=4, fw=12, free_surf=1, L=L, fl=fl, dt=dt, dh=dh, damping=300.0, fpml=fc)
    ctx.set_model(vs, rho, qs)
    ctx.set_receivers(rec_i, rec_j)
    ctx.set_shots(src_i, src_j, sig)
    ctx.set_observed(obs)
    ctx.fwi_start(**opts)
    for it in range(4):
        r = ctx.fwi_iteration()
        w = ref.iteration()
        gv, gr = ctx.fwi_model()
        say("  iter %d  L2t %s | %s  step1 %d/%d  eps_u %.4g/%.4g" % (
            it + 1, " ".join("%.6g" % v for v in r["L2t"]), " ".join("%.6g" % v for v in w["L2t"]),
            r["step1"], w["step1"], r["eps_u"], w["eps_u"]))
        assert (r["step1"], r["step3"], r["next_stage"]) == (w["step1"], w["step3"], w["next_stage"])
        for a, b in zip(r["L2t"], w["L2t"]):
            assert abs(a - b) <= 1e-4 * abs(b) + 1e-12
        dv = rel_l2(gv.astype(np.float64) - vs, ref.vs.astype(np.float64) - vs)
        dr = rel_l2(gr.astype(np.float64) - rho, ref.rho.astype(np.float64) - rho) if np.any(ref.rho != rho) else 0.0
        say("    model update rel L2: vs %.2e  rho %.2e" % (dv, dr))
        if it == 0:
            assert dv == 0.0 and dr == 0.0
        assert dv < 1e-4 and dr < 1e-4
    ctx.close()
