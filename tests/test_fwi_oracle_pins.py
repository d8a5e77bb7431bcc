"""Pins of the oracle's FWI iteration (oracle/fwi_oracle.py) by what its
parts compose to, CPU only.  Each part is pinned on its own in
test_oracle_pins.py; here the composition, with gradients computed by
independent calls of the oracle's gradient at the models the iteration
produced:

  iteration 1 (L-BFGS, empty history): the vs and rho updates point along
      minus the normalised gradients C_vs g_vs and C_rho g_rho
      (extract_LBFGS_SH.c with iter 1, calc_mat_change_test.c);
  iteration 2: the direction is the one-pair BFGS inverse-Hessian
      H = V' (gamma I) V + rho s s', V = I - rho y s', rho = 1 / y's,
      gamma = s'y / y'y, applied to q = [C_rho g_rho; C_vs g_vs], with
      y = q2 - q1 and s = [rho2 - rho1; vs2 - vs1] (store_LBFGS_SH.c,
      LBFGS.c), and scaled back by C (extract_LBFGS_SH.c);
  the line search commits a model whose test-shot misfit is below the
      starting one (line_search.c, calc_opt_step.c);
  PCG: iteration 1 is steepest descent, iteration 2 the Polak-Ribiere
      direction d2 = -g2 + beta d1, beta = g2'(g2 - g1) / g1'g1 per
      parameter class (PCG.c, descent.c).
A wrong sign, a swapped pair (y = q1 - q2), a missing or doubled C factor,
the rho and vs blocks exchanged, or a stale gradient breaks one of them."""
import os
import sys

import numpy as np

from common import ROOT, model, oracle

sys.path.insert(0, os.path.join(ROOT, "oracle"))
from fwi_oracle import FwiOracle, float_mean  # noqa: E402
from pyoracle import make_cfg  # noqa: E402

O = oracle("f32")
NX, NY, NT, DT, DH, FC = 64, 36, 600, 1e-4, 0.25, 40.0


def _problem():
    vs, rho, _ = model(NX, NY, DH, seed=41, vs0=140.0, vs1=260.0, rho0=1600.0, rho1=2000.0)
    j = np.arange(NY)[:, None]
    i = np.arange(NX)[None, :]
    bump = np.exp(-((i - 32.0) ** 2 + (j - 12.0) ** 2) / (2 * 5.0 ** 2))
    vt = (vs * (1.0 + 0.1 * bump)).astype(np.float32)
    rt = (rho * (1.0 - 0.05 * bump)).astype(np.float32)
    rec_i = np.arange(8, NX - 8, 3, dtype=np.int32)
    rec_j = np.full_like(rec_i, 2)
    src_i = np.array([16, 48], np.int32)
    src_j = np.full(2, 3, np.int32)
    t = np.arange(1, NT + 1) * DT
    tau = np.pi * (t - 1.5 / FC) / (1.5 / FC)
    w = ((1 - 4 * tau ** 2) * np.exp(-2 * tau ** 2)).astype(np.float32)
    sig = np.stack([w, w])
    cfg = make_cfg(NX, NY, fdorder=4, fw=10, free_surf=1, L=0, nt=NT, dtinv=2, dt=DT, dh=DH, damping=300.0,
                   fpml=FC)
    obs = np.stack([O.forward(cfg, vt, rt, None, src_i[k:k + 1], src_j[k:k + 1], w, rec_i, rec_j)[0]
                    for k in range(2)]).astype(np.float32)
    return cfg, vs, rho, src_i, src_j, sig, rec_i, rec_j, obs


def _cos(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


def _grad(F, vs, rho):
    g = O.gradient(F.cfg, vs, rho, None, F.src_i, F.src_j, F.sig, F.rec_i, F.rec_j, F.obs)
    return np.asarray(g["grad_vs"], np.float64), np.asarray(g["grad_rho"], np.float64)


def _fwi(grad_method):
    cfg, vs, rho, si, sj, sig, ri, rj, obs = _problem()
    F = FwiOracle(O, cfg, vs, rho, None, si, sj, sig, ri, rj, obs, testshot_start=1, testshot_end=2,
                  testshot_incr=1, vs_lower=1.0, vs_upper=1e4, rho_lower=1.0, rho_upper=1e5, inv_rho_iter=0,
                  grad_method=grad_method)
    models = [(F.vs.copy(), F.rho.copy())]
    recs = []
    for _ in range(3):
        recs.append(F.iteration())
        models.append((F.vs.copy(), F.rho.copy()))
    for r in recs:
        print(r)
    return F, models, recs


def _check_descent_and_misfit(F, models, recs, d_vs, d_rho, k):
    """The update of iteration k + 1 points along -d (one positive scale per
    class) and the misfit fell."""
    (v0, r0), (v1, r1) = models[k], models[k + 1]
    cv, cr = _cos(v1.astype(np.float64) - v0, -d_vs), _cos(r1.astype(np.float64) - r0, -d_rho)
    print("iteration %d: cos(dvs, -d_vs) = %.9f, cos(drho, -d_rho) = %.9f" % (k + 1, cv, cr))
    assert recs[k]["step1"] == 1 and recs[k]["next_stage"] == 0
    assert cv > 1 - 1e-5 and cr > 1 - 1e-5
    assert recs[k + 1]["L2t"][0] < recs[k]["L2t"][0]


def test_lbfgs_iterations_compose_to_bfgs():
    F, models, recs = _fwi(2)
    cvs, crho = np.float32(F.C_vs), np.float32(F.C_rho)
    assert cvs == float_mean(models[0][0]) and crho == float_mean(models[0][1])
    g1v, g1r = _grad(F, *models[0])
    _check_descent_and_misfit(F, models, recs, cvs * g1v, crho * g1r, 0)
    g2v, g2r = _grad(F, *models[1])
    q1 = np.concatenate([crho * g1r.ravel(), cvs * g1v.ravel()])
    q2 = np.concatenate([crho * g2r.ravel(), cvs * g2v.ravel()])
    y = q2 - q1
    s = np.concatenate([(models[1][1].astype(np.float64) - models[0][1]).ravel(),
                        (models[1][0].astype(np.float64) - models[0][0]).ravel()])
    rho_ = 1.0 / (y @ s)
    gam = (s @ y) / (y @ y)
    a = rho_ * (s @ q2)
    z = gam * (q2 - a * y)
    hq = z - rho_ * (y @ z) * s + a * s
    n = NX * NY
    d_rho = crho * hq[:n].reshape(NY, NX)
    d_vs = cvs * hq[n:].reshape(NY, NX)
    # the BFGS direction differs from the gradient's, so the check below is
    # not met by a step along the gradient
    print("cos(BFGS vs direction, gradient) = %.6f" % _cos(d_vs, g2v))
    assert _cos(d_vs, g2v) < 0.999
    _check_descent_and_misfit(F, models, recs, d_vs, d_rho, 1)


def test_pcg_iterations_compose_to_polak_ribiere():
    F, models, recs = _fwi(1)
    g1v, g1r = _grad(F, *models[0])
    _check_descent_and_misfit(F, models, recs, g1v, g1r, 0)
    g2v, g2r = _grad(F, *models[1])
    d = []
    for g1, g2 in ((g1v, g2v), (g1r, g2r)):
        beta = max(0.0, float(g2.ravel() @ (g2 - g1).ravel() / (g1.ravel() @ g1.ravel())))
        print("beta = %.6g" % beta)
        d.append(g2 + beta * g1)     # d2 = -g2 + beta d1 with d1 = -g1, as a gradient (minus d2)
    assert _cos(d[0], g2v) < 0.9999
    _check_descent_and_misfit(F, models, recs, d[0], d[1], 1)
