"""The learned preconditioner end to end on the GPU: training pairs from
finite-difference Hessian probes on randomised models, training of the MFMA
U-Net, and FWI of a held-out model with and without it (the plain
energy-preconditioned L-BFGS judges the learned direction).  Small sizes and
short records; the misfits and wall times are printed, the judge's rules are
asserted on every iteration."""
import importlib.util
import os
import sys

import numpy as np
import pytest

from common import ROOT, product

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a)
    sys.stdout.flush()


def _pt():
    spec = importlib.util.spec_from_file_location("precond_train", os.path.join(ROOT, "denise-sh_amd",
                                                                                 "precond_train.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def drop_of(rec):
    """The relative misfit drop the judge compares, from a record's line
    search: (L2t[1] - min(L2t[2], L2t[3])) / L2t[1] with the difference in
    FP32, 0 when the line search found no decrease."""
    if not rec["step1"]:
        return 0.0
    L = np.asarray(rec["L2t"], np.float32)
    return float(np.float32(np.float64(np.float32(L[0] - min(L[1], L[2]))) / np.float64(L[0])))


def check_judge_rules(recs, judge_every, revert=True):
    """The rules of the learned mode on each record: the learned direction's
    line search runs; the plain one runs on the first iteration of the
    stage, every judge_every-th iteration and whenever the learned search
    found no decrease; when both ran, the learned direction is taken only
    with a decrease at least as large as the plain one's.  With revert, once
    the plain direction has been taken the rest of the stage is plain
    L-BFGS: no learned search and no judge."""
    off = False
    for r in recs:
        it = r["iter"]
        if off:
            assert r["judged"] == 0 and r["precond"] == 0, it
            assert r["drop_learned"] == -1.0 and r["drop_plain"] == -1.0, it
        else:
            due = it == 1 or (judge_every > 0 and (it - 1) % judge_every == 0)
            assert r["drop_learned"] >= 0.0
            if due or r["drop_learned"] == 0.0:
                assert r["judged"] == 1, it
                assert r["drop_plain"] >= 0.0
                assert r["precond"] == int(r["drop_learned"] > 0.0 and r["drop_learned"] >= r["drop_plain"]), it
                off = revert and r["precond"] == 0
            else:
                assert r["judged"] == 0 and r["precond"] == 1 and r["drop_plain"] == -1.0, it
        if r["next_stage"]:
            off = False


def check_against_plain(learned, plain):
    """While the judge keeps the plain direction the learned run follows the
    plain run exactly: the same model at the start of each iteration (the
    same misfit L2t[1]), the judge's plain line search is the plain method's
    own (the same drop, bit for bit), and an iteration that takes the plain
    direction records the plain run's line search."""
    for k in range(min(len(learned), len(plain))):
        assert learned[k]["L2t"][0] == plain[k]["L2t"][0], k + 1
        if learned[k]["judged"]:
            assert learned[k]["drop_plain"] == drop_of(plain[k]), k + 1
        if learned[k]["precond"] == 1:
            break
        assert learned[k]["L2t"] == plain[k]["L2t"], k + 1


def test_learned_preconditioner_trains_and_is_judged():
    dn = product()
    pt = _pt()
    nx, ny = 128, 48
    geom = dict(nshots=6, nt=700, fc=40.0)
    say("training pairs 16 x (1 + 4 probes) gradients at %dx%d" % (nx, ny))
    X, T, sec = pt.make_pairs(dn, 16, nx, ny, nprobes=4, **geom)
    say("  %.1f s" % sec)
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(T))
    net = dn.UNet(seed=1)
    hist = pt.train(net, X[:12], T[:12], steps=150, lr=2e-3, batch=6)
    val0 = float(np.mean((0.0 * T[12:] + np.mean(T[:12]) - T[12:]) ** 2))
    val = net.loss_grad(X[12:], T[12:])[0]
    say("  training loss %.4f -> %.4f; held-out %.4f (constant predictor %.4f)" % (hist[0], hist[-1], val, val0))
    assert hist[-1] < hist[0]
    # the network learned the structure of the Hessian diagonal
    assert val < 0.5 * val0
    vt, v0, rho, qs = pt.random_model(nx, ny, 0.25, 999)
    out = {}
    for mode in ("plain", "learned"):
        sv = pt.Survey(dn, nx, ny, **geom)
        recs, vs = pt.run_fwi(sv, vt, v0, rho, qs, 6, net=net if mode == "learned" else None, judge_every=3)
        out[mode] = recs
        sv.ctx.close()
        say("  %-7s misfit %s  precond %s  %.2f s" % (mode, " ".join("%.4f" % r["L2t"][3] for r in recs),
                                                     "".join(str(r["precond"]) for r in recs), recs[-1]["seconds"]))
        assert recs[-1]["L2t"][3] < recs[0]["L2t"][3]
    # wall time of each run to the plain run's final misfit: printed, the
    # measured result of the method at this size (no bar)
    target = out["plain"][-1]["L2t"][3]
    def time_to(recs):
        for r in recs:
            if r["L2t"][3] <= target:
                return r["seconds"]
        return float("inf")
    tp, tl = time_to(out["plain"]), time_to(out["learned"])
    say("  time to the plain final misfit %.4f: plain %.2f s, learned %.2f s (ratio %.2f)" % (target, tp, tl, tl / tp))
    for r in out["plain"]:
        assert r["judged"] == 0 and r["precond"] == 0 and r["drop_learned"] == -1.0
    check_judge_rules(out["learned"], 3)
    check_against_plain(out["learned"], out["plain"])


def test_network_from_1200x300_judged_on_held_out_4000x800():
    """A network trained on 1200 x 300 models is applied to a held-out
    4000 x 800 model of the same family (the shapes of the learned-
    preconditioner workload, with short records and few shots so that the
    test takes seconds): the judge runs on every iteration and its rules hold
    on each record."""
    dn = product()
    pt = _pt()
    geom = dict(dh=0.25, nshots=4, nt=1000, fc=40.0, dtinv=4)
    lab = pt.Survey(dn, 1200, 300, eprecond=0, **geom)
    models = [pt.random_checkerboard(1200, 300, 0.25, 500 + k) for k in range(8)]
    X, T = pt.pairs_on_context(lab.ctx, models, nprobes=4, seed0=500)
    lab.ctx.close()
    assert np.all(np.isfinite(X)) and np.all(np.isfinite(T))
    net = dn.UNet(seed=2)
    hist = pt.train(net, X[:6], T[:6], steps=120, lr=2e-3, batch=3)
    val0 = float(np.mean((np.mean(T[:6]) - T[6:]) ** 2))
    val = net.loss_grad(X[6:], T[6:])[0]
    say("1200x300: training loss %.4f -> %.4f; held-out %.4f (constant predictor %.4f)" % (hist[0], hist[-1], val,
                                                                                          val0))
    assert hist[-1] < hist[0]
    vt, v0, rho, qs = pt.random_checkerboard(4000, 800, 0.25, 4242)
    out = {}
    for mode in ("plain", "learned", "learned_no_revert"):
        sv = pt.Survey(dn, 4000, 800, **geom)
        recs, vs = pt.run_fwi(sv, vt, v0, rho, qs, 3, net=None if mode == "plain" else net, judge_every=1,
                              revert=mode == "learned")
        sv.ctx.close()
        out[mode] = recs
        say("4000x800 %-17s misfit %s  precond %s  drops %s  %.2f s" % (
            mode, " ".join("%.5g" % r["L2t"][3] for r in recs), "".join(str(r["precond"]) for r in recs),
            " ".join("%.4g/%.4g" % (r["drop_learned"], r["drop_plain"]) for r in recs), recs[-1]["seconds"]))
        assert all(np.isfinite(r["L2t"][3]) for r in recs)
    check_judge_rules(out["learned"], 1, revert=True)
    check_judge_rules(out["learned_no_revert"], 1, revert=False)
    for m in ("learned", "learned_no_revert"):
        check_against_plain(out[m], out["plain"])
