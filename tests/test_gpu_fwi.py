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


def _small_problem(dn, O, L=2):
    nx, ny, nt, dt, dh, fc = 100, 40, 900, 1e-4, 0.25, 40.0
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
    ctx = dn.Context(nx, ny, fdorder=4, fw=12, free_surf=1, L=L, fl=fl, dt=dt, dh=dh, damping=300.0, fpml=fc)
    ctx.set_model(vs, rho, qs)
    ctx.set_receivers(rec_i, rec_j)
    ctx.set_shots(src_i, src_j, sig)
    ctx.set_observed(obs)
    return dict(vs=vs, rho=rho, qs=qs, rec_i=rec_i, rec_j=rec_j, src_i=src_i, src_j=src_j, sig=sig, cfg=cfg,
                obs=obs, ctx=ctx)


@pytest.mark.parametrize("time_filt", [1, 2])
def test_time_filtered_gradient_is_bit_identical(time_filt):
    """dn_set_time_filt filters the signals and observed data as
    timedomain_filt does (low-pass FC, then high-pass FC_START): the gradient
    equals the oracle's on the oracle-filtered inputs bit for bit, and
    TIME_FILT 0 afterwards restores the unfiltered gradient."""
    dn = product()
    O = oracle("f32")
    P = _small_problem(dn, O)
    ctx, cfg = P["ctx"], P["cfg"]
    say("time-filtered gradient, TIME_FILT", time_filt)
    fc, fcs, order = 25.0, 8.0, 6
    sig = O.timedomain_filt(P["sig"], cfg.dt, fc, order, 1)
    obs = np.stack([O.timedomain_filt(o, cfg.dt, fc, order, 1) for o in P["obs"]])
    if time_filt == 2:
        sig = O.timedomain_filt(sig, cfg.dt, fcs, order, 2)
        obs = np.stack([O.timedomain_filt(o, cfg.dt, fcs, order, 2) for o in obs])
    for flt, s, ob in ((time_filt, sig, obs), (0, P["sig"], P["obs"])):
        ctx.set_time_filt(flt, fcs, fc, order)
        l2, en = ctx.gradient()
        gv, gr = ctx.get_gradient()[:2]
        ref = O.gradient(cfg, P["vs"], P["rho"], P["qs"], P["src_i"], P["src_j"], s, P["rec_i"], P["rec_j"],
                         ob.astype(np.float32))
        say("  TIME_FILT %d: l2 %.6e (oracle %.6e)" % (flt, l2, ref["l2"]))
        assert np.array_equal(gv, ref["grad_vs"]) and np.array_equal(gr, ref["grad_rho"])
        assert abs(l2 - ref["l2"]) <= 1e-12 * abs(ref["l2"])
    ctx.close()


WORKFLOW_3 = """PRO\tTIME_FILT\tFC_low\tFC_high\tORDER\tTIME_WIN\tGAMMA\tGAMMA-\tTWIN+\tINV_VP_ITER\tINV_VS_ITER\tINV_RHO_ITER\tSPATFILTER\tWD_DAMPx\tWD_DAMPy\tEPRECOND\tLNORM\tSTF_INV\tOFFSETC_STF\tEPS_STF\tNORMALIZE\tOFFSET_MUTE\tOFFSETC\tSCALERHO
0.01\t1\t0.0\t15.00\t6\t0\t8\t0.0\t0.0\t0\t0\t0\t0\t0.25\t0.25\t1\t2\t0\t100.0\t5e-2\t0\t0\t30.0\t0.5
0.01\t2\t6.0\t25.00\t4\t0\t8\t0.0\t0.0\t0\t0\t0\t0\t0.25\t0.25\t1\t2\t0\t100.0\t5e-2\t0\t0\t30.0\t0.5
0.01\t1\t0.0\t40.00\t6\t0\t8\t0.0\t0.0\t0\t0\t0\t0\t0.25\t0.25\t1\t2\t0\t100.0\t5e-2\t0\t0\t30.0\t0.5
"""


def test_three_stage_workflow_matches_oracle(tmp_path):
    """A three-stage workflow in the reference's table format (low-pass 15
    Hz order 6, band-pass 6-25 Hz order 4, low-pass 40 Hz order 6; the
    canonical workflow's layout), each stage started by dn_fwi_start with the
    values the product's workflow reader returns, two iterations per stage,
    against FwiOracle.start_stage."""
    import ctypes as C
    dn = product()
    O = oracle("f32")
    wf = tmp_path / "workflow.inp"
    wf.write_text(WORKFLOW_3)
    L = dn.lib()
    L.dn_workflow_read.restype = C.c_void_p
    L.dn_workflow_read.argtypes = [C.c_char_p]
    L.dn_workflow_get.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.POINTER(C.c_float)]
    L.dn_workflow_nstages.argtypes = [C.c_void_p]
    L.dn_workflow_free.argtypes = [C.c_void_p]
    w = L.dn_workflow_read(str(wf).encode())
    assert w and L.dn_workflow_nstages(w) == 3

    def col(st, name):
        v = C.c_float()
        assert L.dn_workflow_get(w, st, name.encode(), C.byref(v)) == 0
        return v.value

    P = _small_problem(dn, O)
    ctx = P["ctx"]
    base = dict(testshot_start=1, testshot_end=3, testshot_incr=2, vs_lower=50.0, vs_upper=800.0,
                rho_lower=500.0, rho_upper=4000.0, eps_scale=0.01, stepmax=4, inv_rho_iter=0)
    ref = None
    for st in (1, 2, 3):
        stage = dict(pro=col(st, "PRO"), time_filt=int(col(st, "TIME_FILT")), fc_start=col(st, "FC_START"),
                     fc_end=col(st, "FC_END"), order=int(col(st, "ORDER")), scalerho=col(st, "SCALERHO"))
        say("stage %d: %s" % (st, stage))
        ctx.fwi_start(**dict(base, **stage))
        if ref is None:
            ref = FwiOracle(O, P["cfg"], P["vs"], P["rho"], P["qs"], P["src_i"], P["src_j"], P["sig"], P["rec_i"],
                            P["rec_j"], P["obs"], **base)
        ts = {k: stage.pop(k) for k in ("time_filt", "fc_start", "order")}
        ref.start_stage(**ts, **stage)
        v0 = ref.vs.astype(np.float64).copy()
        r0 = ref.rho.astype(np.float64).copy()
        for it in range(2):
            r = ctx.fwi_iteration()
            wr = ref.iteration()
            gv, gr = ctx.fwi_model()
            say("  iter %d  L2t %s | %s  step1 %d/%d next %d/%d" % (
                it + 1, " ".join("%.6g" % v for v in r["L2t"]), " ".join("%.6g" % v for v in wr["L2t"]),
                r["step1"], wr["step1"], r["next_stage"], wr["next_stage"]))
            assert r["stage"] == st
            assert (r["step1"], r["step3"], r["next_stage"]) == (wr["step1"], wr["step3"], wr["next_stage"])
            for a, b in zip(r["L2t"], wr["L2t"]):
                assert abs(a - b) <= 1e-4 * abs(b) + 1e-12
            dv = rel_l2(gv.astype(np.float64) - v0, ref.vs.astype(np.float64) - v0)
            assert dv < 1e-4
            if st == 1 and it == 0:
                assert dv == 0.0 and rel_l2(gr.astype(np.float64) - r0, ref.rho.astype(np.float64) - r0) == 0.0
            if r["next_stage"]:
                break
    L.dn_workflow_free(w)
    ctx.close()


@pytest.mark.parametrize("grad_method", [2, 1])
def test_fwi_resumes_bit_identically_from_state_files(tmp_path, grad_method):
    """Four iterations in one run equal two iterations, dn_fwi_save_state,
    and two more in a new context after dn_fwi_load_state, bit for bit
    (records and models).  L-BFGS also writes the reference-named files of
    extract_LBFGS_SH.c and resumes with the previous model and gradient read
    from them (name fix on); with the reference's own read names (.old.00)
    the load fails as store_LBFGS_SH.c would."""
    dn = product()
    O = oracle("f32")
    opts = dict(testshot_start=1, testshot_end=3, testshot_incr=2, vs_lower=50.0, vs_upper=800.0,
                rho_lower=500.0, rho_upper=4000.0, eps_scale=0.01, stepmax=4, inv_rho_iter=0, grad_method=grad_method)
    say("resume, GRAD_METHOD", grad_method)
    P = _small_problem(dn, O)
    a = P["ctx"]
    a.fwi_start(**opts)
    full = [a.fwi_iteration() for _ in range(4)]
    model_full = a.fwi_model()
    a.close()

    jac = str(tmp_path / "jacobian")
    state = str(tmp_path / "fwi_state.bin")
    P = _small_problem(dn, O)
    b = P["ctx"]
    b.fwi_start(**opts)
    models = []
    for k in range(2):
        r = b.fwi_iteration()
        assert r == full[k]
        b.fwi_write_files(r, jacobian=jac, lbfgs_old=1)
        models.append(b.fwi_model())
    b.fwi_save_state(state)
    b.close()
    if grad_method == 2:
        # _p_vs.old: the model iteration 2 started from, x-major (writemod.c)
        pv = np.fromfile(jac + "_p_vs.old", np.float32).reshape(P["vs"].shape[1], P["vs"].shape[0]).T
        assert np.array_equal(pv, models[0][0])
        assert np.array_equal(np.fromfile(jac + "_p_vs.old.0.0", np.float32), np.fromfile(jac + "_p_vs.old", np.float32))

    for jacobian in ((None, jac) if grad_method == 2 else (None,)):
        P = _small_problem(dn, O)
        c = P["ctx"]
        c.fwi_start(**opts)
        c.fwi_load_state(state, jacobian=jacobian, name_fix=1)
        for k in range(2, 4):
            r = c.fwi_iteration()
            say("  resumed iter %d (jacobian files %s): L2t %s" % (k + 1, jacobian is not None,
                                                                   " ".join("%.6g" % v for v in r["L2t"])))
            assert r == full[k]
        vs, rho = c.fwi_model()
        assert np.array_equal(vs, model_full[0]) and np.array_equal(rho, model_full[1])
        c.close()
    if grad_method == 2:
        P = _small_problem(dn, O)
        c = P["ctx"]
        c.fwi_start(**opts)
        with pytest.raises(Exception) as e:
            c.fwi_load_state(state, jacobian=jac, name_fix=0)
        assert ".old.00" in str(e.value)
        c.close()


@pytest.mark.parametrize("revert", [True, False], ids=["revert", "no_revert"])
def test_learned_fwi_resumes_bit_identically(tmp_path, revert):
    """The learned preconditioner's state goes into the state file too: the
    stage's fixed diagonal D and whether the stage reverted to plain L-BFGS.
    Four learned iterations in one run equal two, dn_fwi_save_state, and two
    more in a new context after dn_fwi_load_state, bit for bit (an untrained
    network of a fixed seed: the test is about the state, not the gain)."""
    dn = product()
    O = oracle("f32")
    net = dn.UNet(seed=5)
    opts = dict(testshot_start=1, testshot_end=3, testshot_incr=2, vs_lower=50.0, vs_upper=800.0,
                rho_lower=500.0, rho_upper=4000.0, eps_scale=0.01, stepmax=4, inv_rho_iter=0, grad_method=2)
    say("learned resume, revert", revert)

    def start():
        P = _small_problem(dn, O)
        ctx = P["ctx"]
        ctx.set_learned(net, 0, fixed=True, revert=revert)
        ctx.fwi_start(**opts)
        return ctx

    a = start()
    full = [a.fwi_iteration() for _ in range(4)]
    model_full = a.fwi_model()
    a.close()
    say("  directions %s, judged %s" % ("".join(str(r["precond"]) for r in full),
                                       "".join(str(r["judged"]) for r in full)))
    assert full[0]["judged"] == 1
    state = str(tmp_path / "fwi_state.bin")
    b = start()
    for k in range(2):
        assert b.fwi_iteration() == full[k]
    b.fwi_save_state(state)
    b.close()
    c = start()
    c.fwi_load_state(state)
    for k in range(2, 4):
        assert c.fwi_iteration() == full[k], k + 1
    vs, rho = c.fwi_model()
    assert np.array_equal(vs, model_full[0]) and np.array_equal(rho, model_full[1])
    c.close()
