"""Gradient post-processing on the GPU against the oracle (denise.c:1818-1847):
gauss_filt_var (SPATFILTER 4), gauss_filt (3), smooth2 (2) and taper_grad
(sws 1, 2, 3, 5), bit-identical
to the oracle's float restatement, on host arrays and on the context's
gradient, plus an FWI iteration with both switched on."""
import sys

import numpy as np
import pytest

from common import model, oracle, product, rel_l2

pytestmark = pytest.mark.gpu
O32 = oracle("f32")


def say(*a):
    print(*a)
    sys.stdout.flush()


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def vel_of(kind, ny, nx, seed):
    rng = np.random.default_rng(seed)
    if kind == "layers":
        v = np.full((ny, nx), 400.0, np.float32)
        v[ny // 2:] = 100.0
        return v
    if kind == "zeros":
        v = (150.0 + 200.0 * rng.random((ny, nx))).astype(np.float32)
        v[:, : nx // 3] = 0.0          # FILT_SIZE_Z = 0 there: copied
        return v
    v, _, _ = model(nx, ny, 0.25, seed=seed, vs0=100.0, vs1=350.0)
    return v


# (name, ny, nx, dh, WD_DAMP, WD_DAMP1, FC_END, velocity)
FILTER_CASES = [
    ("two layers, ragged tiles", 61, 131, 0.25, 0.25, 0.15, 20.0, "layers"),
    ("smooth model, many widths", 120, 300, 0.2, 0.25, 0.25, 25.0, "smooth"),
    ("half widths above the LDS window", 40, 90, 0.25, 0.8, 0.5, 10.0, "smooth"),
    ("zero velocity copies", 33, 70, 0.25, 0.3, 0.2, 30.0, "zeros"),
    ("WD_DAMP 0 copies", 20, 65, 0.25, 0.0, 0.2, 30.0, "smooth"),
]


@pytest.mark.parametrize("case", FILTER_CASES, ids=[c[0] for c in FILTER_CASES])
def test_gauss_filt_var_bit_identical(case):
    dn = product()
    name, ny, nx, dh, wd, wd1, fc, kind = case
    say("case", name, ny, nx)
    vel = vel_of(kind, ny, nx, 5)
    g = np.random.default_rng(6).standard_normal((ny, nx)).astype(np.float32)
    got = dn.gauss_filt_var(g, vel, dh, wd, wd1, fc)
    want = O32.gauss_filt_var(g, vel, dh, wd, wd1, fc)
    assert np.array_equal(bits(got), bits(want))
    # error against the FP64 oracle, for the record
    w64 = oracle("f64").gauss_filt_var(g.astype(np.float64), vel, dh, wd, wd1, fc)
    say("  rel L2 vs FP64 oracle %.2e" % rel_l2(got, w64))


TAPER_CASES = [
    ("vertical", dict(sws=1, gradt=(5, 15, 60, 70))),
    ("horizontal", dict(sws=2, gradt=(3, 12, 0, 0))),
    ("sources erf", dict(sws=3, srtshape=1, srtradius=2.0, filtsize=1)),
    ("sources log", dict(sws=3, srtshape=2, srtradius=1.5, filtsize=2)),
]


@pytest.mark.parametrize("case", TAPER_CASES, ids=[c[0] for c in TAPER_CASES])
def test_taper_grad_bit_identical(case):
    dn = product()
    name, kw = case
    ny, nx, dh = 40, 80, 0.25
    say("case", name)
    g = np.random.default_rng(7).standard_normal((ny, nx)).astype(np.float32)
    src = dict(src_i=[10, 40, 79], src_j=[3, 3, 20], rec_i=[20, 30, 50], rec_j=[2, 2, 2])
    got, tap = dn.taper_grad(g, dh=dh, **kw, **src)
    want, wtap = O32.taper_grad(g, dh=dh, **kw, **src)
    assert np.array_equal(bits(got), bits(want))
    assert np.array_equal(bits(tap), bits(wtap))


def test_taper_from_map():
    dn = product()
    ny, nx = 25, 33
    g = np.random.default_rng(8).standard_normal((ny, nx)).astype(np.float32)
    m = np.random.default_rng(9).random((ny, nx)).astype(np.float32)
    got, tap = dn.taper_grad(g, 5, 0.25, taper_map=m)
    assert np.array_equal(bits(got), bits(g * m))
    assert np.array_equal(tap, m)


def _ctx_case(dn, nx=160, ny=60, nshots=3, nt=500):
    vs, rho, _ = model(nx, ny, 0.25, seed=21, vs0=120.0, vs1=300.0)
    ctx = dn.Context(nx, ny, fdorder=4, fw=12, free_surf=1, L=0, dt=1e-4, dh=0.25, damping=300.0, fpml=40.0)
    rec_i = np.arange(14, nx - 14, 3, dtype=np.int32)
    ctx.set_receivers(rec_i, np.full_like(rec_i, 2))
    src_i = np.linspace(20, nx - 20, nshots).astype(np.int32)
    sig = np.tile(dn.wavelet(1, nt, 1e-4, 0.0, 40.0, 1.0), (nshots, 1)).astype(np.float32)
    ctx.set_shots(src_i, np.full(nshots, 3, np.int32), sig)
    vt = vs * np.float32(1.05)
    ctx.set_model(vs, rho, None)
    ctx.set_model(vt, rho, None)
    obs = ctx.forward()
    ctx.set_observed(obs)
    ctx.set_model(vs, rho, None)
    return ctx, vs, rho, src_i, rec_i


def test_context_smooth_and_taper_match_oracle():
    dn = product()
    ctx, vs, rho, src_i, rec_i = _ctx_case(dn)
    ctx.gradient()
    gvs, grho, _, _ = ctx.get_gradient()
    ms = ctx.smooth_grad(4, 0.25, 0.25, 40.0)
    ctx.taper_grad(3, 1, gradt=(3, 10, 140, 150))
    svs, srho, _, _ = ctx.get_gradient()
    for g, s in ((gvs, svs), (grho, srho)):
        w = O32.gauss_filt_var(g, vs, 0.25, 0.25, 0.25, 40.0)
        w, _ = O32.taper_grad(w, 1, 0.25, gradt=(3, 10, 140, 150), src_i=src_i, src_j=np.full_like(src_i, 3),
                              rec_i=rec_i, rec_j=np.full_like(rec_i, 2))
        assert np.array_equal(bits(s), bits(w))
    say("  smoothing of both gradients on the device: %.3f ms" % ms)
    ctx.close()


# (name, ny, nx, dh, WD_DAMP, FC_END, C_vs): FILT_SIZE_GRAD 7, 11, 41 and 101
# (the last beyond the LDS window of a 64 x 4 tile), then 0 (nothing done)
FIXED_CASES = [
    ("size 7, ragged tiles", 61, 131, 0.25, 0.35, 40.0, 200.0),
    ("size 11", 70, 150, 0.2, 0.55, 50.0, 200.0),
    ("size 41", 50, 90, 0.25, 1.0, 20.0, 200.0),
    ("size 101, unstaged", 50, 90, 0.25, 5.05, 40.0, 200.0),
    ("size 0 leaves it", 20, 30, 0.25, 0.0, 40.0, 200.0),
]


@pytest.mark.parametrize("case", FIXED_CASES, ids=[c[0] for c in FIXED_CASES])
def test_gauss_filt_bit_identical(case):
    """gauss_filt (SPATFILTER 3) on a host array, bit-identical to the
    oracle's float restatement."""
    dn = product()
    name, ny, nx, dh, wd, fc, cvs = case
    say("case gauss_filt", name)
    g = np.random.default_rng(8).standard_normal((ny, nx)).astype(np.float32)
    got = dn.gauss_filt(g, dh, wd, fc, cvs)
    want, rc = O32.gauss_filt(g, dh, wd, fc, cvs)
    assert rc in (0, 1)
    assert np.array_equal(bits(got), bits(want))
    if rc == 1:
        assert np.array_equal(got, g)
    w64, _ = oracle("f64").gauss_filt(g.astype(np.float64), dh, wd, fc, cvs)
    say("  rel L2 vs FP64 oracle %.2e" % rel_l2(got, w64))


def test_gauss_filt_rejects_what_the_reference_cannot_filter():
    dn = product()
    g = np.ones((10, 12), np.float32)
    for wd in (0.15, -0.25):          # FILT_SIZE_GRAD 3 (sigma 0) and a negative size
        with pytest.raises(dn.DeniseError):
            dn.gauss_filt(g, 0.25, wd, 40.0, 200.0)


# (name, ny, nx, WD_DAMP)
SMOOTH2_CASES = [
    ("61 x 131", 61, 131, 3.0),
    ("smallest grid", 3, 2, 1.5),
    ("300 x 200, strong", 200, 300, 5.0),
    ("40 x 57, weak", 40, 57, 0.7),
]


@pytest.mark.parametrize("case", SMOOTH2_CASES, ids=[c[0] for c in SMOOTH2_CASES])
def test_smooth2_bit_identical(case):
    """smooth2 (SPATFILTER 2) on a host array, bit-identical to the oracle's
    float restatement (the reference's index ranges: the last row and column
    are kept)."""
    dn = product()
    name, ny, nx, wd = case
    say("case smooth2", name)
    g = np.random.default_rng(9).standard_normal((ny, nx)).astype(np.float32)
    got = dn.smooth2(g, wd)
    want = O32.smooth2(g, wd)
    assert np.array_equal(bits(got), bits(want))
    assert np.array_equal(got[-1], g[-1]) and np.array_equal(got[:, -1], g[:, -1])


def test_smooth2_rejects_grids_below_three_rows():
    dn = product()
    with pytest.raises(dn.DeniseError):
        dn.smooth2(np.ones((2, 10), np.float32), 1.0)


def test_context_smooth_grad_spatfilter_2_matches_oracle():
    dn = product()
    ctx, vs, rho, src_i, rec_i = _ctx_case(dn)
    ctx.gradient()
    gvs, grho, _, _ = ctx.get_gradient()
    ctx.smooth_grad(2, 2.5, 0.0, 40.0)
    svs, srho, _, _ = ctx.get_gradient()
    for g, s in ((gvs, svs), (grho, srho)):
        assert np.array_equal(bits(s), bits(O32.smooth2(g, 2.5)))
    ctx.close()


def test_context_smooth_grad_spatfilter_3_matches_oracle():
    """SPATFILTER 3 on the context's gradients uses the context's C_vs (the
    mean of the first gradient's model)."""
    dn = product()
    ctx, vs, rho, src_i, rec_i = _ctx_case(dn)
    assert ctx.means() is None
    ctx.gradient()
    cvs, _ = ctx.means()
    gvs, grho, _, _ = ctx.get_gradient()
    ctx.smooth_grad(3, 0.35, 0.0, 40.0)
    svs, srho, _, _ = ctx.get_gradient()
    for g, s in ((gvs, svs), (grho, srho)):
        w, rc = O32.gauss_filt(g, 0.25, 0.35, 40.0, cvs)
        assert rc == 0
        assert np.array_equal(bits(s), bits(w))
    say("  C_vs %.4f" % cvs)
    ctx.close()


def test_fwi_iteration_with_smoothing_and_taper():
    """The FWI loop applies SPATFILTER 4 and the vertical taper to the summed
    gradient before the optimiser: the first model update is a multiple of
    the oracle's smoothed and tapered gradient (the first L-BFGS direction is
    the scaled gradient), and the misfit falls over three iterations."""
    dn = product()
    ctx, vs, rho, src_i, rec_i = _ctx_case(dn)
    ctx.gradient()
    gvs, _, _, _ = ctx.get_gradient()
    w = O32.gauss_filt_var(gvs, vs, 0.25, 0.25, 0.25, 40.0)
    w, _ = O32.taper_grad(w, 1, 0.25, gradt=(3, 10, 140, 150), src_i=src_i, src_j=np.full_like(src_i, 3),
                          rec_i=rec_i, rec_j=np.full_like(rec_i, 2))
    ctx.fwi_start(testshot_start=1, testshot_end=3, testshot_incr=1, vs_lower=50.0, vs_upper=800.0,
                  rho_lower=500.0, rho_upper=4000.0, inv_rho_iter=1000, spatfilter=4, wd_damp=0.25,
                  wd_damp1=0.25, fc_end=40.0, taper_vert=1, gradt=(3, 10, 140, 150))
    recs = [ctx.fwi_iteration()]
    v1, _ = ctx.fwi_model()
    recs += [ctx.fwi_iteration() for _ in range(2)]
    ctx.close()
    say("  misfit", " ".join("%.4g" % r["L2t"][0] for r in recs), " eps_u %.3g" % recs[0]["eps_u"])
    assert recs[0]["eps_u"] > 0.0
    assert recs[-1]["L2t"][0] < recs[0]["L2t"][0]
    dv = vs.astype(np.float64) - v1.astype(np.float64)
    unit = lambda a: np.asarray(a, np.float64) / np.linalg.norm(a)
    # float rounding of the model update (vs ~ 1e2, steps ~ 1) bounds the match
    assert rel_l2(unit(dv), unit(w)) < 1e-3
    # without the post-processing the update would point elsewhere
    assert rel_l2(unit(dv), unit(gvs)) > 1e-1


def test_fwi_files_have_the_reference_names_and_formats(tmp_path):
    dn = product()
    ctx, vs, rho, _, _ = _ctx_case(dn)
    ctx.fwi_start(testshot_start=1, testshot_end=3, testshot_incr=1, vs_lower=50.0, vs_upper=800.0,
                  rho_lower=500.0, rho_upper=4000.0, pro=1e9)
    log = str(tmp_path / "misfit.dat")
    lines = []
    for it in range(2):
        r = ctx.fwi_iteration()
        ctx.fwi_write_files(r, misfit_log=log, time_filt=1, fc_start=0.0, fc=40.0, jacobian=str(tmp_path / "jac"),
                            inv_modelfile=str(tmp_path / "model"), write_it=1)
        e, L = r["epst1"], r["L2t"]
        lines.append("%e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %e \t %f \t %f \t %d \n" % (
            r["opteps_vs"], e[0], e[1], e[2], L[0], L[1], L[2], L[3], r["M2_vs"], r["M2_rho"], 0.0, 40.0, r["stage"]))
    assert open(log).read() == "".join(lines)
    vs1, rho1 = ctx.fwi_model()
    ny, nx = vs1.shape
    back = np.fromfile(str(tmp_path / "model_vs.bin"), np.float32).reshape(nx, ny).T   # x-major
    assert np.array_equal(back, vs1)
    it_file = tmp_path / ("model_vs_it_%d.bin" % r["iter"])
    assert np.array_equal(np.fromfile(str(it_file), np.float32).reshape(nx, ny).T, vs1)
    gd = np.fromfile(str(tmp_path / "jac_grad_vs.bin"), np.float32).reshape(nx, ny).T
    assert np.abs(gd).max() > 0
    if r["next_stage"]:
        assert (tmp_path / ("model_vs_stage_%d.bin" % r["stage"])).exists()
    ctx.close()
