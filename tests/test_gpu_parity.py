"""GPU parity of the fused HIP time step against the CPU oracle.

Seismograms and gradients from the HIP path are compared with the oracle's
FP32 restatement of the reference (bar: seismograms 1e-5, gradients 1e-4
relative L2; the kernels keep the reference's operation order, so the
expected difference is zero or a few ulps) and the error against the FP64
oracle is printed for every case.
"""
import sys

import numpy as np
import pytest

from common import model, oracle, product, rel_l2, ricker

pytestmark = pytest.mark.gpu

SEIS_TOL = 1e-5
GRAD_TOL = 1e-4


def say(*a):
    print(*a)
    sys.stdout.flush()


# name, nx, ny, fdorder, L, free_surf, fw, nt, fc
FWD_CASES = [
    ("el_fd4_fs", 120, 50, 4, 0, 1, 10, 500, 60.0),
    ("el_fd2_nofs", 90, 40, 2, 0, 0, 8, 300, 40.0),
    ("el_fd6_fs", 100, 45, 6, 0, 1, 10, 300, 60.0),
    ("el_fd8_nofs", 100, 45, 8, 0, 0, 10, 300, 60.0),
    ("el_fd10_fs", 100, 45, 10, 0, 1, 10, 250, 60.0),
    ("el_fd12_nofs", 100, 45, 12, 0, 0, 10, 250, 60.0),
    ("visc_fd8_nofs", 120, 50, 8, 2, 0, 10, 400, 60.0),
    ("visc_fd4_fs", 110, 48, 4, 1, 1, 10, 300, 60.0),
    ("visc_fd12_fs", 100, 45, 12, 4, 1, 10, 250, 60.0),
    ("visc_fd6_nofs_L3", 100, 45, 6, 3, 0, 10, 250, 60.0),
]

DT = 1e-4
DH = 0.25


def setup_case(nx, ny, fdorder, L, free_surf, fw, nt, fc, nshots=2, dtinv=1, snap_precision=0, seed=3,
               nts=None, eprecond=1):
    qs = 20.0 if L > 0 else None
    vs, rho, q = model(nx, ny, DH, seed=seed, qs=qs)
    fl = [6.25, 50.0, 12.0, 30.0][:L]
    dn = product()
    ctx = dn.Context(nx, ny, fdorder=fdorder, maxrelerror=1, fw=fw, free_surf=free_surf, L=L, fl=fl, dt=DT, dh=DH,
                     damping=300.0, fpml=fc, dtinv=dtinv, snap_precision=snap_precision, eprecond=eprecond)
    ctx.set_model(vs, rho, q)
    rec_i = np.arange(fw + 3, nx - fw - 2, 3, dtype=np.int32)
    rec_j = np.full_like(rec_i, 2)
    ctx.set_receivers(rec_i, rec_j)
    src_i = np.linspace(fw + 5, nx - fw - 5, nshots).astype(np.int32)
    src_j = np.full(nshots, 3, np.int32)
    sig = np.stack([ricker(nt, DT, fc, tshift=0.001 * s) for s in range(nshots)])
    ntv = None if nts is None else np.asarray(nts, np.int32)
    ctx.set_shots(src_i, src_j, sig, ntv)
    from pyoracle import make_cfg
    cfg = make_cfg(nx, ny, fdorder=fdorder, maxrelerror=1, fw=fw, free_surf=free_surf, L=L, nt=nt, dtinv=dtinv,
                   dt=DT, dh=DH, damping=300.0, fpml=fc, fl=fl, eprecond=eprecond)
    return ctx, cfg, dict(vs=vs, rho=rho, qs=q, rec_i=rec_i, rec_j=rec_j, src_i=src_i, src_j=src_j, sig=sig)


@pytest.mark.parametrize("case", FWD_CASES, ids=[c[0] for c in FWD_CASES])
def test_forward_seismograms(case):
    name, nx, ny, fdorder, L, fs, fw, nt, fc = case
    say("case", name)
    ctx, cfg, d = setup_case(nx, ny, fdorder, L, fs, fw, nt, fc)
    seis = ctx.forward()
    o32, o64 = oracle("f32"), oracle("f64")
    worst = 0.0
    for s in range(len(d["src_i"])):
        ref32, _, _ = o32.forward(cfg, d["vs"], d["rho"], d["qs"], d["src_i"][s:s + 1], d["src_j"][s:s + 1],
                                  d["sig"][s], d["rec_i"], d["rec_j"])
        ref64, _, _ = o64.forward(cfg, d["vs"], d["rho"], d["qs"], d["src_i"][s:s + 1], d["src_j"][s:s + 1],
                                  d["sig"][s], d["rec_i"], d["rec_j"])
        e32 = rel_l2(seis[s], ref32)
        e64 = rel_l2(seis[s], ref64)
        o_err = rel_l2(ref32, ref64)
        say("  shot %d: rel L2 vs oracle f32 %.3e (max abs %.3e), vs f64 %.3e (oracle f32 vs f64 %.3e)" %
            (s, e32, np.max(np.abs(seis[s] - ref32)), e64, o_err))
        assert np.all(np.isfinite(seis[s]))
        assert np.max(np.abs(ref32)) > 0
        worst = max(worst, e32)
    assert worst <= SEIS_TOL


def test_uneven_record_lengths_in_one_batch():
    say("case uneven nt batch")
    nts = [400, 250, 330]
    ctx, cfg, d = setup_case(110, 46, 8, 2, 1, 10, 400, 60.0, nshots=3, nts=nts)
    seis = ctx.forward()
    o32 = oracle("f32")
    from pyoracle import make_cfg
    for s, n in enumerate(nts):
        c2 = make_cfg(110, 46, fdorder=8, maxrelerror=1, fw=10, free_surf=1, L=2, nt=n, dt=DT, dh=DH, damping=300.0,
                      fpml=60.0, fl=[6.25, 50.0])
        ref, _, _ = o32.forward(c2, d["vs"], d["rho"], d["qs"], d["src_i"][s:s + 1], d["src_j"][s:s + 1],
                                d["sig"][s][:n], d["rec_i"], d["rec_j"])
        e = rel_l2(seis[s][:, :n], ref)
        say("  shot %d nt=%d rel L2 %.3e" % (s, n, e))
        assert e <= SEIS_TOL
        assert np.all(seis[s][:, n:] == 0.0)


def test_batched_equals_sequential():
    say("case batch vs one-by-one")
    ctx, cfg, d = setup_case(100, 44, 8, 2, 1, 10, 300, 60.0, nshots=4)
    a = ctx.forward()
    ctx.set_batch(1)
    b = ctx.forward()
    assert np.array_equal(a, b)


GRAD_CASES = [
    ("el_fd4_dtinv1", 100, 44, 4, 0, 1, 10, 800, 60.0, 1),
    ("visc_fd8_dtinv3", 100, 44, 8, 2, 1, 10, 800, 60.0, 3),
    ("visc_fd8_nofs_dtinv4", 96, 40, 8, 2, 0, 10, 700, 60.0, 4),
    # EPRECOND 3 (eprecond1.c:34-50): the asinh receiver approximation in double
    # on the device and in the oracle's C library
    ("visc_fd8_eprecond3", 100, 44, 8, 2, 1, 10, 800, 60.0, 3, 3),
]


@pytest.mark.parametrize("case", GRAD_CASES, ids=[c[0] for c in GRAD_CASES])
def test_gradient(case):
    name, nx, ny, fdorder, L, fs, fw, nt, fc, dtinv = case[:10]
    eprecond = case[10] if len(case) > 10 else 1
    say("case", name)
    ctx, cfg, d = setup_case(nx, ny, fdorder, L, fs, fw, nt, fc, nshots=2, dtinv=dtinv, eprecond=eprecond)
    # observed data: the true model is 3% faster with an extra +10% box in vs
    vt = d["vs"] * np.float32(1.03)
    vt[ny // 4: ny // 2, nx // 3: 2 * nx // 3] *= np.float32(1.1)
    o32, o64 = oracle("f32"), oracle("f64")
    obs = np.zeros((2, len(d["rec_i"]), nt), np.float32)
    for s in range(2):
        ref, _, _ = o32.forward(cfg, vt, d["rho"], d["qs"], d["src_i"][s:s + 1], d["src_j"][s:s + 1], d["sig"][s],
                                d["rec_i"], d["rec_j"])
        obs[s] = ref
    ctx.set_observed(obs)
    l2, en = ctx.gradient()
    gvs, grho, raw_u, raw_rho = ctx.get_gradient()
    r32 = o32.gradient(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"], d["rec_j"], obs)
    r64 = o64.gradient(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"], d["rec_j"], obs)
    ev = rel_l2(gvs, r32["grad_vs"])
    er = rel_l2(grho, r32["grad_rho"])
    say("  grad_vs rel L2 vs f32 %.3e vs f64 %.3e (oracle f32 vs f64 %.3e)" %
        (ev, rel_l2(gvs, r64["grad_vs"]), rel_l2(r32["grad_vs"], r64["grad_vs"])))
    say("  grad_rho rel L2 vs f32 %.3e vs f64 %.3e (oracle f32 vs f64 %.3e)" %
        (er, rel_l2(grho, r64["grad_rho"]), rel_l2(r32["grad_rho"], r64["grad_rho"])))
    say("  l2 %.9e oracle %.9e; energy %.9e oracle %.9e" % (l2, r32["l2"], en, r32["energy"]))
    assert np.max(np.abs(r32["grad_vs"])) > 0
    assert ev <= GRAD_TOL and er <= GRAD_TOL
    assert abs(l2 - r32["l2"]) <= 1e-9 * abs(r32["l2"])
    assert abs(en - r32["energy"]) <= 1e-9 * abs(r32["energy"])


def _obs_for(ctx, cfg, d, nx, ny, nshots):
    vt = d["vs"] * np.float32(1.03)
    vt[ny // 4: ny // 2, nx // 3: 2 * nx // 3] *= np.float32(1.1)
    o32 = oracle("f32")
    obs = np.zeros((nshots, len(d["rec_i"]), d["sig"].shape[1]), np.float32)
    for s in range(nshots):
        obs[s], _, _ = o32.forward(cfg, vt, d["rho"], d["qs"], d["src_i"][s:s + 1], d["src_j"][s:s + 1],
                                   d["sig"][s], d["rec_i"], d["rec_j"])
    return obs


def test_graph_replay_equals_eager_launches():
    """The captured time-step chunk, replayed many times per pass and across
    passes (with eager work and host synchronisation in between), gives the
    same bits as launching every step from the host."""
    say("case graph replay vs eager")
    nts = [700, 530, 611]
    ctx, cfg, d = setup_case(100, 44, 8, 2, 1, 10, 700, 60.0, nshots=3, dtinv=3, snap_precision=1, nts=nts)
    ctx.set_exec(0)
    eager = ctx.forward()
    ctx.set_exec(-1)
    g1 = ctx.forward()
    g2 = ctx.forward()
    ctx.set_exec(30)            # chunk of 30 steps (lcm(2, dtinv) = 6)
    g3 = ctx.forward()
    for g in (g1, g2, g3):
        assert np.array_equal(g, eager)
    # gradient with bf16 snapshots: the pack kernel reads the step counter too
    from pyoracle import make_cfg
    obs = np.zeros((3, len(d["rec_i"]), 700), np.float32)
    obs[:, :, :] = eager * np.float32(0.9)
    ctx.set_observed(obs)
    ctx.set_exec(0)
    ctx.gradient()
    ge = ctx.get_gradient()
    ctx.set_exec(-1)
    for _ in range(2):
        ctx.gradient()
        gg = ctx.get_gradient()
        for a, b in zip(ge, gg):
            assert np.array_equal(a, b)


def test_autotuner_keeps_results():
    say("case autotune")
    ctx, cfg, d = setup_case(120, 60, 8, 2, 1, 10, 300, 60.0, nshots=4)
    ref = ctx.forward()
    rows, graph = ctx.autotune(4, 30)
    say("  tuned rows %d graph %d" % (rows, graph))
    assert 0 <= rows <= 60 and graph in (-1, 0)   # rows 0: the cost model's choice
    assert np.array_equal(ctx.forward(), ref)


@pytest.mark.parametrize("kind,tol", [(1, 1e-2), (2, 2e-3)], ids=["bf16", "fp16"])
@pytest.mark.parametrize("nx,ny", [(100, 44), (97, 43)], ids=["grid4400", "grid4171"])
def test_reduced_precision_snapshots(kind, tol, nx, ny):
    """Gradient with BF16 / FP16 snapshots (per-snapshot max-abs scaling, FP32
    accumulation): error against FP32 storage and the FP64 oracle stated.  A
    plane of 4400 cells is encoded four values per lane, one of 4171 cells
    one value per lane."""
    nt = 800
    say("case snapshots", "bf16" if kind == 1 else "fp16")
    ctx, cfg, d = setup_case(nx, ny, 8, 2, 1, 10, nt, 60.0, nshots=2, dtinv=2)
    obs = _obs_for(ctx, cfg, d, nx, ny, 2)
    ctx.set_observed(obs)
    ctx.gradient()
    f32 = ctx.get_gradient()
    ctx16, _, _ = setup_case(nx, ny, 8, 2, 1, 10, nt, 60.0, nshots=2, dtinv=2, snap_precision=kind)
    ctx16.set_observed(obs)
    ctx16.gradient()
    low = ctx16.get_gradient()
    r64 = oracle("f64").gradient(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"],
                                 d["rec_j"], obs)
    ev, er = rel_l2(low[0], f32[0]), rel_l2(low[1], f32[1])
    say("  grad_vs rel L2 vs fp32 storage %.3e, vs f64 oracle %.3e (fp32 storage vs f64 %.3e)" %
        (ev, rel_l2(low[0], r64["grad_vs"]), rel_l2(f32[0], r64["grad_vs"])))
    say("  grad_rho rel L2 vs fp32 storage %.3e, vs f64 oracle %.3e" % (er, rel_l2(low[1], r64["grad_rho"])))
    assert 0 < ev <= tol and er <= tol


BF16_CASES = [
    # name, nx, ny, fdorder, L, free_surf, fw, nt, dtinv, nshots, graph
    ("visc_fd8_dtinv3_graph", 100, 44, 8, 2, 1, 10, 800, 3, 3, -1),
    ("el_fd4_nofs_dtinv2_eager", 97, 43, 4, 0, 0, 10, 500, 2, 2, 0),
    ("visc_fd6_dtinv4_graph", 90, 40, 6, 1, 0, 8, 600, 4, 2, -1),
]


@pytest.mark.parametrize("case", BF16_CASES, ids=[c[0] for c in BF16_CASES])
def test_bf16_snapshots_match_oracle_bit_for_bit(case):
    """BF16 snapshots are encoded by the storing step with a power-of-two
    scale per snapshot and field, taken from the previous snapshot's
    maximum.  The oracle holding its snapshots the same way
    (orc_snap_bf16) gives the same gradient bit for bit, over graph replay
    and eager launches."""
    name, nx, ny, fdorder, L, fs, fw, nt, dtinv, nshots, graph = case
    say("case bf16 snapshots", name)
    ctx, cfg, d = setup_case(nx, ny, fdorder, L, fs, fw, nt, 60.0, nshots=nshots, dtinv=dtinv, snap_precision=1)
    ctx.set_exec(graph)
    obs = _obs_for(ctx, cfg, d, nx, ny, nshots)
    ctx.set_observed(obs)
    l2, en = ctx.gradient()
    gvs, grho, raw_u, raw_rho = ctx.get_gradient()
    cfg.snap_bf16 = 1
    r = oracle("f32").gradient(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"],
                               d["rec_j"], obs)
    cfg.snap_bf16 = 0
    r32 = oracle("f32").gradient(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"],
                                 d["rec_j"], obs)
    say("  grad_vs rel L2 vs oracle with bf16 snapshots %.3e (max abs diff %.3e); vs fp32 storage %.3e" %
        (rel_l2(gvs, r["grad_vs"]), np.max(np.abs(gvs - r["grad_vs"])), rel_l2(gvs, r32["grad_vs"])))
    assert np.max(np.abs(r["grad_vs"])) > 0
    assert np.array_equal(gvs, r["grad_vs"]) and np.array_equal(grho, r["grad_rho"])
    assert np.array_equal(raw_u, r["raw_u"]) and np.array_equal(raw_rho, r["raw_rho"])
    # FP64 sums in another order
    assert abs(l2 - r["l2"]) <= 1e-9 * abs(r["l2"]) and abs(en - r["energy"]) <= 1e-9 * abs(r["energy"])
    assert rel_l2(gvs, r32["grad_vs"]) > 0


def test_pseudo_hessian():
    """HESSIAN = 1 path: band-limited spike (QUELLART 6 low-passed at
    FC_HESSIAN), sums of squared forward fields, chain rule.  The reference's
    accumulator runs on across shots in FP32; the batched sum adds per-shot
    sums, so the bar is 1e-5 relative L2 (FP32 summation order)."""
    say("case pseudo-Hessian")
    nx, ny, nt = 100, 44, 500
    ctx, cfg, d = setup_case(nx, ny, 8, 2, 1, 10, nt, 60.0, nshots=3, dtinv=3)
    dn = product()
    spike = np.zeros((3, nt), np.float32)
    spike[:, 0] = 1.0
    sig = dn.timedomain_filt(spike, DT, 80.0, 4, 1)
    ctx.set_signals(sig)
    hv, hr = ctx.pseudo_hessian()
    r32 = oracle("f32").pseudo_hessian(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], sig)
    r64 = oracle("f64").pseudo_hessian(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], sig)
    ev, er = rel_l2(hv, r32[0]), rel_l2(hr, r32[1])
    say("  H_vs rel L2 vs f32 %.3e vs f64 %.3e; H_rho rel L2 vs f32 %.3e vs f64 %.3e" %
        (ev, rel_l2(hv, r64[0]), er, rel_l2(hr, r64[1])))
    assert np.max(r32[0]) > 0
    assert ev <= 1e-5 and er <= 1e-5


def _rademacher(seed, probe, n):
    """numpy restatement of the device's Rademacher signs (32-bit hash)."""
    m = np.uint64(0xFFFFFFFF)
    cell = np.arange(n, dtype=np.uint64)
    x = ((np.uint64(seed) * np.uint64(0x9E3779B9)) & m) ^ ((np.uint64(probe) * np.uint64(0x85EBCA6B)) & m) ^ \
        ((cell * np.uint64(0xC2B2AE35)) & m)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & m
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & m
    x ^= x >> np.uint64(16)
    return np.where(x & np.uint64(0x80000000), -1.0, 1.0).astype(np.float32)


def test_gradient_without_eprecond_and_hessian_probe():
    """EPRECOND 0 gradient against the oracle, energy maps against the
    oracle's Ws/Wr sums, and one finite-difference Hessian-vector probe
    against the same probe formed from oracle gradients."""
    nx, ny, nt = 90, 40, 700
    say("case eprecond 0 + Hessian probe")
    ctx, cfg, d = setup_case(nx, ny, 4, 0, 1, 10, nt, 60.0, nshots=2, dtinv=1, eprecond=0)
    obs = _obs_for(ctx, cfg, d, nx, ny, 2)
    ctx.set_observed(obs)
    ctx.gradient()
    g = ctx.get_gradient()
    o32 = oracle("f32")
    r = o32.gradient(cfg, d["vs"], d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"], d["rec_j"], obs)
    ev = rel_l2(g[0], r["grad_vs"])
    say("  EPRECOND 0 grad_vs rel L2 vs f32 %.3e" % ev)
    assert ev <= GRAD_TOL and rel_l2(g[1], r["grad_rho"]) <= GRAD_TOL
    ws, wr = ctx.energy_maps()
    ws_ref = np.zeros((ny, nx))
    for s in range(2):
        _, _, W = o32.forward(cfg, d["vs"], d["rho"], d["qs"], d["src_i"][s:s + 1], d["src_j"][s:s + 1], d["sig"][s],
                              d["rec_i"], d["rec_j"], want_snaps=True)
        nd = cfg.fdorder // 2 + 1
        ws_ref += W[nd:nd + ny, nd:nd + nx]
    say("  Ws map rel L2 %.3e" % rel_l2(ws, ws_ref))
    assert rel_l2(ws, ws_ref) <= 1e-5 and np.all(wr >= 0) and np.max(wr) > 0
    eps, seed = 1e-2, 7
    diag = ctx.hessian_probe(nprobes=1, eps=eps, seed=seed)
    v = _rademacher(seed, 0, nx * ny).reshape(ny, nx)
    vs1 = (d["vs"] * (np.float32(1.0) + np.float32(eps) * v)).astype(np.float32)
    r1 = o32.gradient(cfg, vs1, d["rho"], d["qs"], d["src_i"], d["src_j"], d["sig"], d["rec_i"], d["rec_j"], obs)
    ref = v * (r1["grad_vs"] - r["grad_vs"]) / (np.float32(eps) * d["vs"])
    e = rel_l2(diag, ref)
    say("  Hessian probe rel L2 vs oracle-built probe %.3e" % e)
    assert e <= 1e-4
    # the model is restored: the gradient is unchanged
    ctx.gradient()
    assert np.array_equal(ctx.get_gradient()[0], g[0])


def test_gradient_after_forward_on_same_shots():
    """Regression: a forward pass followed by a gradient over the same shots
    (the workspace is re-sized for snapshots in between) must give the
    gradient of a fresh context."""
    say("case gradient after forward, same shots")
    nx, ny, nt = 96, 40, 600
    ctx, cfg, d = setup_case(nx, ny, 8, 2, 1, 10, nt, 60.0, nshots=3, dtinv=2)
    obs = _obs_for(ctx, cfg, d, nx, ny, 3)
    ctx.forward()
    ctx.set_observed(obs)
    ctx.gradient()
    g1 = ctx.get_gradient()
    fresh, _, _ = setup_case(nx, ny, 8, 2, 1, 10, nt, 60.0, nshots=3, dtinv=2)
    fresh.set_observed(obs)
    fresh.gradient()
    g2 = fresh.get_gradient()
    for a, b in zip(g1, g2):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("L", [0, 2, 3])
def test_two_shots_per_wave_equals_one(L):
    """dn_set_wave_shots(2) and (1) (the default) give identical seismograms
    and gradients, with an odd shot count and uneven record lengths; elastic
    (packed float2 coefficients), L = 2 (float4 memory variables) and odd L
    (the float2 tail plane of the second shot of a wave)."""
    nx, ny, nt = 110, 48, 300
    say("  L %d" % L)
    ctx, cfg, d = setup_case(nx, ny, 8, L, 1, 10, nt, 60.0, nshots=3, nts=[300, 260, 290])
    obs = _obs_for(ctx, cfg, d, nx, ny, 3)
    for s, n in enumerate([300, 260, 290]):
        obs[s, :, n:] = 0.0
    ctx.set_observed(obs)
    out = {}
    for nsh in (1, 2):
        ctx.set_wave_shots(nsh)
        seis = ctx.forward()
        l2, en = ctx.gradient()
        gvs, grho, _, _ = ctx.get_gradient()
        out[nsh] = (seis, gvs, grho, l2, en)
    a, b = out[1], out[2]
    say("  seismograms equal %s, gradients equal %s" % (np.array_equal(a[0], b[0]), np.array_equal(a[1], b[1])))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert a[3] == b[3] and a[4] == b[4]
    ctx.close()


@pytest.mark.parametrize("snap", [0, 1])
def test_shot_group_layout_equals_strip_groups(snap):
    """dn_set_group_layout 2, 4 and 8 (workgroups of 4 strips x 2 shots, 2
    strips x 4 shots, 1 strip x 8 shots) give the seismograms and gradients
    of the default layout (8 strips of one shot) bit for bit: six shots (the
    last shot group has spare waves for 4 and 8) on 3 strips (the last strip
    group has spare waves for 2 and 4), uneven record lengths (a group's
    shots finish at different steps), with one and two shots per wave on the
    plain steps, FP32 and BF16 snapshots."""
    nx, ny, nt = 130, 52, 300
    nts = [300, 240, 280, 300, 200, 260]
    ctx, cfg, d = setup_case(nx, ny, 8, 2, 1, 10, nt, 60.0, nshots=6, dtinv=3, nts=nts, snap_precision=snap)
    obs = _obs_for(ctx, cfg, d, nx, ny, 6)
    for s, n in enumerate(nts):
        obs[s, :, n:] = 0.0
    ctx.set_observed(obs)
    out = {}
    for layout in (0, 2, 4, 8):
        for nsh in (1, 2):
            ctx.set_group_layout(layout)
            ctx.set_wave_shots(nsh)
            seis = ctx.forward()
            l2, en = ctx.gradient()
            gvs, grho, _, _ = ctx.get_gradient()
            out[(layout, nsh)] = (seis, gvs, grho, l2, en)
    ctx.set_group_layout(0)
    ctx.set_wave_shots(1)
    a = out[(0, 1)]
    for k, b in out.items():
        say("  layout %d, %d shot(s) per wave: seismograms equal %s, gradients equal %s" % (
            k[0], k[1], np.array_equal(a[0], b[0]), np.array_equal(a[1], b[1])))
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
        assert a[3] == b[3] and a[4] == b[4]
    assert np.abs(a[1]).max() > 0
    ctx.close()



def test_batches_by_record_length_give_the_shot_order_sums():
    """Seven shots of uneven record lengths in batches of three: batches
    formed by decreasing length (the default) give the gradient, the misfit
    and the energy maps of batches in shot order, of one shot at a time and
    of one batch of all seven, bit for bit (the per-shot terms are summed in
    shot order)."""
    say("case batches by record length")
    nx, ny, nt = 110, 46, 300
    nts = [300, 180, 260, 120, 300, 220, 150]
    ctx, cfg, d = setup_case(nx, ny, 8, 2, 1, 10, nt, 60.0, nshots=7, dtinv=3, nts=nts)
    obs = _obs_for(ctx, cfg, d, nx, ny, 7)
    for s, n in enumerate(nts):
        obs[s, :, n:] = 0.0
    ctx.set_observed(obs)
    out = {}
    for name, batch, by_len in (("length, 3", 3, True), ("shot order, 3", 3, False), ("one at a time", 1, True),
                                ("all seven", 7, True)):
        ctx.set_batch(batch)
        ctx.set_batch_order(by_len)
        l2, en = ctx.gradient()
        out[name] = (ctx.get_gradient(), ctx.energy_maps(), l2, en)
    ctx.set_batch(0)
    ctx.set_batch_order(True)
    a = out["length, 3"]
    assert np.abs(a[0][0]).max() > 0
    for name, b in out.items():
        same = all(np.array_equal(x, y) for x, y in zip(a[0] + list(a[1]), b[0] + list(b[1])))
        say("  %s: gradients and energy maps equal %s, misfit %.9e" % (name, same, b[2]))
        assert same and a[2] == b[2] and a[3] == b[3], name
    ctx.close()


def test_gradient_means_are_those_of_the_first_gradient():
    """The normalisation means C_vs, C_rho come from the model of the first
    gradient and stay fixed (denise.c:746-819 computes them in the first
    iteration only): another model set before it, as when observed data are
    modelled from a true model, does not count; a later gradient at another
    model keeps the first means."""
    nx, ny, nt = 100, 44, 400
    ctx, cfg, d = setup_case(nx, ny, 4, 0, 1, 10, nt, 60.0, nshots=1)
    obs = _obs_for(ctx, cfg, d, nx, ny, 1)
    o32 = oracle("f32")
    ctx.set_model(d["vs"] * np.float32(1.2), d["rho"] * np.float32(0.9))
    ctx.set_model(d["vs"], d["rho"])
    ctx.set_observed(obs)
    ctx.gradient()
    gvs, grho, _, _ = ctx.get_gradient()
    r = o32.gradient(cfg, d["vs"], d["rho"], None, d["src_i"], d["src_j"], d["sig"], d["rec_i"], d["rec_j"], obs)
    say("  first gradient: grad_vs rel L2 %.3e" % rel_l2(gvs, r["grad_vs"]))
    assert np.array_equal(gvs, r["grad_vs"]) and np.array_equal(grho, r["grad_rho"])
    # a second gradient at another model: the means of the first
    from fwi_oracle import float_mean
    vs2 = d["vs"] * np.float32(1.05)
    ctx.set_model(vs2, d["rho"])
    ctx.gradient()
    g2, gr2, _, _ = ctx.get_gradient()
    cfg.c_vs, cfg.c_rho = float(float_mean(d["vs"])), float(float_mean(d["rho"]))
    r2 = o32.gradient(cfg, vs2, d["rho"], None, d["src_i"], d["src_j"], d["sig"], d["rec_i"], d["rec_j"], obs)
    cfg.c_vs = cfg.c_rho = 0.0
    say("  second gradient, first means: grad_vs rel L2 %.3e" % rel_l2(g2, r2["grad_vs"]))
    assert np.array_equal(g2, r2["grad_vs"]) and np.array_equal(gr2, r2["grad_rho"])
    ctx.close()


def test_segment_heights_give_identical_results():
    """Rows per segment 8, 12, 16, 24 and one segment for the whole grid:
    seismograms and gradients identical (the segments recompute 2K-1 rows
    of the velocity at their edges; small batches default to short
    segments, the measured workload to 16 rows)."""
    nx, ny, nt = 130, 56, 300
    ctx, cfg, d = setup_case(nx, ny, 8, 2, 1, 10, nt, 60.0, nshots=3, dtinv=3)
    obs = _obs_for(ctx, cfg, d, nx, ny, 3)
    ctx.set_observed(obs)
    out = {}
    for rows in (8, 12, 16, 24, ny):
        ctx.set_batch(0, rows)
        seis = ctx.forward()
        ctx.gradient()
        gvs, grho, _, _ = ctx.get_gradient()
        out[rows] = (seis, gvs, grho)
    a = out[ny]
    for rows, b in out.items():
        say("  rows %d: equal %s" % (rows, all(np.array_equal(x, y) for x, y in zip(a, b))))
        assert all(np.array_equal(x, y) for x, y in zip(a, b))
    ctx.close()


@pytest.mark.parametrize("L", [0, 2])
def test_cpml_with_k_max_above_one(L):
    """K_MAX_PML = 2 (CPML stretching K > 1, divided through in the CPML
    cells; every other test uses the default K = 1), top CPML on:
    seismograms and gradients against the FP32 oracle."""
    from pyoracle import make_cfg
    nx, ny, nt, fc = 110, 48, 400, 60.0
    dn = product()
    vs, rho, qs = model(nx, ny, DH, seed=8, qs=20.0 if L else None)
    fl = [6.25, 50.0][:L]
    kw = dict(fdorder=8, maxrelerror=1, fw=10, free_surf=0, L=L, dt=DT, dh=DH, damping=300.0, fpml=fc, dtinv=2)
    ctx = dn.Context(nx, ny, fl=fl, k_max_pml=2.0, **kw)
    cfg = make_cfg(nx, ny, nt=nt, fl=fl, k_max_pml=2.0, **kw)
    rec_i = np.arange(13, nx - 12, 3, dtype=np.int32)
    rec_j = np.full_like(rec_i, 5)
    src_i = np.array([30, 80], np.int32)
    src_j = np.array([6, 40], np.int32)
    sig = np.stack([ricker(nt, DT, fc) for _ in range(2)])
    ctx.set_receivers(rec_i, rec_j)
    ctx.set_shots(src_i, src_j, sig)
    o32 = oracle("f32")
    obs = np.stack([o32.forward(cfg, vs * np.float32(1.05), rho, qs, src_i[k:k + 1], src_j[k:k + 1], sig[k],
                                rec_i, rec_j)[0] for k in range(2)]).astype(np.float32)
    ctx.set_model(vs, rho, qs)
    seis = ctx.forward()
    ref = np.stack([o32.forward(cfg, vs, rho, qs, src_i[k:k + 1], src_j[k:k + 1], sig[k], rec_i, rec_j)[0]
                    for k in range(2)])
    ctx.set_observed(obs)
    ctx.gradient()
    gvs, grho, _, _ = ctx.get_gradient()
    r = o32.gradient(cfg, vs, rho, qs, src_i, src_j, sig, rec_i, rec_j, obs)
    say("  L %d: seismograms rel L2 %.2e, grad_vs rel L2 %.2e" % (L, rel_l2(seis, ref), rel_l2(gvs, r["grad_vs"])))
    assert np.array_equal(seis, ref)
    assert np.array_equal(gvs, r["grad_vs"]) and np.array_equal(grho, r["grad_rho"])
    ctx.close()


# nx, ny, fdorder, L, free_surf, fw: grids smaller than a strip group or a
# segment, rows fewer than the stencil reach, no absorbing frame at all
EDGE_GRIDS = [
    (20, 12, 8, 2, 1, 3),
    (65, 9, 4, 0, 0, 2),
    (130, 16, 12, 1, 1, 4),
    (57, 57, 2, 0, 1, 0),
    (71, 33, 6, 2, 0, 0),
]


@pytest.mark.parametrize("case", EDGE_GRIDS, ids=["%dx%d_fd%d_L%d_fs%d_fw%d" % c for c in EDGE_GRIDS])
def test_edge_grids_match_oracle(case):
    """Seismograms and gradients bit for bit on small and awkward grids, with
    receivers on the first and last interior rows and columns and a source
    in a corner cell."""
    from pyoracle import make_cfg
    nx, ny, fdorder, L, fs, fw = case
    nt, fc = 200, 60.0
    dn = product()
    vs, rho, qs = model(nx, ny, DH, seed=nx + ny, qs=20.0 if L else None)
    fl = [6.25, 50.0, 12.0][:L]
    kw = dict(fdorder=fdorder, maxrelerror=1, fw=fw, free_surf=fs, L=L, dt=DT, dh=DH, damping=300.0, fpml=fc,
              dtinv=2)
    ctx = dn.Context(nx, ny, fl=fl, **kw)
    cfg = make_cfg(nx, ny, nt=nt, fl=fl, **kw)
    rec_i = np.array([1, nx, nx // 2, 1, nx], np.int32)
    rec_j = np.array([1, 1, ny // 2, ny, ny], np.int32)
    src_i = np.array([1, nx // 2], np.int32)
    src_j = np.array([1, max(1, ny // 3)], np.int32)
    sig = np.stack([ricker(nt, DT, fc) for _ in range(2)])
    ctx.set_receivers(rec_i, rec_j)
    ctx.set_shots(src_i, src_j, sig)
    o32 = oracle("f32")
    obs = np.stack([o32.forward(cfg, vs * np.float32(1.05), rho, qs, src_i[k:k + 1], src_j[k:k + 1], sig[k],
                                rec_i, rec_j)[0] for k in range(2)]).astype(np.float32)
    ref = np.stack([o32.forward(cfg, vs, rho, qs, src_i[k:k + 1], src_j[k:k + 1], sig[k], rec_i, rec_j)[0]
                    for k in range(2)])
    ctx.set_model(vs, rho, qs)
    seis = ctx.forward()
    ctx.set_observed(obs)
    ctx.gradient()
    gvs, grho, _, _ = ctx.get_gradient()
    r = o32.gradient(cfg, vs, rho, qs, src_i, src_j, sig, rec_i, rec_j, obs)
    say("  %s: seismograms rel L2 %.2e, grad_vs rel L2 %.2e" % (case, rel_l2(seis, ref), rel_l2(gvs, r["grad_vs"])))
    assert np.all(np.isfinite(seis)) and np.max(np.abs(ref)) > 0
    assert np.array_equal(seis, ref)
    assert np.array_equal(gvs, r["grad_vs"]) and np.array_equal(grho, r["grad_rho"])
    ctx.close()
