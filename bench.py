"""Benchmark of the SH adjoint-gradient hot path on one MI355X.

Four workloads (viscoelastic SH, FDORDER 8 Holberg with MAXRELERROR 1, L = 2
relaxation mechanisms with Q = 20 and FL = 6.25 / 50 Hz, 20-point CPML on
the sides and bottom, free surface on top, snapshots every DTINV = 4 steps):

* ``c4`` (default, the headline): NX = 4000, NY = 800, DH = 0.1 m,
  DT = 5e-5 s, NT = 12 000; vs in four layers 80 / 150 / 250 / 400 m/s with
  twelve seeded Gaussian lenses (+-15 %, sigma 3 m, 2-40 m deep) in the true model, the
  plain layers as the model the gradient is taken at; rho 1800 -> 2000;
  32 shots evenly spread from 10 m to 390 m along the row j = 5 (0.5 m
  deep), 60 Hz Ricker, 394
  receivers every 1 m on the same row.  A shot stores 3000 snapshots of
  three fields (115 GB in FP32), so HBM holds two shots at a time: one step
  is the gradient of one such batch (forward with snapshot storage,
  residuals, adjoint with correlation, energy preconditioning, chain rule),
  the steps walk through the 16 batches of the 32 shots in turn.
* ``c2``: NX = 1200, NY = 300, DH = 0.2 m, DT = 1e-4 s, NT = 10 000; vs
  100 -> 350 m/s with depth times a +-10 % checkerboard of 8 m cells (true)
  and the plain gradient (start); rho 1700 -> 2000; 40 shots, 25 Hz Ricker,
  240 receivers at 1 m spacing.  One step is the gradient over all 40
  shots (batches of as many shots as fit in HBM).
* ``c3``: the batched workload: NX = 2000, NY = 400, DH = 0.2 m, DT = 1e-4 s,
  the c2 model family; 64 shots with seeded record lengths of 0.4-1.2 s and
  wavelets low-passed at 15, 30 or 60 Hz (Butterworth order 4, shot s in
  stage s mod 3), all mixed in the same batches; 390 receivers at 1 m
  spacing.  One step is the gradient over all 64 shots;
  --compare-sequential also times it with one shot per batch.
* ``c5``: the learned preconditioner's held-out workload: NX = 4000,
  NY = 800, DH = 0.2 m, DT = 1e-4 s, NT = 10 000; a model of the training
  family (random_checkerboard of denise-sh_amd/precond_train.py, seed 9999;
  training uses seeds 7000+); 64 shots, 25 Hz Ricker, 790 receivers at 1 m
  spacing.  Meant for --fwi-learned with --train-config c2.

Observed data are synthesised on the device from the true model before
timing (synthetic, seeded).  The metric counts cell-updates of the forward
and adjoint sweeps (NX*NY*NT*2 per shot).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c4|c2|c3|c5] [--compare-sequential]
       [--nt NT] [--shots S] [--snap fp32|bf16|fp16] [--cpu-seconds T] [--no-cpu]
       [--rows R] [--max-batch B] [--autotune] [--exec graph|eager]
       python bench.py --config c2 --fwi N     (N plain L-BFGS iterations, wall time)
       python bench.py --config c2 --fwi-learned N [--pairs P] [--probes Q] [--label-shots S]
                                               (learned preconditioner: training, then N FWI
                                                iterations plain and learned, time to plain's misfit)
       python bench.py --unet N                (N training steps of the preconditioner's U-Net)
       python bench.py --copy-probe            (copy kernel that scales HBM counters)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def _package():
    import importlib.util
    name = "denise_sh_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "denise-sh_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def _precond_train():
    """denise-sh_amd/precond_train.py (training pairs and the model family of
    the learned preconditioner)."""
    import importlib.util
    name = "precond_train"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "denise-sh_amd", "precond_train.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def layered_lenses(nx, ny, dh, seed):
    """vs layers 80 / 150 / 250 / 400 m/s with depth and twelve seeded
    Gaussian lenses (+-15 %, sigma 3 m, centres between 2 % and 50 % of the
    depth) in the true model; the plain layers as the starting model;
    rho 1800 -> 2000; Q 20."""
    j = np.arange(ny)[:, None]
    i = np.arange(nx)[None, :]
    layer = np.minimum(4 * j // ny, 3)
    vs0 = np.array([80.0, 150.0, 250.0, 400.0])[layer] + 0.0 * i
    rng = np.random.default_rng(seed)
    pert = np.zeros((ny, nx))
    for _ in range(12):
        cx, cy = rng.uniform(0, nx * dh), rng.uniform(0.02 * ny * dh, 0.5 * ny * dh)
        pert += rng.choice([-1.0, 1.0]) * 0.15 * np.exp(-(((i * dh - cx) ** 2 + (j * dh - cy) ** 2) / (2 * 3.0 ** 2)))
    rho = (1800.0 + 200.0 * j / (ny - 1) + 0.0 * i).astype(np.float32)
    return (vs0 * (1 + pert)).astype(np.float32), vs0.astype(np.float32), rho, np.full((ny, nx), 20.0, np.float32)


def checkerboard(nx, ny, dh):
    """vs 100 -> 350 m/s with depth times a +-10 % checkerboard of 8 m cells
    (true) and the plain gradient (start); rho 1700 -> 2000; Q 20."""
    j = np.arange(ny, dtype=np.float64)[:, None]
    i = np.arange(nx, dtype=np.float64)[None, :]
    vs1d = 100.0 + 250.0 * j / (ny - 1) + 0.0 * i
    cell = 8.0 / dh
    cb = np.where(((np.floor(i / cell) + np.floor(j / cell)) % 2) == 0, 1.0, -1.0)
    vs_true = (vs1d * (1.0 + 0.1 * cb)).astype(np.float32)
    rho = (1700.0 + 300.0 * j / (ny - 1) + 0.0 * i).astype(np.float32)
    return vs_true, vs1d.astype(np.float32), rho, np.full((ny, nx), 20.0, np.float32)


def workload(config="c4", nt=None):
    """Model, acquisition and solver settings of a workload (see the module
    docstring).  nt overrides the record length."""
    if config == "c2":
        nx, ny, dh, dt = 1200, 300, 0.2, 1e-4
        vs_true, vs_start, rho, qs = checkerboard(nx, ny, dh)
        nshots, nrec, fc = 40, 240, 25.0
        src_x = np.linspace(3.0, nx * dh - 3.0, nshots)
        rec_x = 0.5 + np.arange(nrec) * 1.0
        row = int(np.floor(0.5 / dh + 0.5))
        nt = 10000 if nt is None else nt
        damping = 350.0
    elif config == "c3":
        nx, ny, dh, dt = 2000, 400, 0.2, 1e-4
        vs_true, vs_start, rho, qs = checkerboard(nx, ny, dh)
        nshots, nrec, fc = 64, 390, 25.0
        src_x = np.linspace(3.0, nx * dh - 3.0, nshots)
        rec_x = 5.0 + np.arange(nrec) * 1.0
        row = int(np.floor(0.5 / dh + 0.5))
        rng = np.random.default_rng(3)
        nts = np.floor(rng.uniform(0.4, 1.2, nshots) / dt + 0.5).astype(np.int32)
        nt = int(nts.max()) if nt is None else nt
        nts = np.minimum(nts, nt)
        damping = 350.0
    elif config == "c5":
        # held-out model of the learned preconditioner's training family
        # (random_checkerboard, seeds 7000+ for training) on a 4000 x 800 grid
        nx, ny, dh, dt = 4000, 800, 0.2, 1e-4
        vs_true, vs_start, rho, qs = _precond_train().random_checkerboard(nx, ny, dh, 9999)
        nshots, nrec, fc = 64, 790, 25.0
        src_x = np.linspace(3.0, nx * dh - 3.0, nshots)
        rec_x = 5.0 + np.arange(nrec) * 1.0
        row = int(np.floor(0.5 / dh + 0.5))
        nt = 10000 if nt is None else nt
        damping = 350.0
    elif config == "c4":
        nx, ny, dh, dt = 4000, 800, 0.1, 5e-5
        vs_true, vs_start, rho, qs = layered_lenses(nx, ny, dh, seed=5)
        nshots, nrec, fc = 32, 394, 60.0
        src_x = np.linspace(10.0, nx * dh - 10.0, nshots)
        rec_x = 3.0 + np.arange(nrec) * 1.0
        row = 5
        nt = 12000 if nt is None else nt
        damping = 350.0
    else:
        raise ValueError("unknown workload %r" % config)
    src_i = np.floor(src_x / dh + 0.5).astype(np.int32)          # iround(x/DH), splitsrc.c
    src_j = np.full(nshots, row, np.int32)
    rec_i = np.floor(rec_x / dh + 0.5).astype(np.int32)          # receiver.c:40
    rec_j = np.full(nrec, row, np.int32)
    W = dict(config=config, nx=nx, ny=ny, dh=dh, dt=dt, nt=nt, vs_true=vs_true, vs_start=vs_start, rho=rho,
             qs=qs, src_i=src_i, src_j=src_j, rec_i=rec_i, rec_j=rec_j, fdorder=8, L=2, fl=[6.25, 50.0], fw=20,
             fc=fc, dtinv=4, damping=damping)
    if config == "c3":
        # uneven record lengths and three low-pass stages (Butterworth order
        # 4 at 15 / 30 / 60 Hz, shot s in stage s mod 3) mixed in one batch
        W["nts"] = nts
        W["stages"] = [15.0, 30.0, 60.0]
    return W


def describe(W, shots, snap):
    if "nts" in W:
        nts = W["nts"][:shots]
        return ("visco SH gradient %dx%d FDORDER8 L2, %d shots x %d-%d steps (uneven), low-pass stages %s Hz "
                "mixed, DTINV %d, %s snapshots" % (W["nx"], W["ny"], shots, int(nts.min()), int(nts.max()),
                                                   "/".join("%g" % f for f in W["stages"]), W["dtinv"], snap))
    return ("visco SH gradient %dx%d FDORDER8 L2, %d shots x %d steps, DTINV %d, %s snapshots"
            % (W["nx"], W["ny"], shots, W["nt"], W["dtinv"], snap))


def signals_of(dn, W, nshots):
    """Source signals [nshots][nt]: the Ricker wavelet, for the batched
    workload low-passed per stage and cut at each shot's record length."""
    base = ricker(dn, W["nt"], W["dt"], W["fc"])
    sig = np.tile(base, (nshots, 1)).astype(np.float32)
    if "nts" in W:
        for s in range(nshots):
            sig[s] = dn.timedomain_filt(base[None], W["dt"], W["stages"][s % 3], 4, 1)[0]
            sig[s, W["nts"][s]:] = 0.0
    return sig


def ricker(dn, nt, dt, fc):
    return dn.wavelet(1, nt, dt, 0.0, fc, 1.0)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_baseline(W, seconds):
    """Oracle (FP32 restatement of the reference, OpenMP over rows) on the host
    cores: one shot of the same grid, forward + adjoint gradient.  A short
    probe sizes the sample to about `seconds` of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Oracle, make_cfg
    o = Oracle("f32")

    def run(nt_sample):
        cfg = make_cfg(W["nx"], W["ny"], fdorder=W["fdorder"], fw=W["fw"], free_surf=1, L=W["L"], nt=nt_sample,
                       dtinv=W["dtinv"], dt=W["dt"], dh=W["dh"], damping=W["damping"], fpml=W["fc"], fl=W["fl"])
        sig = o.wavelet(1, nt_sample, W["dt"], 0.0, W["fc"], 1.0).astype(np.float32)[None]
        obs = np.zeros((1, len(W["rec_i"]), nt_sample), np.float32)
        t0 = time.perf_counter()
        o.gradient(cfg, W["vs_start"], W["rho"], W["qs"], W["src_i"][:1], W["src_j"][:1], sig, W["rec_i"],
                   W["rec_j"], obs)
        return time.perf_counter() - t0

    probe = 40
    t = run(probe)
    nt_sample = int(min(W["nt"], max(probe, probe * seconds / max(t, 1e-3))))
    nt_sample -= nt_sample % W["dtinv"]
    sec = run(nt_sample)
    cells = 2.0 * W["nx"] * W["ny"] * nt_sample
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return dict(value=cells / sec / 1e9, unit="Gcell-updates/s", cores=threads, kind="port",
                sample="1 of %d shots, forward+adjoint gradient over %d of %d steps on the same %dx%d grid "
                       "(%.1f s, %d OpenMP threads on %s)"
                       % (len(W["src_i"]), nt_sample, W["nt"], W["nx"], W["ny"], sec, threads, cpu_model()))


def traffic_file(W):
    return os.path.join(ROOT, "profiles", "k_step_visco_gradient_%dx%d_hbm_traffic.md" % (W["nx"], W["ny"]))


def measured_traffic(W, batch, cells_per_launch):
    """HBM bytes per launch of the step kernel (mean over the forward and
    adjoint launches of a gradient), from the FETCH_SIZE / WRITE_SIZE
    profile of this grid and batch shape kept in profiles/ (scaled by the
    copy-probe factors): bytes per cell-update x cell-updates per launch.
    None when no profile of this shape exists."""
    try:
        per, shots = None, None
        for line in open(traffic_file(W)):
            if line.startswith("traffic_bytes_per_cell_update_gradient:"):
                per = float(line.split(":")[1])
            if line.startswith("shots_per_launch:"):
                shots = int(line.split(":")[1])
        if per is not None and (shots is None or shots == batch):
            return round(per * cells_per_launch)
    except OSError:
        pass
    return None


def make_context(dn, W, args, nshots, prec, eprecond=1, shots=None):
    """A context of the workload with the execution settings of the command
    line; shots (indices) selects a subset of the acquisition."""
    ctx = dn.Context(W["nx"], W["ny"], fdorder=W["fdorder"], maxrelerror=1, fw=W["fw"], free_surf=1, L=W["L"],
                     fl=W["fl"], dt=W["dt"], dh=W["dh"], damping=W["damping"], fpml=W["fc"], dtinv=W["dtinv"],
                     snap_precision=prec, eprecond=eprecond)
    ctx.set_exec(-1 if args.exec == "graph" else 0)
    ctx.set_wave_shots(args.wave_shots)
    if args.rows or args.max_batch:
        ctx.set_batch(args.max_batch, args.rows)
    ctx.set_group_layout(args.layout)
    ctx.set_batch_order(args.batch_order == "length")
    ctx.set_receivers(W["rec_i"], W["rec_j"])
    sel = np.arange(nshots) if shots is None else np.asarray(shots)
    sig = signals_of(dn, W, len(W["src_i"]))[sel]
    ctx.set_shots(W["src_i"][sel], W["src_j"][sel], sig, W["nts"][sel] if "nts" in W else None)
    return ctx


def fwi_run(ctx, W, args):
    """Plain L-BFGS FWI of the workload (GRAD_METHOD 2, NLBFGS 20, EPRECOND 1,
    vs only, test shots 1, 11, 21, 31 for the line search) from the starting
    model: wall time of the iterations and the misfit after each, with the
    distance of the model to the true one.  The library's calls return with
    their work done (they synchronise their own stream)."""
    nshots = len(W["src_i"]) if args.shots <= 0 else args.shots
    ctx.fwi_start(testshot_start=1, testshot_end=nshots, testshot_incr=10, eps_scale=0.01, stepmax=6,
                  vs_lower=50.0, vs_upper=600.0)
    hist = []
    t0 = time.perf_counter()
    for _ in range(args.fwi):
        r = ctx.fwi_iteration()
        hist.append(dict(misfit=round(float(r["L2t"][3]), 6), step=round(float(r["opteps_vs"]), 6),
                         test_forwards=int(r["test_forwards"])))
        print("fwi iteration %d: L2t %.6g, step %.4g, %d test forwards, %.1f s" % (
            len(hist), hist[-1]["misfit"], hist[-1]["step"], hist[-1]["test_forwards"],
            time.perf_counter() - t0), file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    vs, _ = ctx.get_model()
    err0 = float(np.linalg.norm(W["vs_start"] - W["vs_true"]) / np.linalg.norm(W["vs_true"]))
    err = float(np.linalg.norm(vs - W["vs_true"]) / np.linalg.norm(W["vs_true"]))
    print(json.dumps({"metric": "FWI wall time, plain L-BFGS", "value": round(wall, 3), "unit": "s",
                      "higher_is_better": False, "iterations": args.fwi, "s_per_iteration": round(wall / args.fwi, 3),
                      "misfit_L2t": [h["misfit"] for h in hist], "steps": [h["step"] for h in hist],
                      "test_forwards": [h["test_forwards"] for h in hist],
                      "vs_rel_error": {"start": round(err0, 5), "end": round(err, 5)},
                      "config": {"workload": "visco SH FWI %dx%d FDORDER8 L2, %d shots x %d steps, DTINV %d"
                                             % (W["nx"], W["ny"], nshots, W["nt"], W["dtinv"]), "n_gpus": 1}}))


LEARNED_MODES = {
    # name: (judge_every, D fixed per stage, plain for the rest of the stage
    # once the plain direction replaced the learned one, power of D)
    "plain": None,
    "learned_fixed": (0, True, True, 1.0),
    "learned_fixed_judge_every": (1, True, True, 1.0),
    "learned_fixed_judge_5": (5, True, True, 1.0),
    "learned_recomputed": (0, False, True, 1.0),
    "learned_fixed_no_revert": (0, True, False, 1.0),
    "learned_fixed_power05": (0, True, True, 0.5),
    "learned_fixed_power15": (0, True, True, 1.5),
}


def fwi_learned_run(dn, W, args, nshots, prec):
    """The learned preconditioner on the workload's FWI: training pairs from
    seeded models of the same family (random_checkerboard) on the same grid,
    with --label-shots of the shots and EPRECOND 0, labels from --probes
    finite-difference Hessian probes; --train-steps Adam steps; then
    --fwi-learned iterations of plain L-BFGS and of each learned mode from
    the same start (fresh contexts, eprecond 1, test shots 1, 11, 21, 31),
    and the wall time each needs to reach the plain run's final misfit."""
    pt = _precond_train()
    log = lambda m: print("fwi-learned: " + m, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    # training pairs on the grid and acquisition of --train-config (default:
    # the FWI workload's own), full record length
    WT = W if args.train_config in (None, W["config"]) else workload(args.train_config)
    nshots_t = len(WT["src_i"])
    lsel = np.linspace(0, nshots_t - 1, min(args.label_shots, nshots_t)).round().astype(int)
    lab = make_context(dn, WT, args, nshots_t, prec, eprecond=0, shots=lsel)
    models = [pt.random_checkerboard(WT["nx"], WT["ny"], WT["dh"], 7000 + k) for k in range(args.pairs)]
    geometry = None
    if args.train_geometry == "mixed":
        # each pair its own acquisition: a random subset of 8..S shots of the
        # workload's positions and a record of 0.4..1 of its length
        sig_all = signals_of(dn, WT, nshots_t)
        grng = np.random.default_rng(4711)
        def geometry(k):
            m = int(grng.integers(8, min(args.label_shots, nshots_t) + 1))
            sel = np.sort(grng.choice(nshots_t, size=m, replace=False))
            nt_k = int(grng.uniform(0.4, 1.0) * WT["nt"])
            lab.set_shots(WT["src_i"][sel], WT["src_j"][sel], sig_all[sel], np.full(m, nt_k, np.int32))
    X, T = pt.pairs_on_context(lab, models, nprobes=args.probes, seed0=100, log=log, geometry=geometry)
    lab.close()
    t_labels = time.perf_counter() - t0
    ntr = max(1, (4 * args.pairs) // 5)
    net = dn.UNet(seed=1)
    t1 = time.perf_counter()
    hist = pt.train(net, X[:ntr], T[:ntr], steps=args.train_steps, lr=2e-3, batch=4)
    t_train = time.perf_counter() - t1
    held = {"pairs": int(args.pairs - ntr)}
    if args.pairs > ntr:
        held["loss"] = round(float(net.loss_grad(X[ntr:], T[ntr:])[0]), 5)
        held["constant_predictor"] = round(float(np.mean((np.mean(T[:ntr]) - T[ntr:]) ** 2)), 5)
    log("labels %.0f s, training %.0f s: loss %.4f -> %.4f, held-out %s" % (t_labels, t_train, hist[0], hist[-1],
                                                                             held))
    runs = {}
    names = [m for m in args.modes.split(",") if m]
    for name in names:
        if name not in LEARNED_MODES:
            raise SystemExit("bench.py: unknown --modes entry %r (choose from %s)" % (name, ", ".join(LEARNED_MODES)))
    if "plain" not in names or "learned_fixed" not in names:
        raise SystemExit("bench.py: --modes must include plain and learned_fixed")
    # with --shots below the acquisition's count, shots spread evenly over it
    fsel = np.linspace(0, len(W["src_i"]) - 1, nshots).round().astype(int)
    for name in names:
        mode = LEARNED_MODES[name]
        ctx = make_context(dn, W, args, nshots, prec, eprecond=1, shots=fsel)
        ctx.set_model(W["vs_true"], W["rho"], W["qs"])
        obs = ctx.forward()
        ctx.set_model(W["vs_start"], W["rho"], W["qs"])
        ctx.set_observed(obs)
        del obs
        if mode is None:
            ctx.set_learned(None)
        else:
            ctx.set_learned(net, mode[0], mode[1], mode[2], mode[3])
        ctx.fwi_start(testshot_start=1, testshot_end=nshots, testshot_incr=10, eps_scale=0.01, stepmax=6,
                      vs_lower=50.0, vs_upper=600.0)
        mis, sec, taken = [], [], []
        ts = time.perf_counter()
        for it in range(args.fwi_learned):
            r = ctx.fwi_iteration()
            mis.append(float(r["L2t"][3]))
            sec.append(time.perf_counter() - ts)
            taken.append(int(r["precond"]))
            log("%s iteration %d: L2t %.6g, direction %s, %.1f s" % (name, it + 1, mis[-1],
                                                                   "learned" if taken[-1] else "plain", sec[-1]))
        vs, _ = ctx.fwi_model()
        ctx.close()
        runs[name] = {"misfit_L2t": [round(m, 4) for m in mis], "seconds": [round(t, 2) for t in sec],
                      "learned_taken": "".join(map(str, taken)),
                      "vs_rel_error": round(float(np.linalg.norm(vs - W["vs_true"]) / np.linalg.norm(W["vs_true"])),
                                            5)}
    target = runs["plain"]["misfit_L2t"][-1]
    t_plain = runs["plain"]["seconds"][-1]
    for r in runs.values():
        t = next((t for m, t in zip(r["misfit_L2t"], r["seconds"]) if m <= target), None)
        r["seconds_to_plain_final_misfit"] = t
        r["ratio_to_plain"] = None if t is None else round(t / t_plain, 4)
    print(json.dumps({
        "metric": "FWI wall time to the plain method's final misfit, learned / plain",
        "value": runs["learned_fixed"]["ratio_to_plain"], "unit": "ratio", "higher_is_better": False,
        "n_gpus": 1, "iterations": args.fwi_learned, "plain_final_misfit": target, "plain_seconds": t_plain,
        "training": {"grid": "%dx%d" % (WT["nx"], WT["ny"]), "model_family": "random_checkerboard seeds 7000+",
                     "geometry": args.train_geometry,
                     "pairs": args.pairs, "train_pairs": ntr, "probes_per_label": args.probes,
                     "shots_per_label": int(len(lsel)), "adam_steps": args.train_steps,
                     "loss_first_last": [round(hist[0], 5), round(hist[-1], 5)], "held_out": held,
                     "label_seconds": round(t_labels, 1), "train_seconds": round(t_train, 1)},
        "modes": {k: ("plain L-BFGS (eprecond 1)" if v is None else
                      "learned, judge_every %d, diagonal %s, %s, power %g" % (
                          v[0], "fixed per stage" if v[1] else "recomputed every iteration",
                          "plain for the rest of the stage once the plain direction is taken" if v[2]
                          else "learned again after a plain iteration", v[3]))
                  for k, v in LEARNED_MODES.items() if k in names},
        "runs": runs,
        "config": {"workload": "visco SH FWI %dx%d FDORDER8 L2, %d shots x %d steps, DTINV %d"
                               % (W["nx"], W["ny"], nshots, W["nt"], W["dtinv"]), "n_gpus": 1}}))


# (Ci, Co, level) of the U-Net's 3x3 convolutions (denise-sh_amd/csrc/dn_unet.cpp),
# level l at 1/2^l of the input resolution
UNET_LAYERS = [(16, 16, 0), (16, 16, 0), (16, 32, 1), (32, 32, 1), (32, 48, 2), (48, 48, 2), (48, 64, 3),
               (64, 64, 3), (112, 48, 2), (48, 48, 2), (80, 32, 1), (32, 32, 1), (48, 16, 0), (16, 16, 0),
               (16, 16, 0)]


def unet_run(dn, args):
    """Training steps of the learned preconditioner's U-Net at the input size
    of the 1200 x 300 workload (4 images of 304 x 1200, seeded inputs and
    targets): steps per second, and the MFMA work of the convolutions per
    second (forward, data gradient except for the first layer, weight
    gradient; padded channels counted as executed)."""
    rng = np.random.default_rng(11)
    n, H, W = 4, 304, 1200
    X = rng.standard_normal((n, 3, H, W)).astype(np.float32)
    T = rng.uniform(-3.0, 0.0, (n, H, W)).astype(np.float32)
    net = dn.UNet(seed=1)
    for _ in range(args.warmup):
        net.train_step(X, T, 1e-3)
    t0 = time.perf_counter()
    for _ in range(args.unet):
        loss = net.train_step(X, T, 1e-3)
    wall = time.perf_counter() - t0
    flops = 0.0
    for k, (ci, co, lev) in enumerate(UNET_LAYERS):
        px = n * (H >> lev) * (W >> lev)
        flops += 2.0 * 9 * ci * co * px * (2 if k == 0 else 3)
    tf = flops * args.unet / wall / 1e12
    print(json.dumps({"metric": "U-Net training steps/s", "value": round(args.unet / wall, 3), "unit": "steps/s",
                      "higher_is_better": True, "n_gpus": 1, "steps": args.unet, "warmup": args.warmup,
                      "ms_per_step": round(wall / args.unet * 1e3, 3), "dtype": "bf16 (fp32 accumulate)",
                      "conv_tflop_per_step": round(flops / 1e12, 4), "conv_tflops": round(tf, 3),
                      "mfma_peak_bf16_tflops": 2500.0, "frac_of_mfma_peak": round(tf / 2500.0, 5),
                      "loss": loss, "config": {"workload": "U-Net train step, batch %d of %dx%d" % (n, H, W)}}))


def copy_probe(nbytes=1 << 30, reps=4):
    """Plain device-to-device copy of known size (torch's copy kernel): the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters of a
    profiling run to bytes."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4", choices=["c4", "c2", "c3", "c5"])
    ap.add_argument("--compare-sequential", action="store_true",
                    help="also time the same gradient with one shot per batch (the per-shot sequential loop)")
    ap.add_argument("--nt", type=int, default=0, help="record length in steps (0: the workload's)")
    ap.add_argument("--shots", type=int, default=0, help="shots of the acquisition used (0: all)")
    ap.add_argument("--snap", default="fp32", choices=["fp32", "bf16", "fp16"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="run only the copy kernel used to scale HBM counters")
    ap.add_argument("--autotune", action="store_true", help="tune rows per segment and graph replay first")
    ap.add_argument("--wave-shots", type=int, default=1, choices=[1, 2],
                    help="shots per wavefront on the plain steps (--autotune chooses)")
    ap.add_argument("--rows", type=int, default=0, help="rows per segment of the step kernel (0: automatic)")
    ap.add_argument("--max-batch", type=int, default=0, help="most shots advanced together (0: from free HBM)")
    ap.add_argument("--layout", type=int, default=0, choices=[0, 1, 2, 4, 8],
                    help="workgroups of 8 adjacent strips of one shot (0), of 8/S strips x S shots (S = 2, 4, 8; "
                         "1 = 8)")
    ap.add_argument("--batch-order", default="length", choices=["length", "shots"],
                    help="uneven record lengths over several batches: batches by decreasing length (default) or "
                         "in shot order (the same bits)")
    ap.add_argument("--exec", default="graph", choices=["graph", "eager"],
                    help="time loop as replayed hipGraph chunks (default) or one host launch per step")
    ap.add_argument("--fwi", type=int, default=0,
                    help="instead of the gradient metric: time this many L-BFGS FWI iterations of the workload")
    ap.add_argument("--fwi-learned", type=int, default=0,
                    help="instead of the gradient metric: train the learned preconditioner, then time this many "
                         "FWI iterations plain and with each learned mode")
    ap.add_argument("--pairs", type=int, default=12, help="--fwi-learned: training pairs (4/5 train, 1/5 held out)")
    ap.add_argument("--probes", type=int, default=4, help="--fwi-learned: Hessian probes per label")
    ap.add_argument("--label-shots", type=int, default=20, help="--fwi-learned: shots of each label's gradients")
    ap.add_argument("--train-steps", type=int, default=300, help="--fwi-learned: Adam steps")
    ap.add_argument("--train-geometry", default="fixed", choices=["fixed", "mixed"],
                    help="--fwi-learned: every pair with the --label-shots acquisition (fixed) or with a random "
                         "subset of 8..label-shots shots and a record of 0.4..1 of the length (mixed)")
    ap.add_argument("--modes", default=",".join(LEARNED_MODES),
                    help="--fwi-learned: comma-separated runs (plain and learned_fixed required)")
    ap.add_argument("--train-config", default=None, choices=["c4", "c2", "c3", "c5"],
                    help="--fwi-learned: workload whose grid and shots the training pairs use (default: --config)")
    ap.add_argument("--unet", type=int, default=0,
                    help="instead of the gradient metric: time this many U-Net training steps")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe()
        return
    if args.unet:
        unet_run(_package(), args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; multi-GPU runs (--gpus %d) are not provided" % args.gpus,
              file=sys.stderr)
        sys.exit(3)

    dn = _package()
    W = workload(args.config, args.nt or None)
    nshots = len(W["src_i"]) if args.shots <= 0 else min(args.shots, len(W["src_i"]))
    prec = {"fp32": 0, "bf16": 1, "fp16": 2}[args.snap]
    if args.fwi_learned:
        fwi_learned_run(dn, W, args, nshots, prec)
        return
    ctx = make_context(dn, W, args, nshots, prec)
    # observed data from the true model (untimed setup, on the device)
    print("bench: %s, observed data from the true model" % describe(W, nshots, args.snap), file=sys.stderr,
          flush=True)
    ctx.set_model(W["vs_true"], W["rho"], W["qs"])
    obs = ctx.forward()
    ctx.set_model(W["vs_start"], W["rho"], W["qs"])
    ctx.set_observed(obs)
    del obs
    if args.fwi:
        fwi_run(ctx, W, args)
        ctx.close()
        return
    tuned = None
    if args.autotune:
        tuned = ctx.autotune(nshots, min(W["nt"], 256))
    # one step: the gradient of one batch (c4) or of all shots (c2, c3)
    if args.config == "c4":
        batch = ctx.batch_shots(True)
        batches = [(s0, min(batch, nshots - s0)) for s0 in range(0, nshots, batch)]
    else:
        batch = ctx.batch_shots(True)
        batches = [(0, nshots)]
    print("bench: %d shots per batch, %d step(s) per pass over the shots" % (batch, len(batches)), file=sys.stderr,
          flush=True)
    for k in range(args.warmup):
        ctx.gradient(*batches[k % len(batches)])
    ms = {0: 0.0, 1: 0.0}
    la = {0: 0, 1: 0}
    ce = {0: 0, 1: 0}
    l2 = 0.0
    t0 = time.perf_counter()
    for k in range(args.steps):
        l2, en = ctx.gradient(*batches[(args.warmup + k) % len(batches)])
        for p in (0, 1):
            m, l, c = ctx.pass_timing(p)
            ms[p] += m
            la[p] += l
            ce[p] += c
    wall = time.perf_counter() - t0
    cells = float(ce[0] + ce[1])
    value = cells / wall / 1e9
    # dominant kernel: the fused step kernel k_step (all but a fraction of a
    # percent of the GPU time), over the launches of both passes: the
    # forward pass moves 84 + 20/DTINV bytes per cell-update, the adjoint
    # 84 + 36/DTINV (14/DTINV and 30/DTINV with 16-bit snapshots)
    nla = max(la[0] + la[1], 1)
    per_launch_ms = (ms[0] + ms[1]) / nla
    units_per_launch = (ce[0] + ce[1]) / nla              # cell-updates per launch
    base = 12.0 * (3 + 2 * W["L"])
    b_fwd = base + (20.0 - (6.0 if prec != 0 else 0.0)) / W["dtinv"]
    b_adj = base + (36.0 - (6.0 if prec != 0 else 0.0)) / W["dtinv"]
    bytes_per_cell = (ce[0] * b_fwd + ce[1] * b_adj) / max(ce[0] + ce[1], 1)
    achieved = units_per_launch * bytes_per_cell / (per_launch_ms * 1e-3) / 1e9
    full = W["nt"] == (12000 if args.config == "c4" else 10000)
    traffic = measured_traffic(W, batch, units_per_launch) if (prec == 0 and full) else None
    out = {
        "metric": "Gcell-updates/s fwd+adj",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded %s true model, observed data modelled on the device)"
                % ("layered-lens" if args.config == "c4" else "checkerboard"),
        "config": {"workload": describe(W, nshots, args.snap), "nx": W["nx"], "ny": W["ny"], "shots": nshots,
                   "nt": W["nt"], "fdorder": 8, "L": 2, "dtinv": W["dtinv"], "shots_per_batch": batch,
                   "step": ("gradient of one %d-shot batch" % batch) if args.config == "c4"
                   else "gradient over all %d shots" % nshots},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_step<4,visco> (forward and adjoint launches of the gradient)",
                     "ms_per_launch": round(per_launch_ms, 4), "bytes_per_cell_update": round(bytes_per_cell, 3),
                     "cell_updates_per_launch": units_per_launch},
        "misfit_l2": l2,
    }
    if tuned is not None:
        shape = ctx.launch_shape()
        out["config"]["rows_per_segment"], out["config"]["graph_steps"] = shape[0], shape[1]
        out["config"]["tuned"] = {"rows_per_segment": shape[0], "graph_steps": shape[1], "wave_shots": shape[2],
                                  "group_shots": shape[3], "note": "rows 0: the cost model's choice per batch"}
    elif args.rows:
        out["config"]["rows_per_segment"] = args.rows
    if args.max_batch:
        out["config"]["max_batch"] = args.max_batch
    out["config"]["exec"] = args.exec
    out["config"]["batch_order"] = args.batch_order
    out["config"]["wave_shots"] = "tuned" if args.autotune else args.wave_shots
    out["config"]["group_layout"] = "tuned" if args.autotune else (
        "8 strips x 1 shot" if args.layout == 0 else "%d strips x %d shots" % (8 // (8 if args.layout == 1 else args.layout),
                                                                          8 if args.layout == 1 else args.layout))
    if args.compare_sequential:
        # the same gradient with one shot per batch: the per-shot sequential
        # GPU loop the batched launch replaces (the reference's shot loop,
        # denise.c:875), with the default launch shape for one shot (eight
        # strips per workgroup, one shot per wave; a layout tuned for the
        # batch could leave waves idle here)
        ctx.set_batch(1, args.rows)
        ctx.set_group_layout(0)
        ctx.set_wave_shots(1)
        t1 = time.perf_counter()
        ce_s = 0
        for k in range(max(1, args.steps)):
            ctx.gradient(*batches[(args.warmup + k) % len(batches)])
            ce_s += sum(ctx.pass_timing(p)[2] for p in (0, 1))
        wall_s = time.perf_counter() - t1
        out["sequential"] = {"value": round(ce_s / wall_s / 1e9, 3), "unit": "Gcell-updates/s",
                             "ms_per_step": round(wall_s / max(1, args.steps) * 1e3, 3),
                             "batched_over_sequential": round(value / (ce_s / wall_s / 1e9), 3)}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(W, args.cpu_seconds)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
