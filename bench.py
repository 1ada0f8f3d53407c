"""Benchmark: photon paths/s of the full-size cylindrical-transducer scene on
one MI355X, next to the CPU restatement of the reference path.

Workload (a "step"): one chunk of B photons of the 1e9-photon render of the
10 mm cube (n0 1.3333, Bessel Delta n 2e-5, 848 kHz, 20 % duty window, HG
sigma_s 1/mm g 0.9, collimated 1 mm disk source, ER step 5 um, 256x256 px x
128 path-length bins over 0-100 mm, FP32) for each of the two transducer
modes m = 0 and m = 2.  Chunk k covers photon indices [k B, (k+1) B) of the
render with weights Li/1e9, so K steps are K/(1e9/B) of the full render; the
per-photon rate is the same as the one-shot render's.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--photons-per-step B]

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "photon paths/s on 1 MI355X; adjoint vs FD gradient time at 1% rel error"
N_TOTAL = 10 ** 9
# algorithmic flops (cost model: + - x = 1, FMA = 2, div/sqrt/rcp = 1,
# transcendental = 8, J0/J1 = 20): per ER step by Bessel mode, per event
F_STEP = {0: 92, 1: 126, 2: 148}
F_EVENT = dict(scatters=80, nee_splats=127, direct_splats=100, face_events=45, photons=40)
FP32_PEAK_TFLOPS = 157.3
# HBM bytes per photon of k_render (mode 0), from rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE of one 2^24-photon launch at the bench's launch shape
# (5,846.9 and 20,943,020.8 KiB), scaled by the counters' reading for a
# 1 GiB device copy (x2.00 fetch, x1.00 write): almost all of it is the
# splat atomics.
TRAFFIC_BYTES_PER_PHOTON = 1278.97
HBM_PEAK_GBS = 8000.0
# HBM bytes per sample of the three prefilter passes on a 512^3 grid
# (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of one call, same scaling):
# y 8.07, x 8.00, z 8.03
PREFILTER_TRAFFIC_BYTES_PER_SAMPLE = 24.10


def eprint(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_flops(st, mode):
    return F_STEP[mode] * st["er_steps"] + sum(F_EVENT[k] * st[k] for k in F_EVENT)


def gpu_bench(args):
    """The timed workload.  Each mode renders its chunks on a context of its
    own (its own HIP stream), so the two modes' launches overlap: a launch's
    last waves (the photons that happen to run longest) share the GPU with
    the other mode's launch instead of leaving it idle.  Before a context
    launches its next chunk it reads the previous launch's event counters,
    which waits for that launch only."""
    import torch

    from simpler_amd import engine, scenes

    B = args.photons_per_step
    modes = [int(m) for m in args.modes.split(",")]
    ctxs = {m: engine.Context(0, use_torch_stream=(i == 0)) for i, m in enumerate(modes)}
    autotuned = None
    if args.tune:
        for c in ctxs.values():
            c.set_tuning(*[int(x) for x in args.tune.split(",")])
    work = {}
    for m in modes:
        d, tex, res = scenes.bessel_cube_full(m)
        sc = engine.Scene(ctxs[m], d, tex)
        X, Y, Z = res
        img = torch.zeros((Z, Y, X), dtype=torch.float32, device="cuda")
        work[m] = (sc, img, res)
    # the autotuner picks the launch shape on the workload itself, on a chunk
    # of up to 2^24 photons (long enough that the launch drain does not decide)
    n_tune = min(args.photons_per_step, 1 << 24) if args.autotune_photons < 0 else args.autotune_photons
    if not args.tune and n_tune > 0:
        autotuned = work[modes[0]][0].autotune(n_tune, work[modes[0]][2])
        shape = ctxs[modes[0]].get_tuning()
        for c in ctxs.values():
            c.set_tuning(shape["block_size"], shape["steps_per_chunk"], shape["blocks_per_cu"], splat=shape["splat"])
        eprint(f"autotune: {autotuned}")
    streams = {m: ctxs[m].stream() for m in modes}
    pending = {m: None for m in modes}  # (events) of the launch in flight per mode
    recs = []

    def collect(m):
        if pending[m] is None:
            return
        e0, e1, timed = pending[m]
        st = ctxs[m].read_stats()  # waits for this mode's launch only
        if timed:
            recs.append((m, st, e0, e1))
        pending[m] = None

    def step(k, timed):
        for m in modes:
            collect(m)
            sc, img, res = work[m]
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(streams[m])
            sc.render(B, res, seed=1, offset=k * B, n_total=N_TOTAL, image=img, accumulate=True, stats=False)
            e1.record(streams[m])
            pending[m] = (e0, e1, timed)

    for w in range(args.warmup):
        step(w, False)
        for m in modes:
            collect(m)
        eprint(f"warmup {w + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(args.warmup + k, True)
        eprint(f"step {k + 1}/{args.steps} launched")
    for m in modes:
        collect(m)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_mode = {}
    tot_flops = tot_ms = 0.0
    tot_steps = 0
    for m, st, e0, e1 in recs:
        ms = e0.elapsed_time(e1)
        f = algorithmic_flops(st, m)
        pm = per_mode.setdefault(m, dict(photons=0, er_steps=0, kernel_ms=0.0, flops=0.0, launches=0))
        pm["photons"] += st["photons"]
        pm["er_steps"] += st["er_steps"]
        pm["kernel_ms"] += ms
        pm["flops"] += f
        pm["launches"] += 1
        tot_flops += f
        tot_ms += ms
        tot_steps += st["er_steps"]
    photons = sum(pm["photons"] for pm in per_mode.values())
    # GPU time of the timed launches: first launch start to last launch end
    first = recs[0][2]
    span_ms = max(first.elapsed_time(e1) for _, _, _, e1 in recs)
    res = dict(wall=wall, photons=photons, er_steps=tot_steps, kernel_ms=tot_ms, flops=tot_flops, span_ms=span_ms,
               per_mode={str(m): dict(photons=pm["photons"], er_steps_per_photon=pm["er_steps"] / pm["photons"],
                                      avg_launch_ms=pm["kernel_ms"] / pm["launches"],
                                      launch_note="launch start to end on its stream; the modes' launches overlap")
                         for m, pm in per_mode.items()},
               tuning=dict(ctxs[modes[0]].get_tuning(), autotuned=autotuned is not None), launches=len(recs))
    for m in modes:
        work[m][0].close()
        ctxs[m].close()
    return res


def batch_bench(args):
    """The batched workload: 256 transducer configurations (mixed modes,
    Delta n, frequency, duty window, sigma_s, g; log-uniform photon budgets)
    rendered by one persistent launch, eager and as a replayed HIP graph."""
    import torch

    from simpler_amd import engine, scenes

    ctx = engine.Context(0)
    d, tex, cfgs, res = scenes.batched_configs(n_cfg=args.batch_configs, seed=0, scale_budgets=args.batch_scale)
    base = engine.Scene(ctx, d, tex)
    # load the kernel with a small batch of the same configurations
    _, _, small, _ = scenes.batched_configs(n_cfg=args.batch_configs, seed=0, scale_budgets=args.batch_scale * 1e-3)
    warm = engine.Batch(base, small, res)
    warm.render()
    torch.cuda.synchronize()
    warm.close()
    batch = engine.Batch(base, cfgs, res)
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, st = batch.render()
    torch.cuda.synchronize()
    eager = time.perf_counter() - t
    batch.capture()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.batch_replays):
        batch.replay()
    torch.cuda.synchronize()
    graph = (time.perf_counter() - t) / args.batch_replays
    eprint(f"batch: {batch.n_cfg} configurations, {batch.photons} photons, eager {eager:.2f}s, graph {graph:.2f}s")
    out = dict(configurations=batch.n_cfg, photons=batch.photons, resolution=list(res),
               photon_paths_per_s_eager=batch.photons / eager, photon_paths_per_s_graph=batch.photons / graph,
               seconds_eager=eager, seconds_graph=graph, er_steps_per_photon=st["er_steps"] / st["photons"],
               budget_scale=args.batch_scale)
    batch.close()
    ctx.close()
    return out


def adjoint_bench(args):
    """Differentiable spline field: 64^3 coefficients on [-5, 5]^3 mm (n0 +
    seeded noise, Gaussian-filtered sigma = 4 cells, max |dn| 5e-5), sigma_s
    2/mm, g 0.9, 5 um step, 64x64x32 sensor.  Objective: throughput into a
    1 mm aperture.  Times the forward render and the path-replay adjoint
    (gradient w.r.t. all 262,144 coefficients) on the same photons; the
    forward render also with every next-event sub-path traced (the next-event
    skip off: same image and splat counts, more ER steps).

    Accuracy: central differences with common random numbers on K
    coefficients drawn uniformly at random from those whose gradient is at
    least 2 % of the largest, on the same scene with index-matched faces and
    no time gate (a small coefficient change then flips no Fresnel or gate
    decision, which the adjoint holds fixed and a central difference would
    pick up as noise)."""
    import numpy as np
    import torch

    from simpler_amd import engine, scenes

    ctx = engine.Context(0)
    ctx.set_adjoint_schedule(args.adjoint_schedule)
    d, tex, res, (coeff, dims, lo, hi) = scenes.spline_cube_ready(n=64)
    sc = engine.Scene(ctx, d, tex, spline=(coeff, dims, lo, hi))
    G = torch.tensor(scenes.aperture_weights(res), device="cuda")
    N = args.adjoint_photons
    stream = torch.cuda.current_stream()

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        out = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return out, e0.elapsed_time(e1) / 1e3

    sc.render(4096, res, seed=11)
    sc.adjoint_render(G, 4096, seed=11)
    (img, fst), t_fwd = timed(lambda: sc.render(N, res, seed=11))
    F_img = float((img.double() * G.double()).sum())
    ctx.set_next_event_skip(False)
    (img_full, fst_full), t_full = timed(lambda: sc.render(N, res, seed=11))
    ctx.set_next_event_skip(True)
    skip = dict(forward_full_trace_s=t_full, forward_s=t_fwd, speedup=t_full / t_fwd,
                er_steps=fst["er_steps"], er_steps_full_trace=fst_full["er_steps"],
                nee_splats=fst["nee_splats"], nee_splats_full_trace=fst_full["nee_splats"],
                image_max_abs_diff=float((img - img_full).abs().max()), image_max=float(img_full.abs().max()))
    del img_full
    eprint(f"next-event skip: {skip}")
    (grad, F, ast), t_adj = timed(lambda: sc.adjoint_render(G, N, seed=11))
    eprint(f"adjoint: {N} paths, forward {t_fwd:.2f}s, adjoint {t_adj:.2f}s, F {F:.6g} (render {F_img:.6g})")
    sc.close()
    # accuracy check on the index-matched scene, at its own photon count
    dm = d.copy()
    dm.set(ior=1.0, pathlength_range=(-1.0, -1.0))
    scm = engine.Scene(ctx, dm, tex, spline=(coeff, dims, lo, hi))
    res1 = (res[0], res[1], 1)
    G1 = G[:1].contiguous()
    Nf = args.adjoint_fd_photons
    gm, _, _ = scm.adjoint_render(G1, Nf, seed=11)
    gm = gm.cpu().numpy().astype(np.float64)
    am = np.abs(gm)
    cand = np.nonzero(am >= 0.02 * am.max())[0]
    picks = np.random.default_rng(args.adjoint_fd_seed).choice(cand, args.adjoint_fd, replace=False)

    def objective(c):
        scm.set_spline(c, dims, lo, hi)
        (im, _), t = timed(lambda: scm.render(Nf, res1, seed=11, stats=False))
        return float((im.double() * G1.double()).sum()), t

    # the objective at c_j + k eps, k = -3..3 (the unperturbed render shared);
    # of the six one-step differences those more than 25 % from their median
    # are dropped (a step that carried a photon across a path discontinuity:
    # a scattering frame switch, the next-event cut-off plane) and the rest
    # averaged
    F0, t_fd = objective(coeff)
    errs = []
    for j in picks:
        xs, Fs = [], []
        for k in (-3, -2, -1, 0, 1, 2, 3):
            c = coeff.copy()
            c[j] = np.float32(coeff[j] + k * args.adjoint_eps)
            xs.append(float(c[j]))
            if k == 0:
                Fs.append(F0)
            else:
                Fk, tk = objective(c)
                Fs.append(Fk)
                t_fd += tk
        steps = np.diff(Fs) / np.diff(xs)
        med = np.median(steps)
        fd = float(steps[np.abs(steps - med) <= 0.25 * abs(med)].mean())
        errs.append(abs(gm[j] - fd) / abs(fd) if fd != 0 else float("inf"))
        eprint(f"  coefficient {j}: adjoint {gm[j]:.5g}  difference {fd:.5g}  rel err {errs[-1]:.3g}")
    scm.close()
    n_par = len(gm)
    # one central difference at the timing run's photon count costs two
    # forward renders of N paths
    per_fd = 2.0 * t_fwd
    out = dict(workload=("64^3 spline field, sigma_s 2/mm, g 0.9, 5 um step, 64x64x32 sensor, "
                         "throughput into a 1 mm aperture"),
               schedule=("deferred: re-trace queueing the backward runs, then full-wave sub-path and main-path "
                         "kernels" if ctx.get_adjoint_schedule()["schedule"] == 1 else "inline"),
               paths=N, parameters=n_par, forward_s=t_fwd, adjoint_s=t_adj, adjoint_over_forward=t_adj / t_fwd,
               objective=F, objective_render=F_img, forward_paths_per_s=N / t_fwd, adjoint_paths_per_s=N / t_adj,
               fd_check=dict(scene="same field and objective, index-matched faces, no time gate",
                             paths=Nf, candidates=int(len(cand)), coefficients=[int(j) for j in picks],
                             coefficient_rel_err=errs, max_rel_err=float(max(errs)),
                             median_rel_err=float(np.median(errs)),
                             within_2pct=float(np.mean(np.array(errs) <= 0.02)),
                             coefficient_eps=args.adjoint_eps, points=7,
                             estimator="mean of the one-step differences within 25 % of their median",
                             seconds=t_fd),
               fd_full_gradient_s=per_fd * n_par, adjoint_speedup_over_fd=per_fd * n_par / t_adj,
               adjoint_stats=ast, forward_er_steps=fst["er_steps"], next_event_skip=skip,
               adjoint_over_full_trace_forward=t_adj / t_full)
    ctx.close()
    return out


def patterns_bench(args):
    """Travelling-wave patterns: 16 point sources on a 10 mm ring (f 1 MHz,
    a_i ~ U[0, 5e-9] m, phi_i ~ U[0, 2 pi)), sigma_s 2/mm, g 0.95, 64x64x32
    sensor.  P patterns in one launch; each Adam step runs the batched
    path-replay adjoint (objective and gradient of every pattern's 1 mm
    aperture throughput w.r.t. its 16 amplitudes and 16 phases) and ascends."""
    import numpy as np
    import torch

    from simpler_amd import engine, scenes

    ctx = engine.Context(0)
    d, tex, _, src = scenes.travelling_wave_cube()
    res = (64, 64, 32)
    base = engine.Scene(ctx, d, tex, sources=src)
    P, n, N = args.patterns, args.pattern_photons, len(src["amp"])
    rng = np.random.default_rng(7)
    x = torch.tensor(np.concatenate([rng.uniform(0, 1, (P, N)), rng.uniform(0, 2 * np.pi, (P, N))], axis=1),
                     dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([x], lr=0.05)
    G = torch.tensor(scenes.aperture_weights(res), device="cuda").expand(P, -1, -1, -1).contiguous()
    seeds = [scenes.splitmix64(11 * 1000003 + c) for c in range(P)]
    hist, times, rev = [], [], 0
    for step in range(args.adam_steps):
        xa = x.detach().numpy()
        b = engine.Batch.patterns(base, xa[:, :N] * 5e-9, xa[:, N:], [n] * P, seeds, res)
        torch.cuda.synchronize()
        t = time.perf_counter()
        grads, F, st = b.adjoint(G)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        b.close()
        g = grads.double().cpu()
        opt.zero_grad()
        x.grad = -torch.cat([g[:, :N] * 5e-9, g[:, N:2 * N]], dim=1)  # ascend the throughput
        opt.step()
        with torch.no_grad():
            x[:, :N].clamp_(min=0.0)
        hist.append(float(np.mean(F)))
        rev += st["reverse_steps"]
        eprint(f"patterns: Adam step {step + 1}/{args.adam_steps}: mean F {hist[-1]:.6g}, {times[-1]:.2f}s")
    base.close()
    ctx.close()
    per_step = float(np.median(times))
    rate = P * n / per_step
    return dict(workload="16-element travelling-wave ring, f 1 MHz, sigma_s 2/mm, g 0.95, 64x64x32 sensor, "
                         "1 mm aperture throughput",
                patterns=P, paths_per_pattern=n, parameters_per_pattern=2 * N, adam_steps=args.adam_steps,
                seconds_per_adam_step=per_step, adjoint_paths_per_s=rate, mean_objective=hist,
                full_size_projection_s=(64 * 1e7 / rate) * 50,
                full_size="64 patterns x 1e7 paths x 50 Adam steps at the measured rate")


def prefilter_bench(args):
    """GPU B-spline prefilter of an n^3 grid of node values (in place, three
    passes y, x, z): HIP-event time per call over the timed calls, and the
    rate against 24 B per sample (each pass reads and writes every sample
    once).  CPU: the oracle's FP64 restatement of the reference's build3d on
    a 256^3 grid, one core."""
    import numpy as np
    import torch

    from oracle import oracle as orc
    from simpler_amd import engine

    ctx = engine.Context(0)
    n = args.prefilter_n
    dims = (n, n, n)
    gen = torch.Generator(device="cuda").manual_seed(0)
    src = 1.3333 + 1e-4 * torch.randn(n ** 3, device="cuda", generator=gen)
    buf = src.clone()
    ctx.spline_prefilter(buf, dims)  # warmup
    times = []
    for _ in range(args.prefilter_reps):
        buf.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        ctx.spline_prefilter(buf, dims)
        e1.record()
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    ms = sum(times) / len(times)
    algo_bytes = 24 * n ** 3
    gbs = algo_bytes / (ms / 1e3) / 1e9
    m = 256
    grid = 1.3333 + 1e-4 * np.random.default_rng(0).standard_normal(m ** 3)
    t = time.perf_counter()
    orc.spline_build3d(grid, [m, m, m])
    cpu_s = time.perf_counter() - t
    eprint(f"prefilter {n}^3: {ms:.3f} ms per call, {gbs:.0f} GB/s algorithmic; CPU {m}^3 in {cpu_s:.2f} s")
    ctx.close()
    return dict(grid=list(dims), ms_per_call=ms, calls=len(times), algorithmic_bytes_per_sample=24,
                roofline={"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": gbs / HBM_PEAK_GBS, "traffic": PREFILTER_TRAFFIC_BYTES_PER_SAMPLE * n ** 3
                          if PREFILTER_TRAFFIC_BYTES_PER_SAMPLE else None},
                samples_per_s=n ** 3 / (ms / 1e3),
                cpu_baseline=dict(value=m ** 3 / cpu_s, unit="samples/s", cores=1, kind="port",
                                  sample=f"oracle build3d (FP64) of a {m}^3 grid, {cpu_s:.2f} s"))


def host_info():
    """The host the CPU baseline ran on: logical CPUs, the CPUs this process
    may use, and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)), model=model)


def cpu_baseline(args):
    """The CPU restatement of the reference path (4 field evaluations per ER
    step, glibc jn, FP32, the reference's -O3 -ffast-math -fopenmp flags) on
    this host, one photon per OpenMP work item.

    The per-photon cost of the reference algorithm is heavy-tailed (a
    next-event ray trapped by total internal reflection can march up to the
    10 km LargeDist), so the wall time of a bounded sample is set by its
    longest photon while the other threads idle.  The rate reported is the
    sample's photons over its CPU-seconds per thread (photons x threads /
    summed per-thread tracing time), the rate a long render keeps while every
    thread is busy; the sample's own wall-clock rate is reported beside it."""
    from oracle import oracle as orc
    from simpler_amd import scenes

    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    host = host_info()
    d, tex, res = scenes.bessel_cube_full(0)
    sc = orc.OracleScene(d, tex, fast=True)
    n = args.cpu_photons or 128 * threads
    t = time.perf_counter()
    _, st = sc.render(n, res, seed=1, offset=0, n_total=N_TOTAL, precision=32, threads=threads, cutoff=False)
    dt = time.perf_counter() - t
    busy = st["busy_seconds"]
    eprint(f"cpu baseline: {n} photons in {dt:.1f}s wall, {busy:.1f} CPU-s on {threads} threads")
    # the same CPU code with the dead-path cutoff the GPU uses: three
    # disjoint photon blocks, median rate
    nc = args.cpu_photons_cutoff or 128 * threads
    cut = []
    for k in range(3):
        t = time.perf_counter()
        _, stc = sc.render(nc, res, seed=1, offset=n + k * nc, n_total=N_TOTAL, precision=32, threads=threads,
                           cutoff=True)
        w = time.perf_counter() - t
        cut.append(dict(rate=nc * threads / stc["busy_seconds"], wall_rate=nc / w, seconds=w,
                        er_steps=stc["er_steps"], busy_seconds=stc["busy_seconds"]))
    cut.sort(key=lambda c: c["rate"])
    med = cut[1]
    eprint(f"cpu with dead-path cutoff: median {med['rate']:.4g} paths/s over 3 blocks of {nc}")
    return dict(value=n * threads / busy, unit="photon paths/s", cores=threads, kind="port",
                sample=(f"photons 0-{n - 1} (seed 1) of the mode-0 1e9-photon render of the 5 um / 256x256x128 "
                        f"scene, reference algorithm in full (4 field evaluations per ER step, glibc jn, no "
                        f"dead-path cutoff), FP32, g++ -O3 -ffast-math -fopenmp, {threads} threads, one photon per "
                        f"work item; {st['er_steps']} ER steps, {busy:.1f} CPU-s, {dt:.1f} s wall. value = photons x "
                        f"threads / CPU-s (all threads busy); wall_rate = photons / wall of this sample"),
                wall_rate=n / dt, single_thread_rate=n / busy, er_steps_per_s_per_thread=st["er_steps"] / busy,
                er_steps_per_photon=st["er_steps"] / n, host=host,
                same_algorithm_as_gpu=dict(value=med["rate"], photons_per_block=nc, blocks=3,
                                           wall_rate=med["wall_rate"], seconds=med["seconds"],
                                           er_steps_per_s_per_thread=med["er_steps"] / med["busy_seconds"],
                                           er_steps_per_photon=med["er_steps"] / nc,
                                           note="same CPU code with the dead-path cutoff the GPU uses; median of 3 "
                                                "disjoint photon blocks"))


def profile_prefilter(args):
    """One in-place prefilter of an n^3 grid (for hardware-counter runs)."""
    import torch

    from simpler_amd import engine
    ctx = engine.Context(0)
    n = args.prefilter_n
    buf = 1.3333 + 1e-4 * torch.randn(n ** 3, device="cuda")
    ctx.spline_prefilter(buf, (n, n, n))
    torch.cuda.synchronize()
    print(json.dumps({"profile": "prefilter", "grid": [n, n, n]}), flush=True)
    ctx.close()


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device copy of known size, for scaling rocprof's FETCH_SIZE /
    WRITE_SIZE counters to bytes (run under rocprofv3 --pmc)."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}), flush=True)


def profile_workload(args):
    """One launch of a named workload, for hardware-counter runs:
    'spline' = the differentiable-spline forward render, 'adjoint' = its
    path-replay adjoint (--profile-photons photons, under
    --adjoint-schedule)."""
    import torch

    from simpler_amd import engine, scenes

    ctx = engine.Context(0)
    d, tex, res, spl = scenes.spline_cube_ready(n=64)
    sc = engine.Scene(ctx, d, tex, spline=spl)
    n = args.profile_photons
    ctx.set_adjoint_schedule(args.adjoint_schedule)
    if args.profile == "spline":
        _, st = sc.render(n, res, seed=11)
    else:
        G = torch.tensor(scenes.aperture_weights(res), device="cuda")
        _, _, st = sc.adjoint_render(G, n, seed=11)
    torch.cuda.synchronize()
    print(json.dumps({"profile": args.profile, "photons": n, "stats": st}), flush=True)
    sc.close()
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons-per-step", type=int, default=1 << 25)
    ap.add_argument("--modes", default="0,2")
    ap.add_argument("--tune", default="", help="block_size,steps_per_chunk,blocks_per_cu (skips the autotuner)")
    ap.add_argument("--autotune-photons", type=int, default=-1,
                    help="photons per autotuner trial (default: the step's chunk size, at most 2^24; 0: no autotuning)")
    ap.add_argument("--cpu-photons", type=int, default=0, help="reference-algorithm sample (default 128 x threads)")
    ap.add_argument("--cpu-photons-cutoff", type=int, default=0, help="cutoff blocks (default 128 x threads)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-prefilter", action="store_true")
    ap.add_argument("--prefilter-n", type=int, default=512)
    ap.add_argument("--prefilter-reps", type=int, default=5)
    ap.add_argument("--profile", choices=["spline", "adjoint", "prefilter"], default="",
                    help="only run one launch of this workload (for hardware-counter runs)")
    ap.add_argument("--profile-photons", type=int, default=1 << 18)
    ap.add_argument("--adjoint-schedule", type=int, default=1, choices=[0, 1],
                    help="image adjoint: 0 backward runs inline, 1 deferred to full-wave kernels")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (for scaling rocprof byte counters)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-configs", type=int, default=256)
    ap.add_argument("--batch-scale", type=float, default=0.2,
                    help="photon budgets of the batched sweep times this factor (1.0: 5.5e8 photons)")
    ap.add_argument("--batch-replays", type=int, default=1)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-patterns", action="store_true")
    ap.add_argument("--patterns", type=int, default=64)
    ap.add_argument("--pattern-photons", type=int, default=1 << 16)
    ap.add_argument("--adam-steps", type=int, default=2)
    ap.add_argument("--adjoint-photons", type=int, default=1 << 22)
    ap.add_argument("--adjoint-fd", type=int, default=8, help="coefficients checked by central differences")
    ap.add_argument("--adjoint-fd-photons", type=int, default=1 << 18)
    ap.add_argument("--adjoint-fd-seed", type=int, default=5)
    ap.add_argument("--adjoint-eps", type=float, default=5e-4)
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.profile == "prefilter":
        profile_prefilter(args)
        return
    if args.profile:
        profile_workload(args)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for {max(args.gpus, world)}); photon shards would be "
              f"rendered independently per GPU and the images reduced once at the end.", file=sys.stderr)
        sys.exit(2)
    g = gpu_bench(args)
    value = g["photons"] / g["wall"]
    # the render kernel keeps the GPU busy over the timed launches (the two
    # modes' launches overlap), so its rate is the algorithmic flops of all
    # timed launches over their GPU time (first start to last end, HIP events)
    achieved = g["flops"] / (g["span_ms"] / 1e3) / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "photon paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": g["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("10 mm cube, Bessel transducer modes 0 and 2, dn 2e-5, 848 kHz, 20% duty, sigma_s 1/mm, "
                         "g 0.9, collimated 1 mm disk, ER step 5 um, 256x256x128 sensor over 0-100 mm, "
                         "chunks of the 1e9-photon render"),
            "photons_per_step_per_mode": args.photons_per_step,
            "modes": [int(m) for m in args.modes.split(",")],
            "n_total": N_TOTAL,
            "seed": 1,
        },
        "er_steps_per_s": g["er_steps"] / g["wall"],
        "per_mode": g["per_mode"],
        "tuning": g["tuning"],
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": TRAFFIC_BYTES_PER_PHOTON * args.photons_per_step,
                     "kernel": "k_render<Bessel> (persistent march + event kernel), modes 0 and 2 on two streams",
                     "gpu_time_ms": g["span_ms"], "launches": g["launches"],
                     "avg_launch_ms": g["kernel_ms"] / g["launches"]},
    }
    if not args.no_batch:
        out["batched"] = batch_bench(args)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench(args)
    if not args.no_patterns:
        out["patterns"] = patterns_bench(args)
    if not args.no_prefilter:
        out["prefilter"] = prefilter_bench(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
