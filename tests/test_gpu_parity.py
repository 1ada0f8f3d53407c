"""Parity of the gfx950 engine (through the C ABI) with the CPU oracle.

* field:   n and grad n at random points vs the FP64 oracle (all field kinds);
* per ray: transparent-medium exit position / direction / optical length vs
           the FP64 oracle, each within 1e-5 relative;
* image:   every bin within 3 sigma of the FP32 oracle render on the same
           inputs (sigma from the spread over photon batches), total
           throughput within 3 sigma, same event counts; an independent-seed
           comparison for the statistics themselves;
* properties at full size: chunked renders add up, renders are
           deterministic, the dead-path cutoff leaves the image unchanged.
"""
import contextlib
import math

import numpy as np
import pytest
import torch

from conftest import gpu_threads
from oracle import oracle as O
from simpler_amd import engine, scenes
from simpler_amd._abi import FLAG_GENERAL_AXIS, FLAG_NO_DEAD_CUT

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


@contextlib.contextmanager
def full_trace(ctx):
    """Forward renders that trace every next-event sub-path, as the oracle
    does: for comparing the work counters (er_steps, face_events), which
    count the sub-paths the next-event skip leaves untraced."""
    ctx.set_next_event_skip(False)
    try:
        yield
    finally:
        ctx.set_next_event_skip(True)


def _pts(n=2000, seed=0, half=0.005):
    return np.random.default_rng(seed).uniform(-half, half, (n, 3))


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("axis,general", [("z", True), ("x", True), ("x", False)])
def test_bessel_field_matches_oracle(gpu_ctx, mode, axis, general):
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-5, mode=mode, axis=axis)
    d.flags = (d.flags | FLAG_GENERAL_AXIS) if general else (d.flags & ~FLAG_GENERAL_AXIS)
    pts = _pts(seed=mode)
    n_o, g_o = O.OracleScene(d, tex).field(pts, scaling=0.7)
    n_g, g_g = engine.Scene(gpu_ctx, d, tex).field(torch.tensor(pts, dtype=torch.float32, device="cuda"), 0.7)
    n_g, g_g = n_g.double().cpu().numpy(), g_g.double().cpu().numpy()
    log(f"bessel m={mode} {axis} general={general}: dn {np.abs(n_g - n_o).max():.2e} "
        f"dg {np.abs(g_g - g_o).max() / np.abs(g_o).max():.2e}")
    assert np.abs(n_g - n_o).max() < 2e-7  # ~1 ulp of n = 1.333
    assert np.abs(g_g - g_o).max() < 1e-5 * np.abs(g_o).max()


def test_sources_field_matches_oracle(gpu_ctx):
    d, tex, _, src = scenes.travelling_wave_cube()
    pts = _pts(seed=7, half=0.0049)
    n_o, g_o = O.OracleScene(d, tex, sources=src).field(pts, tau=1.3)
    n_g, g_g = engine.Scene(gpu_ctx, d, tex, sources=src).field(
        torch.tensor(pts, dtype=torch.float32, device="cuda"), 1.0, 1.3)
    n_g, g_g = n_g.double().cpu().numpy(), g_g.double().cpu().numpy()
    assert np.abs(n_g - n_o).max() < 2e-7
    assert np.abs(g_g - g_o).max() < 1e-4 * np.abs(g_o).max()


def test_spline_field_matches_oracle(gpu_ctx):
    d, tex, _, (vals, dims, lo, hi) = scenes.spline_cube(n=24)
    coeff = O.spline_build3d(vals.reshape(-1), dims)
    spl = (coeff.astype(np.float32), dims, lo, hi)
    pts = _pts(seed=9, half=0.0049)
    n_o, g_o = O.OracleScene(d, tex, spline=spl).field(pts)
    n_g, g_g = engine.Scene(gpu_ctx, d, tex, spline=spl).field(torch.tensor(pts, dtype=torch.float32, device="cuda"))
    n_g, g_g = n_g.double().cpu().numpy(), g_g.double().cpu().numpy()
    assert np.abs(n_g - n_o).max() < 4e-7
    assert np.abs(g_g - g_o).max() < 1e-4 * np.abs(g_o).max()


@pytest.mark.parametrize("mode", [0, 2])
def test_transparent_per_ray_exit_matches_fp64(gpu_ctx, mode):
    """configs: sigma_s = 0, per-ray exit p, d, L within 1e-5 relative."""
    d, tex, _ = scenes.transparent_cube(mode=mode)
    n = 20000
    p, dr, L, s = engine.Scene(gpu_ctx, d, tex).trace_rays(n, seed=1)
    p, dr, L, s = (t.double().cpu().numpy() for t in (p, dr, L, s))
    p64, d64, L64, s64 = O.OracleScene(d, tex).trace_rays(n, seed=1, precision=64, threads=gpu_threads())
    assert np.array_equal(s, s64) and (s == 1).all()
    half = 0.005
    ep = np.abs(p - p64).max() / half
    ed = np.abs(dr - d64).max()
    eL = (np.abs(L - L64) / L64).max()
    log(f"mode {mode}: pos {ep:.2e} dir {ed:.2e} L {eL:.2e}")
    assert ep < 1e-5 and ed < 1e-5 and eL < 1e-5


def test_full_size_reference_sources_per_ray(gpu_ctx):
    """The point-source field at the reference's own size (fus.json: 10,200
    sources in two arc lattices, transparent slab): per-ray exit position,
    direction and optical length against the FP64 oracle within 1e-5
    relative, and the field's n and grad n at random points."""
    import time
    d, tex, _, src = scenes.fus_sources_scene()
    assert len(src["amp"]) == 10200
    sc = engine.Scene(gpu_ctx, d, tex, sources=src)
    osc = O.OracleScene(d, tex, sources=src)
    pts = _pts(n=256, seed=4, half=0.006)
    n_o, g_o = osc.field(pts)
    n_g, g_g = sc.field(torch.tensor(pts, dtype=torch.float32, device="cuda"))
    n_g, g_g = n_g.double().cpu().numpy(), g_g.double().cpu().numpy()
    assert np.abs(n_g - n_o).max() < 4e-7
    assert np.abs(g_g - g_o).max() < 1e-4 * np.abs(g_o).max()
    n = 48
    p, dr, L, s = (t.double().cpu().numpy() for t in sc.trace_rays(n, seed=3))
    p64, d64, L64, s64 = osc.trace_rays(n, seed=3, precision=64, threads=gpu_threads())
    assert np.array_equal(s, s64) and (s == 1).all()
    ep = np.abs(p - p64).max() / 0.00635
    ed = np.abs(dr - d64).max()
    eL = (np.abs(L - L64) / L64).max()
    # the GPU's rate on the same scene, for the log
    torch.cuda.synchronize()
    t = time.perf_counter()
    sc.trace_rays(1 << 16, seed=5)
    torch.cuda.synchronize()
    log(f"10,200 sources: pos {ep:.2e} dir {ed:.2e} L {eL:.2e}; GPU {(1 << 16) / (time.perf_counter() - t):.3g} rays/s")
    assert ep < 1e-5 and ed < 1e-5 and eL < 1e-5


def _batches(render, B):
    return np.stack([render(b) for b in range(B)])


def _bin_check(g, o):
    """g, o: (B, Z, Y, X) batch images.  Every bin within 3 sigma (sigma of
    the difference of the totals from the batch spread) and the totals within
    3 sigma."""
    B = g.shape[0]
    Tg, To = g.sum(0), o.sum(0)
    var = B * (g.var(0, ddof=1) + o.var(0, ddof=1))
    sig = np.sqrt(var)
    scale = max(Tg.max(), To.max())
    bad = np.abs(Tg - To) > 3 * sig + 1e-6 * scale
    tot_sig = math.sqrt(B * (g.sum((1, 2, 3)).var(ddof=1) + o.sum((1, 2, 3)).var(ddof=1)))
    return int(bad.sum()), abs(Tg.sum() - To.sum()), tot_sig, int(((Tg > 0) | (To > 0)).sum())


def _image_parity(gpu_ctx, d, tex, res, seed=3, what="image"):
    gs = engine.Scene(gpu_ctx, d, tex)
    osc = O.OracleScene(d, tex)
    B, n = 8, 48
    N = B * n
    g_st, o_st = [], []

    def rg(b):
        img, st = gs.render(n, res, seed=seed, offset=b * n, n_total=N)
        g_st.append(st)
        return img.double().cpu().numpy()

    def ro(b):
        img, st = osc.render(n, res, seed=seed, offset=b * n, n_total=N, precision=32, threads=gpu_threads(),
                             cutoff=True)
        o_st.append(st)
        return img

    g, o = _batches(rg, B), _batches(ro, B)
    bad, dtot, tot_sig, nbins = _bin_check(g, o)
    log(f"{what}: bins {nbins} bad {bad} total diff {dtot:.4g} sigma {tot_sig:.4g}")
    assert nbins > 0
    assert bad == 0
    assert dtot <= 3 * tot_sig
    for k in ("photons", "scatters", "direct_splats"):
        assert sum(s[k] for s in g_st) == sum(s[k] for s in o_st), k
    ng, no = sum(s["nee_splats"] for s in g_st), sum(s["nee_splats"] for s in o_st)
    assert abs(ng - no) <= 0.002 * no + 2


@pytest.mark.parametrize("mode", [0, 2])
def test_coordinate_frame_kernel_equals_general_axis_kernel(gpu_ctx, mode):
    """Single-scene renders specialise the Bessel evaluation for the axis
    frame uz = +z, ux = +x (rho = (q.x, q.y, 0)).  Against the general-axis
    kernel on the same field written in the frame ux = -x, uy = -y (for even
    m the same field: 2 phi shifts by 2 pi): field values to an ulp, and with
    the same photons the same images and step counts up to float rounding of
    the field."""
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-5, mode=mode)
    d2 = d.copy()
    d2.set(axis_ux=(-1.0, 0.0, 0.0))
    sc1, sc2 = engine.Scene(gpu_ctx, d, tex), engine.Scene(gpu_ctx, d2, tex)
    pts = torch.tensor(_pts(seed=7), dtype=torch.float32, device="cuda")
    (n1, g1), (n2, g2) = sc1.field(pts, 0.7), sc2.field(pts, 0.7)
    assert (n1 - n2).abs().max().item() < 2e-7
    assert (g1 - g2).abs().max().item() < 1e-5 * g1.abs().max().item()
    res = (32, 32, 16)
    (a, s1), (b, s2) = sc1.render(1 << 15, res, seed=3), sc2.render(1 << 15, res, seed=3)
    a, b = a.double().cpu().numpy(), b.double().cpu().numpy()
    log(f"frame kernel m={mode}: L1 {np.abs(a - b).sum() / a.sum():.2e}, steps {s1['er_steps']} {s2['er_steps']}")
    assert a.sum() > 0
    assert abs(a.sum() - b.sum()) <= 1e-4 * a.sum()
    assert np.abs(a - b).sum() <= 2e-3 * a.sum()
    assert abs(s1["er_steps"] - s2["er_steps"]) <= 1e-4 * s2["er_steps"]
    assert s1["photons"] == s2["photons"]
    assert abs(s1["scatters"] - s2["scatters"]) <= 1e-3 * s2["scatters"] + 2


def test_render_every_bin_within_3_sigma_of_oracle(gpu_ctx):
    d, tex, res = scenes.bessel_cube_small()
    _image_parity(gpu_ctx, d, tex, res)


@pytest.mark.parametrize("lenses", ["emitter", "sensor", "emitter_cone"])
def test_lensed_render_within_3_sigma_of_oracle(gpu_ctx, lenses):
    """Thin lenses (lens.cpp:19-33,69-97): the emitter lens focuses the
    collimated beam 15 mm behind the source face, or a 10-degree cone; the
    sensor lens refocuses the exit rays (and stops those outside its 3 mm
    aperture)."""
    d, tex, res = scenes.bessel_cube_small()
    if lenses.startswith("emitter"):
        d.set(emitter_lens_active=1, emitter_lens_focal_length=0.015, emitter_lens_aperture=0.0015)
    if lenses == "emitter_cone":
        d.set(half_theta_limit=math.radians(10.0))
    if lenses == "sensor":
        d.set(sensor_lens_active=1, sensor_lens_focal_length=0.02, sensor_lens_aperture=0.003)
    _image_parity(gpu_ctx, d, tex, res, seed=17, what=lenses)


def test_render_statistics_with_independent_seeds(gpu_ctx):
    d, tex, res = scenes.bessel_cube(step=1e-5, n_max=1e-5, mode=2, res=(16, 16, 8))
    gs = engine.Scene(gpu_ctx, d, tex)
    osc = O.OracleScene(d, tex)
    B, n = 8, 64
    g = _batches(lambda b: gs.render(n, res, seed=101, offset=b * n, n_total=B * n)[0].double().cpu().numpy(), B)
    o = _batches(lambda b: osc.render(n, res, seed=202, offset=b * n, n_total=B * n, precision=32,
                                      threads=gpu_threads(), cutoff=True)[0], B)
    bad, dtot, tot_sig, nbins = _bin_check(g, o)
    log(f"independent seeds: bins {nbins} beyond 3 sigma {bad} total diff {dtot:.4g} sigma {tot_sig:.4g}")
    assert dtot <= 3 * tot_sig
    assert bad <= max(3, 0.03 * nbins)


def test_chunked_render_adds_up_on_gpu(gpu_ctx):
    d, tex, res = scenes.bessel_cube_full(mode=2)
    gs = engine.Scene(gpu_ctx, d, tex)
    N = 1 << 15
    full, st = gs.render(N, res, seed=5, n_total=N)
    full = full.double().cpu().numpy()
    img = torch.zeros((res[2], res[1], res[0]), dtype=torch.float32, device="cuda")
    gs.render(N // 4, res, seed=5, offset=0, n_total=N, image=img)
    for k in range(1, 4):
        gs.render(N // 4, res, seed=5, offset=k * (N // 4), n_total=N, image=img, accumulate=True)
    chunks = img.double().cpu().numpy()
    assert abs(chunks.sum() - full.sum()) < 1e-5 * full.sum()
    assert np.abs(chunks - full).max() < 1e-4 * full.max()
    again, st2 = gs.render(N, res, seed=5, n_total=N)
    assert st == st2  # counters are exactly reproducible
    assert np.abs(again.double().cpu().numpy() - full).max() < 1e-4 * full.max()


def test_stream_switch_is_ordered(gpu_ctx):
    """A context moved to another stream while its previous render is still
    running: the second render must wait for the first (they share the work
    queue and the counters), so two half renders on two streams add up to the
    one-shot render."""
    d, tex, res = scenes.bessel_cube(step=1e-5, n_max=1e-5, mode=0, res=(32, 32, 16))
    gs = engine.Scene(gpu_ctx, d, tex)
    N = 1 << 16
    ref, st_ref = gs.render(N, res, seed=19, n_total=N)
    ref = ref.double().cpu().numpy()
    img = torch.zeros((res[2], res[1], res[0]), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    with torch.cuda.stream(sa):
        gs.render(N // 2, res, seed=19, offset=0, n_total=N, image=img, accumulate=True, stats=False)
    with torch.cuda.stream(sb):
        gs.render(N // 2, res, seed=19, offset=N // 2, n_total=N, image=img, accumulate=True, stats=False)
        st = gpu_ctx.read_stats()
    torch.cuda.synchronize()
    two = img.double().cpu().numpy()
    log(f"stream switch: photons of the second launch {st['photons']}, totals {two.sum():.6g} / {ref.sum():.6g}")
    assert st["photons"] == N // 2
    assert abs(two.sum() - ref.sum()) < 1e-5 * ref.sum()
    assert np.abs(two - ref).max() < 1e-4 * ref.max()


@pytest.mark.parametrize("mode", [0, 2])
def test_wave_aggregated_splat_equals_per_lane_splat(gpu_ctx, mode):
    """The two splat strategies (per-lane atomics, wave-aggregated atomics)
    render the same photons into the same image, up to the order of the
    float additions."""
    d, tex, res = scenes.bessel_cube_full(mode)
    sc = engine.Scene(gpu_ctx, d, tex)
    tun = gpu_ctx.get_tuning()
    n = 1 << 17
    imgs, sts = [], []
    for strategy in (0, 1):
        gpu_ctx.set_tuning(splat=strategy)
        img, st = sc.render(n, res, seed=21)
        imgs.append(img.double().cpu().numpy())
        sts.append(st)
    gpu_ctx.set_tuning(splat=tun["splat"])
    a, b = imgs
    log(f"splat strategies m={mode}: totals {a.sum():.7g} / {b.sum():.7g}, max diff {np.abs(a - b).max():.3g}")
    sc.close()
    assert sts[0] == sts[1]
    assert a.sum() > 0 and abs(a.sum() - b.sum()) <= 1e-5 * a.sum()
    assert np.abs(a - b).max() <= 1e-5 * a.max()


def test_dead_path_cutoff_leaves_gpu_image_unchanged(gpu_ctx):
    d, tex, res = scenes.bessel_cube_small()
    n, seed = 48, 13
    # the oracle runs the same photons first: a bounded full-work render
    _, ost = O.OracleScene(d, tex).render(n, res, seed=seed, precision=32, threads=gpu_threads(), cutoff=False)
    assert ost["er_steps"] < 5e8
    d2 = d.copy()
    d2.flags |= FLAG_NO_DEAD_CUT
    with full_trace(gpu_ctx):
        cut, _ = engine.Scene(gpu_ctx, d, tex).render(n, res, seed=seed)
        full, st = engine.Scene(gpu_ctx, d2, tex).render(n, res, seed=seed)
    cut, full = cut.double().cpu().numpy(), full.double().cpu().numpy()
    assert np.abs(cut - full).max() <= 1e-6 * full.max()
    assert abs(st["er_steps"] - ost["er_steps"]) < 1e-3 * ost["er_steps"]


def test_api_errors_are_reported(gpu_ctx):
    d, tex, res = scenes.bessel_cube_small()
    bad = d.copy()
    bad.er_stepsize = -1.0
    with pytest.raises(engine.SimplerError, match="er_stepsize"):
        engine.Scene(gpu_ctx, bad, tex)
    d2, tex2, res2, src = scenes.travelling_wave_cube()
    sc = engine.Scene(gpu_ctx, d2, tex2)  # no sources attached
    with pytest.raises(engine.SimplerError, match="without sources"):
        sc.render(10, res2)
    with pytest.raises(engine.SimplerError, match="n_total"):
        engine.Scene(gpu_ctx, d, tex).render(100, res, n_total=10)
