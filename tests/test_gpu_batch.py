"""Batched persistent launch: many transducer configurations with uneven
photon budgets in one kernel.  Each configuration's image must equal the
single-scene render of that configuration (same seed, same photons); a
captured HIP graph must replay to the eager result every time."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from conftest import gpu_threads
from simpler_amd import engine, scenes

pytestmark = pytest.mark.gpu


def small_batch(n_cfg=6, seed=3):
    d, tex, cfgs, res = scenes.batched_configs(n_cfg=n_cfg, seed=seed, budget_lo=200, budget_hi=3000)
    return d, tex, cfgs, res


def test_batch_images_equal_single_scene_renders(gpu_ctx):
    d, tex, cfgs, res = small_batch()
    base = engine.Scene(gpu_ctx, d, tex)
    batch = engine.Batch(base, cfgs, res)
    imgs, st = batch.render()
    imgs = imgs.double().cpu().numpy()
    assert st["photons"] == sum(c["n_photons"] for c in cfgs)
    assert len({c["mode"] for c in cfgs}) > 1
    for i, c in enumerate(cfgs):
        sc = engine.Scene(gpu_ctx, scenes.desc_for_config(d, c), tex)
        ref, rst = sc.render(c["n_photons"], res, seed=c["seed"])
        ref = ref.double().cpu().numpy()
        assert ref.sum() > 0
        assert abs(imgs[i].sum() - ref.sum()) <= 1e-5 * ref.sum(), i
        assert np.abs(imgs[i] - ref).max() <= 1e-4 * ref.max(), i


def test_batch_graph_replays_match_eager(gpu_ctx):
    d, tex, cfgs, res = small_batch(n_cfg=5, seed=8)
    base = engine.Scene(gpu_ctx, d, tex)
    batch = engine.Batch(base, cfgs, res)
    eager, _ = batch.render()
    eager = eager.clone()
    batch.capture()
    for k in range(3):
        out = batch.replay()
        torch.cuda.synchronize()
        assert torch.allclose(out, eager, rtol=1e-4, atol=1e-6 * float(eager.abs().max())), k
        junk = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        assert torch.isfinite(junk)


def test_batch_configuration_matches_oracle(gpu_ctx):
    """Each configuration's total against the FP64 oracle on the same photons,
    within 3 sigma of the oracle's own Monte-Carlo spread (8 photon blocks):
    identical paths give equal totals, and a rare photon whose discrete events
    differ between FP32 and FP64 stays well inside the bound."""
    d, tex, cfgs, res = small_batch(n_cfg=3, seed=11)
    n = 1024
    for c in cfgs:
        c["n_photons"] = n
    base = engine.Scene(gpu_ctx, d, tex)
    imgs, _ = engine.Batch(base, cfgs, res).render()
    imgs = imgs.double().cpu().numpy()
    for i, c in enumerate(cfgs):
        osc = O.OracleScene(scenes.desc_for_config(d, c), tex)
        blocks = [osc.render(n // 8, res, seed=c["seed"], offset=k * n // 8, n_total=n, precision=64,
                             threads=gpu_threads(), cutoff=True)[0].sum() for k in range(8)]
        total, sigma = float(np.sum(blocks)), float(np.std(blocks, ddof=1) * np.sqrt(8))
        assert abs(imgs[i].sum() - total) <= 3 * sigma, (i, imgs[i].sum(), total, sigma)


def test_destroy_order_is_enforced():
    ctx = engine.Context(0)
    d, tex, cfgs, res = small_batch(n_cfg=2)
    sc = engine.Scene(ctx, d, tex)
    b = engine.Batch(sc, cfgs, res)
    assert ctx.L.simpler_ctx_destroy(ctx.h) == 3  # SIMPLER_ERR_STATE: scene and batch alive
    assert ctx.L.simpler_scene_destroy(sc.h) == 3  # batch alive
    ctx.close()  # closes the batch, then the scene, then the context
    assert b.h is None and sc.h is None and ctx.h is None


def pattern_setup(n_cfg=4, seed=2, photons=(1500, 400, 2600, 900)):
    d, tex, _, src = scenes.travelling_wave_cube()
    rng = np.random.default_rng(seed)
    N = len(src["amp"])
    amps = rng.uniform(0.0, 5e-9, (n_cfg, N)).astype(np.float32)
    phases = rng.uniform(0.0, 2 * np.pi, (n_cfg, N)).astype(np.float32)
    seeds = [scenes.splitmix64(seed * 1000 + i) for i in range(n_cfg)]
    return d, tex, src, amps, phases, list(photons[:n_cfg]), seeds, (32, 32, 8)


def test_pattern_batch_matches_single_scene_renders_and_adjoints(gpu_ctx):
    d, tex, src, amps, phases, budgets, seeds, res = pattern_setup()
    base = engine.Scene(gpu_ctx, d, tex, sources=src)
    batch = engine.Batch.patterns(base, amps, phases, budgets, seeds, res)
    imgs, st = batch.render()
    imgs = imgs.double().cpu().numpy()
    assert st["photons"] == sum(budgets)
    X, Y, Z = res
    yy, xx = np.mgrid[0:Y, 0:X]
    mask = ((xx + 0.5 - X / 2) ** 2 + (yy + 0.5 - Y / 2) ** 2 <= 36).astype(np.float32)
    G = torch.tensor(np.repeat(mask[None, None], len(budgets), axis=0).repeat(Z, axis=1), device="cuda")
    grads, F, ast = batch.adjoint(G)
    grads = grads.double().cpu().numpy()
    assert ast["dropped"] == 0 and ast["photons"] == sum(budgets)
    for i in range(len(budgets)):
        sc = engine.Scene(gpu_ctx, d, tex, sources=dict(src, amp=amps[i], phase=phases[i]))
        ref, _ = sc.render(budgets[i], res, seed=seeds[i])
        ref = ref.double().cpu().numpy()
        assert ref.sum() > 0
        assert abs(imgs[i].sum() - ref.sum()) <= 1e-5 * ref.sum(), i
        assert np.abs(imgs[i] - ref).max() <= 1e-4 * ref.max(), i
        g, Fi, _ = sc.adjoint_render(G[i], budgets[i], seed=seeds[i])
        g = g.double().cpu().numpy()
        assert abs(F[i] - Fi) <= 1e-6 * abs(Fi), i
        assert np.allclose(grads[i], g, rtol=2e-3, atol=1e-4 * np.abs(g).max()), i


def test_pattern_batch_adjoint_schedules_agree(gpu_ctx):
    """The batched adjoint (16 patterns x 4096 photons) with its backward runs
    deferred (the default) and inline, and deferred with queues sized to 70 %
    of the batch's queued sub-paths: the whole-batch chunk overflows (its
    objectives and counters are restored) and the photons are traced again
    in smaller chunks that start inside a configuration.  Same objectives
    and counts, gradients to the order of float additions."""
    d, tex, src, amps, phases, budgets, seeds, res = pattern_setup(n_cfg=16, seed=5, photons=(4096,) * 16)
    base = engine.Scene(gpu_ctx, d, tex, sources=src)
    batch = engine.Batch.patterns(base, amps, phases, budgets, seeds, res)
    X, Y, Z = res
    yy, xx = np.mgrid[0:Y, 0:X]
    mask = ((xx + 0.5 - X / 2) ** 2 + (yy + 0.5 - Y / 2) ** 2 <= 36).astype(np.float32)
    G = torch.tensor(np.repeat(mask[None, None], len(budgets), axis=0).repeat(Z, axis=1), device="cuda")
    out = {}
    try:
        gpu_ctx.set_adjoint_schedule(1, 1 << 23)
        out["deferred"] = batch.adjoint(G)
        jobs = out["deferred"][2]["sub_paths_reversed"]
        assert jobs > 1000
        gpu_ctx.set_adjoint_schedule(0)
        out["inline"] = batch.adjoint(G)
        gpu_ctx.set_adjoint_schedule(1, max(64, int(0.7 * jobs)))
        out["overflow"] = batch.adjoint(G)
    finally:
        gpu_ctx.set_adjoint_schedule(1, 1 << 23)
        batch.close()
        base.close()
    gi, Fi, si = out["inline"]
    gi = gi.double().cpu().numpy()
    for name in ("deferred", "overflow"):
        g, F, st = out[name]
        g = g.double().cpu().numpy()
        err = np.abs(g - gi).max() / np.abs(gi).max()
        print(f"patterns {name}: {st}, gradient difference {err:.2e} of the largest entry", flush=True)
        assert np.allclose(np.asarray(F), np.asarray(Fi), rtol=1e-7, atol=0.0), name
        for k in ("er_steps", "reverse_steps"):  # kernels compiled apart: rounding may flip a step
            assert abs(st[k] - si[k]) <= 1e-9 * si[k], (name, k)
        for k in ("photons", "sub_paths_reversed", "main_paths_reversed", "dropped"):
            assert st[k] == si[k], (name, k)
        assert err <= 1e-4, name
