"""Whole-render properties of the oracle: Beer-Lambert energy bookkeeping of
the direct term, time binning, chunked additivity, the dead-path cutoff
leaving the image unchanged, and FP32 vs FP64 agreement."""
import math

import numpy as np

from oracle import oracle as O
from simpler_amd import scenes


def _small_scene(**kw):
    d, tex, res = scenes.bessel_cube(step=2e-5, n_max=1e-5, mode=0, res=(32, 32, 16))
    d.set(**kw)
    return d, tex, res


def test_transparent_direct_energy_and_time_bin():
    d, tex, res = _small_scene(sigma_t=0.0)
    sc = O.OracleScene(d, tex)
    img, st = sc.render(200, res, seed=4, precision=64, threads=4)
    # sigma_t = 0 forces albedo 0: only the direct term, weight Li/N each
    assert st["scatters"] == 0 and st["nee_splats"] == 0 and st["direct_splats"] == 200
    assert abs(img.sum() - d.Li) < 1e-6 * d.Li
    # optical path ~ n0 * 10 mm -> bin floor(L / 0.1 m * Z)
    zbin = int(math.floor(d.n_o * 0.010 / 0.1 * res[2]))
    assert abs(img[zbin].sum() - d.Li) < 1e-6 * d.Li


def test_beer_lambert_direct_term():
    d, tex, res = _small_scene(sigma_t=150.0, use_angular_sampling=0, albedo=0.0)
    sc = O.OracleScene(d, tex)
    img, st = sc.render(100, res, seed=2, precision=64, threads=4)
    # the inside test tolerates eps = 1.19209290e-7 m past a face
    # (block.cpp:18-19), so the refined march stops at 10 mm + eps, within
    # one finest step h/10^precision
    eps = 1.19209290e-07
    expect = d.Li * math.exp(-150.0 * (0.010 + eps))
    assert abs(img.sum() - expect) < 150.0 * d.er_stepsize * 1e-4 * expect + 1e-9 * expect


def test_chunked_render_is_additive():
    d, tex, res = _small_scene()
    sc = O.OracleScene(d, tex)
    full, _ = sc.render(60, res, seed=9, precision=32, threads=4)
    a, _ = sc.render(25, res, seed=9, offset=0, n_total=60, precision=32, threads=4)
    b, _ = sc.render(35, res, seed=9, offset=25, n_total=60, precision=32, threads=4)
    np.testing.assert_allclose(a + b, full, rtol=1e-12, atol=1e-12 * full.max())


def test_dead_path_cutoff_leaves_image_unchanged():
    d, tex, res = _small_scene()
    sc = O.OracleScene(d, tex)
    full, st_full = sc.render(40, res, seed=11, precision=32, threads=4, cutoff=False)
    cut, st_cut = sc.render(40, res, seed=11, precision=32, threads=4, cutoff=True)
    assert np.array_equal(full, cut)
    assert st_cut["er_steps"] < st_full["er_steps"]
    # cut marches never splat and dead photons stop scattering: fewer events, same image
    assert st_cut["nee_splats"] <= st_full["nee_splats"] and st_cut["scatters"] <= st_full["scatters"]


def test_dead_path_cutoff_is_disabled_without_gate():
    d, tex, res = _small_scene(pathlength_range=(-1.0, -1.0))
    assert math.isinf(O.OracleScene(d, tex).dead_path_cut())
    d, tex, res = _small_scene(sensor_lens_active=1)
    assert math.isinf(O.OracleScene(d, tex).dead_path_cut())


def test_fp32_and_fp64_renders_agree():
    d, tex, res = _small_scene()
    sc = O.OracleScene(d, tex)
    i32, s32 = sc.render(60, res, seed=5, precision=32, threads=4, cutoff=True)
    i64, s64 = sc.render(60, res, seed=5, precision=64, threads=4, cutoff=True)
    assert abs(i32.sum() - i64.sum()) < 2e-3 * i64.sum()
    assert abs(s32["scatters"] - s64["scatters"]) <= 0.02 * s64["scatters"]


def test_direct_leak_shifts_scattered_time_bins():
    """With the reference's path-length leak the scattered light arrives
    about n0 * 10 mm later than without it."""
    d, tex, res = _small_scene()
    sc = O.OracleScene(d, tex)
    leak, _ = sc.render(60, res, seed=6, precision=32, threads=4, cutoff=True)
    d2 = d.copy()
    d2.flags &= ~1
    nol, _ = O.OracleScene(d2, tex).render(60, res, seed=6, precision=32, threads=4, cutoff=True)
    z = np.arange(res[2])
    mean_leak = (leak.sum((1, 2)) * z).sum() / leak.sum()
    mean_nol = (nol.sum((1, 2)) * z).sum() / nol.sum()
    shift = (mean_leak - mean_nol) * 0.1 / res[2]
    assert 0.6 * d.n_o * 0.010 < shift < 1.4 * d.n_o * 0.010
