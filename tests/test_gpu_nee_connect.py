"""simpler_nee_connect (the reference's computefdfNEE, scene.cpp:512-574)
through the C ABI against the FP64 oracle's restatement of it, for every
field kind, and against the straight-ray closed form in a uniform medium.

Bars: the miss p - p2 within 1e-8 m (the last bisection steps, h / 2^12 and
below, are beneath FP32 resolution of the position, so the GPU's and the
oracle's end points can differ by about 1e-9 m); the Jacobian within 3e-4 of
its largest entry (FP32 accumulation over about 1,800 steps)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from simpler_amd import engine, scenes
from test_oracle_nee_connect import _rays, uniform_closed_form

pytestmark = pytest.mark.gpu


def t32(a):
    return torch.tensor(np.asarray(a), dtype=torch.float32, device="cuda")


def _gpu(sc, p1, v1, p2, scaling=1.0, tau=0.0):
    err, jac = sc.nee_connect(t32(p1), t32(v1), t32(p2), scaling, tau)
    return err.double().cpu().numpy(), jac.double().cpu().numpy()


def _f32(*xs):
    return [np.asarray(x, np.float32).astype(np.float64) for x in xs]


def _compare(tag, ge, gj, oe, oj):
    de = np.abs(ge - oe).max()
    dj = np.abs(gj - oj).max() / np.abs(oj).max()
    print(f"{tag}: miss {de:.2e} m, Jacobian {dj:.2e} of max", flush=True)
    assert de <= 1e-8 and dj <= 3e-4


def test_uniform_medium_closed_form(gpu_ctx):
    d, tex, _ = scenes.transparent_cube(mode=0)
    d.set(egap_end_loc_x=-1.0, sgap_begin_loc_x=1.0)
    sc = engine.Scene(gpu_ctx, d, tex)
    p1, v1, p2 = _f32(*_rays(1024, 3)[:3])
    err, jac = _gpu(sc, p1, v1, p2)
    sc.close()
    Q, J = uniform_closed_form(p1, v1, p2)
    h = float(np.float32(d.er_stepsize))
    assert np.abs(err[:, 0]).max() <= h / 2 ** 12
    assert np.abs((err + p2) - Q).max() <= 2e-8
    assert np.abs(jac - J).max() <= 3e-4 * np.abs(J).max()


@pytest.mark.parametrize("mode", [0, 2])
def test_bessel_matches_the_oracle(gpu_ctx, mode):
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-5, mode=mode)
    sc = engine.Scene(gpu_ctx, d, tex)
    p1, v1, p2 = _f32(*_rays(256, 10 + mode)[:3])
    ge, gj = _gpu(sc, p1, v1, p2, 0.8)
    sc.close()
    oe, oj = O.OracleScene(d, tex).nee_connect(p1, v1, p2, 0.8)
    _compare(f"Bessel m={mode}", ge, gj, oe, oj)


def test_bessel_component_form_rule(gpu_ctx):
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-5, mode=2, axis="x")
    from simpler_amd._abi import FLAG_GENERAL_AXIS
    d.set(axis_uz=(1.0, 0.0, 0.0), flags=d.flags & ~FLAG_GENERAL_AXIS)
    sc = engine.Scene(gpu_ctx, d, tex)
    p1, v1, p2 = _f32(*_rays(4, 1)[:3])
    with pytest.raises(engine.SimplerError, match="component form"):
        _gpu(sc, p1, v1, p2)
    sc.close()


def test_sources_match_the_oracle(gpu_ctx):
    d, tex, _, src = scenes.travelling_wave_cube()
    sc = engine.Scene(gpu_ctx, d, tex, sources=src)
    p1, v1, p2 = _f32(*_rays(64, 4)[:3])
    ge, gj = _gpu(sc, p1, v1, p2, 1.0, 0.3)
    sc.close()
    oe, oj = O.OracleScene(d, tex, sources=src).nee_connect(p1, v1, p2, 1.0, 0.3)
    _compare("sources", ge, gj, oe, oj)


def test_spline_matches_the_oracle(gpu_ctx):
    d, tex, _, (vals, dims, lo, hi) = scenes.spline_cube(n=16)
    coeff = O.spline_build3d(vals.reshape(-1), dims).astype(np.float32)
    sc = engine.Scene(gpu_ctx, d, tex, spline=(coeff, dims, lo, hi))
    p1, v1, p2 = _f32(*_rays(256, 5)[:3])
    ge, gj = _gpu(sc, p1, v1, p2)
    sc.close()
    oe, oj = O.OracleScene(d, tex, spline=(coeff, dims, lo, hi)).nee_connect(p1, v1, p2)
    _compare("spline", ge, gj, oe, oj)
