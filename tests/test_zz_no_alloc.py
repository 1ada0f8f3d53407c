"""The render entry point that allocates no device memory
(simpler_render_host_noalloc) against simpler_render_host, through the C
ABI: same photons, same image and counters, on two calls in a row (the
second smaller, so that the shared fixed device array keeps the first
call's larger contents beyond the second image), plus accumulation and the
size limit."""
import numpy as np
import pytest

from simpler_amd import engine, scenes

pytestmark = pytest.mark.gpu


def _scene(ctx):
    d, tex, _ = scenes.bessel_cube(step=1e-5, n_max=2e-5, mode=2)
    return engine.Scene(ctx, d, tex)


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = engine.Context(0)
    yield c
    c.close()


def _check(a, sa, b, sb):
    assert sa == sb, (sa, sb)
    assert a.sum() > 0
    assert abs(a.sum(dtype=np.float64) - b.sum(dtype=np.float64)) <= 1e-5 * a.sum(dtype=np.float64)
    assert np.abs(a - b).max() <= 1e-5 * a.max()


def test_noalloc_render_equals_render_host_on_two_calls(ctx):
    sc = _scene(ctx)
    for n, res, seed in ((1 << 14, (128, 128, 64), 3), (1 << 12, (32, 48, 8), 4)):
        a, sa = sc.render_host(n, res, seed=seed)
        b, sb = sc.render_host_noalloc(n, res, seed=seed)
        print(f"{n} photons {res}: photons {sb['photons']} er_steps {sb['er_steps']} total {b.sum():.6g}", flush=True)
        _check(a, sa, b, sb)
    sc.close()


def test_noalloc_render_accumulates_chunks(ctx):
    sc = _scene(ctx)
    res, N = (64, 64, 32), 1 << 13
    full, _ = sc.render_host_noalloc(N, res, seed=9, n_total=N)
    img = np.zeros(res[0] * res[1] * res[2], np.float32)
    sc.render_host_noalloc(N // 2, res, seed=9, offset=0, n_total=N, image=img)
    sc.render_host_noalloc(N // 2, res, seed=9, offset=N // 2, n_total=N, image=img, accumulate=True)
    img = img.reshape(full.shape)
    assert abs(img.sum(dtype=np.float64) - full.sum(dtype=np.float64)) <= 1e-5 * full.sum(dtype=np.float64)
    assert np.abs(img - full).max() <= 1e-4 * full.max()
    sc.close()


def test_noalloc_render_rejects_images_beyond_its_buffer(ctx):
    sc = _scene(ctx)
    with pytest.raises(engine.SimplerError, match="staging buffer"):
        sc.render_host_noalloc(16, (256, 256, 129))
    sc.close()
