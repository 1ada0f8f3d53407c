"""The GPU B-spline interpolation prefilter (spline.h:686-764) against the
oracle's FP64 build3d (itself bit-exact with the reference's CSV
coefficients), the reference's own 10^3 golden grid, and at full size the
interpolation property: the spline through the coefficients returns the
samples at every node."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from simpler_amd import engine, scenes

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(__file__), "golden")


def oracle_coeffs(g32):
    """FP64 build3d of float32 samples g32 (shape (N2, N1, N0))."""
    N2, N1, N0 = g32.shape
    return O.spline_build3d(g32.reshape(-1).astype(np.float64), [N0, N1, N2]).reshape(g32.shape)


def spline_scene(ctx, half=0.005):
    d, tex, _, _ = scenes.spline_cube(n=8)
    return engine.Scene(ctx, d, tex), d, tex


# line lengths below, at and above the kernel's 48-sample window and its
# 32-sample steps (48, 49, 50, 80, 81, 97, ...)
@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 5, 4), (17, 3, 30), (7, 200, 9), (49, 48, 50), (97, 81, 33),
                                   (64, 64, 64), (101, 101, 101)])
def test_in_place_prefilter_matches_oracle(gpu_ctx, shape):
    print("prefilter shape", shape, flush=True)
    rng = np.random.default_rng(sum(shape))
    g = (1.3333 + 1e-3 * rng.standard_normal(shape)).astype(np.float32)
    t = torch.tensor(g.reshape(-1), device="cuda")
    gpu_ctx.spline_prefilter(t, (shape[2], shape[1], shape[0]))
    gpu_ctx.synchronize()
    c = t.double().cpu().numpy().reshape(shape)
    ref = oracle_coeffs(g)
    # FP32 recursion on absolute values: a few ulp of the values (~1.3)
    assert np.abs(c - ref).max() <= 3e-6 * np.abs(ref).max()


def test_prefilter_reproduces_reference_golden_coefficients(gpu_ctx):
    gold = np.load(os.path.join(G, "spline_golden.npz"))
    data = gold["i3d_data"]  # 10^3 samples, x fastest
    ref = np.zeros(1000)
    for k in range(10):  # the CSV prints rows (k, i) with columns j
        for i in range(10):
            for j in range(10):
                ref[i + j * 10 + k * 100] = gold["c3d"][k * 10 + i, j]
    sc, _, _ = spline_scene(gpu_ctx)
    sc.set_spline_samples(data.astype(np.float32), (10, 10, 10), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0))
    c = sc.spline_coeffs().astype(np.float64)
    # the reference's FP64 coefficients of the FP64 samples; the samples enter
    # as float32 (rounding 6e-8 relative, amplified ~3x per axis by the filter)
    assert np.abs(c - ref).max() <= 1e-6 * np.abs(ref).max()


def test_scene_from_samples_equals_scene_from_coefficients(gpu_ctx):
    vals, dims, lo, hi = scenes.gaussian_spline_coeffs(n=32, max_dn=5e-5, seed=5)
    g32 = vals.astype(np.float32)
    ref = oracle_coeffs(g32)
    sc_s, d, tex = spline_scene(gpu_ctx)
    sc_s.set_spline_samples(g32.reshape(-1), dims, lo, hi)
    c = sc_s.spline_coeffs().astype(np.float64).reshape(g32.shape)
    shift = float(g32.reshape(-1)[0])
    # the shift keeps FP32 resolution on the 5e-5 variation: ~1e-9 absolute
    assert np.abs(c - ref).max() <= 1e-3 * np.abs(ref - shift).max() + 2e-7
    # field against an FP64 evaluation of the FP64 coefficients; the scene
    # built from float32 absolute coefficients is the comparison: its
    # coefficients carry the float32 rounding of ~1.33 (6e-8), which the
    # gradient divides by the grid spacing
    sc_c = engine.Scene(gpu_ctx, d, tex, spline=(ref.reshape(-1).astype(np.float32), dims, lo, hi))
    pts = np.random.default_rng(1).uniform(-0.0049, 0.0049, (4096, 3)).astype(np.float32)
    pts64 = pts.astype(np.float64)
    c64 = ref.reshape(-1)
    ev = [O.spline_eval3d(c64, list(dims), list(lo), list(hi), *o, pts64)
          for o in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1))]
    n64, g64 = ev[0], np.stack(ev[1:], axis=1)
    tp = torch.tensor(pts, device="cuda")
    (n_s, g_s), (n_c, g_c) = sc_s.field(tp), sc_c.field(tp)
    err = lambda a, b: np.abs(a.double().cpu().numpy() - b).max()
    gmax = np.abs(g64).max()
    assert err(n_s, n64) < 2e-7 and err(n_c, n64) < 3e-7
    assert err(g_s, g64) < 2e-4 * gmax
    assert err(g_s, g64) <= err(g_c, g64)


def test_full_size_grid_interpolates_its_samples(gpu_ctx):
    """256^3 samples from device memory: the spline evaluated at nodes
    (interior and boundary) returns the samples."""
    n = 256
    vals, dims, lo, hi = scenes.gaussian_spline_coeffs(n=n, max_dn=5e-5, sigma_cells=6.0, seed=2)
    g32 = torch.tensor(vals.astype(np.float32).reshape(-1), device="cuda")
    sc, _, _ = spline_scene(gpu_ctx)
    sc.set_spline_samples(g32, dims, lo, hi)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.integers(0, n, (4000, 3)),
                          np.array([[0, 0, 0], [n - 1, n - 1, n - 1], [0, n - 1, 5], [n - 1, 0, n - 2]])])
    h = (hi[0] - lo[0]) / (n - 1)
    pts = torch.tensor(lo[0] + idx * h, dtype=torch.float32, device="cuda")
    nv, _ = sc.field(pts)
    want = vals[idx[:, 2], idx[:, 1], idx[:, 0]]
    assert np.abs(nv.double().cpu().numpy() - want).max() < 3e-7
    # and every coefficient is finite (the minimum reduction saw them all)
    c = sc.spline_coeffs().reshape(n, n, n)
    assert np.isfinite(c).all()


def test_non_finite_sample_is_rejected(gpu_ctx):
    sc, _, _ = spline_scene(gpu_ctx)
    g = np.full(4 * 5 * 6, 1.3333, np.float32)
    g[17] = np.nan
    with pytest.raises(engine.SimplerError):
        sc.set_spline_samples(g, (4, 5, 6), (-1, -1, -1), (1, 1, 1))
    g[17] = np.inf
    g[0] = 1.0
    with pytest.raises(engine.SimplerError):
        sc.set_spline_samples(g, (4, 5, 6), (-1, -1, -1), (1, 1, 1))
    with pytest.raises(engine.SimplerError):
        sc.set_spline_samples(np.ones(8, np.float32), (1, 2, 4), (-1, -1, -1), (1, 1, 1))


def test_reference_vol_grid_as_spline_matches_the_bessel_field(gpu_ctx):
    """renderings/besselriffile.vol (101^3 samples of the testing_splines.sh
    Bessel field; constant along x, so the committed y-z slice rebuilds it
    exactly): the GPU Bessel field reproduces its nodes, and the spline scene
    built from the grid (prefiltered on the GPU) matches the Bessel field
    between the nodes -- the comparison of renderer/samples/testUS.cpp."""
    g = np.load(os.path.join(G, "besselriffile_slice.npz"))
    N, lo, hi, sl = g["N"], g["xmin"], g["xmax"], g["slice_zy"]
    assert float(g["x_variation"]) == 0.0
    grid = np.repeat(sl[:, :, None], int(N[0]), axis=2)  # (z, y, x), x fastest
    d = scenes.vol_bessel_desc(True)
    tex = scenes.disk_texture(11)
    bes = engine.Scene(gpu_ctx, d, tex)
    ys = np.linspace(lo[1], hi[1], N[1])
    zs = np.linspace(lo[2], hi[2], N[2])
    zz, yy = np.meshgrid(zs, ys, indexing="ij")
    nodes = np.stack([np.zeros(zz.size), yy.ravel(), zz.ravel()], axis=1)
    n_nodes, _ = bes.field(torch.tensor(nodes, dtype=torch.float32, device="cuda"))
    assert np.abs(n_nodes.double().cpu().numpy().reshape(sl.shape) - sl).max() < 2.5e-7
    ds = d.copy()
    ds.set(field_kind=2)
    spl = engine.Scene(gpu_ctx, ds, tex)
    spl.set_spline_samples(torch.tensor(grid.reshape(-1), device="cuda"), [int(x) for x in N], list(lo), list(hi))
    # points away from the grid's y / z faces: within a few cells of a face
    # the mirror boundary (zero slope there) departs from the field, in the
    # FP64 oracle spline as well (5e-6 at 2.5 cells)
    rng = np.random.default_rng(8)
    pts = np.stack([rng.uniform(-0.014, 0.014, 4096), rng.uniform(-0.0008, 0.0008, 4096),
                    rng.uniform(-0.0008, 0.0008, 4096)], axis=1).astype(np.float32)
    tp = torch.tensor(pts, dtype=torch.float32, device="cuda")
    (n_s, g_s), (n_b, g_b) = spl.field(tp), bes.field(tp)
    en = (n_s - n_b).abs().max().item()
    eg = ((g_s - g_b).abs().max() / g_b.abs().max()).item()
    print(f"spline vs Bessel between nodes: n {en:.2e}, grad {eg:.2e} of max", flush=True)
    assert en < 3e-7 and eg < 1e-3
    # and against the FP64 oracle spline of the same grid at the same points
    Ni = [int(x) for x in N]
    c64 = O.spline_build3d(grid.reshape(-1).astype(np.float64), Ni)
    p64 = pts.astype(np.float64)
    n64 = O.spline_eval3d(c64, Ni, list(lo.astype(float)), list(hi.astype(float)), 0, 0, 0, p64)
    g64 = np.stack([O.spline_eval3d(c64, Ni, list(lo.astype(float)), list(hi.astype(float)), *o, p64)
                    for o in ((1, 0, 0), (0, 1, 0), (0, 0, 1))], axis=1)
    assert np.abs(n_s.double().cpu().numpy() - n64).max() < 2e-7
    assert np.abs(g_s.double().cpu().numpy() - g64).max() < 1e-4 * np.abs(g64).max()
