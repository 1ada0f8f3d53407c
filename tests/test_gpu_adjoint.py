"""Path-replay adjoint against central finite differences.

Per-ray objectives (transparent medium) are checked against finite
differences of the FP64 oracle's per-ray trace -- a reference gradient with
no Monte-Carlo noise.  Image objectives with scattering are checked against
central differences of the GPU forward render with common random numbers,
and the adjoint's objective value against the render itself."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from simpler_amd import engine, scenes

pytestmark = pytest.mark.gpu


def t32(a):
    return torch.tensor(np.asarray(a), dtype=torch.float32, device="cuda")


def spline_scene(n=16, sigma=0.0, g=0.9, seed=0):
    d, tex, res, (vals, dims, lo, hi) = scenes.spline_cube(n=n, seed=seed)
    d.set(sigma_t=sigma, g=g)
    coeff = O.spline_build3d(vals.reshape(-1), dims).astype(np.float32)
    return d, tex, res, (coeff, dims, lo, hi)


def ray_objective(out, mask, gp, gd, gL):
    p, dv, L, _ = out
    return float((gp[mask] * p[mask]).sum() + (gd[mask] * dv[mask]).sum() + (gL[mask] * L[mask]).sum())


def upstream(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, 3)), rng.standard_normal((n, 3)), rng.standard_normal(n)


def test_ray_adjoint_spline_matches_fp64_finite_differences(gpu_ctx):
    d, tex, _, (coeff, dims, lo, hi) = spline_scene()
    n = 256
    gp, gd, gL = upstream(n, 1)
    sc = engine.Scene(gpu_ctx, d, tex, spline=(coeff, dims, lo, hi))
    grad, F, st = sc.adjoint_rays(t32(gp), t32(gd), t32(gL), seed=3)
    grad = grad.cpu().numpy().astype(np.float64)
    out = [x.cpu().numpy().astype(np.float64) for x in sc.trace_rays(n, seed=3)]
    mask = out[3] == 1
    assert mask.sum() > 0.9 * n
    F_trace = ray_objective(out, mask, gp, gd, gL)
    assert abs(F - F_trace) <= 1e-5 * np.abs(out[2][mask] * gL[mask]).sum() + 1e-6
    assert st["dropped"] == 0 and st["sub_paths_reversed"] == int(mask.sum())
    osc = O.OracleScene(d, tex, spline=(coeff, dims, lo, hi))
    base = osc.trace_rays(n, seed=3)
    omask = mask & (base[3] == 1)
    order = np.argsort(-np.abs(grad))
    rng = np.random.default_rng(7)
    picks = list(order[:4]) + list(rng.choice(order[4:300], 3, replace=False))
    gmax = np.abs(grad).max()
    assert gmax > 0
    for j in picks:
        cp, cm = coeff.copy(), coeff.copy()
        cp[j] = np.float32(coeff[j] + 2e-5)
        cm[j] = np.float32(coeff[j] - 2e-5)
        osc.set_spline(cp, dims, lo, hi)
        Fp = ray_objective(osc.trace_rays(n, seed=3), omask, gp, gd, gL)
        osc.set_spline(cm, dims, lo, hi)
        Fm = ray_objective(osc.trace_rays(n, seed=3), omask, gp, gd, gL)
        fd = (Fp - Fm) / (float(cp[j]) - float(cm[j]))
        assert abs(grad[j] - fd) <= 0.02 * abs(fd) + 2e-3 * gmax, (j, grad[j], fd)


def test_ray_adjoint_sources_matches_fp64_finite_differences(gpu_ctx):
    d, tex, _, src = scenes.travelling_wave_cube()
    d.set(sigma_t=0.0)
    n = 256
    gp, gd, gL = upstream(n, 2)
    sc = engine.Scene(gpu_ctx, d, tex, sources=src)
    N = len(src["amp"])
    assert sc.param_count() == 2 * N + 1
    grad, F, st = sc.adjoint_rays(t32(gp), t32(gd), t32(gL), seed=4)
    grad = grad.cpu().numpy().astype(np.float64)
    out = [x.cpu().numpy().astype(np.float64) for x in sc.trace_rays(n, seed=4)]
    mask = out[3] == 1
    assert st["dropped"] == 0
    osc = O.OracleScene(d, tex, sources=src)
    omask = mask & (osc.trace_rays(n, seed=4)[3] == 1)
    gmax = np.abs(grad).max()

    def fd_param(kind, i):
        s_p = {k: np.array(v, copy=True) for k, v in src.items()}
        s_m = {k: np.array(v, copy=True) for k, v in src.items()}
        dp, dm = d.copy(), d.copy()
        if kind == "amp":
            h = 0.01 * src["amp"][i]
            s_p["amp"][i] = np.float32(src["amp"][i] + h)
            s_m["amp"][i] = np.float32(src["amp"][i] - h)
            den = float(s_p["amp"][i]) - float(s_m["amp"][i])
        elif kind == "phase":
            s_p["phase"][i] = np.float32(src["phase"][i] + 1e-3)
            s_m["phase"][i] = np.float32(src["phase"][i] - 1e-3)
            den = float(s_p["phase"][i]) - float(s_m["phase"][i])
        else:  # wavenumber through f_u: k = 2 pi f / c
            dp.set(f_u=d.f_u * 1.0005)
            dm.set(f_u=d.f_u * 0.9995)
            den = 2 * np.pi * (float(np.float32(dp.f_u)) - float(np.float32(dm.f_u))) / d.speed_u
        Fp = ray_objective(O.OracleScene(dp, tex, sources=s_p).trace_rays(n, seed=4), omask, gp, gd, gL)
        Fm = ray_objective(O.OracleScene(dm, tex, sources=s_m).trace_rays(n, seed=4), omask, gp, gd, gL)
        return (Fp - Fm) / den

    for kind, i, j in (("amp", 0, 0), ("amp", 5, 5), ("phase", 3, N + 3), ("phase", 11, N + 11), ("k", 0, 2 * N)):
        fd = fd_param(kind, i)
        assert abs(grad[j] - fd) <= 0.02 * abs(fd) + 2e-3 * np.abs(grad[:2 * N]).max() * (j < 2 * N), \
            (kind, i, grad[j], fd)
    assert gmax > 0


def aperture_mask(res, radius_px):
    X, Y, Z = res
    yy, xx = np.mgrid[0:Y, 0:X]
    m = (((xx + 0.5 - X / 2) ** 2 + (yy + 0.5 - Y / 2) ** 2) <= radius_px ** 2).astype(np.float32)
    return np.repeat(m[None], Z, axis=0)


def smooth_direction(dims, sigma_cells, amp, seed):
    """A smooth random coefficient perturbation (x fastest), max |.| = amp."""
    N0, N1, N2 = dims
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((N2, N1, N0))
    k = [np.fft.fftfreq(n) for n in (N2, N1, N0)]
    kk = k[0][:, None, None] ** 2 + k[1][None, :, None] ** 2 + k[2][None, None, :] ** 2
    sm = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.exp(-2 * np.pi ** 2 * sigma_cells ** 2 * kk)))
    return (sm * (amp / np.abs(sm).max())).reshape(-1)


def test_render_adjoint_spline_matches_fp64_finite_differences(gpu_ctx):
    """Image objective with scattering, photon by photon: the GPU adjoint of
    each photon's aperture throughput against central differences of the
    FP64 oracle along a smooth perturbation of every coefficient.  Faces are
    index matched (ior 1) and the gate is off, so no photon changes its
    discrete events under the perturbation; a splat that crosses a pixel edge
    of the mask (a kink of the bilinear weights) may differ, so the bar is
    90 % of the contributing photons within 2 %."""
    from concurrent.futures import ThreadPoolExecutor
    from conftest import gpu_threads
    d, tex, _, (coeff, dims, lo, hi) = spline_scene(n=16, sigma=2000.0, g=0.9)
    d.set(ior=1.0, pathlength_range=(-1.0, -1.0))
    res = (32, 32, 16)
    Gn = aperture_mask(res, 6)
    G = t32(Gn)
    NT, NP = 1 << 18, 256
    sc = engine.Scene(gpu_ctx, d, tex, spline=(coeff, dims, lo, hi))
    grad, F, st = sc.adjoint_render(G, NP, seed=5, n_total=NT)
    assert st["dropped"] == 0 and st["sub_paths_reversed"] > 0 and st["main_paths_reversed"] > 0
    img, _ = sc.render(NP, res, seed=5, n_total=NT)
    assert abs(F - float((img.double() * G.double()).sum())) <= 1e-4 * abs(F)
    dc = smooth_direction(dims, 1.5, 1e-5, 4)
    cp = (coeff + dc).astype(np.float32)
    cm = (coeff - dc).astype(np.float32)
    half = (cp.astype(np.float64) - cm.astype(np.float64)) / 2.0
    ad = np.array([float(sc.adjoint_render(G, 1, seed=5, offset=i, n_total=NT)[0].cpu().double().numpy() @ half)
                   for i in range(NP)])
    assert abs(ad.sum() - float(grad.cpu().double().numpy() @ half)) <= 1e-3 * np.abs(ad).sum()
    op = O.OracleScene(d, tex, spline=(cp, dims, lo, hi))
    om = O.OracleScene(d, tex, spline=(cm, dims, lo, hi))

    def fd(i):
        Fp = float((op.render(1, res, seed=5, offset=i, n_total=NT, precision=64, threads=1)[0] * Gn).sum())
        Fm = float((om.render(1, res, seed=5, offset=i, n_total=NT, precision=64, threads=1)[0] * Gn).sum())
        return (Fp - Fm) / 2.0

    with ThreadPoolExecutor(gpu_threads()) as ex:
        fds = np.array(list(ex.map(fd, range(NP))))
    live = np.abs(fds) > 1e-3 * np.abs(fds).max()
    assert live.sum() >= 20
    good = np.abs(ad - fds) <= 0.02 * np.abs(fds)
    assert good[live].mean() >= 0.9, (good[live].mean(), ad[live][~good[live]], fds[live][~good[live]])


def test_adjoint_rejects_bessel_scenes(gpu_ctx):
    d, tex, res = scenes.bessel_cube_small()
    sc = engine.Scene(gpu_ctx, d, tex)
    with pytest.raises(engine.SimplerError):
        sc.param_count()


def test_autograd_function_and_adam_on_sources(gpu_ctx):
    from simpler_amd import autograd
    d, tex, _, src = scenes.travelling_wave_cube()
    res = (32, 32, 8)
    sc = engine.Scene(gpu_ctx, d, tex, sources=src)
    G = t32(aperture_mask(res, 6))
    n = 1 << 15
    params = torch.tensor(sc.params(), dtype=torch.float64, requires_grad=True)
    img = autograd.render(params, sc, res, n, seed=9)
    loss = (img.double() * G.double()).sum()
    loss.backward()
    g_ref, F, _ = sc.adjoint_render(G, n, seed=9)
    g_ref = g_ref.double().cpu()
    assert abs(loss.item() - F) <= 1e-4 * abs(F)
    assert torch.allclose(params.grad, g_ref, rtol=1e-4, atol=1e-6 * float(g_ref.abs().max()))
    # Adam on (amplitude / 5e-9, phase), k fixed, maximising the aperture
    # throughput with common random numbers: the steps run through the
    # autograd binding and move the parameters.
    N = len(src["amp"])
    base = torch.tensor(sc.params(), dtype=torch.float64)
    x = torch.cat([base[:N] / 5e-9, base[N:2 * N]]).clone().requires_grad_(True)
    opt = torch.optim.Adam([x], lr=0.05)
    vals = []
    for _ in range(4):
        opt.zero_grad()
        p = torch.cat([x[:N] * 5e-9, x[N:], base[2 * N:]])
        F_t = (autograd.render(p, sc, res, n, seed=9).double() * G.double()).sum()
        (-F_t).backward()
        opt.step()
        vals.append(F_t.item())
    assert all(np.isfinite(vals))
    assert float((x[:N] * 5e-9 - base[:N]).abs().max()) > 0
