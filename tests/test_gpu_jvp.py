"""Forward mode of the per-ray trace (simpler_jvp_rays; the reference's
er_derivativestep, scene.cpp:351-359, generalised to field parameters).

* uniform medium: the tangents with respect to the initial velocity v0
  (the reference's dp/dv0, dv/dv0) against the closed form of a straight
  ray ending on the sensor face;
* every field kind: the dot-product identity with the path-replay adjoint,
  <g, J theta_dot> = <J^T g, theta_dot>, for random directions theta_dot and
  random per-ray objective weights g;
* Bessel transducer parameters: the tangents against central differences of
  the FP64 oracle's per-ray trace along the same direction.
"""
import numpy as np
import pytest
import torch

from conftest import gpu_threads
from oracle import oracle as O
from simpler_amd import engine, scenes

pytestmark = pytest.mark.gpu


def t32(a):
    return torch.tensor(np.asarray(a), dtype=torch.float32, device="cuda")


def np64(ts):
    return [t.double().cpu().numpy() for t in ts]


def test_v0_tangents_of_a_straight_ray_match_the_closed_form(gpu_ctx):
    d, tex, _ = scenes.transparent_cube(mode=0)
    # n = n_o everywhere: straight rays; the emitter lens tilts them, so the
    # crossing-time term t' below is not zero
    d.set(egap_end_loc_x=-1.0, sgap_begin_loc_x=1.0, emitter_lens_active=1)
    sc = engine.Scene(gpu_ctx, d, tex)
    n = 4096
    vd = np.random.default_rng(3).standard_normal((n, 3))
    p, dr, L, s, tp, td, tL = np64(sc.jvp_rays(n, v0_dot=t32(vd), seed=2))
    assert (s == 1).all()
    assert np.abs(dr[:, 1:]).max() > 1e-2
    n0 = float(np.float32(d.n_o))
    # the ray p0 + t d0 reaches the sensor face x = -h: with v0 = n0 d0,
    # d' = (v0' - d (d . v0')) / n0, t' = -t d'_x / d_x, p' = t d' + t' d.
    # The march advances p by h v / n and L by h n per step, so in the
    # medium L = n0 s with s = n0 (x_face - x0) / v_x the step parameter at
    # the face: L' = -L v0'_x / v_x = -t v0'_x / d_x (the lens's term of L
    # does not depend on v0)
    x0 = float(d.block_r[0])
    t = (p[:, 0] - x0) / dr[:, 0]
    dd = (vd - dr * (dr * vd).sum(1, keepdims=True)) / n0
    tt = -t * dd[:, 0] / dr[:, 0]
    pe = t[:, None] * dd + tt[:, None] * dr
    Le = -t * vd[:, 0] / dr[:, 0]
    ep = np.abs(tp - pe).max() / np.abs(pe).max()
    ed = np.abs(td - dd).max() / np.abs(dd).max()
    eL = np.abs(tL - Le).max() / np.abs(Le).max()
    print(f"v0 tangents of straight rays: position {ep:.2e} direction {ed:.2e} length {eL:.2e}", flush=True)
    sc.close()
    assert ep < 1e-4 and ed < 1e-5 and eL < 1e-4


def _dot_identity(gpu_ctx, d, tex, n, seed, **field):
    sc = engine.Scene(gpu_ctx, d, tex, **field)
    npar = sc.param_count()
    rng = np.random.default_rng(seed)
    th = rng.standard_normal(npar)
    gp, gd, gL = rng.standard_normal((n, 3)), rng.standard_normal((n, 3)), rng.standard_normal(n)
    p, dr, L, s, tp, td, tL = np64(sc.jvp_rays(n, theta_dot=t32(th), seed=seed))
    grad, F, st = sc.adjoint_rays(t32(gp), t32(gd), t32(gL), seed=seed)
    grad = grad.double().cpu().numpy()
    ok = s == 1
    fwd = float((gp[ok] * tp[ok]).sum() + (gd[ok] * td[ok]).sum() + (gL[ok] * tL[ok]).sum())
    rev = float(grad @ th)
    scale = float(np.abs(grad).sum() * np.abs(th).max())
    sc.close()
    return fwd, rev, scale, ok.mean(), st


@pytest.mark.parametrize("mode", [0, 2])
def test_jvp_and_adjoint_agree_bessel(gpu_ctx, mode):
    d, tex, _ = scenes.transparent_cube(mode=mode)
    fwd, rev, scale, frac, st = _dot_identity(gpu_ctx, d, tex, 2048, 11 + mode)
    print(f"dot identity Bessel m={mode}: forward {fwd:.7g} reverse {rev:.7g}", flush=True)
    assert frac == 1.0 and st["dropped"] == 0
    assert abs(fwd - rev) <= 1e-3 * abs(rev)


def test_jvp_and_adjoint_agree_sources(gpu_ctx):
    d, tex, _, src = scenes.travelling_wave_cube()
    d.set(sigma_t=0.0)
    fwd, rev, scale, frac, st = _dot_identity(gpu_ctx, d, tex, 1024, 12, sources=src)
    print(f"dot identity sources: forward {fwd:.7g} reverse {rev:.7g}", flush=True)
    assert frac > 0.9 and st["dropped"] == 0
    assert abs(fwd - rev) <= 1e-3 * abs(rev)


def test_jvp_and_adjoint_agree_spline(gpu_ctx):
    d, tex, _, (vals, dims, lo, hi) = scenes.spline_cube(n=16)
    d.set(sigma_t=0.0)
    coeff = O.spline_build3d(vals.reshape(-1), dims).astype(np.float32)
    fwd, rev, scale, frac, st = _dot_identity(gpu_ctx, d, tex, 1024, 13, spline=(coeff, dims, lo, hi))
    print(f"dot identity spline: forward {fwd:.7g} reverse {rev:.7g}", flush=True)
    assert frac > 0.9 and st["dropped"] == 0
    assert abs(fwd - rev) <= 1e-3 * abs(rev)


@pytest.mark.parametrize("mode", [0, 2])
def test_bessel_tangents_match_fp64_central_differences(gpu_ctx, mode):
    """Along theta_dot = (n_max, k, duty shift) scaled to 1 %, 0.05 % and
    1e-3 rad: the per-ray tangents against the FP64 oracle's central
    differences, summed over the rays (each quantity within 1 %)."""
    d, tex, _ = scenes.transparent_cube(mode=mode)
    n = 512
    sc = engine.Scene(gpu_ctx, d, tex)
    k = 2 * np.pi * d.f_u / d.speed_u
    th = np.array([0.01 * d.n_max, 5e-4 * k, 1e-3])
    p, dr, L, s, tp, td, tL = np64(sc.jvp_rays(n, theta_dot=t32(th), seed=4))
    sc.close()
    assert (s == 1).all()

    def shifted(e):
        q = d.copy()
        q.set(n_max=d.n_max * (1 + 0.01 * e), n_clip=d.n_clip * (1 + 0.01 * e), f_u=d.f_u * (1 + 5e-4 * e),
              phi_min=d.phi_min + 1e-3 * e, phi_max=d.phi_max + 1e-3 * e)
        return O.OracleScene(q, tex).trace_rays(n, seed=4, precision=64, threads=gpu_threads())

    (pp, dp, Lp, sp), (pm, dm, Lm, sm) = shifted(1.0), shifted(-1.0)
    assert (sp == 1).all() and (sm == 1).all()
    # the descriptor holds float32: the step actually taken along theta_dot
    # differs from the nominal 2 eps by float rounding of the inputs, below
    # the 1 % bar
    fp, fdir, fL = (pp - pm) / 2, (dp - dm) / 2, (Lp - Lm) / 2
    for name, a, b in (("position", tp, fp), ("direction", td, fdir), ("length", tL, fL)):
        err = np.abs(a - b).sum() / np.abs(b).sum()
        print(f"Bessel m={mode} tangent vs FP64 central difference, {name}: {err:.2e}", flush=True)
        assert err < 0.01, name
