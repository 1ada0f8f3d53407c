"""The oracle's deterministic next-event connection (the reference's
computefdfNEE, scene.cpp:512-574, with er_derivativestep, scene.cpp:351-359)
and the field Hessians it uses, pinned on the CPU:

* each field's Hessian against central differences of the oracle's own
  gradient (FP64);
* in a uniform medium the march is a straight line: the miss against the
  ray's intersection with the plane, and the Jacobian against the closed form
  (L / |v1|) (I - d e_x^T / d_x)(I - d d^T) of that intersection;
* in the Bessel field the Jacobian against central differences of the miss
  (the reference's scheme takes d2Path at the step's end point, so it is the
  derivative of its own march only to O(h); measured 2e-5 here).
"""
import numpy as np
import pytest

from oracle import oracle as O
from simpler_amd import scenes


def _fd_hessian(osc, pts, scaling=1.0, tau=0.0, d=1e-7):
    H = np.zeros((len(pts), 3, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = d
        _, gp = osc.field(pts + e, scaling, tau)
        _, gm = osc.field(pts - e, scaling, tau)
        H[:, :, k] = (gp - gm) / (2 * d)
    return H


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_bessel_hessian_matches_differences_of_the_gradient(mode):
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-5, mode=mode)
    osc = O.OracleScene(d, tex)
    pts = np.random.default_rng(mode).uniform(-0.004, 0.004, (50, 3))
    H = osc.field_hessian(pts, 0.7)
    Hf = _fd_hessian(osc, pts, 0.7)
    assert np.abs(H - Hf).max() <= 1e-6 * np.abs(Hf).max()
    assert np.abs(H - H.transpose(0, 2, 1)).max() <= 1e-12 * np.abs(H).max()


def test_sources_and_spline_hessians_match_differences_of_the_gradient():
    pts = np.random.default_rng(5).uniform(-0.004, 0.004, (50, 3))
    d, tex, _, src = scenes.travelling_wave_cube()
    osc = O.OracleScene(d, tex, sources=src)
    H, Hf = osc.field_hessian(pts, 1.0, 0.3), _fd_hessian(osc, pts, 1.0, 0.3, d=1e-6)
    assert np.abs(H - Hf).max() <= 2e-5 * np.abs(Hf).max()
    d, tex, _, (vals, dims, lo, hi) = scenes.spline_cube(n=16)
    coeff = O.spline_build3d(vals.reshape(-1), dims)
    osc = O.OracleScene(d, tex, spline=(coeff.astype(np.float32), dims, lo, hi))
    H, Hf = osc.field_hessian(pts), _fd_hessian(osc, pts)
    assert np.abs(H - Hf).max() <= 5e-6 * np.abs(Hf).max()


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    p1 = np.column_stack([np.full(n, 0.004), rng.uniform(-0.002, 0.002, (n, 2))])
    dirs = np.column_stack([-np.ones(n), rng.uniform(-0.2, 0.2, (n, 2))])
    dirs /= np.linalg.norm(dirs, axis=1)[:, None]
    v1 = dirs * rng.uniform(0.5, 2.0, (n, 1))
    p2 = np.column_stack([np.full(n, -0.005), rng.uniform(-0.002, 0.002, (n, 2))])
    return p1, v1, p2, dirs


def uniform_closed_form(p1, v1, p2):
    """Intersection of the straight ray with the plane x = p2.x and the
    transpose of its Jacobian in v1."""
    dirs = v1 / np.linalg.norm(v1, axis=1)[:, None]
    L = (p2[:, 0] - p1[:, 0]) / dirs[:, 0]
    Q = p1 + L[:, None] * dirs
    n1 = np.linalg.norm(v1, axis=1)
    J = np.zeros((len(p1), 3, 3))
    for i, dd in enumerate(dirs):
        A = np.eye(3) - np.outer(dd, [1.0, 0.0, 0.0]) / dd[0]
        J[i] = (A @ ((L[i] / n1[i]) * (np.eye(3) - np.outer(dd, dd)))).T
    return Q, J


def test_connection_in_a_uniform_medium_is_the_straight_ray():
    d, tex, _ = scenes.transparent_cube(mode=0)
    d.set(egap_end_loc_x=-1.0, sgap_begin_loc_x=1.0)  # n = n_o everywhere
    osc = O.OracleScene(d, tex)
    p1, v1, p2, _ = _rays(20, 1)
    err, jac = osc.nee_connect(p1, v1, p2)
    Q, J = uniform_closed_form(p1, v1, p2)
    # the last step is bisected ceil(4 / log10 2) = 14 times: the end point
    # is within h / 2^14 of the plane, on the ray
    h = float(np.float32(d.er_stepsize))
    assert np.abs(err[:, 0]).max() <= h / 2 ** 13
    assert np.abs((err + p2) - Q).max() <= h / 2 ** 13
    assert np.abs(jac - J).max() <= 1e-7 * np.abs(J).max()


def test_connection_jacobian_matches_differences_in_the_bessel_field():
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-5, mode=2)
    osc = O.OracleScene(d, tex)
    p1, v1, p2, _ = _rays(12, 2)
    err, jac = osc.nee_connect(p1, v1, p2, 0.8)
    Jf = np.zeros_like(jac)
    for k in range(3):
        e = np.zeros(3)
        e[k] = 1e-3
        ep, _ = osc.nee_connect(p1, v1 + e, p2, 0.8)
        em, _ = osc.nee_connect(p1, v1 - e, p2, 0.8)
        Jf[:, k, :] = (ep - em) / 2e-3
    assert np.abs(jac - Jf).max() <= 1e-4 * np.abs(Jf).max()
    # the field bends the rays: the miss differs from the straight ray's
    Q, _ = uniform_closed_form(p1, v1, p2)
    assert np.abs((err + p2) - Q)[:, 1:].max() > 1e-7
