"""Pin the oracle's physics by what the mathematics fixes: Bessel values
(scipy), field gradients (finite differences, rotation equivalence, the
reference's axis=x formula), the leapfrog (straight lines, invariants,
reversibility, second-order convergence), Fresnel / Snell closed forms, and
Henyey-Greenstein moments."""
import math

import numpy as np
import pytest
from scipy import special

from oracle import oracle as O
from simpler_amd import scenes
from simpler_amd._abi import FIELD_SOURCES, FLAG_GENERAL_AXIS


def test_bessel_matches_scipy():
    x = np.linspace(0.0, 80.0, 401)
    for m in range(5):
        v = np.array([O.lib().oracle_besselj(m, float(t)) for t in x])
        np.testing.assert_allclose(v, special.jv(m, x), atol=1e-13)


def _fd_grad(sc, pts, h=1e-7, **kw):
    g = np.zeros_like(pts)
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        n_p, _ = sc.field(pts + e, **kw)
        n_m, _ = sc.field(pts - e, **kw)
        g[:, k] = (n_p - n_m) / (2 * h)
    return g


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("axis", ["x", "z"])
def test_bessel_gradient_is_fd_of_value(mode, axis):
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-5, mode=mode, axis=axis)
    sc = O.OracleScene(d, tex)
    rng = np.random.default_rng(mode)
    pts = rng.uniform(-0.0045, 0.0045, (200, 3))
    n, g = sc.field(pts, scaling=0.8)
    gfd = _fd_grad(sc, pts, scaling=0.8)
    assert np.abs(g - gfd).max() < 1e-6 * np.abs(g).max() + 1e-12
    # closed form of the value
    uz = np.array(d.axis_uz)
    ux = np.array(d.axis_ux)
    uy = np.cross(uz, ux)
    rho = pts - np.outer(pts @ uz, uz)
    r = np.linalg.norm(rho, axis=1)
    phi = np.arctan2(rho @ uy, rho @ ux)
    k = 2 * math.pi * d.f_u / d.speed_u
    ref = d.n_o + d.n_max * 0.8 * special.jv(mode, k * r) * np.cos(mode * phi)
    np.testing.assert_allclose(n, ref, atol=1e-12)


def test_reference_axis_formula_equals_general_formula_at_axis_x():
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-5, mode=2, axis="x")
    pts = np.random.default_rng(3).uniform(-0.005, 0.005, (300, 3))
    gen = O.OracleScene(d, tex).field(pts)[1]
    d2 = d.copy()
    d2.flags &= ~FLAG_GENERAL_AXIS
    ref = O.OracleScene(d2, tex).field(pts)[1]
    assert np.abs(ref[:, 0]).max() == 0.0  # the reference hard-codes dn.x = 0
    np.testing.assert_allclose(gen, ref, atol=1e-12 * np.abs(ref).max())


def test_bessel_rotation_equivalence():
    """The field for axis z at point p equals the field for axis x at the
    rotated point; gradients rotate with it."""
    dz, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-5, mode=2, axis="z")
    dx = dz.copy().set(axis_uz=(-1.0, 0.0, 0.0), axis_ux=(0.0, 0.0, 1.0))
    # rotation R with R @ uz_z = uz_x and R @ ux_z = ux_x
    uz1, ux1 = np.array(dz.axis_uz), np.array(dz.axis_ux)
    uz2, ux2 = np.array(dx.axis_uz), np.array(dx.axis_ux)
    A = np.stack([ux1, np.cross(uz1, ux1), uz1])
    B = np.stack([ux2, np.cross(uz2, ux2), uz2])
    R = B.T @ A
    pts = np.random.default_rng(5).uniform(-0.004, 0.004, (100, 3))
    n1, g1 = O.OracleScene(dz, tex).field(pts)
    n2, g2 = O.OracleScene(dx, tex).field(pts @ R.T)
    np.testing.assert_allclose(n1, n2, atol=1e-14)
    np.testing.assert_allclose(g1 @ R.T, g2, atol=1e-10 * np.abs(g1).max())


def test_sources_field_closed_form_and_gradient():
    d, tex, _, src = scenes.travelling_wave_cube()
    sc = O.OracleScene(d, tex, sources=src)
    pts = np.random.default_rng(2).uniform(-0.004, 0.004, (100, 3))
    tau = 0.7
    n, g = sc.field(pts, tau=tau)
    k = 2 * math.pi * d.f_u / d.speed_u
    r = np.linalg.norm(pts[:, None, :] - src["pos"][None].astype(np.float64), axis=2)
    ref = d.n_o + (src["amp"][None] * np.cos(k * r + src["phase"][None] - tau) / r).sum(1)
    np.testing.assert_allclose(n, ref, atol=1e-12)
    gfd = _fd_grad(sc, pts, h=1e-8, tau=tau)
    assert np.abs(g - gfd).max() < 1e-5 * np.abs(g).max()


def test_reference_arc_sources_map_to_the_reference_sum():
    """scene.cpp:140-160 with the arc lattice of rif_sources.h:156-198."""
    src = scenes.reference_arc_sources(radius=0.0508, center1=(-0.0508, 0, 0), center2=(-0.0508, 0, 0),
                                       active1=True, active2=True, phase1=0.3, phase2=1.1, theta_min=-0.12,
                                       theta_max=0.12, theta_sources=5, trans_z_min=-0.006, trans_z_max=0.006,
                                       trans_z_sources=4, n_scaling=5e-5, n_coeff=3)
    d, tex, _ = scenes.bessel_cube()
    d.set(field_kind=FIELD_SOURCES, f_u=5e6)
    sc = O.OracleScene(d, tex, sources=src)
    p = np.array([[0.001, -0.002, 0.0005]])
    n, _ = sc.field(p)
    k = 2 * math.pi * 5e6 / 1500.0
    ns = 20
    tot = 0.0
    for c in range(1, 4):
        sf = 4 * (-1) ** (c + 1) / ((c * math.pi) ** 2 * ns)
        t = 0.0
        for i in range(2 * ns):
            r = np.linalg.norm(p[0] - src["pos"][i])
            ph = 0.3 if i < ns else 1.1
            t += math.cos(k * r * c + ph) / r
        tot += t * sf
    assert abs(n[0] - (d.n_o + 5e-5 * tot)) < 1e-9


def _straight_scene():
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=0.0)
    d.set(n_clip=1.0)
    return d, tex


def test_integrator_straight_line_without_gradient():
    d, tex = _straight_scene()
    sc = O.OracleScene(d, tex)
    v0 = np.array([-0.6, 0.64, 0.48]) * d.n_o
    p, v = sc.er_steps([0.004, -0.001, 0.002], v0, 5e-6, 1000)
    np.testing.assert_allclose(v, v0, atol=1e-15)
    np.testing.assert_allclose(p, np.array([0.004, -0.001, 0.002]) + 1000 * 5e-6 * v0 / d.n_o, atol=1e-15)


def test_integrator_conserves_axial_angular_momentum_and_is_reversible():
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-3, mode=0, axis="z")
    sc = O.OracleScene(d, tex)
    p0 = np.array([0.003, 0.001, -0.002])
    n0, _ = sc.field(p0[None])
    v0 = n0[0] * np.array([-0.8, 0.36, 0.48])
    p, v = sc.er_steps(p0, v0, 5e-6, 1500)
    uz = np.array(d.axis_uz)
    Lz = lambda p, v: np.dot(np.cross(p - np.dot(p, uz) * uz, v), uz)
    assert abs(Lz(p, v) - Lz(p0, v0)) < 1e-15
    # the ray really bent
    assert np.linalg.norm(v / np.linalg.norm(v) - v0 / np.linalg.norm(v0)) > 1e-5
    # time reversal
    pb, vb = sc.er_steps(p, -v, 5e-6, 1500)
    np.testing.assert_allclose(pb, p0, atol=1e-14)
    np.testing.assert_allclose(-vb, v0, atol=1e-12)


def test_integrator_second_order_convergence():
    d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-3, mode=2, axis="z")
    sc = O.OracleScene(d, tex)
    p0 = np.array([0.004, 0.0015, -0.001])
    n0, _ = sc.field(p0[None])
    v0 = n0[0] * np.array([-1.0, 0.0, 0.0])
    length = 0.008
    ref_p, _ = sc.er_steps(p0, v0, length / 64000, 64000)
    errs = []
    for nsteps in (250, 500, 1000):
        p, _ = sc.er_steps(p0, v0, length / nsteps, nsteps)
        errs.append(np.linalg.norm(p - ref_p))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.3 < r1 < 5.0 and 3.3 < r2 < 5.0, errs  # order 2: ratio 4 (order 1: 2, order 4: 16)


def _fresnel_textbook(ni, nt, cos_i):
    sin_t = ni / nt * math.sqrt(max(0.0, 1 - cos_i ** 2))
    if sin_t >= 1:
        return 1.0, None
    cos_t = math.sqrt(1 - sin_t ** 2)
    rs = (ni * cos_i - nt * cos_t) / (ni * cos_i + nt * cos_t)
    rp = (ni * cos_t - nt * cos_i) / (ni * cos_t + nt * cos_i)
    return 0.5 * (rs * rs + rp * rp), sin_t


@pytest.mark.parametrize("theta_deg", [0.0, 10.0, 30.0, 45.0, 48.0, 49.0, 70.0])
def test_dielectric_bsdf_against_fresnel_and_snell(theta_deg):
    ior = 1.3333
    th = math.radians(theta_deg)
    nrm = np.array([1.0, 0.0, 0.0])  # outward normal; ray leaves the medium
    din = np.array([math.cos(th), math.sin(th), 0.0])
    F_ref, sin_t = _fresnel_textbook(ior, 1.0, math.cos(th))
    out = np.zeros(3)
    F = C_double()
    refl = O.lib().oracle_bsdf_sample(ior, din, nrm, 0.999999, out, F)
    if sin_t is None:  # total internal reflection
        assert refl == 1
        np.testing.assert_allclose(out, [-din[0], din[1], 0.0], atol=1e-15)
    else:
        assert abs(F.value - F_ref) < 1e-12
        assert refl == 0
        assert abs(out[1] - sin_t) < 1e-12 and out[0] > 0  # Snell, transmitted outward
        refl2 = O.lib().oracle_bsdf_sample(ior, din, nrm, 1e-9, out, F)
        assert refl2 == 1
        np.testing.assert_allclose(out, [-din[0], din[1], 0.0], atol=1e-15)


def C_double():
    import ctypes
    return ctypes.c_double()


@pytest.mark.parametrize("g", [0.0, 0.5, 0.9, 0.95])
def test_hg_sampling_moments_and_histogram(g):
    n = 200
    u = (np.arange(n) + 0.5) / n
    cos = []
    out = np.zeros(3)
    din = np.array([0.3, -0.5, 0.812403840463596])
    din /= np.linalg.norm(din)
    for a in u:
        O.lib().oracle_hg_sample(g, din, float(a), 0.37, out)
        cos.append(out @ din)
        assert abs(np.linalg.norm(out) - 1) < 1e-12
    cos = np.array(cos)
    assert abs(cos.mean() - g) < 2e-3
    # the sampled cosines follow the HG cdf: bins of equal probability
    edges = np.quantile(cos, np.linspace(0, 1, 11))
    pdf = lambda c: O.lib().oracle_hg_f(g, c) * 2 * math.pi
    masses = [np.trapezoid([pdf(c) for c in np.linspace(a, b, 200)], np.linspace(a, b, 200))
              for a, b in zip(edges[:-1], edges[1:])]
    np.testing.assert_allclose(masses, 0.1, atol=6e-3)
