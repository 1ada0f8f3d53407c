"""Pin the oracle's sampling and lens functions against what the mathematics
fixes: thin-lens focusing (textbook: a parallel bundle meets in one point of
the back focal plane, the chief ray is undeviated, every ray's corrected path
to the focus equals the chief ray's), the uniform hemisphere of the
next-event directions (Archimedes: height uniform on the axis; cap masses
against brute-force sphere sampling), and inverse-CDF sampling of the emitter
texture (u in (F(i-1), F(i)] picks texel i; only lit texels; row masses of
the Disk_501 texture by a chi-square test)."""
import numpy as np
import pytest

from oracle import oracle as O
from simpler_amd import scenes


def f32(x):
    """The lens holds its aperture and focal length in float32."""
    return float(np.float32(x))


def unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("d_in", [(-1.0, 0.0, 0.0), (-1.0, 0.07, -0.04), (-0.8, -0.3, 0.2)])
def test_thin_lens_focuses_a_parallel_bundle(d_in):
    f, ap = f32(0.02), f32(0.003)
    d = unit(d_in)
    rng = np.random.default_rng(1)
    r = ap * np.sqrt(rng.uniform(0, 1, 300))
    phi = rng.uniform(0, 2 * np.pi, 300)
    x0 = 0.004
    # the chief ray through the lens centre is undeviated
    ok, p_c, v_c, t_c = O.lens_deflect(ap, f, [x0, 0, 0], d, 0.25)
    assert ok
    np.testing.assert_allclose(v_c, d, atol=1e-14)
    chief_len = f / abs(d[0])
    focus = np.array([x0, 0, 0]) + chief_len * d  # on the plane x = x0 - f
    assert t_c == pytest.approx(0.25, abs=1e-15)
    for rr, ph in zip(r, phi):
        pos = [x0, rr * np.cos(ph), rr * np.sin(ph)]
        ok, p, v, t = O.lens_deflect(ap, f, pos, d, 0.25)
        assert ok
        np.testing.assert_allclose(p, pos, atol=0)  # deflection happens in the lens plane
        assert abs(np.linalg.norm(v) - 1) < 1e-14
        s = (x0 - f - p[0]) / v[0]  # to the back focal plane
        assert s > 0
        np.testing.assert_allclose(p + s * v, focus, atol=1e-13)
        # equal corrected path: total + distance to the focus = the chief ray's
        assert t + np.linalg.norm(focus - p) == pytest.approx(0.25 + chief_len, abs=1e-14)


def test_thin_lens_normal_incidence_is_isochronous():
    """On axis the correction is f - sqrt(f^2 + r^2): lens plane to focus is f
    for every ray, the optical path of a perfect lens."""
    f, ap = f32(0.015), f32(0.002)
    for rr in np.linspace(0, ap, 9):
        ok, p, v, t = O.lens_deflect(ap, f, [0, rr, 0], [-1, 0, 0], 0.0)
        assert ok
        assert t == pytest.approx(f - np.hypot(f, rr), abs=1e-16)
        np.testing.assert_allclose(v, unit([-f, -rr, 0]), atol=1e-15)


def test_thin_lens_aperture_stops_rays_outside():
    f, ap = f32(0.02), f32(0.001)
    d = unit([-1, 0.1, 0.0])
    for rr, inside in ((0.999 * ap, True), (1.001 * ap, False), (2 * ap, False)):
        for ph in np.linspace(0, 2 * np.pi, 7):
            pos = [0.0, rr * np.cos(ph), rr * np.sin(ph)]
            ok, p, v, t = O.lens_deflect(ap, f, pos, d, 1.0)
            assert ok == inside
            if not inside:  # a stopped ray leaves the lens unchanged
                np.testing.assert_array_equal(v, d)
                assert t == 1.0


def test_hemisphere_directions_are_uniform_on_the_minus_x_half():
    rng = np.random.default_rng(5)
    n = 200_000
    d = O.hemisphere(rng.uniform(size=n), rng.uniform(size=n))
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-14)
    assert np.all(d[:, 0] <= 0)
    # brute force: isotropic directions (normalised Gaussians) folded onto x < 0
    g = rng.standard_normal((n, 3))
    g /= np.linalg.norm(g, axis=1, keepdims=True)
    g[:, 0] = -np.abs(g[:, 0])
    se = np.sqrt(0.25 / n)
    for k in range(3):
        for q in (1, 2):  # first and second moments of each coordinate
            assert abs(np.mean(d[:, k] ** q) - np.mean(g[:, k] ** q)) < 6 * se, (k, q)
    assert abs(d[:, 0].mean() + 0.5) < 6 * se and abs(np.mean(d[:, 1] ** 2) - 1 / 3) < 6 * se
    # solid-angle caps about -x: mass 1 - cos(alpha) of the hemisphere
    for alpha in (0.3, 0.8, 1.2, 1.5):
        frac = np.mean(-d[:, 0] >= np.cos(alpha))
        assert abs(frac - (1 - np.cos(alpha))) < 6 * np.sqrt(frac * (1 - frac) / n)
    # octants around the axis carry equal mass (azimuth uniform)
    az = np.arctan2(d[:, 2], d[:, 1])
    counts = np.histogram(az, bins=8, range=(-np.pi, np.pi))[0]
    chi2 = ((counts - n / 8) ** 2 / (n / 8)).sum()
    assert chi2 < 30  # 7 degrees of freedom


def texture_scene(tex):
    d, _, _ = scenes.bessel_cube_small()
    return O.OracleScene(d, np.asarray(tex, np.float32))


def test_cdf_sampling_follows_the_inverse_cdf_rule():
    # texels row-major (x fastest): pmf 0, 1/2, 0 | 0, 1/4, 1/4
    sc = texture_scene([[0, 2, 0], [0, 1, 1]])
    np.testing.assert_array_equal(sc.cdf(), np.float32([0, 0, 0.5, 0.5, 0.5, 0.75, 1.0]))
    u = np.float32([0.0, 1e-7, 0.25, 0.5, np.nextafter(np.float32(0.5), np.float32(1)), 0.75, 0.76, 1.0])
    np.testing.assert_array_equal(sc.cdf_sample(u), [1, 1, 1, 1, 4, 4, 5, 5])


def test_cdf_sampling_of_the_disk_texture():
    tex = scenes.disk_texture(501)
    sc = texture_scene(tex)
    c = sc.cdf().astype(np.float64)
    rng = np.random.default_rng(11)
    n = 400_000
    u = np.concatenate([rng.uniform(size=n).astype(np.float32), np.float32([0.0, 1.0])])
    idx = sc.cdf_sample(u)
    flat = tex.reshape(-1)
    assert np.all(flat[idx] > 0)  # only lit texels
    inner = u > 0  # u in (F(i-1), F(i)] picks texel i
    assert np.all(c[idx[inner]] < u[inner]) and np.all(u[inner] <= c[idx[inner] + 1])
    rows = idx[:n] // tex.shape[1]
    expect = tex.sum(axis=1) / tex.sum() * n
    lit = expect > 0
    counts = np.bincount(rows, minlength=tex.shape[0])
    assert counts[~lit].sum() == 0
    chi2 = ((counts[lit] - expect[lit]) ** 2 / expect[lit]).sum()
    dof = lit.sum() - 1
    assert chi2 < dof + 5 * np.sqrt(2 * dof), (chi2, dof)
