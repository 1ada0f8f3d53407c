"""Closed-form pins of the oracle's next-event (NEE) weight and scatter walk.

Single scattering in a homogeneous cube has an exact expectation that a
quadrature computes independently of the tracer:

* the field is n0 everywhere (the gap planes exclude every point, so the
  index is the constant n_o: straight rays), the faces do not reflect
  (scene ior 1: the dielectric BSDF passes rays through, so a ray that
  leaves through a side face is lost), and max_depth = 1 keeps one NEE per
  photon (photon.cpp:111-116);
* a point source at the centre of the +x face shoots along -x; the first
  free flight s ~ sigma_t exp(-sigma_t s) (photon.cpp:97-104) reaches the
  vertex x = h - s when s < 2h, and the NEE at that vertex is skipped within
  1e-4 of the sensor face (scene.cpp:875);
* the NEE direction omega is uniform on the hemisphere toward -x (pdf
  1/2pi, scene.cpp:26-39); the ray reaches the sensor face x = -h when its
  transverse drift stays inside the side faces; its weight is
  Li/N 2pi exp(-sigma_t l) f_HG(d_in . omega) fore(omega)
  (scene.cpp:866-946) with l the geometric length, d_in = -x and the
  foreshortening of the exit "refraction" (scene.cpp:897-929):
  fore = 1 / sqrt(cos^2 a + n0^2 sin^2 a) for cos a = -omega_x.

So the image total is Li * int_0^{2h-1e-4} sigma e^{-sigma s}
int_{hits} e^{-sigma l} f_HG(cos a) fore(a) d omega ds.  The quadrature below
evaluates it; the plausible mistakes (a 2pi dropped or doubled, the
foreshortening left out, the HG angle taken against +x, the optical instead
of the geometric length in the attenuation) each move it by many standard
errors, which the test checks as well.

The albedo loop is pinned by its geometric law in an unbounded medium
(photons start at the centre of a 2 m cube, mean free path 1 mm): with
albedo a the scatter count S has P(S >= k) = a^(k+1) for k >= 1
(photon.cpp:19-38,97-116), max_depth D caps it at D, and max_pathlength P
gates the whole loop on the first flight alone (the reference never updates
totalDist, photon.cpp:112-114).  The mean ER step count per flight, with
precision 0 (no refinement: the overshooting step ends the flight,
scene.cpp:404-407), is E floor(s/h) + 1 = 1/(e^{sigma h} - 1) + 1, which pins
the mean of the free-path sampler -mfp ln u.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from simpler_amd import scenes

H = 0.005      # half side of the cube (m)
N0 = 1.3333    # index of the medium
SIG = 100.0    # sigma_t (1/m): mean free path 10 mm
G = 0.5        # Henyey-Greenstein g
LI = 75000.0


def _single_scatter_desc():
    d, tex, _ = scenes.bessel_cube(step=1e-4, n_max=1e-5, mode=0, res=(16, 16, 1))
    d.set(ior=1.0, n_o=N0, egap_end_loc_x=-1.0, sgap_begin_loc_x=1.0,  # n = n_o at every point
          light_plane=(1e-9, 1e-9), half_theta_limit=0.0, light_origin=(H, 0.0, 0.0),
          emitter_lens_origin=(H, 0.0, 0.0), view_plane=(0.1, 0.1), pathlength_range=(-1.0, -1.0),
          sigma_t=SIG, albedo=1.0, g=G, max_depth=1, use_direct=0, use_angular_sampling=1, Li=LI)
    return d, np.ones((1, 1), np.float32), (16, 16, 1)


def hg(cos, g=G):
    return (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * cos) ** 1.5)


def _gl(n, a, b):
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


def single_scatter_total(two_pi=2 * math.pi, fore_on=True, hg_sign=1.0, optical=False):
    """Li * E[NEE weight] by Gauss-Legendre quadrature; the keyword arguments
    plant the plausible mistakes (two_pi replaces the 2 pi factor combined
    with the 1/2pi pdf)."""
    s, ws = _gl(96, 0.0, 2 * H - 1e-4)
    phi, wphi = _gl(48, 0.0, math.pi / 4)  # 8-fold symmetry of the square exit face
    total = 0.0
    for si, wsi in zip(s, ws):
        x = H - si
        dist_x = x + H  # to the sensor face
        acc = 0.0
        for p, wp in zip(phi, wphi):
            amax = math.atan(H / (dist_x * max(abs(math.cos(p)), abs(math.sin(p)))))
            a, wa = _gl(64, 0.0, amax)
            ca, sa = np.cos(a), np.sin(a)
            length = dist_x / ca
            att = np.exp(-SIG * length * (N0 if optical else 1.0))
            fore = 1.0 / np.sqrt(ca * ca + N0 * N0 * sa * sa) if fore_on else 1.0
            acc += wp * np.sum(wa * att * hg(hg_sign * ca) * fore * sa)
        inner = 8.0 * acc * (two_pi / (2 * math.pi))
        total += wsi * SIG * math.exp(-SIG * si) * inner
    return LI * total


@pytest.fixture(scope="module")
def single_scatter_batches():
    d, tex, res = _single_scatter_desc()
    sc = O.OracleScene(d, tex)
    B, n = 10, 40000
    tots = []
    for b in range(B):
        img, st = sc.render(n, res, seed=23, offset=b * n, n_total=n, precision=64, threads=8)
        assert st["direct_splats"] == 0
        tots.append(img.sum())
    return np.array(tots)


def test_single_scatter_nee_total_matches_quadrature(single_scatter_batches):
    t = single_scatter_batches
    mean, se = t.mean(), t.std(ddof=1) / math.sqrt(len(t))
    ref = single_scatter_total()
    print(f"single scatter: oracle {mean:.6g} +- {se:.2g}, quadrature {ref:.6g}")
    assert abs(mean - ref) < 4 * se
    assert se < 0.01 * ref


@pytest.mark.parametrize("mistake", ["2pi_dropped", "2pi_doubled", "no_foreshortening", "hg_against_plus_x",
                                     "optical_length"])
def test_planted_mistakes_are_far_outside_the_error(single_scatter_batches, mistake):
    """The quadrature with each plausible mistake differs from the oracle by
    more than 10 standard errors: the pin above would catch it."""
    t = single_scatter_batches
    mean, se = t.mean(), t.std(ddof=1) / math.sqrt(len(t))
    kw = {"2pi_dropped": dict(two_pi=1.0), "2pi_doubled": dict(two_pi=4 * math.pi),
          "no_foreshortening": dict(fore_on=False), "hg_against_plus_x": dict(hg_sign=-1.0),
          "optical_length": dict(optical=True)}[mistake]
    wrong = single_scatter_total(**kw)
    assert abs(mean - wrong) > 10 * se, (mistake, mean, wrong, se)


def test_max_depth_zero_has_no_next_event():
    d, tex, res = _single_scatter_desc()
    d.set(max_depth=0)
    img, st = O.OracleScene(d, tex).render(2000, res, seed=3, precision=64, threads=8)
    assert st["nee_splats"] == 0 and img.sum() == 0.0


def _walk_desc(albedo, max_depth=-1, max_pathlength=-1.0, precision=4, step=1e-4):
    """2 m cube, photons start at its centre (the source sits on the +x face
    and the inactive emitter lens plane at x = 0 moves them there), mean free
    path 1 mm: no photon of the walk can reach a face."""
    d, tex, _ = scenes.bessel_cube(step=step, n_max=1e-5, mode=0, res=(4, 4, 1))
    half = 1.0
    d.set(block_l=(-half, -half, -half), block_r=(half, half, half), light_origin=(half, 0.0, 0.0),
          emitter_lens_origin=(0.0, 0.0, 0.0), egap_end_loc_x=-2.0, sgap_begin_loc_x=2.0, ior=1.0,
          light_plane=(1e-6, 1e-6), sigma_t=1000.0, albedo=albedo, g=0.7, max_depth=max_depth,
          max_pathlength=max_pathlength, use_direct=0, use_angular_sampling=0, precision=precision,
          view_origin=(-half, 0.0, 0.0), sensor_lens_origin=(-half, 0.0, 0.0))
    return d, np.ones((1, 1), np.float32)


def _walk(d, tex, n=20000, seed=5):
    _, st = O.OracleScene(d, tex).render(n, (4, 4, 1), seed=seed, precision=64, threads=8)
    assert st["aborted"] == 0
    return st


def _moments(pmf):
    k = np.arange(len(pmf))
    m = float((k * pmf).sum())
    return m, float((k * k * pmf).sum()) - m * m


def _walk_pmf(a, kmax=400):
    """Scatter count S and flight count F = S + 1{first albedo pass}: the
    first pass has probability a, then K scatters with P(K >= k) = a^k."""
    k = np.arange(kmax)
    pk = (1 - a) * a ** k  # K = k given the first pass
    ps = np.zeros(kmax + 1)
    ps[0] = 1 - a
    ps[:kmax] += a * pk
    pf = np.zeros(kmax + 2)
    pf[0] = 1 - a
    pf[1:kmax + 1] += a * pk
    return ps, pf


@pytest.mark.parametrize("a", [0.3, 0.7, 0.9])
def test_albedo_loop_follows_the_geometric_law(a):
    d, tex = _walk_desc(a)
    n = 20000
    st = _walk(d, tex, n)
    ps, pf = _walk_pmf(a)
    ms, vs = _moments(ps)
    assert abs(ms - a * a / (1 - a)) < 1e-9
    got = st["scatters"] / n
    assert abs(got - ms) < 4 * math.sqrt(vs / n), (got, ms)
    # each flight is one march segment in the unbounded medium
    mf, vf = _moments(pf)
    flights = st["segments"] / n
    assert abs(flights - mf) < 4 * math.sqrt(vf / n), (flights, mf)


@pytest.mark.parametrize("depth", [1, 2, 5])
def test_max_depth_caps_the_walk(depth):
    a = 0.8
    d, tex = _walk_desc(a, max_depth=depth)
    n = 20000
    st = _walk(d, tex, n)
    # S = min(K, D) with P(K >= k) = a^k, after the first albedo pass (prob a)
    pk = np.array([a ** k for k in range(1, depth + 1)])
    mean = a * pk.sum()
    e2 = a * sum((2 * k - 1) * a ** k for k in range(1, depth + 1))
    sd = math.sqrt(e2 - mean * mean)
    got = st["scatters"] / n
    assert abs(got - mean) < 4 * sd / math.sqrt(n), (got, mean)
    # an off-by-one in the loop bound moves the mean by a^(D+2) or more
    wrong = a * sum(a ** k for k in range(1, depth + 2))
    assert abs(wrong - mean) > 8 * sd / math.sqrt(n)


def test_max_pathlength_gates_on_the_first_flight_only():
    a, sig = 0.8, 1000.0
    P = 0.7e-3  # 0.7 mm
    d, tex = _walk_desc(a, max_pathlength=P)
    n = 20000
    st = _walk(d, tex, n)
    pgate = 1 - math.exp(-sig * P)
    mean = a * pgate * a / (1 - a)
    e2 = a * pgate * a * (1 + a) / (1 - a) ** 2
    sd = math.sqrt(e2 - mean * mean)
    got = st["scatters"] / n
    assert abs(got - mean) < 4 * sd / math.sqrt(n), (got, mean)


def test_free_path_mean_from_the_step_count():
    """precision 0: every flight costs floor(s/h) + 1 ER steps."""
    a, sig, h = 0.5, 1000.0, 1e-4
    d, tex = _walk_desc(a, precision=0, step=h)
    n = 40000
    st = _walk(d, tex, n)
    per_flight = 1.0 / math.expm1(sig * h) + 1.0
    flights = st["segments"]
    got = st["er_steps"] / flights
    # floor(s/h) is geometric with ratio q = e^{-sigma h}: variance q/(1-q)^2
    q = math.exp(-sig * h)
    sd = math.sqrt(q) / (1 - q)
    assert abs(got - per_flight) < 4 * sd / math.sqrt(flights), (got, per_flight)
    # a free path of -mfp ln(1-u) has the same law, but -2 mfp ln u or a
    # wrong mean free path would move it by ~ 10 steps per flight
    assert abs(got - (1.0 / math.expm1(sig * h / 1.1) + 1.0)) > 8 * sd / math.sqrt(flights)
