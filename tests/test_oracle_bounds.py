"""The gradient bounds the next-event skip rests on, against the oracle's
field (FP64).

The engine leaves a next-event sub-path untraced when its straight line,
widened by G s^2 / 2 over a length s and turned by at most G s, cannot reach
the objective; that is exact only if G >= |grad n| (and >= |grad n| / n)
everywhere in the block.  The bounds, as the engine's host code forms them:

* spline fields: |dn/dx_a| <= res_a max|c_{i+1} - c_i| along axis a (the
  cubic B-spline's derivative is a convex combination of coefficient
  differences), G = |(res_a dmax_a)_a|;
* point sources, n = n_o + sum_i a_i sum_c w_c cos(c k r_i + phi_i) / r_i:
  |grad n| <= sum_i |a_i| sum_c |w_c| (1 / d_i^2 + c k / d_i), d_i the
  source's distance from the block, and n >= n_o - sum_i |a_i| sum_c |w_c| / d_i;
* Bessel fields: |grad n| <= 2 a k, a the amplitude times its largest duty
  scaling, k the wavenumber.

Checked at random points of the block and at the block's points nearest to
each source (where a source's term peaks), for the bench's ring and for
random pattern amplitudes and phases."""
import math

import numpy as np

from oracle import oracle as O
from simpler_amd import scenes


def _block_points(d, n, seed):
    lo, hi = np.array(d.block_l, np.float64), np.array(d.block_r, np.float64)
    return np.random.default_rng(seed).uniform(lo, hi, (n, 3)), lo, hi


def _sources_bound(d, pos, amp, coeff):
    lo, hi = np.array(d.block_l, np.float64), np.array(d.block_r, np.float64)
    k = 2.0 * math.pi * d.f_u / d.speed_u
    nearest = np.clip(pos.astype(np.float64), lo, hi)
    dist = np.linalg.norm(pos - nearest, axis=1)
    assert (dist > 0).all()
    c = np.arange(1, len(coeff) + 1)
    w = np.abs(coeff)
    g = sum(abs(a) * (w * (1.0 / di ** 2 + c * k / di)).sum() for a, di in zip(amp, dist))
    dn = sum(abs(a) * w.sum() / di for a, di in zip(amp, dist))
    return g, dn, nearest


def test_point_source_gradient_bound_holds_over_the_block():
    d, tex, _, src = scenes.travelling_wave_cube()
    rng = np.random.default_rng(3)
    cases = [src] + [dict(src, amp=rng.uniform(0.0, 5e-9, len(src["amp"])).astype(np.float32),
                          phase=rng.uniform(0.0, 2 * math.pi, len(src["amp"])).astype(np.float32))
                     for _ in range(4)]
    for j, s in enumerate(cases):
        osc = O.OracleScene(d, tex, sources=s)
        G, dn, nearest = _sources_bound(d, np.asarray(s["pos"], np.float64), np.asarray(s["amp"], np.float64),
                                        np.asarray(s["coeff"], np.float64))
        pts, lo, hi = _block_points(d, 50_000, seed=j)
        # the block's points nearest to each source, and a shell just inside
        # them, where the 1 / r terms are largest
        inward = (lo + hi) / 2 - nearest
        inward /= np.linalg.norm(inward, axis=1, keepdims=True)
        near = np.concatenate([nearest + t * inward for t in (0.0, 1e-5, 1e-4, 1e-3)])
        for tau in (0.0, 1.3, 4.0):
            n, g = osc.field(np.concatenate([pts, near]), tau=tau)
            gmax = np.linalg.norm(g, axis=1).max()
            assert gmax <= G, (j, tau, gmax, G)
            assert n.min() >= d.n_o - dn - 1e-12, (j, tau)
        print(f"sources case {j}: max |grad n| {gmax:.4g} of the bound {G:.4g}", flush=True)
        assert gmax >= 0.05 * G  # the bound is not vacuous


def test_point_source_bound_is_tight_for_one_source():
    """One source at its most favourable phase attains the bound to a few
    per cent at the block's nearest point (|grad n| = a sqrt(1 + (k r)^2) /
    r^2 there, against a (1 + k r) / r^2): a bound that dropped the 1 / d^2
    term, or took the distance to the block's centre, would fall below
    the field's gradient."""
    d, tex, _, src = scenes.travelling_wave_cube()
    pos = np.asarray(src["pos"], np.float64)
    lo, hi = np.array(d.block_l, np.float64), np.array(d.block_r, np.float64)
    i = int(np.argmin(np.linalg.norm(pos - np.clip(pos, lo, hi), axis=1)))
    gmax, G = 0.0, None
    for phase in np.linspace(0.0, 2 * math.pi, 64, endpoint=False):
        s = dict(pos=pos[i:i + 1].astype(np.float32), amp=np.array([5e-9], np.float32),
                 phase=np.array([phase], np.float32), coeff=np.ones(1, np.float32))
        G, _, nearest = _sources_bound(d, pos[i:i + 1], np.array([5e-9]), np.ones(1))
        n, g = O.OracleScene(d, tex, sources=s).field(nearest)
        gmax = max(gmax, float(np.linalg.norm(g, axis=1).max()))
    print(f"one source: max |grad n| {gmax:.5g} of the bound {G:.5g}", flush=True)
    assert 0.9 * G <= gmax <= G


def test_bessel_gradient_bound_holds_over_the_block():
    """|grad n| <= 2 a k for the Bessel field, a = n_max min(1, n_clip /
    n_max), at random points, near the axis (where the gradient is clamped)
    and for every duty scaling, orders 0-3 and both axis forms."""
    for mode in (0, 1, 2, 3):
        for axis in ("z", "x"):
            d, tex, _ = scenes.bessel_cube(step=5e-6, n_max=2e-5, mode=mode, axis=axis)
            osc = O.OracleScene(d, tex)
            a = abs(d.n_max) * min(1.0, abs(d.n_clip / d.n_max))
            k = 2.0 * math.pi * d.f_u / d.speed_u
            G = 2.0 * a * k
            pts, _, _ = _block_points(d, 20_000, seed=mode)
            ax = np.array(d.axis_uz, np.float64)
            t = np.random.default_rng(7).uniform(-0.005, 0.005, (4000, 1))
            off = np.random.default_rng(8).normal(size=(4000, 3)) * 1e-7
            near_axis = t * ax + off - (off @ ax)[:, None] * ax  # within a few 1e-7 m of the axis
            gmax = 0.0
            for scaling in (1.0, -1.0, 0.3):
                n, g = osc.field(np.concatenate([pts, near_axis]), scaling=scaling)
                gmax = max(gmax, float(np.linalg.norm(g, axis=1).max()))
                assert n.min() >= d.n_o - a - 1e-12
            print(f"bessel m={mode} axis {axis}: max |grad n| {gmax:.4g} of the bound {G:.4g}", flush=True)
            assert gmax <= G, (mode, axis)
            assert gmax >= 0.1 * G  # not vacuous (m = 0: max |J_1| / 2 = 0.29)


def test_spline_gradient_bound_holds_over_the_block():
    d, tex, _, (coeff, dims, lo, hi) = scenes.spline_cube_ready(n=64)
    osc = O.OracleScene(d, tex, spline=(coeff, dims, lo, hi))
    c = np.asarray(coeff, np.float64).reshape(dims[2], dims[1], dims[0])
    dmax = np.array([np.abs(np.diff(c, axis=2)).max(), np.abs(np.diff(c, axis=1)).max(),
                     np.abs(np.diff(c, axis=0)).max()])
    res = (np.array(dims, np.float64) - 1) / (np.array(hi, np.float64) - np.array(lo, np.float64))
    G = np.linalg.norm(dmax * res)
    assert c.min() >= 1.0  # then |grad n| / n <= |grad n|
    pts, _, _ = _block_points(d, 100_000, seed=9)
    n, g = osc.field(pts)
    gabs = np.abs(g).max(axis=0)
    gmax = np.linalg.norm(g, axis=1).max()
    print(f"spline: max |grad n| {gmax:.4g} of the bound {G:.4g}; per axis {gabs} of {dmax * res}", flush=True)
    assert (gabs <= dmax * res * (1 + 1e-9)).all()
    assert gmax <= G
    assert n.min() >= c.min() - 1e-9
    assert gmax >= 0.05 * G
