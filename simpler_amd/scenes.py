"""Scene builders for the workloads this engine is measured and tested on.

All lengths in metres.  Geometry follows the reference drivers: the source
sits on the +x face shooting along -x, the orthographic sensor on the -x face
looking along -x with horizontal (0,-1,0), lenses inactive, gap planes on the
two x faces (reference renderer/pybindFiles/tests/test_renderer3d_sample.py:
216-245, renderer/samples/renderer3d_sample.cpp).
"""
import math

import numpy as np

from ._abi import (FIELD_BESSEL, FIELD_SOURCES, FIELD_SPLINE, FLAG_GENERAL_AXIS, FLAGS_DEFAULT,
                   SceneDesc)


def disk_texture(n=501):
    """Inscribed-disk emitter texture, identical to the reference's
    renderer/images/Disk_501.pfm for n = 501: texel (x, y) is 1 where
    (x-c)^2 + (y-c)^2 <= (c+1)^2 with c = (n-1)/2, else 0."""
    c = (n - 1) / 2.0
    yy, xx = np.mgrid[0:n, 0:n]
    return (((xx - c) ** 2 + (yy - c) ** 2) <= (c + 1) ** 2).astype(np.float32)


def base_desc(half=0.005, ior=1.3333, step=1e-5, precision=4,
              sensor_plane=0.010, emitter_size=0.002, pl_range=(0.0, 0.1)):
    """Cube of side 2*half centred at the origin, collimated textured
    source on the +x face, sensor on the -x face."""
    d = SceneDesc()
    d.set(block_l=(-half, -half, -half), block_r=(half, half, half), ior=ior,
          light_origin=(half, 0.0, 0.0), light_dir=(-1.0, 0.0, 0.0),
          half_theta_limit=0.0, light_plane=(emitter_size, emitter_size),
          Li=75000.0, emitter_lens_origin=(half, 0.0, 0.0),
          emitter_lens_aperture=0.015, emitter_lens_focal_length=0.015,
          emitter_lens_active=0,
          view_origin=(-half, 0.0, 0.0), view_dir=(-1.0, 0.0, 0.0),
          view_horizontal=(0.0, -1.0, 0.0),
          view_plane=(sensor_plane, sensor_plane),
          pathlength_range=pl_range, use_bounce_decomposition=0,
          sensor_lens_origin=(-half, 0.0, 0.0), sensor_lens_aperture=0.015,
          sensor_lens_focal_length=0.015, sensor_lens_active=0,
          field_kind=FIELD_BESSEL, n_o=ior, f_u=848e3, speed_u=1500.0,
          n_max=1e-5, n_clip=1e-5,
          phi_min=math.pi / 2 - 0.2 * math.pi,
          phi_max=math.pi / 2 + 0.2 * math.pi, mode=0,
          axis_uz=(0.0, 0.0, 1.0), axis_ux=(1.0, 0.0, 0.0),
          p_u=(0.0, 0.0, 0.0), er_stepsize=step, precision=precision,
          egap_end_loc_x=half, sgap_begin_loc_x=-half, time_coupling=0.0,
          sigma_t=1000.0, albedo=1.0, g=0.9, max_depth=-1,
          max_pathlength=-1.0, use_direct=1, use_angular_sampling=1,
          flags=FLAGS_DEFAULT)
    return d


# Workloads ---------------------------------------------------------------
# Each returns (desc, texture, (X, Y, Z)).

def bessel_cube(step=1e-5, n_max=1e-5, mode=0, res=(128, 128, 64),
                sigma_s=1000.0, g=0.9, axis="z"):
    """10 mm cube, n0 = 1.3333, cylindrical transducer (Bessel J_m) with a
    20 % duty window, 848 kHz, 1500 m/s; HG scattering; collimated 1 mm
    radius disk source; orthographic sensor, path length 0-100 mm."""
    d = base_desc(step=step)
    d.set(n_max=n_max, n_clip=n_max, mode=mode, sigma_t=sigma_s, g=g)
    if axis == "x":  # the reference drivers' axis (rif_base.h:146-148)
        d.set(axis_uz=(-1.0, 0.0, 0.0), axis_ux=(0.0, 0.0, 1.0))
    return d, disk_texture(501), res


def bessel_cube_small():
    """CPU-sized scene: 10 um ER step, 128x128 px x 64 path-length bins,
    Delta n = 1e-5, mode 0."""
    return bessel_cube(step=1e-5, n_max=1e-5, mode=0, res=(128, 128, 64))


def bessel_cube_full(mode=0):
    """Full-size scene: 5 um ER step, 256x256 px x 128 bins,
    Delta n = 2e-5, mode 0 or 2."""
    return bessel_cube(step=5e-6, n_max=2e-5, mode=mode, res=(256, 256, 128))


def transparent_cube(mode=0, step=5e-6):
    """Same cube without scattering (sigma_s = 0: the Medium forces
    albedo 0, so only the direct term remains)."""
    d, tex, res = bessel_cube(step=step, n_max=2e-5, mode=mode,
                              res=(256, 256, 128))
    d.set(sigma_t=0.0)
    return d, tex, res


def ring_sources(n=16, radius=0.010, seed=0, amp_max=5e-9, z=0.0):
    """Point-source ring for the travelling-wave field: n elements on a
    circle of the given radius in the plane z = const, amplitudes
    U[0, amp_max] (units: refractive index x metre), phases U[0, 2 pi)."""
    rng = np.random.default_rng(seed)
    th = 2 * math.pi * np.arange(n) / n
    pos = np.stack([radius * np.cos(th), radius * np.sin(th),
                    np.full(n, z)], axis=1).astype(np.float32)
    amp = rng.uniform(0.0, amp_max, n).astype(np.float32)
    phase = rng.uniform(0.0, 2 * math.pi, n).astype(np.float32)
    return pos, amp, phase


def travelling_wave_cube(n_src=16, f=1e6, sigma_s=2000.0, g=0.95, seed=0,
                         step=5e-6, res=(64, 64, 32)):
    """Travelling-wave field of a 16-element ring: Delta n =
    sum a_i sin(k|x - x_i| - w t + phi_i)/|x - x_i|, f = 1 MHz.  Written as
    cos(k r + phi_i - pi/2 - tau) with the per-photon duty phase tau = w t."""
    d = base_desc(step=step)
    d.set(field_kind=FIELD_SOURCES, f_u=f, sigma_t=sigma_s, g=g,
          time_coupling=1.0, phi_min=0.0, phi_max=2 * math.pi)
    pos, amp, phase = ring_sources(n_src, seed=seed)
    sources = dict(pos=pos, amp=amp, phase=(phase - math.pi / 2).astype(np.float32),
                   coeff=np.ones(1, np.float32))
    return d, disk_texture(501), res, sources


def reference_arc_sources(radius, center1, center2, active1, active2,
                          phase1, phase2, theta_min, theta_max, theta_sources,
                          trans_z_min, trans_z_max, trans_z_sources,
                          n_scaling, n_coeff):
    """The reference's two-group arc lattice of point sources and its
    harmonic weights (rif_sources.h:156-198, scene.cpp:140-160), mapped to
    the general point-source field:
      weight_c = n_scaling * 4 (-1)^(c+1) / ((c pi)^2 * nsources),
    amplitude 1, phase phase1 / phase2 by group."""
    nsrc = theta_sources * trans_z_sources
    th_d = (theta_max - theta_min) / (theta_sources - 1) if theta_sources > 1 else 0.0
    tz_d = (trans_z_max - trans_z_min) / (trans_z_sources - 1) if trans_z_sources > 1 else 0.0
    pos, ph = [], []
    for grp, (active, cen, phs) in enumerate(((active1, center1, phase1),
                                              (active2, center2, phase2))):
        if not active:
            continue
        pts = np.zeros((nsrc, 3), np.float64)
        for i in range(theta_sources):
            theta = theta_min + th_d * i
            xval = radius * (1 - math.cos(theta))
            yval = radius * math.sin(theta)
            for j in range(trans_z_sources):
                idx = i + j * theta_sources
                if grp == 0:
                    pts[idx] = (trans_z_min + tz_d * j + cen[0], yval + cen[1], xval + cen[2])
                else:
                    pts[idx] = (trans_z_min + tz_d * j + cen[0], xval + cen[1], yval + cen[2])
        pos.append(pts)
        ph.append(np.full(nsrc, phs))
    pos = np.concatenate(pos).astype(np.float32)
    ph = np.concatenate(ph).astype(np.float32)
    coeff = np.array([n_scaling * 4 * (-1) ** (c + 1) / ((c * math.pi) ** 2 * nsrc)
                      for c in range(1, n_coeff + 1)], np.float32)
    return dict(pos=pos, amp=np.ones(len(pos), np.float32), phase=ph, coeff=coeff)


def gaussian_spline_coeffs(n=64, half=0.005, n0=1.3333, max_dn=5e-5,
                           sigma_cells=4.0, seed=0):
    """Random smooth index field on an n^3 grid over [-half, half]^3:
    n0 + seeded normal noise Gaussian-filtered (sigma in cells, periodic
    FFT filter) and rescaled to max |dn| = max_dn.  Returns node VALUES
    (x fastest); prefilter them into B-spline coefficients before use."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((n, n, n))
    k = np.fft.fftfreq(n)
    kk = k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2
    filt = np.exp(-2 * (math.pi ** 2) * (sigma_cells ** 2) * kk)
    sm = np.real(np.fft.ifftn(np.fft.fftn(noise) * filt))
    sm *= max_dn / np.abs(sm).max()
    # array index order (z, y, x) so that the flat layout is x fastest
    return (n0 + sm).astype(np.float64), (n, n, n), (-half,) * 3, (half,) * 3


def fus_sources_scene(res=(1024, 1024, 1)):
    """The reference's renderings/settings/fus.json scene: a 12.7 mm slab
    (x in [-6.35, 6.35] mm, y and z +-5 m), the two 100 x 51 arc lattices of
    point sources (radius 50.8 mm, chord 12.7 mm, centres (0, 0, -50.8) and
    (0, -50.8, 0) mm, phase pi each, n_scaling 2.15e-5, one harmonic), 2 MHz,
    1500 m/s, transparent (sigma_t 0: the direct term only), 100 um ER step,
    1 mm emitter and sensor, path length 0-100 m.  Returns (desc, texture,
    res, sources) with the Disk texture at 101 x 101 texels."""
    half = 0.00635
    d = base_desc(half=half, step=1e-4, precision=4, sensor_plane=0.001, emitter_size=0.001,
                  pl_range=(0.0, 100000.0))
    d.set(field_kind=FIELD_SOURCES, f_u=2e6, speed_u=1500.0, n_o=1.3333, sigma_t=0.0, albedo=0.9, g=0.9,
          use_direct=1, time_coupling=0.0, block_l=(-half, -5.0, -5.0), block_r=(half, 5.0, 5.0),
          emitter_lens_aperture=0.3, emitter_lens_focal_length=0.3, sensor_lens_aperture=0.3,
          sensor_lens_focal_length=0.3)
    th = math.asin(0.0127 / (2 * 0.0508))
    src = reference_arc_sources(0.0508, (0.0, 0.0, -0.0508), (0.0, -0.0508, 0.0), True, True, math.pi, math.pi,
                                -th, th, 100, -0.0127 / 2, 0.0127 / 2, 51, 2.15e-5, 1)
    return d, disk_texture(101), res, src


def vol_bessel_desc(general_axis=True):
    """The Bessel field of renderings/testing_splines.sh (f_u 898.1 kHz,
    1500 m/s, n_o 1.3333, n_max 0.0278, n_clip 1, phi = pi/2, m = 0, axis
    -x through the origin) in its 30 mm slab (x in +-15 mm): the field that
    the reference's renderings/besselriffile.vol samples on a 101^3 grid."""
    d = base_desc(half=0.015, step=1e-3, precision=4)
    d.set(block_l=(-0.015, -5.0, -5.0), block_r=(0.015, 5.0, 5.0), egap_end_loc_x=0.015,
          sgap_begin_loc_x=-0.015, f_u=898100.0, speed_u=1500.0, n_o=1.3333, n_max=0.0278, n_clip=1.0,
          phi_min=math.pi / 2, phi_max=math.pi / 2, mode=0, axis_uz=(-1.0, 0.0, 0.0), axis_ux=(0.0, 0.0, 1.0),
          p_u=(0.0, 0.0, 0.0))
    if not general_axis:
        d.set(flags=d.flags & ~FLAG_GENERAL_AXIS)
    return d


def spline_cube(n=64, sigma_s=2000.0, g=0.9, step=5e-6, res=(64, 64, 32), seed=0):
    """Cube with a tricubic B-spline index field (node values from
    gaussian_spline_coeffs)."""
    d = base_desc(step=step)
    d.set(field_kind=FIELD_SPLINE, sigma_t=sigma_s, g=g)
    vals, dims, lo, hi = gaussian_spline_coeffs(n=n, seed=seed)
    return d, disk_texture(501), res, (vals, dims, lo, hi)


def prefilter3d(vals, dims):
    """B-spline coefficients of node values (x fastest, dims (N0, N1, N2))
    with the engine's host prefilter (scene_pybind.prefilter3d)."""
    import sys
    from . import build as _build
    if _build.LIB not in sys.path:
        sys.path.insert(0, _build.LIB)
    import scene_pybind
    N0, N1, N2 = dims
    a = np.ascontiguousarray(np.asarray(vals, np.float32).reshape(N2, N1, N0))
    return np.asarray(scene_pybind.prefilter3d(a), np.float32).reshape(-1)


def spline_cube_ready(n=64, sigma_s=2000.0, g=0.9, step=5e-6, res=(64, 64, 32), seed=0):
    """spline_cube with the node values already prefiltered into
    coefficients: (desc, texture, res, (coeffs, dims, xmin, xmax))."""
    d, tex, res, (vals, dims, lo, hi) = spline_cube(n=n, sigma_s=sigma_s, g=g, step=step, res=res, seed=seed)
    return d, tex, res, (prefilter3d(vals, dims), dims, lo, hi)


def aperture_weights(res, view_plane=0.010, diameter=0.001):
    """dF/dimage of the throughput into a centred circular aperture on the
    sensor: 1 on pixels whose centre lies inside, all time bins; (Z, Y, X)."""
    X, Y, Z = res
    px = view_plane / X
    yy, xx = np.mgrid[0:Y, 0:X]
    r2 = ((xx + 0.5 - X / 2) * px) ** 2 + ((yy + 0.5 - Y / 2) * px) ** 2
    m = (r2 <= (diameter / 2) ** 2).astype(np.float32)
    return np.repeat(m[None], Z, axis=0)


def splitmix64(x):
    """SplitMix64 finaliser (Steele, Lea, Flood 2014): seed derivation."""
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


def batched_configs(n_cfg=256, seed=0, budget_lo=1e5, budget_hi=1e7, scale_budgets=1.0):
    """Transducer sweep: Delta n ~ U[0, 2e-5], f ~ U[0.5, 2] MHz, mode in
    {0, 1, 2}, duty-window start phi ~ U[0, 2 pi) (window length 0.4 pi),
    sigma_s ~ U[0.5, 3] /mm, g ~ U[0.8, 0.95], photon budgets log-uniform in
    [budget_lo, budget_hi] (times scale_budgets), seeds derived from the batch
    seed.  Returns (base desc, texture, per-configuration dicts, resolution
    64 x 64 x 32)."""
    rng = np.random.default_rng(seed)
    d = base_desc(step=5e-6)
    d.set(n_max=1e-5, n_clip=1e-5)
    cfgs = []
    for i in range(n_cfg):
        dn = float(rng.uniform(0.0, 2e-5))
        phi = float(rng.uniform(0.0, 2 * math.pi))
        n = int(round(10 ** rng.uniform(math.log10(budget_lo), math.log10(budget_hi)) * scale_budgets))
        cfgs.append(dict(n_max=max(dn, 1e-12), n_clip=max(dn, 1e-12), f_u=float(rng.uniform(0.5e6, 2e6)),
                         phi_min=phi, phi_max=phi + 0.4 * math.pi, mode=int(rng.integers(0, 3)),
                         sigma_t=float(rng.uniform(500.0, 3000.0)), albedo=1.0, g=float(rng.uniform(0.8, 0.95)),
                         n_photons=max(n, 1), seed=splitmix64(seed * 1000003 + i)))
    return d, disk_texture(501), cfgs, (64, 64, 32)


def desc_for_config(base, cfg):
    """The single-scene descriptor equivalent to one batched configuration."""
    d = base.copy()
    d.set(**{k: v for k, v in cfg.items() if k not in ("n_photons", "seed")})
    return d
