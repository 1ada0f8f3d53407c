"""ctypes wrapper of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from simpler_amd._abi import SceneDesc, Stats

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")
_LIBS = {}


def build():
    """Compile both oracle libraries with the committed Makefile."""
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = os.path.join(BUILD, name)
    if not os.path.exists(path):
        build()
    L = C.CDLL(path)
    P = C.c_void_p
    dp = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
    fp = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    ip = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    L.oracle_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.oracle_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
    L.oracle_uniform.restype = C.c_float
    L.oracle_besselj.argtypes = [C.c_int, C.c_double]
    L.oracle_besselj.restype = C.c_double
    L.oracle_scene_create.argtypes = [C.POINTER(SceneDesc), P, C.c_int, C.c_int]
    L.oracle_scene_create.restype = P
    L.oracle_scene_destroy.argtypes = [P]
    L.oracle_scene_update.argtypes = [P, C.POINTER(SceneDesc)]
    L.oracle_scene_get_cdf.argtypes = [P, fp, C.c_int]
    L.oracle_scene_get_cdf.restype = C.c_int
    L.oracle_scene_set_sources.argtypes = [P, C.c_int, fp, fp, fp, C.c_int, fp]
    L.oracle_scene_set_spline.argtypes = [P, fp, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.oracle_render.argtypes = [P, C.c_int, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, C.c_int,
                                dp, C.c_int, C.c_int, C.c_int, C.POINTER(Stats), C.c_int, C.POINTER(C.c_double)]
    L.oracle_dead_path_cut.argtypes = [P]
    L.oracle_dead_path_cut.restype = C.c_double
    L.oracle_trace_rays.argtypes = [P, C.c_int, C.c_int64, C.c_uint64, C.c_uint64, dp, dp, dp, ip, C.c_int]
    L.oracle_field_eval.argtypes = [P, C.c_int, C.c_int, dp, C.c_double, C.c_double, dp, dp]
    L.oracle_er_steps.argtypes = [P, C.c_int, dp, dp, C.c_double, C.c_long, C.c_double, C.c_double]
    L.oracle_field_hessian.argtypes = [P, C.c_int, dp, C.c_double, C.c_double, dp]
    L.oracle_nee_connect.argtypes = [P, C.c_int, dp, dp, dp, C.c_double, C.c_double, dp, dp]
    L.oracle_hg_sample.argtypes = [C.c_double, dp, C.c_double, C.c_double, dp]
    L.oracle_hg_f.argtypes = [C.c_double, C.c_double]
    L.oracle_hg_f.restype = C.c_double
    L.oracle_bsdf_sample.argtypes = [C.c_double, dp, dp, C.c_double, dp, C.POINTER(C.c_double)]
    L.oracle_bsdf_sample.restype = C.c_int
    L.oracle_spline_kernel.argtypes = [C.c_int, C.c_int, dp, dp]
    L.oracle_spline_build1d.argtypes = [dp, C.c_int, dp]
    L.oracle_spline_eval1d.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp]
    L.oracle_spline_build2d.argtypes = [dp, ip, dp]
    L.oracle_spline_eval2d.argtypes = [dp, ip, dp, dp, C.c_int, C.c_int, C.c_int, dp, dp]
    L.oracle_spline_build3d.argtypes = [dp, ip, dp]
    L.oracle_spline_eval3d.argtypes = [dp, ip, dp, dp, C.c_int, C.c_int, C.c_int, C.c_int, dp, dp]
    L.oracle_lens_deflect.argtypes = [C.c_double, C.c_double, dp, dp, C.POINTER(C.c_double)]
    L.oracle_lens_deflect.restype = C.c_int
    L.oracle_hemisphere.argtypes = [C.c_int, dp, dp, dp]
    L.oracle_cdf_sample.argtypes = [P, C.c_int, fp, ip]
    L.oracle_cdf_sample.restype = C.c_int
    _LIBS[name] = L
    return L


def philox(ctr, key):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.oracle_philox4x32_10(c, k, o)
    return list(o)


def uniform(seed, photon, draw, stream=0):
    return lib().oracle_uniform(seed, photon, draw, stream)


class OracleScene:
    """A scene for the oracle: descriptor + texture (+ sources / spline)."""

    def __init__(self, desc, texture=None, sources=None, spline=None, fast=False):
        self.L = lib(fast)
        self.desc = desc.copy()
        if texture is None:
            self.h = self.L.oracle_scene_create(C.byref(self.desc), None, 0, 0)
        else:
            tex = np.ascontiguousarray(texture, np.float32)
            self.h = self.L.oracle_scene_create(C.byref(self.desc), tex.ctypes.data, tex.shape[1], tex.shape[0])
        if sources is not None:
            self.set_sources(**sources)
        if spline is not None:
            self.set_spline(*spline)

    def set_sources(self, pos, amp, phase, coeff):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
        amp = np.ascontiguousarray(amp, np.float32)
        phase = np.ascontiguousarray(phase, np.float32)
        coeff = np.ascontiguousarray(coeff, np.float32)
        self.L.oracle_scene_set_sources(self.h, len(amp), pos, amp, phase, len(coeff), coeff)

    def set_spline(self, coeffs, dims, xmin, xmax):
        c = np.ascontiguousarray(coeffs, np.float32).reshape(-1)
        self.L.oracle_scene_set_spline(self.h, c, (C.c_int32 * 3)(*dims), (C.c_float * 3)(*xmin),
                                       (C.c_float * 3)(*xmax))

    def update(self, desc):
        self.desc = desc.copy()
        self.L.oracle_scene_update(self.h, C.byref(self.desc))

    def cdf(self):
        n = self.L.oracle_scene_get_cdf(self.h, np.zeros(1, np.float32), 1)
        out = np.zeros(n, np.float32)
        self.L.oracle_scene_get_cdf(self.h, out, n)
        return out

    def render(self, n, res, seed=1, offset=0, n_total=None, precision=32, threads=0, cutoff=False):
        """Render photons [offset, offset+n) of an n_total-photon render;
        returns the image [z][y][x] (float64) and the event counters (plus
        'busy_seconds': the CPU-seconds the threads spent tracing)."""
        X, Y, Z = res
        img = np.zeros(X * Y * Z, np.float64)
        st = Stats()
        busy = C.c_double()
        self.L.oracle_render(self.h, precision, n, n_total or n, seed, offset, int(cutoff), img, X, Y, Z,
                             C.byref(st), threads, C.byref(busy))
        out = st.as_dict()
        out["busy_seconds"] = busy.value
        return img.reshape(Z, Y, X), out

    def cdf_sample(self, u):
        """Texel indices that the source's texture CDF maps the uniforms u to."""
        u = np.ascontiguousarray(u, np.float32)
        out = np.zeros(len(u), np.int32)
        self.L.oracle_cdf_sample(self.h, len(u), u, out)
        return out

    def dead_path_cut(self):
        return self.L.oracle_dead_path_cut(self.h)

    def trace_rays(self, n, seed=1, offset=0, precision=64, threads=0):
        p = np.zeros(3 * n)
        d = np.zeros(3 * n)
        L = np.zeros(n)
        s = np.zeros(n, np.int32)
        self.L.oracle_trace_rays(self.h, precision, n, seed, offset, p, d, L, s, threads)
        return p.reshape(n, 3), d.reshape(n, 3), L, s

    def field(self, pts, scaling=1.0, tau=0.0, precision=64):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        n = np.zeros(len(pts))
        g = np.zeros((len(pts), 3))
        self.L.oracle_field_eval(self.h, precision, len(pts), pts.reshape(-1), scaling, tau, n, g.reshape(-1))
        return n, g

    def field_hessian(self, pts, scaling=1.0, tau=0.0):
        """Hessian of n at points (FP64): (n, 3, 3)."""
        pts = np.ascontiguousarray(np.asarray(pts, np.float64).reshape(-1, 3))
        out = np.zeros((len(pts), 3, 3))
        self.L.oracle_field_hessian(self.h, len(pts), pts.reshape(-1), scaling, tau, out.reshape(-1))
        return out

    def nee_connect(self, p1, v1, p2, scaling=1.0, tau=0.0):
        """The reference's computefdfNEE (FP64) per ray: (p - p2 as (n, 3),
        (dQ/dv1)^T as (n, 3, 3))."""
        a, b, c = (np.ascontiguousarray(np.asarray(x, np.float64).reshape(-1, 3)) for x in (p1, v1, p2))
        err = np.zeros((len(a), 3))
        jac = np.zeros((len(a), 3, 3))
        self.L.oracle_nee_connect(self.h, len(a), a.reshape(-1), b.reshape(-1), c.reshape(-1), scaling, tau,
                                  err.reshape(-1), jac.reshape(-1))
        return err, jac

    def er_steps(self, p, v, h, nsteps, scaling=1.0, tau=0.0, precision=64):
        p = np.array(p, np.float64)
        v = np.array(v, np.float64)
        self.L.oracle_er_steps(self.h, precision, p, v, h, nsteps, scaling, tau)
        return p, v

    def __del__(self):
        try:
            self.L.oracle_scene_destroy(self.h)
        except Exception:
            pass


# single path functions (double precision) --------------------------------
def lens_deflect(aperture, focal, pos, d, total=0.0):
    """Thin-lens deflection of a ray at the lens plane x = const (optical axis
    -x).  Returns (passed, pos, dir, total) with the total path length
    corrected for the lens."""
    p = np.array(pos, np.float64)
    v = np.array(d, np.float64)
    t = C.c_double(total)
    ok = lib().oracle_lens_deflect(aperture, focal, p, v, C.byref(t))
    return bool(ok), p, v, t.value


def hemisphere(u0, u1):
    """Next-event directions for the uniform pairs (u0, u1), shape (n, 3)."""
    u0 = np.ascontiguousarray(u0, np.float64)
    u1 = np.ascontiguousarray(u1, np.float64)
    out = np.zeros((len(u0), 3))
    lib().oracle_hemisphere(len(u0), u0, u1, out.reshape(-1))
    return out


# spline helpers (double precision, for the reference CSV tests) ----------
def spline_kernel(d, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    lib().oracle_spline_kernel(d, len(x), x, out)
    return out


def spline_build1d(data):
    data = np.ascontiguousarray(data, np.float64)
    out = np.zeros_like(data)
    lib().oracle_spline_build1d(data, len(data), out)
    return out


def spline_eval1d(coeff, xmin, xmax, d, t):
    coeff = np.ascontiguousarray(coeff, np.float64)
    t = np.ascontiguousarray(t, np.float64)
    out = np.zeros_like(t)
    lib().oracle_spline_eval1d(coeff, len(coeff), xmin, xmax, d, len(t), t, out)
    return out


def spline_build2d(data, N):
    data = np.ascontiguousarray(data, np.float64).reshape(-1)
    out = np.zeros_like(data)
    lib().oracle_spline_build2d(data, np.array(N, np.int32), out)
    return out


def spline_eval2d(coeff, N, xmin, xmax, dx, dy, pts):
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 2)
    out = np.zeros(len(pts))
    lib().oracle_spline_eval2d(np.ascontiguousarray(coeff, np.float64), np.array(N, np.int32),
                               np.array(xmin, np.float64), np.array(xmax, np.float64), dx, dy, len(pts),
                               pts.reshape(-1), out)
    return out


def spline_build3d(data, N):
    data = np.ascontiguousarray(data, np.float64).reshape(-1)
    out = np.zeros_like(data)
    lib().oracle_spline_build3d(data, np.array(N, np.int32), out)
    return out


def spline_eval3d(coeff, N, xmin, xmax, dx, dy, dz, pts):
    pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
    out = np.zeros(len(pts))
    lib().oracle_spline_eval3d(np.ascontiguousarray(coeff, np.float64), np.array(N, np.int32),
                               np.array(xmin, np.float64), np.array(xmax, np.float64), dx, dy, dz, len(pts),
                               pts.reshape(-1), out)
    return out
