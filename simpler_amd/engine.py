"""Python binding of the C ABI in include/simpler.h (libsimpler_hip.so).

PyTorch is used only for device memory and the HIP stream; every photon is
traced by the gfx950 kernels in csrc/render.hip.  There is no CPU fallback:
constructing a Context without the built library or without a GPU raises.
"""
import ctypes as C
import os
import weakref

import numpy as np

from . import build as _build
from ._abi import AdjointStats, BatchCfg, SceneDesc, Stats

_LIB = None


class SimplerError(RuntimeError):
    pass


class SimplerWarning(RuntimeWarning):
    pass


def _warn_dropped(st):
    """The adjoint leaves out the contributions of photons whose segment log
    overflowed (more than 512 main-path segments, or more than 64 face
    reflections in one next-event / direct sub-path); their objective still
    counts, so the gradient is biased by their share."""
    if st.get("dropped", 0):
        import warnings
        warnings.warn(SimplerWarning(
            f"adjoint: {st['dropped']} splat contribution(s) left out of the gradient (segment log full: "
            f"more than 512 main-path segments or 64 sub-path reflections in one photon)"), stacklevel=3)


def load_library():
    """Load libsimpler_hip.so (built in-tree by simpler_amd.build)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    # SIMPLER_HIP_LIBRARY selects another build of the same ABI (e.g. a
    # different tuning of the kernels); default: the in-tree library
    path = os.environ.get("SIMPLER_HIP_LIBRARY") or _build.hip_library_path()
    if not os.path.exists(path):
        raise SimplerError(f"{path} is missing: build it with `python -m simpler_amd.build`")
    L = C.CDLL(path)
    P = C.c_void_p
    PP = C.POINTER(C.c_void_p)
    fptr = C.POINTER(C.c_float)
    L.simpler_abi_version.restype = C.c_int
    L.simpler_scene_desc_default.argtypes = [C.POINTER(SceneDesc)]
    L.simpler_ctx_create.argtypes = [C.c_int, PP]
    L.simpler_ctx_destroy.argtypes = [P]
    L.simpler_ctx_set_stream.argtypes = [P, P]
    L.simpler_ctx_get_stream.argtypes = [P]
    L.simpler_ctx_get_stream.restype = P
    L.simpler_ctx_own_stream.argtypes = [P]
    L.simpler_ctx_own_stream.restype = P
    L.simpler_ctx_synchronize.argtypes = [P]
    L.simpler_last_error.argtypes = [P]
    L.simpler_last_error.restype = C.c_char_p
    L.simpler_ctx_set_tuning.argtypes = [P, C.c_int, C.c_int, C.c_int]
    L.simpler_ctx_get_tuning.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.simpler_ctx_set_splat.argtypes = [P, C.c_int]
    L.simpler_ctx_get_splat.argtypes = [P, C.POINTER(C.c_int)]
    L.simpler_ctx_set_adjoint_schedule.argtypes = [P, C.c_int, C.c_int64]
    L.simpler_ctx_get_adjoint_schedule.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    L.simpler_ctx_set_next_event_skip.argtypes = [P, C.c_int]
    L.simpler_ctx_get_next_event_skip.argtypes = [P, C.POINTER(C.c_int)]
    L.simpler_scene_create.argtypes = [P, C.POINTER(SceneDesc), fptr, C.c_int, C.c_int, PP]
    L.simpler_scene_destroy.argtypes = [P]
    L.simpler_scene_update.argtypes = [P, C.POINTER(SceneDesc)]
    L.simpler_scene_get_desc.argtypes = [P, C.POINTER(SceneDesc)]
    L.simpler_scene_set_sources.argtypes = [P, C.c_int, fptr, fptr, fptr, C.c_int, fptr]
    L.simpler_scene_set_spline.argtypes = [P, fptr, C.POINTER(C.c_int32), fptr, fptr]
    L.simpler_scene_set_spline_samples.argtypes = [P, P, C.c_int, C.POINTER(C.c_int32), fptr, fptr]
    L.simpler_scene_get_spline.argtypes = [P, fptr]
    L.simpler_spline_prefilter.argtypes = [P, P, C.POINTER(C.c_int32)]
    L.simpler_render.argtypes = [P, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, P, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.POINTER(Stats)]
    L.simpler_render_host.argtypes = [P, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, fptr, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.POINTER(Stats)]
    L.simpler_render_host_noalloc.argtypes = [P, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, fptr, C.c_int,
                                              C.c_int, C.c_int, C.c_int, C.POINTER(Stats)]
    L.simpler_trace_rays.argtypes = [P, C.c_int64, C.c_uint64, C.c_uint64, P, P, P, P]
    L.simpler_jvp_rays.argtypes = [P, C.c_int64, C.c_uint64, C.c_uint64, P, P, P, P, P, P, P, P, P]
    L.simpler_nee_connect.argtypes = [P, C.c_int64, P, P, P, C.c_float, C.c_float, P, P]
    L.simpler_field_eval.argtypes = [P, C.c_int, P, C.c_float, C.c_float, P, P]
    L.simpler_read_stats.argtypes = [P, C.POINTER(Stats)]
    L.simpler_batch_create.argtypes = [P, C.c_int, C.POINTER(BatchCfg), C.c_int, C.c_int, C.c_int, PP]
    L.simpler_batch_render.argtypes = [P, P, C.c_int, C.POINTER(Stats)]
    L.simpler_batch_graph_capture.argtypes = [P, P]
    L.simpler_batch_graph_launch.argtypes = [P]
    L.simpler_batch_destroy.argtypes = [P]
    L.simpler_autotune.argtypes = [P, C.c_int64, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                   C.POINTER(C.c_int), C.POINTER(C.c_float)]
    L.simpler_adjoint_param_count.argtypes = [P, C.POINTER(C.c_int64)]
    L.simpler_batch_create_sources.argtypes = [P, C.c_int, fptr, fptr, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                                               C.c_int, C.c_int, C.c_int, PP]
    L.simpler_batch_adjoint.argtypes = [P, P, P, C.c_int, C.POINTER(C.c_double), C.POINTER(AdjointStats)]
    L.simpler_adjoint_render.argtypes = [P, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, P, C.c_int, C.c_int,
                                         C.c_int, P, C.c_int, C.POINTER(C.c_double), C.POINTER(AdjointStats)]
    L.simpler_adjoint_rays.argtypes = [P, C.c_int64, C.c_uint64, C.c_uint64, P, P, P, P, C.c_int,
                                       C.POINTER(C.c_double), C.POINTER(AdjointStats)]
    for name in ("simpler_ctx_create", "simpler_ctx_destroy", "simpler_ctx_set_stream", "simpler_ctx_synchronize",
                 "simpler_ctx_set_tuning", "simpler_ctx_get_tuning", "simpler_ctx_set_splat", "simpler_ctx_get_splat",
                 "simpler_ctx_set_adjoint_schedule", "simpler_ctx_get_adjoint_schedule",
                 "simpler_ctx_set_next_event_skip", "simpler_ctx_get_next_event_skip", "simpler_scene_create", "simpler_scene_destroy",
                 "simpler_scene_update", "simpler_scene_get_desc", "simpler_scene_set_sources",
                 "simpler_scene_set_spline", "simpler_render", "simpler_render_host", "simpler_trace_rays", "simpler_jvp_rays", "simpler_nee_connect",
                 "simpler_field_eval", "simpler_read_stats", "simpler_render_host_noalloc", "simpler_batch_create", "simpler_batch_render",
                 "simpler_batch_graph_capture", "simpler_batch_graph_launch", "simpler_batch_destroy",
                 "simpler_adjoint_param_count", "simpler_adjoint_render", "simpler_adjoint_rays",
                 "simpler_autotune", "simpler_batch_create_sources", "simpler_batch_adjoint",
                 "simpler_scene_set_spline_samples", "simpler_scene_get_spline", "simpler_spline_prefilter"):
        getattr(L, name).restype = C.c_int
    _LIB = L
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Context:
    """One GPU + one HIP stream.  By default the context follows torch's
    current stream: every call first moves the context to the stream that is
    current at that moment (an ordered switch, see simpler_ctx_set_stream),
    so the tensors it returns are ready in the caller's stream order, also
    inside `with torch.cuda.stream(s)`.  use_torch_stream=False keeps the
    context's own stream; an explicit set_stream pins a stream."""

    def __init__(self, device=0, use_torch_stream=True):
        import torch
        if not torch.cuda.is_available():
            raise SimplerError("no GPU visible: the engine has no CPU fallback")
        self.L = load_library()
        self.device = device
        self._scenes = weakref.WeakSet()
        self._batches = weakref.WeakSet()
        torch.cuda.set_device(device)
        h = C.c_void_p()
        rc = self.L.simpler_ctx_create(device, C.byref(h))
        if rc != 0:
            raise SimplerError(f"simpler_ctx_create failed ({rc})")
        self.h = h
        self.follow_torch = bool(use_torch_stream)
        self._bound = None
        self._bind()

    def _bind(self):
        """Move the context to torch's current stream (when following it)."""
        if not self.follow_torch:
            return
        import torch
        s = torch.cuda.current_stream(self.device).cuda_stream
        if s != self._bound:
            self.check(self.L.simpler_ctx_set_stream(self.h, C.c_void_p(s)))
            self._bound = s

    def check(self, rc):
        if rc != 0:
            msg = self.L.simpler_last_error(self.h).decode()
            raise SimplerError(f"simpler call failed ({rc}): {msg}")

    def spline_prefilter(self, grid, dims):
        """In-place B-spline prefilter of a contiguous float32 CUDA tensor of
        node values (x fastest) on this context's stream."""
        if grid.dtype.itemsize != 4 or not grid.is_contiguous() or not grid.is_cuda or grid.numel() != int(
                np.prod(dims)):
            raise SimplerError("grid must be a contiguous float32 CUDA tensor matching dims")
        self._bind()
        self.check(self.L.simpler_spline_prefilter(self.h, C.c_void_p(grid.data_ptr()), (C.c_int32 * 3)(*dims)))

    def set_stream(self, stream_ptr):
        """Pin the context to a HIP stream (it stops following torch's)."""
        self.follow_torch = False
        self.check(self.L.simpler_ctx_set_stream(self.h, C.c_void_p(stream_ptr)))
        self._bound = stream_ptr

    def synchronize(self):
        self.check(self.L.simpler_ctx_synchronize(self.h))

    def stream(self):
        """The HIP stream the context runs on, as a torch stream (for events)."""
        import torch
        ptr = self.L.simpler_ctx_get_stream(self.h) or 0
        cur = torch.cuda.current_stream(self.device)
        if ptr == cur.cuda_stream:
            return cur
        return torch.cuda.ExternalStream(ptr, device=f"cuda:{self.device}")

    def set_tuning(self, block_size=0, steps_per_chunk=0, blocks_per_cu=0, splat=None):
        self.check(self.L.simpler_ctx_set_tuning(self.h, block_size, steps_per_chunk, blocks_per_cu))
        if splat is not None:
            self.check(self.L.simpler_ctx_set_splat(self.h, int(splat)))

    def get_tuning(self):
        b, s, c, sp = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.check(self.L.simpler_ctx_get_tuning(self.h, C.byref(b), C.byref(s), C.byref(c)))
        self.check(self.L.simpler_ctx_get_splat(self.h, C.byref(sp)))
        return dict(block_size=b.value, steps_per_chunk=s.value, blocks_per_cu=c.value, splat=sp.value)

    def set_adjoint_schedule(self, schedule, queue_capacity=0):
        """Image adjoint: 0 runs the backward runs inline in the re-trace, 1
        defers them to full-wave backward kernels; queue_capacity sizes the
        deferred queues of one chunk of photons (0 keeps it)."""
        self.check(self.L.simpler_ctx_set_adjoint_schedule(self.h, int(schedule), int(queue_capacity)))

    def get_adjoint_schedule(self):
        s, c = C.c_int(), C.c_int64()
        self.check(self.L.simpler_ctx_get_adjoint_schedule(self.h, C.byref(s), C.byref(c)))
        return dict(schedule=s.value, queue_capacity=c.value)

    def set_next_event_skip(self, enable):
        """Forward renders of spline fields: leave untraced the next-event
        sub-paths that provably add nothing to the image (default on; the
        image and splat counts are those of the full trace, er_steps and
        face_events drop)."""
        self.check(self.L.simpler_ctx_set_next_event_skip(self.h, 1 if enable else 0))

    def get_next_event_skip(self):
        e = C.c_int()
        self.check(self.L.simpler_ctx_get_next_event_skip(self.h, C.byref(e)))
        return bool(e.value)

    def read_stats(self):
        self._bind()
        st = Stats()
        self.check(self.L.simpler_read_stats(self.h, C.byref(st)))
        return st.as_dict()

    def close(self):
        """Destroys the context after its batches and scenes."""
        if getattr(self, "h", None):
            for b in list(self._batches):
                b.close()
            for sc in list(self._scenes):
                sc.close()
            self.check(self.L.simpler_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A scene uploaded to the device (descriptor + texture CDF, optional
    point sources or spline coefficients)."""

    def __init__(self, ctx, desc, texture=None, sources=None, spline=None):
        self.ctx = ctx
        self.L = ctx.L
        self.desc = desc.copy()
        h = C.c_void_p()
        if texture is None:
            rc = self.L.simpler_scene_create(ctx.h, C.byref(self.desc), None, 0, 0, C.byref(h))
        else:
            tex = np.ascontiguousarray(texture, np.float32)
            rc = self.L.simpler_scene_create(ctx.h, C.byref(self.desc), _fp(tex), tex.shape[1], tex.shape[0],
                                             C.byref(h))
        ctx.check(rc)
        self.h = h
        self._batches = weakref.WeakSet()
        ctx._scenes.add(self)
        if sources is not None:
            self.set_sources(**sources)
        if spline is not None:
            self.set_spline(*spline)

    def set_sources(self, pos, amp, phase, coeff):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
        amp = np.ascontiguousarray(amp, np.float32)
        phase = np.ascontiguousarray(phase, np.float32)
        coeff = np.ascontiguousarray(coeff, np.float32)
        self.ctx._bind()
        self.ctx.check(self.L.simpler_scene_set_sources(self.h, len(amp), _fp(pos), _fp(amp), _fp(phase), len(coeff),
                                                        _fp(coeff)))
        self._sources = dict(pos=pos, amp=amp, phase=phase, coeff=coeff)

    def set_spline(self, coeffs, dims, xmin, xmax):
        c = np.ascontiguousarray(coeffs, np.float32).reshape(-1)
        self.ctx._bind()
        self.ctx.check(self.L.simpler_scene_set_spline(self.h, _fp(c), (C.c_int32 * 3)(*dims),
                                                       (C.c_float * 3)(*xmin), (C.c_float * 3)(*xmax)))
        self._spline = (tuple(dims), tuple(xmin), tuple(xmax), c)

    def set_spline_samples(self, samples, dims, xmin, xmax):
        """Spline field from grid node values (x fastest): a numpy array or a
        float32 CUDA tensor; the B-spline prefilter runs on the GPU."""
        import torch
        if isinstance(samples, torch.Tensor):
            if samples.dtype != torch.float32 or not samples.is_contiguous() or not samples.is_cuda:
                raise SimplerError("samples must be a contiguous float32 CUDA tensor")
            ptr, on_dev, keep = samples.data_ptr(), 1, samples
        else:
            keep = np.ascontiguousarray(samples, np.float32).reshape(-1)
            ptr, on_dev = keep.ctypes.data, 0
        count = keep.numel() if on_dev else keep.size
        if count != int(np.prod(dims)):
            raise SimplerError("samples do not match dims")
        self.ctx._bind()
        self.ctx.check(self.L.simpler_scene_set_spline_samples(self.h, C.c_void_p(ptr), on_dev,
                                                               (C.c_int32 * 3)(*dims), (C.c_float * 3)(*xmin),
                                                               (C.c_float * 3)(*xmax)))
        self._spline = (tuple(dims), tuple(xmin), tuple(xmax), None)
        self._spline = (tuple(dims), tuple(xmin), tuple(xmax), self.spline_coeffs())

    def spline_coeffs(self):
        """The scene's B-spline coefficients (float32, x fastest)."""
        dims = self._spline[0] if hasattr(self, "_spline") else None
        if dims is None:
            raise SimplerError("the scene has no spline field")
        out = np.zeros(int(np.prod(dims)), np.float32)
        self.ctx._bind()
        self.ctx.check(self.L.simpler_scene_get_spline(self.h, _fp(out)))
        return out

    def params(self):
        """The differentiable field parameters as a float64 vector: spline
        coefficients; or source amplitudes, phases and the wavenumber
        k = 2 pi f_u / speed_u; or, for a Bessel transducer, (n_max, k,
        phi_min) -- the amplitude, the wavenumber and the start of the duty
        window (its length stays fixed, so this is the transducer phase)."""
        k = 2.0 * np.pi * float(self.desc.f_u) / float(self.desc.speed_u)
        if self.desc.field_kind == 2:
            return self._spline[3].astype(np.float64)
        if self.desc.field_kind == 1:
            s = self._sources
            return np.concatenate([s["amp"], s["phase"], [k]]).astype(np.float64)
        return np.array([float(self.desc.n_max), k, float(self.desc.phi_min)], np.float64)

    def set_params(self, vec):
        """Inverse of params()."""
        vec = np.asarray(vec, np.float64).reshape(-1)
        if self.desc.field_kind == 2:
            dims, lo, hi, _ = self._spline
            self.set_spline(vec, dims, lo, hi)
        elif self.desc.field_kind == 1:
            s = self._sources
            N = len(s["amp"])
            if vec.size != 2 * N + 1:
                raise SimplerError(f"expected {2 * N + 1} source parameters")
            self.set_sources(s["pos"], vec[:N], vec[N:2 * N], s["coeff"])
            d = self.desc.copy()
            d.set(f_u=float(vec[2 * N]) * float(d.speed_u) / (2.0 * np.pi))
            self.update(d)
        else:
            if vec.size != 3:
                raise SimplerError("expected 3 Bessel parameters (n_max, k, phi_min)")
            d = self.desc.copy()
            ratio = float(d.n_clip) / float(d.n_max)
            window = float(d.phi_max) - float(d.phi_min)
            d.set(n_max=float(vec[0]), n_clip=float(vec[0]) * ratio,
                  f_u=float(vec[1]) * float(d.speed_u) / (2.0 * np.pi),
                  phi_min=float(vec[2]), phi_max=float(vec[2]) + window)
            self.update(d)

    def update(self, desc):
        self.desc = desc.copy()
        self.ctx.check(self.L.simpler_scene_update(self.h, C.byref(self.desc)))

    def render(self, n, res, seed=1, offset=0, n_total=None, image=None, accumulate=False, stats=True):
        """Render photons [offset, offset+n) of an n_total-photon render into
        a float32 CUDA tensor of shape (Z, Y, X); returns (image, stats)."""
        import torch
        X, Y, Z = res
        if image is None:
            image = torch.empty((Z, Y, X), dtype=torch.float32, device=f"cuda:{self.ctx.device}")
        if image.dtype != torch.float32 or not image.is_contiguous() or image.numel() != X * Y * Z:
            raise SimplerError("image must be a contiguous float32 tensor of X*Y*Z elements")
        st = Stats()
        self.ctx._bind()
        self.ctx.check(self.L.simpler_render(self.h, n, n_total or n, seed, offset, C.c_void_p(image.data_ptr()), X,
                                             Y, Z, int(accumulate), C.byref(st) if stats else None))
        return image, (st.as_dict() if stats else None)

    def render_host(self, n, res, seed=1, offset=0, n_total=None):
        X, Y, Z = res
        img = np.zeros(X * Y * Z, np.float32)
        st = Stats()
        self.ctx._bind()
        self.ctx.check(self.L.simpler_render_host(self.h, n, n_total or n, seed, offset, _fp(img), X, Y, Z, 0,
                                                  C.byref(st)))
        return img.reshape(Z, Y, X), st.as_dict()

    def render_host_noalloc(self, n, res, seed=1, offset=0, n_total=None, image=None, accumulate=False):
        """render_host through the entry point that allocates no device
        memory (its staging lives in the library's fixed device array; at
        most 8,388,608 pixels).  image: optional float32 numpy array (Z, Y, X)
        to fill, or to add to with accumulate."""
        X, Y, Z = res
        img = np.zeros(X * Y * Z, np.float32) if image is None else image
        if img.dtype != np.float32 or not img.flags["C_CONTIGUOUS"] or img.size != X * Y * Z:
            raise SimplerError("image must be a contiguous float32 array of X*Y*Z elements")
        st = Stats()
        self.ctx._bind()
        self.ctx.check(self.L.simpler_render_host_noalloc(self.h, n, n_total or n, seed, offset, _fp(img), X, Y, Z,
                                                          int(accumulate), C.byref(st)))
        return img.reshape(Z, Y, X), st.as_dict()

    def trace_rays(self, n, seed=1, offset=0):
        """Per-ray direct trace: (exit position, unit exit direction, optical
        length, status) as CUDA tensors."""
        import torch
        dev = f"cuda:{self.ctx.device}"
        p = torch.empty((n, 3), dtype=torch.float32, device=dev)
        d = torch.empty((n, 3), dtype=torch.float32, device=dev)
        L = torch.empty((n,), dtype=torch.float32, device=dev)
        s = torch.empty((n,), dtype=torch.int32, device=dev)
        self.ctx._bind()
        self.ctx.check(self.L.simpler_trace_rays(self.h, n, seed, offset, C.c_void_p(p.data_ptr()),
                                                 C.c_void_p(d.data_ptr()), C.c_void_p(L.data_ptr()),
                                                 C.c_void_p(s.data_ptr())))
        return p, d, L, s

    def jvp_rays(self, n, theta_dot=None, v0_dot=None, seed=1, offset=0):
        """Per-ray direct trace with tangents (forward mode): returns
        (exit position, unit direction, optical length, status, and their
        tangents) along theta_dot (a CUDA float32 tensor of param_count()
        values, or None) and the per-ray initial-velocity directions v0_dot
        (CUDA float32 (n, 3), or None)."""
        import torch
        dev = f"cuda:{self.ctx.device}"
        out = [torch.empty(sh, dtype=dt, device=dev) for sh, dt in
               (((n, 3), torch.float32), ((n, 3), torch.float32), ((n,), torch.float32), ((n,), torch.int32),
                ((n, 3), torch.float32), ((n, 3), torch.float32), ((n,), torch.float32))]
        td = None if theta_dot is None else theta_dot.contiguous()
        if td is not None and (td.dtype != torch.float32 or td.numel() != self.param_count()):
            raise SimplerError(f"theta_dot must be {self.param_count()} float32 values")
        vd = None if v0_dot is None else v0_dot.contiguous()
        if vd is not None and (vd.dtype != torch.float32 or vd.numel() != 3 * n):
            raise SimplerError("v0_dot must be (n, 3) float32")
        ptr = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
        self.ctx._bind()
        self.ctx.check(self.L.simpler_jvp_rays(self.h, n, seed, offset, ptr(td), ptr(vd), *[ptr(t) for t in out]))
        return tuple(out)

    def nee_connect(self, p1, v1, p2, scaling=1.0, tau=0.0):
        """Deterministic next-event connection (the reference's
        computefdfNEE): march from p1 with velocity v1 (scaled to n(p1)) to
        the plane x = p2.x; returns (p - p2 as (n, 3), (dQ/dv1)^T as
        (n, 3, 3)).  p1, v1, p2: CUDA float32 tensors (n, 3)."""
        import torch
        a, b, c = (t.contiguous() for t in (p1, v1, p2))
        n = a.shape[0]
        for t in (a, b, c):
            if t.dtype != torch.float32 or tuple(t.shape) != (n, 3) or not t.is_cuda:
                raise SimplerError("p1, v1, p2 must be CUDA float32 tensors of shape (n, 3)")
        err = torch.empty((n, 3), dtype=torch.float32, device=a.device)
        jac = torch.empty((n, 3, 3), dtype=torch.float32, device=a.device)
        ptr = lambda t: C.c_void_p(t.data_ptr())
        self.ctx._bind()
        self.ctx.check(self.L.simpler_nee_connect(self.h, n, ptr(a), ptr(b), ptr(c), scaling, tau, ptr(err), ptr(jac)))
        return err, jac

    def autotune(self, n, res):
        """Pick block size, steps per chunk and blocks per CU for this scene
        by timing an n-photon render of each candidate; the choice stays on
        the context.  Returns it with the best time (ms)."""
        b, s, c, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        X, Y, Z = res
        self.ctx._bind()
        self.ctx.check(self.L.simpler_autotune(self.h, n, X, Y, Z, C.byref(b), C.byref(s), C.byref(c), C.byref(ms)))
        return dict(block_size=b.value, steps_per_chunk=s.value, blocks_per_cu=c.value,
                    splat=self.ctx.get_tuning()["splat"], ms=ms.value)

    def param_count(self):
        """Number of differentiable field parameters (spline coefficients, or
        source amplitudes + phases + wavenumber)."""
        n = C.c_int64()
        self.ctx.check(self.L.simpler_adjoint_param_count(self.h, C.byref(n)))
        return n.value

    def _grad_buffer(self, grad):
        import torch
        n = self.param_count()
        if grad is None:
            return torch.empty((n,), dtype=torch.float32, device=f"cuda:{self.ctx.device}"), False
        if grad.dtype != torch.float32 or not grad.is_contiguous() or grad.numel() != n:
            raise SimplerError(f"gradient buffer must be a contiguous float32 tensor of {n} elements")
        return grad, True

    def adjoint_render(self, dF_dimage, n, seed=1, offset=0, n_total=None, grad=None):
        """Path-replay adjoint of F = sum(dF_dimage * render(n, ...)).
        dF_dimage: CUDA float32 tensor (Z, Y, X).  Returns (gradient tensor,
        F, stats); with grad given, the gradient is added to it."""
        G = dF_dimage.contiguous()
        Z, Y, X = G.shape
        g, acc = self._grad_buffer(grad)
        F = C.c_double()
        st = AdjointStats()
        self.ctx._bind()
        self.ctx.check(self.L.simpler_adjoint_render(self.h, n, n_total or n, seed, offset, C.c_void_p(G.data_ptr()),
                                                     X, Y, Z, C.c_void_p(g.data_ptr()), int(acc), C.byref(F),
                                                     C.byref(st)))
        st = st.as_dict()
        _warn_dropped(st)
        return g, F.value, st

    def adjoint_rays(self, g_pos, g_dir, g_len, seed=1, offset=0, grad=None):
        """Adjoint of F = sum g_pos . exit_pos + g_dir . exit_dir + g_len * len
        over the rays of trace_rays(n, seed, offset), n = len(g_len)."""
        gp, gd, gl = g_pos.contiguous(), g_dir.contiguous(), g_len.contiguous()
        g, acc = self._grad_buffer(grad)
        F = C.c_double()
        st = AdjointStats()
        self.ctx._bind()
        self.ctx.check(self.L.simpler_adjoint_rays(self.h, gl.numel(), seed, offset, C.c_void_p(gp.data_ptr()),
                                                   C.c_void_p(gd.data_ptr()), C.c_void_p(gl.data_ptr()),
                                                   C.c_void_p(g.data_ptr()), int(acc), C.byref(F), C.byref(st)))
        st = st.as_dict()
        _warn_dropped(st)
        return g, F.value, st

    def field(self, pts, scaling=1.0, tau=0.0):
        """n and grad n at points (CUDA tensor (N, 3) float32)."""
        import torch
        pts = pts.contiguous().to(torch.float32)
        n = torch.empty((pts.shape[0],), dtype=torch.float32, device=pts.device)
        g = torch.empty((pts.shape[0], 3), dtype=torch.float32, device=pts.device)
        self.ctx._bind()
        self.ctx.check(self.L.simpler_field_eval(self.h, pts.shape[0], C.c_void_p(pts.data_ptr()), scaling, tau,
                                                 C.c_void_p(n.data_ptr()), C.c_void_p(g.data_ptr())))
        return n, g

    def close(self):
        """Destroys the scene after the batches built on it."""
        if getattr(self, "h", None):
            for b in list(self._batches):
                b.close()
            self.ctx.check(self.L.simpler_scene_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many transducer configurations (list of dicts with the BatchCfg
    fields) over one base Bessel scene, rendered by one persistent launch
    into a (n_cfg, Z, Y, X) float32 CUDA tensor."""

    def __init__(self, scene, configs, res):
        import torch
        self.scene = scene
        self.ctx = scene.ctx
        self.L = scene.L
        self.res = res
        arr = (BatchCfg * len(configs))()
        for i, c in enumerate(configs):
            for k, v in c.items():
                setattr(arr[i], k, v)
        self.n_cfg = len(configs)
        self.photons = sum(int(c["n_photons"]) for c in configs)
        h = C.c_void_p()
        X, Y, Z = res
        self.ctx.check(self.L.simpler_batch_create(scene.h, len(configs), arr, X, Y, Z, C.byref(h)))
        self.h = h
        scene._batches.add(self)
        self.ctx._batches.add(self)
        self.images = torch.zeros((len(configs), Z, Y, X), dtype=torch.float32, device=f"cuda:{self.ctx.device}")
        self.captured = False

    @classmethod
    def patterns(cls, scene, amps, phases, n_photons, seeds, res):
        """Point-source pattern batch: per-pattern amplitudes and phases
        (arrays (n_cfg, n_src)), photon budgets and seeds over one
        point-source base scene."""
        import torch
        self = cls.__new__(cls)
        self.scene, self.ctx, self.L, self.res = scene, scene.ctx, scene.L, res
        amps = np.ascontiguousarray(amps, np.float32)
        phases = np.ascontiguousarray(phases, np.float32)
        n_cfg = amps.shape[0]
        npho = (C.c_int64 * n_cfg)(*[int(x) for x in n_photons])
        sds = (C.c_uint64 * n_cfg)(*[int(x) for x in seeds])
        self.n_cfg = n_cfg
        self.photons = int(sum(int(x) for x in n_photons))
        self.n_src = amps.shape[1]
        h = C.c_void_p()
        X, Y, Z = res
        self.ctx.check(self.L.simpler_batch_create_sources(scene.h, n_cfg, _fp(amps), _fp(phases), npho, sds, X, Y,
                                                           Z, C.byref(h)))
        self.h = h
        scene._batches.add(self)
        self.ctx._batches.add(self)
        self.images = torch.zeros((n_cfg, Z, Y, X), dtype=torch.float32, device=f"cuda:{self.ctx.device}")
        self.captured = False
        return self

    def adjoint(self, dF_dimages, grads=None):
        """Adjoint of every pattern: F_c = sum(dF_dimages[c] * image_c).
        Returns (gradients (n_cfg, 2 n_src + 1), F (n_cfg,), stats)."""
        import torch
        G = dF_dimages.contiguous()
        npar = 2 * self.n_src + 1
        acc = grads is not None
        if grads is None:
            grads = torch.empty((self.n_cfg, npar), dtype=torch.float32, device=G.device)
        F = (C.c_double * self.n_cfg)()
        st = AdjointStats()
        self.ctx._bind()
        self.ctx.check(self.L.simpler_batch_adjoint(self.h, C.c_void_p(G.data_ptr()), C.c_void_p(grads.data_ptr()),
                                                    int(acc), F, C.byref(st)))
        st = st.as_dict()
        _warn_dropped(st)
        return grads, np.array(F[:]), st

    def render(self, accumulate=False, stats=True):
        st = Stats()
        self.ctx._bind()
        self.ctx.check(self.L.simpler_batch_render(self.h, C.c_void_p(self.images.data_ptr()), int(accumulate),
                                                   C.byref(st) if stats else None))
        return self.images, (st.as_dict() if stats else None)

    def capture(self):
        """Capture zero + render of self.images into a HIP graph."""
        self.ctx._bind()
        self.ctx.check(self.L.simpler_batch_graph_capture(self.h, C.c_void_p(self.images.data_ptr())))
        self.captured = True

    def replay(self):
        self.ctx._bind()
        self.ctx.check(self.L.simpler_batch_graph_launch(self.h))
        return self.images

    def close(self):
        if getattr(self, "h", None):
            self.L.simpler_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
