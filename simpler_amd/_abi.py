"""ctypes mirror of include/simpler.h (the C ABI of the gfx950 engine).

Field order and types must match ``simpler_scene_desc`` / ``simpler_stats``
exactly; ``tests/test_abi.py`` checks the sizes against the compiled library.
"""
import ctypes as C

FIELD_BESSEL = 0
FIELD_SOURCES = 1
FIELD_SPLINE = 2

FLAG_DIRECT_LEAK = 1 << 0
FLAG_USE_FRESNEL = 1 << 1
FLAG_GENERAL_AXIS = 1 << 2
FLAG_NO_DEAD_CUT = 1 << 3
FLAGS_DEFAULT = FLAG_DIRECT_LEAK | FLAG_GENERAL_AXIS

SIMPLER_OK = 0

_F3 = C.c_float * 3
_F2 = C.c_float * 2


class SceneDesc(C.Structure):
    """``simpler_scene_desc``: scn::Scene + med::Medium + photon::Renderer
    parameters (reference scene.h:47-168, medium.h:24-38, photon.h:27-32)."""

    _fields_ = [
        ("block_l", _F3),
        ("block_r", _F3),
        ("ior", C.c_float),
        ("light_origin", _F3),
        ("light_dir", _F3),
        ("half_theta_limit", C.c_float),
        ("light_plane", _F2),
        ("Li", C.c_float),
        ("emitter_lens_origin", _F3),
        ("emitter_lens_aperture", C.c_float),
        ("emitter_lens_focal_length", C.c_float),
        ("emitter_lens_active", C.c_int32),
        ("view_origin", _F3),
        ("view_dir", _F3),
        ("view_horizontal", _F3),
        ("view_plane", _F2),
        ("pathlength_range", _F2),
        ("use_bounce_decomposition", C.c_int32),
        ("sensor_lens_origin", _F3),
        ("sensor_lens_aperture", C.c_float),
        ("sensor_lens_focal_length", C.c_float),
        ("sensor_lens_active", C.c_int32),
        ("field_kind", C.c_int32),
        ("n_o", C.c_float),
        ("f_u", C.c_float),
        ("speed_u", C.c_float),
        ("n_max", C.c_float),
        ("n_clip", C.c_float),
        ("phi_min", C.c_float),
        ("phi_max", C.c_float),
        ("mode", C.c_int32),
        ("axis_uz", _F3),
        ("axis_ux", _F3),
        ("p_u", _F3),
        ("er_stepsize", C.c_float),
        ("precision", C.c_int32),
        ("egap_end_loc_x", C.c_float),
        ("sgap_begin_loc_x", C.c_float),
        ("time_coupling", C.c_float),
        ("sigma_t", C.c_float),
        ("albedo", C.c_float),
        ("g", C.c_float),
        ("max_depth", C.c_int32),
        ("max_pathlength", C.c_float),
        ("use_direct", C.c_int32),
        ("use_angular_sampling", C.c_int32),
        ("flags", C.c_uint32),
    ]

    def copy(self):
        d = SceneDesc()
        C.memmove(C.byref(d), C.byref(self), C.sizeof(SceneDesc))
        return d

    def set(self, **kw):
        for k, v in kw.items():
            cur = getattr(self, k)
            if isinstance(cur, C.Array):
                for i, x in enumerate(v):
                    cur[i] = x
            else:
                setattr(self, k, v)
        return self

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if isinstance(v, C.Array) else v
        return out


class Stats(C.Structure):
    """``simpler_stats``: exact event counters of one render."""

    _fields_ = [
        ("photons", C.c_uint64),
        ("er_steps", C.c_uint64),
        ("scatters", C.c_uint64),
        ("nee_splats", C.c_uint64),
        ("direct_splats", C.c_uint64),
        ("face_events", C.c_uint64),
        ("segments", C.c_uint64),
        ("aborted", C.c_uint64),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class AdjointStats(C.Structure):
    """``simpler_adjoint_stats``: work counters of one adjoint call."""

    _fields_ = [
        ("photons", C.c_uint64),
        ("er_steps", C.c_uint64),
        ("reverse_steps", C.c_uint64),
        ("sub_paths_reversed", C.c_uint64),
        ("main_paths_reversed", C.c_uint64),
        ("dropped", C.c_uint64),
    ]

    def as_dict(self):
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


class BatchCfg(C.Structure):
    """``simpler_batch_cfg``: per-configuration overrides of a batched launch."""

    _fields_ = [
        ("n_max", C.c_float),
        ("n_clip", C.c_float),
        ("f_u", C.c_float),
        ("phi_min", C.c_float),
        ("phi_max", C.c_float),
        ("mode", C.c_int32),
        ("sigma_t", C.c_float),
        ("albedo", C.c_float),
        ("g", C.c_float),
        ("reserved", C.c_int32),
        ("n_photons", C.c_int64),
        ("seed", C.c_uint64),
    ]
