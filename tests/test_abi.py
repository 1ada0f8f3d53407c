"""The C-ABI library: it loads without a GPU, exports every entry point that
include/simpler.h declares, and the ctypes mirror matches the C layout."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from simpler_amd import build
from simpler_amd._abi import AdjointStats, BatchCfg, SceneDesc, Stats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "simpler.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(simpler_[a-z_0-9]+)\s*\(", src)))


def lib():
    path = build.hip_library_path()
    if not os.path.exists(path):
        build.build_hip()
    return C.CDLL(path)


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("simpler_ctx_create", "simpler_scene_create", "simpler_render", "simpler_trace_rays",
                     "simpler_render_host", "simpler_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    L = lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def _c_layout():
    fields_desc = [f for f, _ in SceneDesc._fields_]
    fields_stats = [f for f, _ in Stats._fields_]
    fields_batch = [f for f, _ in BatchCfg._fields_]
    fields_adj = [f for f, _ in AdjointStats._fields_]
    prog = ['#include <stdio.h>', '#include <stddef.h>', '#include "simpler.h"', 'int main(void){',
            'printf("%zu %zu\\n", sizeof(simpler_scene_desc), sizeof(simpler_stats));']
    for f in fields_desc:
        prog.append(f'printf("%zu\\n", offsetof(simpler_scene_desc, {f}));')
    for f in fields_stats:
        prog.append(f'printf("%zu\\n", offsetof(simpler_stats, {f}));')
    prog.append('printf("%zu\\n", sizeof(simpler_batch_cfg));')
    for f in fields_batch:
        prog.append(f'printf("%zu\\n", offsetof(simpler_batch_cfg, {f}));')
    prog.append('printf("%zu\\n", sizeof(simpler_adjoint_stats));')
    for f in fields_adj:
        prog.append(f'printf("%zu\\n", offsetof(simpler_adjoint_stats, {f}));')
    prog.append("return 0;}")
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "p.c")
        exe = os.path.join(td, "p")
        open(c, "w").write("\n".join(prog))
        subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    vals = [int(x) for x in out]
    nd, ns, nb = len(fields_desc), len(fields_stats), len(fields_batch)
    o = 2 + nd + ns
    return (vals[0], vals[1], vals[2:2 + nd], vals[2 + nd:o], vals[o], vals[o + 1:o + 1 + nb],
            vals[o + 1 + nb], vals[o + 2 + nb:])


def test_ctypes_mirror_matches_c_layout():
    sd, ss, off_d, off_s, sb, off_b, sa, off_a = _c_layout()
    assert C.sizeof(SceneDesc) == sd
    assert C.sizeof(Stats) == ss
    assert C.sizeof(BatchCfg) == sb
    assert [getattr(SceneDesc, f).offset for f, _ in SceneDesc._fields_] == off_d
    assert [getattr(Stats, f).offset for f, _ in Stats._fields_] == off_s
    assert [getattr(BatchCfg, f).offset for f, _ in BatchCfg._fields_] == off_b
    assert C.sizeof(AdjointStats) == sa
    assert [getattr(AdjointStats, f).offset for f, _ in AdjointStats._fields_] == off_a


def test_default_descriptor_and_version_without_gpu():
    L = lib()
    assert L.simpler_abi_version() == 1
    d = SceneDesc()
    L.simpler_scene_desc_default(C.byref(d))
    assert d.n_o == pytest.approx(1.3333) and d.precision == 4 and d.flags == 5
    assert list(d.axis_uz) == [-1.0, 0.0, 0.0]


def test_entry_points_reject_null_without_gpu():
    L = lib()
    L.simpler_render.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64, C.c_void_p, C.c_int,
                                 C.c_int, C.c_int, C.c_int, C.c_void_p]
    assert L.simpler_render(None, 1, 1, 0, 0, None, 1, 1, 1, 0, None) != 0
    L.simpler_ctx_create.argtypes = [C.c_int, C.c_void_p]
    assert L.simpler_ctx_create(0, None) != 0
    L.simpler_last_error.restype = C.c_char_p
    assert L.simpler_last_error(None) == b"null context"
