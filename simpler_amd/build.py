"""Build the gfx950 engine (libsimpler_hip.so) and the host-side modules.

    python -m simpler_amd.build          # everything
    python -m simpler_amd.build hip      # only the HIP library

Outputs go into simpler_amd/lib/ (in-tree, so they travel with the source
tree).  hipcc cross-compiles for gfx950 on a machine without a GPU.
"""
import os
import shutil
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
INCLUDE = os.path.join(ROOT, "include")
ARCH = os.environ.get("SIMPLER_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fno-gpu-rdc", "-Wno-unused-result",
]

HIP_SOURCES = ["render.hip"]
HIP_DEPS = ["render.hip", "fields.h", "device_math.h", "adjoint.h", "adjoint_kernel.h", "prefilter.h"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _stale(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps)


def hip_library_path():
    return os.path.join(LIB, "libsimpler_hip.so")


def build_hip(force=True):
    os.makedirs(LIB, exist_ok=True)
    out = hip_library_path()
    deps = [os.path.join(CSRC, f) for f in HIP_DEPS] + [os.path.join(INCLUDE, "simpler.h")]
    if not force and not _stale(out, deps):
        return out
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-I", INCLUDE, "-o", tmp] + [os.path.join(CSRC, s) for s in HIP_SOURCES])
    os.replace(tmp, out)
    return out


def build_all(force=True):
    build_hip(force)
    try:
        from . import build_host
    except ImportError:
        return
    build_host.build(force)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what == "hip":
        build_hip()
    else:
        build_all()
