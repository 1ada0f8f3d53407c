"""Build the host-side C++ layer: libsimpler_host.so (the reference's
scn / photon / med / pfunc / image / tvec API on the C ABI) and the six
pybind11 modules with the reference's names (tvector_pybind, scene_pybind,
photon_pybind, medium_pybind, phase_pybind, image_pybind), all into
simpler_amd/lib/, and the two command-line front-ends (simplER,
renderer3d_sample) into simpler_amd/bin/.  Plain g++: no HIP headers are needed above the C ABI."""
import os
import subprocess
import sysconfig

from . import build as hb

HOST = os.path.join(hb.CSRC, "host")
HOST_INC = os.path.join(HOST, "include")
MODULES = ["tvector", "phase", "medium", "image", "scene", "photon"]
CXX = os.environ.get("CXX", "g++")


def _py_includes():
    import pybind11
    return ["-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"]]


def host_library_path():
    return os.path.join(hb.LIB, "libsimpler_host.so")


def module_path(name):
    return os.path.join(hb.LIB, f"{name}_pybind{sysconfig.get_config_var('EXT_SUFFIX')}")


FRONTEND = os.path.join(HOST, "frontend")
FRONTENDS = {"simplER": "simplER_main.cpp", "renderer3d_sample": "renderer3d_sample_main.cpp"}
BIN = os.path.join(os.path.dirname(hb.LIB), "bin")


def frontend_path(name):
    return os.path.join(BIN, name)


def build(force=True):
    os.makedirs(hb.LIB, exist_ok=True)
    hip = hb.hip_library_path()
    inc = ["-I", hb.INCLUDE, "-I", HOST_INC]
    rpath = "-Wl,-rpath,$ORIGIN"
    common = ["-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wno-unused-parameter"]
    lib = host_library_path()
    src = os.path.join(HOST, "compat.cpp")
    hdr = os.path.join(HOST_INC, "simpler_compat.h")
    if force or hb._stale(lib, [src, hdr, hip]):
        hb._run([CXX] + common + inc + [src, "-o", lib + ".tmp", "-L", hb.LIB, "-lsimpler_hip", rpath])
        os.replace(lib + ".tmp", lib)
    for m in MODULES:
        out = module_path(m)
        msrc = os.path.join(HOST, "pybind", f"{m}_pybind.cpp")
        if force or hb._stale(out, [msrc, hdr, lib]):
            hb._run([CXX] + common + ["-fvisibility=hidden"] + inc + _py_includes() +
                    [msrc, "-o", out + ".tmp", "-L", hb.LIB, "-lsimpler_host", "-lsimpler_hip", rpath])
            os.replace(out + ".tmp", out)
    # the reference's command-line front-ends (simplER.cpp, renderer3d_sample.cpp)
    os.makedirs(BIN, exist_ok=True)
    fsrc = [os.path.join(FRONTEND, f) for f in ("settings.cpp", "settings.h", "json_value.h")]
    for name, main in FRONTENDS.items():
        out = frontend_path(name)
        srcs = [os.path.join(FRONTEND, "settings.cpp"), os.path.join(FRONTEND, main)]
        if force or hb._stale(out, fsrc + [os.path.join(FRONTEND, main), hdr, lib]):
            hb._run([CXX, "-O2", "-std=c++17", "-Wall", "-Wno-unused-parameter"] + inc + ["-I", FRONTEND] + srcs +
                    ["-o", out + ".tmp", "-L", hb.LIB, "-lsimpler_host", "-lsimpler_hip",
                     "-Wl,-rpath,$ORIGIN/../lib"])
            os.replace(out + ".tmp", out)


if __name__ == "__main__":
    build()
