"""The reference-shaped host layer (C++ scn / photon / med / pfunc / image /
tvec API and the six same-named pybind modules) without a GPU: module
import, the reference's constructor argument orders, PFM / PFM3D / VOL I/O,
and the C++ B-spline prefilter against the reference's golden coefficients."""
import math
import os
import struct
import sys

import numpy as np
import pytest

from simpler_amd import build, build_host

LIB = build.LIB


@pytest.fixture(scope="module")
def mods():
    if not all(os.path.exists(build_host.module_path(m)) for m in build_host.MODULES):
        build_host.build(force=False)
    if LIB not in sys.path:
        sys.path.insert(0, LIB)
    import image_pybind
    import medium_pybind
    import phase_pybind
    import photon_pybind
    import scene_pybind
    import tvector_pybind
    return dict(tvector=tvector_pybind, scene=scene_pybind, photon=photon_pybind, medium=medium_pybind,
                phase=phase_pybind, image=image_pybind)


def write_pfm_double(path, img):
    """The reference's PFM flavour: 'Pf' header, double raster, rows bottom
    to top (renderer/include/image.h:183-230)."""
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(f"Pf\n{w} {h}\n-1\n".encode())
        f.write(np.ascontiguousarray(img[::-1], np.float64).tobytes())


def reference_scene_args(mods, tex_path, bessel=True, ior=1.3333, half=0.005):
    """Argument list in the order of renderer/pybindFiles/tests/
    test_renderer3d_sample.py:250-306 (sources build) or the Bessel build."""
    tv = mods["tvector"]
    V3, V2 = tv.Vec3f, tv.Vec2f
    common = [ior, V3(-half, -half, -half), V3(half, half, half), V3(half, 0, 0), V3(-1, 0, 0), 0.0, tex_path,
              V2(0.002, 0.002), 75000.0, V3(-half, 0, 0), V3(-1, 0, 0), V3(0, -1, 0), V2(0.01, 0.01), V2(0, 0.1),
              False, "none", 1.0, V3(half, 0, 0), 0.3, 0.3, False, V3(-half, 0, 0), 0.3, 0.3, False]
    tail = [V3(-1, 0, 0), V3(0, 0, 1), V3(0, 0, 0), 1e-5, 1e-6, 1e-3, 4, half, -half, False]
    if bessel:
        field = [848e3, 1500.0, ior, 1e-5, 1e-5, math.pi / 2 - 0.2 * math.pi, math.pi / 2 + 0.2 * math.pi, 0]
    else:
        r = 2 * 25.4e-3
        chord = 0.5 * 25.4e-3
        field = [5e6, 1500.0, ior, 0.05e-3, 1, r, V3(-r, 0, 0), V3(-r, 0, 0), True, True, 0.0, 0.0,
                 -math.asin(chord / (2 * r)), math.asin(chord / (2 * r)), 10, -chord / 2, chord / 2, 12]
    return common + field + tail


def test_modules_import_and_types(mods):
    tv = mods["tvector"]
    v = tv.Vec3f(1.0, 2.0, 3.0)
    assert [v.index(i) for i in range(3)] == [1.0, 2.0, 3.0]
    i3 = tv.Vec3i(128, 64, 32)
    assert (i3.x, i3.y, i3.z) == (128, 64, 32)
    hg = mods["phase"].henyey_greenstein(0.9)
    med = mods["medium"].Medium(1000.0, 1.0, hg)
    assert med.getSigmaT() == 1000.0 and abs(med.getMfp() - 1e-3) < 1e-9
    clear = mods["medium"].Medium(0.0, 1.0, hg)
    assert clear.getAlbedo() == 0.0  # sigma_t <= eps forces albedo 0 (medium.h:33-37)
    img = mods["image"].SmallImage(8, 4, 2)
    assert (img.get_x_res(), img.get_y_res(), img.getZRes()) == (8, 4, 2)
    img.setPixel(7, 3, 1, 2.5)
    assert img.get_pixel(7, 3, 1) == 2.5 and img.numpy()[1, 3, 7] == 2.5
    r = mods["photon"].Renderer(-1, -1.0, True, True, -1)
    assert r.getSeed() == 0


def test_scene_constructors_follow_reference_order(mods, tmp_path):
    tex = str(tmp_path / "disk.pfm")
    from simpler_amd.scenes import disk_texture
    write_pfm_double(tex, disk_texture(51))
    S = mods["scene"].Scene
    s_b = S(*reference_scene_args(mods, tex, bessel=True))
    assert s_b.field_kind() == 0
    s_b.set_n_max(2e-5)
    s_b.set_f_u(1e6)
    s_s = S(*reference_scene_args(mods, tex, bessel=False))
    assert s_s.field_kind() == 1
    s_s.set_n_scaling(1e-4)
    s_s.set_phase1(0.5)
    s_s.set_phase2(1.5)
    with pytest.raises(RuntimeError):
        S(*reference_scene_args(mods, str(tmp_path / "missing.pfm")))


def test_spline_scene_from_vol(mods, tmp_path):
    tex = str(tmp_path / "w.pfm")
    write_pfm_double(tex, np.ones((1, 1)))
    vol = str(tmp_path / "n.vol")
    grid = np.full((6, 5, 4), 1.3333, np.float32)
    mods["scene"].write_vol(vol, grid, [-0.005] * 3, [0.005] * 3)
    a, lo, hi = mods["scene"].read_vol(vol)
    assert np.array_equal(a, grid) and lo == pytest.approx([-0.005] * 3) and hi == pytest.approx([0.005] * 3)
    s = mods["scene"].Scene(*(reference_scene_args(mods, tex) + [vol]))
    assert s.field_kind() == 2


def test_cpp_prefilter_matches_reference_coefficients(mods):
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "spline_golden.npz"))
    data = g["i3d_data"].reshape(10, 10, 10)  # (z, y, x), x fastest
    c = mods["scene"].prefilter3d(data).astype(np.float64)
    printed = np.zeros((100, 10))
    for k in range(10):
        for i in range(10):
            printed[k * 10 + i] = c[k, :, i]
    # FP32 output of a double filter: agree to float rounding
    np.testing.assert_allclose(printed, g["c3d"], rtol=1e-6, atol=1e-6)


def test_pfm3d_round_trip(mods, tmp_path):
    img = mods["image"].SmallImage(5, 4, 3)
    for z in range(3):
        for y in range(4):
            for x in range(5):
                img.setPixel(x, y, z, x + 10 * y + 100 * z)
    f = str(tmp_path / "o.pfm3d")
    img.writePFM3D(f)
    raw = open(f, "rb").read()
    assert raw.startswith(b"Pf\n5 4 3\n-1\n")
    vals = np.frombuffer(raw[len(b"Pf\n5 4 3\n-1\n"):], "<f8")
    assert vals[1 + 5 * 2 + 20 * 1] == 1 + 10 * 2 + 100 * 1
    img2 = mods["image"].SmallImage(1, 1, 1)
    img2.readPFM3D(f)
    assert np.array_equal(img2.numpy(), img.numpy())
