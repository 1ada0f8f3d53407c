"""The reference-shaped pybind modules render on the GPU exactly what the C
ABI renders for the same scene: Bessel and point-source builds, driven the
way the reference's own Python driver drives them."""
import math
import os
import sys

import numpy as np
import pytest

from simpler_amd import build, build_host, engine, scenes
from simpler_amd._abi import FIELD_SOURCES
from test_host_layer import reference_scene_args, write_pfm_double

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods():
    if not all(os.path.exists(build_host.module_path(m)) for m in build_host.MODULES):
        pytest.fail("pybind modules are not built")
    if build.LIB not in sys.path:
        sys.path.insert(0, build.LIB)
    import image_pybind
    import medium_pybind
    import phase_pybind
    import photon_pybind
    import scene_pybind
    import tvector_pybind
    return dict(tvector=tvector_pybind, scene=scene_pybind, photon=photon_pybind, medium=medium_pybind,
                phase=phase_pybind, image=image_pybind)


def matching_desc(bessel=True, half=0.005):
    d = scenes.base_desc(half=half, step=1e-5, sensor_plane=0.01, emitter_size=0.002)
    d.set(axis_uz=(-1.0, 0.0, 0.0), axis_ux=(0.0, 0.0, 1.0), precision=4, sigma_t=1000.0, albedo=1.0, g=0.9,
          emitter_lens_aperture=0.3, emitter_lens_focal_length=0.3, sensor_lens_aperture=0.3,
          sensor_lens_focal_length=0.3)
    if not bessel:
        d.set(field_kind=FIELD_SOURCES, f_u=5e6, time_coupling=0.0)
    return d


@pytest.mark.parametrize("bessel", [True, False])
def test_pybind_render_equals_c_abi_render(gpu_ctx, mods, tmp_path, bessel):
    tex = scenes.disk_texture(51)
    tex_path = str(tmp_path / "disk.pfm")
    write_pfm_double(tex_path, tex)
    scene = mods["scene"].Scene(*reference_scene_args(mods, tex_path, bessel=bessel))
    hg = mods["phase"].henyey_greenstein(0.9)
    medium = mods["medium"].Medium(1000.0, 1.0, hg)
    renderer = mods["photon"].Renderer(-1, -1.0, True, True, -1)
    img = mods["image"].SmallImage(32, 32, 16)
    n = 4000
    renderer.renderImage(img, medium, scene, n)
    got = img.numpy().copy()

    d = matching_desc(bessel)
    src = None
    if not bessel:
        r = 2 * 25.4e-3
        chord = 0.5 * 25.4e-3
        src = scenes.reference_arc_sources(radius=r, center1=(-r, 0, 0), center2=(-r, 0, 0), active1=True,
                                           active2=True, phase1=0.0, phase2=0.0,
                                           theta_min=-math.asin(chord / (2 * r)), theta_max=math.asin(chord / (2 * r)),
                                           theta_sources=10, trans_z_min=-chord / 2, trans_z_max=chord / 2,
                                           trans_z_sources=12, n_scaling=0.05e-3, n_coeff=1)
    ref, st = engine.Scene(gpu_ctx, d, tex, sources=src).render(n, (32, 32, 16), seed=0)
    ref = ref.double().cpu().numpy()
    assert st["photons"] == n
    assert abs(got.sum() - ref.sum()) <= 1e-5 * ref.sum()
    assert np.abs(got - ref).max() <= 1e-4 * ref.max()
