"""The command-line front-ends render on the GPU: `simplER settings.json`
and `renderer3d_sample key=value ...` write a PFM3D image equal to the C-ABI
render of the scene they describe (same photons, seed 0), for the Bessel,
point-source and spline-grid fields."""
import copy
import math
import os
import subprocess

import numpy as np
import pytest

from simpler_amd import build_host, engine, scenes
from simpler_amd._abi import SceneDesc
from test_frontend import FUS, load_describe, run_sample
from test_host_layer import write_pfm_double

pytestmark = pytest.mark.gpu

DESC_FIELDS = [f for f, _ in SceneDesc._fields_]


def read_pfm3d(path):
    """The reference's PFM3D (image.cpp:161-206): 'Pf', 'X Y Z', scale, then
    X*Y*Z doubles, z-major; the raster is the last 8*X*Y*Z bytes."""
    raw = open(path, "rb").read()
    tok = raw.split(maxsplit=5)
    assert tok[0] == b"Pf"
    X, Y, Z = int(tok[1]), int(tok[2]), int(tok[3])
    data = np.frombuffer(raw[len(raw) - 8 * X * Y * Z:], np.float64)
    return data.reshape(Z, Y, X)


def desc_from(d):
    sd = SceneDesc()
    sd.set(**{k: d[k] for k in DESC_FIELDS if k in d})
    return sd


def engine_render(gpu_ctx, d, tex, sources=None, spline=None):
    sc = engine.Scene(gpu_ctx, desc_from(d), tex, sources=sources, spline=spline)
    img, _ = sc.render(d["num_photons"], (d["x_res"], d["y_res"], d["z_res"]), seed=0)
    return img.double().cpu().numpy()


def run_exe(tmp_path, cmd):
    print("front-end run:", " ".join(cmd[:3]), "...", flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r


def compare(img, ref):
    assert img.shape == ref.shape
    assert ref.sum() > 0
    # one engine, the same photons: only the order of the splat atomics differs
    assert abs(img.sum() - ref.sum()) <= 1e-5 * ref.sum()
    assert np.abs(img - ref).max() <= 1e-4 * ref.max()


def test_simpler_json_bessel_render(gpu_ctx, tmp_path):
    tex_path = str(tmp_path / "disk.pfm")
    tex = scenes.disk_texture(101)
    write_pfm_double(tex_path, tex)
    s = copy.deepcopy(FUS)
    s.update(projector_texture=tex_path, rendering_type="rif_analytical", print_inputs=False)
    s["film_parameters"] = {"path_length": {"min": 0.0, "max": 0.1, "bins": 16}, "spatial": {"x": 64, "y": 64}}
    s["adoc_geometry_parameters"].update(emitter_size=0.002, sensor_size=0.01)
    s["scattering_parameters"].update(sigma_t=1000.0, albedo=1.0)
    s["rendering_parameters"]["num_photons"] = 4096
    s["scene_parameters"] = {"medium_lx": -0.005, "medium_rx": 0.005}
    rp = {k: FUS["rif_parameters"][k] for k in ("speed_u", "precision", "rr_weight", "direct_to_l",
                                                 "use_initialization_hack", "n_o")}
    rp.update(f_u=848e3, er_stepsize=2e-5, n_max=2e-5, n_clip=2e-5, phi_min=1.0, phi_max=2.0, mode=2)
    s["rif_parameters"] = rp
    import json
    cfg = tmp_path / "bessel.json"
    cfg.write_text(json.dumps(s))
    env = dict(os.environ, SIMPLER_DESCRIBE=str(tmp_path / "d.json"))
    r = subprocess.run([build_host.frontend_path("simplER"), str(cfg)], cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = load_describe(tmp_path / "d.json")
    run_exe(tmp_path, [build_host.frontend_path("simplER"), str(cfg)])
    out = [f for f in os.listdir(tmp_path) if f.endswith(".pfm3d")]
    assert len(out) == 1 and out[0].startswith("rif_analytical_")
    img = read_pfm3d(str(tmp_path / out[0]))
    compare(img, engine_render(gpu_ctx, d, tex))


def test_renderer3d_sample_sources_render(gpu_ctx, tmp_path):
    tex_path = str(tmp_path / "disk.pfm")
    tex = scenes.disk_texture(101)
    write_pfm_double(tex_path, tex)
    args = [f"projectorTexture={tex_path}", "outFilePrefix=src", "numPhotons=4096", "spatialX=48", "spatialY=48",
            "pathLengthMin=0", "pathLengthMax=0.1", "pathLengthBins=8", "mediumLx=-0.005", "mediumRx=0.005",
            "sigmaT=1000", "albedo=1", "gVal=0.9", "er_stepsize=2e-5", "precision=4", "useDirect=true",
            "emitter_size=0.002", "sensor_size=0.01", "halfThetaLimit=0", "theta_sources=4",
            "trans_z_sources=3", "n_scaling=2e-5", "n_coeff=2", "phase1=0.3", "phase2=1.1", "printInputs=false"]
    r, d = run_sample(tmp_path, args)
    assert r.returncode == 0, r.stderr
    run_exe(tmp_path, [build_host.frontend_path("renderer3d_sample")] + args)
    img = read_pfm3d(str(tmp_path / "src.pfm3d"))
    rad, chord = 2 * 25.4e-3, 0.5 * 25.4e-3
    th = math.asin(chord / (2 * rad))
    src = scenes.reference_arc_sources(rad, (-rad, 0.0, 0.0), (-rad, 0.0, 0.0), True, True, 0.3, 1.1, -th, th, 4,
                                       -chord / 2, chord / 2, 3, 2e-5, 2)
    np.testing.assert_allclose(np.array(d["source_pos"]).reshape(-1, 3), src["pos"], atol=1e-8)
    src["pos"] = np.array(d["source_pos"], np.float32).reshape(-1, 3)  # the lattice the front-end built
    compare(img, engine_render(gpu_ctx, d, tex, sources=src))


def test_renderer3d_sample_spline_grid_render(gpu_ctx, tmp_path):
    import sys
    from simpler_amd import build
    if build.LIB not in sys.path:
        sys.path.insert(0, build.LIB)
    import scene_pybind
    tex_path = str(tmp_path / "disk.pfm")
    tex = scenes.disk_texture(101)
    write_pfm_double(tex_path, tex)
    vals, dims, lo, hi = scenes.gaussian_spline_coeffs(n=24, half=0.005, max_dn=5e-5, seed=3)
    vol = str(tmp_path / "grid.vol")
    grid = np.asarray(vals, np.float32).reshape(dims[2], dims[1], dims[0])
    scene_pybind.write_vol(vol, grid, list(lo), list(hi))
    args = [f"projectorTexture={tex_path}", f"rifgridFile={vol}", "outFilePrefix=spl", "numPhotons=2048",
            "spatialX=48", "spatialY=48", "pathLengthMin=0", "pathLengthMax=0.1", "pathLengthBins=8",
            "mediumLx=-0.005", "mediumRx=0.005", "sigmaT=1000", "albedo=1", "gVal=0.9", "er_stepsize=2e-5",
            "precision=4", "useDirect=true", "emitter_size=0.002", "sensor_size=0.01", "halfThetaLimit=0",
            "printInputs=false"]
    r, d = run_sample(tmp_path, args)
    assert r.returncode == 0, r.stderr
    run_exe(tmp_path, [build_host.frontend_path("renderer3d_sample")] + args)
    img = read_pfm3d(str(tmp_path / "spl.pfm3d"))
    # the front-end prefilters the VOL grid on the GPU, as set_spline_samples does
    sc = engine.Scene(gpu_ctx, desc_from(d), tex)
    sc.set_spline_samples(grid.reshape(-1), list(dims), list(lo), list(hi))
    ref, _ = sc.render(d["num_photons"], (d["x_res"], d["y_res"], d["z_res"]), seed=0)
    compare(img, ref.double().cpu().numpy())
    # and its coefficients are the host prefilter's to FP32 accuracy
    host = scene_pybind.prefilter3d(grid).reshape(-1)
    assert np.abs(sc.spline_coeffs() - host).max() <= 1e-6 * np.abs(host).max()
