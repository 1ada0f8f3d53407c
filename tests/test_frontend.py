"""The command-line front-ends (no GPU): `simplER settings.json` and
`renderer3d_sample key=value ...` parse the reference's settings, apply its
error rules and build the scene it builds (reference simplER.cpp:21-164,
settings_parser.cpp, json_helper.h, renderer3d_sample.cpp:41-870).  With
SIMPLER_DESCRIBE set the executables write the engine descriptor of the run
instead of rendering; these tests read it back."""
import copy
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

from simpler_amd import build_host, scenes
from test_host_layer import write_pfm_double

# The reference's renderings/settings/fus.json, as a dict (texture path set
# per test).
FUS = {
    "stricts": True, "rendering_type": "rif_sources", "threads": -1, "emitter_gap": 0.0, "sensor_gap": 0.0,
    "use_bounce_decomposition": False, "print_inputs": True, "projector_texture": "",
    "film_parameters": {"path_length": {"min": 0.0, "max": 100000.0, "bins": 1}, "spatial": {"x": 1024, "y": 1024}},
    "adoc_geometry_parameters": {"half_theta_limit": 0.0, "emitter_size": 0.001, "sensor_size": 0.001,
                                 "emitter_distance": 0.0, "sensor_distance": 0.0},
    "scattering_parameters": {"sigma_t": 0.0, "albedo": 0.9, "g_val": 0.9},
    "rendering_parameters": {"num_photons": 10000, "use_direct": True, "max_depth": -1, "max_pathlength": 100000.0,
                             "use_angular_sampling": True},
    "importance_sampling_parameters": {"distribution": "none", "g_or_kappa": 1.0},
    "scene_parameters": {"medium_lx": -0.00635, "medium_rx": 0.00635},
    "rif_parameters": {"f_u": 2000000.0, "speed_u": 1500.0, "precision": 4, "rr_weight": 0.001,
                       "er_stepsize": 0.0001, "direct_to_l": 0.000001, "use_initialization_hack": False,
                       "n_o": 1.3333, "n_scaling": 0.0000215, "n_coeff": 1, "radius": 0.0508,
                       "center1": [0.0, 0.0, -0.0508], "center2": [0.0, -0.0508, 0.0], "active1": True,
                       "active2": True, "phase1": 3.141592653589793, "phase2": 3.141592653589793,
                       "chordlength": 0.0127, "theta_sources": 100, "trans_z_sources": 51},
    "lens_parameters": {"emitter_lens_aperture": 0.3, "emitter_lens_focal_length": 0.3,
                        "emitter_lens_active": False, "sensor_lens_aperture": 0.3,
                        "sensor_lens_focal_length": 0.3, "sensor_lens_active": False},
}


def exe(name):
    path = build_host.frontend_path(name)
    if not os.path.exists(path):
        build_host.build(force=False)
    return path


@pytest.fixture()
def texture(tmp_path):
    p = str(tmp_path / "disk.pfm")
    write_pfm_double(p, scenes.disk_texture(101))
    return p


def load_describe(path):
    """The descriptor prints float32 values with 9 significant digits; read
    them back as float32 (the double sums stay double)."""
    d = json.loads(path.read_text())
    for k, v in d.items():
        if k in ("texture_sum", "spline_coeff_sum", "source_pos"):
            continue
        if isinstance(v, float):
            d[k] = f32(v)
        elif isinstance(v, list):
            d[k] = [f32(x) if isinstance(x, float) else x for x in v]
    return d


def run_simpler(tmp_path, settings, describe=True):
    cfg = tmp_path / "settings.json"
    cfg.write_text(json.dumps(settings) if isinstance(settings, dict) else settings)
    env = dict(os.environ)
    d = tmp_path / "describe.json"
    if describe:
        env["SIMPLER_DESCRIBE"] = str(d)
    r = subprocess.run([exe("simplER"), str(cfg)], cwd=tmp_path, env=env, capture_output=True, text=True,
                       timeout=120)
    desc = load_describe(d) if describe and r.returncode == 0 else None
    return r, desc


def run_sample(tmp_path, args):
    env = dict(os.environ)
    d = tmp_path / "describe.json"
    env["SIMPLER_DESCRIBE"] = str(d)
    r = subprocess.run([exe("renderer3d_sample")] + args, cwd=tmp_path, env=env, capture_output=True, text=True,
                       timeout=120)
    return r, (load_describe(d) if r.returncode == 0 else None)


def f32(x):
    return float(np.float32(x))


def test_fus_settings_build_the_reference_scene(tmp_path, texture):
    s = copy.deepcopy(FUS)
    s["projector_texture"] = texture
    r, d = run_simpler(tmp_path, s)
    assert r.returncode == 0, r.stderr
    assert "simplER renderer started with the following parameters" in r.stdout
    assert re.fullmatch(r"rif_sources_\d{4}_\d{2}_\d{2}_\d{2}_\d{2}\.pfm3d", d["output_file"])
    # geometry (simplER.cpp:60-85): only x comes from the settings
    assert d["block_l"] == [f32(-0.00635), -5, -5] and d["block_r"] == [f32(0.00635), 5, 5]
    assert d["light_origin"] == [f32(0.00635), 0, 0] and d["view_origin"] == [f32(-0.00635), 0, 0]
    assert d["emitter_lens_origin"] == [f32(0.00635), 0, 0] and d["sensor_lens_origin"] == [f32(-0.00635), 0, 0]
    assert d["egap_end_loc_x"] == f32(0.00635) and d["sgap_begin_loc_x"] == f32(-0.00635)
    assert d["light_dir"][0] == -1.0 and abs(d["light_dir"][1]) < 1e-7  # cos / sin of float(pi)
    assert d["view_dir"] == [-1, 0, 0] and d["view_horizontal"] == [0, -1, 0]
    assert d["light_plane"] == [f32(0.001)] * 2 and d["view_plane"] == [f32(0.001)] * 2
    assert d["pathlength_range"] == [0, 100000] and (d["x_res"], d["y_res"], d["z_res"]) == (1024, 1024, 1)
    assert d["ior"] == f32(1.3333) and d["Li"] == 75000
    assert d["axis_uz"] == [-1, 0, 0] and d["axis_ux"] == [0, 0, 1]
    # medium: sigma_t = 0 forces albedo 0 (medium.h:24-38)
    assert d["sigma_t"] == 0 and d["albedo"] == 0 and d["g"] == f32(0.9)
    assert d["use_direct"] == 1 and d["max_depth"] == -1 and d["num_photons"] == 10000
    assert d["er_stepsize"] == f32(1e-4) and d["precision"] == 4 and d["f_u"] == 2e6
    # the source lattice (rif_sources.h:156-198): 100 x 51 per group, two groups
    assert d["field_kind"] == 1 and d["n_sources"] == 2 * 100 * 51
    th = math.asin(0.0127 / (2 * 0.0508))
    ref = scenes.reference_arc_sources(0.0508, (0.0, 0.0, -0.0508), (0.0, -0.0508, 0.0), True, True, math.pi,
                                       math.pi, -th, th, 100, -0.0127 / 2, 0.0127 / 2, 51, 2.15e-5, 1)
    pos = np.array(d["source_pos"]).reshape(-1, 3)
    # the lattice is computed in float32 (Float), the restatement in double: 2 ulp at 5 cm
    np.testing.assert_allclose(pos, ref["pos"], rtol=0, atol=1e-8)
    # texture: the PFM file, rows flipped back to top-to-bottom
    assert (d["tex_x"], d["tex_y"]) == (101, 101)
    assert d["texture_sum"] == float(scenes.disk_texture(101).sum())


def test_json_bessel_settings(tmp_path, texture):
    s = copy.deepcopy(FUS)
    s["projector_texture"] = texture
    s["rendering_type"] = "rif_analytical"
    s["emitter_gap"] = 0.002
    s["sensor_gap"] = 0.001
    s["adoc_geometry_parameters"].update(emitter_distance=0.001, sensor_distance=0.003)
    s["lens_parameters"].update(emitter_lens_active=True, sensor_lens_active=True)
    rp = s["rif_parameters"]
    for k in ("n_scaling", "n_coeff", "radius", "center1", "center2", "active1", "active2", "phase1", "phase2",
              "chordlength", "theta_sources", "trans_z_sources"):
        rp.pop(k)
    rp.update(n_max=2e-5, n_clip=1e-5, phi_min=1.0, phi_max=2.0, mode=2)
    r, d = run_simpler(tmp_path, s)
    assert r.returncode == 0, r.stderr
    assert d["field_kind"] == 0 and d["mode"] == 2 and d["n_max"] == f32(2e-5) and d["n_clip"] == f32(1e-5)
    assert d["phi_min"] == 1.0 and d["phi_max"] == 2.0
    assert d["light_origin"] == [f32(f32(0.00635) + f32(0.001)), 0, 0]
    assert d["view_origin"] == [f32(f32(-0.00635) - f32(0.003)), 0, 0]
    assert d["egap_end_loc_x"] == f32(f32(0.00635) - f32(0.002))
    assert d["sgap_begin_loc_x"] == f32(f32(-0.00635) + f32(0.001))
    assert d["emitter_lens_active"] == 1 and d["sensor_lens_active"] == 1 and d["n_sources"] == 0
    assert re.fullmatch(r"rif_analytical_\d{4}(_\d{2}){4}\.pfm3d", d["output_file"])


def test_json_interpolated_settings(tmp_path, texture, mods_scene):
    vol = str(tmp_path / "n.vol")
    grid = 1.3333 + 1e-4 * np.random.default_rng(0).standard_normal((9, 8, 7)).astype(np.float32)
    mods_scene.write_vol(vol, grid, [-0.005, -0.004, -0.003], [0.005, 0.004, 0.003])
    s = copy.deepcopy(FUS)
    s["projector_texture"] = texture
    s["rendering_type"] = "rif_interpolated"
    s["rif_parameters"] = {k: FUS["rif_parameters"][k] for k in
                           ("f_u", "speed_u", "precision", "rr_weight", "er_stepsize", "direct_to_l",
                            "use_initialization_hack", "n_o")}
    s["rif_parameters"]["rif_grid_file"] = vol
    r, d = run_simpler(tmp_path, s)
    assert r.returncode == 0, r.stderr
    assert d["field_kind"] == 2 and d["spline_dims"] == [7, 8, 9]
    coeff = mods_scene.prefilter3d(grid)
    assert d["spline_coeff_sum"] == pytest.approx(float(coeff.astype(np.float64).sum()), rel=1e-12)


@pytest.fixture()
def mods_scene():
    import sys
    from simpler_amd import build
    if not os.path.exists(build_host.module_path("scene")):
        build_host.build(force=False)
    if build.LIB not in sys.path:
        sys.path.insert(0, build.LIB)
    import scene_pybind
    return scene_pybind


def test_json_error_rules(tmp_path, texture):
    base = copy.deepcopy(FUS)
    base["projector_texture"] = texture
    # malformed JSON and a non-object top level
    r, _ = run_simpler(tmp_path, '{"stricts": true,')
    assert r.returncode == 1 and "parse" in r.stderr
    r, _ = run_simpler(tmp_path, "[1, 2]")
    assert r.returncode == 1
    # a missing settings file
    r = subprocess.run([exe("simplER"), str(tmp_path / "nope.json")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Failed to open" in r.stderr
    # "stricts" itself missing always fails (settings_parser.cpp:136-141)
    s = copy.deepcopy(base)
    s.pop("stricts")
    r, _ = run_simpler(tmp_path, s)
    assert r.returncode == 1 and "stricts" in r.stderr
    # unknown rendering type
    s = copy.deepcopy(base)
    s["rendering_type"] = "rif_magic"
    r, _ = run_simpler(tmp_path, s)
    assert r.returncode == 1 and "Invalid rendering_type" in r.stderr
    # strict run with a missing nested key fails ...
    s = copy.deepcopy(base)
    s["film_parameters"]["spatial"].pop("y")
    r, _ = run_simpler(tmp_path, s)
    assert r.returncode == 1 and "Missing key" in r.stderr
    # ... a lenient one keeps the defaults of that block and stops reading there
    s["stricts"] = False
    s["film_parameters"]["spatial"]["x"] = 64
    r, d = run_simpler(tmp_path, s)
    assert r.returncode == 1 and "Failed to create scene" in r.stderr  # the field block was never reached
    # a mistyped boolean (json_helper.h get_exact: no conversion)
    s = copy.deepcopy(base)
    s["rendering_parameters"]["use_direct"] = 1
    r, _ = run_simpler(tmp_path, s)
    assert r.returncode == 1 and "use_direct" in r.stderr
    # numbers convert freely (get_num): an integer where a float is expected
    s = copy.deepcopy(base)
    s["scattering_parameters"]["sigma_t"] = 2000
    s["scattering_parameters"]["albedo"] = 1
    r, d = run_simpler(tmp_path, s)
    assert r.returncode == 0 and d["sigma_t"] == 2000 and d["albedo"] == 1
    # the reference's range checks (settings_parser.cpp:152-185)
    s = copy.deepcopy(base)
    s["lens_parameters"]["emitter_lens_active"] = True
    r, _ = run_simpler(tmp_path, s)
    assert r.returncode == 1 and "emitter_distance" in r.stderr
    s = copy.deepcopy(base)
    s["emitter_gap"] = 0.01
    s["sensor_gap"] = 0.01
    r, _ = run_simpler(tmp_path, s)
    assert r.returncode == 1 and "gaps" in r.stderr


# renderings/testing_splines.sh's arguments (the rifgridFile is written per test)
SPLINE_ARGS = ["threads=-1", "stricts=true", "precision=4", "numPhotons=1000000", "outFilePrefix=withsplines",
               "sigmaT=0", "albedo=.9", "gVal=.9", "f_u=898100", "speed_u=1500", "n_o=1.3333", "n_max=0.0278",
               "n_clip=1", "phi_min=+1.5708", "phi_max=+1.5708", "mode=0", "er_stepsize=1e-3", "directTol=1e-6",
               "rr_weight=.001", "useDirect=true", "useAngularSampling=True",
               "use_bounce_decomposition=False", "maxDepth=-1", "maxPathlength=100000", "pathLengthMin=0",
               "pathLengthMax=100000", "pathLengthBins=1", "spatialX=1024", "spatialY=1024", "mediumLx=-.015",
               "mediumRx=.015", "distribution=none", "gOrKappa=1", "halfThetaLimit=0", "emitter_gap=0.0",
               "sensor_gap=0.0", "emitter_size=0.0005", "emitter_distance=0", "emitter_lens_aperture=.3",
               "emitter_lens_focalLength=.3", "emitter_lens_active=false", "sensor_size=0.005",
               "sensor_distance=0", "sensor_lens_aperture=.3", "sensor_lens_focalLength=.3",
               "sensor_lens_active=false", "useInitializationHack=false", "printInputs=true"]


def test_renderer3d_sample_spline_arguments(tmp_path, texture, mods_scene):
    vol = str(tmp_path / "besselriffile.vol")
    grid = np.full((5, 6, 7), 1.3333, np.float32)
    mods_scene.write_vol(vol, grid, [-0.015, -0.001, -0.001], [0.015, 0.001, 0.001])
    args = SPLINE_ARGS + [f"projectorTexture={texture}", f"rifgridFile={vol}"]
    r, d = run_sample(tmp_path, args)
    assert r.returncode == 0, r.stderr
    assert d["output_file"] == "withsplines.pfm3d" and d["field_kind"] == 2 and d["spline_dims"] == [7, 6, 5]
    assert d["er_stepsize"] == f32(1e-3) and d["n_max"] == f32(0.0278) and d["f_u"] == f32(898100)
    assert d["light_plane"] == [f32(0.0005)] * 2 and d["view_plane"] == [f32(0.005)] * 2
    assert d["block_l"][0] == f32(-0.015) and d["use_bounce_decomposition"] == 0 and d["num_photons"] == 1000000
    # the same arguments without the grid build the Bessel field
    r, d = run_sample(tmp_path, [a for a in args if not a.startswith("rifgridFile")] + ["stricts=false"])
    assert r.returncode == 0, r.stderr
    assert d["field_kind"] == 0


def test_renderer3d_sample_defaults_and_errors(tmp_path, texture):
    # defaults of renderer3d_sample.cpp:50-183 (Bessel build)
    r, d = run_sample(tmp_path, [f"projectorTexture={texture}"])
    assert r.returncode == 0, r.stderr
    assert d["output_file"] == "USOCTRendering.pfm3d"
    assert (d["x_res"], d["y_res"], d["z_res"]) == (128, 128, 128) and d["pathlength_range"] == [0, 64]
    assert d["precision"] == 8 and d["er_stepsize"] == f32(1e-3) and d["f_u"] == 848000
    assert d["n_max"] == f32(1e-3) and d["half_theta_limit"] == f32(12.8e-3) and d["num_photons"] == 500
    assert d["use_bounce_decomposition"] == 1 and d["block_l"] == [f32(-0.015), -5, -5]
    # point-source keys select the point-source build with its defaults
    r, d = run_sample(tmp_path, [f"projectorTexture={texture}", "theta_sources=4", "trans_z_sources=3"])
    assert r.returncode == 0, r.stderr
    assert d["field_kind"] == 1 and d["n_sources"] == 2 * 4 * 3 and d["f_u"] == 5e6
    # errors
    for bad, msg in ((["foo"], "format arg=value"), (["a=b=c"], "format arg=value"),
                     (["colour=red"], "Unknown variable"), (["useDirect=maybe"], "true or false"),
                     (["numPhotons=lots"], "Bad number"), (["center1=1,2"], "Three input coordinates"),
                     (["n_max=1e-3", "radius=0.05"], "cannot be mixed"),
                     (["stricts=true", "numPhotons=10"], "crashing as one or more inputs is absent"),
                     (["emitter_gap=1.0"], "invalid gap")):
        r, _ = run_sample(tmp_path, [f"projectorTexture={texture}"] + bad)
        assert r.returncode == 1 and msg in r.stderr, (bad, r.stderr)


def test_json_reader_escapes_numbers_and_malformed_documents(tmp_path):
    """The settings reader: \\u and other escapes in strings (a texture path
    with a space and slashes written as \\u escapes), exponents and integral
    floats where integers are read, a repeated key (the last one counts), and
    malformed documents rejected with the position."""
    sub = tmp_path / "tex dir"
    sub.mkdir()
    tex = sub / "d.pfm"
    write_pfm_double(str(tex), scenes.disk_texture(11))
    escaped = str(tex).replace("/", "\\u002F").replace(" ", "\\u0020")
    s = copy.deepcopy(FUS)
    s["projector_texture"] = "PLACEHOLDER"
    s["film_parameters"] = {"path_length": {"min": 0.0, "max": 1e5, "bins": 2.0}, "spatial": {"x": 32, "y": 16}}
    text = json.dumps(s).replace('"PLACEHOLDER"', '"' + escaped + '"')
    # a repeated key: the later value is the one read
    text = text.replace('"threads": -1', '"threads": 3, "threads": -1')
    r, d = run_simpler(tmp_path, text)
    assert r.returncode == 0, r.stderr
    assert (d["tex_x"], d["tex_y"]) == (11, 11)
    assert d["pathlength_range"] == [0, 100000] and (d["x_res"], d["y_res"], d["z_res"]) == (32, 16, 2)
    assert d["threads"] == -1
    for bad, what in (('{"stricts": true,}', "expected a key string"), ('{"stricts": tru}', "unexpected character"),
                      ('{"a": "no end}', "unterminated string"), ('{"a": 1} x', "trailing characters"),
                      ('{"a": 01}', "expected ',' or '}'"), ('{"a": "\\q"}', "bad escape"),
                      ('{"a": "\\ud800"}', "lone high surrogate"), ('{"a": -}', "bad number")):
        r, _ = run_simpler(tmp_path, bad)
        assert r.returncode == 1 and "JSON parse error at line 1" in r.stderr and what in r.stderr, (bad, r.stderr)
