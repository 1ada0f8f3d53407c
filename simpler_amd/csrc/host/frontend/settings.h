// settings.h -- the reference's two command-line front-ends on the gfx950
// engine: `simplER <settings.json>` (reference simplER.cpp:21-164 with
// settings_parser.{h,cpp} and json_helper.h) and `renderer3d_sample
// key=value ...` (reference renderer/samples/renderer3d_sample.cpp:41-870).
//
// Both fill one Settings record, build the scene through the reference-shaped
// C++ API (scn::Scene, med::Medium, photon::Renderer, image::SmallImage; see
// simpler_compat.h), render on the GPU and write the image as PFM3D.
#pragma once
#include <cstdint>
#include <ostream>
#include <string>

#include "simpler_compat.h"

namespace simpler_frontend {

enum class Field { Analytical, Sources, Interpolated };

// One render's parameters.  Member defaults are the reference's "vanilla"
// initialisers (settings_parser.h:24-236, execution_parameters at 238-330);
// the key=value front-end starts from its own defaults instead
// (renderer3d_sample.cpp:50-183), set by argv_defaults().
struct Settings {
    // execution / system
    bool stricts = false;
    std::string rendering_type = "analytic_rif";
    std::string output_file_name;  // without the .pfm3d extension
    int threads = -1;
    float emitter_gap = 0.f, sensor_gap = 0.f;
    bool use_bounce_decomposition = true;
    bool print_inputs = true;
    std::string projector_texture = "White.pfm";

    // adoc geometry
    float half_theta_limit = 12.8e-3f, emitter_size = 0.002f, sensor_size = 0.002f;
    float emitter_distance = 0.f, sensor_distance = 0.f;
    // film
    float path_length_min = 0.f, path_length_max = 64.f;
    int path_length_bins = 128, spatial_x = 128, spatial_y = 128;
    // scattering
    float sigma_t = 0.f, albedo = 1.f, g_val = 0.f;
    // rendering
    int64_t num_photons = 500;
    int max_depth = -1;
    float max_pathlength = -1.f;
    bool use_direct = false, use_angular_sampling = true;
    // scene: only the x extent is a setting; y and z keep +-5 m
    tvec::Vec3f medium_l{-0.015f, -5.f, -5.f}, medium_r{0.015f, 5.f, 5.f};
    // final-path importance sampling (accepted, unused by the forward path)
    std::string distribution = "vmf";
    float g_or_kappa = 4.f;
    // lenses
    float emitter_lens_aperture = 0.015f, emitter_lens_focal_length = 0.015f;
    bool emitter_lens_active = false;
    float sensor_lens_aperture = 0.015f, sensor_lens_focal_length = 0.015f;
    bool sensor_lens_active = false;

    // refractive-index field ("rif_parameters")
    Field field = Field::Analytical;
    bool have_rif = false;  // false when the settings stopped before the field block
    float f_u = 848e3f, speed_u = 1500.f, n_o = 1.3333f;
    float direct_to_l = 1e-5f, rr_weight = 1e-2f, er_stepsize = 1e-3f;
    int precision = 8;
    bool use_initialization_hack = true;
    // Bessel (rif_analytical); the interpolated field samples its duty phase
    // with the same parameters
    float n_max = 1e-3f, n_clip = 1e-3f, phi_min = 1.57079633f, phi_max = 1.57079633f;
    int mode = 0;
    // point sources (rif_sources)
    float n_scaling = 0.05e-3f;
    int n_coeff = 1;
    float radius = 2 * 25.4e-3f;
    tvec::Vec3f center1{-2 * 25.4e-3f, 0.f, 0.f}, center2{-2 * 25.4e-3f, 0.f, 0.f};
    bool active1 = true, active2 = true;
    float phase1 = 0.f, phase2 = 0.f, chordlength = 0.5f * 25.4e-3f;
    int theta_sources = 100, trans_z_sources = 501;
    // spline grid (rif_interpolated)
    std::string rif_grid_file = "us";
};

// Status codes of the front-ends (process exit codes).
constexpr int kOk = 0;
constexpr int kFail = 1;

// simplER.cpp: read settings.json.  Follows execution_parameters(AnyMap)
// (settings_parser.cpp:62-148): a missing or mistyped key stops the parse
// with a message; the run then continues with the values read so far unless
// "stricts" is true or the failing key is "stricts" itself.  An unknown
// rendering_type always fails.  Returns kOk / kFail.
int settings_from_json_file(const std::string &path, Settings &s, std::ostream &err);
int settings_from_json_text(const std::string &text, Settings &s, std::ostream &err);

// renderer3d_sample.cpp: key=value arguments (argv[1..argc-1]).  The field
// kind follows the keys: rifgridFile selects the spline grid, any of the
// point-source keys selects point sources, otherwise Bessel (the reference
// selects it at build time with USE_RIF_SOURCES / USE_RIF_INTERPOLATED).
// Returns kOk / kFail (bad token, unknown key, mixed field keys, a strict run
// with missing keys, or a failed range check).
int settings_from_argv(int argc, const char *const *argv, Settings &s, std::ostream &err);

// settings_parser.cpp:152-185 and renderer3d_sample.cpp:788-797.  The
// reference states these as assert()s (compiled out in its release build);
// here they are errors.  Returns kOk / kFail.
int sanity_checks(const Settings &s, std::ostream &err);

void print_settings(std::ostream &os, const Settings &s);

// Timestamped output name of simplER.cpp: <rendering_type>_YYYY_MM_DD_HH_MM.
std::string timestamped_name(const std::string &rendering_type);

// Build the scene exactly as simplER.cpp:57-151 / renderer3d_sample.cpp:
// 806-863 do, render num_photons on the GPU and write <out>.pfm3d.  With
// describe_path non-empty nothing is rendered: the engine descriptor and the
// run's sizes are written there as JSON (no GPU needed).  Returns kOk / kFail.
int run(const Settings &s, const std::string &out_prefix, const std::string &describe_path, std::ostream &err);

}  // namespace simpler_frontend
