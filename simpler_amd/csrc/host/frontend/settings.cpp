// settings.cpp -- the simplER (JSON) and renderer3d_sample (key=value)
// front-ends: parsing, checks, scene construction and the render.
#include "settings.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <set>
#include <sstream>
#include <vector>

#include "json_value.h"

namespace simpler_frontend {

namespace {

const char *field_name(Field f) {
    switch (f) {
        case Field::Analytical: return "rif_analytical";
        case Field::Sources: return "rif_sources";
        default: return "rif_interpolated";
    }
}

}  // namespace

std::string timestamped_name(const std::string &rendering_type) {
    // settings_parser.cpp:89-95
    std::time_t t = std::time(nullptr);
    std::tm tm = *std::localtime(&t);
    char buf[32];
    std::strftime(buf, sizeof(buf), "%Y_%m_%d_%H_%M", &tm);
    return rendering_type + "_" + buf;
}

// ---------------------------------------------------------------- JSON
int settings_from_json_text(const std::string &text, Settings &s, std::ostream &err) {
    using namespace jsonv;
    Value j;
    try {
        j = parse(text);
    } catch (const std::exception &e) {
        err << "Failed to parse settings: " << e.what() << "\n";
        return kFail;
    }
    if (j.kind != Value::Object) {  // json_helper.h:47-64
        err << "Failed to parse settings: the top level must be an object\n";
        return kFail;
    }
    try {
        // flat / top-level (settings_parser.cpp:67-77)
        s.stricts = boolean(j, "stricts");
        s.rendering_type = str(j, "rendering_type");
        s.threads = num<int>(j, "threads");
        s.emitter_gap = num<float>(j, "emitter_gap");
        s.sensor_gap = num<float>(j, "sensor_gap");
        s.use_bounce_decomposition = boolean(j, "use_bounce_decomposition");
        s.print_inputs = boolean(j, "print_inputs");
        s.projector_texture = str(j, "projector_texture");
        if (s.rendering_type != "rif_analytical" && s.rendering_type != "rif_sources" &&
            s.rendering_type != "rif_interpolated") {
            err << "Invalid rendering_type: '" << s.rendering_type
                << "'. Must be one of 'rif_analytical', 'rif_sources', or 'rif_interpolated'.\n";
            return kFail;
        }
        s.output_file_name = timestamped_name(s.rendering_type);

        // Each nested block is read whole and only then stored, like the
        // reference's per-block constructors (settings_parser.cpp:98-145).
        {
            const Value &f = object(j, "film_parameters");
            Settings b = s;
            b.path_length_min = num<float>(f, "path_length.min");
            b.path_length_max = num<float>(f, "path_length.max");
            b.path_length_bins = num<int>(f, "path_length.bins");
            b.spatial_x = num<int>(f, "spatial.x");
            b.spatial_y = num<int>(f, "spatial.y");
            s = b;
        }
        {
            const Value &g = object(j, "adoc_geometry_parameters");
            Settings b = s;
            b.half_theta_limit = num<float>(g, "half_theta_limit");
            b.emitter_size = num<float>(g, "emitter_size");
            b.sensor_size = num<float>(g, "sensor_size");
            b.emitter_distance = num<float>(g, "emitter_distance");
            b.sensor_distance = num<float>(g, "sensor_distance");
            s = b;
        }
        {
            const Value &m = object(j, "scattering_parameters");
            Settings b = s;
            b.sigma_t = num<float>(m, "sigma_t");
            b.albedo = num<float>(m, "albedo");
            b.g_val = num<float>(m, "g_val");
            s = b;
        }
        {
            const Value &r = object(j, "rendering_parameters");
            Settings b = s;
            b.num_photons = num<int64_t>(r, "num_photons");
            b.max_depth = num<int>(r, "max_depth");
            b.max_pathlength = num<float>(r, "max_pathlength");
            b.use_direct = boolean(r, "use_direct");
            b.use_angular_sampling = boolean(r, "use_angular_sampling");
            s = b;
        }
        {
            const Value &i = object(j, "importance_sampling_parameters");
            Settings b = s;
            b.distribution = str(i, "distribution");
            b.g_or_kappa = num<float>(i, "g_or_kappa");
            s = b;
        }
        {
            const Value &sc = object(j, "scene_parameters");
            Settings b = s;
            b.medium_l.x = num<float>(sc, "medium_lx");
            b.medium_r.x = num<float>(sc, "medium_rx");
            s = b;
        }
        {
            const Value &l = object(j, "lens_parameters");
            Settings b = s;
            b.emitter_lens_aperture = num<float>(l, "emitter_lens_aperture");
            b.emitter_lens_focal_length = num<float>(l, "emitter_lens_focal_length");
            b.emitter_lens_active = boolean(l, "emitter_lens_active");
            b.sensor_lens_aperture = num<float>(l, "sensor_lens_aperture");
            b.sensor_lens_focal_length = num<float>(l, "sensor_lens_focal_length");
            b.sensor_lens_active = boolean(l, "sensor_lens_active");
            s = b;
        }
        {
            // rif_base.h:62-75 + rif_analytical.h:24-35 / rif_sources.h:58-84 /
            // rif_interpolated.h:16-21
            const Value &p = object(j, "rif_parameters");
            Settings b = s;
            b.f_u = num<float>(p, "f_u");
            b.speed_u = num<float>(p, "speed_u");
            b.n_o = num<float>(p, "n_o");
            b.direct_to_l = num<float>(p, "direct_to_l");
            b.rr_weight = num<float>(p, "rr_weight");
            b.er_stepsize = num<float>(p, "er_stepsize");
            b.precision = num<int>(p, "precision");
            b.use_initialization_hack = boolean(p, "use_initialization_hack");
            if (b.rendering_type == "rif_analytical") {
                b.field = Field::Analytical;
                b.n_max = num<float>(p, "n_max");
                b.n_clip = num<float>(p, "n_clip");
                b.phi_min = num<float>(p, "phi_min");
                b.phi_max = num<float>(p, "phi_max");
                b.mode = num<int>(p, "mode");
                if (b.phi_max < b.phi_min)  // rif_analytical.h:66
                    throw std::runtime_error("phi_max must be greater than or equal to phi_min");
            } else if (b.rendering_type == "rif_sources") {
                b.field = Field::Sources;
                b.n_scaling = num<float>(p, "n_scaling");
                b.n_coeff = (int)num<float>(p, "n_coeff");
                b.radius = num<float>(p, "radius");
                const std::vector<float> c1 = num_array<float>(p, "center1");
                const std::vector<float> c2 = num_array<float>(p, "center2");
                if (c1.size() < 3 || c2.size() < 3) throw std::runtime_error("center1/center2 need 3 coordinates");
                b.center1 = tvec::Vec3f(c1[0], c1[1], c1[2]);
                b.center2 = tvec::Vec3f(c2[0], c2[1], c2[2]);
                b.active1 = boolean(p, "active1");
                b.active2 = boolean(p, "active2");
                b.phase1 = num<float>(p, "phase1");
                b.phase2 = num<float>(p, "phase2");
                b.chordlength = num<float>(p, "chordlength");
                b.theta_sources = num<int>(p, "theta_sources");
                b.trans_z_sources = num<int>(p, "trans_z_sources");
            } else {
                b.field = Field::Interpolated;
                b.rif_grid_file = str(p, "rif_grid_file");
                // the duty-phase parameters are optional for the spline grid
                auto opt = [&](const char *k, float &dst) {
                    if (p.o.count(k)) dst = num<float>(p, k);
                };
                opt("n_max", b.n_max);
                opt("n_clip", b.n_clip);
                opt("phi_min", b.phi_min);
                opt("phi_max", b.phi_max);
                if (p.o.count("mode")) b.mode = num<int>(p, "mode");
            }
            b.have_rif = true;
            s = b;
        }
    } catch (const std::exception &e) {
        // settings_parser.cpp:132-146
        err << "Error occurred while starting simplER renderer: " << e.what() << "\n";
        if (std::string(e.what()).find("stricts") != std::string::npos) {
            err << "Error with 'stricts' parameter: " << e.what() << "\n";
            return kFail;
        }
        if (s.stricts) return kFail;
    }
    return kOk;
}

int settings_from_json_file(const std::string &path, Settings &s, std::ostream &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f.is_open()) {
        err << "Failed to open " << path << "\n";
        return kFail;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return settings_from_json_text(ss.str(), s, err);
}

// ----------------------------------------------------------- key=value
namespace {

// renderer3d_sample.cpp:26-38: split on any run of '=' characters
std::vector<std::string> tokenize(const std::string &str, const std::string &delim) {
    std::vector<std::string> tokens;
    std::string::size_type last = str.find_first_not_of(delim, 0);
    std::string::size_type pos = str.find_first_of(delim, last);
    while (std::string::npos != pos || std::string::npos != last) {
        tokens.push_back(str.substr(last, pos - last));
        last = str.find_first_not_of(delim, pos);
        pos = str.find_first_of(delim, last);
    }
    return tokens;
}

bool parse_bool(const std::string &key, std::string v, bool &out, std::ostream &err) {
    std::transform(v.begin(), v.end(), v.begin(), ::tolower);
    if (v == "true") {
        out = true;
        return true;
    }
    if (v == "false") {
        out = false;
        return true;
    }
    err << key << " should be either true or false; Argument " << v << " not recognized\n";
    return false;
}

bool parse_vec3(const std::string &key, const std::string &v, tvec::Vec3f &out, std::ostream &err) {
    std::vector<std::string> parts;
    std::string cur;
    for (char c : v) {
        if (c == ',') {
            parts.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    parts.push_back(cur);
    if (parts.size() < 3) {
        err << "Three input coordinates required for " << key << "\n";
        return false;
    }
    if (parts.size() > 3) {
        err << "More than three input coordinates entered for " << key << "\n";
        return false;
    }
    for (int i = 0; i < 3; ++i) out[i] = std::stof(parts[i]);
    return true;
}

const std::set<std::string> kSourceKeys = {"n_scaling", "n_coeff",  "radius", "center1",     "center2",
                                           "active1",   "active2",  "phase1", "phase2",      "chordlength",
                                           "theta_sources", "trans_z_sources"};
const std::set<std::string> kBesselKeys = {"n_max", "n_clip", "phi_min", "phi_max", "mode"};
const std::vector<std::string> kCommonKeys = {
    "threads", "precision", "numPhotons", "outFilePrefix", "sigmaT", "albedo", "gVal", "f_u", "speed_u", "n_o",
    "emitter_gap", "sensor_gap", "er_stepsize", "directTol", "useInitializationHack", "rr_weight",
    "projectorTexture", "useDirect", "useAngularSampling", "use_bounce_decomposition", "maxDepth",
    "maxPathlength", "pathLengthMin", "pathLengthMax", "pathLengthBins", "spatialX", "spatialY",
    "halfThetaLimit", "emitter_size", "sensor_size", "mediumLx", "mediumRx", "distribution", "gOrKappa",
    "emitter_distance", "emitter_lens_aperture", "emitter_lens_focalLength", "emitter_lens_active",
    "sensor_distance", "sensor_lens_aperture", "sensor_lens_focalLength", "sensor_lens_active", "printInputs"};

}  // namespace

int settings_from_argv(int argc, const char *const *argv, Settings &s, std::ostream &err) {
    // renderer3d_sample.cpp:50-183 defaults (they differ from the JSON ones)
    s = Settings();
    s.output_file_name = "USOCTRendering";
    s.use_bounce_decomposition = true;
    s.print_inputs = true;
    std::set<std::string> given;
    std::vector<std::pair<std::string, std::string>> kv;
    for (int i = 1; i < argc; ++i) {
        const std::vector<std::string> param = tokenize(argv[i], "=");
        if (param.size() != 2) {
            err << "Input argument " << argv[i] << " should be in the format arg=value\n";
            return kFail;
        }
        kv.emplace_back(param[0], param[1]);
        given.insert(param[0]);
    }
    bool src = false, bes = false;
    for (const auto &k : given) {
        src |= kSourceKeys.count(k) > 0;
        bes |= kBesselKeys.count(k) > 0;
    }
    if (src && (bes || given.count("rifgridFile"))) {
        err << "point-source keys cannot be mixed with Bessel or rifgridFile keys (the reference builds one "
               "field per binary)\n";
        return kFail;
    }
    s.field = src ? Field::Sources : (given.count("rifgridFile") ? Field::Interpolated : Field::Analytical);
    s.rendering_type = field_name(s.field);
    s.have_rif = true;
    if (s.field == Field::Sources) {
        s.f_u = 5e6f;  // renderer3d_sample.cpp:121-141
        s.center1 = s.center2 = tvec::Vec3f(-s.radius, 0.f, 0.f);
    }
    try {
        for (const auto &p : kv) {
            const std::string &k = p.first, &v = p.second;
            bool ok = true;
            if (k == "stricts") ok = parse_bool(k, v, s.stricts, err);
            else if (k == "threads") s.threads = std::stoi(v);
            else if (k == "precision") s.precision = std::stoi(v);
            else if (k == "numPhotons") s.num_photons = std::stoll(v);
            else if (k == "outFilePrefix") s.output_file_name = v;
            else if (k == "sigmaT") s.sigma_t = std::stof(v);
            else if (k == "albedo") s.albedo = std::stof(v);
            else if (k == "gVal") s.g_val = std::stof(v);
            else if (k == "f_u") s.f_u = std::stof(v);
            else if (k == "speed_u") s.speed_u = std::stof(v);
            else if (k == "n_o") s.n_o = std::stof(v);
            else if (k == "n_scaling") s.n_scaling = std::stof(v);
            else if (k == "n_coeff") s.n_coeff = std::stoi(v);
            else if (k == "radius") s.radius = std::stof(v);
            else if (k == "center1") ok = parse_vec3(k, v, s.center1, err);
            else if (k == "center2") ok = parse_vec3(k, v, s.center2, err);
            else if (k == "active1") ok = parse_bool(k, v, s.active1, err);
            else if (k == "active2") ok = parse_bool(k, v, s.active2, err);
            else if (k == "phase1") s.phase1 = std::stof(v);
            else if (k == "phase2") s.phase2 = std::stof(v);
            else if (k == "chordlength") s.chordlength = std::stof(v);
            else if (k == "theta_sources") s.theta_sources = std::stoi(v);
            else if (k == "trans_z_sources") s.trans_z_sources = std::stoi(v);
            else if (k == "n_max") s.n_max = std::stof(v);
            else if (k == "n_clip") s.n_clip = std::stof(v);
            else if (k == "phi_min") s.phi_min = std::stof(v);
            else if (k == "phi_max") s.phi_max = std::stof(v);
            else if (k == "mode") s.mode = std::stoi(v);
            else if (k == "emitter_gap") s.emitter_gap = std::stof(v);
            else if (k == "sensor_gap") s.sensor_gap = std::stof(v);
            else if (k == "er_stepsize") s.er_stepsize = std::stof(v);
            else if (k == "directTol") s.direct_to_l = std::stof(v);
            else if (k == "useInitializationHack") ok = parse_bool(k, v, s.use_initialization_hack, err);
            else if (k == "rr_weight") s.rr_weight = std::stof(v);
            else if (k == "projectorTexture") s.projector_texture = v;
            else if (k == "useDirect") ok = parse_bool(k, v, s.use_direct, err);
            else if (k == "useAngularSampling") ok = parse_bool(k, v, s.use_angular_sampling, err);
            else if (k == "use_bounce_decomposition") ok = parse_bool(k, v, s.use_bounce_decomposition, err);
            else if (k == "maxDepth") s.max_depth = std::stoi(v);
            else if (k == "maxPathlength") s.max_pathlength = std::stof(v);
            else if (k == "pathLengthMin") s.path_length_min = std::stof(v);
            else if (k == "pathLengthMax") s.path_length_max = std::stof(v);
            else if (k == "pathLengthBins") s.path_length_bins = std::stoi(v);
            else if (k == "spatialX") s.spatial_x = std::stoi(v);
            else if (k == "spatialY") s.spatial_y = std::stoi(v);
            else if (k == "halfThetaLimit") s.half_theta_limit = std::stof(v);
            else if (k == "emitter_size") s.emitter_size = std::stof(v);
            else if (k == "sensor_size") s.sensor_size = std::stof(v);
            else if (k == "mediumLx") s.medium_l.x = std::stof(v);
            else if (k == "mediumRx") s.medium_r.x = std::stof(v);
            else if (k == "distribution") s.distribution = v;
            else if (k == "gOrKappa") s.g_or_kappa = std::stof(v);
            else if (k == "emitter_distance") s.emitter_distance = std::stof(v);
            else if (k == "emitter_lens_aperture") s.emitter_lens_aperture = std::stof(v);
            else if (k == "emitter_lens_focalLength") s.emitter_lens_focal_length = std::stof(v);
            else if (k == "emitter_lens_active") ok = parse_bool(k, v, s.emitter_lens_active, err);
            else if (k == "sensor_distance") s.sensor_distance = std::stof(v);
            else if (k == "sensor_lens_aperture") s.sensor_lens_aperture = std::stof(v);
            else if (k == "sensor_lens_focalLength") s.sensor_lens_focal_length = std::stof(v);
            else if (k == "sensor_lens_active") ok = parse_bool(k, v, s.sensor_lens_active, err);
            else if (k == "printInputs") ok = parse_bool(k, v, s.print_inputs, err);
            else if (k == "rifgridFile") s.rif_grid_file = v;
            else {
                err << "Unknown variable in the input argument:" << k << "\n";
                return kFail;
            }
            if (!ok) return kFail;
        }
    } catch (const std::exception &e) {  // std::stoi / std::stof on a bad number
        err << "Bad number in the input arguments: " << e.what() << "\n";
        return kFail;
    }
    if (s.stricts) {  // renderer3d_sample.cpp:650-717
        std::vector<std::string> need = kCommonKeys;
        if (s.field == Field::Sources)
            need.insert(need.end(), kSourceKeys.begin(), kSourceKeys.end());
        else
            need.insert(need.end(), kBesselKeys.begin(), kBesselKeys.end());
        if (s.field == Field::Interpolated) need.push_back("rifgridFile");
        bool missing = false;
        for (const auto &k : need)
            if (!given.count(k)) {
                err << k << " is not specified\n";
                missing = true;
            }
        if (missing) {
            err << "crashing as one or more inputs is absent\n";
            return kFail;
        }
    }
    // renderer3d_sample.cpp:788-797: the gap checks apply to given gaps only
    const float len = s.medium_r.x - s.medium_l.x;
    if (given.count("emitter_gap") && (s.emitter_gap < 0 || s.emitter_gap > len)) {
        err << "invalid gap between the emitter and the US:" << s.emitter_gap << "\n";
        return kFail;
    }
    if (given.count("sensor_gap") && (s.sensor_gap < 0 || s.sensor_gap > len)) {
        err << "invalid gap between the sensor and the US:" << s.sensor_gap << "\n";
        return kFail;
    }
    return kOk;
}

int sanity_checks(const Settings &s, std::ostream &err) {
    const float len = s.medium_r.x - s.medium_l.x;
    auto bad = [&](const char *msg) {
        err << msg << "\n";
        return kFail;
    };
    if (s.emitter_distance < 0) return bad("emitter_distance must not be negative");
    if (s.sensor_distance < 0) return bad("sensor_distance must not be negative");
    if (s.emitter_lens_active && s.emitter_distance < 1e-4f)
        return bad("emitter lens active: emitter_distance should be strictly positive (>1e-4)");
    if (s.sensor_lens_active && s.sensor_distance < 1e-4f)
        return bad("sensor lens active: sensor_distance should be strictly positive (>1e-4)");
    if (s.emitter_gap < 0 || s.emitter_gap > len) return bad("invalid gap between the emitter and the US");
    if (s.sensor_gap < 0 || s.sensor_gap > len) return bad("invalid gap between the sensor and the US");
    if (s.emitter_gap + s.sensor_gap > len) return bad("sum of sensor and emitter gaps is more than the medium size");
    if (s.field == Field::Analytical && s.phi_max < s.phi_min)
        return bad("phi_max must be greater than or equal to phi_min");
    if (s.spatial_x <= 0 || s.spatial_y <= 0 || s.path_length_bins <= 0) return bad("film resolution must be positive");
    return kOk;
}

void print_settings(std::ostream &os, const Settings &s) {
    os << "execution_parameters { rendering_type: " << s.rendering_type << ", output_file_name: "
       << s.output_file_name << ", threads: " << s.threads << ", emitter_gap: " << s.emitter_gap
       << ", sensor_gap: " << s.sensor_gap << ", use_bounce_decomposition: " << s.use_bounce_decomposition
       << ", print_inputs: " << s.print_inputs << ", projector_texture: " << s.projector_texture << " }\n";
    os << "adoc_geometry_parameters { half_theta_limit: " << s.half_theta_limit << ", emitter_size: "
       << s.emitter_size << ", sensor_size: " << s.sensor_size << ", emitter_distance: " << s.emitter_distance
       << ", sensor_distance: " << s.sensor_distance << " }\n";
    os << "scattering_parameters { sigma_t: " << s.sigma_t << ", albedo: " << s.albedo << ", g_val: " << s.g_val
       << " }\n";
    os << "rendering_parameters { num_photons: " << s.num_photons << ", max_depth: " << s.max_depth
       << ", max_pathlength: " << s.max_pathlength << ", use_direct: " << s.use_direct
       << ", use_angular_sampling: " << s.use_angular_sampling << " }\n";
    os << "scene_parameters { medium_l: [" << s.medium_l.x << ", " << s.medium_l.y << ", " << s.medium_l.z
       << "], medium_r: [" << s.medium_r.x << ", " << s.medium_r.y << ", " << s.medium_r.z << "] }\n";
    os << "film_parameters { path_length_min: " << s.path_length_min << ", path_length_max: " << s.path_length_max
       << ", path_length_bins: " << s.path_length_bins << ", spatial_x: " << s.spatial_x
       << ", spatial_y: " << s.spatial_y << " }\n";
    os << "importance_sampling_parameters { distribution: " << s.distribution << ", g_or_kappa: " << s.g_or_kappa
       << " }\n";
    os << "lens_parameters { emitter_lens_aperture: " << s.emitter_lens_aperture
       << ", emitter_lens_focal_length: " << s.emitter_lens_focal_length
       << ", emitter_lens_active: " << s.emitter_lens_active << ", sensor_lens_aperture: " << s.sensor_lens_aperture
       << ", sensor_lens_focal_length: " << s.sensor_lens_focal_length
       << ", sensor_lens_active: " << s.sensor_lens_active << " }\n";
    os << field_name(s.field) << " { f_u: " << s.f_u << ", speed_u: " << s.speed_u << ", n_o: " << s.n_o
       << ", er_stepsize: " << s.er_stepsize << ", precision: " << s.precision;
    if (s.field == Field::Sources)
        os << ", n_scaling: " << s.n_scaling << ", n_coeff: " << s.n_coeff << ", radius: " << s.radius
           << ", center1: (" << s.center1.x << ", " << s.center1.y << ", " << s.center1.z << "), center2: ("
           << s.center2.x << ", " << s.center2.y << ", " << s.center2.z << "), active1: " << s.active1
           << ", active2: " << s.active2 << ", phase1: " << s.phase1 << ", phase2: " << s.phase2
           << ", chordlength: " << s.chordlength << ", theta_sources: " << s.theta_sources
           << ", trans_z_sources: " << s.trans_z_sources;
    else
        os << ", n_max: " << s.n_max << ", n_clip: " << s.n_clip << ", phi_min: " << s.phi_min
           << ", phi_max: " << s.phi_max << ", mode: " << s.mode;
    if (s.field == Field::Interpolated) os << ", rif_grid_file: " << s.rif_grid_file;
    os << " }\n";
}

// ---------------------------------------------------------------- run
namespace {

void json_floats(std::ostream &o, const char *key, const float *v, int n) {
    o << "  \"" << key << "\": [";
    for (int i = 0; i < n; ++i) o << (i ? ", " : "") << std::setprecision(9) << v[i];
    o << "],\n";
}

void write_describe(const std::string &path, const Settings &s, const simpler_scene_desc &d,
                    const scn::Scene<tvec::TVector3> &scene, const std::string &out_file) {
    std::ofstream o(path);
    if (!o) throw std::runtime_error("cannot write " + path);
    o << "{\n";
    json_floats(o, "block_l", d.block_l, 3);
    json_floats(o, "block_r", d.block_r, 3);
    json_floats(o, "light_origin", d.light_origin, 3);
    json_floats(o, "light_dir", d.light_dir, 3);
    json_floats(o, "light_plane", d.light_plane, 2);
    json_floats(o, "emitter_lens_origin", d.emitter_lens_origin, 3);
    json_floats(o, "view_origin", d.view_origin, 3);
    json_floats(o, "view_dir", d.view_dir, 3);
    json_floats(o, "view_horizontal", d.view_horizontal, 3);
    json_floats(o, "view_plane", d.view_plane, 2);
    json_floats(o, "pathlength_range", d.pathlength_range, 2);
    json_floats(o, "sensor_lens_origin", d.sensor_lens_origin, 3);
    json_floats(o, "axis_uz", d.axis_uz, 3);
    json_floats(o, "axis_ux", d.axis_ux, 3);
    json_floats(o, "p_u", d.p_u, 3);
    const std::pair<const char *, float> scalars[] = {
        {"ior", d.ior}, {"half_theta_limit", d.half_theta_limit}, {"Li", d.Li},
        {"emitter_lens_aperture", d.emitter_lens_aperture}, {"emitter_lens_focal_length", d.emitter_lens_focal_length},
        {"sensor_lens_aperture", d.sensor_lens_aperture}, {"sensor_lens_focal_length", d.sensor_lens_focal_length},
        {"n_o", d.n_o}, {"f_u", d.f_u}, {"speed_u", d.speed_u}, {"n_max", d.n_max}, {"n_clip", d.n_clip},
        {"phi_min", d.phi_min}, {"phi_max", d.phi_max}, {"er_stepsize", d.er_stepsize},
        {"egap_end_loc_x", d.egap_end_loc_x}, {"sgap_begin_loc_x", d.sgap_begin_loc_x},
        {"time_coupling", d.time_coupling}, {"sigma_t", d.sigma_t}, {"albedo", d.albedo}, {"g", d.g},
        {"max_pathlength", d.max_pathlength}};
    for (const auto &kv : scalars) o << "  \"" << kv.first << "\": " << std::setprecision(9) << kv.second << ",\n";
    const std::pair<const char *, long long> ints[] = {
        {"emitter_lens_active", d.emitter_lens_active}, {"sensor_lens_active", d.sensor_lens_active},
        {"use_bounce_decomposition", d.use_bounce_decomposition}, {"field_kind", d.field_kind},
        {"mode", d.mode}, {"precision", d.precision}, {"max_depth", d.max_depth}, {"use_direct", d.use_direct},
        {"use_angular_sampling", d.use_angular_sampling}, {"flags", (long long)d.flags},
        {"x_res", s.spatial_x}, {"y_res", s.spatial_y}, {"z_res", s.path_length_bins},
        {"num_photons", (long long)s.num_photons}, {"threads", s.threads}, {"tex_x", scene.tex_x()},
        {"tex_y", scene.tex_y()}, {"n_sources", (long long)scene.source_amp().size()}};
    for (const auto &kv : ints) o << "  \"" << kv.first << "\": " << kv.second << ",\n";
    double tex_sum = 0.0;
    for (float t : scene.texture()) tex_sum += t;
    o << "  \"texture_sum\": " << std::setprecision(17) << tex_sum << ",\n";
    const std::vector<float> &sp = scene.source_pos();
    o << "  \"source_pos\": [";
    for (size_t i = 0; i < sp.size(); ++i) o << (i ? ", " : "") << std::setprecision(9) << sp[i];
    o << "],\n";
    const spline::VolGrid &g = scene.spline_coeffs();
    o << "  \"spline_dims\": [" << g.N[0] << ", " << g.N[1] << ", " << g.N[2] << "],\n";
    double csum = 0.0;
    for (float c : g.data) csum += c;
    o << "  \"spline_coeff_sum\": " << std::setprecision(17) << csum << ",\n";
    o << "  \"output_file\": \"" << out_file << "\"\n}\n";
}

}  // namespace

int run(const Settings &s, const std::string &out_prefix, const std::string &describe_path, std::ostream &err) {
    if (!s.have_rif) {
        err << "Failed to create scene: no " << field_name(s.field) << " parameters were read\n";
        return kFail;
    }
    if (sanity_checks(s, err) != kOk) return kFail;
    try {
        const float ior = 1.3333f;  // simplER.cpp:28, renderer3d_sample.cpp:83
        pfunc::henyey_greenstein phase(s.g_val);
        // simplER.cpp:60-85 (renderer3d_sample.cpp:808-832)
        const tvec::Vec3f emitter_lens_origin(s.medium_r.x, 0.f, 0.f);
        const float egap_end_loc_x = emitter_lens_origin.x - s.emitter_gap;
        const tvec::Vec3f sensor_lens_origin(s.medium_l.x, 0.f, 0.f);
        const float sgap_begin_loc_x = sensor_lens_origin.x + s.sensor_gap;
        const tvec::Vec3f light_origin(s.medium_r.x + s.emitter_distance, 0.f, 0.f);
        const float light_angle = (float)M_PI;
        const tvec::Vec3f light_dir(std::cos(light_angle), std::sin(light_angle), 0.f);
        const tvec::Vec2f light_plane(s.emitter_size, s.emitter_size);
        const float Li = 75000.f;
        const tvec::Vec3f view_origin(s.medium_l.x - s.sensor_distance, 0.f, 0.f);
        const tvec::Vec3f view_dir(-1.f, 0.f, 0.f);
        const tvec::Vec3f view_x(0.f, -1.f, 0.f);
        const tvec::Vec2f view_plane(s.sensor_size, s.sensor_size);
        const tvec::Vec2f pathlength_range(s.path_length_min, s.path_length_max);
        // the drivers' transducer frame (renderer3d_sample.cpp:838-840, rif_base.h:146-148)
        const tvec::Vec3f axis_uz(-1.f, 0.f, 0.f), axis_ux(0.f, 0.f, 1.f), p_u(0.f, 0.f, 0.f);
        const med::Medium medium(s.sigma_t, s.albedo, &phase);

        std::unique_ptr<scn::Scene<tvec::TVector3>> scene;
        if (s.field == Field::Sources) {
            // rif_sources.h:196-201 / renderer3d_sample.cpp:800-805
            const float theta_max = std::asin(s.chordlength / (2 * s.radius));
            const float trans_z_max = s.chordlength / 2;
            scene.reset(new scn::Scene<tvec::TVector3>(
                ior, s.medium_l, s.medium_r, light_origin, light_dir, s.half_theta_limit, s.projector_texture,
                light_plane, Li, view_origin, view_dir, view_x, view_plane, pathlength_range,
                s.use_bounce_decomposition, s.distribution, s.g_or_kappa, emitter_lens_origin,
                s.emitter_lens_aperture, s.emitter_lens_focal_length, s.emitter_lens_active, sensor_lens_origin,
                s.sensor_lens_aperture, s.sensor_lens_focal_length, s.sensor_lens_active, s.f_u, s.speed_u, s.n_o,
                s.n_scaling, s.n_coeff, s.radius, s.center1, s.center2, s.active1, s.active2, s.phase1, s.phase2,
                -theta_max, theta_max, s.theta_sources, -trans_z_max, trans_z_max, s.trans_z_sources, axis_uz,
                axis_ux, p_u, s.er_stepsize, s.direct_to_l, s.rr_weight, s.precision, egap_end_loc_x,
                sgap_begin_loc_x, s.use_initialization_hack));
        } else if (s.field == Field::Analytical) {
            scene.reset(new scn::Scene<tvec::TVector3>(
                ior, s.medium_l, s.medium_r, light_origin, light_dir, s.half_theta_limit, s.projector_texture,
                light_plane, Li, view_origin, view_dir, view_x, view_plane, pathlength_range,
                s.use_bounce_decomposition, s.distribution, s.g_or_kappa, emitter_lens_origin,
                s.emitter_lens_aperture, s.emitter_lens_focal_length, s.emitter_lens_active, sensor_lens_origin,
                s.sensor_lens_aperture, s.sensor_lens_focal_length, s.sensor_lens_active, s.f_u, s.speed_u, s.n_o,
                s.n_max, s.n_clip, s.phi_min, s.phi_max, s.mode, axis_uz, axis_ux, p_u, s.er_stepsize,
                s.direct_to_l, s.rr_weight, s.precision, egap_end_loc_x, sgap_begin_loc_x,
                s.use_initialization_hack));
        } else {
            scene.reset(new scn::Scene<tvec::TVector3>(
                ior, s.medium_l, s.medium_r, light_origin, light_dir, s.half_theta_limit, s.projector_texture,
                light_plane, Li, view_origin, view_dir, view_x, view_plane, pathlength_range,
                s.use_bounce_decomposition, s.distribution, s.g_or_kappa, emitter_lens_origin,
                s.emitter_lens_aperture, s.emitter_lens_focal_length, s.emitter_lens_active, sensor_lens_origin,
                s.sensor_lens_aperture, s.sensor_lens_focal_length, s.sensor_lens_active, s.f_u, s.speed_u, s.n_o,
                s.n_max, s.n_clip, s.phi_min, s.phi_max, s.mode, axis_uz, axis_ux, p_u, s.er_stepsize,
                s.direct_to_l, s.rr_weight, s.precision, egap_end_loc_x, sgap_begin_loc_x,
                s.use_initialization_hack, s.rif_grid_file));
        }
        photon::Renderer<tvec::TVector3> renderer(s.max_depth, s.max_pathlength, s.use_direct,
                                                  s.use_angular_sampling, s.threads);
        const std::string out_file = out_prefix + ".pfm3d";
        if (!describe_path.empty()) {
            write_describe(describe_path, s, renderer.full_desc(medium, *scene), *scene, out_file);
            return kOk;
        }
        image::SmallImage img(s.spatial_x, s.spatial_y, s.path_length_bins);
        renderer.renderImage(img, medium, *scene, s.num_photons);
        img.writePFM3D(out_file);
    } catch (const std::exception &e) {
        err << "simplER: " << e.what() << "\n";
        return kFail;
    }
    return kOk;
}

}  // namespace simpler_frontend
