// compat.cpp -- the reference renderer's C++ API on the gfx950 engine.
// See simpler_compat.h.  Host code only: every render goes through the C ABI
// of libsimpler_hip.so.
#include "simpler_compat.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <atomic>
#include <memory>
#include <mutex>
#include <sstream>

namespace {

const float kEps = 1.19209290e-07f;  // reference constants.h:38

void check(simpler_ctx *ctx, int rc, const char *what) {
    if (rc != SIMPLER_OK)
        throw std::runtime_error(std::string(what) + ": " + (ctx ? simpler_last_error(ctx) : "error"));
}

template <class V>
void put3(float dst[3], const V &v) {
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
}

}  // namespace

namespace pfunc {
Float henyey_greenstein::f(const tvec::Vec3f &in, const tvec::Vec3f &out) const {
    const double c = (double)in.x * out.x + (double)in.y * out.y + (double)in.z * out.z;
    const double g = m_g;
    return (Float)((1.0 / (4.0 * M_PI)) * (1 - g * g) / std::pow(1 + g * g - 2 * g * c, 1.5));
}
}  // namespace pfunc

namespace med {
Medium::Medium(const Float sigmaT, const Float albedo, pfunc::henyey_greenstein *phase)
    : m_sigmaT(sigmaT), m_albedo(albedo), m_sigmaS(albedo * sigmaT), m_sigmaA((1 - albedo) * sigmaT),
      m_mfp(Float(1) / sigmaT), m_phase(phase) {
    if (!phase) throw std::invalid_argument("Medium: null phase function");
    if (m_sigmaT <= kEps) {  // medium.h:33-37
        m_sigmaT = 0;
        m_mfp = 1;
        m_albedo = 0;
    }
}
}  // namespace med

namespace image {

template <typename T>
Image3<T>::Image3(int xRes, int yRes, int zRes) : m_xRes(0), m_yRes(0), m_zRes(0) {
    setSize(xRes, yRes, zRes);
}

template <typename T>
void Image3<T>::setSize(int xRes, int yRes, int zRes) {
    if (xRes <= 0 || yRes <= 0 || zRes <= 0) throw std::invalid_argument("Image3: resolution must be positive");
    m_xRes = xRes;
    m_yRes = yRes;
    m_zRes = zRes;
    m_pixels.assign((size_t)xRes * yRes * zRes, T(0));
}

template <typename T>
void Image3<T>::writePFM3D(const std::string &fileName) const {
    std::ofstream f(fileName, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + fileName);
    f << "Pf\n" << m_xRes << ' ' << m_yRes << ' ' << m_zRes << "\n" << -1 << "\n";
    std::vector<double> d(m_pixels.begin(), m_pixels.end());
    f.write(reinterpret_cast<const char *>(d.data()), (std::streamsize)(d.size() * sizeof(double)));
    if (!f) throw std::runtime_error("error writing " + fileName);
}

template <typename T>
void Image3<T>::readPFM3D(const std::string &fileName) {
    std::ifstream f(fileName, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + fileName);
    std::string magic;
    int x, y, z;
    double scale;
    f >> magic >> x >> y >> z >> scale;
    if (magic != "Pf" || x <= 0 || y <= 0 || z <= 0) throw std::runtime_error(fileName + ": not a PFM3D file");
    f.get();
    setSize(x, y, z);
    std::vector<double> d(m_pixels.size());
    f.read(reinterpret_cast<char *>(d.data()), (std::streamsize)(d.size() * sizeof(double)));
    if (!f) throw std::runtime_error(fileName + ": truncated raster");
    std::copy(d.begin(), d.end(), m_pixels.begin());
}

template class Image3<float>;
template class Image3<double>;

std::vector<float> readPFM(const std::string &fileName, int &xRes, int &yRes) {
    std::ifstream f(fileName, std::ios::binary);
    if (!f) throw std::runtime_error("Cannot open file " + fileName + ", or it does not exist");
    std::string magic;
    double scale;
    f >> magic >> xRes >> yRes >> scale;
    if (magic != "Pf") throw std::runtime_error(fileName + ": invalid magic number (grayscale \"Pf\" expected)");
    if (xRes <= 0 || yRes <= 0) throw std::runtime_error(fileName + ": bad resolution");
    f.seekg(0, std::ios::end);
    const long size = (long)f.tellg();
    const long n = (long)xRes * yRes;
    // the reference writes double rasters (image.h:203-208); standard PFM
    // holds float32; both sit at the end of the file
    const bool dbl = size >= n * 8 + 8;
    std::vector<double> raw(n);
    if (dbl) {
        f.seekg(size - n * 8);
        f.read(reinterpret_cast<char *>(raw.data()), n * 8);
    } else {
        std::vector<float> r32(n);
        f.seekg(size - n * 4);
        f.read(reinterpret_cast<char *>(r32.data()), n * 4);
        std::copy(r32.begin(), r32.end(), raw.begin());
    }
    if (!f) throw std::runtime_error(fileName + ": truncated raster");
    if (scale > 0) throw std::runtime_error(fileName + ": big-endian PFM is not supported");
    std::vector<float> out(n);
    for (int r = 0; r < yRes; ++r)  // bottom-to-top -> top-to-bottom
        for (int c = 0; c < xRes; ++c) out[(size_t)(yRes - 1 - r) * xRes + c] = (float)raw[(size_t)r * xRes + c];
    return out;
}

}  // namespace image

namespace spline {

VolGrid readVol(const std::string &fileName) {
    std::ifstream f(fileName, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open VOL file " + fileName);
    char hdr[4];
    f.read(hdr, 4);
    if (!f || hdr[0] != 'V' || hdr[1] != 'O' || hdr[2] != 'L') throw std::runtime_error(fileName + " is not a VOL file");
    if (hdr[3] != 3) throw std::runtime_error(fileName + ": VOL version must be 3");
    int32_t type, ch;
    VolGrid g;
    f.read(reinterpret_cast<char *>(&type), 4);
    if (type != 1) throw std::runtime_error(fileName + ": VOL type must be 1 (float32)");
    f.read(reinterpret_cast<char *>(g.N), 12);
    f.read(reinterpret_cast<char *>(&ch), 4);
    if (ch != 1) throw std::runtime_error(fileName + ": VOL must have one channel");
    f.read(reinterpret_cast<char *>(g.xmin), 12);
    f.read(reinterpret_cast<char *>(g.xmax), 12);
    for (int i = 0; i < 3; ++i)
        if (g.N[i] < 2 || g.N[i] > 4096) throw std::runtime_error(fileName + ": bad VOL resolution");
    const size_t n = (size_t)g.N[0] * g.N[1] * g.N[2];
    g.data.resize(n);
    f.read(reinterpret_cast<char *>(g.data.data()), (std::streamsize)(n * 4));
    if (!f) throw std::runtime_error(fileName + ": truncated VOL data");
    return g;
}

void writeVol(const std::string &fileName, const VolGrid &g) {
    std::ofstream f(fileName, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + fileName);
    const char hdr[4] = {'V', 'O', 'L', 3};
    const int32_t type = 1, ch = 1;
    f.write(hdr, 4);
    f.write(reinterpret_cast<const char *>(&type), 4);
    f.write(reinterpret_cast<const char *>(g.N), 12);
    f.write(reinterpret_cast<const char *>(&ch), 4);
    f.write(reinterpret_cast<const char *>(g.xmin), 12);
    f.write(reinterpret_cast<const char *>(g.xmax), 12);
    f.write(reinterpret_cast<const char *>(g.data.data()), (std::streamsize)(g.data.size() * 4));
}

// Mirror-boundary cubic B-spline interpolation prefilter along one line
// (Unser's recursive filter with pole z1 = sqrt(3) - 2 and gain 6; the
// causal initial value sums the mirrored sequence of period 2n-2).
static void prefilter_line(double *c, long stride, int n) {
    const double z1 = std::sqrt(3.0) - 2.0;
    if (n == 1) return;
    std::vector<double> a(n);
    for (int i = 0; i < n; ++i) a[i] = c[i * stride];
    double c0 = 0.0, zk = 1.0;
    for (int i = 0; i < n; ++i) {
        c0 += a[i] * zk;
        zk *= z1;
    }
    // zk = z1^n; continue with the mirrored part i = n-2 .. 1 (powers n .. 2n-3)
    for (int i = n - 2; i > 0; --i) {
        c0 += a[i] * zk;
        zk *= z1;
    }
    c0 /= (1.0 - std::pow(z1, 2 * n - 2));
    std::vector<double> cp(n), cn(n);
    cp[0] = c0;
    for (int i = 1; i < n; ++i) cp[i] = a[i] + z1 * cp[i - 1];
    cn[n - 1] = z1 / (z1 * z1 - 1.0) * (cp[n - 1] + z1 * cp[n - 2]);
    for (int i = n - 2; i >= 0; --i) cn[i] = z1 * (cn[i + 1] - cp[i]);
    for (int i = 0; i < n; ++i) c[i * stride] = 6.0 * cn[i];
}

std::vector<float> prefilter3d(const std::vector<float> &data, const int N[3]) {
    const long nx = N[0], ny = N[1], nz = N[2];
    std::vector<double> c(data.begin(), data.end());
    for (long k = 0; k < nz; ++k)
        for (long i = 0; i < nx; ++i) prefilter_line(&c[k * nx * ny + i], nx, (int)ny);
    for (long k = 0; k < nz; ++k)
        for (long j = 0; j < ny; ++j) prefilter_line(&c[(k * ny + j) * nx], 1, (int)nx);
    for (long j = 0; j < ny; ++j)
        for (long i = 0; i < nx; ++i) prefilter_line(&c[j * nx + i], nx * ny, (int)nz);
    return std::vector<float>(c.begin(), c.end());
}

}  // namespace spline

namespace scn {

// Content revisions are drawn from one process-wide counter, so a revision
// names the texture / source lattice / grid content of a scene uniquely (a
// copy of a scene shares its content and its revision); descriptor-only
// changes (set_f_u, set_n_max, mutable_desc) keep the revision and are found
// by comparing descriptors.
static uint64_t next_revision() {
    static std::atomic<uint64_t> counter{0};
    return ++counter;
}

template <template <typename> class V>
void Scene<V>::common(Float ior, const V<Float> &blockL, const V<Float> &blockR, const V<Float> &lightOrigin,
                      const V<Float> &lightDir, const Float &halfThetaLimit, const std::string &lightTextureFile,
                      const tvec::Vec2f &lightPlane, const Float Li, const V<Float> &viewOrigin,
                      const V<Float> &viewDir, const V<Float> &viewHorizontal, const tvec::Vec2f &viewPlane,
                      const tvec::Vec2f &pathlength_range, const bool &use_bounce_decomposition,
                      const V<Float> &emitter_lens_origin, const Float &emitter_lens_aperture,
                      const Float &emitter_lens_focalLength, const bool &emitter_lens_active,
                      const V<Float> &sensor_lens_origin, const Float &sensor_lens_aperture,
                      const Float &sensor_lens_focalLength, const bool &sensor_lens_active, const Float &f_u,
                      const Float &speed_u, const Float &n_o, const V<Float> &axis_uz, const V<Float> &axis_ux,
                      const V<Float> &p_u, const Float &er_stepsize, const int &precision,
                      const Float &EgapEndLocX, const Float &SgapBeginLocX) {
    simpler_scene_desc_default(&m_desc);
    simpler_scene_desc &d = m_desc;
    put3(d.block_l, blockL);
    put3(d.block_r, blockR);
    d.ior = ior;
    put3(d.light_origin, lightOrigin);
    put3(d.light_dir, lightDir);
    d.half_theta_limit = halfThetaLimit;
    d.light_plane[0] = lightPlane.x;
    d.light_plane[1] = lightPlane.y;
    d.Li = Li;
    put3(d.emitter_lens_origin, emitter_lens_origin);
    d.emitter_lens_aperture = emitter_lens_aperture;
    d.emitter_lens_focal_length = emitter_lens_focalLength;
    d.emitter_lens_active = emitter_lens_active;
    put3(d.view_origin, viewOrigin);
    put3(d.view_dir, viewDir);
    put3(d.view_horizontal, viewHorizontal);
    d.view_plane[0] = viewPlane.x;
    d.view_plane[1] = viewPlane.y;
    d.pathlength_range[0] = pathlength_range.x;
    d.pathlength_range[1] = pathlength_range.y;
    d.use_bounce_decomposition = use_bounce_decomposition;
    put3(d.sensor_lens_origin, sensor_lens_origin);
    d.sensor_lens_aperture = sensor_lens_aperture;
    d.sensor_lens_focal_length = sensor_lens_focalLength;
    d.sensor_lens_active = sensor_lens_active;
    d.f_u = f_u;
    d.speed_u = speed_u;
    d.n_o = n_o;
    put3(d.axis_uz, axis_uz);
    put3(d.axis_ux, axis_ux);
    put3(d.p_u, p_u);
    d.er_stepsize = er_stepsize;
    d.precision = precision;
    d.egap_end_loc_x = EgapEndLocX;
    d.sgap_begin_loc_x = SgapBeginLocX;
    d.flags = SIMPLER_FLAGS_DEFAULT;
    m_texture = image::readPFM(lightTextureFile, m_tex_x, m_tex_y);
    m_revision = next_revision();
}

template <template <typename> class V>
Scene<V>::Scene(Float ior, const V<Float> &blockL, const V<Float> &blockR, const V<Float> &lightOrigin,
                const V<Float> &lightDir, const Float &halfThetaLimit, const std::string &lightTextureFile,
                const tvec::Vec2f &lightPlane, const Float Li, const V<Float> &viewOrigin, const V<Float> &viewDir,
                const V<Float> &viewHorizontal, const tvec::Vec2f &viewPlane, const tvec::Vec2f &pathlength_range,
                const bool &use_bounce_decomposition, const std::string &, const Float &,
                const V<Float> &emitter_lens_origin, const Float &emitter_lens_aperture,
                const Float &emitter_lens_focalLength, const bool &emitter_lens_active,
                const V<Float> &sensor_lens_origin, const Float &sensor_lens_aperture,
                const Float &sensor_lens_focalLength, const bool &sensor_lens_active, const Float &f_u,
                const Float &speed_u, const Float &n_o, const Float &n_max, const Float &n_clip,
                const Float &phi_min, const Float &phi_max, const int &mode, const V<Float> &axis_uz,
                const V<Float> &axis_ux, const V<Float> &p_u, const Float &er_stepsize, const Float &,
                const Float &, const int &precision, const Float &EgapEndLocX, const Float &SgapBeginLocX,
                const bool &) {
    common(ior, blockL, blockR, lightOrigin, lightDir, halfThetaLimit, lightTextureFile, lightPlane, Li, viewOrigin,
           viewDir, viewHorizontal, viewPlane, pathlength_range, use_bounce_decomposition, emitter_lens_origin,
           emitter_lens_aperture, emitter_lens_focalLength, emitter_lens_active, sensor_lens_origin,
           sensor_lens_aperture, sensor_lens_focalLength, sensor_lens_active, f_u, speed_u, n_o, axis_uz, axis_ux,
           p_u, er_stepsize, precision, EgapEndLocX, SgapBeginLocX);
    m_desc.field_kind = SIMPLER_FIELD_BESSEL;
    m_desc.n_max = n_max;
    m_desc.n_clip = n_clip;
    m_desc.phi_min = phi_min;
    m_desc.phi_max = phi_max;
    m_desc.mode = mode;
}

template <template <typename> class V>
Scene<V>::Scene(Float ior, const V<Float> &blockL, const V<Float> &blockR, const V<Float> &lightOrigin,
                const V<Float> &lightDir, const Float &halfThetaLimit, const std::string &lightTextureFile,
                const tvec::Vec2f &lightPlane, const Float Li, const V<Float> &viewOrigin, const V<Float> &viewDir,
                const V<Float> &viewHorizontal, const tvec::Vec2f &viewPlane, const tvec::Vec2f &pathlength_range,
                const bool &use_bounce_decomposition, const std::string &distribution, const Float &gOrKappa,
                const V<Float> &emitter_lens_origin, const Float &emitter_lens_aperture,
                const Float &emitter_lens_focalLength, const bool &emitter_lens_active,
                const V<Float> &sensor_lens_origin, const Float &sensor_lens_aperture,
                const Float &sensor_lens_focalLength, const bool &sensor_lens_active, const Float &f_u,
                const Float &speed_u, const Float &n_o, const Float &n_max, const Float &n_clip,
                const Float &phi_min, const Float &phi_max, const int &mode, const V<Float> &axis_uz,
                const V<Float> &axis_ux, const V<Float> &p_u, const Float &er_stepsize, const Float &tol,
                const Float &rrWeight, const int &precision, const Float &EgapEndLocX, const Float &SgapBeginLocX,
                const bool &useInitializationHack, const std::string &rifgridFile)
    : Scene(ior, blockL, blockR, lightOrigin, lightDir, halfThetaLimit, lightTextureFile, lightPlane, Li, viewOrigin,
            viewDir, viewHorizontal, viewPlane, pathlength_range, use_bounce_decomposition, distribution, gOrKappa,
            emitter_lens_origin, emitter_lens_aperture, emitter_lens_focalLength, emitter_lens_active,
            sensor_lens_origin, sensor_lens_aperture, sensor_lens_focalLength, sensor_lens_active, f_u, speed_u,
            n_o, n_max, n_clip, phi_min, phi_max, mode, axis_uz, axis_ux, p_u, er_stepsize, tol, rrWeight,
            precision, EgapEndLocX, SgapBeginLocX, useInitializationHack) {
    // spline.h:88-162 + rif_interpolated.h:101-104: read the VOL grid of n;
    // the prefilter into B-spline coefficients runs on the GPU when the scene
    // is uploaded (simpler_scene_set_spline_samples)
    m_spline_samples = spline::readVol(rifgridFile);
    m_desc.field_kind = SIMPLER_FIELD_SPLINE;
    m_revision = next_revision();
}

template <template <typename> class V>
Scene<V>::Scene(Float ior, const V<Float> &blockL, const V<Float> &blockR, const V<Float> &lightOrigin,
                const V<Float> &lightDir, const Float &halfThetaLimit, const std::string &lightTextureFile,
                const tvec::Vec2f &lightPlane, const Float Li, const V<Float> &viewOrigin, const V<Float> &viewDir,
                const V<Float> &viewHorizontal, const tvec::Vec2f &viewPlane, const tvec::Vec2f &pathlength_range,
                const bool &use_bounce_decomposition, const std::string &, const Float &,
                const V<Float> &emitter_lens_origin, const Float &emitter_lens_aperture,
                const Float &emitter_lens_focalLength, const bool &emitter_lens_active,
                const V<Float> &sensor_lens_origin, const Float &sensor_lens_aperture,
                const Float &sensor_lens_focalLength, const bool &sensor_lens_active, const Float &f_u,
                const Float &speed_u, const Float &n_o, const Float &n_scaling, const int &n_coeff,
                const Float &radius, const V<Float> &center1, const V<Float> &center2, const bool &active1,
                const bool &active2, const Float &phase1, const Float &phase2, const Float &theta_min,
                const Float &theta_max, const int &theta_sources, const Float &trans_z_min,
                const Float &trans_z_max, const int &trans_z_sources, const V<Float> &axis_uz,
                const V<Float> &axis_ux, const V<Float> &p_u, const Float &er_stepsize, const Float &,
                const Float &, const int &precision, const Float &EgapEndLocX, const Float &SgapBeginLocX,
                const bool &) {
    common(ior, blockL, blockR, lightOrigin, lightDir, halfThetaLimit, lightTextureFile, lightPlane, Li, viewOrigin,
           viewDir, viewHorizontal, viewPlane, pathlength_range, use_bounce_decomposition, emitter_lens_origin,
           emitter_lens_aperture, emitter_lens_focalLength, emitter_lens_active, sensor_lens_origin,
           sensor_lens_aperture, sensor_lens_focalLength, sensor_lens_active, f_u, speed_u, n_o, axis_uz, axis_ux,
           p_u, er_stepsize, precision, EgapEndLocX, SgapBeginLocX);
    m_desc.field_kind = SIMPLER_FIELD_SOURCES;
    m_desc.time_coupling = 0.0f;  // the reference's sources field ignores the time phase
    m_n_scaling = n_scaling;
    m_n_coeff = n_coeff;
    m_radius = radius;
    m_center1 = tvec::Vec3f(center1.x, center1.y, center1.z);
    m_center2 = tvec::Vec3f(center2.x, center2.y, center2.z);
    m_active1 = active1;
    m_active2 = active2;
    m_phase1 = phase1;
    m_phase2 = phase2;
    m_theta_min = theta_min;
    m_theta_max = theta_max;
    m_theta_sources = theta_sources;
    m_trans_z_min = trans_z_min;
    m_trans_z_max = trans_z_max;
    m_trans_z_sources = trans_z_sources;
    if (n_coeff < 1 || n_coeff > 8) throw std::invalid_argument("Scene: n_coeff must be in [1, 8]");
    if (theta_sources < 1 || trans_z_sources < 1) throw std::invalid_argument("Scene: source counts must be >= 1");
    if (!active1 && !active2) throw std::invalid_argument("Scene: no active transducer");
    rebuild_sources();
}

// rif_sources.h:156-198 (arc lattice, index i + j*theta_sources) and
// scene.cpp:140-160 (harmonic weights 4(-1)^(c+1)/((c pi)^2 nsources))
template <template <typename> class V>
void Scene<V>::rebuild_sources() {
    const int nsrc = m_theta_sources * m_trans_z_sources;
    const double th_d = m_theta_sources > 1 ? (m_theta_max - m_theta_min) / (double)(m_theta_sources - 1) : 0.0;
    const double tz_d = m_trans_z_sources > 1 ? (m_trans_z_max - m_trans_z_min) / (double)(m_trans_z_sources - 1) : 0.0;
    m_src_pos.clear();
    m_src_amp.clear();
    m_src_phase.clear();
    for (int grp = 0; grp < 2; ++grp) {
        const bool active = grp == 0 ? m_active1 : m_active2;
        if (!active) continue;
        const tvec::Vec3f &cen = grp == 0 ? m_center1 : m_center2;
        std::vector<float> pos((size_t)3 * nsrc);
        for (int i = 0; i < m_theta_sources; ++i) {
            const double theta = m_theta_min + th_d * i;
            const double xval = m_radius * (1 - std::cos(theta));
            const double yval = m_radius * std::sin(theta);
            for (int j = 0; j < m_trans_z_sources; ++j) {
                const int idx = i + j * m_theta_sources;
                const double px = m_trans_z_min + tz_d * j + cen.x;
                pos[3 * idx] = (float)px;
                if (grp == 0) {
                    pos[3 * idx + 1] = (float)(yval + cen.y);
                    pos[3 * idx + 2] = (float)(xval + cen.z);
                } else {
                    pos[3 * idx + 1] = (float)(xval + cen.y);
                    pos[3 * idx + 2] = (float)(yval + cen.z);
                }
            }
        }
        m_src_pos.insert(m_src_pos.end(), pos.begin(), pos.end());
        m_src_amp.insert(m_src_amp.end(), nsrc, 1.0f);
        m_src_phase.insert(m_src_phase.end(), nsrc, grp == 0 ? m_phase1 : m_phase2);
    }
    m_src_coeff.clear();
    for (int c = 1; c <= m_n_coeff; ++c)
        m_src_coeff.push_back((float)(m_n_scaling * 4.0 * ((c + 1) % 2 == 0 ? 1.0 : -1.0) /
                                      (std::pow(c * M_PI, 2) * nsrc)));
    m_revision = next_revision();
}

template <template <typename> class V>
void Scene<V>::set_f_u(Float value) {
    m_desc.f_u = value;
}
template <template <typename> class V>
void Scene<V>::set_n_max(Float value) {
    m_desc.n_max = value;
}
template <template <typename> class V>
void Scene<V>::set_n_scaling(Float value) {
    m_n_scaling = value;
    if (m_desc.field_kind == SIMPLER_FIELD_SOURCES) rebuild_sources();
}
template <template <typename> class V>
void Scene<V>::set_phase1(Float value) {
    m_phase1 = value;
    if (m_desc.field_kind == SIMPLER_FIELD_SOURCES) rebuild_sources();
}
template <template <typename> class V>
void Scene<V>::set_phase2(Float value) {
    m_phase2 = value;
    if (m_desc.field_kind == SIMPLER_FIELD_SOURCES) rebuild_sources();
}
// scene.h:141-164: transmission into the medium of the source direction,
// ior^2 (1 - F), when Fresnel weighting is on; 1 otherwise (the reference's
// CMake build compiles Fresnel out, which SIMPLER_FLAG_USE_FRESNEL undoes).
template <template <typename> class V>
Float Scene<V>::getFresnelTrans() const {
    const simpler_scene_desc &d = m_desc;
    if (!(d.flags & SIMPLER_FLAG_USE_FRESNEL) || !(d.ior > 1.0f)) return 1.0f;
    const float ry = d.light_dir[1] / d.ior, rz = d.light_dir[2] / d.ior;
    float rx = std::sqrt(1.0f - (ry * ry + rz * rz));
    if (d.light_dir[0] < 0.0f) rx = -rx;
    const float eta = d.ior, cI = d.light_dir[0];
    float F = 0.0f;
    if (!(std::fabs(eta - 1.0f) < 1.19209290e-7f * std::max(1.0f, std::fabs(eta)))) {  // util.h:42-56
        const float Rs = (cI - eta * rx) / (cI + eta * rx);
        const float Rp = (rx - eta * cI) / (rx + eta * cI);
        F = 0.5f * (Rs * Rs + Rp * Rp);
    }
    return eta * eta * (1.0f - F);
}

template <template <typename> class V>
const spline::VolGrid &Scene<V>::spline_coeffs() const {
    if (m_spline_coeffs.data.size() != m_spline_samples.data.size()) {
        m_spline_coeffs = m_spline_samples;
        m_spline_coeffs.data = spline::prefilter3d(m_spline_samples.data, m_spline_samples.N);
    }
    return m_spline_coeffs;
}

template class Scene<tvec::TVector3>;

}  // namespace scn

namespace photon {

simpler_ctx *engine_context() {
    static std::mutex mu;
    static simpler_ctx *ctx = nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ctx) {
        const char *dev = std::getenv("SIMPLER_DEVICE");
        const int rc = simpler_ctx_create(dev ? std::atoi(dev) : 0, &ctx);
        if (rc != SIMPLER_OK) throw std::runtime_error("no usable GPU for the photon engine (simpler_ctx_create)");
    }
    return ctx;
}

// The device copy of the last scene a renderer rendered: a render of the
// same content (same revision) reuses it, with at most a descriptor update
// (a cheap call unless the Bessel table must be rebuilt), so a parameter
// sweep through the mutators does not re-upload the texture CDF, the source
// lattice or the grid on every renderImage.
struct DeviceSceneCache {
    std::mutex mu;
    simpler_scene *sc = nullptr;
    uint64_t revision = 0;
    simpler_scene_desc desc{};
    ~DeviceSceneCache() {
        if (sc) simpler_scene_destroy(sc);
    }
};

template <template <typename> class V>
Renderer<V>::Renderer(const int maxDepth, const Float maxPathlength, const bool useDirect,
                      const bool useAngularSampling, const int64_t threads)
    : m_maxDepth(maxDepth), m_maxPathlength(maxPathlength), m_useDirect(useDirect),
      m_useAngularSampling(useAngularSampling), m_threads(threads),
      m_cache(std::make_shared<DeviceSceneCache>()) {}

template <template <typename> class V>
simpler_scene_desc Renderer<V>::full_desc(const med::Medium &medium, const scn::Scene<V> &scene) const {
    simpler_scene_desc d = scene.desc();
    d.sigma_t = medium.getSigmaT();
    d.albedo = medium.getAlbedo();
    d.g = medium.getPhaseFunction()->getG();
    d.max_depth = m_maxDepth;
    d.max_pathlength = m_maxPathlength;
    d.use_direct = m_useDirect;
    d.use_angular_sampling = m_useAngularSampling;
    return d;
}

template <template <typename> class V>
void Renderer<V>::renderChunk(image::SmallImage &img0, const med::Medium &medium, const scn::Scene<V> &scene,
                              int64 offset, int64 numPhotons, int64 nTotal, bool accumulate,
                              simpler_stats *stats) const {
    simpler_ctx *ctx = engine_context();
    const simpler_scene_desc d = full_desc(medium, scene);
    DeviceSceneCache &cache = *m_cache;
    std::lock_guard<std::mutex> lock(cache.mu);
    if (cache.sc && cache.revision == scene.revision() && cache.desc.field_kind == d.field_kind) {
        if (std::memcmp(&cache.desc, &d, sizeof(d)) != 0) {
            check(ctx, simpler_scene_update(cache.sc, &d), "scene update");
            cache.desc = d;
        }
    } else {
        if (cache.sc) simpler_scene_destroy(cache.sc);
        cache.sc = nullptr;
        simpler_scene *fresh = nullptr;
        check(ctx, simpler_scene_create(ctx, &d, scene.texture().data(), scene.tex_x(), scene.tex_y(), &fresh),
              "scene upload");
        struct Guard {
            simpler_scene *s;
            ~Guard() {
                if (s) simpler_scene_destroy(s);
            }
        } guard{fresh};
        if (d.field_kind == SIMPLER_FIELD_SOURCES)
            check(ctx,
                  simpler_scene_set_sources(fresh, (int)scene.source_amp().size(), scene.source_pos().data(),
                                            scene.source_amp().data(), scene.source_phase().data(),
                                            (int)scene.source_coeff().size(), scene.source_coeff().data()),
                  "sources upload");
        if (d.field_kind == SIMPLER_FIELD_SPLINE) {
            const spline::VolGrid &g = scene.spline_samples();
            check(ctx, simpler_scene_set_spline_samples(fresh, g.data.data(), 0, g.N, g.xmin, g.xmax),
                  "spline upload");
        }
        guard.s = nullptr;
        cache.sc = fresh;
        cache.revision = scene.revision();
        cache.desc = d;
    }
    simpler_scene *sc = cache.sc;
    check(ctx,
          simpler_render_host(sc, numPhotons, nTotal, m_seed, (uint64_t)offset, img0.data(), img0.get_x_res(),
                              img0.get_y_res(), img0.getZRes(), accumulate ? 1 : 0, stats),
          "render");
}

template <template <typename> class V>
void Renderer<V>::renderImage(image::SmallImage &img0, const med::Medium &medium, const scn::Scene<V> &scene,
                              const int64 numPhotons) const {
    if (numPhotons <= 0) {
        img0.zero();
        return;
    }
    renderChunk(img0, medium, scene, 0, numPhotons, numPhotons, false, nullptr);
}

template <template <typename> class V>
void Renderer<V>::renderImage(image::Image3<double> &img0, const med::Medium &medium, const scn::Scene<V> &scene,
                              const int64 numPhotons) const {
    image::SmallImage tmp(img0.get_x_res(), img0.get_y_res(), img0.getZRes());
    renderImage(tmp, medium, scene, numPhotons);
    std::copy(tmp.data(), tmp.data() + (size_t)tmp.get_x_res() * tmp.get_y_res() * tmp.getZRes(), img0.data());
}

template class Renderer<tvec::TVector3>;

}  // namespace photon
