// simpler_compat.h -- the reference renderer's C++ scene / render API
// (namespaces tvec, pfunc, med, image, scn, photon) implemented on the
// gfx950 engine's C ABI (include/simpler.h).
//
// A C++ caller of the reference path
//   scn::Scene<tvec::TVector3> scene(...);           (reference scene.h:47-114)
//   photon::Renderer<tvec::TVector3> r(maxDepth, maxPathlength, useDirect,
//                                      useAngularSampling, threads);
//                                                    (reference photon.h:27-32)
//   r.renderImage(img, medium, scene, numPhotons);   (reference photon.h:82-84)
// compiles unchanged against these headers; every photon is traced on the GPU.
// Errors throw std::runtime_error instead of exit() (reference image.h:231-244,
// spline.h:123-135).
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "simpler.h"

typedef float Float;  // reference constants.h:21 (USE_DOUBLE_PRECISION 0)
typedef long int64;   // reference constants.h:16

namespace tvec {

template <typename T>
struct TVector2 {
    T x, y;
    TVector2() : x(0), y(0) {}
    explicit TVector2(T v) : x(v), y(v) {}
    TVector2(T a, T b) : x(a), y(b) {}
    T operator[](int i) const { return i == 0 ? x : y; }
    T &operator[](int i) { return i == 0 ? x : y; }
};

template <typename T>
struct TVector3 {
    static const int dim = 3;
    T x, y, z;
    TVector3() : x(0), y(0), z(0) {}
    explicit TVector3(T v) : x(v), y(v), z(v) {}
    TVector3(T a, T b, T c) : x(a), y(b), z(c) {}
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    T index(int i) const { return (*this)[i]; }  // tvector_pybind.cpp:189
};

typedef TVector2<Float> Vec2f;
typedef TVector3<Float> Vec3f;
typedef TVector3<int> Vec3i;

}  // namespace tvec

namespace pfunc {

// reference phase.h / phase.cpp:16-76 (the kernels implement f and sample)
class henyey_greenstein {
public:
    explicit henyey_greenstein(const Float g) : m_g(g) {}
    Float getG() const { return m_g; }
    Float f(const tvec::Vec3f &in, const tvec::Vec3f &out) const;

private:
    Float m_g;
};

}  // namespace pfunc

namespace med {

// reference medium.h:22-73
struct Medium {
    Medium(const Float sigmaT, const Float albedo, pfunc::henyey_greenstein *phase);
    Float getSigmaT() const { return m_sigmaT; }
    Float getSigmaS() const { return m_sigmaS; }
    Float getSigmaA() const { return m_sigmaA; }
    Float getMfp() const { return m_mfp; }
    Float getAlbedo() const { return m_albedo; }
    const pfunc::henyey_greenstein *getPhaseFunction() const { return m_phase; }
    virtual ~Medium() {}

protected:
    Float m_sigmaT, m_albedo, m_sigmaS, m_sigmaA, m_mfp;
    pfunc::henyey_greenstein *m_phase;
};

}  // namespace med

namespace image {

// reference image.h:263-457 (storage [z][y][x], addEnergy index
// z*X*Y + y*X + x).  get_pixel / setPixel use the same index (the
// reference's y*yRes there, image.h:311,317, differs only for X != Y).
template <typename T>
class Image3 {
public:
    Image3() : m_xRes(0), m_yRes(0), m_zRes(0) {}
    Image3(int xRes, int yRes, int zRes);
    void setSize(int xRes, int yRes, int zRes);
    void zero() { std::fill(m_pixels.begin(), m_pixels.end(), T(0)); }
    T get_pixel(int x, int y, int z) const { return m_pixels.at(idx(x, y, z)); }
    void setPixel(int x, int y, int z, T val) { m_pixels.at(idx(x, y, z)) = val; }
    void addEnergy(int x, int y, int z, T val) { m_pixels.at(idx(x, y, z)) += val; }
    int get_x_res() const { return m_xRes; }
    int get_y_res() const { return m_yRes; }
    int getZRes() const { return m_zRes; }
    T *data() { return m_pixels.data(); }
    const T *data() const { return m_pixels.data(); }
    // reference image.cpp:161-206: "Pf\n X Y Z\n-1\n" + doubles, z-major
    void writePFM3D(const std::string &fileName) const;
    void readPFM3D(const std::string &fileName);

private:
    size_t idx(int x, int y, int z) const {
        if (x < 0 || x >= m_xRes || y < 0 || y >= m_yRes || z < 0 || z >= m_zRes)
            throw std::out_of_range("Image3 pixel index out of range");
        return ((size_t)z * m_yRes + y) * m_xRes + x;
    }
    int m_xRes, m_yRes, m_zRes;
    std::vector<T> m_pixels;
};

typedef Image3<Float> SmallImage;

// reference image.h:183-245: grayscale PFM ("Pf"), stored bottom-to-top; the
// reader returns rows top-to-bottom.  Reads float or double rasters.
std::vector<float> readPFM(const std::string &fileName, int &xRes, int &yRes);

}  // namespace image

namespace spline {
// VOL grid ("VOL", version 3, type 1 = float32, 1 channel; reference
// spline.h:115-162) and the mirror-boundary cubic B-spline prefilter
// (spline.h:686-764), in double.
struct VolGrid {
    int N[3];
    float xmin[3], xmax[3];
    std::vector<float> data;  // x fastest
};
VolGrid readVol(const std::string &fileName);
void writeVol(const std::string &fileName, const VolGrid &g);
std::vector<float> prefilter3d(const std::vector<float> &data, const int N[3]);
}  // namespace spline

namespace scn {

// reference scene.h:40-397.  The refractive-index field is chosen by the
// constructor used: Bessel (8 field arguments), point sources (18 field
// arguments) or Bessel arguments + a VOL grid file (tricubic spline).
template <template <typename> class vector_type>
class Scene {
public:
    // Bessel field (USE_RIF_SOURCES 0 build)
    Scene(Float ior, const vector_type<Float> &blockL, const vector_type<Float> &blockR,
          const vector_type<Float> &lightOrigin, const vector_type<Float> &lightDir, const Float &halfThetaLimit,
          const std::string &lightTextureFile, const tvec::Vec2f &lightPlane, const Float Li,
          const vector_type<Float> &viewOrigin, const vector_type<Float> &viewDir,
          const vector_type<Float> &viewHorizontal, const tvec::Vec2f &viewPlane,
          const tvec::Vec2f &pathlength_range, const bool &use_bounce_decomposition, const std::string &distribution,
          const Float &gOrKappa, const vector_type<Float> &emitter_lens_origin, const Float &emitter_lens_aperture,
          const Float &emitter_lens_focalLength, const bool &emitter_lens_active,
          const vector_type<Float> &sensor_lens_origin, const Float &sensor_lens_aperture,
          const Float &sensor_lens_focalLength, const bool &sensor_lens_active, const Float &f_u,
          const Float &speed_u, const Float &n_o, const Float &n_max, const Float &n_clip, const Float &phi_min,
          const Float &phi_max, const int &mode, const vector_type<Float> &axis_uz,
          const vector_type<Float> &axis_ux, const vector_type<Float> &p_u, const Float &er_stepsize,
          const Float &tol, const Float &rrWeight, const int &precision, const Float &EgapEndLocX,
          const Float &SgapBeginLocX, const bool &useInitializationHack);

    // Bessel arguments + VOL grid of n (USE_RIF_INTERPOLATED build)
    Scene(Float ior, const vector_type<Float> &blockL, const vector_type<Float> &blockR,
          const vector_type<Float> &lightOrigin, const vector_type<Float> &lightDir, const Float &halfThetaLimit,
          const std::string &lightTextureFile, const tvec::Vec2f &lightPlane, const Float Li,
          const vector_type<Float> &viewOrigin, const vector_type<Float> &viewDir,
          const vector_type<Float> &viewHorizontal, const tvec::Vec2f &viewPlane,
          const tvec::Vec2f &pathlength_range, const bool &use_bounce_decomposition, const std::string &distribution,
          const Float &gOrKappa, const vector_type<Float> &emitter_lens_origin, const Float &emitter_lens_aperture,
          const Float &emitter_lens_focalLength, const bool &emitter_lens_active,
          const vector_type<Float> &sensor_lens_origin, const Float &sensor_lens_aperture,
          const Float &sensor_lens_focalLength, const bool &sensor_lens_active, const Float &f_u,
          const Float &speed_u, const Float &n_o, const Float &n_max, const Float &n_clip, const Float &phi_min,
          const Float &phi_max, const int &mode, const vector_type<Float> &axis_uz,
          const vector_type<Float> &axis_ux, const vector_type<Float> &p_u, const Float &er_stepsize,
          const Float &tol, const Float &rrWeight, const int &precision, const Float &EgapEndLocX,
          const Float &SgapBeginLocX, const bool &useInitializationHack, const std::string &rifgridFile);

    // point-source field (USE_RIF_SOURCES 1 build)
    Scene(Float ior, const vector_type<Float> &blockL, const vector_type<Float> &blockR,
          const vector_type<Float> &lightOrigin, const vector_type<Float> &lightDir, const Float &halfThetaLimit,
          const std::string &lightTextureFile, const tvec::Vec2f &lightPlane, const Float Li,
          const vector_type<Float> &viewOrigin, const vector_type<Float> &viewDir,
          const vector_type<Float> &viewHorizontal, const tvec::Vec2f &viewPlane,
          const tvec::Vec2f &pathlength_range, const bool &use_bounce_decomposition, const std::string &distribution,
          const Float &gOrKappa, const vector_type<Float> &emitter_lens_origin, const Float &emitter_lens_aperture,
          const Float &emitter_lens_focalLength, const bool &emitter_lens_active,
          const vector_type<Float> &sensor_lens_origin, const Float &sensor_lens_aperture,
          const Float &sensor_lens_focalLength, const bool &sensor_lens_active, const Float &f_u,
          const Float &speed_u, const Float &n_o, const Float &n_scaling, const int &n_coeff, const Float &radius,
          const vector_type<Float> &center1, const vector_type<Float> &center2, const bool &active1,
          const bool &active2, const Float &phase1, const Float &phase2, const Float &theta_min,
          const Float &theta_max, const int &theta_sources, const Float &trans_z_min, const Float &trans_z_max,
          const int &trans_z_sources, const vector_type<Float> &axis_uz, const vector_type<Float> &axis_ux,
          const vector_type<Float> &p_u, const Float &er_stepsize, const Float &tol, const Float &rrWeight,
          const int &precision, const Float &EgapEndLocX, const Float &SgapBeginLocX,
          const bool &useInitializationHack);

    // mutators (scene.h:179-219)
    void set_f_u(Float value);
    void set_n_max(Float value);
    void set_n_scaling(Float value);
    void set_phase1(Float value);
    void set_phase2(Float value);

    Float get_medium_ior() const { return m_desc.ior; }
    Float getFresnelTrans() const;
    const simpler_scene_desc &desc() const { return m_desc; }
    simpler_scene_desc &mutable_desc() { return m_desc; }
    const std::vector<float> &texture() const { return m_texture; }
    int tex_x() const { return m_tex_x; }
    int tex_y() const { return m_tex_y; }
    // point-source lattice (rif_sources.h:156-198) in the engine's layout
    const std::vector<float> &source_pos() const { return m_src_pos; }
    const std::vector<float> &source_amp() const { return m_src_amp; }
    const std::vector<float> &source_phase() const { return m_src_phase; }
    const std::vector<float> &source_coeff() const { return m_src_coeff; }
    // the VOL grid's node values, and its B-spline coefficients (computed on
    // the host on first use; renders prefilter on the GPU instead)
    const spline::VolGrid &spline_samples() const { return m_spline_samples; }
    const spline::VolGrid &spline_coeffs() const;
    uint64_t revision() const { return m_revision; }

private:
    void common(Float ior, const vector_type<Float> &blockL, const vector_type<Float> &blockR,
                const vector_type<Float> &lightOrigin, const vector_type<Float> &lightDir, const Float &halfThetaLimit,
                const std::string &lightTextureFile, const tvec::Vec2f &lightPlane, const Float Li,
                const vector_type<Float> &viewOrigin, const vector_type<Float> &viewDir,
                const vector_type<Float> &viewHorizontal, const tvec::Vec2f &viewPlane,
                const tvec::Vec2f &pathlength_range, const bool &use_bounce_decomposition,
                const vector_type<Float> &emitter_lens_origin, const Float &emitter_lens_aperture,
                const Float &emitter_lens_focalLength, const bool &emitter_lens_active,
                const vector_type<Float> &sensor_lens_origin, const Float &sensor_lens_aperture,
                const Float &sensor_lens_focalLength, const bool &sensor_lens_active, const Float &f_u,
                const Float &speed_u, const Float &n_o, const vector_type<Float> &axis_uz,
                const vector_type<Float> &axis_ux, const vector_type<Float> &p_u, const Float &er_stepsize,
                const int &precision, const Float &EgapEndLocX, const Float &SgapBeginLocX);
    void rebuild_sources();

    simpler_scene_desc m_desc;
    std::vector<float> m_texture;
    int m_tex_x = 1, m_tex_y = 1;
    // point-source parameters kept for the mutators
    Float m_n_scaling = 0, m_radius = 0, m_phase1 = 0, m_phase2 = 0;
    Float m_theta_min = 0, m_theta_max = 0, m_trans_z_min = 0, m_trans_z_max = 0;
    int m_n_coeff = 0, m_theta_sources = 0, m_trans_z_sources = 0;
    bool m_active1 = false, m_active2 = false;
    tvec::Vec3f m_center1, m_center2;
    std::vector<float> m_src_pos, m_src_amp, m_src_phase, m_src_coeff;
    spline::VolGrid m_spline_samples;
    mutable spline::VolGrid m_spline_coeffs;
    uint64_t m_revision = 0;
};

}  // namespace scn

namespace photon {

struct DeviceSceneCache;

// reference photon.h:24-116 (forward path only)
template <template <typename> class vector_type>
class Renderer {
public:
    Renderer(const int maxDepth, const Float maxPathlength, const bool useDirect, const bool useAngularSampling,
             const int64_t threads);

    // reference photon.cpp:266-340: img0 is overwritten with the render of
    // numPhotons photons.  Synchronous, like the reference backends.
    void renderImage(image::SmallImage &img0, const med::Medium &medium, const scn::Scene<vector_type> &scene,
                     const int64 numPhotons) const;
    // the same into a double image (the reference pybind build's SmallImage)
    void renderImage(image::Image3<double> &img0, const med::Medium &medium, const scn::Scene<vector_type> &scene,
                     const int64 numPhotons) const;

    // extensions: seed of the counter-based random streams (default 0) and
    // chunked rendering (photons [offset, offset+n) of an n_total render)
    void setSeed(uint64_t seed) { m_seed = seed; }
    uint64_t getSeed() const { return m_seed; }
    void renderChunk(image::SmallImage &img0, const med::Medium &medium, const scn::Scene<vector_type> &scene,
                     int64 offset, int64 numPhotons, int64 nTotal, bool accumulate, simpler_stats *stats) const;

    int getMaxDepth() const { return m_maxDepth; }
    Float getMaxPathlength() const { return m_maxPathlength; }
    bool getUseDirect() const { return m_useDirect; }

    // the engine descriptor a render of (medium, scene) uploads: the scene's
    // descriptor with the medium's and this renderer's settings filled in
    simpler_scene_desc full_desc(const med::Medium &medium, const scn::Scene<vector_type> &scene) const;

private:
    int m_maxDepth;
    Float m_maxPathlength;
    bool m_useDirect, m_useAngularSampling;
    int64_t m_threads;
    uint64_t m_seed = 0;
    std::shared_ptr<DeviceSceneCache> m_cache;  // device copy of the last scene rendered
};

// the process-wide engine context (device SIMPLER_DEVICE, default 0)
simpler_ctx *engine_context();

}  // namespace photon
