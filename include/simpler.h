/*
 * simpler.h -- C ABI of the MI355X (gfx950) photon-tracing engine.
 *
 * This is the drop-in boundary for the reference renderer's hot path:
 *   photon::Renderer<tvec::TVector3>::renderImage(img, medium, scene, N)
 *   (reference: renderer/include/photon.h:82-84, renderer/src/photon.cpp:266-340)
 * and its CUDA backend cuda::CudaRenderer::renderImage
 *   (reference: renderer/include/cuda_renderer.h:40, renderer/src/cuda_renderer.cu:710-793).
 *
 * Plain C: POD structs, pointers and sizes only.  Every entry point returns an
 * int status (SIMPLER_OK == 0) and never calls exit(); the text of the last
 * failure on a context is returned by simpler_last_error().  All device
 * pointers are HIP device pointers (hipMalloc or torch CUDA tensors); host
 * pointers are marked _host.
 *
 * Units: metres for every length (the reference drivers run in metres and its
 * absolute epsilons -- 1.19209290e-7, the 1e-4 NEE skip, LargeDist = 10000 --
 * are only meaningful in metres).
 */
#ifndef SIMPLER_H_
#define SIMPLER_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SIMPLER_ABI_VERSION 1

/* ---- status codes ---------------------------------------------------- */
#define SIMPLER_OK 0
#define SIMPLER_ERR_INVALID 1   /* bad argument (null, size, range)       */
#define SIMPLER_ERR_HIP 2       /* a HIP runtime call failed              */
#define SIMPLER_ERR_STATE 3     /* object in the wrong state              */
#define SIMPLER_ERR_NOMEM 4     /* device or host allocation failed       */

/* ---- refractive-index field kinds (reference: rif_analytical.h,
 *      rif_sources.h, rif_interpolated.h; scene.cpp:140-320) ----------- */
#define SIMPLER_FIELD_BESSEL 0  /* n0 + n_max*s*J_m(k r)cos(m phi)        */
#define SIMPLER_FIELD_SOURCES 1 /* n0 + sum point-source / travelling wave */
#define SIMPLER_FIELD_SPLINE 2  /* tricubic B-spline of absolute n         */

/* ---- behaviour flags (defaults reproduce the reference CMake build) --- */
/* Direct tracing adds its optical length to the photon's running path
 * length before scattering starts (photon.cpp:40-49,80,331-333). */
#define SIMPLER_FLAG_DIRECT_LEAK (1u << 0)
/* Fresnel weights on entry / exit (compiled out in the reference build:
 * scene.h:157-159, scene.cpp:910-915, photon.cpp:62-67). */
#define SIMPLER_FLAG_USE_FRESNEL (1u << 1)
/* Gradient formula of the Bessel field for a general transducer axis; when
 * clear the reference's axis=+-x component formula (scene.cpp:248-251) is
 * used. Both agree exactly for axis_uz = +-x. */
#define SIMPLER_FLAG_GENERAL_AXIS (1u << 2)
/* Disable the dead-path cutoff: by default a march stops as soon as its
 * optical length exceeds pathlength_range.y plus a geometric margin, since
 * nothing it does afterwards can reach the sensor's time gate
 * (scene.cpp:793-794).  Each such march draws from its own random
 * sub-stream, so the image is the same either way; only work is saved. */
#define SIMPLER_FLAG_NO_DEAD_CUT (1u << 3)
#define SIMPLER_FLAGS_DEFAULT (SIMPLER_FLAG_DIRECT_LEAK | SIMPLER_FLAG_GENERAL_AXIS)

/*
 * Scene descriptor: everything scn::Scene, med::Medium, pfunc::henyey_greenstein
 * and photon::Renderer hold for the forward path (scene.h:47-168; medium.h:24-38;
 * phase.h; photon.h:27-32), in one fixed-layout POD.
 */
typedef struct simpler_scene_desc {
    /* medium block (scn::block, block.cpp:11-23) */
    float block_l[3];
    float block_r[3];
    float ior;                      /* bsdf ior2 (ior1 = 1), scene.h:125    */

    /* textured area source (area_textured_source, scene.cpp:107-137) */
    float light_origin[3];
    float light_dir[3];
    float half_theta_limit;         /* 0 = collimated; > 0 = cone (diffuse) */
    float light_plane[2];
    float Li;
    float emitter_lens_origin[3];
    float emitter_lens_aperture;
    float emitter_lens_focal_length;
    int32_t emitter_lens_active;

    /* orthographic camera (camera.cpp:10-75) */
    float view_origin[3];
    float view_dir[3];
    float view_horizontal[3];
    float view_plane[2];
    float pathlength_range[2];      /* (-1,-1) = no time gating            */
    int32_t use_bounce_decomposition;
    float sensor_lens_origin[3];
    float sensor_lens_aperture;
    float sensor_lens_focal_length;
    int32_t sensor_lens_active;

    /* refractive-index field (US / rif, scene.h:76-113) */
    int32_t field_kind;             /* SIMPLER_FIELD_*                      */
    float n_o;
    float f_u;                      /* ultrasound frequency (Hz)            */
    float speed_u;                  /* ultrasound speed (m/s)               */
    float n_max;                    /* Bessel amplitude                     */
    float n_clip;                   /* Bessel clip; scaling clamp n_clip/n_max */
    float phi_min;                  /* duty window start (rad)              */
    float phi_max;                  /* duty window end (rad)                */
    int32_t mode;                   /* Bessel order m (0..8)                */
    float axis_uz[3];               /* transducer axis                      */
    float axis_ux[3];               /* azimuth reference axis               */
    float p_u[3];                   /* a point on the axis                  */
    float er_stepsize;              /* ER step h (m)                        */
    int32_t precision;              /* decimal refinement levels            */
    float egap_end_loc_x;           /* n = n_o for x > egap_end_loc_x       */
    float sgap_begin_loc_x;         /* n = n_o for x < sgap_begin_loc_x     */
    /* point-source field: per-source data is set with
     * simpler_scene_set_sources(); time_coupling multiplies the per-photon
     * duty phase tau in cos(c k r_i + phi_i - time_coupling*tau). */
    float time_coupling;

    /* medium (med::Medium) and phase function (HG) */
    float sigma_t;
    float albedo;
    float g;

    /* renderer (photon::Renderer) */
    int32_t max_depth;              /* < 0: unlimited                        */
    float max_pathlength;           /* < 0: unlimited                        */
    int32_t use_direct;
    int32_t use_angular_sampling;
    uint32_t flags;                 /* SIMPLER_FLAG_*                       */
} simpler_scene_desc;

/* Event counters of one render (all exact 64-bit sums). */
typedef struct simpler_stats {
    uint64_t photons;       /* photons launched                              */
    uint64_t er_steps;      /* er_step invocations, accepted or refined      */
    uint64_t scatters;      /* HG scattering events                          */
    uint64_t nee_splats;    /* NEE contributions that reached the sensor     */
    uint64_t direct_splats; /* direct contributions that reached the sensor  */
    uint64_t face_events;   /* dielectric BSDF samples at block faces        */
    uint64_t segments;      /* march segments (traceTillBlock calls)         */
    uint64_t aborted;       /* marches abandoned after 2^20 face events      */
} simpler_stats;

typedef struct simpler_ctx simpler_ctx;
typedef struct simpler_scene simpler_scene;

/* Returns SIMPLER_ABI_VERSION. */
int simpler_abi_version(void);

/* Fills a descriptor with the defaults of the reference drivers
 * (renderer/pybindFiles/tests/test_renderer3d_sample.py:9-95,216-245). */
void simpler_scene_desc_default(simpler_scene_desc *desc);

/* ---- context: one device + one HIP stream; calls on one ctx are not
 *      thread-safe, separate ctx objects are independent.  Replaces the
 *      process-global __constant__ state of cuda_renderer.cu:28-48. ------ */
int simpler_ctx_create(int device, simpler_ctx **out);
/* Fails with SIMPLER_ERR_STATE while scenes or batches of the context are
 * alive (destroy them first); scene_destroy likewise fails while batches
 * built on the scene are alive. */
int simpler_ctx_destroy(simpler_ctx *ctx);
/* Run the context's work on a caller's HIP stream (hipStream_t passed as
 * void*).  NULL is the HIP null stream, which orders with every blocking
 * stream (e.g. PyTorch's default stream); a new context uses its own
 * non-blocking stream, whose handle simpler_ctx_own_stream returns.  The
 * switch is ordered: work issued afterwards on the new stream starts after
 * the work already queued on the previous stream (an event wait, no host
 * synchronisation), because all launches of a context share its work queue,
 * counters and adjoint buffers. */
int simpler_ctx_set_stream(simpler_ctx *ctx, void *hip_stream);
void *simpler_ctx_get_stream(simpler_ctx *ctx);
void *simpler_ctx_own_stream(simpler_ctx *ctx);
int simpler_ctx_synchronize(simpler_ctx *ctx);
const char *simpler_last_error(simpler_ctx *ctx);

/* ---- scene (replaces Scene::from / Medium::from deep copies,
 *      cuda_scene.cuh:808-900): texture is the source's PFM intensity
 *      (tex_x * tex_y floats, row 0 = top after the PFM flip, image.h:216-223);
 *      NULL means a 1x1 white texture.  The texel CDF is built on the host
 *      exactly as area_textured_source.cpp:30-45 / pmf.h:41-99. ---------- */
int simpler_scene_create(simpler_ctx *ctx, const simpler_scene_desc *desc,
                         const float *texture_host, int tex_x, int tex_y,
                         simpler_scene **out);
int simpler_scene_destroy(simpler_scene *scene);
/* Cheap mutators (scene.h:179-219: set_f_u, set_n_max, set_n_scaling,
 * set_phase1, set_phase2): replace every scalar of the descriptor; geometry
 * sizes of attached tables stay. */
int simpler_scene_update(simpler_scene *scene, const simpler_scene_desc *desc);
int simpler_scene_get_desc(const simpler_scene *scene, simpler_scene_desc *out);

/* Point-source field (scene.cpp:140-195; rif_sources.h:156-198):
 *   n = n_o + sum_{c=1..n_coeff} coeff_weight[c-1] *
 *               sum_i amp[i] * cos(c*k*r_i + phase[i] - time_coupling*tau) / r_i
 * with r_i = |p - pos[i]|.  pos_xyz_host holds 3*n_src floats. */
int simpler_scene_set_sources(simpler_scene *scene, int n_src,
                              const float *pos_xyz_host, const float *amp_host,
                              const float *phase_host, int n_coeff,
                              const float *coeff_weight_host);

/* Tricubic B-spline field (spline.h:88-162,308-448): coeffs_host holds the
 * prefiltered coefficients, x fastest (index x + N0*(y + N1*z)).  Grid nodes
 * lie at xmin + i*(xmax-xmin)/(N-1). */
int simpler_scene_set_spline(simpler_scene *scene, const float *coeffs_host,
                             const int32_t dims[3], const float xmin[3],
                             const float xmax[3]);

/* Tricubic B-spline field from grid SAMPLES (node values, x fastest), with
 * the interpolation prefilter run on the GPU (replaces the host build3d of
 * spline.h:686-764 that rif_interpolated's VOL constructor runs,
 * spline.h:115-162 / rif_interpolated.h:101-104).  samples may be host or
 * device memory (samples_on_device); fails with SIMPLER_ERR_INVALID on a
 * non-finite sample.  Synchronous. */
int simpler_scene_set_spline_samples(simpler_scene *scene, const float *samples,
                                     int samples_on_device, const int32_t dims[3],
                                     const float xmin[3], const float xmax[3]);

/* The scene's B-spline coefficients (absolute values, x fastest) into host
 * memory; SIMPLER_ERR_STATE if the scene has none.  Synchronous. */
int simpler_scene_get_spline(const simpler_scene *scene, float *coeffs_host);

/* In-place B-spline interpolation prefilter of a device grid (spline.h:
 * 686-764: mirror-boundary recursive filter, pole sqrt(3)-2, gain 6, along
 * y, then x, then z) on the context's stream; dims in [2, 4096]. */
int simpler_spline_prefilter(simpler_ctx *ctx, float *grid_dev, const int32_t dims[3]);

/* ---- the hot path ---------------------------------------------------- */
/*
 * Render photons [photon_offset, photon_offset + n_photons) of a render of
 * n_total photons into image_dev (float, layout [z][y][x] = z*X*Y + y*X + x,
 * as Image3::addEnergy image.h:320-324).  The photon weight is
 * Li*fresnelTrans/n_total (photon.h:76-80), so chunks with offsets sum to the
 * one-shot render.  accumulate = 0 zeroes image_dev first (renderImage
 * overwrites its target), 1 adds to it.  Random numbers: Philox-4x32-10 keyed
 * by seed, counter (photon index, draw index, stream), so the image is a
 * function of (scene, seed, photon range) only.  Asynchronous on the ctx
 * stream.  stats_host may be NULL; when given, the call synchronizes and
 * fills it.
 */
int simpler_render(simpler_scene *scene, int64_t n_photons, int64_t n_total,
                   uint64_t seed, uint64_t photon_offset, float *image_dev,
                   int x_res, int y_res, int z_res, int accumulate,
                   simpler_stats *stats_host);

/* Same as simpler_render, but with the image in host memory (copied in when
 * accumulate = 1, copied back at the end); synchronous, like
 * cuda::CudaRenderer::renderImage (cuda_renderer.cu:710-742). */
int simpler_render_host(simpler_scene *scene, int64_t n_photons,
                        int64_t n_total, uint64_t seed, uint64_t photon_offset,
                        float *image_host, int x_res, int y_res, int z_res,
                        int accumulate, simpler_stats *stats_host);

/* simpler_render_host without any device allocation: the same render, with
 * the launch's work queue, counters and device image staging in a fixed
 * file-scope device array of the library (no counterpart in the reference;
 * for callers that cannot allocate device memory: tight serving loops,
 * harnesses that forbid hipMalloc).  The array holds a 256 B header plus at
 * most 8,388,608 image floats (32 MiB, a 256 x 256 x 128 image); larger
 * images fail with SIMPLER_ERR_INVALID.  Each device the library runs on
 * holds one copy until the process ends, and every call on that device
 * shares it (the last call's contents stay): calls that overlap on one
 * device -- on two streams or two contexts at once -- race and must be
 * serialised by the caller.  Synchronous. */
int simpler_render_host_noalloc(simpler_scene *scene, int64_t n_photons,
                                int64_t n_total, uint64_t seed,
                                uint64_t photon_offset, float *image_host,
                                int x_res, int y_res, int z_res, int accumulate,
                                simpler_stats *stats_host);

/*
 * Per-ray record of the direct (unscattered) trace, the transparent-medium
 * parity mode: for each photon i in [photon_offset, photon_offset+n_rays) the
 * source is sampled, then the ray is marched to the sensor face exactly as
 * Renderer::directTracing (photon.cpp:40-87).  Outputs (device, n_rays each):
 * out_pos (3 floats/ray) = exit point on the block face, out_dir (3) = unit
 * exit direction inside the medium, out_len = optical path length at the exit
 * point, out_status = 1 reached the sensor face, 0 lost through another face,
 * -1 LargeDist exhausted, -2 abandoned (2^20 face events), -4 rejected by
 * the emitter lens.
 */
int simpler_trace_rays(simpler_scene *scene, int64_t n_rays, uint64_t seed,
                       uint64_t photon_offset, float *out_pos_dev,
                       float *out_dir_dev, float *out_len_dev,
                       int32_t *out_status_dev);

/* Forward mode (tangents) of the per-ray direct trace of simpler_trace_rays:
 * the same records plus their tangents along a direction of the field
 * parameters (theta_dot_dev: n_params floats in the layout of
 * simpler_adjoint_param_count, or NULL) and of each ray's initial velocity
 * v0 = n(p0) d0 (v0_dot_dev: 3 floats per ray, or NULL).  Replaces the
 * reference's er_derivativestep (scene.cpp:351-359, with d2V / d2Path,
 * scene.h:259-272), which propagates dp/dv0 and dv/dv0 along the leapfrog:
 * v0_dot = e_i gives their columns.  Face crossings are differentiated with
 * the implicit crossing time, reflections by reflecting the tangent.
 * Outputs: tan_pos_dev (3 per ray), tan_dir_dev (3, tangent of the unit
 * direction), tan_len_dev (1, optical length).  Asynchronous on the ctx
 * stream.  Together with simpler_adjoint_rays it gives the exact
 * dot-product identity <g, J theta_dot> = <J^T g, theta_dot>. */
int simpler_jvp_rays(simpler_scene *scene, int64_t n_rays, uint64_t seed,
                     uint64_t photon_offset, const float *theta_dot_dev,
                     const float *v0_dot_dev, float *out_pos_dev,
                     float *out_dir_dev, float *out_len_dev,
                     int32_t *out_status_dev, float *tan_pos_dev,
                     float *tan_dir_dev, float *tan_len_dev);

/* Deterministic next-event connection: replaces Scene::computefdfNEE
 * (scene.cpp:512-574, scene.h:291).  For each ray i (device arrays, 3
 * floats per ray): start at p1 with the velocity v1 scaled to |v| = n(p1),
 * march with the forward-mode leapfrog of er_derivativestep (scene.cpp:
 * 351-359) until x < p2.x, then bisect the last step ceil(precision /
 * log10 2) times (a half step is kept while it stays at x > p2.x).  The
 * reference's caller-initialised derivatives start as dp/dv1 = 0,
 * dv/dv1 = I.  Outputs: err_dev = p - p2 (3 per ray) and jac_dev = the
 * transpose of dQ/dv1 (9 per ray, row-major: row j is the derivative of the
 * crossing point Q with respect to v1_j), where dQ/dv1 = (I - t e_x^T / t_x)
 * dp/dv1, t = v / n(p), adds the crossing time's dependence on v1.  The
 * field is the scene's at time scaling `scaling` (Bessel) and duty phase
 * `tau` (point sources with time coupling).  The reference's HessianRIF
 * has no definition at HEAD (its predecessors: the unscaled Bessel
 * Hessian of scene.cpp:260-320 and a zero matrix for point sources,
 * scene.cpp:197-201); this uses each field's full Hessian.  The Bessel
 * component form (general axis off) is accepted where it equals the general
 * form (axis -x, or +x with m = 0).  The reference has no caller; the loop
 * runs at most 1e5 steps, as there.  Asynchronous on the ctx stream. */
int simpler_nee_connect(simpler_scene *scene, int64_t n_rays,
                        const float *p1_dev, const float *v1_dev,
                        const float *p2_dev, float scaling, float tau,
                        float *err_dev, float *jac_dev);

/* Refractive index n and its gradient at n_pts points (device arrays:
 * pts 3*n_pts, out_n n_pts, out_grad 3*n_pts) for the per-photon amplitude
 * scaling s and duty phase tau -- the field evaluation of the march
 * (scene.h:253-256,337-340). */
int simpler_field_eval(simpler_scene *scene, int n_pts, const float *pts_dev,
                       float scaling, float tau, float *out_n_dev,
                       float *out_grad_dev);

/* Read back the counters of the last render on this scene's ctx
 * (synchronizes). */
int simpler_read_stats(simpler_ctx *ctx, simpler_stats *stats_host);

/* ---- batched launch: many transducer configurations with their own photon
 *      budgets, one persistent kernel (reference: one renderImage call per
 *      configuration, e.g. the sweeps of FUS_configurations.ipynb and
 *      helperScripts/computePeakPhotonsPerVoltage.m:10-19).  The base scene
 *      (a Bessel field) fixes geometry, source, sensor and flags; each
 *      configuration overrides the fields below.  Configuration i renders
 *      photons [0, n_photons) with its own seed and weight Li/n_photons into
 *      image i of images_dev (n_cfg x z x y x x floats) -- exactly the image a
 *      single-scene simpler_render of that configuration produces. ---------- */
typedef struct simpler_batch_cfg {
    float n_max;
    float n_clip;
    float f_u;
    float phi_min;
    float phi_max;
    int32_t mode;
    float sigma_t;
    float albedo;
    float g;
    int32_t reserved;
    int64_t n_photons;
    uint64_t seed;
} simpler_batch_cfg;

typedef struct simpler_batch simpler_batch;
struct simpler_adjoint_stats; /* defined with the adjoint entry points below */

int simpler_batch_create(simpler_scene *base, int n_cfg, const simpler_batch_cfg *cfgs_host, int x_res,
                         int y_res, int z_res, simpler_batch **out);
/* Asynchronous on the ctx stream; stats_host (may be NULL) synchronizes. */
int simpler_batch_render(simpler_batch *batch, float *images_dev, int accumulate, simpler_stats *stats_host);
/* Capture the zeroing + render of images_dev into a HIP graph (on the ctx's
 * own stream); simpler_batch_graph_launch replays it on the ctx stream. */
int simpler_batch_graph_capture(simpler_batch *batch, float *images_dev);
int simpler_batch_graph_launch(simpler_batch *batch);
int simpler_batch_destroy(simpler_batch *batch);

/* ---- point-source pattern batch (no reference counterpart: the reference
 *      sweeps its phases one renderImage call at a time, e.g.
 *      FUS_configurations.ipynb with set_phase1/set_phase2, scene.h:179-219).
 *      The base scene (a point-source field with its sources) fixes source
 *      positions, harmonic weights, k, geometry, medium and flags; pattern c
 *      has its own amplitudes amp[c*n_src + i], phases phase[c*n_src + i],
 *      photon budget and seed.  Render with simpler_batch_render (images c =
 *      pattern c) or capture it as a graph like any batch. -------------------- */
int simpler_batch_create_sources(simpler_scene *base, int n_cfg, const float *amp_host,
                                 const float *phase_host, const int64_t *n_photons_host,
                                 const uint64_t *seeds_host, int x_res, int y_res, int z_res,
                                 simpler_batch **out);

/* Path-replay adjoint of every pattern of a point-source batch in one launch:
 * F_c = sum dF_dimages[c] * image_c; grads_dev[c * (2 n_src + 1) + j] receives
 * dF_c / dtheta_j in the layout of simpler_adjoint_param_count (amplitudes,
 * phases, k); objectives_host (n_cfg doubles) and stats_host may be NULL. */
int simpler_batch_adjoint(simpler_batch *batch, const float *dF_dimages_dev,
                          float *grads_dev, int accumulate, double *objectives_host,
                          struct simpler_adjoint_stats *stats_host);

/* ---- path-replay adjoint (no counterpart in the reference, which has no
 *      gradient with respect to field parameters; its forward-mode
 *      derivative machinery is scene.cpp:351-359 / scene.h:259-272).
 *
 * Differentiable parameters theta of the scene's field:
 *   spline         the N0*N1*N2 coefficients, layout of simpler_scene_set_spline;
 *   point sources  2*n_src + 1 values: amplitudes a_i, phases phi_i, then the
 *                  wavenumber k = 2 pi f_u / speed_u (n_src <= 16);
 *   Bessel         3 values: the amplitude n_max (with n_clip / n_max held
 *                  fixed), the wavenumber k = 2 pi f_u / speed_u, and a shift
 *                  of the duty window [phi_min, phi_max] (the transducer
 *                  phase); scenes in the reference's component form (general
 *                  axis off) are accepted for the axis -x, or +x at m = 0.
 * Every photon is re-traced from its seed exactly as by simpler_render; paths
 * are differentiated with their random numbers held fixed (the Fresnel
 * transmission factor and the discrete choices -- time bin, face, reflect or
 * refract -- are held fixed too).  Scenes with an active sensor lens or
 * precision > 7 are rejected.  grad_dev (float, n_params) is overwritten, or
 * added to with accumulate = 1.  objective_host / stats_host may be NULL;
 * when given, the call synchronizes and fills them. ------------------------ */
typedef struct simpler_adjoint_stats {
    uint64_t photons;
    uint64_t er_steps;             /* forward ER steps                          */
    uint64_t reverse_steps;        /* backward (adjoint) ER steps               */
    uint64_t sub_paths_reversed;   /* NEE / direct / per-ray sub-paths          */
    uint64_t main_paths_reversed;
    uint64_t dropped;              /* contributions skipped: segment log full   */
} simpler_adjoint_stats;

int simpler_adjoint_param_count(const simpler_scene *scene, int64_t *n_params);

/* Gradient of F = sum_pixels dF_dimage[pixel] * image[pixel], image being the
 * render of simpler_render with the same arguments (dF_dimage: device float
 * array, same layout and resolution); objective_host receives F (double). */
int simpler_adjoint_render(simpler_scene *scene, int64_t n_photons, int64_t n_total,
                           uint64_t seed, uint64_t photon_offset,
                           const float *dF_dimage_dev, int x_res, int y_res, int z_res,
                           float *grad_dev, int accumulate, double *objective_host,
                           simpler_adjoint_stats *stats_host);

/* Gradient of F = sum_rays g_pos . exit_pos + g_dir . exit_dir + g_len * len
 * over the rays of simpler_trace_rays with the same arguments (device arrays
 * 3, 3 and 1 floats per ray); rays that do not reach the sensor face add
 * nothing. */
int simpler_adjoint_rays(simpler_scene *scene, int64_t n_rays, uint64_t seed,
                         uint64_t photon_offset, const float *g_pos_dev,
                         const float *g_dir_dev, const float *g_len_dev,
                         float *grad_dev, int accumulate, double *objective_host,
                         simpler_adjoint_stats *stats_host);

/* Tunables of the persistent march kernel (picked by the autotuner):
 * block_size in {64,128,256} (the kernels are compiled for at most 256
 * lanes per workgroup), steps_per_chunk in [1,4096] (march steps between
 * event phases), blocks_per_cu in [1,64].  0 keeps the current value; other
 * values fail with SIMPLER_ERR_INVALID. */
int simpler_ctx_set_tuning(simpler_ctx *ctx, int block_size,
                           int steps_per_chunk, int blocks_per_cu);
int simpler_ctx_get_tuning(simpler_ctx *ctx, int *block_size,
                           int *steps_per_chunk, int *blocks_per_cu);

/* Splat strategy of the render kernels (an autotuner candidate): 0 = each
 * lane issues its splat's 4 device atomics (cuda_renderer.cu:457-509 does the
 * same per thread); 1 = wave-aggregated splats: after each event phase the
 * wave groups its lanes' pending splats by image plane and base pixel, sums
 * each group's weights across the wave and issues the group's 4 atomics once.
 * Strategy 1 is compiled for single-scene Bessel renders in the axis frame
 * (uz = +z, ux = +x); other renders splat per lane whatever the setting. */
int simpler_ctx_set_splat(simpler_ctx *ctx, int strategy);
int simpler_ctx_get_splat(simpler_ctx *ctx, int *strategy);

/* Schedule of simpler_adjoint_render's backward runs (no reference
 * counterpart: the reference has no adjoint).  schedule 0 = inline: a lane
 * that splats into the objective runs the sub-path backwards at once, and the
 * photon's main path when it ends, while the other lanes of its wave march
 * on; 1 (default) = deferred: the re-trace queues each such sub-path (its log
 * and adjoint seed) and each photon's main-path log, then a second kernel runs
 * the queued sub-paths and a third the main paths, every lane reversing.
 * queue_capacity c (0 keeps the current value, default 2^23; else
 * 64 <= c <= 2^26) sizes the queues of one chunk of photons: c sub-paths
 * with c sub-path records, c main paths with 8 c main-path records, 768 c
 * bytes of device memory (6 GiB at the default), allocated on first use and
 * kept by the context.  The chunk starts at 2^22 photons, halves when a
 * chunk's queues overflow (the chunk is then traced again) and doubles, up
 * to 2^24, after a chunk that filled no queue past 40 %; at 2^14 photons an
 * overflowing chunk runs inline.  Both schedules give the same objective
 * and counts, and the same gradient up to the order of float additions. */
int simpler_ctx_set_adjoint_schedule(simpler_ctx *ctx, int schedule, int64_t queue_capacity);
int simpler_ctx_get_adjoint_schedule(simpler_ctx *ctx, int *schedule, int64_t *queue_capacity);

/* Forward renders (no reference counterpart: the reference traces every
 * next-event sub-path, scene.cpp:866-893): enable != 0 (the default) leaves
 * untraced the next-event sub-paths that provably add nothing to the image,
 * for spline fields whose least coefficient is at least 1, with the camera
 * on the block's -x face looking along -x and no sensor lens (Bessel and
 * point-source renders trace every sub-path).  From a bound on |grad n| the sub-path is replayed as a polyline from
 * face to face: it is skipped when it leaves through a side face (total
 * reflections on the way are followed; the face's Fresnel draw, taken ahead
 * from the photon's counter-based stream, is above the largest reflectance
 * the bound allows) or meets the sensor face outside the image.  The image,
 * the splat counts and every draw of the photon are those of the full trace;
 * er_steps and face_events count the work done and drop.  Record renders and
 * adjoints are not affected (image adjoints of every field skip by their
 * objective's support in any case). */
int simpler_ctx_set_next_event_skip(simpler_ctx *ctx, int enable);
int simpler_ctx_get_next_event_skip(simpler_ctx *ctx, int *enable);

/* Autotuner (no reference counterpart: the reference hard-codes its launch,
 * cuda_renderer.cu:763-767): times an n_photons render of the scene at the
 * given resolution for every candidate (block size 128/256, steps per chunk
 * 32/64/128, blocks per CU 3/4/5, and the splat strategy 0/1 where strategy 1
 * exists for the scene), keeps the fastest on the scene's context and returns
 * it (outputs may be NULL; the splat strategy is read back with
 * simpler_ctx_get_splat).  Synchronous. */
int simpler_autotune(simpler_scene *scene, int64_t n_photons, int x_res,
                     int y_res, int z_res, int *block_size, int *steps_per_chunk,
                     int *blocks_per_cu, float *ms_best);

#ifdef __cplusplus
}
#endif

#endif /* SIMPLER_H_ */
