/*
 * oracle.cpp -- CPU restatement of simplEr's photon-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product path (simpler_amd, libsimpler_hip.so) never links or calls it.
 *
 * It restates, function by function, the reference CPU path
 *   photon::Renderer::renderImage ... Scene::er_step
 * (reference files cited per function below) in a template over the real
 * type R (float = the reference build's Float, double = the FP64 oracle for
 * per-ray checks).  Differences from the reference, all deliberate:
 *   * random numbers: Philox-4x32-10 (Salmon et al., SC'11), keyed by the
 *     seed, counter (photon, draw>>2, sub-stream), word draw&3, converted to
 *     (2k+1)*2^-24 with k the top 23 bits (the reference's SFMT stream is
 *     re-seeded per photon, photon.cpp:313, and yields k*2^-23,
 *     rng_sse.h:40-50).  Same draw ORDER as the reference within each
 *     sub-stream: 0 = main path, 1 = direct trace, 1+depth = NEE at that
 *     vertex (the reference uses one sequential stream for all of them).
 *   * optional dead-path cutoff (oracle_render cutoff=1, off by default):
 *     marches whose optical length can no longer pass the time gate stop;
 *     the image is unchanged (tests compare both modes).
 *   * 3-D spline indices are mirror-wrapped like the 1-D/2-D code
 *     (spline.h:617-620,636-644); the 3-D reference reads out of bounds
 *     within two cells of the edge (spline.h:313-316).  The reference's own
 *     golden vectors (splinepractice/interpolate3d.csv) include such points
 *     and agree with mirror wrapping.
 *   * images accumulate in double.
 *   * Bessel gradient: the reference's axis=x component formula
 *     (scene.cpp:248-251) or, with SIMPLER_FLAG_GENERAL_AXIS, the same
 *     gradient written for any axis (identical at axis = +-x).
 *
 * Parity pins (tests/test_oracle_*.py): Philox known-answer vectors; spline
 * kernels / prefilter / interpolation against the reference CSVs; Bessel
 * values against scipy and closed-form FD; the integrator against straight
 * lines (dn = 0), second-order convergence and reversibility; HG / hemisphere
 * / CDF / free path by moments and chi-square; the dielectric BSDF against the
 * Fresnel closed form; whole renders by energy bookkeeping (transparent
 * slab), chunked additivity and FP32-vs-FP64 agreement.
 */
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/simpler.h"

namespace oracle {

/* ------------------------------------------------------------------ RNG */
/* Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85. */
static inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                                 uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

static inline float u32_to_uniform(uint32_t w) {
    /* (2k+1) * 2^-24, k = top 23 bits: exact in FP32, in (0,1). */
    return (float)(((w >> 9) << 1) | 1u) * (1.0f / 16777216.0f);
}

/* Per-photon sampler: the reference re-seeds smp::Sampler with the photon
 * index for every photon (photon.cpp:313); here the stream is the Philox
 * counter sequence of that photon. */
struct Sampler {
    uint32_t key[2];
    uint64_t photon;
    uint32_t stream;
    uint32_t draw;
    uint64_t *draw_counter;
    float next() {
        const uint32_t ctr[4] = {(uint32_t)photon, (uint32_t)(photon >> 32), draw >> 2, stream};
        uint32_t out[4];
        philox4x32_10(ctr, key, out);
        const float u = u32_to_uniform(out[draw & 3u]);
        ++draw;
        return u;
    }
};

/* -------------------------------------------------------------- vectors */
template <class R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    R &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3 &o) const { return V3(x + o.x, y + o.y, z + o.z); }
    V3 operator-(const V3 &o) const { return V3(x - o.x, y - o.y, z - o.z); }
    V3 operator-() const { return V3(-x, -y, -z); }
    V3 operator*(R s) const { return V3(x * s, y * s, z * s); }
    V3 operator/(R s) const { return V3(x / s, y / s, z / s); }
    V3 &operator+=(const V3 &o) {
        x += o.x;
        y += o.y;
        z += o.z;
        return *this;
    }
    V3 &operator*=(R s) {
        x *= s;
        y *= s;
        z *= s;
        return *this;
    }
    R lengthSquared() const { return x * x + y * y + z * z; }
    R length() const { return std::sqrt(lengthSquared()); }
    void normalize() {
        const R l = length();
        x /= l;
        y /= l;
        z /= l;
    }
};
template <class R>
static inline V3<R> operator*(R s, const V3<R> &v) {
    return V3<R>(v.x * s, v.y * s, v.z * s);
}
template <class R>
static inline R dot(const V3<R> &a, const V3<R> &b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <class R>
static inline V3<R> cross(const V3<R> &a, const V3<R> &b) {
    return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R>
static inline V3<R> v3(const float *f) {
    return V3<R>((R)f[0], (R)f[1], (R)f[2]);
}

/* constants.h:38-41,60-68 */
template <class R>
struct K {
    static constexpr R EPS = (R)1.19209290e-07f;
    static constexpr R MAXV = (R)3.40282347e+38f;
    static constexpr R PI = (R)3.14159265358979323846;
};

/* --------------------------------------------------------------- spline */
/* spline.h:33-62 */
template <class R>
static inline R bs_kernel0(R x) {
    x = std::abs(x);
    if (x > 2) return R(0);
    if (x > 1) return (R(1) / R(6)) * (2 - x) * (2 - x) * (2 - x);
    return (R(2) / R(3)) - x * x + R(0.5) * x * x * x;
}
template <class R>
static inline R bs_kernel1(R x) {
    const int s = (R(0) < x) - (x < R(0));
    x = std::abs(x);
    if (x > 2) return R(0);
    if (x > 1) return s * (R(-0.5) * (2 - x) * (2 - x));
    return s * ((R(1.5) * x - 2) * x);
}
template <class R>
static inline R bs_kernel2(R x) {
    x = std::abs(x);
    if (x > 2) return R(0);
    if (x > 1) return 2 - x;
    return 3 * x - 2;
}
template <class R>
static inline R bs_kernel(int d, R x) {
    return d == 0 ? bs_kernel0(x) : (d == 1 ? bs_kernel1(x) : bs_kernel2(x));
}
/* mirror wrap of spline.h:617-620 */
static inline int mirror_index(int index, int n) {
    if (n == 1) return 0;
    const int period = 2 * n - 2;
    int w = index % period;
    if (w < 0) w += period;
    if (w >= n) w = period - w;
    return w;
}

/* spline.h:686-714: causal/anti-causal recursive prefilter with mirror
 * boundary, pole z1 = sqrt(3) - 2. */
template <class R>
static void build1d(const R *data, long offset, long stride, int size, R *out) {
    const R z1 = -2 + std::sqrt((R)3);
    std::vector<R> cp(size), cn(size);
    cp[0] = 0;
    for (int i = 0; i < size; i++) cp[0] += data[offset + i * stride] * std::pow(z1, (R)i);
    for (int i = size - 2; i > 0; i--)
        cp[0] += data[offset + i * stride] * std::pow(z1, (R)(2 * size - 2 - i));
    cp[0] /= (1 - std::pow(z1, (R)(2 * size - 2)));
    for (int i = 1; i < size; i++) cp[i] = data[offset + i * stride] + z1 * cp[i - 1];
    cn[size - 1] = z1 / (z1 * z1 - 1) * (cp[size - 1] + z1 * cp[size - 2]);
    for (int i = size - 2; i >= 0; i--) cn[i] = z1 * (cn[i + 1] - cp[i]);
    for (int i = 0; i < size; i++) out[i] = 6 * cn[i];
}

/* spline.h:739-764 (y pass, x pass, z pass) */
template <class R>
static void build3d(const R *data, const int N[3], R *coeff) {
    std::vector<R> temp(std::max(std::max(N[0], N[1]), N[2]));
    const long NX = N[0], NXY = (long)N[0] * N[1];
    for (int k = 0; k < N[2]; k++)
        for (int i = 0; i < N[0]; i++) {
            build1d(data, k * NXY + i, NX, N[1], temp.data());
            for (int t = 0; t < N[1]; t++) coeff[i + t * NX + k * NXY] = temp[t];
        }
    for (int k = 0; k < N[2]; k++)
        for (int j = 0; j < N[1]; j++) {
            build1d(coeff, k * NXY + j * NX, 1, N[0], temp.data());
            for (int t = 0; t < N[0]; t++) coeff[t + j * NX + k * NXY] = temp[t];
        }
    for (int i = 0; i < N[0]; i++)
        for (int j = 0; j < N[1]; j++) {
            build1d(coeff, j * NX + i, NXY, N[2], temp.data());
            for (int t = 0; t < N[2]; t++) coeff[i + j * NX + t * NXY] = temp[t];
        }
}

template <class R>
struct Spline3 {
    int N[3];
    R xmin[3], xres[3];
    const float *coeff; /* FP32 coefficients, as the device holds them */
    /* spline.h:600-603 */
    void to_grid(const V3<R> &p, R x[3]) const {
        for (int i = 0; i < 3; i++) x[i] = (p[i] - xmin[i]) * xres[i];
    }
    /* spline.h:659-683 (value3d<DX,DY,DZ>) with mirror-wrapped indices */
    R value(const V3<R> &p, int dx, int dy, int dz) const {
        R x[3];
        to_grid(p, x);
        R v = 0;
        for (int i1 = (int)std::ceil(x[0] - 2); i1 <= (int)std::floor(x[0] + 2); i1++) {
            const int w1 = mirror_index(i1, N[0]);
            const R k1 = bs_kernel(dx, x[0] - i1);
            for (int i2 = (int)std::ceil(x[1] - 2); i2 <= (int)std::floor(x[1] + 2); i2++) {
                const int w2 = mirror_index(i2, N[1]);
                const R k2 = bs_kernel(dy, x[1] - i2);
                for (int i3 = (int)std::ceil(x[2] - 2); i3 <= (int)std::floor(x[2] + 2); i3++) {
                    const int w3 = mirror_index(i3, N[2]);
                    v += (R)coeff[w1 + (long)N[0] * (w2 + (long)N[1] * w3)] * k1 * k2 *
                         bs_kernel(dz, x[2] - i3);
                }
            }
        }
        const int d[3] = {dx, dy, dz};
        for (int i = 0; i < 3; i++) {
            if (d[i] == 1) v *= xres[i];
            if (d[i] == 2) v *= xres[i] * xres[i];
        }
        return v;
    }
};

/* ---------------------------------------------------------------- scene */
template <class R>
struct Scene {
    simpler_scene_desc d;
    /* derived (rif_base.h:144-154; rif_analytical.h:64-68; camera.cpp:33-39;
     * area_textured_source.cpp:36-39) */
    V3<R> blockL, blockR, camOrigin, camDir, camH, camV;
    R k_r, n_maxScaling, ct, pixsize[2];
    R weight, fresnelTrans;
    std::vector<float> cdf; /* FP32 texel CDF, pmf.h:41-99 */
    int tex_x, tex_y;
    /* point sources */
    int n_src, n_coeff;
    std::vector<float> src_pos, src_amp, src_phase, coeff_w;
    /* spline */
    Spline3<R> spl;
    bool has_spline;
};

template <class R>
static void scene_init(Scene<R> &S, const simpler_scene_desc &d, const float *cdf, int ncdf,
                       int tex_x, int tex_y, int64_t n_total) {
    S.d = d;
    S.blockL = v3<R>(d.block_l);
    S.blockR = v3<R>(d.block_r);
    S.camOrigin = v3<R>(d.view_origin);
    S.camDir = v3<R>(d.view_dir);
    S.camDir.normalize();
    S.camH = v3<R>(d.view_horizontal);
    S.camH.normalize();
    S.camV = cross(S.camDir, S.camH);
    /* rif_base.h:149-153: lambda = c/f, k = 2 pi / lambda */
    const R lambda = (R)d.speed_u / (R)d.f_u;
    S.k_r = (2 * K<R>::PI) / lambda;
    S.n_maxScaling = (R)d.n_clip / (R)d.n_max;
    S.ct = std::cos((R)d.half_theta_limit);
    S.tex_x = tex_x;
    S.tex_y = tex_y;
    S.pixsize[0] = (R)d.light_plane[0] / tex_x;
    S.pixsize[1] = (R)d.light_plane[1] / tex_y;
    S.cdf.assign(cdf, cdf + ncdf);
    /* scene.h:141-164: fresnelTrans (= 1 with Fresnel compiled out) */
    S.fresnelTrans = 1;
    if ((d.flags & SIMPLER_FLAG_USE_FRESNEL) && (R)d.ior > 1) {
        const V3<R> ld = v3<R>(d.light_dir);
        R sumSqr = 0;
        R ry = ld.y / (R)d.ior, rz = ld.z / (R)d.ior;
        sumSqr = ry * ry + rz * rz;
        R rx = std::sqrt(1 - sumSqr);
        if (ld.x < 0) rx = -rx;
        const R eta = (R)d.ior;
        R Rs = (ld.x - eta * rx) / (ld.x + eta * rx);
        R Rp = (rx - eta * ld.x) / (rx + eta * ld.x);
        R F = (std::abs(eta - 1) < K<R>::EPS * std::max((R)1, std::abs(eta))) ? R(0)
                                                                                 : R(0.5) * (Rs * Rs + Rp * Rp);
        S.fresnelTrans = eta * eta * (1 - F);
    }
    /* photon.h:76-80 */
    S.weight = (R)d.Li * S.fresnelTrans / (R)n_total;
    S.n_src = 0;
    S.n_coeff = 0;
    S.has_spline = false;
}

/* --------------------------------------------------------------- fields */
/* Bessel J_m in double: glibc jn, as the reference CPU path
 * (scene.cpp:214,238,240). */
static inline double besselj(int m, double x) { return jn(m, x); }

/* cuda_scene.cuh:463-481 / rif_base.h:134-137: gap rule */
template <class R>
static inline bool outside_gap(const Scene<R> &S, const V3<R> &p) {
    return p.x > (R)S.d.egap_end_loc_x || p.x < (R)S.d.sgap_begin_loc_x;
}

/* scene.cpp:205-215 */
template <class R>
static R bessel_RIF(const Scene<R> &S, const V3<R> &p, R scaling) {
    const V3<R> uz = v3<R>(S.d.axis_uz), ux = v3<R>(S.d.axis_ux), pu = v3<R>(S.d.p_u);
    const V3<R> p_axis = pu + dot(p - pu, uz) * uz;
    const R r = (p - p_axis).length();
    const R dotp = dot(p - p_axis, ux);
    const R detp = dot(cross(ux, p - p_axis), uz);
    const R phi = std::atan2(detp, dotp);
    const int m = S.d.mode;
    return (R)((double)S.d.n_o + (double)((R)S.d.n_max * scaling) * besselj(m, (double)(S.k_r * r)) *
                                     (double)std::cos((R)m * phi));
}

/* scene.cpp:217-257 (reference components, axis = +-x) and the same
 * gradient for a general axis (SIMPLER_FLAG_GENERAL_AXIS). */
template <class R>
static V3<R> bessel_dRIF(const Scene<R> &S, const V3<R> &q, R scaling) {
    const V3<R> uz = v3<R>(S.d.axis_uz), ux = v3<R>(S.d.axis_ux), pu = v3<R>(S.d.p_u);
    const V3<R> p_axis = pu + dot(q - pu, uz) * uz;
    V3<R> p = q - p_axis;
    R r = p.length();
    const R dotp = dot(p, ux);
    const R detp = dot(cross(ux, p), uz);
    const R phi = std::atan2(detp, dotp);
    const bool general = (S.d.flags & SIMPLER_FLAG_GENERAL_AXIS) != 0;
    const V3<R> uy = cross(uz, ux);
    if (r < K<R>::EPS) {
        if (general) {
            p = K<R>::EPS * (ux + uy);
        } else {
            p.y = K<R>::EPS;
            p.z = K<R>::EPS;
        }
        r = K<R>::EPS;
    }
    const int m = S.d.mode;
    const R krr = S.k_r * r;
    const R J = (R)besselj(m, (double)krr);
    const R dJ = (R)m / krr * J - (R)besselj(m + 1, (double)krr);
    const R invr = R(1) / r;
    const R invr2 = invr * invr;
    const R cosmp = std::cos((R)m * phi);
    const R sinmp = std::sin((R)m * phi);
    const R a = (R)S.d.n_max * scaling;
    if (!general) {
        return V3<R>(0, a * (dJ * S.k_r * p.y * invr * cosmp - J * (R)m * sinmp * p.z * invr2),
                     a * (dJ * S.k_r * p.z * invr * cosmp + J * (R)m * sinmp * p.y * invr2));
    }
    const V3<R> tphi = cross(uz, p); /* r * (azimuthal unit vector) */
    const R cr = a * dJ * S.k_r * invr * cosmp;
    const R ct = a * J * (R)m * sinmp * invr2;
    return V3<R>(cr * p.x - ct * tphi.x, cr * p.y - ct * tphi.y, cr * p.z - ct * tphi.z);
}

/* scene.cpp:140-160 generalised: per-source amplitude and phase, per
 * harmonic weight, optional time phase (see simpler.h). */
template <class R>
static R sources_RIF(const Scene<R> &S, const V3<R> &p, R tau) {
    R rif = 0;
    for (int c = 1; c <= S.n_coeff; c++) {
        R t = 0;
        for (int n = 0; n < S.n_src; n++) {
            const V3<R> ci((R)S.src_pos[3 * n], (R)S.src_pos[3 * n + 1], (R)S.src_pos[3 * n + 2]);
            const R r = (p - ci).length();
            t += (R)S.src_amp[n] *
                 std::cos(S.k_r * r * (R)c + (R)S.src_phase[n] - (R)S.d.time_coupling * tau) / r;
        }
        rif += t * (R)S.coeff_w[c - 1];
    }
    return (R)S.d.n_o + rif;
}

/* scene.cpp:162-195 generalised likewise */
template <class R>
static V3<R> sources_dRIF(const Scene<R> &S, const V3<R> &p, R tau) {
    V3<R> dn(0, 0, 0);
    for (int c = 1; c <= S.n_coeff; c++) {
        V3<R> t(0, 0, 0);
        for (int n = 0; n < S.n_src; n++) {
            const V3<R> ci((R)S.src_pos[3 * n], (R)S.src_pos[3 * n + 1], (R)S.src_pos[3 * n + 2]);
            const V3<R> dv = p - ci;
            const R r = dv.length();
            const R krrc = S.k_r * r * (R)c;
            const R ph = krrc + (R)S.src_phase[n] - (R)S.d.time_coupling * tau;
            const R common = -(std::cos(ph) + krrc * std::sin(ph)) / (r * r * r);
            t += ((R)S.src_amp[n] * common) * dv;
        }
        dn += t * (R)S.coeff_w[c - 1];
    }
    return dn;
}

/* Per-photon time state: scaling s (Bessel amplitude, photon.cpp:324) and
 * the duty phase tau it is the sine of. */
template <class R>
struct TimeState {
    R scaling;
    R tau;
};

template <class R>
static R RIF(const Scene<R> &S, const V3<R> &p, const TimeState<R> &ts) {
    if (outside_gap(S, p)) return (R)S.d.n_o;
    switch (S.d.field_kind) {
        case SIMPLER_FIELD_BESSEL: return bessel_RIF(S, p, ts.scaling);
        case SIMPLER_FIELD_SOURCES: return sources_RIF(S, p, ts.tau);
        default: return S.spl.value(p, 0, 0, 0); /* rif_interpolated.h:46-60 */
    }
}

template <class R>
static V3<R> dRIF(const Scene<R> &S, const V3<R> &p, const TimeState<R> &ts) {
    if (outside_gap(S, p)) return V3<R>(0, 0, 0);
    switch (S.d.field_kind) {
        case SIMPLER_FIELD_BESSEL: return bessel_dRIF(S, p, ts.scaling);
        case SIMPLER_FIELD_SOURCES: return sources_dRIF(S, p, ts.tau);
        default: /* rif_interpolated.h:62-78; spline.h:324-356 */
            return V3<R>(S.spl.value(p, 1, 0, 0), S.spl.value(p, 0, 1, 0), S.spl.value(p, 0, 0, 1));
    }
}

/* Hessian of n (symmetric, H[i][j] = d2n / dx_i dx_j), for the
 * deterministic next-event connection.  The reference's HessianRIF has no
 * definition at HEAD (its predecessors in scene.cpp: an unscaled Bessel
 * Hessian, 260-320, and a zero matrix for point sources, 197-201); here each
 * field's full second derivative, derived from its closed form:
 *   Bessel  f = A J_m(k r) cos(m phi), A = n_max s; in the frame (r^, t^ =
 *           uz x r^): H = f_rr r^r^T + (f_r / r + f_pp / r^2) t^t^T
 *           + (f_rp / r - f_p / r^2)(r^t^T + t^r^T), with J'' from Bessel's
 *           equation J'' = -J'/x - (1 - m^2 / x^2) J (glibc jn);
 *   sources f = cos(psi) / r per source and harmonic: H = f'' r^r^T
 *           + (f' / r)(I - r^r^T);
 *   spline  the B-spline kernels' second derivatives (spline.h:33-53). */
template <class R>
struct M3 {
    R m[3][3];
};

template <class R>
static M3<R> bessel_HRIF(const Scene<R> &S, const V3<R> &q, R scaling) {
    const V3<R> uz = v3<R>(S.d.axis_uz), ux = v3<R>(S.d.axis_ux), pu = v3<R>(S.d.p_u);
    const V3<R> p_axis = pu + dot(q - pu, uz) * uz;
    V3<R> p = q - p_axis;
    R r = p.length();
    const R phi = std::atan2(dot(cross(ux, p), uz), dot(p, ux));
    if (r < K<R>::EPS) {
        p = K<R>::EPS * (ux + cross(uz, ux));
        r = K<R>::EPS;
    }
    const int m = S.d.mode;
    const R x = S.k_r * r;
    const R J = (R)besselj(m, (double)x);
    const R dJ = (R)m / x * J - (R)besselj(m + 1, (double)x);
    const R d2J = -dJ / x - (R(1) - (R)(m * m) / (x * x)) * J;
    const R A = (R)S.d.n_max * scaling, C = std::cos((R)m * phi), Sn = std::sin((R)m * phi), k = S.k_r;
    const R f_r = A * k * dJ * C, f_p = -A * (R)m * J * Sn;
    const R f_rr = A * k * k * d2J * C, f_pp = -A * (R)(m * m) * J * C, f_rp = -A * (R)m * k * dJ * Sn;
    const V3<R> rh = p / r, th = cross(uz, rh);
    const R a = f_rr, b = f_r / r + f_pp / (r * r), c = f_rp / r - f_p / (r * r);
    M3<R> H;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H.m[i][j] = a * rh[i] * rh[j] + b * th[i] * th[j] + c * (rh[i] * th[j] + th[i] * rh[j]);
    return H;
}

template <class R>
static M3<R> sources_HRIF(const Scene<R> &S, const V3<R> &p, R tau) {
    M3<R> H{};
    for (int c = 1; c <= S.n_coeff; c++) {
        for (int n = 0; n < S.n_src; n++) {
            const V3<R> ci((R)S.src_pos[3 * n], (R)S.src_pos[3 * n + 1], (R)S.src_pos[3 * n + 2]);
            const V3<R> dv = p - ci;
            const R r = dv.length();
            const R kc = S.k_r * (R)c;
            const R psi = kc * r + (R)S.src_phase[n] - (R)S.d.time_coupling * tau;
            const R g = std::cos(psi) + kc * r * std::sin(psi);
            const R f1 = -g / (r * r);
            const R f2 = R(2) * g / (r * r * r) - kc * kc * std::cos(psi) / r;
            const R w = (R)S.src_amp[n] * (R)S.coeff_w[c - 1];
            const V3<R> rh = dv / r;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    H.m[i][j] += w * (f2 * rh[i] * rh[j] + (f1 / r) * ((i == j ? R(1) : R(0)) - rh[i] * rh[j]));
        }
    }
    return H;
}

template <class R>
static M3<R> HRIF(const Scene<R> &S, const V3<R> &p, const TimeState<R> &ts) {
    M3<R> H{};
    if (outside_gap(S, p)) return H;
    switch (S.d.field_kind) {
        case SIMPLER_FIELD_BESSEL: return bessel_HRIF(S, p, ts.scaling);
        case SIMPLER_FIELD_SOURCES: return sources_HRIF(S, p, ts.tau);
        default:
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    int d[3] = {0, 0, 0};
                    d[i]++;
                    d[j]++;
                    H.m[i][j] = S.spl.value(p, d[0], d[1], d[2]);
                }
            return H;
    }
}

/* scene.cpp:512-574 computefdfNEE with er_derivativestep (351-359) and
 * dV / d2V / d2Path (scene.h:253-272); dpdv0 and dvdv0 enter as 0 and I.
 * Matrices: P[i][j] = dp_i / dv1_j, V[i][j] = dv_i / dv1_j. */
template <class R>
static void nee_connect(const Scene<R> &S, const TimeState<R> &ts, const V3<R> &v_i, const V3<R> &p1,
                        const V3<R> &p2, int nbis, V3<R> &error, M3<R> &derror) {
    R h = (R)S.d.er_stepsize;
    V3<R> p = p1, v = v_i;
    M3<R> P{}, V{};
    const R r = RIF(S, p, ts);
    const R n1 = v.length(), n2 = n1 * n1, n3 = n2 * n1;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) V.m[i][j] = ((i == j ? n2 : R(0)) - v[i] * v[j]) / n3 * r;
    v = v / n1 * r;
    auto step = [&](V3<R> &pp, V3<R> &vv, M3<R> &PP, M3<R> &VV, R hs) {
        const R hh = R(0.5) * hs;
        vv += hh * dRIF(S, pp, ts);
        M3<R> H = HRIF(S, pp, ts);
        M3<R> T{};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) T.m[i][j] += H.m[i][k] * PP.m[k][j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) VV.m[i][j] += hh * T.m[i][j];
        pp += hs * vv / RIF(S, pp, ts);
        const R n = RIF(S, pp, ts);
        const V3<R> g = dRIF(S, pp, ts);
        R gP[3] = {0, 0, 0};
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) gP[j] += g[k] * PP.m[k][j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) PP.m[i][j] += hs * (-(vv[i] * gP[j]) / (n * n) + VV.m[i][j] / n);
        vv += hh * dRIF(S, pp, ts);
        H = HRIF(S, pp, ts);
        M3<R> T2{};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) T2.m[i][j] += H.m[i][k] * PP.m[k][j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) VV.m[i][j] += hh * T2.m[i][j];
    };
    V3<R> oldp = p, oldv = v;
    M3<R> oldP = P, oldV = V;
    for (int it = 0; it < 100000; it++) {
        oldp = p;
        oldv = v;
        oldP = P;
        oldV = V;
        step(p, v, P, V, h);
        if (p.x < p2.x) {
            while (nbis > 0) {
                nbis--;
                p = oldp;
                v = oldv;
                P = oldP;
                V = oldV;
                h = h / 2;
                step(p, v, P, V, h);
                if (p.x > p2.x) {
                    oldp = p;
                    oldv = v;
                    oldP = P;
                    oldV = V;
                }
            }
            break;
        }
    }
    const V3<R> t = v / RIF(S, p, ts);
    /* dQdV = (I - t dgdQ^T / (dgdQ . t)) P, dgdQ = (-1, 0, 0); derror = dQdV^T */
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) derror.m[j][i] = P.m[i][j] - t[i] * P.m[0][j] / t.x;
    error = p - p2;
}

/* ------------------------------------------------------------- counters */
struct Counters {
    uint64_t photons = 0, steps = 0, scatters = 0, nee = 0, direct = 0, faces = 0, segments = 0, aborted = 0;
};
/* A march that takes more than this many face events is abandoned (treated
 * like a cut march); the reference would loop without bound. */
static const uint32_t kMaxFaceEvents = 1u << 20;

/* ----------------------------------------------------------- transport */
template <class R>
struct Tracer {
    const Scene<R> &S;
    Counters &C;
    Sampler *smp;
    TimeState<R> ts;
    /* Dead-path cutoff (off unless requested): a march whose optical length
     * exceeds Lcut can no longer reach the time gate, see dead_path_cut(). */
    R Lcut;
    bool dead;

    R rnd() { return (R)smp->next(); }

    /* Each march that may be cut short draws from its own Philox sub-stream:
     * 0 = main path, 1 = direct trace, 1 + depth = NEE at that vertex, so a
     * cut march never shifts the draws of any other segment. */
    Sampler substream(uint32_t id) const {
        Sampler s = *smp;
        s.stream = id;
        s.draw = 0;
        return s;
    }

    /* block.cpp:11-23 */
    bool inside(const V3<R> &p) const {
        for (int i = 0; i < 3; i++) {
            if (!(p[i] - S.blockL[i] > -K<R>::EPS)) return false;
            if (!(S.blockR[i] - p[i] > -K<R>::EPS)) return false;
        }
        return true;
    }

    /* scene.cpp:323-330 (OMEGA_TRACKING undefined) */
    void er_step(V3<R> &p, V3<R> &d, R h) {
        ++C.steps;
        d += (R(0.5) * h) * dRIF(S, p, ts);
        p += (h * d) / RIF(S, p, ts);
        d += (R(0.5) * h) * dRIF(S, p, ts);
    }

    /* scene.cpp:385-425 */
    void traceTillBlock(V3<R> &p, V3<R> &d, R dist, R &disy, R &L) {
        ++C.segments;
        V3<R> oldp, oldd;
        R distance = 0;
        const R q = dist / (R)S.d.er_stepsize + 1;
        long maxsteps = q > (R)4.0e18 ? (long)4000000000000000000L : (long)q;
        long precision = S.d.precision;
        R h = (R)S.d.er_stepsize;
        for (long i = 0; i < maxsteps; i++) {
            oldp = p;
            oldd = d;
            er_step(p, d, h);
            if (!inside(p) || (distance + h) > dist) {
                precision--;
                if (precision < 0) break;
                p = oldp;
                d = oldd;
                h = h / 10;
                i = 0;
                maxsteps = 11;
            } else {
                distance += h;
                L += h * RIF(S, p, ts);
                if (L > Lcut) {
                    dead = true;
                    break;
                }
            }
        }
        disy = distance;
    }

    /* util.h:42-88 */
    static R fresnelDielectric(R cosI, R cosT, R eta) {
        if (std::abs(eta - 1) < K<R>::EPS * std::max((R)1, std::abs(eta))) return 0;
        const R Rs = (cosI - eta * cosT) / (cosI + eta * cosT);
        const R Rp = (cosT - eta * cosI) / (cosT + eta * cosI);
        return R(0.5) * (Rs * Rs + Rp * Rp);
    }
    static void reflect(const V3<R> &a, const V3<R> &n, V3<R> &b) { b = -(R(2) * dot(a, n)) * n + a; }
    static bool refract(const V3<R> &a, const V3<R> &n, R eta, V3<R> &b) {
        V3<R> q = dot(a, n) * n;
        const V3<R> p = (a - q) / eta;
        if (p.length() > R(1)) {
            reflect(a, n, b);
            return false;
        }
        q.normalize();
        const R t = R(1) - p.lengthSquared();
        q *= (t > 0 ? std::sqrt(t) : R(0));
        b = p + q;
        return true;
    }

    /* bsdf.cpp:25-58 (ior1 = 1, ior2 = scene ior) */
    void bsdf_sample(const V3<R> &in, const V3<R> &n, V3<R> &out) {
        ++C.faces;
        const R ior1 = 1, ior2 = (R)S.d.ior;
        if (std::abs(ior1 - ior2) < K<R>::EPS) {
            out = in;
            return;
        }
        const R eta = dot(in, n) < -K<R>::EPS ? ior2 / ior1 : ior1 / ior2;
        V3<R> outT;
        if (!refract(in, n, eta, outT)) {
            out = outT;
        } else {
            V3<R> outR;
            reflect(in, n, outR);
            const R cosI = std::abs(dot(n, in)), cosT = std::abs(dot(n, outT));
            const R F = fresnelDielectric(cosI, cosT, eta);
            out = (rnd() < F ? outR : outT);
        }
    }

    /* scene.cpp:655-676 / 736-757: outward normal of the nearest face */
    V3<R> nearest_face_normal(const V3<R> &p) const {
        R minDiff = K<R>::MAXV, minDir = 0;
        int chosen = 3;
        for (int i = 0; i < 3; ++i) {
            R diff = std::abs(S.blockL[i] - p[i]);
            if (diff < minDiff) {
                minDiff = diff;
                chosen = i;
                minDir = -1;
            }
            diff = std::abs(S.blockR[i] - p[i]);
            if (diff < minDiff) {
                minDiff = diff;
                chosen = i;
                minDir = 1;
            }
        }
        V3<R> nrm(0, 0, 0);
        if (chosen < 3) nrm[chosen] = minDir;
        return nrm;
    }

    /* scene.cpp:622-703; returns 1 sensor reached, 0 lost, -1 LargeDist
     * exhausted, -2 cut as dead */
    int movePhotonTillSensor(V3<R> &p, V3<R> &d, R &distToSensor, R &L) {
        R LargeDist = (R)10000;
        R disy;
        dead = false;
        traceTillBlock(p, d, LargeDist, disy, L);
        if (dead) return -2;
        distToSensor = disy;
        LargeDist -= disy;
        uint32_t faces = 0;
        while (true) {
            if (LargeDist < 0) return -1;
            if (++faces > kMaxFaceEvents) {
                ++C.aborted;
                return -2;
            }
            const V3<R> norm = nearest_face_normal(p);
            if (std::abs(S.camDir.x - norm.x) < K<R>::EPS && std::abs(S.camDir.y - norm.y) < K<R>::EPS &&
                std::abs(S.camDir.z - norm.z) < K<R>::EPS)
                return 1;
            V3<R> d1;
            bsdf_sample(d, norm, d1); /* unnormalised d: reference quirk */
            if (dot(d1, norm) < R(0))
                d = d1;
            else
                return 0;
            traceTillBlock(p, d, LargeDist, disy, L);
            if (dead) return -2;
            distToSensor += disy;
            LargeDist -= disy;
        }
    }

    /* scene.cpp:706-780 */
    bool movePhoton(V3<R> &p, V3<R> &d, R dist, R &L) {
        R disy;
        dead = false;
        traceTillBlock(p, d, dist, disy, L);
        if (dead) return false;
        dist -= disy;
        uint32_t faces = 0;
        while (dist > K<R>::EPS) {
            if (++faces > kMaxFaceEvents) {
                ++C.aborted;
                return false;
            }
            const V3<R> norm = nearest_face_normal(p);
            const R magnitude = d.length();
            V3<R> d1;
            bsdf_sample(d / magnitude, norm, d1);
            if (dot(d1, norm) < R(0))
                d = d1 * magnitude;
            else
                return false;
            traceTillBlock(p, d, dist, disy, L);
            if (dead) return false;
            dist -= disy;
        }
        return true;
    }

    /* camera.cpp:53-75 with lens.cpp:69-97 */
    static bool lens_deflect(const float *origin, float aperture, float focal, V3<R> &pos, V3<R> &dir,
                             R &total) {
        (void)origin;
        R sq = pos.y * pos.y + pos.z * pos.z;
        if (sq > (R)aperture * (R)aperture) return false;
        const R invd = -1 / dir.x;
        dir.x = -(R)focal;
        dir.y = dir.y * invd * (R)focal - pos.y;
        dir.z = dir.z * invd * (R)focal - pos.z;
        total += (R)focal * invd - dir.length();
        dir.normalize();
        return true;
    }
    bool propagate_till_sensor(V3<R> &pos, V3<R> &dir, R &total) const {
        if (S.d.sensor_lens_active &&
            !lens_deflect(S.d.sensor_lens_origin, S.d.sensor_lens_aperture, S.d.sensor_lens_focal_length,
                          pos, dir, total))
            return false;
        const R dist = (S.camOrigin.x - pos.x) / dir.x;
        pos += dist * dir;
        total += dist;
        return true;
    }

    /* scene.cpp:782-829 */
    void addEnergyToImage(double *img, int X, int Y, int Z, const V3<R> &p, R pathlength, int depth,
                          R val) const {
        R x = dot(S.camH, p) - S.camOrigin.y;
        R y = dot(S.camV, p) - S.camOrigin.z;
        const R pl0 = (R)S.d.pathlength_range[0], pl1 = (R)S.d.pathlength_range[1];
        const bool nogate = (pl0 == -1) && (pl1 == -1);
        if (!(nogate || (pathlength > pl0 && pathlength < pl1))) return;
        x = (x / (R)S.d.view_plane[0] + R(0.5)) * (R)X;
        y = (y / (R)S.d.view_plane[1] + R(0.5)) * (R)Y;
        const R fx0 = std::floor(x), fy0 = std::floor(y);
        /* guard the int conversion (out-of-range floats are UB in C++) */
        if (!(fx0 > -4 && fx0 < (R)X + 4 && fy0 > -4 && fy0 < (R)Y + 4)) return;
        const int ix = (int)fx0, iy = (int)fy0;
        int iz;
        if (S.d.use_bounce_decomposition) {
            iz = depth;
        } else if (nogate) {
            iz = 0;
        } else {
            R z = pathlength - pl0;
            const R range = pl1 - pl0;
            z = (z / range) * (R)Z;
            const R fz = std::floor(z);
            if (!(fz > -4 && fz < (R)Z + 4)) return;
            iz = (int)fz;
        }
        const R fx = x - fx0, fy = y - fy0;
        auto add = [&](int xx, int yy, R v) {
            if (xx >= 0 && xx < X && yy >= 0 && yy < Y && iz >= 0 && iz < Z)
                img[(long)iz * X * Y + (long)yy * X + xx] += (double)v;
        };
        add(ix, iy, val * (R(1) - fx) * (R(1) - fy));
        add(ix + 1, iy, val * fx * (R(1) - fy));
        add(ix, iy + 1, val * (R(1) - fx) * fy);
        add(ix + 1, iy + 1, val * fx * fy);
    }

    /* phase.cpp:16-21 */
    R hg_f(const V3<R> &in, const V3<R> &out) const {
        const R g = (R)S.d.g;
        const R cosTheta = dot(in, out);
        return (R)(1.0 / (4.0 * 3.14159265358979323846)) * (R(1) - g * g) /
               std::pow(R(1) + g * g - R(2) * g * cosTheta, R(1.5));
    }
    /* util.h:17-27 */
    static void coordinateSystem(const V3<R> &a, V3<R> &b, V3<R> &c) {
        if (std::abs(a.x) > std::abs(a.y)) {
            const R invLen = R(1) / std::sqrt(a.x * a.x + a.z * a.z);
            c = V3<R>(a.z * invLen, R(0), -a.x * invLen);
        } else {
            const R invLen = R(1) / std::sqrt(a.y * a.y + a.z * a.z);
            c = V3<R>(R(0), a.z * invLen, -a.y * invLen);
        }
        b = cross(c, a);
    }
    /* phase.cpp:39-76 */
    void hg_sample(const V3<R> &in, V3<R> &out) {
        const R g = (R)S.d.g;
        const R sx = rnd(), sy = rnd();
        R cosTheta;
        if (std::abs(g) < K<R>::EPS) {
            cosTheta = 1 - 2 * sx;
        } else {
            const R sqrTerm = (1 - g * g) / (1 - g + 2 * g * sx);
            cosTheta = (1 + g * g - sqrTerm * sqrTerm) / (2 * g);
        }
        const R sinTheta = std::sqrt(std::max(R(0), R(1) - cosTheta * cosTheta));
        const R phi = (R)(2.0 * 3.14159265358979323846) * sy;
        const R sinPhi = std::sin(phi), cosPhi = std::cos(phi);
        V3<R> ax, ay;
        coordinateSystem(in, ax, ay);
        out = (sinTheta * cosPhi) * ax + (sinTheta * sinPhi) * ay + cosTheta * in;
    }

    /* medium.h:24-38 */
    R sigmaT() const { return (R)S.d.sigma_t <= K<R>::EPS ? R(0) : (R)S.d.sigma_t; }
    R mfp() const { return (R)S.d.sigma_t <= K<R>::EPS ? R(1) : R(1) / (R)S.d.sigma_t; }
    R albedo() const { return (R)S.d.sigma_t <= K<R>::EPS ? R(0) : (R)S.d.albedo; }
    /* photon.h:72-74 */
    R getMoveStep() { return -mfp() * std::log(rnd()); }

    /* exit-side weighting shared by directTracing / NEE
     * (photon.cpp:53-80, scene.cpp:897-929) */
    bool exit_to_sensor(V3<R> &p1, V3<R> &dirToSensor, R &L, R &foreshortening, R &fresnelWeight) {
        dirToSensor.normalize();
        V3<R> refr = dirToSensor;
        fresnelWeight = 1;
        const R ior = RIF(S, p1, ts);
        if (ior > R(1)) {
            refr.x = refr.x / ior;
            refr.normalize();
            if (S.d.flags & SIMPLER_FLAG_USE_FRESNEL)
                fresnelWeight = (R(1) - fresnelDielectric(dirToSensor.x, refr.x, R(1) / ior)) / ior / ior;
        }
        foreshortening = dot(refr, S.camDir) / dot(dirToSensor, S.camDir);
        R distanceToSensor = 0;
        if (!propagate_till_sensor(p1, refr, distanceToSensor)) return false;
        L += distanceToSensor;
        return true;
    }

    /* photon.cpp:40-87; returns false when the trace was cut as dead */
    bool directTracing(const V3<R> &p, const V3<R> &d, double *img, int X, int Y, int Z, R &L) {
        Sampler *main = smp;
        Sampler sub = substream(1);
        smp = &sub;
        V3<R> p1 = p, d1 = d;
        R distToSensor;
        const int res = movePhotonTillSensor(p1, d1, distToSensor, L);
        smp = main;
        if (res == -2) return false;
        if (res != 1) return true;
        R fore, fw;
        if (!exit_to_sensor(p1, d1, L, fore, fw)) return true;
        const R val = S.weight * std::exp(-sigmaT() * distToSensor) * fw;
        ++C.direct;
        addEnergyToImage(img, X, Y, Z, p1, L, 0, val);
        return true;
    }

    /* scene.cpp:20-39 with warp.cpp:22-34 */
    V3<R> sampleRandomDirection() {
        const R s0 = rnd(), s1 = rnd();
        return hemisphere_dir(s0, s1);
    }
    static V3<R> hemisphere_dir(R s0, R s1) {
        const R z = s0;
        const R tmp = std::sqrt(R(1) - z * z);
        const R ang = R(2) * K<R>::PI * s1;
        const V3<R> h(std::cos(ang) * tmp, std::sin(ang) * tmp, z);
        return V3<R>(-h.z, h.x, h.y);
    }

    /* scene.cpp:866-946 (distTravelled by value) */
    void addEnergyInParticle(double *img, int X, int Y, int Z, const V3<R> &p, const V3<R> &d,
                             R distTravelled, int depth) {
        if ((double)(p.x - S.camOrigin.x) < 1e-4) return;
        Sampler *main = smp;
        Sampler sub = substream(1u + (uint32_t)depth);
        smp = &sub;
        V3<R> p1 = p;
        V3<R> dirToSensor = sampleRandomDirection();
        dirToSensor *= RIF(S, p1, ts);
        R distToSensor;
        const int res = movePhotonTillSensor(p1, dirToSensor, distToSensor, distTravelled);
        smp = main;
        if (res != 1) return;
        R fore, fw;
        R L = distTravelled;
        /* dirToSensor normalised inside exit_to_sensor, as scene.cpp:897 */
        V3<R> dts = dirToSensor;
        if (!exit_to_sensor(p1, dts, L, fore, fw)) return;
        const R val = S.weight * (R)(2 * K<R>::PI) * std::exp(-sigmaT() * distToSensor) *
                      hg_f(d / d.length(), dts) * fore * fw;
        ++C.nee;
        addEnergyToImage(img, X, Y, Z, p1, L, depth, val);
    }

    /* photon.cpp:19-38 */
    bool scatterOnce(V3<R> &p, V3<R> &d, R &dist, R &L) {
        const R a = albedo();
        if ((a > R(0)) && ((a >= R(1)) || (rnd() < a))) {
            const R magnitude = d.length();
            V3<R> d1;
            hg_sample(d / magnitude, d1);
            ++C.scatters;
            d = magnitude * d1;
            dist = getMoveStep();
            return movePhoton(p, d, dist, L);
        }
        dist = 0;
        return false;
    }

    /* photon.cpp:89-136 */
    void scatter(const V3<R> &p, const V3<R> &d, double *img, int X, int Y, int Z, R &L) {
        const R a = albedo();
        if ((a > R(0)) && ((a >= R(1)) || (rnd() < a))) {
            V3<R> pos(p), dir(d);
            R dist = getMoveStep();
            if (!movePhoton(pos, dir, dist, L)) return;
            int depth = 1;
            const R totalDist = dist;
            const int maxDepth = S.d.max_depth;
            const R maxPL = (R)S.d.max_pathlength;
            while ((maxDepth < 0 || depth <= maxDepth) && (maxPL < 0 || totalDist <= maxPL)) {
                if (S.d.use_angular_sampling) addEnergyInParticle(img, X, Y, Z, pos, dir, L, depth);
                if (!scatterOnce(pos, dir, dist, L)) break;
                ++depth;
            }
        }
    }

    /* pmf.h:109-121 */
    int cdf_sample(R u) const { return cdf_search(S.cdf.data(), (long)S.cdf.size(), (float)u); }
    static int cdf_search(const float *c, long n, float u) {
        const long pos = std::lower_bound(c, c + n, u) - c;
        long index = std::min(n - 2, std::max(0L, pos - 1));
        while (c[index + 1] - c[index] == 0 && index < n - 2) ++index;
        return (int)index;
    }

    /* scene.cpp:107-137 + area_textured_source.cpp:107-116 + lens.cpp:19-33 */
    bool sample_ray(V3<R> &pos, V3<R> &dir, R &total) {
        pos = v3<R>(S.d.light_origin);
        const int pixel = cdf_sample(rnd());
        const int py = pixel / S.tex_x; /* image.h:78-82 */
        const int px = pixel - py * S.tex_x;
        const int pix[2] = {px, py};
        for (int it = 1; it < 3; ++it)
            pos[it] += -(R)S.d.light_plane[it - 1] / R(2) + (R)pix[it - 1] * S.pixsize[it - 1] +
                       rnd() * S.pixsize[it - 1];
        const R z = rnd() * (1 - S.ct) + S.ct;
        const R zt = std::sqrt(1 - z * z);
        const R phi = rnd() * 2 * K<R>::PI;
        dir = V3<R>(-z, zt * std::cos(phi), zt * std::sin(phi));
        /* Lens::propagate_till_lens */
        const R dist = -(pos.x - (R)S.d.emitter_lens_origin[0]) / dir.x;
        pos += dist * dir;
        total += dist;
        if (S.d.emitter_lens_active)
            return lens_deflect(S.d.emitter_lens_origin, S.d.emitter_lens_aperture,
                                S.d.emitter_lens_focal_length, pos, dir, total);
        return true;
    }

    /* photon.cpp:324 (Bessel: clamp(sin(.)); the sources field ignores the
     * amplitude and uses the phase tau itself) */
    void sample_time() {
        const R u = rnd();
        ts.tau = (R)S.d.phi_min + ((R)S.d.phi_max - (R)S.d.phi_min) * u;
        const R m = S.n_maxScaling;
        ts.scaling = std::max(std::min(std::sin(ts.tau), m), -m);
    }

    /* photon.cpp:300-335 (one loop iteration) */
    void photon(double *img, int X, int Y, int Z) {
        ++C.photons;
        V3<R> pos, dir;
        R total = 0;
        if (!sample_ray(pos, dir, total)) return;
        sample_time();
        dir *= RIF(S, pos, ts);
        R L = total;
        if (S.d.use_direct) {
            R Ld = total;
            const bool alive = directTracing(pos, dir, img, X, Y, Z, Ld);
            if (S.d.flags & SIMPLER_FLAG_DIRECT_LEAK) {
                if (!alive) return; /* the leaked length is already dead */
                L = Ld;
            }
        }
        scatter(pos, dir, img, X, Y, Z, L);
    }
};

/* Dead-path cutoff length.  Optical length only grows along a march
 * (n > 0), and the camera step at the sensor face (camera.cpp:63-73) can
 * shorten it by at most the step overshoot plus the transverse distance the
 * splat point may travel and still land on the sensor; margin bounds both.
 * A march longer than pathlength_range.y + margin can never pass the time
 * gate of addEnergyToImage (scene.cpp:793-794).  Disabled (infinite) without
 * a time gate, with an active sensor lens, or when n > 0 is not guaranteed. */
template <class R>
static R dead_path_cut(const Scene<R> &S) {
    const simpler_scene_desc &d = S.d;
    if (d.pathlength_range[0] == -1 && d.pathlength_range[1] == -1) return (R)INFINITY;
    if (d.sensor_lens_active) return (R)INFINITY;
    double nlow = d.n_o;
    if (d.field_kind == SIMPLER_FIELD_BESSEL) {
        nlow = d.n_o - std::abs((double)d.n_max * (double)S.n_maxScaling);
    } else if (d.field_kind == SIMPLER_FIELD_SOURCES) {
        double s = 0, w = 0;
        for (int c = 0; c < S.n_coeff; c++) w += std::abs((double)S.coeff_w[c]);
        for (int i = 0; i < S.n_src; i++) {
            double dist2 = 0;
            for (int k = 0; k < 3; k++) {
                const double x = S.src_pos[3 * i + k];
                const double lo = d.block_l[k], hi = d.block_r[k];
                const double e = x < lo ? lo - x : (x > hi ? x - hi : 0.0);
                dist2 += e * e;
            }
            if (dist2 <= 0) return (R)INFINITY;
            s += std::abs((double)S.src_amp[i]) / std::sqrt(dist2);
        }
        nlow = d.n_o - w * s;
    } else if (S.has_spline) {
        const long n = (long)S.spl.N[0] * S.spl.N[1] * S.spl.N[2];
        nlow = *std::min_element(S.spl.coeff, S.spl.coeff + n);
        nlow = std::min(nlow, (double)d.n_o);
    }
    if (!(nlow > 0)) return (R)INFINITY;
    auto norm3 = [](double a, double b, double c) { return std::sqrt(a * a + b * b + c * c); };
    const double margin =
        2.0 * (std::sqrt((double)d.view_plane[0] * d.view_plane[0] + (double)d.view_plane[1] * d.view_plane[1]) +
               norm3(d.block_r[0] - d.block_l[0], d.block_r[1] - d.block_l[1], d.block_r[2] - d.block_l[2]) +
               norm3(d.view_origin[0], d.view_origin[1], d.view_origin[2])) +
        (double)d.er_stepsize;
    return (R)((double)d.pathlength_range[1] + margin);
}

/* busy (optional): per-thread seconds spent tracing photons, summed over
 * the threads (CPU-seconds of the render, independent of load balance). */
template <class R>
static int render(const Scene<R> &S, int64_t n, uint64_t seed, uint64_t offset, int cutoff,
                  double *img, int X, int Y, int Z, simpler_stats *st, int threads, double *busy) {
    const R Lcut = cutoff ? dead_path_cut(S) : (R)INFINITY;
    const long npix = (long)X * Y * Z;
    int nt = 1;
#ifdef _OPENMP
    nt = threads > 0 ? threads : omp_get_num_procs();
#endif
    std::vector<std::vector<double>> imgs(nt, std::vector<double>(npix, 0.0));
    std::vector<Counters> cs(nt);
    std::vector<double> tbusy(nt, 0.0);
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        const double t0 = omp_get_wtime();
#endif
        Sampler smp;
        smp.key[0] = (uint32_t)seed;
        smp.key[1] = (uint32_t)(seed >> 32);
        smp.photon = offset + (uint64_t)i;
        smp.stream = 0;
        smp.draw = 0;
        Tracer<R> t{S, cs[tid], &smp, {0, 0}, Lcut, false};
        t.photon(imgs[tid].data(), X, Y, Z);
#ifdef _OPENMP
        tbusy[tid] += omp_get_wtime() - t0;
#endif
    }
    if (busy) {
        *busy = 0.0;
        for (double b : tbusy) *busy += b;
    }
    for (int t = 0; t < nt; ++t)
        for (long k = 0; k < npix; ++k) img[k] += imgs[t][k];
    if (st) {
        std::memset(st, 0, sizeof(*st));
        for (auto &c : cs) {
            st->photons += c.photons;
            st->er_steps += c.steps;
            st->scatters += c.scatters;
            st->nee_splats += c.nee;
            st->direct_splats += c.direct;
            st->face_events += c.faces;
            st->segments += c.segments;
            st->aborted += c.aborted;
        }
    }
    return 0;
}

template <class R>
static int trace_rays(const Scene<R> &S, int64_t n, uint64_t seed, uint64_t offset, double *out_p,
                      double *out_d, double *out_L, int32_t *status, int threads) {
    int nt = 1;
#ifdef _OPENMP
    nt = threads > 0 ? threads : omp_get_num_procs();
#endif
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        Counters c;
        Sampler smp;
        smp.key[0] = (uint32_t)seed;
        smp.key[1] = (uint32_t)(seed >> 32);
        smp.photon = offset + (uint64_t)i;
        smp.stream = 0;
        smp.draw = 0;
        Tracer<R> t{S, c, &smp, {0, 0}, (R)INFINITY, false};
        V3<R> pos, dir;
        R total = 0;
        int stt = 0;
        if (t.sample_ray(pos, dir, total)) {
            t.sample_time();
            dir *= RIF(S, pos, t.ts);
            R dist;
            Sampler sub = t.substream(1); /* the direct trace's sub-stream */
            t.smp = &sub;
            stt = t.movePhotonTillSensor(pos, dir, dist, total);
            t.smp = &smp;
            dir.normalize();
        }
        out_p[3 * i] = pos.x;
        out_p[3 * i + 1] = pos.y;
        out_p[3 * i + 2] = pos.z;
        out_d[3 * i] = dir.x;
        out_d[3 * i + 1] = dir.y;
        out_d[3 * i + 2] = dir.z;
        out_L[i] = total;
        status[i] = stt;
    }
    return 0;
}

/* pmf.h:41-99 with the texture of area_textured_source.cpp:30-45 */
static void build_cdf(const float *tex, int n, std::vector<float> &cdf) {
    cdf.assign(1, 0.0f);
    cdf.reserve(n + 1);
    for (int i = 0; i < n; ++i) cdf.push_back(cdf.back() + tex[i]);
    const float sum = cdf.back();
    if (sum > 0) {
        const float norm = 1.0f / sum;
        for (size_t i = 1; i < cdf.size(); ++i) cdf[i] *= norm;
        cdf.back() = 1.0f;
    }
}

} /* namespace oracle */

using namespace oracle;

/* ------------------------------------------------------------ C entries */
struct OracleSceneHandle {
    simpler_scene_desc desc;
    std::vector<float> cdf;
    int tex_x, tex_y;
    int n_src, n_coeff;
    std::vector<float> pos, amp, phase, coeffw;
    std::vector<float> spline;
    int N[3];
    float xmin[3], xmax[3];
};

template <class R>
static void fill_scene(Scene<R> &S, const OracleSceneHandle &h, int64_t n_total) {
    scene_init(S, h.desc, h.cdf.data(), (int)h.cdf.size(), h.tex_x, h.tex_y, n_total);
    S.n_src = h.n_src;
    S.n_coeff = h.n_coeff;
    S.src_pos = h.pos;
    S.src_amp = h.amp;
    S.src_phase = h.phase;
    S.coeff_w = h.coeffw;
    if (!h.spline.empty()) {
        S.has_spline = true;
        for (int i = 0; i < 3; i++) {
            S.spl.N[i] = h.N[i];
            S.spl.xmin[i] = (R)h.xmin[i];
            S.spl.xres[i] = (R)(h.N[i] - 1) / ((R)h.xmax[i] - (R)h.xmin[i]);
        }
        S.spl.coeff = h.spline.data();
    }
}

extern "C" {

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox4x32_10(ctr, key, out);
}

float oracle_uniform(uint64_t seed, uint64_t photon, uint32_t draw, uint32_t stream) {
    Sampler s;
    s.key[0] = (uint32_t)seed;
    s.key[1] = (uint32_t)(seed >> 32);
    s.photon = photon;
    s.stream = stream;
    s.draw = draw;
    return s.next();
}

double oracle_besselj(int m, double x) { return besselj(m, x); }

void *oracle_scene_create(const simpler_scene_desc *desc, const float *texture, int tex_x, int tex_y) {
    OracleSceneHandle *h = new OracleSceneHandle();
    h->desc = *desc;
    static const float white = 1.0f;
    if (!texture || tex_x <= 0 || tex_y <= 0) {
        texture = &white;
        tex_x = tex_y = 1;
    }
    h->tex_x = tex_x;
    h->tex_y = tex_y;
    build_cdf(texture, tex_x * tex_y, h->cdf);
    h->n_src = 0;
    h->n_coeff = 0;
    return h;
}

void oracle_scene_destroy(void *h) { delete (OracleSceneHandle *)h; }

void oracle_scene_update(void *hv, const simpler_scene_desc *desc) { ((OracleSceneHandle *)hv)->desc = *desc; }

int oracle_scene_get_cdf(void *hv, float *out, int n) {
    OracleSceneHandle *h = (OracleSceneHandle *)hv;
    const int m = std::min(n, (int)h->cdf.size());
    std::memcpy(out, h->cdf.data(), sizeof(float) * m);
    return (int)h->cdf.size();
}

void oracle_scene_set_sources(void *hv, int n_src, const float *pos, const float *amp, const float *phase,
                              int n_coeff, const float *coeffw) {
    OracleSceneHandle *h = (OracleSceneHandle *)hv;
    h->n_src = n_src;
    h->n_coeff = n_coeff;
    h->pos.assign(pos, pos + 3 * n_src);
    h->amp.assign(amp, amp + n_src);
    h->phase.assign(phase, phase + n_src);
    h->coeffw.assign(coeffw, coeffw + n_coeff);
}

void oracle_scene_set_spline(void *hv, const float *coeffs, const int32_t dims[3], const float xmin[3],
                             const float xmax[3]) {
    OracleSceneHandle *h = (OracleSceneHandle *)hv;
    const long n = (long)dims[0] * dims[1] * dims[2];
    h->spline.assign(coeffs, coeffs + n);
    for (int i = 0; i < 3; i++) {
        h->N[i] = dims[i];
        h->xmin[i] = xmin[i];
        h->xmax[i] = xmax[i];
    }
}

/* precision: 32 = float restatement, 64 = double */
/* cutoff = 1 enables the dead-path cutoff (identical image, less work);
 * 0 runs the reference algorithm in full. */
int oracle_render(void *hv, int precision, int64_t n_photons, int64_t n_total, uint64_t seed,
                  uint64_t offset, int cutoff, double *image, int X, int Y, int Z, simpler_stats *st,
                  int threads, double *busy_seconds) {
    const OracleSceneHandle &h = *(OracleSceneHandle *)hv;
    if (precision == 64) {
        Scene<double> S;
        fill_scene(S, h, n_total);
        return render(S, n_photons, seed, offset, cutoff, image, X, Y, Z, st, threads, busy_seconds);
    }
    Scene<float> S;
    fill_scene(S, h, n_total);
    return render(S, n_photons, seed, offset, cutoff, image, X, Y, Z, st, threads, busy_seconds);
}

/* Single path functions in FP64, for the distribution / closed-form pins. */
int oracle_lens_deflect(double aperture, double focal, double pos[3], double dir[3], double *total) {
    V3<double> p(pos[0], pos[1], pos[2]), d(dir[0], dir[1], dir[2]);
    double t = *total;
    const bool ok = Tracer<double>::lens_deflect(nullptr, (float)aperture, (float)focal, p, d, t);
    for (int i = 0; i < 3; ++i) {
        pos[i] = p[i];
        dir[i] = d[i];
    }
    *total = t;
    return ok ? 1 : 0;
}

void oracle_hemisphere(int n, const double *u0, const double *u1, double *out) {
    for (int i = 0; i < n; ++i) {
        const V3<double> d = Tracer<double>::hemisphere_dir(u0[i], u1[i]);
        for (int k = 0; k < 3; ++k) out[3 * i + k] = d[k];
    }
}

int oracle_cdf_sample(void *hv, int n, const float *u, int *out) {
    const OracleSceneHandle &h = *(OracleSceneHandle *)hv;
    const long m = (long)h.cdf.size();
    for (int i = 0; i < n; ++i) out[i] = Tracer<double>::cdf_search(h.cdf.data(), m, u[i]);
    return 0;
}

double oracle_dead_path_cut(void *hv) {
    const OracleSceneHandle &h = *(OracleSceneHandle *)hv;
    Scene<double> S;
    fill_scene(S, h, 1);
    return dead_path_cut(S);
}

int oracle_trace_rays(void *hv, int precision, int64_t n, uint64_t seed, uint64_t offset, double *out_p,
                      double *out_d, double *out_L, int32_t *status, int threads) {
    const OracleSceneHandle &h = *(OracleSceneHandle *)hv;
    if (precision == 64) {
        Scene<double> S;
        fill_scene(S, h, n);
        return trace_rays(S, n, seed, offset, out_p, out_d, out_L, status, threads);
    }
    Scene<float> S;
    fill_scene(S, h, n);
    return trace_rays(S, n, seed, offset, out_p, out_d, out_L, status, threads);
}

/* n and grad n at points (for field tests); scaling and tau per call */
int oracle_field_eval(void *hv, int precision, int npts, const double *pts, double scaling, double tau,
                      double *out_n, double *out_g) {
    const OracleSceneHandle &h = *(OracleSceneHandle *)hv;
    for (int i = 0; i < npts; i++) {
        if (precision == 64) {
            Scene<double> S;
            fill_scene(S, h, 1);
            TimeState<double> ts{scaling, tau};
            V3<double> p(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
            out_n[i] = RIF(S, p, ts);
            V3<double> g = dRIF(S, p, ts);
            out_g[3 * i] = g.x;
            out_g[3 * i + 1] = g.y;
            out_g[3 * i + 2] = g.z;
        } else {
            Scene<float> S;
            fill_scene(S, h, 1);
            TimeState<float> ts{(float)scaling, (float)tau};
            V3<float> p((float)pts[3 * i], (float)pts[3 * i + 1], (float)pts[3 * i + 2]);
            out_n[i] = RIF(S, p, ts);
            V3<float> g = dRIF(S, p, ts);
            out_g[3 * i] = g.x;
            out_g[3 * i + 1] = g.y;
            out_g[3 * i + 2] = g.z;
        }
    }
    return 0;
}

/* Hessian of n at points (H row-major, 9 per point). */
int oracle_field_hessian(void *hv, int npts, const double *pts, double scaling, double tau, double *out_h) {
    const OracleSceneHandle &h = *(OracleSceneHandle *)hv;
    Scene<double> S;
    fill_scene(S, h, 1);
    TimeState<double> ts{scaling, tau};
    for (int i = 0; i < npts; i++) {
        const M3<double> H = HRIF(S, V3<double>(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]), ts);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) out_h[9 * i + 3 * a + b] = H.m[a][b];
    }
    return 0;
}

/* computefdfNEE for n rays (FP64): err 3 and derror 9 (row-major) per ray. */
int oracle_nee_connect(void *hv, int n, const double *p1, const double *v1, const double *p2, double scaling,
                       double tau, double *err, double *jac) {
    const OracleSceneHandle &h = *(OracleSceneHandle *)hv;
    Scene<double> S;
    fill_scene(S, h, 1);
    TimeState<double> ts{scaling, tau};
    const int nbis = (int)std::ceil((double)S.d.precision / std::log10(2.0)); /* scene.cpp:520 */
    for (int i = 0; i < n; i++) {
        V3<double> e;
        M3<double> J;
        nee_connect(S, ts, V3<double>(v1[3 * i], v1[3 * i + 1], v1[3 * i + 2]),
                    V3<double>(p1[3 * i], p1[3 * i + 1], p1[3 * i + 2]),
                    V3<double>(p2[3 * i], p2[3 * i + 1], p2[3 * i + 2]), nbis, e, J);
        err[3 * i] = e.x;
        err[3 * i + 1] = e.y;
        err[3 * i + 2] = e.z;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) jac[9 * i + 3 * a + b] = J.m[a][b];
    }
    return 0;
}

/* One ER step pair (for integrator tests): runs nsteps leapfrog steps of
 * size h from (p, v) with no boundary logic. */
int oracle_er_steps(void *hv, int precision, double *p, double *v, double h, long nsteps, double scaling,
                    double tau) {
    const OracleSceneHandle &hh = *(OracleSceneHandle *)hv;
    Counters c;
    Sampler smp{};
    if (precision == 64) {
        Scene<double> S;
        fill_scene(S, hh, 1);
        Tracer<double> t{S, c, &smp, {scaling, tau}, INFINITY, false};
        V3<double> P(p[0], p[1], p[2]), Vv(v[0], v[1], v[2]);
        for (long i = 0; i < nsteps; i++) t.er_step(P, Vv, h);
        p[0] = P.x; p[1] = P.y; p[2] = P.z;
        v[0] = Vv.x; v[1] = Vv.y; v[2] = Vv.z;
    } else {
        Scene<float> S;
        fill_scene(S, hh, 1);
        Tracer<float> t{S, c, &smp, {(float)scaling, (float)tau}, INFINITY, false};
        V3<float> P((float)p[0], (float)p[1], (float)p[2]), Vv((float)v[0], (float)v[1], (float)v[2]);
        for (long i = 0; i < nsteps; i++) t.er_step(P, Vv, (float)h);
        p[0] = P.x; p[1] = P.y; p[2] = P.z;
        v[0] = Vv.x; v[1] = Vv.y; v[2] = Vv.z;
    }
    return 0;
}

/* Sampling primitives driven by explicit uniforms (distribution tests). */
void oracle_hg_sample(double g, const double in[3], double u1, double u2, double out[3]) {
    /* phase.cpp:39-76 with the two uniforms given */
    double cosTheta;
    if (std::abs(g) < 1.19209290e-07) {
        cosTheta = 1 - 2 * u1;
    } else {
        const double sq = (1 - g * g) / (1 - g + 2 * g * u1);
        cosTheta = (1 + g * g - sq * sq) / (2 * g);
    }
    const double sinTheta = std::sqrt(std::max(0.0, 1 - cosTheta * cosTheta));
    const double phi = 2 * 3.14159265358979323846 * u2;
    V3<double> a(in[0], in[1], in[2]), ax, ay;
    Tracer<double>::coordinateSystem(a, ax, ay);
    V3<double> o = (sinTheta * std::cos(phi)) * ax + (sinTheta * std::sin(phi)) * ay + cosTheta * a;
    out[0] = o.x;
    out[1] = o.y;
    out[2] = o.z;
}

double oracle_hg_f(double g, double cosTheta) {
    return (1.0 / (4.0 * 3.14159265358979323846)) * (1 - g * g) / std::pow(1 + g * g - 2 * g * cosTheta, 1.5);
}

/* bsdf sample for a given uniform: returns 1 reflected, 0 refracted; out =
 * direction */
int oracle_bsdf_sample(double ior, const double in[3], const double n[3], double u, double out[3],
                       double *fresnelR) {
    V3<double> a(in[0], in[1], in[2]), nn(n[0], n[1], n[2]), o;
    *fresnelR = 1.0;
    if (std::abs(1.0 - ior) < 1.19209290e-07) {
        out[0] = a.x; out[1] = a.y; out[2] = a.z;
        *fresnelR = 0;
        return 0;
    }
    const double eta = dot(a, nn) < -1.19209290e-07 ? ior : 1.0 / ior;
    V3<double> outT;
    int refl;
    if (!Tracer<double>::refract(a, nn, eta, outT)) {
        o = outT;
        refl = 1;
    } else {
        V3<double> outR;
        Tracer<double>::reflect(a, nn, outR);
        const double F = Tracer<double>::fresnelDielectric(std::abs(dot(nn, a)), std::abs(dot(nn, outT)), eta);
        *fresnelR = F;
        refl = u < F;
        o = refl ? outR : outT;
    }
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
    return refl;
}

/* Spline entry points for the golden-vector tests (double). */
void oracle_spline_kernel(int d, int n, const double *x, double *out) {
    for (int i = 0; i < n; i++) out[i] = bs_kernel(d, x[i]);
}
void oracle_spline_build1d(const double *data, int n, double *out) { build1d(data, 0, 1, n, out); }
void oracle_spline_eval1d(const double *coeff, int n, double xmin, double xmax, int d, int npts,
                          const double *t, double *out) {
    const double xres = (n - 1) / (xmax - xmin);
    for (int k = 0; k < npts; k++) {
        const double x = (t[k] - xmin) * xres;
        double v = 0;
        for (int i = (int)std::ceil(x - 2); i <= (int)std::floor(x + 2); i++)
            v += coeff[mirror_index(i, n)] * bs_kernel(d, x - i);
        if (d == 1) v *= xres;
        if (d == 2) v *= xres * xres;
        out[k] = v;
    }
}
/* spline.h:716-737 (build2d): rows then columns */
void oracle_spline_build2d(const double *data, const int N[2], double *coeff) {
    std::vector<double> temp(std::max(N[0], N[1]));
    for (int i = 0; i < N[0]; i++) {
        build1d(data, i, N[0], N[1], temp.data());
        for (int j = 0; j < N[1]; j++) coeff[j * N[0] + i] = temp[j];
    }
    for (int i = 0; i < N[1]; i++) {
        build1d(coeff, (long)i * N[0], 1, N[0], temp.data());
        for (int j = 0; j < N[0]; j++) coeff[i * N[0] + j] = temp[j];
    }
}
/* spline.h:630-657 (value2d<DX,DY>) */
void oracle_spline_eval2d(const double *coeff, const int N[2], const double xmin[2], const double xmax[2],
                          int dx, int dy, int npts, const double *pts, double *out) {
    double xres[2];
    for (int i = 0; i < 2; i++) xres[i] = (N[i] - 1) / (xmax[i] - xmin[i]);
    for (int k = 0; k < npts; k++) {
        const double x0 = (pts[2 * k] - xmin[0]) * xres[0], x1 = (pts[2 * k + 1] - xmin[1]) * xres[1];
        double v = 0;
        for (int i1 = (int)std::ceil(x0 - 2); i1 <= (int)std::floor(x0 + 2); i1++)
            for (int i2 = (int)std::ceil(x1 - 2); i2 <= (int)std::floor(x1 + 2); i2++)
                v += coeff[mirror_index(i1, N[0]) + mirror_index(i2, N[1]) * N[0]] * bs_kernel(dx, x0 - i1) *
                     bs_kernel(dy, x1 - i2);
        if (dx == 1) v *= xres[0];
        if (dx == 2) v *= xres[0] * xres[0];
        if (dy == 1) v *= xres[1];
        if (dy == 2) v *= xres[1] * xres[1];
        out[k] = v;
    }
}
void oracle_spline_build3d(const double *data, const int N[3], double *coeff) { build3d(data, N, coeff); }
void oracle_spline_eval3d(const double *coeff, const int N[3], const double xmin[3], const double xmax[3],
                          int dx, int dy, int dz, int npts, const double *pts, double *out) {
    /* value3d with mirror wrap, in double, coefficients in double */
    double xres[3];
    for (int i = 0; i < 3; i++) xres[i] = (N[i] - 1) / (xmax[i] - xmin[i]);
    for (int k = 0; k < npts; k++) {
        double x[3];
        for (int i = 0; i < 3; i++) x[i] = (pts[3 * k + i] - xmin[i]) * xres[i];
        double v = 0;
        for (int i1 = (int)std::ceil(x[0] - 2); i1 <= (int)std::floor(x[0] + 2); i1++)
            for (int i2 = (int)std::ceil(x[1] - 2); i2 <= (int)std::floor(x[1] + 2); i2++)
                for (int i3 = (int)std::ceil(x[2] - 2); i3 <= (int)std::floor(x[2] + 2); i3++)
                    v += coeff[mirror_index(i1, N[0]) +
                               (long)N[0] * (mirror_index(i2, N[1]) + (long)N[1] * mirror_index(i3, N[2]))] *
                         bs_kernel(dx, x[0] - i1) * bs_kernel(dy, x[1] - i2) * bs_kernel(dz, x[2] - i3);
        const int d[3] = {dx, dy, dz};
        for (int i = 0; i < 3; i++) {
            if (d[i] == 1) v *= xres[i];
            if (d[i] == 2) v *= xres[i] * xres[i];
        }
        out[k] = v;
    }
}

} /* extern "C" */
