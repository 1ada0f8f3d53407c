// fields.h -- refractive-index fields on gfx950: value and gradient fused in
// one evaluation (the leapfrog reuses the end-of-step evaluation as the next
// step's start, so one call per ER step).
//
//   Bessel:  n = n0 + A s J_m(k r) cos(m phi)       (reference scene.cpp:205-257)
//   Sources: n = n0 + sum_c w_c sum_i a_i cos(c k r_i + phi_i - tc tau)/r_i
//                                                    (scene.cpp:140-195, generalised)
//   Spline:  n = tricubic B-spline of the node coefficients (spline.h:308-448,
//            rif_interpolated.h:46-78), mirror-wrapped indices.
// All share the gap rule n = n0, grad n = 0 for x outside
// [sgap_begin_loc_x, egap_end_loc_x] (cuda_scene.cuh:463-481).
#pragma once
#include "device_math.h"

namespace simpler {

// Scene as the kernels see it: one POD passed by value as a kernel argument
// (scalar registers / constant cache), so launches on different scenes can
// run concurrently (the reference keeps a single __constant__ copy,
// cuda_renderer.cu:48).
struct DevScene {
    f3 blockL, blockR;
    float ior;
    // source
    f3 light_origin;
    float ct;  // cos(half_theta_limit)
    float plane[2], pixsize[2];
    int tex_x, ncdf;
    const float *cdf;
    float elens_x, elens_ap2, elens_f;
    int elens_active;
    // camera
    f3 cam_o, cam_dir, cam_h, cam_v;
    float view_plane[2], pl0, pl1;
    int nogate, bounce;
    int slens_active;
    float slens_ap2, slens_f;
    // field
    float n_o, k_r, amp, n_max_scaling, phi_min, phi_range, time_coupling;
    int mode, general_axis;
    f3 uz, ux, uy, pu;
    float h0;
    int precision;
    float egap, sgap;
    int n_src, n_coeff;
    const float4 *src;  // (x, y, z, amplitude)
    const float *src_phase;
    float coeff_w[8];
    const float *spline;  // coefficients minus spline_ref
    float spline_ref;
    int sN[3];
    float sxmin[3], sxres[3];
    // medium
    float sigma_t, mfp, albedo, g;
    // renderer
    int max_depth;
    float max_pathlength;
    int use_direct, use_nee, direct_leak, use_fresnel;
    float weight, Lcut;
    uint32_t seed_lo, seed_hi;
};

enum FieldKind { kBessel = 0, kSources = 1, kSpline = 2 };

// ---------------------------------------------------------------- Bessel
// J_m(x) and J_{m+1}(x), 0 <= m <= 7, FP32, absolute error < ~1e-6 for the
// arguments the engine meets (x = k r <= ~100).
//   x < 3: power series J_n(x) = sum_j (-t)^j (x/2)^n / (j! (n+j)!), t = x^2/4
//          (11 terms: last term < 2.25^11/(11!)^2 ~ 6e-12).
//   x >= 3: Abramowitz & Stegun 9.4.3 / 9.4.6 for J0, J1 (rational
//          asymptotic forms, |eps| < 1e-7), then upward recurrence
//          J_{n+1} = (2n/x) J_n - J_{n-1}, stable for n < x.
__device__ __forceinline__ float bessel_series(int n, float x) {
    const float hx = 0.5f * x;
    const float t = hx * hx;
    float pw = 1.0f, fact = 1.0f;
    for (int i = 1; i <= n; ++i) {
        pw *= hx;
        fact *= (float)i;
    }
    // Horner in t over j = 10..0 of (-1)^j / (j! (n+j)!) * n!
    float s = 1.0f;
#pragma unroll
    for (int j = 10; j >= 1; --j) s = 1.0f - s * __fdividef(t, (float)(j * (n + j)));
    return s * pw / fact;
}

__device__ __forceinline__ void bessel_j01_large(float x, float &j0, float &j1) {
    const float y = 3.0f / x;
    const float f0 = 0.79788456f +
                     y * (-0.00000077f +
                          y * (-0.00552740f +
                               y * (-0.00009512f + y * (0.00137237f + y * (-0.00072805f + y * 0.00014476f)))));
    const float t0 = -0.78539816f +
                     y * (-0.04166397f +
                          y * (-0.00003954f +
                               y * (0.00262573f + y * (-0.00054125f + y * (-0.00029333f + y * 0.00013558f)))));
    const float f1 = 0.79788456f +
                     y * (0.00000156f +
                          y * (0.01659667f +
                               y * (0.00017105f + y * (-0.00249511f + y * (0.00113653f + y * -0.00020033f)))));
    const float t1 = -2.35619449f +
                     y * (0.12499612f +
                          y * (0.00005650f +
                               y * (-0.00637879f + y * (0.00074348f + y * (0.00079824f + y * -0.00029166f)))));
    const float rs = rsqrtf(x);
    // cos(x + t) with x up to ~100 rad: reduce x in revolutions first
    const float inv2pi = 0.15915494309189535f;
    const float xr = x * inv2pi;
    const float xf = xr - rintf(xr);
    const float a0 = xf * (2.0f * kPi) + t0;
    const float a1 = xf * (2.0f * kPi) + t1;
    j0 = rs * f0 * __cosf(a0);
    j1 = rs * f1 * __cosf(a1);
}

__device__ __forceinline__ void bessel_pair(int m, float x, float &jm, float &jm1) {
    if (x < 3.0f) {
        jm = bessel_series(m, x);
        jm1 = bessel_series(m + 1, x);
    } else {
        float a, b;
        bessel_j01_large(x, a, b);  // J0, J1
        const float ix = 1.0f / x;
        for (int n = 1; n <= m; ++n) {
            const float c = 2.0f * (float)n * ix * b - a;
            a = b;
            b = c;
        }
        jm = a;
        jm1 = b;
    }
}

template <int FIELD>
__device__ __forceinline__ void field_eval(const DevScene &S, f3 p, float scaling, float tau, float &n, f3 &g);

// reference scene.cpp:205-257 (axis = +-x component form) or the same
// gradient for a general axis: grad n = A s [J'(kr) k cos(m phi) rho/r
//                                           - J m sin(m phi) (uz x rho)/r^2]
template <>
__device__ __forceinline__ void field_eval<kBessel>(const DevScene &S, f3 p, float scaling, float, float &n,
                                                    f3 &g) {
    if (p.x > S.egap || p.x < S.sgap) {
        n = S.n_o;
        g = mk3(0.f, 0.f, 0.f);
        return;
    }
    const f3 q = p - S.pu;
    const float t = dot(q, S.uz);
    f3 rho = q - t * S.uz;
    float r = len(rho);
    const float cx = dot(rho, S.ux), cy = dot(rho, S.uy);
    // cos(m phi), sin(m phi) from phi = atan2(cy, cx) by the angle-addition
    // recurrence (atan2(0,0) = 0 at r = 0)
    float c1 = 1.0f, s1 = 0.0f;
    if (r > 0.0f) {
        const float ir = 1.0f / r;
        c1 = cx * ir;
        s1 = cy * ir;
    }
    float cm = 1.0f, sm = 0.0f;
    for (int k = 0; k < S.mode; ++k) {
        const float cn = cm * c1 - sm * s1;
        sm = sm * c1 + cm * s1;
        cm = cn;
    }
    const float A = S.amp * scaling;
    float jm, jm1;
    const bool clamp = r < kEps;
    if (clamp) {
        // the value uses the true radius; the gradient the clamped one
        // (reference scene.cpp:209-214 vs 229-234)
        float dummy;
        bessel_pair(S.mode, S.k_r * r, jm, dummy);
        n = S.n_o + A * jm * cm;
        if (S.general_axis) {
            rho = kEps * (S.ux + S.uy);
        } else {
            rho.y = kEps;
            rho.z = kEps;
        }
        r = kEps;
    }
    const float x = S.k_r * r;
    bessel_pair(S.mode, x, jm, jm1);
    if (!clamp) n = S.n_o + A * jm * cm;
    const float dJ = (float)S.mode / x * jm - jm1;
    const float invr = 1.0f / r;
    const float cr = A * dJ * S.k_r * invr * cm;
    const float ctan = A * jm * (float)S.mode * sm * invr * invr;
    if (S.general_axis) {
        const f3 tp = cross(S.uz, rho);
        g = mk3(cr * rho.x - ctan * tp.x, cr * rho.y - ctan * tp.y, cr * rho.z - ctan * tp.z);
    } else {
        g = mk3(0.f, cr * rho.y - ctan * rho.z, cr * rho.z + ctan * rho.y);
    }
}

// reference scene.cpp:140-195, generalised (see simpler.h)
template <>
__device__ __forceinline__ void field_eval<kSources>(const DevScene &S, f3 p, float, float tau, float &n,
                                                     f3 &g) {
    if (p.x > S.egap || p.x < S.sgap) {
        n = S.n_o;
        g = mk3(0.f, 0.f, 0.f);
        return;
    }
    float acc = 0.f;
    f3 gacc = mk3(0.f, 0.f, 0.f);
    const float tshift = S.time_coupling * tau;
    for (int i = 0; i < S.n_src; ++i) {
        const float4 s = S.src[i];
        const f3 dv = mk3(p.x - s.x, p.y - s.y, p.z - s.z);
        const float r2 = dot(dv, dv);
        const float ir = rsqrtf(r2);
        const float r = r2 * ir;
        const float ph0 = S.src_phase[i] - tshift;
        float vs = 0.f, gs = 0.f;
        for (int c = 1; c <= S.n_coeff; ++c) {
            const float krc = S.k_r * r * (float)c;
            float sn, cs;
            __sincosf(krc + ph0, &sn, &cs);
            vs += S.coeff_w[c - 1] * cs;
            gs += S.coeff_w[c - 1] * (cs + krc * sn);
        }
        acc += s.w * vs * ir;
        const float common = -s.w * gs * ir * ir * ir;
        gacc = gacc + common * dv;
    }
    n = S.n_o + acc;
    g = gacc;
}

__device__ __forceinline__ int mirror_idx(int i, int n) {
    const int period = 2 * n - 2;
    int w = i % period;
    w += (w < 0) ? period : 0;
    return (w >= n) ? period - w : w;
}

// Cubic B-spline weights of the 4 taps floor(x)-1 .. floor(x)+2 and their
// derivatives (spline.h:33-53 evaluated at x - index).
__device__ __forceinline__ void bspline_w(float t, float w[4], float dw[4]) {
    const float t2 = t * t, t3 = t2 * t;
    const float u = 1.0f - t;
    w[0] = (1.0f / 6.0f) * u * u * u;
    w[1] = (2.0f / 3.0f) - t2 + 0.5f * t3;
    w[2] = (2.0f / 3.0f) - u * u + 0.5f * u * u * u;
    w[3] = (1.0f / 6.0f) * t3;
    dw[0] = -0.5f * u * u;
    dw[1] = (1.5f * t - 2.0f) * t;
    dw[2] = -(1.5f * u - 2.0f) * u;
    dw[3] = 0.5f * t2;
}

template <>
__device__ __forceinline__ void field_eval<kSpline>(const DevScene &S, f3 p, float, float, float &n, f3 &g) {
    if (p.x > S.egap || p.x < S.sgap) {
        n = S.n_o;
        g = mk3(0.f, 0.f, 0.f);
        return;
    }
    const float gx = (p.x - S.sxmin[0]) * S.sxres[0];
    const float gy = (p.y - S.sxmin[1]) * S.sxres[1];
    const float gz = (p.z - S.sxmin[2]) * S.sxres[2];
    const float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    float wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
    bspline_w(gx - fx, wx, dwx);
    bspline_w(gy - fy, wy, dwy);
    bspline_w(gz - fz, wz, dwz);
    const int ix = (int)fx - 1, iy = (int)fy - 1, iz = (int)fz - 1;
    const int NX = S.sN[0], NY = S.sN[1], NZ = S.sN[2];
    int xi[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) xi[a] = mirror_idx(ix + a, NX);
    float v = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int zi = mirror_idx(iz + c, NZ);
        float s0 = 0.f, sx = 0.f, sy0 = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int yi = mirror_idx(iy + b, NY);
            const float *row = S.spline + ((long)zi * NY + yi) * NX;
            const float c0 = row[xi[0]], c1 = row[xi[1]], c2 = row[xi[2]], c3 = row[xi[3]];
            const float a0 = c0 * wx[0] + c1 * wx[1] + c2 * wx[2] + c3 * wx[3];
            const float a1 = c0 * dwx[0] + c1 * dwx[1] + c2 * dwx[2] + c3 * dwx[3];
            s0 += a0 * wy[b];
            sx += a1 * wy[b];
            sy0 += a0 * dwy[b];
        }
        v += s0 * wz[c];
        vx += sx * wz[c];
        vy += sy0 * wz[c];
        vz += s0 * dwz[c];
    }
    // the weights sum to 1: n = ref + sum (c - ref) w keeps the rounding on
    // the perturbation; the gradient weights sum to 0
    n = S.spline_ref + v;
    g = mk3(vx * S.sxres[0], vy * S.sxres[1], vz * S.sxres[2]);
}

}  // namespace simpler
