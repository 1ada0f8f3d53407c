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
    f3 inL, inR;  // blockL - eps, blockR + eps (rounded from double): the inside test's bounds
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
    const float4 *jtab;  // Bessel table (see bessel_tab)
    int jtab_n;
    f3 uz, ux, uy, pu;
    float h0;
    int precision;
    float egap, sgap;
    int n_src, n_coeff;
    const float4 *src;  // (x, y, z, amplitude)
    const float *src_phase;
    float coeff_w[8];
    const float *spline;  // coefficients minus spline_ref
    const float *spline_pad;  // the same with a mirrored halo (see kSplineHalo), or null
    float spline_ref;
    int sN[3], sP[3];  // grid and padded grid dimensions
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

// Field parameters that differ between the configurations of a batched
// launch (uniform in a single-scene launch, per lane in a batched one).
struct FieldP {
    float amp, k_r;
    int mode, jtab_n;
    const float4 *jtab;   // Bessel J_m / J_m' table
    const float4 *src;    // point sources (x, y, z, amplitude)
    const float *phase;   // point-source phases
};

__device__ __forceinline__ FieldP field_params(const DevScene &S) {
    FieldP F;
    F.amp = S.amp;
    F.k_r = S.k_r;
    F.mode = S.mode;
    F.jtab_n = S.jtab_n;
    F.jtab = S.jtab;
    F.src = S.src;
    F.phase = S.src_phase;
    return F;
}

// ---------------------------------------------------------------- Bessel
// J_m(x) and J_m'(x) from a per-scene table built on the host: on every unit
// interval [i, i+1) of x = k r, degree-7 polynomials in u = x - (i + 1/2),
// fitted in double at Chebyshev nodes to glibc jn (interpolation error
// < 1e-9 for |J^(8)| <= 1).  Layout: 4 float4 per interval = 8 coefficients
// of J_m, then 8 of J_m'.  One branch-free evaluation per step, no division,
// no transcendental (the reference calls jn three times per evaluation,
// scene.cpp:238-240).
__device__ __forceinline__ void bessel_tab(const FieldP &F, float x, float &J, float &dJ) {
    const int i = min((int)x, F.jtab_n - 1);
    const float u = x - ((float)i + 0.5f);
    const float4 *t = F.jtab + 4 * i;
    const float4 a = t[0], b = t[1], c = t[2], d = t[3];
    J = a.x + u * (a.y + u * (a.z + u * (a.w + u * (b.x + u * (b.y + u * (b.z + u * b.w))))));
    dJ = c.x + u * (c.y + u * (c.z + u * (c.w + u * (d.x + u * (d.y + u * (d.z + u * d.w))))));
}

template <int FIELD>
__device__ __forceinline__ void field_eval(const DevScene &S, const FieldP &F, f3 p, float scaling, float tau,
                                           float &n, f3 &g);

// reference scene.cpp:205-257 (axis = +-x component form) or the same
// gradient for a general axis: grad n = A s [J'(kr) k cos(m phi) rho/r
//                                           - J m sin(m phi) (uz x rho)/r^2]
// MM >= 0: the order m known at compile time (m = 0 drops the angular
// terms; the angle-addition recurrence unrolls); MM < 0: m = F.mode.
// GA >= 0: the axis form known at compile time (0: the reference's
// component form, 1: general axis, 2: general axis with the frame
// uz = +z, ux = +x, uy = +y, where the projection onto the plane normal to
// the axis is rho = (q.x, q.y, 0) and uz x rho = (-rho.y, rho.x, 0));
// GA < 0: S.general_axis.
template <int MM, int GA = -1>
__device__ __forceinline__ void bessel_eval(const DevScene &S, const FieldP &F, f3 p, float scaling, float &n,
                                            f3 &g) {
    const int mode = MM >= 0 ? MM : F.mode;
    const bool general = GA >= 0 ? (GA != 0) : (S.general_axis != 0);
    // outside the gap planes n = n_o and grad n = 0: the amplitude is zero
    // there (n_o + 0 J = n_o exactly, the table values being finite), which
    // costs one select instead of a branch around the evaluation
    const bool in_gap = !(p.x > S.egap || p.x < S.sgap);
    const float A = in_gap ? F.amp * scaling : 0.f;
    f3 rho;
    float r2;
    if (GA == 2) {
        rho = mk3(p.x - S.pu.x, p.y - S.pu.y, 0.f);
        r2 = rho.x * rho.x + rho.y * rho.y;
    } else {
        const f3 q = p - S.pu;
        const float t = dot(q, S.uz);
        rho = q - t * S.uz;
        r2 = dot(rho, rho);
    }
    const bool nz = r2 > 0.f;
    float ir = nz ? __builtin_amdgcn_rsqf(r2) : 0.f;
    float r = r2 * ir;
    // cos(m phi), sin(m phi) with phi = atan2(rho.uy, rho.ux) by the
    // angle-addition recurrence (atan2(0,0) = 0 at r = 0)
    float cm = 1.0f, sm = 0.0f;
    if (MM != 0) {
        const float c1 = nz ? (GA == 2 ? rho.x : dot(rho, S.ux)) * ir : 1.0f;
        const float s1 = nz ? (GA == 2 ? rho.y : dot(rho, S.uy)) * ir : 0.0f;
        if (MM > 0) {  // the first product with (1, 0) is (c1, s1)
            cm = c1;
            sm = s1;
        }
#pragma unroll
        for (int k = MM > 0 ? 1 : 0; k < (MM >= 0 ? MM : mode); ++k) {
            const float cn = cm * c1 - sm * s1;
            sm = sm * c1 + cm * s1;
            cm = cn;
        }
    }
    float J, dJ;
    const bool clamp = r < kEps;
    if (clamp) {
        // the value uses the true radius, the gradient the clamped one
        // (reference scene.cpp:209-214 vs 229-234)
        bessel_tab(F, F.k_r * r, J, dJ);
        n = S.n_o + A * J * cm;
        if (GA == 2) {
            rho = mk3(kEps, kEps, 0.f);
        } else if (general) {
            rho = kEps * (S.ux + S.uy);
        } else {
            rho.y = kEps;
            rho.z = kEps;
        }
        r = kEps;
        ir = 1.0f / kEps;
    }
    bessel_tab(F, F.k_r * r, J, dJ);
    if (!clamp) n = S.n_o + A * J * cm;
    const float cr = (A * F.k_r) * dJ * ir * cm;
    if (MM == 0) {
        if (GA == 2)
            g = mk3(cr * rho.x, cr * rho.y, 0.f);
        else
            g = general ? cr * rho : mk3(0.f, cr * rho.y, cr * rho.z);
        return;
    }
    const float ctan = A * J * (float)mode * sm * ir * ir;
    if (GA == 2) {
        g = mk3(cr * rho.x + ctan * rho.y, cr * rho.y - ctan * rho.x, 0.f);
    } else if (general) {
        const f3 tp = cross(S.uz, rho);
        g = mk3(cr * rho.x - ctan * tp.x, cr * rho.y - ctan * tp.y, cr * rho.z - ctan * tp.z);
    } else {
        g = mk3(0.f, cr * rho.y - ctan * rho.z, cr * rho.z + ctan * rho.y);
    }
}

template <>
__device__ __forceinline__ void field_eval<kBessel>(const DevScene &S, const FieldP &F, f3 p, float scaling,
                                                    float, float &n, f3 &g) {
    bessel_eval<-1>(S, F, p, scaling, n, g);
}

// reference scene.cpp:140-195, generalised (see simpler.h)
template <>
__device__ __forceinline__ void field_eval<kSources>(const DevScene &S, const FieldP &F, f3 p, float, float tau,
                                                     float &n, f3 &g) {
    if (p.x > S.egap || p.x < S.sgap) {
        n = S.n_o;
        g = mk3(0.f, 0.f, 0.f);
        return;
    }
    float acc = 0.f;
    f3 gacc = mk3(0.f, 0.f, 0.f);
    const float tshift = S.time_coupling * tau;
    for (int i = 0; i < S.n_src; ++i) {
        const float4 s = F.src[i];
        const f3 dv = mk3(p.x - s.x, p.y - s.y, p.z - s.z);
        const float r2 = dot(dv, dv);
        const float ir = rsqrtf(r2);
        const float r = r2 * ir;
        const float ph0 = F.phase[i] - tshift;
        float vs = 0.f, gs = 0.f;
        for (int c = 1; c <= S.n_coeff; ++c) {
            const float krc = F.k_r * r * (float)c;
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

// Mirror-wrapped node index (spline.h:617-620, 636-644 for 1-D / 2-D; also
// applied in 3-D).  One reflection covers i in [-(n-1), 2n-2], i.e. every
// point within the grid's box and up to a grid length beyond it, without the
// integer division of the general period formula, which is kept for the rest.
__device__ __forceinline__ int mirror_idx(int i, int n) {
    int r = i < 0 ? -i : i;
    r = r >= n ? 2 * n - 2 - r : r;
    if ((unsigned)r < (unsigned)n) return r;
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

// Mirrored halo of the padded coefficient copy: kSplineHalo cells below and
// kSplineHalo + 1 above each axis, so that points up to a cell outside the
// grid (a step overshooting the block) still read their taps in place.
constexpr int kSplineHalo = 2;

// One row of 4 x-taps: a0 (value weights) and a1 (x-derivative weights)
// accumulated with the row's y weight and y-derivative weight.
__device__ __forceinline__ void spline_row(float c0, float c1, float c2, float c3, const float wx[4],
                                           const float dwx[4], float wyb, float dwyb, float &s0, float &sx,
                                           float &sy0) {
    const float a0 = c0 * wx[0] + c1 * wx[1] + c2 * wx[2] + c3 * wx[3];
    const float a1 = c0 * dwx[0] + c1 * dwx[1] + c2 * dwx[2] + c3 * dwx[3];
    s0 += a0 * wyb;
    sx += a1 * wyb;
    sy0 += a0 * dwyb;
}

// The 4 x 4 x 4 taps of an evaluation, tap (c, b, q) at z, y, x offsets
// c, b, q from the base corner: read in place from the padded copy, or from
// a register copy of the cell (SplineCell).
struct PadTaps {
    const float *base;
    long row, slab;
    __device__ __forceinline__ float operator()(int c, int b, int q) const { return base[c * slab + b * row + q]; }
};
struct CellTaps {
    const float *t;
    __device__ __forceinline__ float operator()(int c, int b, int q) const { return t[(c * 4 + b) * 4 + q]; }
};

// Are the taps of base corner (jx, jy, jz) (padded indices) all inside the
// padded copy?
__device__ __forceinline__ bool spline_padded(const DevScene &S, int jx, int jy, int jz) {
    return S.spline_pad && (unsigned)jx <= (unsigned)(S.sP[0] - 4) && (unsigned)jy <= (unsigned)(S.sP[1] - 4) &&
           (unsigned)jz <= (unsigned)(S.sP[2] - 4);
}

// Value and gradient sums over the taps (x, then y, then z passes).
template <class Taps>
__device__ __forceinline__ void spline_sum(const Taps &T, const float wx[4], const float dwx[4], const float wy[4],
                                           const float dwy[4], const float wz[4], const float dwz[4], float &v,
                                           float &vx, float &vy, float &vz) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        float s0 = 0.f, sx = 0.f, sy0 = 0.f;
#pragma unroll
        for (int b = 0; b < 4; ++b)
            spline_row(T(c, b, 0), T(c, b, 1), T(c, b, 2), T(c, b, 3), wx, dwx, wy[b], dwy[b], s0, sx, sy0);
        v += s0 * wz[c];
        vx += sx * wz[c];
        vy += sy0 * wz[c];
        vz += s0 * dwz[c];
    }
}

template <>
__device__ __forceinline__ void field_eval<kSpline>(const DevScene &S, const FieldP &, f3 p, float, float,
                                                    float &n, f3 &g) {
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
    float v = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
    // taps (ix..ix+3, iy.., iz..) inside the padded grid: fixed strides
    // from one corner, the halo holding the mirrored values
    const int jx = ix + kSplineHalo, jy = iy + kSplineHalo, jz = iz + kSplineHalo;
    if (spline_padded(S, jx, jy, jz)) {
        const long row = S.sP[0], slab = (long)S.sP[0] * S.sP[1];
        spline_sum(PadTaps{S.spline_pad + (long)jz * slab + (long)jy * row + jx, row, slab}, wx, dwx, wy, dwy, wz,
                   dwz, v, vx, vy, vz);
    } else {
        int xi[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) xi[a] = mirror_idx(ix + a, NX);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int zi = mirror_idx(iz + c, NZ);
            float s0 = 0.f, sx = 0.f, sy0 = 0.f;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int yi = mirror_idx(iy + b, NY);
                const float *row = S.spline + ((long)zi * NY + yi) * NX;
                spline_row(row[xi[0]], row[xi[1]], row[xi[2]], row[xi[3]], wx, dwx, wy[b], dwy[b], s0, sx, sy0);
            }
            v += s0 * wz[c];
            vx += sx * wz[c];
            vy += sy0 * wz[c];
            vz += s0 * dwz[c];
        }
    }
    // the weights sum to 1: n = ref + sum (c - ref) w keeps the rounding on
    // the perturbation; the gradient weights sum to 0
    n = S.spline_ref + v;
    g = mk3(vx * S.sxres[0], vy * S.sxres[1], vz * S.sxres[2]);
}

// The 64 taps of one base cell of the padded coefficient copy, held in
// registers by a march loop.  At 5 um steps in cells of about 160 um a path
// stays in one cell for tens of consecutive steps, so the taps are read from
// memory once per cell instead of once per step (16 L2 reads per step were
// the spline march's bound).
struct SplineCell {
    float c[64];  // taps [z][y][x], 4 x 4 x 4
    long key;     // offset of the cell's first tap in the padded copy; -1 none
};

// Loads the cell of padded offset key into the register copy when it holds
// another one.
__device__ __forceinline__ void spline_cell_load(const DevScene &S, long key, SplineCell &cell) {
    if (key == cell.key) return;
    const long row = S.sP[0], slab = (long)S.sP[0] * S.sP[1];
    const float *base = S.spline_pad + key;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) cell.c[(c * 4 + b) * 4 + q] = base[c * slab + b * row + q];
    cell.key = key;
}

// field_eval<kSpline> with the taps from the cell cache, loaded into it when
// the base cell changes; the same sums as the padded path.  Points whose taps
// leave the padded copy take the uncached wrapping path.
__device__ __forceinline__ void spline_eval_cached(const DevScene &S, const FieldP &F, f3 p, float &n, f3 &g,
                                                   SplineCell &cell) {
    if (p.x > S.egap || p.x < S.sgap) {
        n = S.n_o;
        g = mk3(0.f, 0.f, 0.f);
        return;
    }
    const float gx = (p.x - S.sxmin[0]) * S.sxres[0];
    const float gy = (p.y - S.sxmin[1]) * S.sxres[1];
    const float gz = (p.z - S.sxmin[2]) * S.sxres[2];
    const float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    const int jx = (int)fx - 1 + kSplineHalo, jy = (int)fy - 1 + kSplineHalo, jz = (int)fz - 1 + kSplineHalo;
    if (!spline_padded(S, jx, jy, jz)) {
        field_eval<kSpline>(S, F, p, 0.f, 0.f, n, g);
        return;
    }
    float wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
    bspline_w(gx - fx, wx, dwx);
    bspline_w(gy - fy, wy, dwy);
    bspline_w(gz - fz, wz, dwz);
    spline_cell_load(S, (long)jz * S.sP[0] * S.sP[1] + (long)jy * S.sP[0] + jx, cell);
    float v = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
    spline_sum(CellTaps{cell.c}, wx, dwx, wy, dwy, wz, dwz, v, vx, vy, vz);
    n = S.spline_ref + v;
    g = mk3(vx * S.sxres[0], vy * S.sxres[1], vz * S.sxres[2]);
}

// ====================================================================
// Second derivatives and parameter derivatives for the adjoint.
//
// field_eval_h: n, grad n and the Hessian of n (the adjoint of a leapfrog
// step needs H v products).  field_vjp: for scalars a and vector b at a
// point, adds  a dn/dtheta + b . d(grad n)/dtheta  to the gradient of every
// field parameter theta the point depends on.
//   Spline:  theta = the coefficients; into a device grid with the layout of
//            the coefficients (64 taps, float atomics).
//   Sources: theta = (a_0..a_{N-1}, phi_0..phi_{N-1}, k); into per-lane
//            accumulators acc[j * stride] (lane-private LDS).
// The gap rule makes all of them zero outside [sgap, egap].

struct Sym3 {
    float xx, yy, zz, xy, xz, yz;
};
__device__ __forceinline__ f3 symmul(const Sym3 &H, f3 v) {
    return mk3(H.xx * v.x + H.xy * v.y + H.xz * v.z, H.xy * v.x + H.yy * v.y + H.yz * v.z,
               H.xz * v.x + H.yz * v.y + H.zz * v.z);
}
__device__ __forceinline__ Sym3 sym_zero() { return Sym3{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}; }

// The XCD (accelerator complex die) the calling wave runs on: hardware
// register XCC_ID (id 20, bits 3:0).  A wave never moves between XCDs.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_id() { return __builtin_amdgcn_s_getreg((3 << 11) | 20) & (kXcds - 1); }

struct GradSink {
    // spline coefficient gradient (device): in the image and per-ray
    // adjoints, the copy of the wave's XCD (the host sums the copies): a
    // multi-XCD device performs device atomics beyond the L2, at the memory
    // side, where atomics to one address serialise, and the reversed
    // next-event sub-paths all converge on the cells in front of the aperture
    float *grid;
    float *acc;   // source-parameter accumulators (LDS, lane-private)
    int stride;
    // spline: the lane's sums for a 4x4x4 window of taps (LDS, 64 slots
    // `stride` floats apart), tap (x, y, z) in slot (z&3)*16 + (y&3)*4 +
    // (x&3); the window follows the lane's base cell, and only the sums of
    // taps that leave it go to grid (16 of 64 for a move to a neighbouring
    // cell; a reverse path moves to a neighbour about every ten steps).
    // (kx, ky, kz): the window's base tap, kx = -1 when empty; cell null for
    // kernels without the buffer
    float *cell;
    int kx, ky, kz;
};

template <int FIELD>
__device__ __forceinline__ void field_eval_h(const DevScene &S, const FieldP &F, f3 p, float scaling, float tau,
                                             float &n, f3 &g, Sym3 &H);
template <int FIELD>
__device__ __forceinline__ void field_vjp(const DevScene &S, const FieldP &F, f3 p, float scaling, float tau,
                                          float a, f3 b, GradSink &G);

// Bessel.  With f = A s J_m(k r) cos(m phi) (A = n_max, s = the photon's
// duty scaling), r and phi the polar coordinates of p about the transducer
// axis, rh = grad r (the transverse unit vector) and th = uz x rh = r grad phi:
//   grad f = A s [k J' C rh - (m J S / r) th]            (C, S = cos, sin m phi)
//   H f    = A s [a rh rh^T + b th th^T + c (rh th^T + th rh^T)],
//     a = k^2 J'' C,  b = k J' C / r - m^2 J C / r^2,  c = m J S / r^2 - m k J' S / r
// (the reference's bessel_HessianRIF, scene.cpp:259-320, is the same
// expansion for its axis = x frame without the scaling factor; here the
// Hessian is that of the gradient the march uses).  J'' is the derivative of
// the table's J' polynomial.  The general-axis frame is used throughout; the
// adjoint accepts the reference's component form only where it equals it.
// Parameters (field_vjp): theta = (A, k, tau shift), accumulated per lane in
// acc[0], acc[stride], acc[2 stride]:
//   d/dA:   n: s J C,                grad: s [k J' C rh - (m J S / r) th]
//   d/dk:   n: A s J' r C,           grad: A s [(J' + k r J'') C rh - m J' S th]
//   d/dtau: n: A s' J C,             grad: A s' [k J' C rh - (m J S / r) th]
// with s' = ds/dtau = cos tau while the clamp |sin tau| <= n_clip / n_max is
// inactive (0 when it clips); tau shift moves the whole duty window
// [phi_min, phi_max].  The amplitude derivative holds n_clip / n_max fixed.
struct BesselLocal {
    f3 rh, th;
    float r, J, dJ, d2J, C, Sn, As;
    int m;
};

__device__ __forceinline__ void bessel_tab2(const FieldP &F, float x, float &J, float &dJ, float &d2J) {
    const int i = min((int)x, F.jtab_n - 1);
    const float u = x - ((float)i + 0.5f);
    const float4 *t = F.jtab + 4 * i;
    const float4 a = t[0], b = t[1], c = t[2], d = t[3];
    J = a.x + u * (a.y + u * (a.z + u * (a.w + u * (b.x + u * (b.y + u * (b.z + u * b.w))))));
    dJ = c.x + u * (c.y + u * (c.z + u * (c.w + u * (d.x + u * (d.y + u * (d.z + u * d.w))))));
    d2J = c.y + u * (2.f * c.z + u * (3.f * c.w + u * (4.f * d.x + u * (5.f * d.y + u * (6.f * d.z + u * (7.f * d.w))))));
}

// The local quantities at p (general-axis frame; r < eps uses the clamped
// transverse vector of the gradient, reference scene.cpp:229-234).
__device__ __forceinline__ BesselLocal bessel_local(const DevScene &S, const FieldP &F, f3 p, float scaling) {
    BesselLocal b;
    const bool in_gap = !(p.x > S.egap || p.x < S.sgap);
    b.As = in_gap ? F.amp * scaling : 0.f;
    b.m = F.mode;
    const f3 q = p - S.pu;
    f3 rho = q - dot(q, S.uz) * S.uz;
    float r2 = dot(rho, rho);
    float ir = r2 > 0.f ? __builtin_amdgcn_rsqf(r2) : 0.f;
    b.r = r2 * ir;
    float c1 = r2 > 0.f ? dot(rho, S.ux) * ir : 1.0f, s1 = r2 > 0.f ? dot(rho, S.uy) * ir : 0.0f;
    if (b.r < kEps) {
        rho = kEps * (S.ux + S.uy);
        b.r = kEps;
        ir = 1.0f / kEps;
    }
    float cm = 1.f, sm = 0.f;
    for (int k = 0; k < b.m; ++k) {
        const float cn = cm * c1 - sm * s1;
        sm = sm * c1 + cm * s1;
        cm = cn;
    }
    b.C = cm;
    b.Sn = sm;
    b.rh = ir * rho;
    b.th = cross(S.uz, b.rh);
    bessel_tab2(F, F.k_r * b.r, b.J, b.dJ, b.d2J);
    return b;
}

template <>
__device__ __forceinline__ void field_eval_h<kBessel>(const DevScene &S, const FieldP &F, f3 p, float scaling,
                                                      float tau, float &n, f3 &g, Sym3 &H) {
    field_eval<kBessel>(S, F, p, scaling, tau, n, g);
    const BesselLocal b = bessel_local(S, F, p, scaling);
    const float k = F.k_r, m = (float)b.m, ir = 1.0f / b.r;
    const float a = b.As * k * k * b.d2J * b.C;
    const float bb = b.As * (k * b.dJ * b.C * ir - m * m * b.J * b.C * ir * ir);
    const float c = b.As * (m * b.J * b.Sn * ir * ir - m * k * b.dJ * b.Sn * ir);
    const f3 R = b.rh, T = b.th;
    H.xx = a * R.x * R.x + bb * T.x * T.x + 2.f * c * R.x * T.x;
    H.yy = a * R.y * R.y + bb * T.y * T.y + 2.f * c * R.y * T.y;
    H.zz = a * R.z * R.z + bb * T.z * T.z + 2.f * c * R.z * T.z;
    H.xy = a * R.x * R.y + bb * T.x * T.y + c * (R.x * T.y + T.x * R.y);
    H.xz = a * R.x * R.z + bb * T.x * T.z + c * (R.x * T.z + T.x * R.z);
    H.yz = a * R.y * R.z + bb * T.y * T.z + c * (R.y * T.z + T.y * R.z);
}

template <>
__device__ __forceinline__ void field_vjp<kBessel>(const DevScene &S, const FieldP &F, f3 p, float scaling, float tau,
                                                   float a, f3 bv, GradSink &G) {
    if (p.x > S.egap || p.x < S.sgap) return;
    const BesselLocal b = bessel_local(S, F, p, scaling);
    const float k = F.k_r, m = (float)b.m, ir = 1.0f / b.r;
    const float br = dot(bv, b.rh), bt = dot(bv, b.th);
    const float unit = a * b.J * b.C + k * b.dJ * b.C * br - m * b.J * b.Sn * ir * bt;  // per unit of A s
    const float dsdt = fabsf(sinf(tau)) < S.n_max_scaling ? cosf(tau) : 0.f;
    G.acc[0] += scaling * unit;
    G.acc[G.stride] += b.As * (a * b.dJ * b.r * b.C + (b.dJ + k * b.r * b.d2J) * b.C * br - m * b.dJ * b.Sn * bt);
    G.acc[2 * G.stride] += F.amp * dsdt * unit;
}

// Point sources.  Per source, with r = |p - x_i|, psi_c = c k r + phi_i - t:
//   f  = sum_c w_c cos(psi_c) / r
//   f' = -sum_c w_c (cos + c k r sin) / r^2 = -gs / r^2
//   f''= (2 gs - hs) / r^3,  hs = sum_c w_c (c k r)^2 cos
//   grad = a f' rhat,   H = a / r^3 [(3 gs - hs) rhat rhat^T - gs I]
template <>
__device__ __forceinline__ void field_eval_h<kSources>(const DevScene &S, const FieldP &F, f3 p, float, float tau,
                                                       float &n, f3 &g, Sym3 &H) {
    H = sym_zero();
    if (p.x > S.egap || p.x < S.sgap) {
        n = S.n_o;
        g = mk3(0.f, 0.f, 0.f);
        return;
    }
    float acc = 0.f;
    f3 gacc = mk3(0.f, 0.f, 0.f);
    const float tshift = S.time_coupling * tau;
    for (int i = 0; i < S.n_src; ++i) {
        const float4 s = F.src[i];
        const f3 dv = mk3(p.x - s.x, p.y - s.y, p.z - s.z);
        const float r2 = dot(dv, dv);
        const float ir = rsqrtf(r2);
        const float r = r2 * ir;
        const float ph0 = F.phase[i] - tshift;
        float vs = 0.f, gs = 0.f, hs = 0.f;
        for (int c = 1; c <= S.n_coeff; ++c) {
            const float krc = F.k_r * r * (float)c;
            float sn, cs;
            __sincosf(krc + ph0, &sn, &cs);
            const float w = S.coeff_w[c - 1];
            vs += w * cs;
            gs += w * (cs + krc * sn);
            hs += w * krc * krc * cs;
        }
        acc += s.w * vs * ir;
        const float ir3 = ir * ir * ir;
        const float common = -s.w * gs * ir3;
        gacc = gacc + common * dv;
        const float crr = s.w * (3.f * gs - hs) * ir3 * ir * ir;  // times dv dv^T
        const float cid = -s.w * gs * ir3;                        // times I
        H.xx += crr * dv.x * dv.x + cid;
        H.yy += crr * dv.y * dv.y + cid;
        H.zz += crr * dv.z * dv.z + cid;
        H.xy += crr * dv.x * dv.y;
        H.xz += crr * dv.x * dv.z;
        H.yz += crr * dv.y * dv.z;
    }
    n = S.n_o + acc;
    g = gacc;
}

// d/da_i:   n: f,            grad: f' rhat
// d/dphi_i: n: -a ss / r,    grad: a (ss - kc) / r^2 rhat,  ss = sum w sin,
//                                                           kc = sum w c k r cos
// d/dk:     n: -a sum w c sin, grad: -a k sum w c^2 cos rhat
template <>
__device__ __forceinline__ void field_vjp<kSources>(const DevScene &S, const FieldP &F, f3 p, float, float tau,
                                                    float a, f3 b, GradSink &G) {
    if (p.x > S.egap || p.x < S.sgap) return;
    const float tshift = S.time_coupling * tau;
    const int N = S.n_src;
    float gk = 0.f;
    for (int i = 0; i < N; ++i) {
        const float4 s = F.src[i];
        const f3 dv = mk3(p.x - s.x, p.y - s.y, p.z - s.z);
        const float r2 = dot(dv, dv);
        const float ir = rsqrtf(r2);
        const float r = r2 * ir;
        const float ph0 = F.phase[i] - tshift;
        float vs = 0.f, gs = 0.f, ss = 0.f, kc = 0.f, cs1 = 0.f, c2c = 0.f;
        for (int c = 1; c <= S.n_coeff; ++c) {
            const float fc = (float)c;
            const float krc = F.k_r * r * fc;
            float sn, cs;
            __sincosf(krc + ph0, &sn, &cs);
            const float w = S.coeff_w[c - 1];
            vs += w * cs;
            gs += w * (cs + krc * sn);
            ss += w * sn;
            kc += w * krc * cs;
            cs1 += w * fc * sn;
            c2c += w * fc * fc * cs;
        }
        const float bdr = dot(b, dv) * ir;  // b . rhat
        // amplitude
        G.acc[i * G.stride] += a * vs * ir - bdr * gs * ir * ir;
        // phase
        G.acc[(N + i) * G.stride] += s.w * (-a * ss * ir + bdr * (ss - kc) * ir * ir);
        // wavenumber
        gk += s.w * (-a * cs1 - bdr * F.k_r * c2c);
    }
    G.acc[2 * N * G.stride] += gk;
}

__device__ __forceinline__ void bspline_w2(float t, float w[4], float dw[4], float d2w[4]) {
    bspline_w(t, w, dw);
    const float u = 1.0f - t;
    d2w[0] = u;
    d2w[1] = 3.0f * t - 2.0f;
    d2w[2] = 3.0f * u - 2.0f;
    d2w[3] = t;
}

// Value, gradient and Hessian sums over the taps: o = v, vx, vy, vz, vxx, vyy,
// vzz, vxy, vxz, vyz (derivatives in grid units).
template <class Taps>
__device__ __forceinline__ void spline_sum_h(const Taps &T, const float wx[4], const float dwx[4],
                                             const float ddx[4], const float wy[4], const float dwy[4],
                                             const float ddy[4], const float wz[4], const float dwz[4],
                                             const float ddz[4], float (&o)[10]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        // s_ab: derivative orders a in x, b in y after the x and y passes
        float s00 = 0, s10 = 0, s01 = 0, s20 = 0, s02 = 0, s11 = 0;
#pragma unroll
        for (int bq = 0; bq < 4; ++bq) {
            const float c0 = T(c, bq, 0), c1 = T(c, bq, 1), c2 = T(c, bq, 2), c3 = T(c, bq, 3);
            const float a0 = c0 * wx[0] + c1 * wx[1] + c2 * wx[2] + c3 * wx[3];
            const float a1 = c0 * dwx[0] + c1 * dwx[1] + c2 * dwx[2] + c3 * dwx[3];
            const float a2 = c0 * ddx[0] + c1 * ddx[1] + c2 * ddx[2] + c3 * ddx[3];
            s00 += a0 * wy[bq];
            s10 += a1 * wy[bq];
            s01 += a0 * dwy[bq];
            s20 += a2 * wy[bq];
            s02 += a0 * ddy[bq];
            s11 += a1 * dwy[bq];
        }
        o[0] += s00 * wz[c];
        o[1] += s10 * wz[c];
        o[2] += s01 * wz[c];
        o[3] += s00 * dwz[c];
        o[4] += s20 * wz[c];
        o[5] += s02 * wz[c];
        o[6] += s00 * ddz[c];
        o[7] += s11 * wz[c];
        o[8] += s10 * dwz[c];
        o[9] += s01 * dwz[c];
    }
}

// Mirror-wrapped taps read from the unpadded grid (points whose taps leave
// the padded copy, or scenes without it).
struct WrapTaps {
    const float *grid;
    long NX, NY;
    int xi[4], yi[4], zi[4];
    __device__ __forceinline__ float operator()(int c, int b, int q) const {
        return grid[((long)zi[c] * NY + yi[b]) * NX + xi[q]];
    }
};

__device__ __forceinline__ WrapTaps wrap_taps(const DevScene &S, int ix, int iy, int iz) {
    WrapTaps T;
    T.grid = S.spline;
    T.NX = S.sN[0];
    T.NY = S.sN[1];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        T.xi[a] = mirror_idx(ix + a, S.sN[0]);
        T.yi[a] = mirror_idx(iy + a, S.sN[1]);
        T.zi[a] = mirror_idx(iz + a, S.sN[2]);
    }
    return T;
}

// n, grad n and H from the tap sums o (spline_sum_h).
__device__ __forceinline__ void spline_h_out(const DevScene &S, const float (&o)[10], float &n, f3 &g, Sym3 &H) {
    const float rx = S.sxres[0], ry = S.sxres[1], rz = S.sxres[2];
    n = S.spline_ref + o[0];
    g = mk3(o[1] * rx, o[2] * ry, o[3] * rz);
    H = Sym3{o[4] * rx * rx, o[5] * ry * ry, o[6] * rz * rz, o[7] * rx * ry, o[8] * rx * rz, o[9] * ry * rz};
}

// CELL: taps from the register copy of the current cell (loaded when the
// base cell changes), for the points whose taps lie in the padded copy.
template <bool CELL>
__device__ __forceinline__ void spline_eval_h(const DevScene &S, f3 p, float &n, f3 &g, Sym3 &H, SplineCell *cell) {
    if (p.x > S.egap || p.x < S.sgap) {
        n = S.n_o;
        g = mk3(0.f, 0.f, 0.f);
        H = sym_zero();
        return;
    }
    const float gx = (p.x - S.sxmin[0]) * S.sxres[0];
    const float gy = (p.y - S.sxmin[1]) * S.sxres[1];
    const float gz = (p.z - S.sxmin[2]) * S.sxres[2];
    const float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    float wx[4], dwx[4], ddx[4], wy[4], dwy[4], ddy[4], wz[4], dwz[4], ddz[4];
    bspline_w2(gx - fx, wx, dwx, ddx);
    bspline_w2(gy - fy, wy, dwy, ddy);
    bspline_w2(gz - fz, wz, dwz, ddz);
    const int ix = (int)fx - 1, iy = (int)fy - 1, iz = (int)fz - 1;
    float o[10] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    // taps inside the padded grid: fixed strides from one corner (see
    // field_eval<kSpline>)
    const int jx = ix + kSplineHalo, jy = iy + kSplineHalo, jz = iz + kSplineHalo;
    if (spline_padded(S, jx, jy, jz)) {
        const long row = S.sP[0], slab = (long)S.sP[0] * S.sP[1];
        const long key = (long)jz * slab + (long)jy * row + jx;
        if constexpr (CELL) {
            spline_cell_load(S, key, *cell);
            spline_sum_h(CellTaps{cell->c}, wx, dwx, ddx, wy, dwy, ddy, wz, dwz, ddz, o);
        } else {
            spline_sum_h(PadTaps{S.spline_pad + key, row, slab}, wx, dwx, ddx, wy, dwy, ddy, wz, dwz, ddz, o);
        }
    } else {
        spline_sum_h(wrap_taps(S, ix, iy, iz), wx, dwx, ddx, wy, dwy, ddy, wz, dwz, ddz, o);
    }
    spline_h_out(S, o, n, g, H);
}

template <>
__device__ __forceinline__ void field_eval_h<kSpline>(const DevScene &S, const FieldP &, f3 p, float, float,
                                                      float &n, f3 &g, Sym3 &H) {
    spline_eval_h<false>(S, p, n, g, H, nullptr);
}

// Moves the lane's tap window to base (nx, ny, nz); nx = -1 empties it.  The
// sums of the taps that leave the window go to the gradient.  Per axis, the
// old window's tap in slot residue r is X = kx + ((r - kx) & 3); it leaves
// when X is outside [nx, nx + 4).
__device__ __forceinline__ void spline_window_move(const DevScene &S, GradSink &G, int nx, int ny, int nz) {
    if (G.kx >= 0) {
        const long NX = S.sN[0], NXY = (long)S.sN[0] * S.sN[1];
        int X[4], Y[4], Z[4];
        unsigned lx = 0u, ly = 0u, lz = 0u;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            X[r] = G.kx + ((r - G.kx) & 3);
            Y[r] = G.ky + ((r - G.ky) & 3);
            Z[r] = G.kz + ((r - G.kz) & 3);
            if (nx < 0 || (unsigned)(X[r] - nx) > 3u) lx |= 1u << r;
            if (nx < 0 || (unsigned)(Y[r] - ny) > 3u) ly |= 1u << r;
            if (nx < 0 || (unsigned)(Z[r] - nz) > 3u) lz |= 1u << r;
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const bool plane = ((lz >> c) | (ly >> b)) & 1u;
                float *row = G.grid + Z[c] * NXY + Y[b] * NX;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (!(plane || ((lx >> q) & 1u))) continue;
                    float *slot = G.cell + (c * 16 + b * 4 + q) * G.stride;
                    const float v = *slot;
                    if (v != 0.f) unsafeAtomicAdd(row + X[q], v);
                    *slot = 0.f;
                }
            }
    }
    G.kx = nx;
    G.ky = ny;
    G.kz = nz;
}

// Adds the lane's window sums to the coefficient gradient and empties them.
__device__ __forceinline__ void spline_cell_flush(const DevScene &S, GradSink &G) {
    spline_window_move(S, G, -1, -1, -1);
}

// coefficient (x, y, z) receives a wx wy wz + bx rx dwx wy wz + by ry wx dwy wz
// + bz rz wx wy dwz.  Taps go to the lane's window sums (see GradSink); near
// the faces the mirror-wrapped taps of a point still lie in one 4-wide window
// per axis, clamped to the grid (coinciding taps share a slot, so their sums
// add), which covers every point of the grid's box; points whose taps do not
// fit one window (outside the box, or grids under 4 per axis) add them to
// the gradient directly.
template <>
__device__ __forceinline__ void field_vjp<kSpline>(const DevScene &S, const FieldP &, f3 p, float, float, float a,
                                                   f3 b, GradSink &G) {
    if (p.x > S.egap || p.x < S.sgap) return;
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
    const float bx = b.x * S.sxres[0], by = b.y * S.sxres[1], bz = b.z * S.sxres[2];
    int xi[4], yi[4], zi[4];
    const bool inside = (unsigned)ix <= (unsigned)(NX - 4) && (unsigned)iy <= (unsigned)(NY - 4) &&
                        (unsigned)iz <= (unsigned)(NZ - 4);
    if (inside) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xi[q] = ix + q;
            yi[q] = iy + q;
            zi[q] = iz + q;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            xi[q] = mirror_idx(ix + q, NX);
            yi[q] = mirror_idx(iy + q, NY);
            zi[q] = mirror_idx(iz + q, NZ);
        }
    }
    const int kx = min(max(ix, 0), NX - 4), ky = min(max(iy, 0), NY - 4), kz = min(max(iz, 0), NZ - 4);
    bool fit = G.cell != nullptr && kx >= 0 && ky >= 0 && kz >= 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        fit = fit && (unsigned)(xi[q] - kx) < 4u && (unsigned)(yi[q] - ky) < 4u && (unsigned)(zi[q] - kz) < 4u;
    if (fit) {
        if (kx != G.kx || ky != G.ky || kz != G.kz) spline_window_move(S, G, kx, ky, kz);
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int bq = 0; bq < 4; ++bq) {
                const float Pw = (a * wy[bq] + by * dwy[bq]) * wz[c] + bz * wy[bq] * dwz[c];
                const float Pd = bx * wy[bq] * wz[c];
                float *row = G.cell + (((zi[c] & 3) << 4) | ((yi[bq] & 3) << 2)) * G.stride;
                // an LDS add without return (ds_add_f32): the lane owns its
                // slots, so it needs neither the read nor its wait
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    __hip_atomic_fetch_add(row + (xi[q] & 3) * G.stride, Pw * wx[q] + Pd * dwx[q], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        return;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
#pragma unroll
        for (int bq = 0; bq < 4; ++bq) {
            float *row = G.grid + ((long)zi[c] * NY + yi[bq]) * NX;
            const float Pw = (a * wy[bq] + by * dwy[bq]) * wz[c] + bz * wy[bq] * dwz[c];  // times wx
            const float Pd = bx * wy[bq] * wz[c];                                          // times dwx
#pragma unroll
            for (int q = 0; q < 4; ++q) unsafeAtomicAdd(row + xi[q], Pw * wx[q] + Pd * dwx[q]);
        }
    }
}


// A lane's register sums of the 64 tap contributions of one base cell
// (unclamped base tap (kx, ky, kz); kx = kAccEmpty when empty).  A reverse path
// stays about ten steps in a cell: summing there and adding the sums to the
// LDS window once per cell cuts the LDS adds about tenfold.
constexpr int kAccEmpty = -0x7fffffff - 1;
struct SplineAcc {
    float v[64];  // [c][b][q] as field_vjp's taps
    int kx, ky, kz;
};

// Adds the register sums to the lane's window (or, where the taps do not fit
// one window, to the gradient) and empties them.
__device__ __forceinline__ void spline_acc_commit(const DevScene &S, GradSink &G, SplineAcc &acc) {
    if (acc.kx == kAccEmpty) return;
    const int ix = acc.kx, iy = acc.ky, iz = acc.kz;
    const int NX = S.sN[0], NY = S.sN[1], NZ = S.sN[2];
    int xi[4], yi[4], zi[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        xi[q] = mirror_idx(ix + q, NX);
        yi[q] = mirror_idx(iy + q, NY);
        zi[q] = mirror_idx(iz + q, NZ);
    }
    const int kx = min(max(ix, 0), NX - 4), ky = min(max(iy, 0), NY - 4), kz = min(max(iz, 0), NZ - 4);
    bool fit = G.cell != nullptr && kx >= 0 && ky >= 0 && kz >= 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        fit = fit && (unsigned)(xi[q] - kx) < 4u && (unsigned)(yi[q] - ky) < 4u && (unsigned)(zi[q] - kz) < 4u;
    if (fit) {
        if (kx != G.kx || ky != G.ky || kz != G.kz) spline_window_move(S, G, kx, ky, kz);
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                float *row = G.cell + (((zi[c] & 3) << 4) | ((yi[b] & 3) << 2)) * G.stride;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    __hip_atomic_fetch_add(row + (xi[q] & 3) * G.stride, acc.v[(c * 4 + b) * 4 + q],
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
    } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                float *row = G.grid + ((long)zi[c] * NY + yi[b]) * NX;
#pragma unroll
                for (int q = 0; q < 4; ++q) unsafeAtomicAdd(row + xi[q], acc.v[(c * 4 + b) * 4 + q]);
            }
    }
#pragma unroll
    for (int t = 0; t < 64; ++t) acc.v[t] = 0.f;
    acc.kx = kAccEmpty;
}

// field_vjp<kSpline> into the register sums (see SplineAcc).
__device__ __forceinline__ void spline_vjp_acc(const DevScene &S, f3 p, float a, f3 b, GradSink &G, SplineAcc &acc) {
    if (p.x > S.egap || p.x < S.sgap) return;
    const float gx = (p.x - S.sxmin[0]) * S.sxres[0];
    const float gy = (p.y - S.sxmin[1]) * S.sxres[1];
    const float gz = (p.z - S.sxmin[2]) * S.sxres[2];
    const float fx = floorf(gx), fy = floorf(gy), fz = floorf(gz);
    float wx[4], dwx[4], wy[4], dwy[4], wz[4], dwz[4];
    bspline_w(gx - fx, wx, dwx);
    bspline_w(gy - fy, wy, dwy);
    bspline_w(gz - fz, wz, dwz);
    const int ix = (int)fx - 1, iy = (int)fy - 1, iz = (int)fz - 1;
    if (ix != acc.kx || iy != acc.ky || iz != acc.kz) {
        spline_acc_commit(S, G, acc);
        acc.kx = ix;
        acc.ky = iy;
        acc.kz = iz;
    }
    const float bx = b.x * S.sxres[0], by = b.y * S.sxres[1], bz = b.z * S.sxres[2];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int bq = 0; bq < 4; ++bq) {
            const float Pw = (a * wy[bq] + by * dwy[bq]) * wz[c] + bz * wy[bq] * dwz[c];
            const float Pd = bx * wy[bq] * wz[c];
#pragma unroll
            for (int q = 0; q < 4; ++q) acc.v[(c * 4 + bq) * 4 + q] += Pw * wx[q] + Pd * dwx[q];
        }
}

// ====================================================================
// Tangent of the field along a parameter direction theta_dot (the forward
// mode of field_vjp, same parameter layout): nt = sum_j theta_dot_j dn /
// dtheta_j and gt = sum_j theta_dot_j d(grad n) / dtheta_j at p.
template <int FIELD>
__device__ __forceinline__ void field_jvp(const DevScene &S, const FieldP &F, f3 p, float scaling, float tau,
                                          const float *td, float &nt, f3 &gt);

// spline: the field is linear in the coefficients, so its tangent is the
// spline of the coefficient direction
template <>
__device__ __forceinline__ void field_jvp<kSpline>(const DevScene &S, const FieldP &F, f3 p, float, float,
                                                   const float *td, float &nt, f3 &gt) {
    nt = 0.f;
    gt = mk3(0.f, 0.f, 0.f);
    if (p.x > S.egap || p.x < S.sgap) return;
    DevScene T = S;
    T.spline = td;
    T.spline_pad = nullptr;
    T.spline_ref = 0.f;
    field_eval<kSpline>(T, F, p, 0.f, 0.f, nt, gt);
}

template <>
__device__ __forceinline__ void field_jvp<kSources>(const DevScene &S, const FieldP &F, f3 p, float, float tau,
                                                    const float *td, float &nt, f3 &gt) {
    nt = 0.f;
    gt = mk3(0.f, 0.f, 0.f);
    if (p.x > S.egap || p.x < S.sgap) return;
    const float tshift = S.time_coupling * tau;
    const int N = S.n_src;
    const float kd = td[2 * N];
    for (int i = 0; i < N; ++i) {
        const float4 s = F.src[i];
        const f3 dv = mk3(p.x - s.x, p.y - s.y, p.z - s.z);
        const float r2 = dot(dv, dv);
        const float ir = rsqrtf(r2);
        const float r = r2 * ir;
        const float ph0 = F.phase[i] - tshift;
        float vs = 0.f, gs = 0.f, ss = 0.f, kc = 0.f, cs1 = 0.f, c2c = 0.f;
        for (int c = 1; c <= S.n_coeff; ++c) {
            const float fc = (float)c;
            const float krc = F.k_r * r * fc;
            float sn, cs;
            __sincosf(krc + ph0, &sn, &cs);
            const float w = S.coeff_w[c - 1];
            vs += w * cs;
            gs += w * (cs + krc * sn);
            ss += w * sn;
            kc += w * krc * cs;
            cs1 += w * fc * sn;
            c2c += w * fc * fc * cs;
        }
        const float ad = td[i], pd = td[N + i];
        // n-part and the radial factor of the gradient part (times rhat)
        nt += ad * vs * ir + pd * s.w * (-ss * ir) + kd * s.w * (-cs1);
        const float radial = ad * (-gs * ir * ir) + pd * s.w * (ss - kc) * ir * ir + kd * s.w * (-F.k_r * c2c);
        gt = gt + (radial * ir) * dv;
    }
}

template <>
__device__ __forceinline__ void field_jvp<kBessel>(const DevScene &S, const FieldP &F, f3 p, float scaling, float tau,
                                                   const float *td, float &nt, f3 &gt) {
    nt = 0.f;
    gt = mk3(0.f, 0.f, 0.f);
    if (p.x > S.egap || p.x < S.sgap) return;
    const BesselLocal b = bessel_local(S, F, p, scaling);
    const float k = F.k_r, m = (float)b.m, ir = 1.0f / b.r;
    const float dsdt = fabsf(sinf(tau)) < S.n_max_scaling ? cosf(tau) : 0.f;
    // per unit of A s: n = J C, grad = k J' C rh - (m J S / r) th
    const float cA = td[0] * scaling + td[2] * F.amp * dsdt;
    const float ck = td[1] * b.As;
    nt = cA * b.J * b.C + ck * b.dJ * b.r * b.C;
    const float gr = cA * k * b.dJ * b.C + ck * (b.dJ + k * b.r * b.d2J) * b.C;
    const float gtn = -cA * m * b.J * b.Sn * ir - ck * m * b.dJ * b.Sn;
    gt = gr * b.rh + gtn * b.th;
}
}  // namespace simpler
