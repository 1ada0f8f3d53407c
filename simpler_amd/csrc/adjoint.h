// adjoint.h -- path-replay adjoint of the photon tracer (included by
// render.hip after the shared path functions).
//
// Objective: F = sum over splats of  val * sum_taps G[tap] * w_tap  for an
// image-space weight G (= dF/dimage; a mask gives aperture throughput), or,
// for the per-ray mode, F = sum_rays gp . p_exit + gd . d_exit + gL L_exit.
// The gradient of F with respect to the field parameters is accumulated by
// re-tracing every photon from its counter-based seed:
//
//   forward   the photon is traced exactly as by the render kernel; every
//             main-path segment end (position, velocity, step pattern, event)
//             is logged in a lane-private record list, and every NEE / direct
//             sub-path segment end likewise in a short second list;
//   sub-path  when a sub-path reaches the sensor and its splat touches G, the
//             splat, the sensor-side exit map and the time-of-flight weight
//             are differentiated, and the sub-path is run BACKWARDS: the
//             leapfrog is inverted step by step from the exact end state
//             (no stored trajectory), the discrete adjoint of each step is
//             applied, and every visited point adds its parameter
//             derivatives (field_vjp); the result at the vertex (adjoint of
//             the vertex position and incoming velocity) is added to the
//             vertex's main-path record;
//   main      when the photon ends, the main path is run backwards from the
//             last vertex that received an adjoint, through Henyey-Greenstein
//             and face-reflection events, to the source.
//
// Each march segment ends on a face (refined to h / 10^precision) or when its
// distance budget is spent.  A face end is differentiated with the implicit
// crossing time: a perturbation (dp, dv) of the end state moves the crossing
// by ds = -nrm.dp / nrm.u (u = v/n, w = grad n), the end state by (u, w) ds
// and the path length by ds, which the adjoint applies as
//   lp <- lp - nrm (lp.u + lv.w + mu) / (nrm.u),
// with mu the adjoint of the path length (sub-paths: of the exp(-sigma d)
// weight; flights: minus the adjoint of the remaining distance budget).
#pragma once

namespace simpler {

enum { RP_NEE = 0, RP_DIRECT = 1, RP_RAY = 2, RP_MAIN = 3 };
enum { K_DIST = 0, K_FACE = 1 };
enum { A_REV_STEPS = 0, A_SUB_REV, A_MAIN_REV, A_DROPPED, A_N };
constexpr int kMaxAdjSources = 16;
constexpr int kRecFloats = 16;  // one 64-byte record

struct AdjState {
    Sym3 H;        // Hessian of n at the current reverse point
    f3 lp, lv;     // adjoint of (p, v)
    float pa;      // pending parameter derivative at the current point:
    f3 pb;         //   pa dn/dtheta + pb . dgrad/dtheta
    float mu, gL;  // path-length adjoint (see header), optical-length adjoint
    f3 ldin;       // NEE: adjoint of the main-path incoming velocity
    f3 src;        // the photon's source point: the exact start of the main path
    float rh;      // reverse iterator: step size, level, steps left, counts
    int rlev, rleft, rfinal, racc;
    uint32_t rcnt0, rpat;
    int rphase, rseg, nsub, nmain, last_lam;
    int slot;       // deferred re-trace: the photon's main-path slot, -1 none
    int cfg;        // batched kernels: the lane's configuration
    bool pending, trunc_main, trunc_sub;
    double F;
    uint32_t rsteps, nsubrev, nmainrev, dropped;
    float *log_main, *log_sub;
    GradSink G;
};

__device__ __forceinline__ int face_code(f3 nrm) {
    const int axis = nrm.x != 0.f ? 0 : (nrm.y != 0.f ? 1 : 2);
    const float s = comp(nrm, axis);
    return axis * 2 + (s > 0.f ? 1 : 0);
}
__device__ __forceinline__ f3 face_normal(int code) {
    const int axis = code >> 1;
    const float s = (code & 1) ? 1.f : -1.f;
    return mk3(axis == 0 ? s : 0.f, axis == 1 ? s : 0.f, axis == 2 ? s : 0.f);
}

// record: p(3) v(3) lam_p(3) lam_v(3) cnt0 pat kind unused
__device__ __forceinline__ void rec_write(float *base, int idx, f3 p, f3 v, uint32_t cnt0, uint32_t pat,
                                          uint32_t kind) {
    float4 *r = reinterpret_cast<float4 *>(base + (size_t)idx * kRecFloats);
    r[0] = make_float4(p.x, p.y, p.z, v.x);
    r[1] = make_float4(v.y, v.z, 0.f, 0.f);
    r[2] = make_float4(0.f, 0.f, 0.f, 0.f);
    r[3] = make_float4(__uint_as_float(cnt0), __uint_as_float(pat), __uint_as_float(kind), 0.f);
}
__device__ __forceinline__ void rec_read(const float *base, int idx, f3 &p, f3 &v, f3 &lamp, f3 &lamv,
                                         uint32_t &cnt0, uint32_t &pat, uint32_t &kind) {
    const float4 *r = reinterpret_cast<const float4 *>(base + (size_t)idx * kRecFloats);
    const float4 a = r[0], b = r[1], c = r[2], d = r[3];
    p = mk3(a.x, a.y, a.z);
    v = mk3(a.w, b.x, b.y);
    lamp = mk3(b.z, b.w, c.x);
    lamv = mk3(c.y, c.z, c.w);
    cnt0 = __float_as_uint(d.x);
    pat = __float_as_uint(d.y);
    kind = __float_as_uint(d.z);
}
__device__ __forceinline__ void rec_add_lambda(float *base, int idx, f3 lamp, f3 lamv) {
    float4 *r = reinterpret_cast<float4 *>(base + (size_t)idx * kRecFloats);
    float4 b = r[1], c = r[2];
    b.z += lamp.x;
    b.w += lamp.y;
    c.x += lamp.z;
    c.y += lamv.x;
    c.z += lamv.y;
    c.w += lamv.z;
    r[1] = b;
    r[2] = c;
}

__device__ __forceinline__ float level_h(float h0, int lev) {
    float h = h0;
    for (int i = 0; i < lev; ++i) h = div10(h);  // the forward refinement chain
    return h;
}

// Next reverse step of the current segment: the kept overshoot step first
// (not counted in the optical length), then the accepted steps level by level
// down to level 0.  Returns false at the segment start.
__device__ __forceinline__ bool rev_next(AdjState &a, float h0) {
    if (a.rfinal) {
        a.rfinal = 0;
        a.racc = 0;
        return true;
    }
    while (a.rleft == 0) {
        if (a.rlev == 0) return false;
        --a.rlev;
        a.rleft = a.rlev == 0 ? (int)a.rcnt0 : (int)((a.rpat >> (4 * (a.rlev - 1))) & 15u);
        a.rh = level_h(h0, a.rlev);
    }
    --a.rleft;
    a.racc = 1;
    return true;
}

__device__ __forceinline__ void rev_init(Hot &h, AdjState &a, float h0, uint32_t cnt0, uint32_t pat) {
    a.rcnt0 = cnt0;
    a.rpat = pat;
    a.rlev = (int)((pat >> 28) & 7u);
    a.rfinal = (int)(pat >> 31);
    a.rleft = a.rlev == 0 ? (int)cnt0 : (int)((pat >> (4 * (a.rlev - 1))) & 15u);
    a.rh = level_h(h0, a.rlev);
    h.state = rev_next(a, h0) ? ST_RMARCH : ST_REND;
}

// lp <- lp - nrm (lp.u + lv.w + mu) / nrm.u  (implicit face-crossing time)
__device__ __forceinline__ void face_correct(AdjState &a, f3 nrm, f3 u, f3 w, float mu) {
    const float nu = dot(nrm, u);
    if (fabsf(nu) < 1e-20f) return;
    const float c = (dot(a.lp, u) + dot(a.lv, w) + mu) / nu;
    a.lp = a.lp - c * nrm;
}

__device__ __forceinline__ f3 reflect_vec(f3 x, f3 nrm) { return x - (2.f * dot(x, nrm)) * nrm; }

// One reverse leapfrog step from the current point q (state p = q, v, with
// n, grad n, H at q) back to the step's start point, applying the step's
// discrete adjoint:
//   forward   vh = v + h/2 g(p);  q = p + h vh / n(p);  v' = vh + h/2 g(q)
//   adjoint   lp += h/2 H(q) lv;           (end kick; pending b += h/2 lv)
//             [accepted step: L += h n(q):  lp += gL h g(q); pending a += gL h]
//             lvh = lv + h/n(p) lp;  s = h/n(p)^2 (vh . lp);  lp -= s g(p)
//             lp += h/2 H(p) lvh;  lv = lvh  (start kick; at p: a = -s, b = h/2 lvh)
// The start point is p = q - h vh / n(p), solved with n evaluated at
// q - h vh / n(q) (the error is O(h^3 |grad n|^2), far below FP32).
// Returns false when the segment start is reached.
// CELL: the spline taps come from the lane's register copy of the current cell;
// ACC: the spline parameter derivatives go to the lane's register sums.
template <int FIELD, bool CELL = false, bool ACC = false>
__device__ __forceinline__ bool rev_step(Hot &h, AdjState &a, const DevScene &S, const FieldP &F,
                                         SplineCell *cell = nullptr, SplineAcc *acc = nullptr) {
    const float hs = a.rh;
    const float hh = 0.5f * hs;
    a.lp = a.lp + hh * symmul(a.H, a.lv);
    a.pb = a.pb + hh * a.lv;
    if (a.racc && a.gL != 0.f) {
        a.lp = a.lp + (a.gL * hs) * h.g0;
        a.pa += a.gL * hs;
    }
    if constexpr (FIELD == kSpline && ACC)
        spline_vjp_acc(S, h.p, a.pa, a.pb, a.G, *acc);
    else
        field_vjp<FIELD>(S, F, h.p, h.scaling, h.tau, a.pa, a.pb, a.G);
    const f3 vh = h.v - hh * h.g0;
    const f3 pt = h.p - (hs / h.n0) * vh;
    float n1;
    f3 g1;
    Sym3 H1;
    if constexpr (FIELD == kSpline && CELL)
        spline_eval_h<true>(S, pt, n1, g1, H1, cell);
    else
        field_eval_h<FIELD>(S, F, pt, h.scaling, h.tau, n1, g1, H1);
    const float inv = 1.0f / n1;
    const f3 p0 = h.p - (hs * inv) * vh;
    const float s = hs * inv * inv * dot(vh, a.lp);
    const f3 lvh = a.lv + (hs * inv) * a.lp;
    a.lp = a.lp - s * g1 + hh * symmul(H1, lvh);
    a.lv = lvh;
    a.pa = -s;
    a.pb = hh * lvh;
    h.v = vh - hh * g1;
    h.p = p0;
    h.n0 = n1;
    h.g0 = g1;
    a.H = H1;
    ++a.rsteps;
    return rev_next(a, S.h0);
}

// Henyey-Greenstein scatter v_out = |v_in| T(v_in / |v_in|) with
// T(d) = alpha (c x d) + beta c + gamma d, c = c(d) the frame vector of
// hg_sample; (alpha, beta, gamma) are read back from d and v_out.  Returns
// the adjoint of v_in for the adjoint lvo of v_out.
__device__ __forceinline__ f3 hg_vjp(f3 vin, f3 vout, f3 lvo) {
    const float mag = len(vin);
    const f3 d = vin / mag;
    const f3 d1 = normalized(vout);
    const bool bx = fabsf(d.x) > fabsf(d.y);
    const f3 chat = bx ? mk3(d.z, 0.f, -d.x) : mk3(0.f, d.z, -d.y);
    const float cl = len(chat);
    const f3 c = chat / cl;
    const f3 b = cross(c, d);
    const float al = dot(d1, b), be = dot(d1, c), ga = dot(d1, d);
    const f3 ld1 = mag * lvo;
    const float lmag = dot(d1, lvo);
    f3 ld = ga * ld1 + al * cross(ld1, c);
    const f3 lc = al * cross(d, ld1) + be * ld1;
    const f3 lch = (lc - dot(c, lc) * c) / cl;
    if (bx) {
        ld.z += lch.x;
        ld.x -= lch.z;
    } else {
        ld.z += lch.y;
        ld.y -= lch.z;
    }
    return (ld - dot(d, ld) * d) / mag + lmag * d;
}

// Splat weights of scene.cpp:782-829 against G: returns false when the splat
// misses the gate / image or every tap of G is zero.  Wsum = sum G w,
// lps = dF/dp_sensor for F = val * Wsum.
__device__ __forceinline__ bool splat_adj(const DevScene &S, const RenderArgs &A, const float *G, f3 p, float L,
                                          int depth, float val, float &Wsum, f3 &lps) {
    float x = dot(S.cam_h, p) - S.cam_o.y;
    float y = dot(S.cam_v, p) - S.cam_o.z;
    if (!(S.nogate || (L > S.pl0 && L < S.pl1))) return false;
    x = (x / S.view_plane[0] + 0.5f) * (float)A.X;
    y = (y / S.view_plane[1] + 0.5f) * (float)A.Y;
    const float fx0 = floorf(x), fy0 = floorf(y);
    if (!(fx0 > -4.f && fx0 < (float)A.X + 4.f && fy0 > -4.f && fy0 < (float)A.Y + 4.f)) return false;
    const int ix = (int)fx0, iy = (int)fy0;
    int iz;
    if (S.bounce) {
        iz = depth;
    } else if (S.nogate) {
        iz = 0;
    } else {
        const float z = ((L - S.pl0) / (S.pl1 - S.pl0)) * (float)A.Z;
        const float fz = floorf(z);
        if (!(fz > -4.f && fz < (float)A.Z + 4.f)) return false;
        iz = (int)fz;
    }
    if (iz < 0 || iz >= A.Z) return false;
    const float *plane = G + (size_t)iz * A.X * A.Y;
    float g[4];
    const int tx[4] = {ix, ix + 1, ix, ix + 1};
    const int ty[4] = {iy, iy, iy + 1, iy + 1};
    bool any = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        g[k] = (tx[k] >= 0 && tx[k] < A.X && ty[k] >= 0 && ty[k] < A.Y) ? plane[(size_t)ty[k] * A.X + tx[k]] : 0.f;
        any = any || g[k] != 0.f;
    }
    if (!any) return false;
    const float fx = x - fx0, fy = y - fy0;
    Wsum = g[0] * (1.f - fx) * (1.f - fy) + g[1] * fx * (1.f - fy) + g[2] * (1.f - fx) * fy + g[3] * fx * fy;
    const float dfx = val * ((1.f - fy) * (g[1] - g[0]) + fy * (g[3] - g[2]));
    const float dfy = val * ((1.f - fx) * (g[2] - g[0]) + fx * (g[3] - g[1]));
    lps = (dfx * (float)A.X / S.view_plane[0]) * S.cam_h + (dfy * (float)A.Y / S.view_plane[1]) * S.cam_v;
    return true;
}

// Adjoint of exit_to_sensor (photon.cpp:53-80, no sensor lens) for
// F(p_sensor, dts, refr): given lps = dF/dp_sensor and the adjoints of the
// unit directions (ldts, lrefr), returns the adjoint of the face point, of
// the velocity v and of the local index n used by the refraction.
__device__ __forceinline__ void exit_adj(const DevScene &S, f3 pface, f3 v, float nloc, f3 lps, f3 ldts, f3 lrefr,
                                         f3 &lp, f3 &lv, float &ln) {
    const float vl = len(v);
    const f3 dts = v / vl;
    f3 rt = dts;
    if (nloc > 1.f) rt.x = rt.x / nloc;
    const float rl = len(rt);
    const f3 refr = rt / rl;
    const float dist = (S.cam_o.x - pface.x) / refr.x;
    const float lr = dot(lps, refr) / refr.x;
    lp = lps - mk3(lr, 0.f, 0.f);
    lrefr = lrefr + dist * lps - mk3(dist * lr, 0.f, 0.f);
    const f3 lrt = (lrefr - dot(refr, lrefr) * refr) / rl;
    ln = 0.f;
    if (nloc > 1.f) {
        ldts = ldts + mk3(lrt.x / nloc, lrt.y, lrt.z);
        ln = -lrt.x * dts.x / (nloc * nloc);
    } else {
        ldts = ldts + lrt;
    }
    lv = (ldts - dot(dts, ldts) * dts) / vl;
}

__device__ __forceinline__ bool adj_main_pending(const AdjState *a) { return a->pending && !a->trunc_main; }

__device__ __forceinline__ void adj_photon_begin(AdjState &a) {
    a.nmain = 0;
    a.nsub = 0;
    a.last_lam = -1;
    a.slot = -1;
    a.pending = false;
    a.trunc_main = false;
    a.trunc_sub = false;
    a.gL = 0.f;
}

__device__ __forceinline__ void adj_sub_begin(AdjState &a) {
    a.nsub = 0;
    a.trunc_sub = false;
}

// Log the end of a main-path segment (before its face reflection / vertex).
__device__ __forceinline__ void adj_log_main(Lane &L, int kind, f3 nrm) {
    AdjState &a = *L.a;
    if (a.trunc_main) return;
    if (a.nmain >= L.A.main_cap) {
        a.trunc_main = true;
        return;
    }
    const uint32_t k = kind == K_FACE ? (uint32_t)(K_FACE | (face_code(nrm) << 1)) : (uint32_t)K_DIST;
    rec_write(a.log_main, a.nmain, L.h.p, L.h.v, L.h.fcnt0, L.h.fpat, k);
    ++a.nmain;
}

// Log the end of an NEE / direct sub-path segment that reflects on a face.
__device__ __forceinline__ void adj_log_sub(Lane &L, f3 nrm) {
    AdjState &a = *L.a;
    if (a.trunc_sub) return;
    if (a.nsub >= L.A.sub_cap) {
        a.trunc_sub = true;
        return;
    }
    rec_write(a.log_sub, a.nsub, L.h.p, L.h.v, L.h.fcnt0, L.h.fpat, (uint32_t)(K_FACE | (face_code(nrm) << 1)));
    ++a.nsub;
}

// ---------------------------------------------------------------------------
// Deferred image adjoint.  The re-trace (M_ADJDEF) does not run backward
// runs itself: a splat that touches G queues a sub-path job (its end state,
// step pattern, face records, adjoint seeds and start point), and a photon
// with a queued next-event sub-path stores its main-path log in a slot when
// it ends.  k_adj_rev then runs the jobs (M_ADJSUB: each adds its vertex
// adjoint to the slot's record) and afterwards the slots (M_ADJMAIN), every
// lane of a wave running backwards, instead of the few reversing lanes of a
// mostly marching wave.  A queue that fills sets qcnt[Q_OVER]; the host then
// re-runs that chunk of photons in smaller pieces.

__device__ __forceinline__ void adj_overflow(const RenderArgs &A) { atomicMax(A.qcnt + Q_OVER, 1ull); }

__device__ __forceinline__ void copy_records(float *dst, const float *src, int n) {
    float4 *d = reinterpret_cast<float4 *>(dst);
    const float4 *s = reinterpret_cast<const float4 *>(src);
    for (int i = 0; i < 4 * n; ++i) d[i] = s[i];
}

// Queue the current sub-path's backward run (the state sub_rev_start takes).
__device__ __forceinline__ void adj_defer_sub(Lane &L, f3 lp, f3 lv, float mu, float ln, int phase) {
    AdjState &a = *L.a;
    const RenderArgs &A = L.A;
    const Hot &h = L.h;
    int slot = -1;
    if (phase == RP_NEE) {
        if (a.slot < 0) {
            const unsigned long long sl = atomicAdd(A.qcnt + Q_SLOTS, 1ull);
            if (sl >= A.slot_cap) {
                adj_overflow(A);
                return;
            }
            a.slot = (int)sl;
        }
        slot = a.slot;
        a.last_lam = a.nmain - 1;
    }
    const unsigned long long j = atomicAdd(A.qcnt + Q_JOBS, 1ull);
    if (j >= A.job_cap) {
        adj_overflow(A);
        return;
    }
    unsigned long long sb = 0;
    if (a.nsub > 0) {
        sb = atomicAdd(A.qcnt + Q_JOBREC, (unsigned long long)a.nsub);
        if (sb + (unsigned long long)a.nsub > A.jobrec_cap) {
            adj_overflow(A);
            return;
        }
        copy_records(A.jobrec + sb * kRecFloats, a.log_sub, a.nsub);
    }
    float4 *r = reinterpret_cast<float4 *>(A.jobs + j * 32);
    r[0] = make_float4(h.p.x, h.p.y, h.p.z, h.v.x);
    r[1] = make_float4(h.v.y, h.v.z, __uint_as_float(h.fcnt0), __uint_as_float(h.fpat));
    r[2] = make_float4(lp.x, lp.y, lp.z, mu);
    r[3] = make_float4(lv.x, lv.y, lv.z, ln);
    r[4] = make_float4(a.ldin.x, a.ldin.y, a.ldin.z, h.scaling);
    r[5] = make_float4(L.c.mp.x, L.c.mp.y, L.c.mp.z, h.tau);
    r[6] = make_float4(__int_as_float(slot), __int_as_float(a.nmain - 1), __int_as_float(phase),
                       __int_as_float(a.nsub));
    r[7] = make_float4(__uint_as_float((uint32_t)sb), __uint_as_float((uint32_t)(sb >> 32)), __int_as_float(a.cfg),
                       0.f);
}

// Photon end of the deferred re-trace: store the main-path log of a photon
// with queued next-event sub-paths (slot layout: source, scaling; tau,
// record count, last vertex with an adjoint, valid; record base,
// configuration).
__device__ __forceinline__ void adj_store_slot(Lane &L) {
    AdjState &a = *L.a;
    const RenderArgs &A = L.A;
    float4 *r = reinterpret_cast<float4 *>(A.slots + (size_t)a.slot * 16);
    int valid = 0;
    unsigned long long base = 0;
    if (!a.trunc_main && a.nmain > 0) {
        base = atomicAdd(A.qcnt + Q_SLOTREC, (unsigned long long)a.nmain);
        if (base + (unsigned long long)a.nmain > A.slotrec_cap) {
            adj_overflow(A);
        } else {
            copy_records(A.slotrec + base * kRecFloats, a.log_main, a.nmain);
            valid = 1;
        }
    }
    r[0] = make_float4(a.src.x, a.src.y, a.src.z, L.h.scaling);
    r[1] = make_float4(L.h.tau, __int_as_float(a.nmain), __int_as_float(a.last_lam), __int_as_float(valid));
    r[2] = make_float4(__uint_as_float((uint32_t)base), __uint_as_float((uint32_t)(base >> 32)), __int_as_float(a.cfg),
                       0.f);
}

// Start the backward run of the current sub-path from its sensor-face end
// state (h.p, h.v): lp, lv = adjoint of the end state, mu = adjoint of the
// travelled distance, ln = adjoint of n at the exit point.
template <int FIELD>
__device__ __forceinline__ void sub_rev_start(Lane &L, f3 lp, f3 lv, float mu, float ln, int phase) {
    AdjState &a = *L.a;
    Hot &h = L.h;
    const DevScene &S = L.S;
    float n;
    f3 g;
    field_eval_h<FIELD>(S, L.v->f, h.p, h.scaling, h.tau, n, g, a.H);
    h.n0 = n;
    h.g0 = g;
    a.lp = lp + ln * g;
    a.lv = lv;
    a.mu = mu;
    face_correct(a, S.cam_dir, h.v / n, g, mu + a.gL * n);
    a.pa = ln;
    a.pb = mk3(0.f, 0.f, 0.f);
    a.rseg = a.nsub;
    a.rphase = phase;
    ++a.nsubrev;
    rev_init(h, a, S.h0, h.fcnt0, h.fpat);
}

// NEE splat (scene.cpp:897-929) differentiated:
//   val = W 2 pi exp(-sigma d) f_HG(din . dts) fore fw,
//   fore = (refr . c) / (dts . c); fw (Fresnel transmission) is held fixed.
// DEFER: queue the backward run (adj_defer_sub) and return false, so that the
// forward run continues; else start it here.
template <int FIELD, bool DEFER = false>
__device__ __forceinline__ bool nee_adj_start(Lane &L, f3 ps, float Lv, float val, f3 din, f3 dts, float fore) {
    AdjState &a = *L.a;
    const DevScene &S = L.S;
    float Wsum;
    f3 lps;
    if (!splat_adj(S, L.A, L.v->image, ps, Lv, L.c.depth, val, Wsum, lps)) return false;
    a.F += (double)val * (double)Wsum;
    if (a.trunc_sub || a.trunc_main) {
        ++a.dropped;
        return false;
    }
    const float g = L.v->g;
    const float x = 1.f + g * g - 2.f * g * dot(din, dts);
    const float kh = Wsum * val * 3.f * g / x;  // dF / d(din . dts)
    f3 ldts = kh * din;
    const f3 ldin = kh * dts;
    const float dc = dot(dts, S.cam_dir);
    const float lfore = fore != 0.f ? Wsum * val / fore : 0.f;
    const f3 lrefr = (lfore / dc) * S.cam_dir;
    ldts = ldts - (lfore * fore / dc) * S.cam_dir;
    f3 lp, lv;
    float ln;
    exit_adj(S, L.h.p, L.h.v, L.h.n0, lps, ldts, lrefr, lp, lv, ln);
    const float ml = len(L.c.mv);
    a.ldin = (ldin - dot(din, ldin) * din) / ml;
    if constexpr (DEFER) {
        adj_defer_sub(L, lp, lv, -L.v->sigma_t * Wsum * val, ln, RP_NEE);
        return false;
    }
    sub_rev_start<FIELD>(L, lp, lv, -L.v->sigma_t * Wsum * val, ln, RP_NEE);
    return true;
}

// Direct splat (photon.cpp:53-80): val = W exp(-sigma d) fw.
template <int FIELD, bool DEFER = false>
__device__ __forceinline__ bool direct_adj_start(Lane &L, f3 ps, float Lv, float val) {
    AdjState &a = *L.a;
    const DevScene &S = L.S;
    float Wsum;
    f3 lps;
    if (!splat_adj(S, L.A, L.v->image, ps, Lv, 0, val, Wsum, lps)) return false;
    a.F += (double)val * (double)Wsum;
    if (a.trunc_sub) {
        ++a.dropped;
        return false;
    }
    f3 lp, lv;
    float ln;
    exit_adj(S, L.h.p, L.h.v, L.h.n0, lps, mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f), lp, lv, ln);
    if constexpr (DEFER) {
        adj_defer_sub(L, lp, lv, -L.v->sigma_t * Wsum * val, ln, RP_DIRECT);
        return false;
    }
    sub_rev_start<FIELD>(L, lp, lv, -L.v->sigma_t * Wsum * val, ln, RP_DIRECT);
    return true;
}

// Per-ray objective gp . p + gd . d + gL L at the sensor-face exit.
template <int FIELD>
__device__ __forceinline__ bool ray_adj_start(Lane &L) {
    AdjState &a = *L.a;
    const RenderArgs &A = L.A;
    const Hot &h = L.h;
    const size_t i = (size_t)((((uint64_t)L.c.pid_hi << 32) | L.c.pid_lo) - A.offset);
    const f3 gp = mk3(A.up_p[3 * i], A.up_p[3 * i + 1], A.up_p[3 * i + 2]);
    const f3 gd = mk3(A.up_d[3 * i], A.up_d[3 * i + 1], A.up_d[3 * i + 2]);
    const float gL = A.up_L[i];
    const float vl = len(h.v);
    const f3 d = h.v / vl;
    a.F += (double)dot(gp, h.p) + (double)dot(gd, d) + (double)gL * (double)h.L;
    if (a.trunc_sub) {
        ++a.dropped;
        return false;
    }
    a.gL = gL;
    sub_rev_start<FIELD>(L, gp, (gd - dot(d, gd) * d) / vl, 0.f, 0.f, RP_RAY);
    return true;
}

}  // namespace simpler
