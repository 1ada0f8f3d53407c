// adjoint_kernel.h -- the backward-run events and the persistent adjoint
// kernel (included by render.hip after the path events it resumes).
#pragma once

namespace simpler {

// Load a logged segment end as the current reverse point.
template <int FIELD>
__device__ __forceinline__ void rev_load(Lane &L, const float *base, int idx, f3 &lamp, f3 &lamv, uint32_t &cnt0,
                                         uint32_t &pat, uint32_t &kind) {
    AdjState &a = *L.a;
    Hot &h = L.h;
    f3 p, v;
    rec_read(base, idx, p, v, lamp, lamv, cnt0, pat, kind);
    h.p = p;
    h.v = v;
    field_eval_h<FIELD>(L.S, L.v->f, p, h.scaling, h.tau, h.n0, h.g0, a.H);
    a.pa = 0.f;
    a.pb = mk3(0.f, 0.f, 0.f);
}

// Photon end with adjoints on its vertices: start the backward run of the main
// path at the last such vertex.  mu is minus the adjoint of the flight's
// distance budget, which moves the vertex along the path.
template <int FIELD>
__device__ __forceinline__ void main_rev_init(Lane &L) {
    AdjState &a = *L.a;
    Hot &h = L.h;
    a.rphase = RP_MAIN;
    a.rseg = a.last_lam;
    a.gL = 0.f;
    f3 lamp, lamv;
    uint32_t cnt0, pat, kind;
    rev_load<FIELD>(L, a.log_main, a.rseg, lamp, lamv, cnt0, pat, kind);
    a.lp = lamp;
    a.lv = lamv;
    a.mu = -(dot(a.lp, h.v / h.n0) + dot(a.lv, h.g0));
    ++a.nmainrev;
    rev_init(h, a, L.S.h0, cnt0, pat);
}

// The start point of a logged segment end (the next segment's start).
__device__ __forceinline__ f3 rec_point(const float *base, int idx) {
    const float4 r = *reinterpret_cast<const float4 *>(base + (size_t)idx * kRecFloats);
    return mk3(r.x, r.y, r.z);
}

// The backward run reached the start of a segment.  The pending parameter
// derivatives of the start point are added at the point the forward run
// started from -- the logged segment end, the NEE vertex or the source --
// not at the backward run's reconstruction of it, which drifts by rounding
// (a source on a gap plane could otherwise fall just outside the gap).
template <int FIELD, int MODE>
__device__ __forceinline__ void rev_end(Lane &L) {
    AdjState &a = *L.a;
    Hot &h = L.h;
    const DevScene &S = L.S;
    const FieldP &F = L.v->f;
    f3 lamp, lamv;
    uint32_t cnt0, pat, kind;
    if (a.rphase == RP_MAIN) {
        if (a.rseg == 0) {  // the source: v0 = n(p0) d0 with p0, d0 fixed
            h.p = a.src;
            field_vjp<FIELD>(S, F, h.p, h.scaling, h.tau, a.pa, a.pb, a.G);
            const float vl = len(h.v);
            field_vjp<FIELD>(S, F, h.p, h.scaling, h.tau, dot(a.lv, h.v) / vl, mk3(0.f, 0.f, 0.f), a.G);
            a.pending = false;
            h.state = ST_NEED;
            return;
        }
        const f3 vout = h.v;
        --a.rseg;
        field_vjp<FIELD>(S, F, rec_point(a.log_main, a.rseg), h.scaling, h.tau, a.pa, a.pb, a.G);
        rev_load<FIELD>(L, a.log_main, a.rseg, lamp, lamv, cnt0, pat, kind);
        if (kind & K_FACE) {  // mirror reflection, then the face-crossing time
            const f3 nrm = face_normal((int)(kind >> 1));
            a.lv = reflect_vec(a.lv, nrm);
            face_correct(a, nrm, h.v / h.n0, h.g0, a.mu);
        } else {  // Henyey-Greenstein vertex: a new flight begins here
            a.lv = hg_vjp(h.v, vout, a.lv);
            a.lp = a.lp + lamp;
            a.lv = a.lv + lamv;
            a.mu = -(dot(a.lp, h.v / h.n0) + dot(a.lv, h.g0));
        }
        rev_init(h, a, S.h0, cnt0, pat);
        return;
    }
    if (a.rseg > 0) {  // a face reflection inside the sub-path
        --a.rseg;
        field_vjp<FIELD>(S, F, rec_point(a.log_sub, a.rseg), h.scaling, h.tau, a.pa, a.pb, a.G);
        rev_load<FIELD>(L, a.log_sub, a.rseg, lamp, lamv, cnt0, pat, kind);
        const f3 nrm = face_normal((int)(kind >> 1));
        a.lv = reflect_vec(a.lv, nrm);
        face_correct(a, nrm, h.v / h.n0, h.g0, a.mu + a.gL * h.n0);
        rev_init(h, a, S.h0, cnt0, pat);
        return;
    }
    // sub-path start (the NEE vertex, or the source for the direct trace):
    // v0 = n(p0) omega with omega fixed
    h.p = L.c.mp;
    field_vjp<FIELD>(S, F, h.p, h.scaling, h.tau, a.pa, a.pb, a.G);
    field_eval<FIELD>(S, F, h.p, h.scaling, h.tau, h.n0, h.g0);
    const float vl = len(h.v);
    const float lw = dot(a.lv, h.v) / vl;
    field_vjp<FIELD>(S, F, h.p, h.scaling, h.tau, lw, mk3(0.f, 0.f, 0.f), a.G);
    a.gL = 0.f;
    if constexpr (MODE == M_ADJ) {
        if (a.rphase == RP_NEE) {
            rec_add_lambda(a.log_main, a.nmain - 1, a.lp + lw * h.g0, a.ldin);
            a.pending = true;
            a.last_lam = a.nmain - 1;
            nee_continue<FIELD>(L);
        } else {
            direct_continue<FIELD>(L, 1);
        }
    } else if constexpr (MODE == M_ADJSUB) {
        // a queued sub-path: its vertex adjoint goes to the photon's stored
        // main-path record (none when the log did not fit)
        if (a.rphase == RP_NEE && a.log_main) rec_add_lambda(a.log_main, a.nmain - 1, a.lp + lw * h.g0, a.ldin);
        h.state = ST_NEED;
    } else {
        photon_end(L);
    }
}

// A lane's source-parameter accumulators and objective, added to the
// configuration's gradient and objective (batched adjoint: when the lane moves
// to another configuration, and at the end).
__device__ __forceinline__ void adj_flush_cfg(AdjState &a, const RenderArgs &A, int cfg, int npar) {
    if (cfg < 0) return;
    for (int j = 0; j < npar; ++j) {
        const float x = a.G.acc[j * a.G.stride];
        if (x != 0.f) atomicAdd(A.grad + (size_t)cfg * npar + j, x);
        a.G.acc[j * a.G.stride] = 0.f;
    }
    if (a.F != 0.0) atomicAdd(A.obj + cfg, a.F);
    a.F = 0.0;
}

// The persistent adjoint kernel: the render kernel's fetch / events / march
// loop plus the backward events (ST_REND, ST_MREV) and backward steps.
// BATCH: photons of many point-source patterns (per-pattern amplitudes,
// phases, photon budget, seed, objective weights, gradient and objective).
// M_ADJDEF: the deferred image adjoint's re-trace: forward runs and logs
// only, backward runs queued for k_adj_rev (no parameter derivatives here).
// The spline re-trace is held to 256 registers (2 waves per SIMD, as the
// spline render); it needs a few more, which spill.
template <int FIELD, int MODE, bool BATCH = false>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(MODE == M_ADJDEF && FIELD == kSpline ? 2 : 1)))
    k_adjoint(DevScene S, RenderArgs A) {
    constexpr bool DEF = MODE == M_ADJDEF;
    __shared__ Cold cold[kBlock];
    __shared__ uint32_t cnt[C_N];
    __shared__ float sacc[DEF ? 1
                              : (FIELD == kSources ? (2 * kMaxAdjSources + 1) * kBlock
                                                   : (FIELD == kBessel ? 3 * kBlock : 1))];
    __shared__ float vcell[FIELD == kSpline && !DEF ? 64 * kBlock : 1];  // spline VJP cell sums
    if (FIELD == kSpline && !DEF)
        for (int t = 0; t < 64; ++t) vcell[t * kBlock + threadIdx.x] = 0.f;
    __shared__ Var vars[BATCH ? kBlock : 1];
    if (threadIdx.x < C_N) cnt[threadIdx.x] = 0u;
    const int npar = DEF ? 0 : (FIELD == kSources ? 2 * S.n_src + 1 : (FIELD == kBessel ? 3 : 0));
    for (int j = 0; j < npar; ++j) sacc[j * kBlock + threadIdx.x] = 0.f;
    __syncthreads();
    Var vsingle = var_from(S, A);
    vsingle.image = const_cast<float *>(A.G);  // the objective weights, read by the splat adjoint
    Var &v = BATCH ? vars[threadIdx.x] : vsingle;
    int cur_cfg = -1;
    Hot h;
    h.steps = 0;
    h.state = ST_NEED;
    h.dead = false;
    AdjState a;
    SplineCell scell;  // spline kernels: register copy of the current cell's taps
    scell.key = -1;
    a.F = 0.0;
    a.rsteps = a.nsubrev = a.nmainrev = a.dropped = 0;
    a.pending = a.trunc_main = a.trunc_sub = false;
    a.slot = -1;
    a.cfg = 0;
    a.gL = 0.f;
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    a.log_main = A.log + gid * (size_t)(A.main_cap + A.sub_cap) * kRecFloats;
    a.log_sub = a.log_main + (size_t)A.main_cap * kRecFloats;
    a.G.grid = FIELD == kSpline && A.gx ? A.gx + (size_t)xcd_id() * A.gx_stride : A.grad;
    a.G.acc = sacc + threadIdx.x;
    a.G.stride = kBlock;
    a.G.cell = FIELD == kSpline && !DEF ? vcell + threadIdx.x : nullptr;
    a.G.kx = -1;
    Lane L{h, cold[threadIdx.x], &v, S, A, cnt, &a, nullptr, nullptr, nullptr, true};
    const int lane = __lane_id();
    while (true) {
        const unsigned long long need = __ballot(h.state == ST_NEED);
        if (need) {
            const int leader = __ffsll((long long)need) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(A.work, (unsigned long long)__popcll(need));
            base = __shfl(base, leader);
            if (h.state == ST_NEED) {
                const unsigned long long rank = __popcll(need & ((1ull << lane) - 1ull));
                const unsigned long long idx = base + rank;
                if (idx < A.n) {
                    uint64_t pid = A.offset + idx;
                    if (BATCH) {
                        // the photon's index over all configurations (A.offset:
                        // the start of a deferred chunk)
                        const unsigned long long gi = A.offset + idx;
                        int lo = 0, hi = A.n_cfg;  // largest c with prefix[c] <= gi
                        while (hi - lo > 1) {
                            const int mid = (lo + hi) >> 1;
                            if (A.prefix[mid] <= gi)
                                lo = mid;
                            else
                                hi = mid;
                        }
                        pid = gi - A.prefix[lo];
                        if (lo != cur_cfg) {
                            adj_flush_cfg(a, A, cur_cfg, npar);
                            cur_cfg = lo;
                            a.cfg = lo;
                            v = A.cfgs[lo];
                            v.image = const_cast<float *>(A.G) + (size_t)lo * A.X * A.Y * A.Z;
                        }
                    }
                    L.c.pid_lo = (uint32_t)pid;
                    L.c.pid_hi = (uint32_t)(pid >> 32);
                    h.state = ST_INIT;
                } else {
                    h.state = ST_DONE;
                }
            }
        }
        if (h.state == ST_INIT) photon_init<FIELD, MODE>(L);
        if (h.state == ST_SEG_END) seg_end<FIELD, MODE>(L);
        if constexpr (DEF) {  // a photon with queued next-event sub-paths ended: store its main path
            if (h.state == ST_NEED && a.slot >= 0) {
                adj_store_slot(L);
                a.slot = -1;
            }
        } else {
            if (h.state == ST_REND) rev_end<FIELD, MODE>(L);
            if (h.state == ST_MREV) main_rev_init<FIELD>(L);
        }
        if (__ballot(h.state != ST_DONE) == 0ull) break;
        if (__ballot(h.state == ST_MARCH || h.state == ST_RMARCH) == 0ull) continue;
        const FieldP F = v.f;
        const float Lcut = v.Lcut;
        bool m = h.state == ST_MARCH, r = h.state == ST_RMARCH;
        uint32_t sl = h.steps_left;
        for (int k = 0; k < A.spc; ++k) {
            if (m)
                m = !march_step<FIELD, true, -1, true, -1, FIELD == kSpline>(h, sl, S, F, Lcut, &scell);
            else if (!DEF && r)
                r = rev_step<FIELD, FIELD == kSpline>(h, a, S, F, &scell);
        }
        h.steps_left = sl;
        if (h.state == ST_MARCH && !m) h.state = ST_SEG_END;
        if (h.state == ST_RMARCH && !r) h.state = ST_REND;
    }
    if (BATCH) adj_flush_cfg(a, A, cur_cfg, npar);
    if (FIELD == kSpline && !DEF) spline_cell_flush(S, a.G);
    unsigned long long st = h.steps;
    double fo = a.F;
    unsigned long long ac[A_N] = {a.rsteps, a.nsubrev, a.nmainrev, a.dropped};
    for (int off = 32; off > 0; off >>= 1) {
        st += __shfl_xor(st, off);
        fo += __shfl_xor(fo, off);
#pragma unroll
        for (int k = 0; k < A_N; ++k) ac[k] += __shfl_xor(ac[k], off);
    }
    if (lane == 0) {
        if (st) atomicAdd(A.counters + C_STEPS, st);
        if (fo != 0.0) atomicAdd(A.obj, fo);
#pragma unroll
        for (int k = 0; k < A_N; ++k)
            if (ac[k]) atomicAdd(A.acnt + k, ac[k]);
    }
    __syncthreads();
    if (threadIdx.x < C_N && threadIdx.x != C_STEPS && cnt[threadIdx.x])
        atomicAdd(A.counters + threadIdx.x, (unsigned long long)cnt[threadIdx.x]);
    for (int j = threadIdx.x; j < (BATCH ? 0 : npar); j += kBlock) {
        float sum = 0.f;
        for (int t = 0; t < kBlock; ++t) sum += sacc[j * kBlock + t];
        if (sum != 0.f) atomicAdd(A.grad + j, sum);
    }
}

// The configuration of a queued job or stored slot (batched adjoint).
template <int MODE>
__device__ __forceinline__ int item_cfg(const RenderArgs &A, unsigned long long i) {
    const float *r = MODE == M_ADJSUB ? A.jobs + i * 32 + 30 : A.slots + i * 16 + 10;
    return __float_as_int(*r);
}

// A queued sub-path (adj_defer_sub's job layout) as the lane's backward run.
template <int FIELD>
__device__ __forceinline__ void job_load(Lane &L, unsigned long long j) {
    AdjState &a = *L.a;
    Hot &h = L.h;
    const RenderArgs &A = L.A;
    const float4 *r = reinterpret_cast<const float4 *>(A.jobs + j * 32);
    const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4], r5 = r[5], r6 = r[6], r7 = r[7];
    h.p = mk3(r0.x, r0.y, r0.z);
    h.v = mk3(r0.w, r1.x, r1.y);
    h.fcnt0 = __float_as_uint(r1.z);
    h.fpat = __float_as_uint(r1.w);
    h.scaling = r4.w;
    h.tau = r5.w;
    a.ldin = mk3(r4.x, r4.y, r4.z);
    L.c.mp = mk3(r5.x, r5.y, r5.z);
    const int slot = __float_as_int(r6.x);
    a.nmain = __float_as_int(r6.y) + 1;
    const int phase = __float_as_int(r6.z);
    a.nsub = __float_as_int(r6.w);
    const unsigned long long sb = (unsigned long long)__float_as_uint(r7.x) |
                                  ((unsigned long long)__float_as_uint(r7.y) << 32);
    a.log_sub = A.jobrec + sb * kRecFloats;
    a.log_main = nullptr;
    if (slot >= 0) {  // the photon's main-path log, when it was stored
        const float4 s1 = reinterpret_cast<const float4 *>(A.slots + (size_t)slot * 16)[1];
        const float4 s2 = reinterpret_cast<const float4 *>(A.slots + (size_t)slot * 16)[2];
        if (__float_as_int(s1.w)) {
            const unsigned long long base = (unsigned long long)__float_as_uint(s2.x) |
                                            ((unsigned long long)__float_as_uint(s2.y) << 32);
            a.log_main = A.slotrec + base * kRecFloats;
        }
    }
    a.gL = 0.f;
    sub_rev_start<FIELD>(L, mk3(r2.x, r2.y, r2.z), mk3(r3.x, r3.y, r3.z), r2.w, r3.w, phase);
}

// A stored main path (adj_store_slot's layout) as the lane's backward run;
// false for a photon whose log was not stored.
template <int FIELD>
__device__ __forceinline__ bool slot_load(Lane &L, unsigned long long s) {
    AdjState &a = *L.a;
    Hot &h = L.h;
    const RenderArgs &A = L.A;
    const float4 *r = reinterpret_cast<const float4 *>(A.slots + s * 16);
    const float4 r0 = r[0], r1 = r[1], r2 = r[2];
    if (!__float_as_int(r1.w)) return false;
    a.src = mk3(r0.x, r0.y, r0.z);
    h.scaling = r0.w;
    h.tau = r1.x;
    a.nmain = __float_as_int(r1.y);
    a.last_lam = __float_as_int(r1.z);
    const unsigned long long base = (unsigned long long)__float_as_uint(r2.x) |
                                    ((unsigned long long)__float_as_uint(r2.y) << 32);
    a.log_main = A.slotrec + base * kRecFloats;
    main_rev_init<FIELD>(L);
    return true;
}

// The deferred image adjoint's backward kernel.  M_ADJSUB: the queued
// next-event and direct sub-paths, each adding its vertex adjoint to the
// photon's stored main-path log; M_ADJMAIN (launched after it): the stored
// main paths.  Every lane runs backwards: items are fetched wave by wave from
// the queue's own fetch counter, so a wave's lanes step together instead of
// waiting on the few lanes of a marching wave that reverse.
// BATCH: items of many point-source patterns; a lane switches to an item's
// configuration (flushing its parameter sums to the previous one's
// gradient) as the re-trace does.
constexpr bool kRevAcc = true;  // spline backward kernels: register tap sums (SplineAcc)

template <int FIELD, int MODE, bool BATCH = false>
__global__ void __launch_bounds__(kBlock) k_adj_rev(DevScene S, RenderArgs A) {
    static_assert(MODE == M_ADJSUB || MODE == M_ADJMAIN, "backward modes only");
    __shared__ uint32_t cnt[C_N];
    __shared__ Var vars[BATCH ? kBlock : 1];
    __shared__ float sacc[FIELD == kSources ? (2 * kMaxAdjSources + 1) * kBlock : (FIELD == kBessel ? 3 * kBlock : 1)];
    __shared__ float vcell[FIELD == kSpline ? 64 * kBlock : 1];  // spline VJP cell sums
    if (FIELD == kSpline)
        for (int t = 0; t < 64; ++t) vcell[t * kBlock + threadIdx.x] = 0.f;
    if (threadIdx.x < C_N) cnt[threadIdx.x] = 0u;
    const int npar = FIELD == kSources ? 2 * S.n_src + 1 : (FIELD == kBessel ? 3 : 0);
    for (int j = 0; j < npar; ++j) sacc[j * kBlock + threadIdx.x] = 0.f;
    __syncthreads();
    const unsigned long long n_items =
        MODE == M_ADJSUB ? min(A.qcnt[Q_JOBS], A.job_cap) : min(A.qcnt[Q_SLOTS], A.slot_cap);
    unsigned long long *fetch = A.qcnt + (MODE == M_ADJSUB ? Q_WSUB : Q_WMAIN);
    Var vsingle = var_from(S, A);
    Var &v = BATCH ? vars[threadIdx.x] : vsingle;
    int cur_cfg = -1;
    Hot h;
    h.steps = 0;
    h.state = ST_NEED;
    h.dead = false;
    Cold c;  // only the sub-path start point (mp) is used
    AdjState a;
    SplineCell scell;
    scell.key = -1;
    constexpr bool ACC = FIELD == kSpline && kRevAcc;
    SplineAcc racc;
    racc.kx = kAccEmpty;
    if constexpr (ACC)
#pragma unroll
        for (int t = 0; t < 64; ++t) racc.v[t] = 0.f;
    a.F = 0.0;
    a.rsteps = a.nsubrev = a.nmainrev = a.dropped = 0;
    a.pending = a.trunc_main = a.trunc_sub = false;
    a.slot = -1;
    a.gL = 0.f;
    a.log_main = a.log_sub = nullptr;
    a.G.grid = FIELD == kSpline && A.gx ? A.gx + (size_t)xcd_id() * A.gx_stride : A.grad;
    a.G.acc = sacc + threadIdx.x;
    a.G.stride = kBlock;
    a.G.cell = FIELD == kSpline ? vcell + threadIdx.x : nullptr;
    a.G.kx = -1;
    Lane L{h, c, &v, S, A, cnt, &a, nullptr, nullptr, nullptr, true};
    const int lane = __lane_id();
    while (true) {
        const unsigned long long need = __ballot(h.state == ST_NEED);
        if (need) {
            const int leader = __ffsll((long long)need) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(fetch, (unsigned long long)__popcll(need));
            base = __shfl(base, leader);
            if (h.state == ST_NEED) {
                const unsigned long long idx = base + __popcll(need & ((1ull << lane) - 1ull));
                if (idx >= n_items) {
                    h.state = ST_DONE;
                } else {
                    if constexpr (BATCH) {
                        const int cfg = item_cfg<MODE>(A, idx);
                        if (cfg != cur_cfg) {
                            adj_flush_cfg(a, A, cur_cfg, npar);
                            cur_cfg = cfg;
                            v = A.cfgs[cfg];
                        }
                    }
                    if constexpr (MODE == M_ADJSUB)
                        job_load<FIELD>(L, idx);
                    else if (!slot_load<FIELD>(L, idx))
                        h.state = ST_NEED;
                }
            }
        }
        if (h.state == ST_REND) rev_end<FIELD, MODE>(L);
        if (__ballot(h.state != ST_DONE) == 0ull) break;
        if (__ballot(h.state == ST_RMARCH) == 0ull) continue;
        const FieldP F = v.f;
        bool r = h.state == ST_RMARCH;
        for (int k = 0; k < A.spc; ++k)
            if (r) r = rev_step<FIELD, FIELD == kSpline, ACC>(h, a, S, F, &scell, &racc);
        if (h.state == ST_RMARCH && !r) h.state = ST_REND;
    }
    if constexpr (ACC) spline_acc_commit(S, a.G, racc);
    if (FIELD == kSpline) spline_cell_flush(S, a.G);
    unsigned long long ac[A_N] = {a.rsteps, a.nsubrev, a.nmainrev, a.dropped};
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < A_N; ++k) ac[k] += __shfl_xor(ac[k], off);
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < A_N; ++k)
            if (ac[k]) atomicAdd(A.acnt + k, ac[k]);
    }
    if (BATCH) adj_flush_cfg(a, A, cur_cfg, npar);
    __syncthreads();
    for (int j = threadIdx.x; j < (BATCH ? 0 : npar); j += kBlock) {
        float sum = 0.f;
        for (int t = 0; t < kBlock; ++t) sum += sacc[j * kBlock + t];
        if (sum != 0.f) atomicAdd(A.grad + j, sum);
    }
}

}  // namespace simpler
