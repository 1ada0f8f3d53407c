// prefilter.h -- cubic B-spline interpolation prefilter of a 3-D grid on the
// GPU (included by render.hip).
//
// Reference: spline.h:686-764 (build1d / build3d; readVolAndBuild at
// 115-162 calls it once per VOL grid).  Per line of n samples the mirror-
// boundary recursive filter with pole z1 = sqrt(3) - 2 and gain 6:
//   causal:      c+[0] = sum_m s(m) z1^m / (1 - z1^(2n-2)) over the mirrored
//                sequence s of period 2n-2, c+[i] = x[i] + z1 c+[i-1]
//   anticausal:  c-[n-1] = z1 / (z1^2 - 1) (c+[n-1] + z1 c+[n-2]),
//                c-[i] = z1 (c-[i+1] - c+[i]),  coefficient = 6 c-[i]
// applied along y, then x, then z (the reference's order).
//
// One lane per line, in place, one read and one write per sample.  Along y
// and z consecutive lanes take consecutive x, so every load and store of a
// wave is one contiguous run; along x each lane walks its own row in 128-byte
// runs.  The anticausal recursion forgets where it started as |z1|^k, so a
// lane streams its line in windows of B = 32 outputs: it holds the causal
// values of the window plus D = 16 samples of lookahead in registers, runs the
// anticausal recursion backwards from zero at the end of the lookahead
// (|z1|^17 < 2e-10 relative) and writes the B outputs; only the last window
// starts from the exact end condition.  The causal initial value keeps the
// first 24 terms of the mirrored sum: |z1|^24 < 2e-14, far below FP32
// resolution.  Both truncations are below FP32 rounding, so the coefficients
// equal the exact recursion's to FP32 accuracy.
//
// The first pass subtracts `shift` from every sample (the filter maps
// constants to themselves, so this is the coefficient grid minus shift, with
// FP32 resolution spent on the variation instead of the offset) and flags
// non-finite samples; the last pass reduces the minimum coefficient.
#pragma once

namespace simpler {

constexpr int kPrefBlock = 256;
constexpr int kPrefInitTerms = 24;
// window per axis: B outputs, D samples of lookahead (along x, D = 32 keeps
// every 32-float read of a row on one 128-byte line)
constexpr int kPrefB = 32, kPrefDyz = 16, kPrefDx = 32;

// orderable unsigned key of a float (monotone), for atomicMin
__device__ __forceinline__ unsigned float_key(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

template <int AXIS, int B, int D>
__global__ void __launch_bounds__(kPrefBlock)
    k_prefilter(float *__restrict__ c, int N0, int N1, int N2, float shift, unsigned *__restrict__ flags,
                unsigned *__restrict__ minkey) {
    const size_t NXY = (size_t)N0 * N1;
    const int n = AXIS == 0 ? N0 : (AXIS == 1 ? N1 : N2);
    const long lines = AXIS == 0 ? (long)N1 * N2 : (AXIS == 1 ? (long)N0 * N2 : (long)N0 * N1);
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float lmin = INFINITY;
    float chk = 0.f;  // x - x summed: stays 0 unless a sample is inf or NaN
    constexpr bool first = AXIS == 1;
    constexpr int W = B + D;  // causal values held per lane
    if (l < lines) {
        size_t base, stride;
        if (AXIS == 0) {  // row l: (y, z) = (l % N1, l / N1)
            base = (size_t)l * N0;
            stride = 1;
        } else if (AXIS == 1) {  // (x, z) = (l % N0, l / N0)
            base = (size_t)(l / N0) * NXY + (size_t)(l % N0);
            stride = (size_t)N0;
        } else {  // (x, y) = l
            base = (size_t)l;
            stride = NXY;
        }
        float *p = c + base;
        auto load = [&](int i) {
            float x = p[(size_t)i * stride];
            if (first) {
                chk += x - x;
                x -= shift;
            }
            return x;
        };
        if (n == 1) {
            const float x = load(0);
            if (first) p[0] = x;
            lmin = x;
        } else {
            const float z1 = -0.267949192431122706f;  // sqrt(3) - 2
            // causal initial value over the mirrored sequence (period 2n-2)
            // Every load below has an in-bounds index (clamped or mirrored)
            // and no branch around it, so a lane's loads issue back to back.
            const int period = 2 * n - 2;
            float xi[kPrefInitTerms];
#pragma unroll
            for (int m = 0; m < kPrefInitTerms; ++m) {
                const int mm = m < period ? m : 0;
                xi[m] = p[(size_t)(mm < n ? mm : period - mm) * stride];
            }
            float c0 = 0.f, zk = 1.f;
#pragma unroll
            for (int m = 0; m < kPrefInitTerms; ++m) {
                const float x = first ? xi[m] - shift : xi[m];
                c0 = fmaf(m < period ? x : 0.f, zk, c0);
                zk *= z1;
            }
            // 1 - z1^(2n-2): the power underflows to 0 long before m = 64
            float zp = 1.f;
            for (int m = 0; m < period && m < 64; ++m) zp *= z1;
            // w holds c+[a .. a+W) for the window of outputs [a, a+B)
            float w[W];
            w[0] = c0 / (1.f - zp);
            (void)load(0);  // flags a non-finite first sample
            float cprev = w[0];
            {
                float x[W];
#pragma unroll
                for (int k = 1; k < W; ++k) x[k] = load(k < n ? k : n - 1);
#pragma unroll
                for (int k = 1; k < W; ++k) {
                    cprev = fmaf(z1, cprev, x[k]);  // past n-1: unused values
                    w[k] = cprev;
                }
            }
            int a = 0;
            while (n - a > W) {
                // outputs [a, a+B): anticausal recursion started D samples
                // ahead at zero (its start is forgotten as |z1|^(D+1) < 2e-10)
                float cn = 0.f;
                float o[B];
#pragma unroll
                for (int k = W - 1; k >= 0; --k) {
                    cn = z1 * (cn - w[k]);
                    if (k < B) o[k] = 6.f * cn;
                }
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    p[(size_t)(a + k) * stride] = o[k];
                    lmin = fminf(lmin, o[k]);
                }
#pragma unroll
                for (int k = 0; k < D; ++k) w[k] = w[B + k];
                // the next B causal values (all loads issued before use)
                float x[B];
#pragma unroll
                for (int k = 0; k < B; ++k) x[k] = load(a + W + k < n ? a + W + k : n - 1);
#pragma unroll
                for (int k = 0; k < B; ++k) {
                    cprev = fmaf(z1, cprev, x[k]);
                    w[D + k] = cprev;
                }
                a += B;
            }
            // last window: v = n - a values (2 <= v <= W), exact end condition
            const int v = n - a;
            float last = 0.f, prev = 0.f;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                if (k == v - 1) last = w[k];
                if (k == v - 2) prev = w[k];
            }
            float cn = z1 / (z1 * z1 - 1.f) * fmaf(z1, prev, last);
#pragma unroll
            for (int k = W - 1; k >= 0; --k) {
                if (k < v - 1) cn = z1 * (cn - w[k]);
                if (k < v) {
                    const float out = 6.f * cn;
                    p[(size_t)(a + k) * stride] = out;
                    lmin = fminf(lmin, out);
                }
            }
        }
    }
    if (AXIS == 1 && flags) {
        if (__ballot(!(chk == 0.f)) && __lane_id() == 0) atomicOr(flags, 1u);
    }
    if (AXIS == 2 && minkey) {
        for (int off = 32; off > 0; off >>= 1) lmin = fminf(lmin, __shfl_xor(lmin, off));
        if (__lane_id() == 0 && lmin < INFINITY) atomicMin(minkey, float_key(lmin));
    }
}

// The x pass with LDS staging: a one-wave workgroup owns 64 rows (one per
// lane) and a 64 x B tile of LDS.  Loads and stores move through the tile so
// that every global instruction covers 8 rows x 128 contiguous bytes (whole
// cache lines); each lane runs its row's recursions on the tile's row, with
// the same windows as k_prefilter.  With one wave per workgroup the barriers
// order the tile's writes and reads without making waves wait on each other.
constexpr int kPrefRowBlock = 64;

template <int B, int D>
__global__ void __launch_bounds__(kPrefRowBlock)
    k_prefilter_rows(float *__restrict__ c, int n, long rows) {
    constexpr int W = B + D;
    constexpr int H = (W + B - 1) / B;  // tiles per window
    constexpr int Q = B / 4;            // float4 per row and tile
    static_assert(B % 4 == 0 && D <= B, "window layout");
    __shared__ float t[64][B + 1];
    const int lane = threadIdx.x;
    const long row0 = (long)blockIdx.x * 64;
    const bool vec = (n % 4) == 0;  // every row starts 16-byte aligned
    // tile <- columns [c0, c0 + B) of the wave's rows (columns past n clamp to n - 1)
    auto stage_in = [&](int c0) {
#pragma unroll
        for (int it = 0; it < Q; ++it) {
            const int f = it * 64 + lane, r = f / Q, q = f % Q, col = c0 + 4 * q;
            const long rr = row0 + r < rows ? row0 + r : rows - 1;
            const float *rp = c + (size_t)rr * n;
            if (vec && col + 3 < n) {
                const float4 v = *reinterpret_cast<const float4 *>(rp + col);
                t[r][4 * q] = v.x;
                t[r][4 * q + 1] = v.y;
                t[r][4 * q + 2] = v.z;
                t[r][4 * q + 3] = v.w;
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) t[r][4 * q + e] = rp[col + e < n ? col + e : n - 1];
            }
        }
    };
    // columns [c0, c0 + B) of the wave's rows <- tile (only columns < n, rows < rows)
    auto stage_out = [&](int c0) {
#pragma unroll
        for (int it = 0; it < Q; ++it) {
            const int f = it * 64 + lane, r = f / Q, q = f % Q, col = c0 + 4 * q;
            if (row0 + r >= rows) continue;
            float *rp = c + (size_t)(row0 + r) * n;
            if (vec && col + 3 < n) {
                *reinterpret_cast<float4 *>(rp + col) =
                    make_float4(t[r][4 * q], t[r][4 * q + 1], t[r][4 * q + 2], t[r][4 * q + 3]);
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (col + e < n) rp[col + e] = t[r][4 * q + e];
            }
        }
    };
    const float z1 = -0.267949192431122706f;  // sqrt(3) - 2
    if (n == 1) return;                       // a one-sample row is its own coefficient
    float w[W];
    float cprev = 0.f;
    // first window: columns [0, W) in H tiles; the causal start from the
    // first tile (every index of the truncated mirrored sum is below 24)
    const int period = 2 * n - 2;
#pragma unroll
    for (int h = 0; h < H; ++h) {
        __syncthreads();
        stage_in(h * B);
        __syncthreads();
        if (h == 0) {
            float c0 = 0.f, zk = 1.f;
#pragma unroll
            for (int m = 0; m < kPrefInitTerms; ++m) {
                const int mm = m < period ? m : 0;
                const float x = t[lane][mm < n ? mm : period - mm];
                c0 = fmaf(m < period ? x : 0.f, zk, c0);
                zk *= z1;
            }
            float zp = 1.f;
            for (int m = 0; m < period && m < 64; ++m) zp *= z1;
            cprev = c0 / (1.f - zp);
            w[0] = cprev;
        }
#pragma unroll
        for (int k = (h == 0 ? 1 : 0); k < B; ++k) {
            if (h * B + k < W) {
                cprev = fmaf(z1, cprev, t[lane][k]);  // past n - 1: unused values
                w[h * B + k] = cprev;
            }
        }
    }
    int a = 0;
    while (n - a > W) {
        float cn = 0.f;
        __syncthreads();  // the tile's last reads are done: outputs go straight into it
#pragma unroll
        for (int k = W - 1; k >= 0; --k) {
            cn = z1 * (cn - w[k]);
            if (k < B) t[lane][k] = 6.f * cn;
        }
        __syncthreads();
        stage_out(a);
#pragma unroll
        for (int k = 0; k < D; ++k) w[k] = w[B + k];
        __syncthreads();
        stage_in(a + W);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < B; ++k) {
            cprev = fmaf(z1, cprev, t[lane][k]);
            w[D + k] = cprev;
        }
        a += B;
    }
    // last window: v = n - a values (2 <= v <= W), exact end condition
    const int v = n - a;
    float last = 0.f, prev = 0.f;
#pragma unroll
    for (int k = 0; k < W; ++k) {
        if (k == v - 1) last = w[k];
        if (k == v - 2) prev = w[k];
    }
    float cn = z1 / (z1 * z1 - 1.f) * fmaf(z1, prev, last);
    // backwards, one tile of B outputs at a time (the highest columns first)
#pragma unroll
    for (int h = H - 1; h >= 0; --h) {
        __syncthreads();
#pragma unroll
        for (int k = B - 1; k >= 0; --k) {
            if (h * B + k < W && h * B + k < v - 1) cn = z1 * (cn - w[h * B + k]);
            t[lane][k] = 6.f * cn;  // columns at or past n are not stored
        }
        __syncthreads();
        stage_out(a + h * B);  // stores only columns < n
    }
}

__global__ void k_prefilter_init(unsigned *w) {
    if (threadIdx.x == 0) {
        w[0] = 0u;           // non-finite flag
        w[1] = 0xFFFFFFFFu;  // min key
    }
}

// Prefilter the grid c (N0 x N1 x N2, x fastest) in place on stream s:
// y pass (minus shift, finiteness flag into w[0]), x pass, z pass (minimum
// coefficient key into w[1]).  w: two unsigned on the device.
inline hipError_t launch_prefilter3d(float *c, const int N[3], float shift, unsigned *w, hipStream_t s) {
    hipLaunchKernelGGL(k_prefilter_init, dim3(1), dim3(64), 0, s, w);
    const long ly = (long)N[0] * N[2], lx = (long)N[1] * N[2], lz = (long)N[0] * N[1];
    hipLaunchKernelGGL((k_prefilter<1, kPrefB, kPrefDyz>), dim3((unsigned)((ly + kPrefBlock - 1) / kPrefBlock)), dim3(kPrefBlock), 0, s,
                       c, N[0], N[1], N[2], shift, w, nullptr);
    hipLaunchKernelGGL((k_prefilter_rows<kPrefB, kPrefDx>), dim3((unsigned)((lx + kPrefRowBlock - 1) / kPrefRowBlock)),
                       dim3(kPrefRowBlock), 0, s, c, N[0], lx);
    hipLaunchKernelGGL((k_prefilter<2, kPrefB, kPrefDyz>), dim3((unsigned)((lz + kPrefBlock - 1) / kPrefBlock)), dim3(kPrefBlock), 0, s,
                       c, N[0], N[1], N[2], 0.f, nullptr, w + 1);
    return hipGetLastError();
}

inline float key_to_float(unsigned k) {
    const unsigned u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    std::memcpy(&f, &u, sizeof(f));
    return f;
}

}  // namespace simpler
