// device_math.h -- small vector type, Philox-4x32-10 and uniforms for the
// gfx950 engine.  Everything here is wave64-friendly straight-line code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace simpler {

constexpr float kEps = 1.19209290e-07f;  // M_EPSILON (reference constants.h:38)
constexpr float kPi = 3.14159265358979323846f;
constexpr float kLargeDist = 10000.0f;   // scene.cpp:627

struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len(f3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalized(f3 a) {
    const float l = len(a);
    return f3{a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// Philox-4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
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

// (2k+1) * 2^-24 with k the top 23 bits: exact in FP32, strictly in (0,1).
__device__ __forceinline__ float to_uniform(uint32_t w) {
    return (float)(((w >> 9) << 1) | 1u) * (1.0f / 16777216.0f);
}

// x / 10 correctly rounded without the IEEE division sequence: q0 = x *
// RN(0.1), the exact residual x - 10 q0 by one fma, one correction step.
// Equal to the FP32 quotient x / 10.0f for every |x| >= 1e-30 (checked over
// all such floats); below that, in the subnormal range, it can differ by an
// ulp.
__device__ __forceinline__ float div10(float x) {
    const float q0 = x * 0.1f;
    return __builtin_fmaf(__builtin_fmaf(-q0, 10.0f, x), 0.1f, q0);
}

// Two-sum accumulation hi + lo += d (Kahan's compensated step).
__device__ __forceinline__ void comp_add(float &hi, float &lo, float d) {
    const float y = d + lo;
    const float s = hi + y;
    lo = y - (s - hi);
    hi = s;
}

}  // namespace simpler
