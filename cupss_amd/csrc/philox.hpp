// philox.hpp -- counter-based Philox4x32-10 and Box-Muller normals, generated
// in-kernel (replaces the cuRAND generator of src/field_init.cpp:130-137 and
// curandGenerateNormal at src/field.cpp:307).
//
// Philox4x32-10 is written from Salmon, Moraes, Dror and Shaw, "Parallel
// Random Numbers: As Easy as 1, 2, 3" (SC'11): multipliers 0xD2511F53 and
// 0xCD9E8D57, Weyl key increments 0x9E3779B9 and 0xBB67AE85, ten rounds.
// Deviate number `cell` of a noise stream uses counter
// (cell/4, step_lo, step_hi, stream) and key (seed_lo, seed_hi); the four
// 32-bit outputs become two Box-Muller pairs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cupss {

struct Philox4 {
    uint32_t x[4];
};

__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                 uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        // one 32x32->64 multiply gives both halves of each product
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    Philox4 o;
    o.x[0] = c0;
    o.x[1] = c1;
    o.x[2] = c2;
    o.x[3] = c3;
    return o;
}

// uniform in (0,1), exactly representable in float: (2*(x>>9)+1) * 2^-24
__device__ __forceinline__ float u01(uint32_t x) { return (float)(2u * (x >> 9) + 1u) * 5.9604644775390625e-08f; }

// (cos, sin)(2 pi w) for w in (0, 1): quadrant q = rint(4w), x = w - q/4 in
// [-1/8, 1/8] (exact in float), theta = 2 pi x in [-pi/4, pi/4], then the
// Taylor series of sin to theta^9 and of cos to theta^8 (truncation below
// 2e-9 and 3e-8 on that interval) rotated back by q quarter turns.
__device__ __forceinline__ float2 cossin_2pi(float w)
{
    const float q = rintf(4.0f * w);
    const float th = 6.2831853071795865f * (w - 0.25f * q);
    const float t2 = th * th;
    const float s = th * (1.0f + t2 * (-1.6666667e-1f + t2 * (8.3333333e-3f + t2 * (-1.9841270e-4f +
                                                                                       t2 * 2.7557319e-6f))));
    const float c = 1.0f + t2 * (-0.5f + t2 * (4.1666667e-2f + t2 * (-1.3888889e-3f + t2 * 2.4801587e-5f)));
    switch ((int)q & 3) {
    case 0:
        return make_float2(c, s);
    case 1:
        return make_float2(-s, c);
    case 2:
        return make_float2(-c, -s);
    default:
        return make_float2(s, -c);
    }
}

// Box-Muller: r = sqrt(-2 ln u_a) = sqrt(-2 ln2 * log2 u_a) with the
// hardware log2 and square root, angle 2 pi u_b
__device__ __forceinline__ float2 box_muller(uint32_t a, uint32_t b)
{
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u01(a)));
    const float2 cs = cossin_2pi(u01(b));
    return make_float2(r * cs.x, r * cs.y);
}

// Box-Muller with the hardware sine and cosine (argument in revolutions,
// absolute error of order 1e-6): the spectral-space draw, whose checks are
// statistical (its law), uses it; the real-space draw, which is compared
// value by value with the oracle's stream, uses box_muller().
__device__ __forceinline__ float2 box_muller_hw(uint32_t a, uint32_t b)
{
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u01(a)));
    const float w = u01(b);
    return make_float2(r * __builtin_amdgcn_cosf(w), r * __builtin_amdgcn_sinf(w));
}

// The normals of cells `cell` and `cell+1` (cell even) of one stream.
__device__ __forceinline__ float2 philox_normal_pair(uint32_t cell, uint32_t step_lo, uint32_t step_hi,
                                                     uint32_t stream, uint32_t key0, uint32_t key1)
{
    const Philox4 o = philox4x32_10(cell >> 2, step_lo, step_hi, stream, key0, key1);
    return (cell & 2) ? box_muller(o.x[2], o.x[3]) : box_muller(o.x[0], o.x[1]);
}

}  // namespace cupss
