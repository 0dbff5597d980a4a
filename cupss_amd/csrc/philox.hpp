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
        const uint32_t lo0 = 0xD2511F53u * c0;
        const uint32_t hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2;
        const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
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

__device__ __forceinline__ float2 box_muller(uint32_t a, uint32_t b)
{
    const float r = sqrtf(-2.0f * logf(u01(a)));
    float s, c;
    sincospif(2.0f * u01(b), &s, &c);
    return make_float2(r * c, r * s);
}

// The normals of cells `cell` and `cell+1` (cell even) of one stream.
__device__ __forceinline__ float2 philox_normal_pair(uint32_t cell, uint32_t step_lo, uint32_t step_hi,
                                                     uint32_t stream, uint32_t key0, uint32_t key1)
{
    const Philox4 o = philox4x32_10(cell >> 2, step_lo, step_hi, stream, key0, key1);
    return (cell & 2) ? box_muller(o.x[2], o.x[3]) : box_muller(o.x[0], o.x[1]);
}

}  // namespace cupss
