// update_math.hpp -- per-bin spectral arithmetic shared by the update
// kernel and the fused step kernels.
//
//   q grid (src/field_init.cpp:207-213), term multipliers (S2,
//   src/term_init.cpp:135-200), the projection onto real fields of the
//   reference's comp->real->comp round trip (src/field.cpp:63-89, see
//   multiplier()), noise amplitude (src/field_init.cpp:268-278) and the
//   semi-implicit update of setDynamic_k / setNotDynamic_k
//   (src/field_kernels.cu:130-227).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_core.hpp"
#include "internal.hpp"
#include "philox.hpp"

namespace cupss {

struct Bin {
    long h;          // flat index in the half spectrum
    int kx, ky, kz;
    float qx, qy, qz, q2;
    float iq;        // 1/|q|, 0 at q = 0 (terms, noise, constraint division)
    float iqd;       // 1/|q|, 0 on the i = j = 0 line (dynamic implicit, Q8)
    bool nyx, nyy, nyz;
};

__device__ __forceinline__ float ipow(float x, int n)
{
    float r = 1.0f;
    for (int i = 0; i < n; i++) r *= x;
    return r;
}

__device__ __forceinline__ Bin make_bin(const Geom &g, int kx, int ky, int kz)
{
    Bin b;
    b.kx = kx;
    b.ky = ky;
    b.kz = kz;
    b.h = ((long)kz * g.sy + ky) * g.nc1 + kx;
    b.qx = (float)(kx < (g.sx + 1) / 2 ? kx : kx - g.sx) * g.sqx;
    b.qy = (float)(ky < (g.sy + 1) / 2 ? ky : ky - g.sy) * g.sqy;
    b.qz = (float)(kz < (g.sz + 1) / 2 ? kz : kz - g.sz) * g.sqz;
    b.q2 = b.qx * b.qx + b.qy * b.qy + b.qz * b.qz;
    b.nyx = g.sx > 1 && kx == g.sx / 2;
    b.nyy = g.sy > 1 && ky == g.sy / 2;
    b.nyz = g.sz > 1 && kz == g.sz / 2;
    // hardware reciprocal square root (q2 = 0 gives +inf, masked below)
    const float r = __builtin_amdgcn_rsqf(b.q2);
    b.iq = b.h > 0 ? r : 0.0f;
    b.iqd = (kx > 0 || ky > 0) ? r : 0.0f;
    return b;
}

// M(q) of a term, with the real-part projection folded in: a prefactor
// monomial i^m c q^2n qx^a qy^b qz^c |q|^-k survives the reference's
// IFFT/.y=0/FFT round trip iff m + (sum of its q_d powers over the axes
// where this bin is not a Nyquist index) is even; otherwise its image
// cancels against its Hermitian partner and it contributes nothing.  The
// parity test is a per-bin select (no divergent branch); the exponent
// loops have uniform trip counts.
__device__ __forceinline__ float multiplier(const PresDev *__restrict__ pres, int p0, int np, int mult_i,
                                            const Bin &b)
{
    float m = 0.0f;
    for (int p = p0; p < p0 + np; p++) {
        const PresDev P = pres[p];
        const int par = mult_i + (b.nyx ? 0 : P.a) + (b.nyy ? 0 : P.b) + (b.nyz ? 0 : P.cz);
        float v = (par & 1) ? 0.0f : P.c;
        if (P.q2n > 0) v *= ipow(b.q2, P.q2n);
        if (P.a > 0) v *= ipow(b.qx, P.a);
        if (P.b > 0) v *= ipow(b.qy, P.b);
        if (P.cz > 0) v *= ipow(b.qz, P.cz);
        if (P.invq > 0) v *= ipow(b.iq, P.invq);
        m += v;
    }
    return m;
}

// i^mult_i * m * X
__device__ __forceinline__ float2 apply_mult(float m, int mult_i, float2 X)
{
    const float2 v = make_float2(X.x * m, X.y * m);
    return mult_i ? make_float2(-v.y, v.x) : v;
}

__device__ __forceinline__ float noise_amplitude(const UpdateArgs &a, const Bin &b)
{
    float pn = a.noise_base;
    if (a.noise_q2n != 0) pn *= sqrtf(ipow(b.q2, a.noise_q2n));
    if (a.noise_invq != 0) pn *= sqrtf(ipow(b.iq, a.noise_invq));
    return pn;
}

// The value of term t at this bin, for TERM_FIELD / TERM_CONST terms and
// for TERM_BUFFER terms whose transform is read from memory.
__device__ __forceinline__ float2 term_value(const UpdateArgs &a, int t, const Bin &b, const Geom &g)
{
    const TermRefDev &T = a.terms[t];
    float2 X;
    if (T.kind == TERM_CONST)
        X = b.h == 0 ? make_float2((float)g.n, 0.f) : make_float2(0.f, 0.f);
    else
        X = T.X[b.h];
    return apply_mult(multiplier(a.pres, T.pres0, T.npres, T.mult_i, b), T.mult_i, X);
}

// New spectrum value of the field from its current value, the sum of its
// explicit terms and its noise transform.
__device__ __forceinline__ float2 finish_update(const UpdateArgs &a, const Bin &b, float2 cur, float2 acc,
                                                float2 eta)
{
    float2 s;
    if (a.dynamic) {
        s = cur;
        s.x += a.dt * acc.x;
        s.y += a.dt * acc.y;
        if (a.noisy) {
            const float pn = noise_amplitude(a, b);
            s.x += pn * eta.x;
            s.y += pn * eta.y;
        }
        if (a.nimp > 0) {
            float D = 1.0f;
            for (int i = 0; i < a.nimp; i++) {
                const PresDev P = a.imp[i];
                float th = P.c;
                if (P.q2n != 0) th *= ipow(b.q2, P.q2n);
                if (P.invq != 0) th *= ipow(b.iqd, P.invq);
                D -= a.dt * th;
            }
            const float rD = __builtin_amdgcn_rcpf(D);  // hardware reciprocal, once per bin
            s.x *= rD;
            s.y *= rD;
        }
    } else {
        s = a.nterms > 0 ? acc : cur;
        if (a.noisy) {
            const float f = a.sdt * noise_amplitude(a, b);
            if (a.nterms == 0) {
                s.x = f * eta.x;
                s.y = f * eta.y;
            } else {
                s.x += f * eta.x;
                s.y += f * eta.y;
            }
        }
        if (a.nimp > 0 && b.h > 0) {
            float L = 0.0f;
            for (int i = 0; i < a.nimp; i++) {
                const PresDev P = a.imp[i];
                float th = P.c;
                if (P.q2n != 0) th *= ipow(b.q2, P.q2n);
                if (P.invq != 0) th *= ipow(b.iq, P.invq);
                L += th;
            }
            const float rL = __builtin_amdgcn_rcpf(L);
            s.x *= rL;
            s.y *= rL;
        }
    }
    return s;
}

// ---- the same arithmetic for the E bins (kx, t + e*T), e < E, that one
// lane of a column kernel holds (2D instances).  Loops over prefactors run
// once per lane with all E bins inside them, so every prefactor is loaded
// once and the loop overhead is shared.

template <int E>
struct ColBins {
    int kx;
    float qx;
    bool nyx;
    float qy[E], q2[E];
    bool nyy[E], zero[E];  // Nyquist row; the q = 0 bin
};

template <int E, int T>
__device__ __forceinline__ void col_bins(const Geom &g, int kx, int t, ColBins<E> &cb)
{
    cb.kx = kx;
    cb.qx = (float)(kx < (g.sx + 1) / 2 ? kx : kx - g.sx) * g.sqx;
    cb.nyx = g.sx > 1 && kx == g.sx / 2;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int ky = t + e * T;
        cb.qy[e] = (float)(ky < (g.sy + 1) / 2 ? ky : ky - g.sy) * g.sqy;
        cb.q2[e] = cb.qx * cb.qx + cb.qy[e] * cb.qy[e];
        cb.nyy[e] = g.sy > 1 && ky == g.sy / 2;
        cb.zero[e] = kx == 0 && ky == 0;
    }
}

// 1/|q| of bin e (0 at q = 0; in 2D the i = j = 0 line is that bin alone)
template <int E>
__device__ __forceinline__ float col_iq(const ColBins<E> &cb, int e)
{
    return cb.zero[e] ? 0.0f : __builtin_amdgcn_rsqf(cb.q2[e]);
}

// multiplier() for the E bins of a column
template <int E>
__device__ __forceinline__ void multiplier_col(const PresDev *__restrict__ pres, int p0, int np, int mult_i,
                                               const ColBins<E> &cb, float (&m)[E])
{
#pragma unroll
    for (int e = 0; e < E; e++) m[e] = 0.0f;
    for (int p = p0; p < p0 + np; p++) {
        const PresDev P = pres[p];
        if (P.cz > 0) continue;  // qz = 0 in 2D
        float base = P.c;
        for (int i = 0; i < P.a; i++) base *= cb.qx;
        const int parx = mult_i + (cb.nyx ? 0 : P.a);
        float v[E];
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = base;
        for (int i = 0; i < P.q2n; i++)
#pragma unroll
            for (int e = 0; e < E; e++) v[e] *= cb.q2[e];
        for (int i = 0; i < P.b; i++)
#pragma unroll
            for (int e = 0; e < E; e++) v[e] *= cb.qy[e];
        if (P.invq > 0) {
#pragma unroll
            for (int e = 0; e < E; e++) {
                const float iq = col_iq(cb, e);
                for (int i = 0; i < P.invq; i++) v[e] *= iq;
            }
        }
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int par = parx + (cb.nyy[e] ? 0 : P.b);
            m[e] += (par & 1) ? 0.0f : v[e];
        }
    }
}

// D(q) = 1 - dt * sum_imp (setDynamic_k) and the noise amplitude, per bin
template <int E>
__device__ __forceinline__ void implicit_col(const UpdateArgs &a, const ColBins<E> &cb, float (&rD)[E])
{
    float D[E];
#pragma unroll
    for (int e = 0; e < E; e++) D[e] = 1.0f;
    for (int i = 0; i < a.nimp; i++) {
        const PresDev P = a.imp[i];
        float th[E];
#pragma unroll
        for (int e = 0; e < E; e++) th[e] = P.c;
        for (int k = 0; k < P.q2n; k++)
#pragma unroll
            for (int e = 0; e < E; e++) th[e] *= cb.q2[e];
        if (P.invq != 0) {
#pragma unroll
            for (int e = 0; e < E; e++) {
                const float iq = col_iq(cb, e);
                for (int k = 0; k < P.invq; k++) th[e] *= iq;
            }
        }
#pragma unroll
        for (int e = 0; e < E; e++) D[e] -= a.dt * th[e];
    }
#pragma unroll
    for (int e = 0; e < E; e++) rD[e] = __builtin_amdgcn_rcpf(D[e]);
}

template <int E>
__device__ __forceinline__ float noise_amplitude_col(const UpdateArgs &a, const ColBins<E> &cb, int e)
{
    float pn = a.noise_base;
    if (a.noise_q2n != 0) pn *= sqrtf(ipow(cb.q2[e], a.noise_q2n));
    if (a.noise_invq != 0) pn *= sqrtf(ipow(col_iq(cb, e), a.noise_invq));
    return pn;
}

// dealias_k (src/field_kernels.cu:229-256): keep |n_d| <= s_d/(p+1) on every axis
// (limits precomputed: lx = sx/(p+1), ly = sy/(p+1), 2D)
__device__ __forceinline__ bool keep_mode2(const Geom &g, int lx, int ly, int kx, int ky)
{
    const int ny = ky > g.sy / 2 ? ky - g.sy : ky;
    return kx <= lx && abs(ny) <= ly;
}

__device__ __forceinline__ bool keep_mode(const Geom &g, int p, int kx, int ky, int kz)
{
    if (p <= 0) return true;
    const int p1 = p + 1;
    const int ny = ky > g.sy / 2 ? ky - g.sy : ky;
    const int nz = kz > g.sz / 2 ? kz - g.sz : kz;
    return kx <= g.sx / p1 && abs(ny) <= g.sy / p1 && abs(nz) <= g.sz / p1;
}

// The law of R2C(xi) for i.i.d. xi ~ N(0,1) on N cells: at each bin a
// complex Gaussian with E|eta|^2 = N (re and im independent, variance N/2),
// independent between conjugate pairs, real with variance N at
// self-conjugate bins.  Bins with kx = 0 or nc have their partner
// (kx, -ky, -kz) inside the half spectrum: the partner with the smaller
// (kz, ky) draws, the other takes the conjugate.
//
// Draws are paired along y: bins (kx, j, kz) and (kx, j + hy, kz), with
// hy = sy/2 (1 if sy = 1) and j < hy, share one Philox4x32-10 call with
// counter (pair index, step_lo, step_hi, stream | 0x80000000); outputs 0,1
// give the first bin's Box-Muller pair and outputs 2,3 the second's
// (box_muller_hw).
__device__ __forceinline__ int noise_hy(const Geom &g) { return g.sy >= 2 ? g.sy / 2 : 1; }

__device__ __forceinline__ Philox4 noise_philox(const Geom &g, int kx, int j, int kz, uint32_t slo, uint32_t shi,
                                                uint32_t stream, uint32_t k0, uint32_t k1)
{
    const uint32_t pair = (uint32_t)(((long)kz * g.nc1 + kx) * noise_hy(g) + j);
    return philox4x32_10(pair, slo, shi, stream | 0x80000000u, k0, k1);
}

__device__ __forceinline__ float2 spectral_noise(const Geom &g, int kx, int ky, int kz, uint32_t slo, uint32_t shi,
                                                 uint32_t stream, uint32_t k0, uint32_t k1)
{
    bool conj = false, real = false;
    int cy = ky, cz = kz;
    if (kx == 0 || kx == g.nc) {
        const int py = (g.sy - ky) % g.sy, pz = (g.sz - kz) % g.sz;
        if (py == ky && pz == kz) {
            real = true;
        } else if (pz < kz || (pz == kz && py < ky)) {
            cy = py;
            cz = pz;
            conj = true;
        }
    }
    const int hy = noise_hy(g);
    const Philox4 o = noise_philox(g, kx, cy % hy, cz, slo, shi, stream, k0, k1);
    const float2 z = (cy / hy) ? box_muller_hw(o.x[2], o.x[3]) : box_muller_hw(o.x[0], o.x[1]);
    const float n = (float)g.n;
    if (real) return make_float2(__builtin_amdgcn_sqrtf(n) * z.x, 0.0f);
    const float s = __builtin_amdgcn_sqrtf(0.5f * n);
    return make_float2(s * z.x, conj ? -s * z.y : s * z.y);
}

// Both bins of the pair (kx, j, kz), (kx, j + hy, kz) from one Philox call
// (sy >= 2); equal to spectral_noise() of each bin.
__device__ __forceinline__ void spectral_noise_pair(const Geom &g, int kx, int j, int kz, uint32_t slo, uint32_t shi,
                                                    uint32_t stream, uint32_t k0, uint32_t k1, float2 &lo,
                                                    float2 &hi)
{
    if (kx == 0 || kx == g.nc) {
        lo = spectral_noise(g, kx, j, kz, slo, shi, stream, k0, k1);
        hi = spectral_noise(g, kx, j + noise_hy(g), kz, slo, shi, stream, k0, k1);
        return;
    }
    const Philox4 o = noise_philox(g, kx, j, kz, slo, shi, stream, k0, k1);
    const float2 z0 = box_muller_hw(o.x[0], o.x[1]);
    const float2 z1 = box_muller_hw(o.x[2], o.x[3]);
    const float s = __builtin_amdgcn_sqrtf(0.5f * (float)g.n);
    lo = make_float2(s * z0.x, s * z0.y);
    hi = make_float2(s * z1.x, s * z1.y);
}

}  // namespace cupss
