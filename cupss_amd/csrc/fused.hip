// fused.hip -- the fused two-kernel step for 2D systems with one dynamic
// field phi whose other fields are "derived": constraint fields
// psi = M_psi(q) phi with no implicit operator and no noise (e.g. iqxh =
// iqx*h).  Covers Cahn-Hilliard, Edwards-Wilkinson and KPZ.
//
// Pipeline state between the kernels (natural half-spectrum layout):
//   Z_f  = IFFT_y(mask_f * L_f * phi_hat)   for each real field f needed by a
//          product (f = phi, L = 1; f = psi, L = M_psi)
//   X_g  = R2C_x(sum of the products of group g)  (rows transformed only)
//   eta  = R2C_x(xi)                        (Philox noise, rows only)
// Per step:
//   row kernel:    Z_f rows -C2R-> real dealiased fields -> products,
//                  noise -R2C-> X_g, eta rows
//   column kernel: X_g, eta columns -FFT_y-> spectral terms; update of
//                  phi (setDynamic_k); mask_f * L_f * phi_new -IFFT_y-> Z_f
// Terms with equal prefactor lists form one group: M*FFT(a) + M*FFT(b) =
// M*FFT(a + b).  The order of phases A-D of the reference
// (src/evolver.cpp:199-226) is kept: derived fields are evaluated from phi
// at the start of the step (phase A/B) and used by phi's products in phase
// C; phi's own dealiased copy is zero before its first update (step 0).
#include <hip/hip_runtime.h>

#include "fft_core.hpp"
#include "fused.hpp"
#include "philox.hpp"
#include "update_math.hpp"

namespace cupss {

// ------------------------------------------------------------ row kernel

template <int NC>
struct FRowCfg {
    static constexpr int E = row_elems(NC);
    static constexpr int T = NC / E;
    static constexpr int LPB = T >= 64 ? 1 : 64 / T;
    static constexpr int THREADS = LPB * T;
    static constexpr int LL = lds_len(NC);
};

// C2R of one row of a half spectrum: v[e] = (x[2m], x[2m+1]) * scale, m = t+e*T
template <int NC, int E, int T>
__device__ __forceinline__ void row_c2r(const float2 *__restrict__ x, bool active, float scale, float2 (&v)[E],
                                        float2 *L, int t, const float2 *__restrict__ tw)
{
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int k = t + e * T;
        float2 X = make_float2(0.f, 0.f), Y = make_float2(0.f, 0.f);
        if (active) {
            X = x[k];
            Y = x[NC - k];
            if (k == 0) {
                X.y = 0.f;
                Y.y = 0.f;
            }
        }
        const float2 Yc = cconj(Y);
        const float2 xe = cadd(X, Yc);
        const float2 xo = cmul(csub(X, Yc), tw_lookup<+1>(tw, k, 2 * NC));
        v[e] = cadd(xe, mul_i<+1>(xo));
    }
    fft_line<NC, E, +1>(v, L, t, tw);
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = cscale(v[e], scale);
}

// R2C of one row whose real values are packed as v[e] = (x[2m], x[2m+1])
template <int NC, int E, int T>
__device__ __forceinline__ void row_r2c_store(float2 (&v)[E], bool active, float2 *__restrict__ o, float2 *L, int t,
                                              const float2 *__restrict__ tw)
{
    fft_line<NC, E, -1>(v, L, t, tw);
    if (active) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = t + e * T;
            const float2 A = v[e];
            const float2 B = cconj(L[lpad((NC - k) & (NC - 1))]);
            const float2 xe = cscale(cadd(A, B), 0.5f);
            const float2 xo = cscale(mul_i<-1>(csub(A, B)), 0.5f);
            o[k] = k == 0 ? make_float2(xe.x + xo.x, 0.0f) : cadd(xe, cmul(tw_lookup<-1>(tw, k, 2 * NC), xo));
        }
        if (t == 0) {
            const float2 z0 = L[lpad(0)];
            o[NC] = make_float2(z0.x - z0.y, 0.0f);
        }
    }
    __syncthreads();  // the LDS line is reused by the next transform
}

template <int NC>
__global__ __launch_bounds__(256) void k_fused_rows(FusedRowArgs a, const float2 *__restrict__ tw)
{
    using C = FRowCfg<NC>;
    constexpr int E = C::E, T = C::T;
    __shared__ float2 lds[C::LPB * C::LL];
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    const long row = (long)blockIdx.x * C::LPB + l;
    const bool active = row < a.nrows;
    float2 *L = lds + l * C::LL;
    const long hrow = row * (NC + 1);

    float2 r[FUSED_MAX_REAL][E];
#pragma unroll
    for (int f = 0; f < FUSED_MAX_REAL; f++) {
        if (f < a.nreal) {
            row_c2r<NC, E, T>(a.Z[f] + hrow, active, a.scale, r[f], L, t, tw);
            __syncthreads();
        }
    }
    for (int gi = 0; gi < a.ngroups; gi++) {
        float2 v[E];
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = make_float2(0.f, 0.f);
        for (int pi = a.g0[gi]; pi < a.g0[gi + 1]; pi++) {
            float2 p[E];
#pragma unroll
            for (int e = 0; e < E; e++) p[e] = make_float2(1.f, 1.f);
            for (int c = 0; c < a.nfac[pi]; c++) {
                const int f = a.fac[pi][c];
#pragma unroll
                for (int e = 0; e < E; e++) {
                    // select without dynamic indexing (keeps r in registers)
                    const float2 x = f == 0 ? r[0][e] : (f == 1 ? r[1][e] : r[2][e]);
                    p[e].x *= x.x;
                    p[e].y *= x.y;
                }
            }
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = cadd(v[e], p[e]);
        }
        row_r2c_store<NC, E, T>(v, active, a.X[gi] + hrow, L, t, tw);
    }
    if (a.noisy) {
        float2 v[E];
        const uint32_t base = (uint32_t)(row * (2L * NC));
#pragma unroll
        for (int e = 0; e < E; e++)
            v[e] = active ? philox_normal_pair(base + 2u * (uint32_t)(t + e * T), a.step_lo, a.step_hi, a.stream,
                                               a.key0, a.key1)
                          : make_float2(0.f, 0.f);
        row_r2c_store<NC, E, T>(v, active, a.eta + hrow, L, t, tw);
    }
}

// ------------------------------------------------------------ column kernel

template <int N>
struct FColCfg {
    static constexpr int E = N >= 512 ? 8 : elems_for(N);
    static constexpr int T = N / E;
    static constexpr int LL = lds_len(N);
    static constexpr bool valid(int twd) { return twd * T >= 64 && twd * T <= 512; }
};

template <int N, int TWD>
__global__ __launch_bounds__(512) void k_fused_cols(FusedColArgs a, const float2 *__restrict__ tw)
{
    using C = FColCfg<N>;
    constexpr int E = C::E, T = C::T;
    __shared__ float2 lds[TWD * C::LL];
    const Geom &g = a.u.g;
    const long ntiles = (g.nc1 + TWD - 1) / TWD;
    const long item = xcd_item(blockIdx.x, gridDim.x);
    if (item >= ntiles) return;
    const int c = threadIdx.x % TWD, t = threadIdx.x / TWD;
    const int kx = (int)(item * TWD + c);
    const bool active = kx < g.nc1;
    float2 *L = lds + c * C::LL;
    const long stride = g.nc1;
    float2 cur[E];
#pragma unroll
    for (int e = 0; e < E; e++)
        cur[e] = active ? a.spec_in[(long)(t + e * T) * stride + kx] : make_float2(0.f, 0.f);

    if (a.do_update) {
        float2 acc[E], eta[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            acc[e] = make_float2(0.f, 0.f);
            eta[e] = make_float2(0.f, 0.f);
        }
        for (int gi = 0; gi < a.ngroups; gi++) {
            float2 v[E];
#pragma unroll
            for (int e = 0; e < E; e++)
                v[e] = active ? a.X[gi][(long)(t + e * T) * stride + kx] : make_float2(0.f, 0.f);
            fft_line<N, E, -1>(v, L, t, tw);
            const TermRefDev &TR = a.u.terms[a.gterm[gi]];
#pragma unroll
            for (int e = 0; e < E; e++) {
                const Bin b = make_bin(g, kx, t + e * T, 0);
                acc[e] = cadd(acc[e], apply_mult(multiplier(a.u.pres, TR.pres0, TR.npres, TR.mult_i, b), TR.mult_i, v[e]));
            }
            __syncthreads();
        }
        if (a.u.noisy) {
#pragma unroll
            for (int e = 0; e < E; e++)
                eta[e] = active ? a.u.noise[(long)(t + e * T) * stride + kx] : make_float2(0.f, 0.f);
            fft_line<N, E, -1>(eta, L, t, tw);
            __syncthreads();
        }
#pragma unroll
        for (int e = 0; e < E; e++) {
            const Bin b = make_bin(g, kx, t + e * T, 0);
            float2 s = acc[e];
            for (int ti = 0; ti < a.u.nterms; ti++) {
                const TermRefDev &TR = a.u.terms[ti];
                if (TR.kind == TERM_GROUPED) continue;
                // single-field terms of phi read its current value
                const float2 X = TR.kind == TERM_CONST
                                     ? (b.h == 0 ? make_float2((float)g.n, 0.f) : make_float2(0.f, 0.f))
                                     : cur[e];
                s = cadd(s, apply_mult(multiplier(a.u.pres, TR.pres0, TR.npres, TR.mult_i, b), TR.mult_i, X));
            }
            cur[e] = finish_update(a.u, b, cur[e], s, eta[e]);
        }
        if (active) {
#pragma unroll
            for (int e = 0; e < E; e++) a.spec_out[(long)(t + e * T) * stride + kx] = cur[e];
        }
    }
    for (int f = 0; f < a.nreal; f++) {
        float2 v[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int ky = t + e * T;
            float2 w = cur[e];
            if (a.zero[f] || !keep_mode(g, a.p[f], kx, ky, 0)) {
                w = make_float2(0.f, 0.f);
            } else if (a.lmul[f].npres > 0) {
                const Bin b = make_bin(g, kx, ky, 0);
                const TermRefDev &LM = a.lmul[f];
                w = apply_mult(multiplier(a.lpres, LM.pres0, LM.npres, LM.mult_i, b), LM.mult_i, w);
            }
            v[e] = w;
        }
        fft_line<N, E, +1>(v, L, t, tw);
        if (active) {
#pragma unroll
            for (int e = 0; e < E; e++) a.Z[f][(long)(t + e * T) * stride + kx] = v[e];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ dispatch

#define FUSED_ROW_CASES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)
#define FUSED_COL_CASES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

bool fused_supported(int sx, int sy, int sz)
{
    return sz == 1 && sx >= 16 && sx <= 4096 && sy >= 8 && sy <= 4096;
}

hipError_t fused_rows(const FusedRowArgs &a, int nc, hipStream_t s)
{
    const float2 *tw = twiddles(s);
    if (!tw) return hipErrorOutOfMemory;
    switch (nc) {
#define X(NC)                                                                                       \
    case NC: {                                                                                      \
        dim3 grid((unsigned)((a.nrows + FRowCfg<NC>::LPB - 1) / FRowCfg<NC>::LPB));                 \
        hipLaunchKernelGGL((k_fused_rows<NC>), grid, dim3(FRowCfg<NC>::THREADS), 0, s, a, tw);      \
        return hipGetLastError();                                                                   \
    }
        FUSED_ROW_CASES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

template <int N, int TWD>
static void launch_fcols(const FusedColArgs &a, hipStream_t s, const float2 *tw)
{
    const long ntiles = (a.u.g.nc1 + TWD - 1) / TWD;
    hipLaunchKernelGGL((k_fused_cols<N, TWD>), dim3((unsigned)xcd_blocks(ntiles)), dim3(TWD * FColCfg<N>::T), 0, s,
                       a, tw);
}

template <int N>
static void fcols_n(const FusedColArgs &a, hipStream_t s, const float2 *tw)
{
    // narrowest valid tile: the column count of a 2D grid is small
    if constexpr (FColCfg<N>::valid(1)) launch_fcols<N, 1>(a, s, tw);
    else if constexpr (FColCfg<N>::valid(2)) launch_fcols<N, 2>(a, s, tw);
    else if constexpr (FColCfg<N>::valid(4)) launch_fcols<N, 4>(a, s, tw);
    else if constexpr (FColCfg<N>::valid(16)) launch_fcols<N, 16>(a, s, tw);
    else launch_fcols<N, 64>(a, s, tw);
}

hipError_t fused_cols(const FusedColArgs &a, int n, hipStream_t s)
{
    const float2 *tw = twiddles(s);
    if (!tw) return hipErrorOutOfMemory;
    switch (n) {
#define X(N)                  \
    case N:                   \
        fcols_n<N>(a, s, tw); \
        return hipGetLastError();
        FUSED_COL_CASES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

}  // namespace cupss
