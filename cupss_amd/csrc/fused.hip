// fused.hip -- the fused two-kernel step for 2D systems with one dynamic
// field phi whose other fields are "derived": constraint fields
// psi = M_psi(q) phi with no implicit operator and no noise (e.g. iqxh =
// iqx*h).  Covers Cahn-Hilliard, Edwards-Wilkinson and KPZ.  Every launch
// serves any number of instances (systems) of one size class through a
// table of instance descriptors (FusedInst) and a per-workgroup block table.
//
// Pipeline state between the kernels (natural half-spectrum layout):
//   Z_f  = IFFT_y(mask_f * L_f * phi_hat)   for each real field f needed by a
//          product (f = phi, L = 1; f = psi, L = M_psi)
//   X_g  = R2C_x(sum of the products of group g)  (rows transformed only)
//   eta  = R2C_x(xi)                        (real-space noise, rows only)
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
//
// Each workgroup copies the twiddle table of its transform length into LDS
// and issues its global loads before the first transform, so that their
// latencies overlap.
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

template <int NC, int E, int T>
__device__ __forceinline__ void load_half_row(const float2 *__restrict__ x, bool active, int t, float2 (&xa)[E],
                                              float2 (&ya)[E])
{
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int k = t + e * T;
        xa[e] = active ? x[k] : make_float2(0.f, 0.f);
        ya[e] = active ? x[NC - k] : make_float2(0.f, 0.f);
    }
}

// C2R pre-processing: Z'[k] from X[k] and X[NC-k]; the imaginary parts of
// the self-conjugate bins k = 0, NC are dropped (real-part projection)
template <int NC, int E, int T>
__device__ __forceinline__ void c2r_prep(const float2 (&xa)[E], const float2 (&ya)[E], int t, float2 (&v)[E],
                                         const float2 *tws)
{
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int k = t + e * T;
        float2 X = xa[e], Y = ya[e];
        if (k == 0) {
            X.y = 0.f;
            Y.y = 0.f;
        }
        const float2 Yc = cconj(Y);
        const float2 xe = cadd(X, Yc);
        const float2 xo = cmul(csub(X, Yc), tw_lookup<+1, 2 * NC>(tws, k, 2 * NC));
        v[e] = cadd(xe, mul_i<+1>(xo));
    }
}

// R2C of one row whose real values are packed as v[e] = (x[2m], x[2m+1])
template <int NC, int E, int T>
__device__ __forceinline__ void r2c_store(float2 (&v)[E], bool active, float2 *__restrict__ o, float2 *L, int t,
                                          const float2 *tws)
{
    fft_line<NC, E, -1, 2 * NC>(v, L, t, tws);
    if (active) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = t + e * T;
            const float2 A = v[e];
            const float2 B = cconj(L[lpad((NC - k) & (NC - 1))]);
            const float2 xe = cscale(cadd(A, B), 0.5f);
            const float2 xo = cscale(mul_i<-1>(csub(A, B)), 0.5f);
            o[k] = k == 0 ? make_float2(xe.x + xo.x, 0.0f)
                          : cadd(xe, cmul(tw_lookup<-1, 2 * NC>(tws, k, 2 * NC), xo));
        }
        if (t == 0) {
            const float2 z0 = L[lpad(0)];
            o[NC] = make_float2(z0.x - z0.y, 0.0f);
        }
    }
    __syncthreads();  // the LDS line is reused by the next transform
}

template <int NC>
__global__ __launch_bounds__(256) void k_frows(const FusedInst *__restrict__ insts,
                                               const FusedBlock *__restrict__ blocks, uint32_t step_rel,
                                               const float2 *__restrict__ sets)
{
    using C = FRowCfg<NC>;
    constexpr int E = C::E, T = C::T;
    __shared__ float2 tws[2 * NC];
    __shared__ float2 lds[C::LPB * C::LL];
    const FusedBlock blk = blocks[blockIdx.x];
    if (blk.inst < 0) return;
    const FusedInst &a = insts[blk.inst];
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    const long row = (long)blk.line + l;
    const bool active = row < (long)a.g.sy * a.g.sz;
    float2 *L = lds + l * C::LL;
    const long hrow = row * (NC + 1);
    const int nreal = a.nreal;

    float2 xa[E], ya[E];
    if (nreal > 0) load_half_row<NC, E, T>(a.Z[0] + hrow, active, t, xa, ya);
    load_twiddles<2 * NC>(tws, sets);
    __syncthreads();

    float2 r[FUSED_MAX_REAL][E];
#pragma unroll
    for (int f = 0; f < FUSED_MAX_REAL; f++) {
        if (f < nreal) {
            float2 v[E];
            c2r_prep<NC, E, T>(xa, ya, t, v, tws);
            if (f + 1 < nreal) load_half_row<NC, E, T>(a.Z[f + 1] + hrow, active, t, xa, ya);
            fft_line<NC, E, +1, 2 * NC>(v, L, t, tws);
#pragma unroll
            for (int e = 0; e < E; e++) r[f][e] = cscale(v[e], a.scale);
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
        r2c_store<NC, E, T>(v, active, a.X[gi] + hrow, L, t, tws);
    }
    if (a.noise_mode == NOISE_REAL) {
        const uint64_t step = a.step_base + step_rel;
        float2 v[E];
        const uint32_t base = (uint32_t)(row * (2L * NC));
#pragma unroll
        for (int e = 0; e < E; e++)
            v[e] = active ? philox_normal_pair(base + 2u * (uint32_t)(t + e * T), (uint32_t)step,
                                               (uint32_t)(step >> 32), a.stream, a.key0, a.key1)
                          : make_float2(0.f, 0.f);
        r2c_store<NC, E, T>(v, active, a.eta + hrow, L, t, tws);
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
__global__ __launch_bounds__(512) void k_fcols(const FusedInst *__restrict__ insts,
                                               const FusedBlock *__restrict__ blocks, int parity,
                                               uint32_t step_rel, int mode, const float2 *__restrict__ sets)
{
    using C = FColCfg<N>;
    constexpr int E = C::E, T = C::T;
    __shared__ float2 tws[N];
    __shared__ float2 lds[TWD * C::LL];
    const FusedBlock blk = blocks[blockIdx.x];
    if (blk.inst < 0) return;
    const FusedInst &a = insts[blk.inst];
    const Geom &g = a.g;
    const int c = threadIdx.x % TWD, t = threadIdx.x / TWD;
    const int kx = blk.line + c;
    const bool active = kx < g.nc1;
    float2 *L = lds + c * C::LL;
    const long stride = g.nc1;
    const bool upd = mode == 1;
    const int ngroups = upd ? a.ngroups : 0;
    const float2 *__restrict__ sin = a.spec[parity];

    float2 cur[E], xv[E], eta[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const long i = (long)(t + e * T) * stride + kx;
        cur[e] = active ? sin[i] : make_float2(0.f, 0.f);
        xv[e] = (active && ngroups > 0) ? a.X[0][i] : make_float2(0.f, 0.f);
        eta[e] = (active && upd && a.noise_mode == NOISE_REAL) ? a.eta[i] : make_float2(0.f, 0.f);
    }
    load_twiddles<N>(tws, sets);
    __syncthreads();

    if (upd) {
        const uint64_t step = a.step_base + step_rel;
        float2 acc[E];
#pragma unroll
        for (int e = 0; e < E; e++) acc[e] = make_float2(0.f, 0.f);
        for (int gi = 0; gi < ngroups; gi++) {
            if (gi > 0) {
#pragma unroll
                for (int e = 0; e < E; e++)
                    xv[e] = active ? a.X[gi][(long)(t + e * T) * stride + kx] : make_float2(0.f, 0.f);
            }
            fft_line<N, E, -1, N>(xv, L, t, tws);
            const TermRefDev &TR = a.u.terms[a.gterm[gi]];
#pragma unroll
            for (int e = 0; e < E; e++) {
                const Bin b = make_bin(g, kx, t + e * T, 0);
                acc[e] = cadd(acc[e],
                              apply_mult(multiplier(a.u.pres, TR.pres0, TR.npres, TR.mult_i, b), TR.mult_i, xv[e]));
            }
            __syncthreads();
        }
        if (a.noise_mode == NOISE_REAL) {
            fft_line<N, E, -1, N>(eta, L, t, tws);
            __syncthreads();
        } else if (a.noise_mode == NOISE_SPECTRAL) {
#pragma unroll
            for (int e = 0; e < E; e++)
                eta[e] = spectral_noise(g, kx, t + e * T, 0, (uint32_t)step, (uint32_t)(step >> 32), a.stream,
                                        a.key0, a.key1);
        }
#pragma unroll
        for (int e = 0; e < E; e++) {
            const Bin b = make_bin(g, kx, t + e * T, 0);
            float2 s = acc[e];
            for (int ti = 0; ti < a.u.nterms; ti++) {
                const TermRefDev &TR = a.u.terms[ti];
                if (TR.kind == TERM_GROUPED) continue;
                // the remaining terms are constants or linear in phi itself
                const float2 X = TR.kind == TERM_CONST
                                     ? (b.h == 0 ? make_float2((float)g.n, 0.f) : make_float2(0.f, 0.f))
                                     : cur[e];
                s = cadd(s, apply_mult(multiplier(a.u.pres, TR.pres0, TR.npres, TR.mult_i, b), TR.mult_i, X));
            }
            cur[e] = finish_update(a.u, b, cur[e], s, eta[e]);
        }
        if (active) {
            float2 *__restrict__ sout = a.spec[parity ^ 1];
#pragma unroll
            for (int e = 0; e < E; e++) sout[(long)(t + e * T) * stride + kx] = cur[e];
        }
    }
#pragma unroll
    for (int f = 0; f < FUSED_MAX_REAL; f++) {
        if (f < a.nreal) {
            float2 v[E];
            const bool zero = !upd && a.zero[f];
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int ky = t + e * T;
                float2 w = cur[e];
                if (zero || !keep_mode(g, a.p[f], kx, ky, 0)) {
                    w = make_float2(0.f, 0.f);
                } else if (a.lmul[f].npres > 0) {
                    const Bin b = make_bin(g, kx, ky, 0);
                    const TermRefDev &LM = a.lmul[f];
                    w = apply_mult(multiplier(a.lpres, LM.pres0, LM.npres, LM.mult_i, b), LM.mult_i, w);
                }
                v[e] = w;
            }
            fft_line<N, E, +1, N>(v, L, t, tws);
            if (active) {
#pragma unroll
                for (int e = 0; e < E; e++) a.Z[f][(long)(t + e * T) * stride + kx] = v[e];
            }
            __syncthreads();
        }
    }
}

// ------------------------------------------------------------ elementwise

// Instances whose step needs no transform (no products, noise drawn in
// spectral space or absent): one workgroup per spectral row.
__global__ __launch_bounds__(256) void k_felem(const FusedInst *__restrict__ insts,
                                               const FusedBlock *__restrict__ blocks, int parity, uint32_t step_rel)
{
    const FusedBlock blk = blocks[blockIdx.x];
    if (blk.inst < 0) return;
    const FusedInst &a = insts[blk.inst];
    const Geom &g = a.g;
    const int ky = blk.line % g.sy, kz = blk.line / g.sy;
    const uint64_t step = a.step_base + step_rel;
    const float2 *__restrict__ sin = a.spec[parity];
    float2 *__restrict__ sout = a.spec[parity ^ 1];
    for (int kx = threadIdx.x; kx < g.nc1; kx += blockDim.x) {
        const Bin b = make_bin(g, kx, ky, kz);
        const float2 cur = sin[b.h];
        float2 s = make_float2(0.f, 0.f);
        for (int ti = 0; ti < a.u.nterms; ti++) {
            const TermRefDev &TR = a.u.terms[ti];
            const float2 X = TR.kind == TERM_CONST ? (b.h == 0 ? make_float2((float)g.n, 0.f) : make_float2(0.f, 0.f))
                                                   : cur;
            s = cadd(s, apply_mult(multiplier(a.u.pres, TR.pres0, TR.npres, TR.mult_i, b), TR.mult_i, X));
        }
        float2 eta = make_float2(0.f, 0.f);
        if (a.noise_mode == NOISE_SPECTRAL)
            eta = spectral_noise(g, kx, ky, kz, (uint32_t)step, (uint32_t)(step >> 32), a.stream, a.key0, a.key1);
        sout[b.h] = finish_update(a.u, b, cur, s, eta);
    }
}

// Derived constraint fields after a run: psi = P(M_psi phi) from phi's
// spectrum at the start of the last step (setNotDynamic_k with one
// single-field term, no implicit operator, no noise).
__global__ __launch_bounds__(256) void k_fderived(const FusedInst *__restrict__ insts,
                                                  const FusedBlock *__restrict__ blocks, int parity)
{
    const FusedBlock blk = blocks[blockIdx.x];
    if (blk.inst < 0) return;
    const FusedInst &a = insts[blk.inst];
    const Geom &g = a.g;
    const int ky = blk.line % g.sy, kz = blk.line / g.sy;
    const float2 *__restrict__ src = a.spec[parity];
    for (int kx = threadIdx.x; kx < g.nc1; kx += blockDim.x) {
        const Bin b = make_bin(g, kx, ky, kz);
        const float2 v = src[b.h];
        for (int d = 0; d < a.nder; d++) {
            const TermRefDev &M = a.dmul[d];
            a.dout[d][b.h] = apply_mult(multiplier(a.dpres, M.pres0, M.npres, M.mult_i, b), M.mult_i, v);
        }
    }
}

hipError_t fused_derived_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int parity,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_fderived, dim3(nblocks), dim3(256), 0, s, inst, blocks, parity);
    return hipGetLastError();
}

// ------------------------------------------------------------ dispatch

#define FUSED_ROW_CASES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)
#define FUSED_COL_CASES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

bool fused_supported(int sx, int sy, int sz) { return sz == 1 && sx >= 16 && sx <= 4096 && sy >= 8 && sy <= 4096; }

int fused_rows_per_block(int nc)
{
    switch (nc) {
#define X(NC) \
    case NC:  \
        return FRowCfg<NC>::LPB;
        FUSED_ROW_CASES(X)
#undef X
    default:
        return 1;
    }
}

// tile width: narrowest valid (a single 2D instance has few columns);
// batches of many instances get the same and rely on their count
int fused_cols_per_block(int ny)
{
    switch (ny) {
#define X(N)                             \
    case N:                              \
        for (int w : {1, 2, 4, 16, 64})  \
            if (FColCfg<N>::valid(w)) return w; \
        return 64;
        FUSED_COL_CASES(X)
#undef X
    default:
        return 1;
    }
}

hipError_t fused_rows_launch(int nc, const FusedInst *inst, const FusedBlock *blocks, int nblocks, uint32_t step_rel,
                             hipStream_t s)
{
    const float2 *sets = twiddle_sets(s);
    if (!sets) return hipErrorOutOfMemory;
    switch (nc) {
#define X(NC)                                                                                               \
    case NC:                                                                                                \
        hipLaunchKernelGGL((k_frows<NC>), dim3(nblocks), dim3(FRowCfg<NC>::THREADS), 0, s, inst, blocks,   \
                           step_rel, sets);                                                                 \
        return hipGetLastError();
        FUSED_ROW_CASES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

template <int N>
static hipError_t fcols_n(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int parity,
                          uint32_t step_rel, int mode, hipStream_t s, const float2 *sets)
{
    const int w = fused_cols_per_block(N);
#define L(W)                                                                                                 \
    if constexpr (FColCfg<N>::valid(W)) {                                                                    \
        if (w == W) {                                                                                        \
            hipLaunchKernelGGL((k_fcols<N, W>), dim3(nblocks), dim3(W * FColCfg<N>::T), 0, s, inst, blocks,  \
                               parity, step_rel, mode, sets);                                                \
            return hipGetLastError();                                                                        \
        }                                                                                                    \
    }
    L(1) L(2) L(4) L(16) L(64)
#undef L
    return hipErrorInvalidValue;
}

hipError_t fused_cols_launch(int ny, const FusedInst *inst, const FusedBlock *blocks, int nblocks, int parity,
                             uint32_t step_rel, int mode, hipStream_t s)
{
    const float2 *sets = twiddle_sets(s);
    if (!sets) return hipErrorOutOfMemory;
    switch (ny) {
#define X(N) \
    case N:  \
        return fcols_n<N>(inst, blocks, nblocks, parity, step_rel, mode, s, sets);
        FUSED_COL_CASES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t fused_elem_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int parity,
                             uint32_t step_rel, hipStream_t s)
{
    hipLaunchKernelGGL(k_felem, dim3(nblocks), dim3(256), 0, s, inst, blocks, parity, step_rel);
    return hipGetLastError();
}

}  // namespace cupss
