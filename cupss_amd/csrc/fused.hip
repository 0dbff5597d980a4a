// fused.hip -- the fused two-kernel step for 2D systems: the first dynamic
// field phi, fields linear in a dynamic field ("derived": psi = M_psi(q) u
// with no implicit operator and no noise, e.g. iqxh = iqx*h) and stored
// fields -- further dynamic fields, constraints of dynamic-field products,
// other constraints and coefficient fields (fused.hpp).  Covers diffusion,
// Cahn-Hilliard, Edwards-Wilkinson, KPZ, model H and reaction-diffusion
// systems.  Every launch serves any number of instances (systems) of one
// size class through a table of instance descriptors (FusedInst) and a
// per-workgroup block table.
//
// Pipeline state between the kernels, all column-major (element (kx, y) at
// kx*sy + y) so that the column kernel streams whole contiguous columns:
//   phi_hat  two spectra, [parity] in and [parity^1] out
//   Z_f  = IFFT_y(mask_f * L_f * phi_hat)   for each real field f needed by a
//          product (f = phi, L = 1; f = psi, L = M_psi); only the columns
//          kx <= sx/(p_f+1) that the dealias mask keeps are stored
//   X_g  = R2C_x(sum of the products of group g)  (rows transformed only)
//   eta  = R2C_x(xi)                        (real-space noise, rows only)
// Per step:
//   row kernel:    a tile of LPB rows of each Z_f is staged through LDS
//                  (runs of LPB*8 contiguous bytes per column), C2R along x
//                  -> real dealiased fields -> products, noise -R2C-> X_g,
//                  eta rows, written back through the same tile
//   column kernel: X_g, eta columns -FFT_y-> spectral terms; constraints of
//                  dynamic-field products; update of phi (setDynamic_k);
//                  the stored fields; mask_f * L_f * u_new -IFFT_y-> Z_f
// Terms with equal prefactor lists form one group: M*FFT(a) + M*FFT(b) =
// M*FFT(a + b).  The order of phases A-D of the reference
// (src/evolver.cpp:199-226) is kept: derived fields are evaluated from phi
// at the start of the step (phase A/B) and used by phi's products in phase
// C; phi's own dealiased copy is zero before its first update (step 0).
//
// A run starts with a tiled transpose of the system's spectrum into the
// column-major buffer and ends with the transpose back (plus the derived
// fields); instances whose step needs no transform advance any number of
// steps per launch in registers (k_felem).
#include <hip/hip_runtime.h>

#include "fft_core.hpp"
#include "fused.hpp"
#include "philox.hpp"
#include "update_math.hpp"

namespace cupss {

// ------------------------------------------------------------ row kernel

// MAXR: real fields a row holds at once (2, 3, or FUSED_MAX_REAL for systems
// with more; those use 4 elements per lane to stay within the registers).
// Rows of 1024 real points also take 4 per lane (more waves for one large
// system).
// Tile row stride: at least lds_len(nc), since each row's transform works
// in place, and chosen for the LDS banks (32 of 4 bytes; a store is served
// 16 lanes at a time, a load 32).  A tile copy moves row pairs: the 16
// lanes of a store cover LPB/2 pairs, 32/(LPB/2) dwords of each, and
// successive pairs lie 4*RS dwords apart, so the runs interleave without
// sharing a bank when RS is odd (LPB 16), 2 mod 4 (LPB 8) or 4 mod 8 (LPB 4).
__host__ __device__ constexpr int tile_stride(int nc, int lpb)
{
    int rs = lds_len(nc);
    if (lpb == 16)
        while (rs % 2 != 1) rs++;
    else if (lpb == 8)
        while (rs % 4 != 2) rs++;
    else if (lpb == 4)
        while (rs % 8 != 4) rs++;
    return rs;
}

template <int NC, int LPB, int MAXR = 3>
struct FRowCfg {
    static constexpr int E = MAXR > 3 ? (NC >= 4 ? 4 : NC) : (NC >= 1024 ? 16 : (NC >= 512 ? 4 : (NC >= 64 ? 8 : (NC >= 4 ? 4 : NC))));
    static constexpr int T = NC / E;
    static constexpr int THREADS = LPB * T;
    static constexpr int RS = tile_stride(NC, LPB);
    static constexpr bool valid()
    {
        return THREADS >= 16 && THREADS <= 512 && (LPB * RS + 2 * NC + tw_stage_alloc(NC, E)) * 8 <= 150 * 1024;
    }
};

// four floats as one 16-byte access (a native vector: HIP's float4 copies
// through a generic pointer, which would make the access flat)
typedef float vf4 __attribute__((ext_vector_type(4)));

// rows y0 .. y0+LPB-1 of a column-major half spectrum into the tile (row l
// at l*RS, entry k at k); columns k > kmax read as zero.  Pairs of rows are
// moved with one 16-byte access.  issue() starts a thread's loads (all at
// once where it has at most CH), commit() writes them to the tile (and does
// the whole copy in rounds of CH loads where there are more), so that the
// loads are in flight together instead of one round trip each.
template <int NC, int LPB, int RS, int TH, int CH = 4>
struct TileLoad {
    static constexpr int HP = LPB / 2;
    static constexpr int ITEMS = (NC + 1) * (LPB >= 2 ? HP : 1);
    static constexpr int IT = (ITEMS + TH - 1) / TH;
    static constexpr bool ONE = IT <= CH;
    vf4 w[ONE ? IT : 1];
    __device__ __forceinline__ vf4 item(const float2 *__restrict__ src, int sy, int y0, int kmax, int i) const
    {
        vf4 v = {0.f, 0.f, 0.f, 0.f};
        if constexpr (LPB >= 2) {
            const int k = i / HP, yl = 2 * (i % HP);
            if (k <= kmax) v = gload(reinterpret_cast<const vf4 *>(src + (long)k * sy + y0 + yl), 0);
        } else {
            if (i <= kmax) {
                const float2 z = gload(src, (long)i * sy + y0);
                v.x = z.x;
                v.y = z.y;
            }
        }
        return v;
    }
    __device__ __forceinline__ void put(float2 *__restrict__ tile, int i, vf4 v) const
    {
        if constexpr (LPB >= 2) {
            const int k = i / HP, yl = 2 * (i % HP);
            tile[yl * RS + k] = make_float2(v.x, v.y);
            tile[(yl + 1) * RS + k] = make_float2(v.z, v.w);
        } else {
            tile[i] = make_float2(v.x, v.y);
        }
    }
    __device__ __forceinline__ void issue(const float2 *__restrict__ src, int sy, int y0, int kmax)
    {
        if constexpr (ONE) {
#pragma unroll
            for (int p = 0; p < IT; p++) {
                const int i = threadIdx.x + p * TH;
                if (i < ITEMS) w[p] = item(src, sy, y0, kmax, i);
            }
        }
    }
    __device__ __forceinline__ void commit(float2 *__restrict__ tile, const float2 *__restrict__ src, int sy, int y0,
                                           int kmax)
    {
        if constexpr (ONE) {
#pragma unroll
            for (int p = 0; p < IT; p++) {
                const int i = threadIdx.x + p * TH;
                if (i < ITEMS) put(tile, i, w[p]);
            }
        } else {
#pragma unroll
            for (int p0 = 0; p0 < IT; p0 += CH) {
                vf4 v[CH];
#pragma unroll
                for (int q = 0; q < CH; q++) {
                    const int i = threadIdx.x + (p0 + q) * TH;
                    if (p0 + q < IT && i < ITEMS) v[q] = item(src, sy, y0, kmax, i);
                }
#pragma unroll
                for (int q = 0; q < CH; q++) {
                    const int i = threadIdx.x + (p0 + q) * TH;
                    if (p0 + q < IT && i < ITEMS) put(tile, i, v[q]);
                }
            }
        }
    }
};

template <int NC, int LPB, int RS, int TH>
__device__ __forceinline__ void tile_load(float2 *__restrict__ tile, const float2 *__restrict__ src, int sy, int y0,
                                          int kmax)
{
    TileLoad<NC, LPB, RS, TH> tl;
    tl.issue(src, sy, y0, kmax);
    tl.commit(tile, src, sy, y0, kmax);
}

template <int NC, int LPB, int RS, int TH>
__device__ __forceinline__ void tile_store(float2 *__restrict__ dst, const float2 *__restrict__ tile, int sy, int y0)
{
    if constexpr (LPB >= 2) {
        constexpr int HP = LPB / 2;
        for (int i = threadIdx.x; i < (NC + 1) * HP; i += TH) {
            const int k = i / HP, yl = 2 * (i % HP);
            const float2 a = tile[yl * RS + k], b = tile[(yl + 1) * RS + k];
            const vf4 w = {a.x, a.y, b.x, b.y};
            gstore(reinterpret_cast<vf4 *>(dst + (long)k * sy + y0 + yl), 0, w);
        }
    } else {
        for (int k = threadIdx.x; k <= NC; k += TH) gstore(dst, (long)k * sy + y0, tile[k]);
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

// R2C of one row whose real values are packed as v[e] = (x[2m], x[2m+1]);
// the NC+1 bins are left in L[0..NC] (unpadded).  Ends with a barrier.
// tws: the 2NC-point table (post-twiddles); the NC-point transform takes
// its twiddles from ftw (TWF = TW_STAGE: the stage tables, else tws).
template <int NC, int E, int T, int TWF>
__device__ __forceinline__ void r2c_tile(float2 (&v)[E], float2 *L, int t, const float2 *tws, const float2 *ftw)
{
    fft_line<NC, E, -1, TWF, true, (T <= 64)>(v, L, t, ftw);
    float2 o[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int k = t + e * T;
        const float2 A = v[e];
        const float2 B = cconj(L[lpad((NC - k) & (NC - 1))]);
        const float2 xe = cscale(cadd(A, B), 0.5f);
        const float2 xo = cscale(mul_i<-1>(csub(A, B)), 0.5f);
        o[e] = k == 0 ? make_float2(xe.x + xo.x, 0.0f) : cadd(xe, cmul(tw_lookup<-1, 2 * NC>(tws, k, 2 * NC), xo));
    }
    const float2 z0 = L[lpad(0)];
    line_sync<(T <= 64)>();  // every read of the transform's LDS line is done
#pragma unroll
    for (int e = 0; e < E; e++) L[t + e * T] = o[e];
    if (t == 0) L[NC] = make_float2(z0.x - z0.y, 0.0f);
    __syncthreads();
}

// Rows y0 .. y0+LPB-1 of one instance (the whole workgroup).  load_tw:
// copy the twiddle tables into tws (2NC points) and tst (the NC-point
// transform's stage tables) first (otherwise they are there already).
// Ends with a workgroup barrier: the tile and the rows written are free.
template <int NC, int LPB, int MAXR>
__device__ __forceinline__ void frows_tile(const FusedInst &a, int y0, uint32_t step_rel,
                                           const float2 *__restrict__ sets, bool load_tw, float2 *tws, float2 *tst,
                                           float2 *tile)
{
    using C = FRowCfg<NC, LPB, MAXR>;
    constexpr int E = C::E, T = C::T, RS = C::RS, TH = C::THREADS;
    static_assert(T % 2 == 0 && E % 2 == 0, "noise lanes are paired");
    const int sy = a.g.sy;
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    float2 *L = tile + l * RS;
    const int nreal = a.nreal;
    // twiddle loads first, then the first tile's: their stores then wait
    // for the twiddles only
    SizeTwiddles<2 * NC, TH> twl;
    StageTwiddles<NC, E, TH> tsl;
    if (load_tw) {
        twl.load(sets);
        tsl.load(sets);
    }

    float2 r[MAXR][E];
#pragma unroll
    for (int f = 0; f < MAXR; f++) {
        if (f < nreal) {
            // wide rows hold up to FUSED_MAX_REAL fields in registers: one
            // load in flight there
            TileLoad<NC, LPB, RS, TH, (MAXR > 3 ? 1 : 4)> tl;
            tl.issue(a.Z[f], sy, y0, a.kmax[f]);
            if (f == 0 && load_tw) {
                twl.store(tws, sets);
                tsl.store(tst, sets);
            }
            tl.commit(tile, a.Z[f], sy, y0, a.kmax[f]);
            __syncthreads();
            float2 xa[E], ya[E], v[E];
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = t + e * T;
                xa[e] = L[k];
                ya[e] = L[NC - k];
            }
            // the row's part of the tile becomes its transform's LDS (the
            // row's lanes are the only ones that read or write it from here)
            line_sync<(T <= 64)>();
            c2r_prep<NC, E, T>(xa, ya, t, v, tws);
            fft_line<NC, E, +1, TW_STAGE, false, (T <= 64)>(v, L, t, tst);
#pragma unroll
            for (int e = 0; e < E; e++) r[f][e] = cscale(v[e], a.scale);
            __syncthreads();
        }
    }
    if (nreal == 0) {
        if (load_tw) {
            twl.store(tws, sets);
            tsl.store(tst, sets);
        }
        __syncthreads();  // twiddles
    }
    for (int gi = 0; gi < a.ngroups; gi++) {
        float2 v[E];
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = make_float2(0.f, 0.f);
        for (int pi = a.pg0[gi]; pi < a.pg0[gi + 1]; pi++) {
            float2 p[E];
            const float c0 = a.coef[pi];
#pragma unroll
            for (int e = 0; e < E; e++) p[e] = make_float2(c0, c0);
            // product = coef * prod_f r_f^cnt_f: static field index, uniform
            // powers (no dynamic indexing, so r stays in registers)
#pragma unroll
            for (int f = 0; f < MAXR; f++)
                for (int c = 0; c < a.cnt[pi][f]; c++) {
#pragma unroll
                    for (int e = 0; e < E; e++) {
                        p[e].x *= r[f][e].x;
                        p[e].y *= r[f][e].y;
                    }
                }
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = cadd(v[e], p[e]);
        }
        r2c_tile<NC, E, T, TW_STAGE>(v, L, t, tws, tst);
        tile_store<NC, LPB, RS, TH>(a.X[gi], tile, sy, y0);
        __syncthreads();
    }
    // real-space noise of phi (source -1) and of the stored dynamic fields
    // that draw it (their own key, stream and buffer)
    for (int src = -1; src < a.ncons; src++) {
        uint32_t stream, key0, key1;
        float2 *eta;
        if (src < 0) {
            if (a.noise_mode != NOISE_REAL) continue;
            stream = a.stream;
            key0 = a.key0;
            key1 = a.key1;
            eta = a.eta;
        } else {
            const FusedTarget &TG = a.tgt[1 + src];
            if (TG.noisy != NOISE_REAL) continue;
            stream = TG.stream;
            key0 = TG.key0;
            key1 = TG.key1;
            eta = TG.eta;
        }
        // Cells 2k, 2k+1 of row y0+l, k = t + e*T.  The Philox call with
        // counter (row*NC + k) >> 1 serves k and k^1, which sits in lane
        // t^1 (T is even): the even lane computes the calls of elements
        // e < E/2, the odd lane those of e >= E/2, and each hands its
        // partner the Box-Muller pair that the partner's cells take.
        const uint64_t step = a.step_base + step_rel + (src < 0 ? 0 : a.tgt[1 + src].step_off);
        const uint32_t rowbase = (uint32_t)((long)(y0 + l) * NC);
        const bool even = (t & 1) == 0;
        float2 v[E];
#pragma unroll
        for (int e2 = 0; e2 < E / 2; e2++) {
            const uint32_t k = (uint32_t)((t & ~1) + (even ? e2 : e2 + E / 2) * T);
            const Philox4 o = philox4x32_10((rowbase + k) >> 1, (uint32_t)step, (uint32_t)(step >> 32), stream,
                                            key0, key1);
            const float2 ze = box_muller(o.x[0], o.x[1]), zo = box_muller(o.x[2], o.x[3]);
            const float2 own = even ? ze : zo, snd = even ? zo : ze;
            const float2 rcv = make_float2(__shfl_xor(snd.x, 1), __shfl_xor(snd.y, 1));
            v[e2] = even ? own : rcv;
            v[e2 + E / 2] = even ? rcv : own;
        }
        r2c_tile<NC, E, T, TW_STAGE>(v, L, t, tws, tst);
        tile_store<NC, LPB, RS, TH>(eta, tile, sy, y0);
        __syncthreads();  // the tile is the next source's transform LDS
    }
}

template <int NC, int LPB, int MAXR>
__global__ __launch_bounds__((FRowCfg<NC, LPB, MAXR>::THREADS)) void k_frows(const FusedInst *__restrict__ insts,
                                                                          const FusedBlock *__restrict__ blocks,
                                                                          uint32_t step_rel,
                                                                          const float2 *__restrict__ sets)
{
    using C = FRowCfg<NC, LPB, MAXR>;
    __shared__ __align__(16) float2 tws[2 * NC];
    __shared__ __align__(16) float2 tst[tw_stage_alloc(NC, C::E)];
    __shared__ float2 tile[LPB * C::RS];
    const FusedBlock blk = blocks[blockIdx.x];
    if (blk.inst < 0) return;
    frows_tile<NC, LPB, MAXR>(insts[blk.inst], blk.line, step_rel, sets, true, tws, tst, tile);
}

// ------------------------------------------------------------ column kernel

// v[e] = p[col + t + e*T] for the E bins of a lane's column, or zeros for
// an inactive column (one uniform branch for the column, the pointer read
// once, constant offsets between the elements)
template <int E, int T>
__device__ __forceinline__ void load_col(float2 (&v)[E], const float2 *__restrict__ p, long col, int t, bool active)
{
    if (active) {
        const float2 *__restrict__ q = p + col + t;
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = gload(q, (long)e * T);
    } else {
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = make_float2(0.f, 0.f);
    }
}

// Sum of the terms of one target field over the E bins of a lane: product
// groups (FFT_y of X_g times the group multiplier) and single-field or
// constant terms.  xv holds X of the first group on entry when `pre`.
template <int N, int E, int T, bool SKIP>
__device__ __forceinline__ void target_sum(const FusedInst &a, const FusedTarget &TG, const ColBins<E> &cb,
                                           const float2 (&old)[E], const float2 (&nw)[E], float2 (&xv)[E],
                                           bool pre, int parity, bool active, long col, float2 *L, int t,
                                           const float2 *tws, float2 (&acc)[E])
{
    float m[E];
#pragma unroll
    for (int e = 0; e < E; e++) acc[e] = make_float2(0.f, 0.f);
    for (int gi = TG.g0; gi < TG.g1; gi++) {
        if (!(pre && gi == TG.g0)) load_col<E, T>(xv, a.X[gi], col, t, active);
        fft_line<N, E, -1, TW_STAGE, !SKIP, (T <= 64)>(xv, L, t, tws);
        const TermRefDev &GM = a.gmul[gi];
        multiplier_col<E>(a.pool, GM.pres0, GM.npres, GM.mult_i, cb, m);
#pragma unroll
        for (int e = 0; e < E; e++) acc[e] = cadd(acc[e], apply_mult(m[e], GM.mult_i, xv[e]));
        line_sync<(T <= 64)>();
    }
    for (int ti = 0; ti < TG.nterms; ti++) {
        const FusedTerm &TR = TG.term[ti];
        multiplier_col<E>(a.pool, TR.pres0, TR.npres, TR.mult_i, cb, m);
        float2 X[E];
        if (TR.src == SRC_OLD || TR.src == SRC_NEW) {
#pragma unroll
            for (int e = 0; e < E; e++) X[e] = TR.src == SRC_OLD ? old[e] : nw[e];
        } else if (TR.src == SRC_CONST) {
#pragma unroll
            for (int e = 0; e < E; e++) X[e] = cb.zero[e] ? make_float2((float)a.g.n, 0.f) : make_float2(0.f, 0.f);
        } else if (TR.src == SRC_DERIVED) {
            const TermRefDev &DM = a.dmul[TR.idx];
            float dm[E];
            multiplier_col<E>(a.pool, DM.pres0, DM.npres, DM.mult_i, cb, dm);
            const int b = a.dbase[TR.idx];
            if (b < 0) {
#pragma unroll
                for (int e = 0; e < E; e++) X[e] = old[e];
            } else {
                const float2 *__restrict__ src = a.cs[b][parity];
                load_col<E, T>(X, src, col, t, active);
            }
#pragma unroll
            for (int e = 0; e < E; e++) X[e] = apply_mult(dm[e], DM.mult_i, X[e]);
        } else {
            // SRC_DYN_NEW: a stored dynamic field's value after the step, which
            // this lane wrote to the other buffer earlier in the same pass
            const float2 *__restrict__ src = a.cs[TR.idx][TR.src == SRC_DYN_NEW ? parity ^ 1 : parity];
            load_col<E, T>(X, src, col, t, active);
        }
#pragma unroll
        for (int e = 0; e < E; e++) acc[e] = cadd(acc[e], apply_mult(m[e], TR.mult_i, X[e]));
    }
}

// Z_f = IFFT_y(mask_f * val) for the columns the mask keeps
template <int N, int E, int T, bool SKIP>
__device__ __forceinline__ void z_out(const FusedInst &a, int f, const Geom &g, const ColBins<E> &cb,
                                      const float2 (&val)[E], bool zero, int line, bool active, long col, float2 *L,
                                      int t, const float2 *tws)
{
    if (line > a.kmax[f]) return;  // uniform: every column of the workgroup is masked
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; e++)
        v[e] = (zero || !keep_mode2(g, a.kmax[f], a.kymax[f], cb.kx, t + e * T)) ? make_float2(0.f, 0.f) : val[e];
    fft_line<N, E, +1, TW_STAGE, !SKIP, (T <= 64)>(v, L, t, tws);
    if (active && cb.kx <= a.kmax[f]) {
        float2 *__restrict__ z = a.Z[f] + col + t;  // read once: the stores could alias the descriptor
#pragma unroll
        for (int e = 0; e < E; e++) gstore(z, (long)e * T, v[e]);
    }
    line_sync<(T <= 64)>();
}

// sum over the pool prefactors [p0, p0 + n) of c q^(2 q2n) |q|^-invq per bin
// (the implicit operator of a stored field)
template <int E>
__device__ __forceinline__ void pool_sum_col(const PresDev *pool, int p0, int n, const ColBins<E> &cb,
                                             float (&sum)[E])
{
#pragma unroll
    for (int e = 0; e < E; e++) sum[e] = 0.0f;
    for (int i = 0; i < n; i++) {
        const PresDev P = pool[p0 + i];
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
        for (int e = 0; e < E; e++) sum[e] += th[e];
    }
}

// noise of a stored field over the E bins of a lane, times its amplitude
// (times 1/sqrt(dt) for a constraint): real-space rows transformed along y,
// or drawn in spectral space, at the pass's step + step_off
template <int N, int E, int T, bool SKIP>
__device__ __forceinline__ void stored_noise(const FusedInst &a, const FusedTarget &TG, const ColBins<E> &cb, int kx,
                                             uint64_t step, bool active, long col, float2 *L, int t,
                                             const float2 *tws, float2 (&v)[E])
{
    float2 ue[E];
#pragma unroll
    for (int e = 0; e < E; e++) ue[e] = make_float2(0.f, 0.f);
    if (TG.noisy == NOISE_REAL) {
        load_col<E, T>(ue, TG.eta, col, t, active);
        fft_line<N, E, -1, TW_STAGE, !SKIP, (T <= 64)>(ue, L, t, tws);
        line_sync<(T <= 64)>();
    } else {
        const uint64_t s = step + (uint64_t)TG.step_off;
#pragma unroll
        for (int e = 0; e < E / 2; e++)
            spectral_noise_pair(a.g, kx, t + e * T, 0, (uint32_t)s, (uint32_t)(s >> 32), TG.stream, TG.key0, TG.key1,
                                ue[e], ue[e + E / 2]);
    }
#pragma unroll
    for (int e = 0; e < E; e++) {
        float pn = TG.noise_base;
        if (TG.noise_q2n != 0) pn *= sqrtf(ipow(cb.q2[e], TG.noise_q2n));
        if (TG.noise_invq != 0) pn *= sqrtf(ipow(col_iq(cb, e), TG.noise_invq));
        if (!TG.dyn) pn = TG.sdt * pn;
        v[e].x += pn * ue[e].x;
        v[e].y += pn * ue[e].y;
    }
}

// constraint value: the term sum divided by the implicit operator's sum
// where q != 0 (setNotDynamic_k, src/field_kernels.cu:130-197)
template <int E>
__device__ __forceinline__ void constraint_solve(const FusedInst &a, const FusedTarget &TG, const ColBins<E> &cb,
                                                 float2 (&v)[E])
{
    if (TG.nimp <= 0) return;
    float lsum[E];
    pool_sum_col<E>(a.pool, TG.imp0, TG.nimp, cb, lsum);
#pragma unroll
    for (int e = 0; e < E; e++)
        if (!cb.zero[e]) {
            const float rl = __builtin_amdgcn_rcpf(lsum[e]);
            v[e].x *= rl;
            v[e].y *= rl;
        }
}

// 1/D(q), D = 1 - dt sum of the pool prefactors [p0, p0 + n) (the implicit
// operator of a stored dynamic field, summed in implicit_col's order)
template <int E>
__device__ __forceinline__ void pool_implicit_col(const PresDev *pool, int p0, int n, float dt, const ColBins<E> &cb,
                                                  float (&rD)[E])
{
    float D[E];
#pragma unroll
    for (int e = 0; e < E; e++) D[e] = 1.0f;
    for (int i = 0; i < n; i++) {
        float th[E];
        pool_sum_col<E>(pool, p0 + i, 1, cb, th);
#pragma unroll
        for (int e = 0; e < E; e++) D[e] -= dt * th[e];
    }
#pragma unroll
    for (int e = 0; e < E; e++) rD[e] = __builtin_amdgcn_rcpf(D[e]);
}

// Z_f = IFFT_y(mask_f * L_f u) of a real field that is a dynamic field u or
// derived from it
template <int N, int E, int T, bool SKIP>
__device__ __forceinline__ void lz_out(const FusedInst &a, int f, const Geom &g, const ColBins<E> &cb,
                                       const float2 (&u)[E], bool zero, int line, bool active, long col, float2 *L,
                                       int t, const float2 *tws)
{
    float2 v[E];
    const TermRefDev &LM = a.lmul[f];
    if (LM.npres > 0) {
        float lm[E];
        multiplier_col<E>(a.pool, LM.pres0, LM.npres, LM.mult_i, cb, lm);
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = apply_mult(lm[e], LM.mult_i, u[e]);
    } else {
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = u[e];
    }
    z_out<N, E, T, SKIP>(a, f, g, cb, v, zero, line, active, col, L, t, tws);
}

// Hermitian projection of the columns kx = 0 and kx = sx/2 (2D): bin ky
// and its partner N - ky of such a column both become their Hermitian mean,
// as the reference's IFFT / real part / FFT round trip of every field every
// step makes them (src/field.cpp:59-89; the interior columns stand for their
// partners implicitly).  Without it the anti-Hermitian part of these columns
// -- invisible to C2R, fed by rounding -- grows in linearly unstable bands
// until it leaks into the real fields.  v[e] holds bin t + e*T of the lane's
// column; the whole workgroup calls this (barriers), only `edge` lines
// change.
template <int N, int E, int T>
__device__ __forceinline__ void hermitian_col(float2 (&v)[E], float2 *L, int t, bool edge)
{
#pragma unroll
    for (int e = 0; e < E; e++) L[lpad(t + e * T)] = v[e];
    line_sync<(T <= 64)>();
    if (edge) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            const float2 p = L[lpad((N - (t + e * T)) & (N - 1))];
            v[e] = make_float2(0.5f * (v[e].x + p.x), 0.5f * (v[e].y - p.y));
        }
    }
    line_sync<(T <= 64)>();
}

template <int N, int TWD>
struct FColCfg {
    // 1024- and 2048-point columns: 4 per lane, so that a single large
    // system still puts several waves on every SIMD
    static constexpr int E = N >= 4096 ? 16 : (N >= 1024 ? 4 : (N >= 8 ? 8 : N));
    static constexpr int T = N / E;
    static constexpr int THREADS = TWD * T;
    static constexpr int LL = lds_len(N);
    static constexpr bool valid() { return THREADS >= 16 && THREADS <= 512; }
};

// CONS: the instances may have stored constraint fields (compiled out
// otherwise, which keeps the kernel within 128 VGPRs: 4 waves per SIMD)
// Columns line .. line+TWD-1 of one instance (the whole workgroup); mode
// 0: prologue (the real fields of the current state), 1: step.  load_tw as
// for frows_tile.  Each column's LDS region is used by its own lanes only.
template <int N, int TWD, bool CONS, bool XDYN>
__device__ __forceinline__ void fcols_tile(const FusedInst &a, int line, int parity, uint32_t step_rel, int mode,
                                           const float2 *__restrict__ sets, bool load_tw, float2 *tws,
                                           float2 *lds)
{
    using C = FColCfg<N, TWD>;
    constexpr int E = C::E, T = C::T;
    const FusedBlock blk{0, line};
    const Geom &g = a.g;
    const int c = threadIdx.x / T, t = threadIdx.x % T;
    const int kx = blk.line + c;
    const bool active = kx < g.nc1;
    float2 *L = lds + c * C::LL;
    const long col = (long)kx * N;
    const bool upd = mode == 1;
    // phi has product groups: load the first early (unless fresh constraints,
    // solved first, need the register block)
    const bool pre = upd && a.tgt[0].g1 > a.tgt[0].g0 && !(XDYN && a.nfresh > 0);
    const float2 *__restrict__ sin = a.spec[parity];

    // the twiddle loads first: waiting for them then leaves the state's
    // loads in flight
    StageTwiddles<N, E, C::THREADS> twr;
    if (load_tw) twr.load(sets);
    float2 cur[E], xv[E];
    load_col<E, T>(cur, sin, col, t, active);
    load_col<E, T>(xv, pre ? a.X[a.tgt[0].g0] : sin, col, t, active && pre);

    ColBins<E> cb;
    col_bins<E, T>(g, kx, t, cb);
    const uint64_t step = a.step_base + step_rel;
    // phi's spectral noise, drawn while the state's loads are in flight and
    // added to phi_m at once (no register block held across the transforms)
    // where nothing else reads phi_m: no single-field terms, no stored
    // fields.  phi_{m+1} = (phi_m + pn eta + dt acc) / D then.  ky and
    // ky + N/2 are elements e and e + E/2 of this lane.
    const bool early = upd && a.noise_mode == NOISE_SPECTRAL && a.u.noisy && a.tgt[0].nterms == 0 &&
                       (!CONS || a.ncons == 0);
    if (early) {
        float2 eta[E];
#pragma unroll
        for (int e = 0; e < E / 2; e++)
            spectral_noise_pair(g, kx, t + e * T, 0, (uint32_t)step, (uint32_t)(step >> 32), a.stream, a.key0,
                                a.key1, eta[e], eta[e + E / 2]);
#pragma unroll
        for (int e = 0; e < E; e++) {
            const float pn = noise_amplitude_col<E>(a.u, cb, e);
            cur[e].x += pn * eta[e].x;
            cur[e].y += pn * eta[e].y;
        }
    }
    if (load_tw) {
        twr.store(tws, sets);
        __syncthreads();
    }
    float2 nw[E];  // phi_{m+1}
    // workgroups holding the column kx = 0 or kx = sx/2 (uniform test)
    const bool edge_wg = blk.line == 0 || blk.line + TWD > g.nc;
    const bool edge = active && (kx == 0 || kx == g.nc);
    if (upd) {
        // ---- fresh constraints (setNotDynamic_k of phase A, src/field_kernels.cu:
        // 130-197): their values of step m, from the dynamic fields of step m,
        // into the buffer the dynamic fields' terms read
        if (XDYN && a.nfresh > 0)
            for (int ci = 0; ci < a.ncons; ci++) {
                const FusedTarget &TG = a.tgt[1 + ci];
                if (!TG.fresh) continue;
                float2 v[E];
                target_sum<N, E, T, !CONS>(a, TG, cb, cur, cur, xv, false, parity, active, col, L, t, tws, v);
                if (TG.noisy) stored_noise<N, E, T, !CONS>(a, TG, cb, kx, step, active, col, L, t, tws, v);
                constraint_solve<E>(a, TG, cb, v);
                if (edge_wg) hermitian_col<N, E, T>(v, L, t, edge);
                if (active) {
                    float2 *__restrict__ cout = a.cs[ci][parity];
#pragma unroll
                    for (int e = 0; e < E; e++) gstore(cout, col + t + e * T, v[e]);
                }
            }
        // ---- phi (setDynamic_k, src/field_kernels.cu:199-227)
        float2 acc[E];
        target_sum<N, E, T, !CONS>(a, a.tgt[0], cb, cur, cur, xv, pre, parity, active, col, L, t, tws, acc);
        float2 eta[E];
#pragma unroll
        for (int e = 0; e < E; e++) eta[e] = make_float2(0.f, 0.f);
        if (a.noise_mode == NOISE_REAL) {
            // loaded here rather than with phi: fewer registers live across
            // the transforms of the terms
            load_col<E, T>(eta, a.eta, col, t, active);
            fft_line<N, E, -1, TW_STAGE, !CONS, (T <= 64)>(eta, L, t, tws);
            line_sync<(T <= 64)>();
        } else if (a.noise_mode == NOISE_SPECTRAL && !early) {
#pragma unroll
            for (int e = 0; e < E / 2; e++)
                spectral_noise_pair(g, kx, t + e * T, 0, (uint32_t)step, (uint32_t)(step >> 32), a.stream, a.key0,
                                    a.key1, eta[e], eta[e + E / 2]);
        }
        float rD[E];
        if (a.u.nimp > 0) implicit_col<E>(a.u, cb, rD);
#pragma unroll
        for (int e = 0; e < E; e++) {
            float2 s = cur[e];
            s.x += a.u.dt * acc[e].x;
            s.y += a.u.dt * acc[e].y;
            if (a.u.noisy && !early) {
                const float pn = noise_amplitude_col<E>(a.u, cb, e);
                s.x += pn * eta[e].x;
                s.y += pn * eta[e].y;
            }
            if (a.u.nimp > 0) {
                s.x *= rD[e];
                s.y *= rD[e];
            }
            nw[e] = s;
        }
        if (edge_wg) hermitian_col<N, E, T>(nw, L, t, edge);
        if (active) {
            float2 *__restrict__ sout = a.spec[parity ^ 1];
#pragma unroll
            for (int e = 0; e < E; e++) gstore(sout, col + t + e * T, nw[e]);
        }
        // ---- stored fields: further dynamic fields (setDynamic_k) from their
        // values before the step, then the constraints (setNotDynamic_k,
        // src/field_kernels.cu:130-197): their values of step m+1, from the
        // pre-step state of step m+1
        for (int ci = 0; ci < (CONS ? a.ncons : 0); ci++) {
            const FusedTarget &TG = a.tgt[1 + ci];
            if (XDYN && TG.fresh) continue;
            float2 v[E];
            target_sum<N, E, T, !CONS>(a, TG, cb, cur, nw, xv, false, parity, active, col, L, t, tws, v);
            if (XDYN && TG.dyn) {
                float2 u0[E];
                load_col<E, T>(u0, a.cs[ci][parity], col, t, active);
                const float dt = a.u.dt;
#pragma unroll
                for (int e = 0; e < E; e++) v[e] = make_float2(u0[e].x + dt * v[e].x, u0[e].y + dt * v[e].y);
                if (TG.noisy) stored_noise<N, E, T, !CONS>(a, TG, cb, kx, step, active, col, L, t, tws, v);
                if (TG.nimp > 0) {
                    float rD[E];
                    pool_implicit_col<E>(a.pool, TG.imp0, TG.nimp, a.u.dt, cb, rD);
#pragma unroll
                    for (int e = 0; e < E; e++) {
                        v[e].x *= rD[e];
                        v[e].y *= rD[e];
                    }
                }
            } else if (TG.keep) {
                // no terms: setNotDynamic_k leaves the value as it is, then
                // divides it by the implicit operator's sum (if any)
                const float2 *__restrict__ kin = a.cs[ci][parity];
                load_col<E, T>(v, kin, col, t, active);
                constraint_solve<E>(a, TG, cb, v);
            } else {
                if (XDYN && TG.noisy) stored_noise<N, E, T, !CONS>(a, TG, cb, kx, step, active, col, L, t, tws, v);
                constraint_solve<E>(a, TG, cb, v);
            }
            if (edge_wg) hermitian_col<N, E, T>(v, L, t, edge);
            if (active) {
                float2 *__restrict__ cout = a.cs[ci][parity ^ 1];
#pragma unroll
                for (int e = 0; e < E; e++) gstore(cout, col + t + e * T, v[e]);
            }
            if (TG.z >= 0) z_out<N, E, T, !CONS>(a, TG.z, g, cb, v, false, blk.line, active, col, L, t, tws);
            if (XDYN && TG.dyn)
                for (int f = 0; f < a.nreal; f++)
                    if (a.zsrc[f] < 0 && a.lbase[f] == ci)
                        lz_out<N, E, T, !CONS>(a, f, g, cb, v, false, blk.line, active, col, L, t, tws);
        }
    } else {
        // prologue: the real fields of the current state
#pragma unroll
        for (int e = 0; e < E; e++) nw[e] = cur[e];
        for (int ci = 0; ci < (CONS ? a.ncons : 0); ci++) {
            const int z = a.tgt[1 + ci].z;
            const bool dyn = XDYN && a.tgt[1 + ci].dyn != 0;
            if (z < 0 && !dyn) continue;
            float2 v[E];
            load_col<E, T>(v, a.cs[ci][0], col, t, active);
            if (z >= 0) z_out<N, E, T, !CONS>(a, z, g, cb, v, a.zero[z] != 0, blk.line, active, col, L, t, tws);
            if (dyn)
                for (int f = 0; f < a.nreal; f++)
                    if (a.zsrc[f] < 0 && a.lbase[f] == ci)
                        lz_out<N, E, T, !CONS>(a, f, g, cb, v, a.zero[f] != 0, blk.line, active, col, L, t, tws);
        }
    }
    // real fields that are phi or derived from it: L_f phi_{m+1}
    for (int f = 0; f < a.nreal; f++) {
        if (a.zsrc[f] >= 0 || a.lbase[f] >= 0) continue;
        lz_out<N, E, T, !CONS>(a, f, g, cb, nw, !upd && a.zero[f] != 0, blk.line, active, col, L, t, tws);
    }
}

template <int N, int TWD, bool CONS, bool XDYN>
__global__ __launch_bounds__((FColCfg<N, TWD>::THREADS)) void k_fcols(const FusedInst *__restrict__ insts,
                                                                   const FusedBlock *__restrict__ blocks,
                                                                   int parity, uint32_t step_rel, int mode,
                                                                   const float2 *__restrict__ sets,
                                                                   unsigned long long *__restrict__ ts)
{
    using C = FColCfg<N, TWD>;
    __shared__ __align__(16) float2 tws[tw_stage_alloc(N, C::E)];
    __shared__ float2 lds[TWD * C::LL];
    const FusedBlock blk = blocks[blockIdx.x];
    if (blk.inst < 0) return;
    if (mode == 0 && insts[blk.inst].resume) return;  // Z_f of this state left by the previous run
    // launch timing (ts != null): earliest workgroup start and latest end in
    // ticks of the constant-rate wall clock, spread over FUSED_TS_SLOTS
    // cache lines so that the workgroups' atomics do not contend
    unsigned long long *slot = ts ? ts + (size_t)(blockIdx.x % FUSED_TS_SLOTS) * FUSED_TS_STRIDE : nullptr;
    if (slot && threadIdx.x == 0) atomicMin(&slot[0], (unsigned long long)wall_clock64());
    fcols_tile<N, TWD, CONS, XDYN>(insts[blk.inst], blk.line, parity, step_rel, mode, sets, true, tws, lds);
    if (slot) {
        __syncthreads();
        if (threadIdx.x == 0) atomicMax(&slot[1], (unsigned long long)wall_clock64());
    }
}

// ------------------------------------------------------------ small instances

// Instances small enough for one workgroup (64^2, 128^2): K steps in one
// launch, one workgroup per instance, each step the same row and column
// tiles as the two-kernel step (so the same results bit for bit) in turn,
// separated by workgroup barriers -- no launch per pass, and the
// instance's buffers stay in the L2 of its workgroup's XCD.
template <int NC, int N, bool CONS, bool XDYN>
struct FSmallCfg {
    static constexpr int THREADS = 512;
    static constexpr int LPB = THREADS / FRowCfg<NC, 1, 3>::T;
    static constexpr int TWD = THREADS / FColCfg<N, 1>::T;
    static constexpr int LDSN = LPB * FRowCfg<NC, LPB, 3>::RS > TWD * FColCfg<N, TWD>::LL
                                    ? LPB * FRowCfg<NC, LPB, 3>::RS
                                    : TWD * FColCfg<N, TWD>::LL;
    static constexpr bool valid()
    {
        return FRowCfg<NC, LPB, 3>::valid() && FColCfg<N, TWD>::valid() && LPB <= N &&
               (2 * NC + tw_stage_alloc(NC, FRowCfg<NC, LPB, 3>::E) + tw_stage_alloc(N, FColCfg<N, TWD>::E) + LDSN) *
                       8 <=
                   150 * 1024;
    }
};

template <int NC, int N, bool CONS, bool XDYN>
__global__ __launch_bounds__(512) void k_fsmall(const FusedInst *__restrict__ insts, const int *__restrict__ ids,
                                                int n0, int k, int rel0, const float2 *__restrict__ sets,
                                                unsigned long long *__restrict__ ts)
{
    using C = FSmallCfg<NC, N, CONS, XDYN>;
    __shared__ float2 twr[2 * NC];
    __shared__ __align__(16) float2 tsr[tw_stage_alloc(NC, FRowCfg<NC, C::LPB, 3>::E)];
    __shared__ __align__(16) float2 twc[tw_stage_alloc(N, FColCfg<N, C::TWD>::E)];
    __shared__ float2 buf[C::LDSN];
    unsigned long long *slot = ts ? ts + (size_t)(blockIdx.x % FUSED_TS_SLOTS) * FUSED_TS_STRIDE : nullptr;
    if (slot && threadIdx.x == 0) atomicMin(&slot[0], (unsigned long long)wall_clock64());
    const FusedInst &a = insts[ids[blockIdx.x]];
    load_twiddles<2 * NC>(twr, sets);
    load_stage_twiddles<NC, FRowCfg<NC, C::LPB, 3>::E>(tsr, sets);
    load_stage_twiddles<N, FColCfg<N, C::TWD>::E>(twc, sets);
    __syncthreads();
    for (int n = n0; n < n0 + k; n++) {
        const uint32_t rel = (uint32_t)(n - rel0);
        for (int y0 = 0; y0 < N; y0 += C::LPB) frows_tile<NC, C::LPB, 3>(a, y0, rel, sets, false, twr, tsr, buf);
        // rows ended with a barrier: X_g and eta are written, the buffer is free
        for (int x0 = 0; x0 <= NC; x0 += C::TWD) fcols_tile<N, C::TWD, CONS, XDYN>(a, x0, n & 1, rel, 1, sets, false, twc, buf);
        __syncthreads();  // Z_f and the spectra written; the buffer is free for the rows
    }
    if (slot && threadIdx.x == 0) atomicMax(&slot[1], (unsigned long long)wall_clock64());
}

// ------------------------------------------------------------ transposes

// tile (tx, ty) of an instance: columns kx in [tx*S, tx*S+S), rows ky in
// [ty*S, ty*S+S) of the half spectrum
__device__ __forceinline__ void tile_origin(const Geom &g, int line, int &x0, int &y0)
{
    const int ntx = (g.nc1 + FUSED_TILE - 1) / FUSED_TILE;
    x0 = (line % ntx) * FUSED_TILE;
    y0 = (line / ntx) * FUSED_TILE;
}

// one tile of a natural half spectrum into a column-major one
__device__ __forceinline__ void tile_to_cm(float2 (&s)[FUSED_TILE][FUSED_TILE + 1], const Geom &g, int x0, int y0,
                                           const float2 *__restrict__ src, float2 *__restrict__ dst)
{
    constexpr int S = FUSED_TILE;
    for (int q = threadIdx.x; q < S * S; q += blockDim.x) {
        const int i = q % S, j = q / S;
        if (x0 + i < g.nc1 && y0 + j < g.sy) s[j][i] = gload(src, (long)(y0 + j) * g.nc1 + x0 + i);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < S * S; q += blockDim.x) {
        const int j = q % S, i = q / S;
        if (x0 + i < g.nc1 && y0 + j < g.sy) gstore(dst, (long)(x0 + i) * g.sy + y0 + j, s[j][i]);
    }
    __syncthreads();
}

// one tile of a column-major half spectrum into a natural one, each value
// times M (npres = 0: copied)
__device__ __forceinline__ void tile_to_nat(float2 (&s)[FUSED_TILE][FUSED_TILE + 1], const FusedInst &a, int x0,
                                            int y0, const float2 *__restrict__ src, float2 *__restrict__ dst,
                                            const TermRefDev &M)
{
    constexpr int S = FUSED_TILE;
    const Geom &g = a.g;
    for (int q = threadIdx.x; q < S * S; q += blockDim.x) {
        const int j = q % S, i = q / S;
        if (x0 + i < g.nc1 && y0 + j < g.sy) s[j][i] = gload(src, (long)(x0 + i) * g.sy + y0 + j);
    }
    __syncthreads();
    for (int q = threadIdx.x; q < S * S; q += blockDim.x) {
        const int i = q % S, j = q / S;
        const int kx = x0 + i, ky = y0 + j;
        if (kx < g.nc1 && ky < g.sy) {
            const long h = (long)ky * g.nc1 + kx;
            if (M.npres == 0) {
                gstore(dst, h, s[j][i]);
            } else {
                const Bin b = make_bin(g, kx, ky, 0);
                gstore(dst, h, apply_mult(multiplier(a.pool, M.pres0, M.npres, M.mult_i, b), M.mult_i, s[j][i]));
            }
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_ftrans_in(const FusedInst *__restrict__ insts,
                                                   const FusedBlock *__restrict__ blocks)
{
    __shared__ float2 s[FUSED_TILE][FUSED_TILE + 1];
    const FusedBlock blk = blocks[blockIdx.x];
    if (blk.inst < 0) return;
    const FusedInst &a = insts[blk.inst];
    int x0, y0;
    tile_origin(a.g, blk.line, x0, y0);
    if (!a.resume) tile_to_cm(s, a.g, x0, y0, a.nat, a.spec[0]);
    if (!a.cont)
        for (int c = 0; c < a.ncons; c++) tile_to_cm(s, a.g, x0, y0, a.cnat[c], a.cs[c][0]);
}

// column-major phi and stored dynamic fields [fin] -> natural; stored
// constraints of the last step's start ([fin^1]) -> natural; the derived
// fields of the last step's starting values of their base ([fin^1]) in
// natural layout
__global__ __launch_bounds__(256) void k_fsettle(const FusedInst *__restrict__ insts,
                                                 const FusedBlock *__restrict__ blocks, int fin)
{
    __shared__ float2 s[FUSED_TILE][FUSED_TILE + 1];
    const FusedBlock blk = blocks[blockIdx.x];
    if (blk.inst < 0) return;
    const FusedInst &a = insts[blk.inst];
    int x0, y0;
    tile_origin(a.g, blk.line, x0, y0);
    const TermRefDev copy{nullptr, 0, 0, 0, 0};
    tile_to_nat(s, a, x0, y0, a.spec[fin], a.nat, copy);
    for (int c = 0; c < a.ncons; c++)
        tile_to_nat(s, a, x0, y0, a.cs[c][a.tgt[1 + c].dyn ? fin : fin ^ 1], a.cnat[c], copy);
    for (int d = 0; d < a.nder; d++)
        tile_to_nat(s, a, x0, y0, a.dbase[d] < 0 ? a.spec[fin ^ 1] : a.cs[a.dbase[d]][fin ^ 1], a.dout[d], a.dmul[d]);
}

// ------------------------------------------------------------ elementwise

// Instances whose step needs no transform (no products, noise drawn in
// spectral space or absent): one thread per pair of bins (kx, j, kz),
// (kx, j + sy/2, kz) -- the pair that shares a noise draw -- advances
// nsteps steps in registers.  The explicit terms are linear in phi or
// constant, so their sum is (ar + i ai) phi_hat + cst; the step is the
// dynamic branch of finish_update with D(q) and the noise amplitude
// evaluated once.
__global__ __launch_bounds__(FUSED_ELEM_BLOCK) void k_felem(const FusedInst *__restrict__ insts,
                                                            const FusedBlock *__restrict__ blocks, int parity,
                                                            uint32_t step_rel, int nsteps)
{
    const FusedBlock blk = blocks[blockIdx.x];
    if (blk.inst < 0) return;
    const FusedInst &a = insts[blk.inst];
    const Geom &g = a.g;
    const int hy = g.sy / 2;  // sy >= 8 for fused instances
    const long pidx = (long)blk.line + threadIdx.x;
    if (pidx >= g.nh / 2) return;  // no barriers below
    // Pairs of the interior columns 1..nc-1 first, then those of the kx = 0
    // and kx = nc planes, whose noise takes two Philox calls (Hermitian
    // partners): only the last waves take that branch, none diverges on it.
    const long prow = (long)hy * g.sz, nint = (long)(g.nc - 1) * prow;
    int kx;
    long r;
    if (pidx < nint) {
        kx = 1 + (int)(pidx % (g.nc - 1));
        r = pidx / (g.nc - 1);
    } else {
        const long q = pidx - nint;
        kx = (q & 1) ? g.nc : 0;
        r = q >> 1;
    }
    const int j = (int)(r % hy), kz = (int)(r / hy);
    const UpdateArgs &u = a.u;
    const bool noisy = u.noisy && a.noise_mode == NOISE_SPECTRAL;
    Bin b[2];
    float ar[2], ai[2], D[2], pn[2];
    float2 cst[2], cur[2];
#pragma unroll
    for (int w = 0; w < 2; w++) {
        b[w] = make_bin(g, kx, j + w * hy, kz);
        ar[w] = 0.f;
        ai[w] = 0.f;
        cst[w] = make_float2(0.f, 0.f);
        for (int ti = 0; ti < u.nterms; ti++) {
            const TermRefDev &TR = u.terms[ti];
            const float m = multiplier(u.pres, TR.pres0, TR.npres, TR.mult_i, b[w]);
            if (TR.kind == TERM_CONST) {
                if (b[w].h == 0) cst[w] = cadd(cst[w], apply_mult(m, TR.mult_i, make_float2((float)g.n, 0.f)));
            } else if (TR.mult_i) {
                ai[w] += m;
            } else {
                ar[w] += m;
            }
        }
        D[w] = 1.0f;
        for (int i = 0; i < u.nimp; i++) {
            const PresDev P = u.imp[i];
            float th = P.c;
            if (P.q2n != 0) th *= ipow(b[w].q2, P.q2n);
            if (P.invq != 0) th *= ipow(b[w].iqd, P.invq);
            D[w] -= u.dt * th;
        }
        D[w] = __builtin_amdgcn_rcpf(D[w]);  // 1/D(q): hardware reciprocal, once per bin
        pn[w] = noisy ? noise_amplitude(u, b[w]) : 0.f;
        cur[w] = gload(a.spec[parity], b[w].h);
    }
    const uint64_t s0 = a.step_base + step_rel;
    for (int n = 0; n < nsteps; n++) {
        if (n == nsteps - 1) {
            for (int d = 0; d < a.nder; d++) {
                const TermRefDev &M = a.dmul[d];
#pragma unroll
                for (int w = 0; w < 2; w++)
                    gstore(a.dout[d], b[w].h,
                           apply_mult(multiplier(a.pool, M.pres0, M.npres, M.mult_i, b[w]), M.mult_i, cur[w]));
            }
        }
        float2 eta[2] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
        if (noisy) {
            const uint64_t step = s0 + (uint64_t)n;
            spectral_noise_pair(g, kx, j, kz, (uint32_t)step, (uint32_t)(step >> 32), a.stream, a.key0, a.key1,
                                eta[0], eta[1]);
        }
#pragma unroll
        for (int w = 0; w < 2; w++) {
            const float2 c = cur[w];
            float2 s = c;
            s.x += u.dt * (ar[w] * c.x - ai[w] * c.y + cst[w].x);
            s.y += u.dt * (ar[w] * c.y + ai[w] * c.x + cst[w].y);
            if (noisy) {
                s.x += pn[w] * eta[w].x;
                s.y += pn[w] * eta[w].y;
            }
            if (u.nimp > 0) {
                s.x *= D[w];
                s.y *= D[w];
            }
            cur[w] = s;
        }
    }
#pragma unroll
    for (int w = 0; w < 2; w++) gstore(a.spec[parity ^ (nsteps & 1)], b[w].h, cur[w]);
}

// ------------------------------------------------------------ 3D systems
//
// A 3D system runs alone (insts[0]) with every buffer in the natural half-
// spectrum layout, element (kx, ky, kz) at (kz*sy + ky)*nc1 + kx, so no
// transposes are needed: rows are contiguous, and lines along y and z are
// read TWD adjacent kx at a time.  Per step:
//   rows  (k_f3rows): Z_f rows -C2R_x-> real fields -> products -R2C_x-> X_g
//   y     (lines_y):  X_g = FFT_y(X_g)
//   z     (k_f3z):    X_g -FFT_z-> terms; update of phi (setDynamic_k);
//                     mask_f * L_f * phi_new -IFFT_z-> W_f
//   y     (lines_y):  Z_f = IFFT_y(W_f), lines kx <= kmax_f
// W_f is written only on the (kx, ky) lines the dealias mask of f keeps; the
// others stay zero from allocation.  The systems have no stored constraint
// fields and no real-space noise (the host checks).

template <int NC, int LPB, int MAXR>
__global__ __launch_bounds__((FRowCfg<NC, LPB, MAXR>::THREADS)) void k_f3rows(const FusedInst *__restrict__ insts,
                                                                           const float2 *__restrict__ sets)
{
    using C = FRowCfg<NC, LPB, MAXR>;
    constexpr int E = C::E, T = C::T, RS = C::RS;
    __shared__ float2 tws[2 * NC];
    __shared__ float2 tile[LPB * RS];
    const FusedInst &a = insts[0];
    const int nc1 = a.g.nc1;
    const long nrows = (long)a.g.sy * a.g.sz;
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    const long row = (long)blockIdx.x * LPB + l;
    const bool active = row < nrows;
    float2 *L = tile + l * RS;
    const int nreal = a.nreal;
    load_twiddles<2 * NC>(tws, sets);
    __syncthreads();

    float2 r[MAXR][E];
#pragma unroll
    for (int f = 0; f < MAXR; f++) {
        if (f < nreal) {
            const float2 *__restrict__ z = a.Z[f] + row * nc1;
            const int km = a.kmax[f];
            float2 xa[E], ya[E], v[E];
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = t + e * T;
                xa[e] = (active && k <= km) ? gload(z, k) : make_float2(0.f, 0.f);
                ya[e] = (active && NC - k <= km) ? gload(z, NC - k) : make_float2(0.f, 0.f);
            }
            c2r_prep<NC, E, T>(xa, ya, t, v, tws);
            fft_line<NC, E, +1, 2 * NC, false, (T <= 64)>(v, L, t, tws);
#pragma unroll
            for (int e = 0; e < E; e++) r[f][e] = cscale(v[e], a.scale);
            line_sync<(T <= 64)>();
        }
    }
    for (int gi = 0; gi < a.ngroups; gi++) {
        float2 v[E];
#pragma unroll
        for (int e = 0; e < E; e++) v[e] = make_float2(0.f, 0.f);
        for (int pi = a.pg0[gi]; pi < a.pg0[gi + 1]; pi++) {
            float2 p[E];
            const float c0 = a.coef[pi];
#pragma unroll
            for (int e = 0; e < E; e++) p[e] = make_float2(c0, c0);
#pragma unroll
            for (int f = 0; f < MAXR; f++)
                for (int c = 0; c < a.cnt[pi][f]; c++) {
#pragma unroll
                    for (int e = 0; e < E; e++) {
                        p[e].x *= r[f][e].x;
                        p[e].y *= r[f][e].y;
                    }
                }
#pragma unroll
            for (int e = 0; e < E; e++) v[e] = cadd(v[e], p[e]);
        }
        r2c_tile<NC, E, T, 2 * NC>(v, L, t, tws, tws);
        if (active) {
            float2 *__restrict__ x = a.X[gi] + row * nc1;
            for (int k = t; k <= NC; k += T) gstore(x, k, L[k]);
        }
        __syncthreads();
    }
}

// q of the E bins (kx, ky, t + e*T) that one lane of a z-line kernel holds
template <int E>
struct LineBins {
    float qx, qy;
    bool nyx, nyy;
    bool axis;              // kx = ky = 0 (the reference's i = j = 0 line, Q8)
    float qz[E], q2[E];
    bool nyz[E], zero[E];   // Nyquist plane; the q = 0 bin
};

template <int E, int T>
__device__ __forceinline__ void line_bins(const Geom &g, int kx, int ky, int t, LineBins<E> &lb)
{
    lb.qx = (float)(kx < (g.sx + 1) / 2 ? kx : kx - g.sx) * g.sqx;
    lb.qy = (float)(ky < (g.sy + 1) / 2 ? ky : ky - g.sy) * g.sqy;
    lb.nyx = g.sx > 1 && kx == g.sx / 2;
    lb.nyy = g.sy > 1 && ky == g.sy / 2;
    lb.axis = kx == 0 && ky == 0;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int kz = t + e * T;
        lb.qz[e] = (float)(kz < (g.sz + 1) / 2 ? kz : kz - g.sz) * g.sqz;
        lb.q2[e] = lb.qx * lb.qx + lb.qy * lb.qy + lb.qz[e] * lb.qz[e];
        lb.nyz[e] = g.sz > 1 && kz == g.sz / 2;
        lb.zero[e] = lb.axis && kz == 0;
    }
}

// multiplier() for the E bins of a z line (same operation order)
template <int E>
__device__ __forceinline__ void multiplier_line(const PresDev *__restrict__ pres, int p0, int np, int mult_i,
                                                const LineBins<E> &lb, float (&m)[E])
{
#pragma unroll
    for (int e = 0; e < E; e++) m[e] = 0.0f;
    for (int p = p0; p < p0 + np; p++) {
        const PresDev P = pres[p];
        const int parl = mult_i + (lb.nyx ? 0 : P.a) + (lb.nyy ? 0 : P.b);
        const float qxa = ipow(lb.qx, P.a), qyb = ipow(lb.qy, P.b);
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int par = parl + (lb.nyz[e] ? 0 : P.cz);
            float v = (par & 1) ? 0.0f : P.c;
            if (P.q2n > 0) v *= ipow(lb.q2[e], P.q2n);
            if (P.a > 0) v *= qxa;
            if (P.b > 0) v *= qyb;
            if (P.cz > 0) v *= ipow(lb.qz[e], P.cz);
            if (P.invq > 0) v *= ipow(lb.zero[e] ? 0.0f : __builtin_amdgcn_rsqf(lb.q2[e]), P.invq);
            m[e] += v;
        }
    }
}

template <int N, int TWD>
struct F3zCfg {
    static constexpr int E = N >= 8 ? 8 : N;
    static constexpr int T = N / E;
    static constexpr int THREADS = TWD * T;
    static constexpr int LL = lds_len(N);
    static constexpr bool valid() { return THREADS >= 64 && THREADS <= 512; }
};

// Lines along z of TWD adjacent kx at one ky (thread = t*TWD + c: a row of
// TWD values is one contiguous access).  Workgroups are placed XCD-aware.
// Register budget for 4 waves per SIMD (128 VGPRs): the pass waits on
// strided HBM accesses, which more resident waves hide.
template <int N, int TWD>
__global__ __launch_bounds__((F3zCfg<N, TWD>::THREADS)) __attribute__((amdgpu_waves_per_eu(4))) void k_f3z(const FusedInst *__restrict__ insts, int parity,
                                                                 uint32_t step_rel, int mode,
                                                                 const float2 *__restrict__ sets,
                                                                 unsigned long long *__restrict__ ts)
{
    using C = F3zCfg<N, TWD>;
    constexpr int E = C::E, T = C::T;
    __shared__ float2 tws[N];
    __shared__ float2 lds[TWD * C::LL];
    const FusedInst &a = insts[0];
    const Geom &g = a.g;
    const int ntx = (g.nc1 + TWD - 1) / TWD;
    const long item = xcd_item(blockIdx.x, gridDim.x);
    if (item >= (long)ntx * g.sy) return;  // whole workgroup idle
    unsigned long long *slot = ts ? ts + (size_t)(blockIdx.x % FUSED_TS_SLOTS) * FUSED_TS_STRIDE : nullptr;
    if (slot && threadIdx.x == 0) atomicMin(&slot[0], (unsigned long long)wall_clock64());
    // Each XCD takes a contiguous range of items; the lines the dealias masks
    // keep (which also transform W_f) have small |ky|, so ky is dealt with
    // stride 8 over the ranges to give every XCD its share of them.
    const int kr = (int)(item / ntx), yq = g.sy >> 3;
    const int kx0 = (int)(item % ntx) * TWD, ky = (kr % yq) * 8 + kr / yq;
    const int c = threadIdx.x % TWD, t = threadIdx.x / TWD;
    const int kx = kx0 + c;
    const bool active = kx < g.nc1;
    float2 *L = lds + c * C::LL;
    const long plane = (long)g.nc1 * g.sy;
    const long base = (long)ky * g.nc1 + kx;  // + kz * plane
    const bool upd = mode != 0;
    const FusedTarget &TG = a.tgt[0];
    const float2 *__restrict__ sin = a.spec[parity];

    float2 cur[E], xv[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const long i = base + (long)(t + e * T) * plane;
        cur[e] = active ? gload(sin, i) : make_float2(0.f, 0.f);
        xv[e] = (active && upd && TG.g1 > TG.g0) ? gload(a.X[TG.g0], i) : make_float2(0.f, 0.f);
    }
    load_twiddles<N>(tws, sets);
    __syncthreads();

    LineBins<E> lb;
    line_bins<E, T>(g, kx, ky, t, lb);
    float2 nw[E];
    if (upd) {
        float m[E];
        float2 acc[E];
#pragma unroll
        for (int e = 0; e < E; e++) acc[e] = make_float2(0.f, 0.f);
        // product groups: FFT_z of X_g times the group multiplier
        for (int gi = TG.g0; gi < TG.g1; gi++) {
            if (gi != TG.g0) {
#pragma unroll
                for (int e = 0; e < E; e++)
                    xv[e] = active ? gload(a.X[gi], base + (long)(t + e * T) * plane) : make_float2(0.f, 0.f);
            }
            fft_line<N, E, -1, N, true, (TWD * T <= 64)>(xv, L, t, tws);
            const TermRefDev &GM = a.gmul[gi];
            multiplier_line<E>(a.pool, GM.pres0, GM.npres, GM.mult_i, lb, m);
#pragma unroll
            for (int e = 0; e < E; e++) acc[e] = cadd(acc[e], apply_mult(m[e], GM.mult_i, xv[e]));
            line_sync<(TWD * T <= 64)>();
        }
        // single-field and constant terms (phi or a derived field of phi)
        for (int ti = 0; ti < TG.nterms; ti++) {
            const FusedTerm &TR = TG.term[ti];
            multiplier_line<E>(a.pool, TR.pres0, TR.npres, TR.mult_i, lb, m);
            float2 X[E];
            if (TR.src == SRC_CONST) {
#pragma unroll
                for (int e = 0; e < E; e++) X[e] = lb.zero[e] ? make_float2((float)g.n, 0.f) : make_float2(0.f, 0.f);
            } else if (TR.src == SRC_DERIVED) {
                const TermRefDev &DM = a.dmul[TR.idx];
                float dm[E];
                multiplier_line<E>(a.pool, DM.pres0, DM.npres, DM.mult_i, lb, dm);
#pragma unroll
                for (int e = 0; e < E; e++) X[e] = apply_mult(dm[e], DM.mult_i, cur[e]);
            } else {
#pragma unroll
                for (int e = 0; e < E; e++) X[e] = cur[e];
            }
#pragma unroll
            for (int e = 0; e < E; e++) acc[e] = cadd(acc[e], apply_mult(m[e], TR.mult_i, X[e]));
        }
        const UpdateArgs &u = a.u;
        const uint64_t step = a.step_base + step_rel;
#pragma unroll
        for (int e = 0; e < E; e++) {
            float2 s = cur[e];
            s.x += u.dt * acc[e].x;
            s.y += u.dt * acc[e].y;
            if (u.noisy && a.noise_mode == NOISE_SPECTRAL) {
                const float2 eta = active ? spectral_noise(g, kx, ky, t + e * T, (uint32_t)step, (uint32_t)(step >> 32),
                                                           a.stream, a.key0, a.key1)
                                          : make_float2(0.f, 0.f);
                float pn = u.noise_base;
                if (u.noise_q2n != 0) pn *= sqrtf(ipow(lb.q2[e], u.noise_q2n));
                if (u.noise_invq != 0)
                    pn *= sqrtf(ipow(lb.zero[e] ? 0.0f : __builtin_amdgcn_rsqf(lb.q2[e]), u.noise_invq));
                s.x += pn * eta.x;
                s.y += pn * eta.y;
            }
            if (u.nimp > 0) {
                float D = 1.0f;
                for (int i = 0; i < u.nimp; i++) {
                    const PresDev P = u.imp[i];
                    float th = P.c;
                    if (P.q2n != 0) th *= ipow(lb.q2[e], P.q2n);
                    if (P.invq != 0) th *= ipow(lb.axis ? 0.0f : __builtin_amdgcn_rsqf(lb.q2[e]), P.invq);
                    D -= u.dt * th;
                }
                const float rD = __builtin_amdgcn_rcpf(D);
                s.x *= rD;
                s.y *= rD;
            }
            nw[e] = s;
        }
        if (active) {
            float2 *__restrict__ sout = a.spec[parity ^ 1];
#pragma unroll
            for (int e = 0; e < E; e++) gstore(sout, base + (long)(t + e * T) * plane, nw[e]);
        }
    } else {
#pragma unroll
        for (int e = 0; e < E; e++) nw[e] = cur[e];
    }
    if (mode != 2) {
        // W_f = IFFT_z(mask_f * L_f * phi) on the lines the mask keeps
        const int ny = ky > g.sy / 2 ? ky - g.sy : ky;
        for (int f = 0; f < a.nreal; f++) {
            if (kx0 > a.kmax[f] || abs(ny) > a.kymax[f]) continue;  // uniform over the workgroup
            const int lz = a.p[f] > 0 ? g.sz / (a.p[f] + 1) : g.sz;
            const bool zero = !upd && a.zero[f] != 0;
            float2 v[E];
            const TermRefDev &LM = a.lmul[f];
            float lm[E];
            if (LM.npres > 0) multiplier_line<E>(a.pool, LM.pres0, LM.npres, LM.mult_i, lb, lm);
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int kz = t + e * T;
                const int nz = kz > g.sz / 2 ? kz - g.sz : kz;
                const bool keep = !zero && kx <= a.kmax[f] && abs(nz) <= lz;
                const float2 w = LM.npres > 0 ? apply_mult(lm[e], LM.mult_i, nw[e]) : nw[e];
                v[e] = keep ? w : make_float2(0.f, 0.f);
            }
            fft_line<N, E, +1, N, true, (TWD * T <= 64)>(v, L, t, tws);
            if (active && kx <= a.kmax[f]) {
                float2 *__restrict__ w = a.W[f];
#pragma unroll
                for (int e = 0; e < E; e++) gstore(w, base + (long)(t + e * T) * plane, v[e]);
            }
            line_sync<(TWD * T <= 64)>();
        }
    }
    if (slot) {
        __syncthreads();
        if (threadIdx.x == 0) atomicMax(&slot[1], (unsigned long long)wall_clock64());
    }
}

// derived fields psi_d = M_d phi from phi's spectrum [parity] (natural layout)
__global__ __launch_bounds__(256) void k_f3der(const FusedInst *__restrict__ insts, int parity)
{
    const FusedInst &a = insts[0];
    const Geom &g = a.g;
    const long h = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= g.nh) return;
    const int kx = (int)(h % g.nc1);
    const long r = h / g.nc1;
    const Bin b = make_bin(g, kx, (int)(r % g.sy), (int)(r / g.sy));
    const float2 X = gload(a.spec[parity], h);
    for (int d = 0; d < a.nder; d++) {
        const TermRefDev &M = a.dmul[d];
        gstore(a.dout[d], h, apply_mult(multiplier(a.pool, M.pres0, M.npres, M.mult_i, b), M.mult_i, X));
    }
}

// ------------------------------------------------------------ dispatch

#define FUSED_ROW_CASES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)
#define FUSED_COL_CASES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)
#define FUSED_LPB(X) X(1) X(2) X(4) X(8) X(16)
#define FUSED_TWD(X) X(1) X(2) X(4) X(8) X(16)

bool fused_supported(int sx, int sy, int sz)
{
    if (sz > 1) return sz >= 8 && sz <= 1024 && sx >= 16 && sx <= 2048 && sy >= 8 && sy <= 4096;  // 3D
    return sx >= 16 && sx <= 4096 && sy >= 8 && sy <= 4096;
}

// wide row kernels (more than 3 real fields) for the sizes such systems use;
// narrow ones (at most 2: fewer registers, one more wave per SIMD) for the
// sizes of large instances (others run the 3-field kernel)
#define FUSED_WIDE_NC(NC) ((NC) >= 32 && (NC) <= 1024)
#define FUSED_NARROW_NC(NC) ((NC) >= 32 && (NC) <= 1024)

int fused_row_width(int nreal) { return nreal <= 2 ? 2 : (nreal == 3 ? 3 : FUSED_MAX_REAL); }

bool fused_shapes_exist(int sx, int sy, int sz, int nreal)
{
    const int nc = sx / 2;
    if (sz > 1) return nreal <= 3 || (nc >= 32 && nc <= 1024);
    return !fused_row_shapes(nc, sy, fused_row_width(nreal)).empty() && !fused_col_shapes(sy).empty();
}

template <int NC, int MAXR>
static std::vector<int> row_shapes(int ny)
{
    std::vector<int> v;
#define X(LPB) \
    if (FRowCfg<NC, LPB, MAXR>::valid() && LPB <= ny) v.push_back(LPB);
    FUSED_LPB(X)
#undef X
    return v;
}

std::vector<int> fused_row_shapes(int nc, int ny, int width)
{
    switch (nc) {
#define X(NC)                                                                                            \
    case NC:                                                                                             \
        if (width > 3) return FUSED_WIDE_NC(NC) ? row_shapes<NC, FUSED_MAX_REAL>(ny) : std::vector<int>(); \
        if (width == 2 && FUSED_NARROW_NC(NC)) return row_shapes<NC, 2>(ny);                              \
        return row_shapes<NC, 3>(ny);
        FUSED_ROW_CASES(X)
#undef X
    default:
        return {};
    }
}

template <int N>
static std::vector<int> col_shapes()
{
    std::vector<int> v;
#define X(TWD) \
    if (FColCfg<N, TWD>::valid()) v.push_back(TWD);
    FUSED_TWD(X)
#undef X
    return v;
}

std::vector<int> fused_col_shapes(int ny)
{
    switch (ny) {
#define X(N) \
    case N:  \
        return col_shapes<N>();
        FUSED_COL_CASES(X)
#undef X
    default:
        return {};
    }
}

template <int NC, int MAXR>
static hipError_t frows_n(int lpb, const FusedInst *inst, const FusedBlock *blocks, int nblocks, uint32_t step_rel,
                          hipStream_t s, const float2 *sets)
{
#define X(LPB)                                                                                                   \
    if constexpr (FRowCfg<NC, LPB, MAXR>::valid()) {                                                            \
        if (lpb == LPB) {                                                                                        \
            hipLaunchKernelGGL((k_frows<NC, LPB, MAXR>), dim3(nblocks), dim3(FRowCfg<NC, LPB, MAXR>::THREADS), 0, \
                               s, inst, blocks, step_rel, sets);                                                 \
            return hipGetLastError();                                                                            \
        }                                                                                                        \
    }
    FUSED_LPB(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t fused_rows_launch(int nc, int lpb, int width, const FusedInst *inst, const FusedBlock *blocks,
                             int nblocks, uint32_t step_rel, hipStream_t s)
{
    const float2 *sets = twiddle_sets(s);
    if (!sets) return hipErrorOutOfMemory;
    switch (nc) {
#define X(NC)                                                                                      \
    case NC:                                                                                       \
        if (width > 3) {                                                                           \
            if constexpr (FUSED_WIDE_NC(NC))                                                       \
                return frows_n<NC, FUSED_MAX_REAL>(lpb, inst, blocks, nblocks, step_rel, s, sets); \
            return hipErrorInvalidValue;                                                           \
        }                                                                                          \
        if constexpr (FUSED_NARROW_NC(NC))                                                         \
            if (width == 2) return frows_n<NC, 2>(lpb, inst, blocks, nblocks, step_rel, s, sets);  \
        return frows_n<NC, 3>(lpb, inst, blocks, nblocks, step_rel, s, sets);
        FUSED_ROW_CASES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

template <int N>
static hipError_t fcols_n(int twd, int stored, const FusedInst *inst, const FusedBlock *blocks, int nblocks,
                          int parity, uint32_t step_rel, int mode, hipStream_t s, const float2 *sets,
                          unsigned long long *ts)
{
#define L(CONS, XDYN, TWD)                                                                                    \
    hipLaunchKernelGGL((k_fcols<N, TWD, CONS, XDYN>), dim3(nblocks), dim3(FColCfg<N, TWD>::THREADS), 0, s, inst, \
                       blocks, parity, step_rel, mode, sets, ts)
#define X(TWD)                                   \
    if constexpr (FColCfg<N, TWD>::valid()) {   \
        if (twd == TWD) {                        \
            if (stored == 2)                     \
                L(true, true, TWD);              \
            else if (stored == 1)                \
                L(true, false, TWD);             \
            else                                 \
                L(false, false, TWD);            \
            return hipGetLastError();            \
        }                                        \
    }
    FUSED_TWD(X)
#undef X
#undef L
    return hipErrorInvalidValue;
}

hipError_t fused_cols_launch(int ny, int twd, int stored, const FusedInst *inst, const FusedBlock *blocks,
                             int nblocks, int parity, uint32_t step_rel, int mode, hipStream_t s,
                             unsigned long long *ts)
{
    const float2 *sets = twiddle_sets(s);
    if (!sets) return hipErrorOutOfMemory;
    switch (ny) {
#define X(N) \
    case N:  \
        return fcols_n<N>(twd, stored, inst, blocks, nblocks, parity, step_rel, mode, s, sets, ts);
        FUSED_COL_CASES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

bool fused_small_class(int nc, int ny, int width)
{
    return width <= 3 && ny == 2 * nc && (ny == 64 || ny == 128);
}

template <int NC, int N>
static hipError_t fsmall_n(int stored, const FusedInst *inst, const int *ids, int nids, int n0, int k, int rel0,
                           hipStream_t s, const float2 *sets, unsigned long long *ts)
{
#define L(CONS, XDYN)                                                                                         \
    static_assert(FSmallCfg<NC, N, CONS, XDYN>::valid(), "small-instance tile");                           \
    hipLaunchKernelGGL((k_fsmall<NC, N, CONS, XDYN>), dim3(nids), dim3(FSmallCfg<NC, N, CONS, XDYN>::THREADS), \
                       0, s, inst, ids, n0, k, rel0, sets, ts)
    if (stored == 2) {
        L(true, true);
    } else if (stored == 1) {
        L(true, false);
    } else {
        L(false, false);
    }
#undef L
    return hipGetLastError();
}

hipError_t fused_small_launch(int nc, int ny, int stored, const FusedInst *inst, const int *ids, int nids, int n0,
                              int k, int rel0, hipStream_t s, unsigned long long *ts)
{
    const float2 *sets = twiddle_sets(s);
    if (!sets) return hipErrorOutOfMemory;
    if (nids <= 0 || k <= 0) return hipSuccess;
    if (nc == 32 && ny == 64) return fsmall_n<32, 64>(stored, inst, ids, nids, n0, k, rel0, s, sets, ts);
    if (nc == 64 && ny == 128) return fsmall_n<64, 128>(stored, inst, ids, nids, n0, k, rel0, s, sets, ts);
    return hipErrorInvalidValue;
}

hipError_t fused_transpose_in_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_ftrans_in, dim3(nblocks), dim3(256), 0, s, inst, blocks);
    return hipGetLastError();
}

hipError_t fused_settle_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int fin, hipStream_t s)
{
    hipLaunchKernelGGL(k_fsettle, dim3(nblocks), dim3(256), 0, s, inst, blocks, fin);
    return hipGetLastError();
}

hipError_t fused_elem_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int parity,
                             uint32_t step_rel, int nsteps, hipStream_t s)
{
    hipLaunchKernelGGL(k_felem, dim3(nblocks), dim3(FUSED_ELEM_BLOCK), 0, s, inst, blocks, parity, step_rel,
                       nsteps);
    return hipGetLastError();
}

// ---- 3D

template <int NC, int MAXR>
static hipError_t f3rows_n(const FusedInst *inst, long nrows, hipStream_t s, const float2 *sets)
{
    constexpr int T = FRowCfg<NC, 1, MAXR>::T;
    constexpr int LPB = T >= 256 ? 1 : 256 / T;  // 256 threads a workgroup
    static_assert(FRowCfg<NC, LPB, MAXR>::valid(), "3D row tile");
    hipLaunchKernelGGL((k_f3rows<NC, LPB, MAXR>), dim3((unsigned)((nrows + LPB - 1) / LPB)),
                       dim3(FRowCfg<NC, LPB, MAXR>::THREADS), 0, s, inst, sets);
    return hipGetLastError();
}

hipError_t fused3d_rows_launch(int nc, int maxr, const FusedInst *inst, long nrows, hipStream_t s)
{
    const float2 *sets = twiddle_sets(s);
    if (!sets) return hipErrorOutOfMemory;
    switch (nc) {
#define X(NC)                                                                 \
    case NC:                                                                  \
        if (maxr > 3) {                                                       \
            if constexpr (FUSED_WIDE_NC(NC))                                  \
                return f3rows_n<NC, FUSED_MAX_REAL>(inst, nrows, s, sets);    \
            return hipErrorInvalidValue;                                      \
        }                                                                     \
        return f3rows_n<NC, 3>(inst, nrows, s, sets);
        X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

#define F3Z_CASES(X) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024)
#define F3Z_TWD(X) X(4) X(8) X(16) X(32) X(64)

template <int N>
static std::vector<int> z_shapes()
{
    std::vector<int> v;
#define X(TWD) \
    if (F3zCfg<N, TWD>::valid()) v.push_back(TWD);
    F3Z_TWD(X)
#undef X
    return v;
}

std::vector<int> fused3d_z_shapes(int nz)
{
    switch (nz) {
#define X(N) \
    case N:  \
        return z_shapes<N>();
        F3Z_CASES(X)
#undef X
    default:
        return {};
    }
}

template <int N>
static hipError_t f3z_n(int twd, const FusedInst *inst, int nc1, int sy, int parity, uint32_t step_rel, int mode,
                        hipStream_t s, const float2 *sets, unsigned long long *ts)
{
#define X(TWD)                                                                                           \
    if constexpr (F3zCfg<N, TWD>::valid()) {                                                            \
        if (twd == TWD) {                                                                                \
            const long total = (long)((nc1 + TWD - 1) / TWD) * sy;                                       \
            hipLaunchKernelGGL((k_f3z<N, TWD>), dim3((unsigned)xcd_blocks(total)),                      \
                               dim3(F3zCfg<N, TWD>::THREADS), 0, s, inst, parity, step_rel, mode, sets, ts); \
            return hipGetLastError();                                                                    \
        }                                                                                                \
    }
    F3Z_TWD(X)
#undef X
    return hipErrorInvalidValue;
}

hipError_t fused3d_z_launch(int nz, int twd, const FusedInst *inst, int nc1, int sy, int parity, uint32_t step_rel,
                            int mode, hipStream_t s, unsigned long long *ts)
{
    const float2 *sets = twiddle_sets(s);
    if (!sets) return hipErrorOutOfMemory;
    switch (nz) {
#define X(N) \
    case N:  \
        return f3z_n<N>(twd, inst, nc1, sy, parity, step_rel, mode, s, sets, ts);
        F3Z_CASES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t fused3d_derived_launch(const FusedInst *inst, long nh, int parity, hipStream_t s)
{
    hipLaunchKernelGGL(k_f3der, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, inst, parity);
    return hipGetLastError();
}

}  // namespace cupss

namespace cupss {

__global__ void k_ts_reset(unsigned long long *ts, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n * FUSED_TS_SLOTS) {
        ts[(size_t)i * FUSED_TS_STRIDE] = ~0ull;
        ts[(size_t)i * FUSED_TS_STRIDE + 1] = 0ull;
    }
}

hipError_t fused_ts_reset_launch(unsigned long long *ts, int n, hipStream_t s)
{
    const int m = n * FUSED_TS_SLOTS;
    hipLaunchKernelGGL(k_ts_reset, dim3((m + 255) / 256), dim3(256), 0, s, ts, n);
    return hipGetLastError();
}

}  // namespace cupss
