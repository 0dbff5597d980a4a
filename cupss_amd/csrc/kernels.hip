// kernels.hip -- the per-pass FFT kernels of the plain (per-solve) path and
// the fused semi-implicit update.
//
// Layout: real fields are float[sz][sy][sx] (x fastest, the reference's
// idx = k*sx*sy + j*sx + i); spectra are the Hermitian half
// float2[sz][sy][sx/2+1].  The C2C passes along y and z work in place on
// the half spectrum.
//
//   forward (R2C):  rows_r2c (x, with a fused source: a real field, the
//                   product of several real fields, or Philox noise)
//                   -> cols y -> cols z
//   inverse (C2R):  cols z (first pass carries the dealias mask) -> cols y
//                   -> rows_c2r (x, 1/N scaling; mask here in 1D)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>
#include <cmath>

#include "fft_core.hpp"
#include "internal.hpp"
#include "philox.hpp"
#include "update_math.hpp"

namespace cupss {

// ------------------------------------------------------------ twiddles

static std::mutex g_tw_mutex;
static std::vector<float2 *> g_tw;    // per device
static std::vector<float2 *> g_tws;   // per device

void twiddle_table(float2 *h)
{
    for (int j = 0; j < TW_N; j++) {
        double a = -2.0 * M_PI * (double)j / (double)TW_N;
        h[j] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
}

const float2 *twiddles(hipStream_t s)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    if ((int)g_tw.size() <= dev) g_tw.resize(dev + 1, nullptr);
    if (!g_tw[dev]) {
        std::vector<float2> h(TW_N);
        twiddle_table(h.data());
        float2 *d = nullptr;
        if (hipMalloc(&d, sizeof(float2) * TW_N) != hipSuccess) return nullptr;
        if (hipMemcpy(d, h.data(), sizeof(float2) * TW_N, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        g_tw[dev] = d;
    }
    (void)s;
    return g_tw[dev];
}

const float2 *twiddle_sets(hipStream_t s)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lock(g_tw_mutex);
    if ((int)g_tws.size() <= dev) g_tws.resize(dev + 1, nullptr);
    if (!g_tws[dev]) {
        std::vector<float2> h(TW_SETS_LEN, make_float2(0.f, 0.f));
        auto w = [](long num, long den) {
            const double a = -2.0 * M_PI * (double)num / (double)den;
            return make_float2((float)std::cos(a), (float)std::sin(a));
        };
        for (int M = 1; M <= TW_N; M *= 2)
            for (int j = 0; j < M; j++) h[M + j] = w(j, M);
        // stage-ordered tables (fft_core.hpp, TW_STAGE) of every (N, E)
        for (int E = 2; E <= 16; E *= 2)
            for (int N = 2; N <= TW_N; N *= 2) {
                float2 *st = h.data() + tw_stage_offset(N, E);
                for (int NS = tw_stage_radix(N, E, 1); NS < N; NS *= tw_stage_radix(N, E, NS)) {
                    const int R = tw_stage_radix(N, E, NS);
                    float2 *sb = st + tw_stage_base(N, E, NS);
                    for (int r = 1; r < R; r++)
                        for (int k = 0; k < NS; k++) sb[(r - 1) * NS + k] = w((long)r * k, (long)NS * R);
                }
            }
        float2 *d = nullptr;
        if (hipMalloc(&d, sizeof(float2) * TW_SETS_LEN) != hipSuccess) return nullptr;
        if (hipMemcpy(d, h.data(), sizeof(float2) * TW_SETS_LEN, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        g_tws[dev] = d;
    }
    (void)s;
    return g_tws[dev];
}

// ------------------------------------------------------------ sources

struct SrcArgs {
    const float *f[MAX_FACTORS];
    int nf;
    uint32_t key0, key1;
    uint32_t step_lo, step_hi, stream;
};

// z[m] = x[2m] + i x[2m+1] of row `row` (real row length 2*nc)
template <int KIND>
__device__ __forceinline__ float2 load_src(const SrcArgs &a, long row, int m, int nc)
{
    const long idx = row * (2L * nc) + 2L * m;
    if constexpr (KIND == LOAD_REAL) {
        return *reinterpret_cast<const float2 *>(a.f[0] + idx);
    } else if constexpr (KIND == LOAD_PRODUCT) {
        float2 r = make_float2(1.0f, 1.0f);
        for (int c = 0; c < a.nf; c++) {
            float2 v = *reinterpret_cast<const float2 *>(a.f[c] + idx);
            r.x *= v.x;
            r.y *= v.y;
        }
        return r;
    } else {
        // two consecutive cells share one Philox block (idx even)
        float2 z = philox_normal_pair((uint32_t)idx, a.step_lo, a.step_hi, a.stream, a.key0, a.key1);
        return z;
    }
}

// ------------------------------------------------------------ rows R2C

// rows per workgroup: one wavefront at least
template <int NC>
struct RowCfg {
    static constexpr int E = row_elems(NC);
    static constexpr int T = NC / E;
    static constexpr int LPB = T >= 64 ? 1 : 64 / T;
    static constexpr int THREADS = LPB * T;
    static constexpr int LL = lds_len(NC);
};

template <int NC, int KIND>
__global__ __launch_bounds__(256) void k_rows_r2c(SrcArgs src, float2 *__restrict__ out, long nrows,
                                                  const float2 *__restrict__ tw)
{
    using C = RowCfg<NC>;
    constexpr int E = C::E, T = C::T;
    __shared__ float2 lds[C::LPB * C::LL];
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    const long row = (long)blockIdx.x * C::LPB + l;
    const bool active = row < nrows;
    float2 *L = lds + l * C::LL;
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = active ? load_src<KIND>(src, row, t + e * T, NC) : make_float2(0.f, 0.f);
    fft_line<NC, E, -1, TW_N, true, (T <= 64)>(v, L, t, tw);
    if (!active) return;
    float2 *o = out + row * (NC + 1);
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int k = t + e * T;
        const float2 A = v[e];
        const float2 B = cconj(L[lpad((NC - k) & (NC - 1))]);
        const float2 xe = cscale(cadd(A, B), 0.5f);
        const float2 xo = cscale(mul_i<-1>(csub(A, B)), 0.5f);  // (A-B)/(2i)
        float2 X;
        if (k == 0) {
            X = make_float2(xe.x + xo.x, 0.0f);
        } else {
            X = cadd(xe, cmul(tw_lookup<-1>(tw, k, 2 * NC), xo));
        }
        o[k] = X;
    }
    if (t == 0) {
        const float2 z0 = L[lpad(0)];
        o[NC] = make_float2(z0.x - z0.y, 0.0f);
    }
}

// ------------------------------------------------------------ rows C2R

template <int NC, bool MASK>
__global__ __launch_bounds__(256) void k_rows_c2r(const float2 *__restrict__ in, float *__restrict__ out,
                                                  long nrows, float scale, int cut,
                                                  const float2 *__restrict__ tw)
{
    using C = RowCfg<NC>;
    constexpr int E = C::E, T = C::T;
    __shared__ float2 lds[C::LPB * C::LL];
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    const long row = (long)blockIdx.x * C::LPB + l;
    const bool active = row < nrows;
    const float2 *x = in + row * (NC + 1);
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int k = t + e * T;
        float2 X = make_float2(0.f, 0.f), Y = make_float2(0.f, 0.f);
        if (active) {
            X = x[k];
            Y = x[NC - k];
            if (k == 0) {  // self-conjugate bins: keep the real part (projection)
                X.y = 0.f;
                Y.y = 0.f;
            }
            if (MASK) {
                if (k > cut) X = make_float2(0.f, 0.f);
                if (NC - k > cut) Y = make_float2(0.f, 0.f);
            }
        }
        const float2 Yc = cconj(Y);
        const float2 xe = cadd(X, Yc);
        const float2 xo = cmul(csub(X, Yc), tw_lookup<+1>(tw, k, 2 * NC));
        v[e] = cadd(xe, mul_i<+1>(xo));
    }
    fft_line<NC, E, +1, TW_N, false, (T <= 64)>(v, lds + l * C::LL, t, tw);
    if (!active) return;
    float2 *o = reinterpret_cast<float2 *>(out + row * (2L * NC));
#pragma unroll
    for (int e = 0; e < E; e++) o[t + e * T] = cscale(v[e], scale);
}

// ------------------------------------------------------------ columns

// Column tile: TWD adjacent lines of one workgroup (one line per TWD-th
// thread, so that a row segment of TWD values is read with consecutive
// lanes).  Valid tiles keep 64 <= TWD*T <= 512 threads.
template <int N>
struct ColCfg {
    static constexpr int E = elems_for(N);
    static constexpr int T = N / E;
    static constexpr int LL = lds_len(N);
    static constexpr bool valid(int twd) { return twd * T >= 64 && twd * T <= 512; }
};

// C2C along a strided axis.  Line `col` of batch `b` is
//   p[b*bstride + col + y*stride], y = 0..N-1, col = 0..ncols-1.
// Workgroups are placed XCD-aware over the (batch, tile) items.
// MASK: dealias mask of order p on the half spectrum (first inverse pass).
template <int N, int TWD, int DIR, bool MASK>
__global__ __launch_bounds__(512) void k_cols(const float2 *__restrict__ in, float2 *__restrict__ out, long ncols,
                                              long stride, long bstride, long nbatch, Geom g, int p,
                                              const float2 *__restrict__ tw)
{
    using C = ColCfg<N>;
    constexpr int E = C::E, T = C::T;
    __shared__ float2 lds[TWD * C::LL];
    const long ntiles = (ncols + TWD - 1) / TWD;
    const long item = xcd_item(blockIdx.x, gridDim.x);
    if (item >= ntiles * nbatch) return;   // whole workgroup idle
    const long b = item / ntiles;
    const long tile = item - b * ntiles;
    const int c = threadIdx.x % TWD, t = threadIdx.x / TWD;
    const long col = tile * TWD + c;
    const bool active = col < ncols;
    const float2 *src = in + b * bstride + col;
    float2 v[E];
#pragma unroll
    for (int e = 0; e < E; e++) v[e] = active ? src[(long)(t + e * T) * stride] : make_float2(0.f, 0.f);
    if constexpr (MASK) {
        // bin indices: 2D lines run along y (col = kx); the 3D first inverse
        // pass runs along z (col = ky*nc1 + kx)
        const int kx = (int)(col % g.nc1);
        const int ky0 = (int)(col / g.nc1);
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int y = t + e * T;
            const bool keep = g.sz > 1 ? keep_mode(g, p, kx, ky0, y) : keep_mode(g, p, kx, y, 0);
            if (!keep) v[e] = make_float2(0.f, 0.f);
        }
    }
    fft_line<N, E, DIR, TW_N, false, (TWD * T <= 64)>(v, lds + c * C::LL, t, tw);
    if (!active) return;
    float2 *dst = out + b * bstride + col;
#pragma unroll
    for (int e = 0; e < E; e++) dst[(long)(t + e * T) * stride] = v[e];
}

// ------------------------------------------------------------ dispatch

#define CUPSS_POW2_CASES(X) \
    X(1) X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048)

static hipError_t rows_r2c(const Geom &g, const SrcArgs &a, int kind, float2 *out, hipStream_t s,
                           const float2 *tw)
{
    const long nrows = (long)g.sy * g.sz;
    switch (g.nc) {
#define X(NC)                                                                                        \
    case NC: {                                                                                       \
        dim3 grid((unsigned)((nrows + RowCfg<NC>::LPB - 1) / RowCfg<NC>::LPB));                      \
        dim3 block(RowCfg<NC>::THREADS);                                                             \
        if (kind == LOAD_REAL)                                                                       \
            hipLaunchKernelGGL((k_rows_r2c<NC, LOAD_REAL>), grid, block, 0, s, a, out, nrows, tw);   \
        else if (kind == LOAD_PRODUCT)                                                               \
            hipLaunchKernelGGL((k_rows_r2c<NC, LOAD_PRODUCT>), grid, block, 0, s, a, out, nrows, tw); \
        else                                                                                         \
            hipLaunchKernelGGL((k_rows_r2c<NC, LOAD_NOISE>), grid, block, 0, s, a, out, nrows, tw);  \
        return hipGetLastError();                                                                    \
    }
        CUPSS_POW2_CASES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

static hipError_t rows_c2r(const Geom &g, const float2 *in, float *out, float scale, int mask_p, hipStream_t s,
                           const float2 *tw)
{
    const long nrows = (long)g.sy * g.sz;
    const int cut = mask_p > 0 ? g.sx / (mask_p + 1) : 1 << 30;
    switch (g.nc) {
#define X(NC)                                                                                              \
    case NC: {                                                                                             \
        dim3 grid((unsigned)((nrows + RowCfg<NC>::LPB - 1) / RowCfg<NC>::LPB));                            \
        dim3 block(RowCfg<NC>::THREADS);                                                                   \
        if (mask_p > 0)                                                                                    \
            hipLaunchKernelGGL((k_rows_c2r<NC, true>), grid, block, 0, s, in, out, nrows, scale, cut, tw); \
        else                                                                                               \
            hipLaunchKernelGGL((k_rows_c2r<NC, false>), grid, block, 0, s, in, out, nrows, scale, cut, tw); \
        return hipGetLastError();                                                                          \
    }
        CUPSS_POW2_CASES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

template <int N, int TWD>
static void launch_cols(int dir, bool mask, int p, const float2 *in, float2 *out, long ncols, long stride,
                        long bstride, long nbatch, const Geom &g, hipStream_t s, const float2 *tw)
{
    const long total = (ncols + TWD - 1) / TWD * nbatch;
    dim3 grid((unsigned)xcd_blocks(total)), block(TWD * ColCfg<N>::T);
    if (dir < 0)
        hipLaunchKernelGGL((k_cols<N, TWD, -1, false>), grid, block, 0, s, in, out, ncols, stride, bstride, nbatch,
                           g, p, tw);
    else if (mask)
        hipLaunchKernelGGL((k_cols<N, TWD, 1, true>), grid, block, 0, s, in, out, ncols, stride, bstride, nbatch, g,
                           p, tw);
    else
        hipLaunchKernelGGL((k_cols<N, TWD, 1, false>), grid, block, 0, s, in, out, ncols, stride, bstride, nbatch,
                           g, p, tw);
}

// widest valid tile that still gives >= 1024 workgroups, else the narrowest
template <int N>
static void cols_n(int dir, bool mask, int p, const float2 *in, float2 *out, long ncols, long stride, long bstride,
                   long nbatch, const Geom &g, hipStream_t s, const float2 *tw)
{
    constexpr int widths[5] = {64, 16, 4, 2, 1};
    int pick = 0;
    for (int w : widths) {
        if (!ColCfg<N>::valid(w)) continue;
        pick = w;
        if ((ncols + w - 1) / w * nbatch >= 1024) break;
    }
    switch (pick) {
    case 64:
        if constexpr (ColCfg<N>::valid(64)) launch_cols<N, 64>(dir, mask, p, in, out, ncols, stride, bstride, nbatch, g, s, tw);
        break;
    case 16:
        if constexpr (ColCfg<N>::valid(16)) launch_cols<N, 16>(dir, mask, p, in, out, ncols, stride, bstride, nbatch, g, s, tw);
        break;
    case 4:
        if constexpr (ColCfg<N>::valid(4)) launch_cols<N, 4>(dir, mask, p, in, out, ncols, stride, bstride, nbatch, g, s, tw);
        break;
    case 2:
        if constexpr (ColCfg<N>::valid(2)) launch_cols<N, 2>(dir, mask, p, in, out, ncols, stride, bstride, nbatch, g, s, tw);
        break;
    case 1:
        if constexpr (ColCfg<N>::valid(1)) launch_cols<N, 1>(dir, mask, p, in, out, ncols, stride, bstride, nbatch, g, s, tw);
        break;
    }
}

#define CUPSS_COL_CASES(X) \
    X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

static hipError_t cols(const Geom &g, int n, int dir, bool mask, int p, const float2 *in, float2 *out, long ncols,
                       long stride, long bstride, long nbatch, hipStream_t s, const float2 *tw)
{
    switch (n) {
#define X(N)                                                                              \
    case N:                                                                               \
        cols_n<N>(dir, mask, p, in, out, ncols, stride, bstride, nbatch, g, s, tw);      \
        return hipGetLastError();
        CUPSS_COL_CASES(X)
#undef X
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t r2c(const Geom &g, const R2CSource &src, float2 *out, hipStream_t s)
{
    const float2 *tw = twiddles(s);
    if (!tw) return hipErrorOutOfMemory;
    SrcArgs a{};
    for (int i = 0; i < MAX_FACTORS; i++) a.f[i] = src.f[i];
    a.nf = src.nf;
    a.key0 = (uint32_t)src.seed;
    a.key1 = (uint32_t)(src.seed >> 32);
    a.step_lo = (uint32_t)src.step;
    a.step_hi = (uint32_t)(src.step >> 32);
    a.stream = src.stream;
    hipError_t e = rows_r2c(g, a, src.kind, out, s, tw);
    if (e != hipSuccess) return e;
    if (g.sy > 1) {
        e = cols(g, g.sy, -1, false, 0, out, out, g.nc1, g.nc1, (long)g.nc1 * g.sy, g.sz, s, tw);
        if (e != hipSuccess) return e;
    }
    if (g.sz > 1) {
        const long plane = (long)g.nc1 * g.sy;
        e = cols(g, g.sz, -1, false, 0, out, out, plane, plane, 0, 1, s, tw);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t c2r(const Geom &g, const float2 *in, int mask_p, float scale, float *out, float2 *scratch,
               hipStream_t s)
{
    const float2 *tw = twiddles(s);
    if (!tw) return hipErrorOutOfMemory;
    return c2r_tw(g, in, mask_p, scale, out, scratch, s, tw);
}

hipError_t c2r_tw(const Geom &g, const float2 *in, int mask_p, float scale, float *out, float2 *scratch,
                  hipStream_t s, const float2 *tw)
{
    hipError_t e;
    const long plane = (long)g.nc1 * g.sy;
    const float2 *cur = in;
    bool masked = false;
    if (g.sz > 1) {
        e = cols(g, g.sz, +1, mask_p > 0, mask_p, cur, scratch, plane, plane, 0, 1, s, tw);
        if (e != hipSuccess) return e;
        cur = scratch;
        masked = true;
    }
    if (g.sy > 1) {
        const bool m = !masked && mask_p > 0;
        e = cols(g, g.sy, +1, m, mask_p, cur, scratch, g.nc1, g.nc1, plane, g.sz, s, tw);
        if (e != hipSuccess) return e;
        cur = scratch;
        masked = masked || m;
    }
    return rows_c2r(g, cur, out, scale, masked ? 0 : mask_p, s, tw);
}

hipError_t lines_y(const Geom &g, int dir, const float2 *in, float2 *out, long ncols, hipStream_t s)
{
    const float2 *tw = twiddles(s);
    if (!tw) return hipErrorOutOfMemory;
    return cols(g, g.sy, dir, false, 0, in, out, ncols, g.nc1, (long)g.nc1 * g.sy, g.sz, s, tw);
}

// ------------------------------------------------------------ update

// Semi-implicit Euler update of one field on the half spectrum (one
// workgroup per spectral row, lanes over kx):
//   dynamic:    s = (cur + dt*sum_t T_t + P_eta*eta) / D(q)          (setDynamic_k)
//   constraint: s = (sum_t T_t [+ dt^-1/2 P_eta eta]) / L(q), q != 0  (setNotDynamic_k)
__global__ __launch_bounds__(256) void k_update(UpdateArgs a)
{
    const Geom &g = a.g;
    const long row = blockIdx.x;
    const int ky = (int)(row % g.sy);
    const int kz = (int)(row / g.sy);
    for (int kx = threadIdx.x; kx < g.nc1; kx += blockDim.x) {
        const Bin b = make_bin(g, kx, ky, kz);
        float2 acc = make_float2(0.f, 0.f);
        for (int t = 0; t < a.nterms; t++) acc = cadd(acc, term_value(a, t, b, g));
        const float2 eta = a.noisy ? a.noise[b.h] : make_float2(0.f, 0.f);
        a.out[b.h] = finish_update(a, b, a.cur[b.h], acc, eta);
    }
}

// eta_hat drawn directly in spectral space (see spectral_noise)
__global__ __launch_bounds__(256) void k_spectral_noise(float2 *__restrict__ eta, Geom g, uint32_t slo, uint32_t shi,
                                                        uint32_t stream, uint32_t k0, uint32_t k1)
{
    const long row = blockIdx.x;
    const int ky = (int)(row % g.sy), kz = (int)(row / g.sy);
    for (int kx = threadIdx.x; kx < g.nc1; kx += blockDim.x)
        eta[row * g.nc1 + kx] = spectral_noise(g, kx, ky, kz, slo, shi, stream, k0, k1);
}

hipError_t spectral_noise_fill(const Geom &g, float2 *eta, uint64_t seed, uint64_t step, uint32_t stream,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_spectral_noise, dim3((unsigned)((long)g.sy * g.sz)), dim3(256), 0, s, eta, g,
                       (uint32_t)step, (uint32_t)(step >> 32), stream, (uint32_t)seed, (uint32_t)(seed >> 32));
    return hipGetLastError();
}

// Hermitian projection of the kx = 0 and kx = sx/2 planes of a half
// spectrum: bin (kx, ky, kz) and its partner (kx, -ky, -kz) both become
// their Hermitian mean (a self-partnered bin keeps its real part).  The
// interior planes stand for their partners implicitly; on these two planes
// both partners are stored, and the reference's IFFT / real part / FFT round
// trip of every field every step (src/field.cpp:59-89) keeps them conjugate.
// Without it the anti-Hermitian part -- invisible to C2R, fed by rounding in
// the transforms -- grows unchecked in linearly unstable bands (e.g. the
// spinodal band of Cahn-Hilliard, x12 per 1000 steps at dt = 0.01) until it
// leaks into the real fields.  One thread per bin of the two planes.
__global__ __launch_bounds__(256) void k_hermitian_planes(float2 *__restrict__ h, Geom g)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long per = (long)g.sy * g.sz;
    if (i >= 2 * per) return;
    const int kx = i < per ? 0 : g.nc;
    const long r = i % per;
    const int ky = (int)(r % g.sy), kz = (int)(r / g.sy);
    const int py = (g.sy - ky) % g.sy, pz = (g.sz - kz) % g.sz;
    if (pz < kz || (pz == kz && py < ky)) return;  // the partner's thread writes both
    const long a = ((long)kz * g.sy + ky) * g.nc1 + kx, b = ((long)pz * g.sy + py) * g.nc1 + kx;
    const float2 u = h[a], v = h[b];
    const float2 m = make_float2(0.5f * (u.x + v.x), 0.5f * (u.y - v.y));
    h[a] = m;
    if (b != a) h[b] = make_float2(m.x, -m.y);
}

hipError_t hermitian_planes(const Geom &g, float2 *spec, hipStream_t s)
{
    const long n = 2L * g.sy * g.sz;
    hipLaunchKernelGGL(k_hermitian_planes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, spec, g);
    return hipGetLastError();
}

hipError_t update(const UpdateArgs &a, hipStream_t s)
{
    const unsigned rows = (unsigned)((long)a.g.sy * a.g.sz);
    const unsigned threads = a.g.nc1 >= 256 ? 256 : (a.g.nc1 >= 128 ? 128 : 64);
    hipLaunchKernelGGL(k_update, dim3(rows), dim3(threads), 0, s, a);
    if (hipError_t e = hipGetLastError()) return e;
    return hermitian_planes(a.g, a.out, s);
}

}  // namespace cupss
