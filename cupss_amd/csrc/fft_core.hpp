// fft_core.hpp -- device building blocks for the LDS-resident Stockham FFTs.
//
// A line of N complex values is held by T = N/E threads of one workgroup,
// E values per thread in the "strided" distribution: thread t owns
// x[t + e*T], e = 0..E-1.  Each Stockham stage of radix R (R <= E, R | E)
// does E/R radix-R butterflies per thread in registers, writes the outputs
// to the line's LDS region at their Stockham positions and reads back the
// strided distribution for the next stage.  After the last stage the
// transform is in natural order both in LDS and in the registers
// (v[e] = X[t + e*T]).  Radix schedule: stages of radix E while the
// remaining length allows it, then one stage with the remainder.
//
// Twiddles come from a table of exp(-2*pi*i*j/TW_N), j < TW_N, computed in
// double precision on the host and rounded to float (twiddle_table()).
//
// Sign convention (the reference's cuFFT/FFTW C2C, unnormalised):
// DIR = -1 forward, exp(-2*pi*i*k*n/N); DIR = +1 inverse.
#pragma once
#include <hip/hip_runtime.h>

namespace cupss {

constexpr int TW_LOG = 13;
constexpr int TW_N = 1 << TW_LOG;  // largest supported line length (complex points)

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by i^DIR: DIR = +1 -> i, DIR = -1 -> -i
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a)
{
    return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// Global-memory access through a pointer read from a descriptor.  The
// compiler cannot tell the address space of such a pointer and emits flat
// instructions, which also count against the LDS wait counter (every LDS
// wait of a transform would then wait for outstanding stores); through
// these the accesses are global loads and stores.  (The host pass of the
// compiler parses device functions too but has no such address space.)
template <class T>
__device__ __forceinline__ T gload(const T *p, long i)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return ((const __attribute__((address_space(1))) T *)p)[i];
#else
    return p[i];
#endif
}
template <class T>
__device__ __forceinline__ void gstore(T *p, long i, const T &v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    ((__attribute__((address_space(1))) T *)p)[i] = v;
#else
    p[i] = v;
#endif
}

// cos(2*pi*j/32), j = 0..31, and the matching sines, for the in-register
// DFT twiddles (R <= 32).
__device__ constexpr float kC32[32] = {
    1.0f, 0.98078528040323044f, 0.92387953251128674f, 0.83146961230254524f, 0.70710678118654752f,
    0.55557023301960218f, 0.38268343236508977f, 0.19509032201612826f, 0.0f, -0.19509032201612826f,
    -0.38268343236508977f, -0.55557023301960218f, -0.70710678118654752f, -0.83146961230254524f,
    -0.92387953251128674f, -0.98078528040323044f, -1.0f, -0.98078528040323044f, -0.92387953251128674f,
    -0.83146961230254524f, -0.70710678118654752f, -0.55557023301960218f, -0.38268343236508977f,
    -0.19509032201612826f, 0.0f, 0.19509032201612826f, 0.38268343236508977f, 0.55557023301960218f,
    0.70710678118654752f, 0.83146961230254524f, 0.92387953251128674f, 0.98078528040323044f};

// a * exp(DIR*2*pi*i*k/R) with compile-time k, R
template <int R, int K, int DIR>
__device__ __forceinline__ float2 twc(float2 a)
{
    constexpr int j = (K * (32 / R)) & 31;
    if constexpr (j == 0) {
        return a;
    } else if constexpr (j == 8) {
        return mul_i<DIR>(a);
    } else if constexpr (j == 16) {
        return make_float2(-a.x, -a.y);
    } else if constexpr (j == 24) {
        return mul_i<-DIR>(a);
    } else {
        constexpr float c = kC32[j];
        constexpr float s = kC32[(j + 24) & 31] * (float)DIR;  // sin(2 pi j/32) * DIR
        return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

// In-register radix-R DFT, natural order in and out:
//   a[k] <- sum_n a[n] exp(DIR*2*pi*i*k*n/R).
template <int R, int DIR>
struct Dft {
    template <int K>
    __device__ __forceinline__ static void combine(float2 *a, const float2 *e, const float2 *o)
    {
        if constexpr (K < R / 2) {
            float2 t = twc<R, K, DIR>(o[K]);
            a[K] = cadd(e[K], t);
            a[K + R / 2] = csub(e[K], t);
            combine<K + 1>(a, e, o);
        }
    }
    __device__ __forceinline__ static void run(float2 *a)
    {
        float2 e[R / 2], o[R / 2];
#pragma unroll
        for (int n = 0; n < R / 2; n++) {
            e[n] = a[2 * n];
            o[n] = a[2 * n + 1];
        }
        Dft<R / 2, DIR>::run(e);
        Dft<R / 2, DIR>::run(o);
        combine<0>(a, e, o);
    }
};
template <int DIR>
struct Dft<1, DIR> {
    __device__ __forceinline__ static void run(float2 *) {}
};
template <int DIR>
struct Dft<2, DIR> {
    __device__ __forceinline__ static void run(float2 *a)
    {
        float2 t = a[1];
        a[1] = csub(a[0], t);
        a[0] = cadd(a[0], t);
    }
};
template <int DIR>
struct Dft<4, DIR> {
    __device__ __forceinline__ static void run(float2 *a)
    {
        float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        float2 t2 = cadd(a[1], a[3]), t3 = mul_i<DIR>(csub(a[1], a[3]));
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    }
};

// LDS index with padding: one extra slot every 32 entries.  This is the
// layout of a finished transform in LDS (and of the last Stockham stage's
// exchange); the earlier stages use their own (stage_pad).
__device__ __forceinline__ int lpad(int i) { return i + (i >> 5); }

// Exchange layout of a Stockham stage that is not the last.  The LDS has
// 32 banks of 4 bytes and serves an 8-byte store 16 lanes at a time, an
// 8-byte load 32 lanes at a time.  Loads read consecutive entries, which
// any of these layouts keeps at two dwords per bank for 32 lanes (the
// least 64 dwords allow).
// Stores put lane j's R outputs at (j - k)*R + k + r*NS, k = j mod NS:
// * NS = 1 (stride R): one slot every 16 entries spreads 16 lanes over
//   all 16 bank pairs (for R = 2, 4, 8, 16);
// * NS = 8, R = 8 (runs of 8 lanes, 64 entries apart): 8 slots every 64
//   entries put the two runs of a quarter-wave on opposite halves of the
//   banks;
// * otherwise runs of 16 or more consecutive entries: lpad.
// Each form is i + c*floor(i/2^s), so pad(base + r*NS) = pad(base) +
// pad(r*NS) for the bases a stage uses, as for lpad.
template <int NS, int R>
__device__ __forceinline__ int stage_pad(int i)
{
    if constexpr (NS == 1) return i + (i >> 4);
    else if constexpr (NS == 8 && R == 8) return i + 8 * (i >> 6);
    else return lpad(i);
}
// entries of one line's LDS region: the largest of the layouts (i + i/8)
__host__ __device__ constexpr int lds_len(int n) { return n + (n >> 3) + 1; }

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n / 2); }

// elements per thread for a line of n points in a column (strided) pass
__host__ __device__ constexpr int elems_for(int n) { return n >= 256 ? 16 : (n >= 16 ? 8 : n); }

// elements per thread in a row pass: one wavefront (64 lanes) per row from
// 128 complex points up, 16 elements per lane at most
__host__ __device__ constexpr int row_elems(int nc)
{
    return nc >= 1024 ? 16 : (nc >= 128 ? nc / 64 : (nc >= 4 ? 4 : nc));
}

// XCD-aware workgroup placement: the hardware deals workgroups to the 8
// XCDs round-robin, so consecutive linear ids would put neighbouring tiles
// (which share cache lines) on different L2s.  Linear id b is remapped to
// work item (b % 8) * (nblk / 8) + b / 8, so each XCD sweeps one contiguous
// range of tiles.  nblk is a multiple of 8; items >= total are idle.
__host__ __device__ constexpr long xcd_blocks(long total) { return (total + 7) / 8 * 8; }
__device__ __forceinline__ long xcd_item(long b, long nblk) { return (b & 7) * (nblk >> 3) + (b >> 3); }

// a table entry exp(-2*pi*i*x) as exp(DIR*2*pi*i*x)
template <int DIR>
__device__ __forceinline__ float2 tw_dir(float2 w)
{
    return DIR > 0 ? make_float2(w.x, -w.y) : w;
}

// exp(DIR*2*pi*i*m/L) from a table of exp(-2*pi*i*j/TWS), j < TWS
// (L | TWS, 0 <= m).  The table lives in global memory (TWS = TW_N) or in
// the workgroup's LDS (TWS = the largest length the kernel transforms).
template <int DIR, int TWS = TW_N>
__device__ __forceinline__ float2 tw_lookup(const float2 *__restrict__ tw, int m, int L)
{
    int j = (m * (TWS / L)) & (TWS - 1);
    float2 w = tw[j];
    return DIR > 0 ? make_float2(w.x, -w.y) : w;
}

// The device keeps one table per power-of-two size M at offset M of one
// array (twiddle_sets()); a kernel copies table M into LDS with this.
template <int M>
__device__ __forceinline__ void load_twiddles(float2 *__restrict__ dst, const float2 *__restrict__ sets)
{
    for (int i = threadIdx.x; i < M; i += blockDim.x) dst[i] = sets[M + i];
}

// The same copy (table M, 16-byte aligned in LDS) by a workgroup of TH
// threads in two halves, as StageTwiddles below: all of a thread's loads
// are in flight at once (up to 2 each; more fall back to a loop in store()).
template <int M, int TH>
struct SizeTwiddles {
    static constexpr int M4 = M / 2;
    static constexpr int PER = M4 > 0 ? (M4 + TH - 1) / TH : 1;
    static constexpr bool REG = M % 2 == 0 && PER <= 2;
    float4 r[REG ? PER : 1];
    __device__ __forceinline__ void load(const float2 *__restrict__ sets)
    {
        if constexpr (REG) {
            const float4 *__restrict__ src = (const float4 *)(sets + M);
#pragma unroll
            for (int p = 0; p < PER; p++) {
                const int i = threadIdx.x + p * TH;
                if (i < M4) r[p] = gload(src, i);
            }
        }
    }
    __device__ __forceinline__ void store(float2 *__restrict__ dst, const float2 *__restrict__ sets) const
    {
        if constexpr (REG) {
#pragma unroll
            for (int p = 0; p < PER; p++) {
                const int i = threadIdx.x + p * TH;
                if (i < M4) ((float4 *)dst)[i] = r[p];
            }
        } else {
            for (int i = threadIdx.x; i < M; i += TH) dst[i] = sets[M + i];
        }
    }
};

// ---- stage-ordered twiddles (TWS = TW_STAGE)
//
// For the transform of N points with E per lane, the stage of span NS > 1
// and radix R needs w^(r*k) = exp(-2*pi*i*r*k/(NS*R)) for r = 1..R-1 and
// k = j mod NS.  The stage table holds them as R-1 rows of NS: entry
// (r-1)*NS + k is w^(r*k), so a lane fetches its R-1 twiddles at one
// address (k) plus constant offsets, each as an 8-byte load in which
// consecutive lanes read consecutive entries: the LDS then serves the 32
// lanes of a half-wave from distinct bank pairs at full rate (a 16-byte
// load per lane would be served as two 8-byte passes of stride 16 bytes,
// half the rate).  The stages' tables follow each other in stage order,
// N - R0 entries in all for a first radix R0.  The tables of every (N, E) sit in the same
// device array as the per-size tables (twiddle_sets()), at tw_stage_offset.
constexpr int TW_STAGE = 0;

__host__ __device__ constexpr int tw_stage_radix(int N, int E, int NS)
{
    return N / NS >= E ? (E >= 2 ? E : 2) : N / NS;
}
__host__ __device__ constexpr int tw_stage_len(int N, int E, int NS)
{
    return (tw_stage_radix(N, E, NS) - 1) * NS;
}
// offset of the table of the stage with span NS (its sum of the earlier
// twiddled stages' lengths); tw_stage_base(N, E, N) is the whole size
__host__ __device__ constexpr int tw_stage_base(int N, int E, int NS)
{
    int off = 0;
    for (int ns = tw_stage_radix(N, E, 1); ns < NS && ns < N; ns *= tw_stage_radix(N, E, ns))
        off += tw_stage_len(N, E, ns);
    return off;
}
__host__ __device__ constexpr int tw_stage_size(int N, int E) { return tw_stage_base(N, E, N); }
// LDS entries to declare for them (a one-stage transform has none)
__host__ __device__ constexpr int tw_stage_alloc(int N, int E) { return tw_stage_size(N, E) > 0 ? tw_stage_size(N, E) : 2; }
// Position of the (N, E) tables in the twiddle-set array: after the 2*TW_N
// entries of the per-size tables, one region of 4*TW_N per E = 1, 2, 4, 8,
// 16 holding the table of N at [2N, 4N) of the region (every table of N
// points is shorter than 2N entries)
__host__ __device__ constexpr long tw_stage_offset(int N, int E)
{
    return 2L * TW_N + (long)ilog2c(E) * 4 * TW_N + 2L * N;
}
constexpr long TW_SETS_LEN = 2L * TW_N + 5L * 4 * TW_N;

// The stage tables of (N, E), copied into LDS (16-byte aligned) by a
// workgroup of TH threads in two halves: load() issues the global loads
// into registers, store() writes them to LDS.  A kernel issues its own
// loads between the two, so that waiting for the table does not also wait
// for them (loads complete in order for the wait counter).
template <int N, int E, int TH>
struct StageTwiddles {
    static constexpr int M4 = tw_stage_size(N, E) / 2;  // 16-byte pairs
    static constexpr int PER = M4 > 0 ? (M4 + TH - 1) / TH : 1;
    static constexpr bool REG = PER <= 2;  // more: a loop in store() (registers)
    static_assert(tw_stage_size(N, E) % 2 == 0, "stage tables hold whole pairs");
    float4 r[REG ? PER : 1];
    __device__ __forceinline__ void load(const float2 *__restrict__ sets)
    {
        if constexpr (REG) {
            const float4 *__restrict__ src = (const float4 *)(sets + tw_stage_offset(N, E));
#pragma unroll
            for (int p = 0; p < PER; p++) {
                const int i = threadIdx.x + p * TH;
                if (i < M4) r[p] = gload(src, i);
            }
        }
    }
    __device__ __forceinline__ void store(float2 *__restrict__ dst, const float2 *__restrict__ sets) const
    {
        if constexpr (REG) {
#pragma unroll
            for (int p = 0; p < PER; p++) {
                const int i = threadIdx.x + p * TH;
                if (i < M4) ((float4 *)dst)[i] = r[p];
            }
        } else {
            const float4 *__restrict__ src = (const float4 *)(sets + tw_stage_offset(N, E));
            for (int i = threadIdx.x; i < M4; i += TH) ((float4 *)dst)[i] = gload(src, i);
        }
    }
};

// the same in one call (any workgroup size)
template <int N, int E>
__device__ __forceinline__ void load_stage_twiddles(float2 *__restrict__ dst, const float2 *__restrict__ sets)
{
    constexpr int M4 = tw_stage_size(N, E) / 2;
    const float4 *__restrict__ src = (const float4 *)(sets + tw_stage_offset(N, E));
    float4 *__restrict__ d = (float4 *)dst;
    for (int i = threadIdx.x; i < M4; i += blockDim.x) d[i] = gload(src, i);
}

// Synchronisation of one line's LDS exchanges.  WL (wave-local): the line's
// lanes lie inside one wavefront and its LDS region is its own.  A wave's
// LDS accesses are performed in program order, so its lanes then exchange
// values through LDS with no workgroup barrier: the fences (wave scope,
// which emit no instruction) only keep the compiler from moving accesses
// across the exchange.  Otherwise a workgroup barrier.
template <bool WL>
__device__ __forceinline__ void line_sync()
{
    if constexpr (WL) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// One Stockham stage of radix R with span NS, then recurse.  OUT_LDS: the
// last stage also leaves the transform in LDS (natural order); without it
// the last stage's butterfly outputs go straight back to the registers --
// in the strided distribution they already sit in the slots their inputs
// came from (output t + T*(b + r*E/R) of butterfly b) -- and the final LDS
// round trip is skipped.
template <int N, int E, int DIR, int NS, int TWS = TW_N, bool OUT_LDS = true, bool WL = false>
struct Stockham {
    static constexpr int T = N / E;
    static constexpr int REM = N / NS;
    static constexpr int R = REM >= E ? E : REM;
    static constexpr bool LAST = NS * R == N;
    // the last stage leaves the transform in the lpad layout
    __device__ __forceinline__ static int pad(int i)
    {
        if constexpr (LAST) return lpad(i);
        else return stage_pad<NS, R>(i);
    }
    __device__ __forceinline__ static void run(float2 (&v)[E], float2 *lds, int t, const float2 *__restrict__ tw)
    {
        if constexpr (NS < N) {
#pragma unroll
            for (int b = 0; b < E / R; b++) {
                float2 a[R];
#pragma unroll
                for (int r = 0; r < R; r++) a[r] = v[b + r * (E / R)];
                const int j = t + b * T;
                const int k = j & (NS - 1);
                if constexpr (NS > 1 && TWS == TW_STAGE) {
                    const float2 *__restrict__ st = tw + tw_stage_base(N, E, NS) + k;
#pragma unroll
                    for (int r = 1; r < R; r++) a[r] = cmul(a[r], tw_dir<DIR>(st[(r - 1) * NS]));
                } else if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; r++) a[r] = cmul(a[r], tw_lookup<DIR, TWS>(tw, r * k, NS * R));
                }
                Dft<R, DIR>::run(a);
                if constexpr (LAST && !OUT_LDS) {
#pragma unroll
                    for (int r = 0; r < R; r++) v[b + r * (E / R)] = a[r];
                } else {
                    // pad(base + r*NS) = pad(base) + pad(r*NS): base = (j-k)*R + k
                    // with (j-k)*R a multiple of NS*R, so the R outputs either stay
                    // inside one padding block (NS*R <= its size) or step across
                    // blocks exactly as r*NS does.  One address register, constant
                    // offsets.
                    const int wb = pad((j - k) * R + k);
#pragma unroll
                    for (int r = 0; r < R; r++) lds[wb + pad(r * NS)] = a[r];
                }
            }
            if constexpr (!(LAST && !OUT_LDS)) {
                line_sync<WL>();
                // pad(t + e*T) = pad(t) + pad(e*T) for t < T, T a power of two
                const int rb = pad(t);
#pragma unroll
                for (int e = 0; e < E; e++) v[e] = lds[rb + pad(e * T)];
                if constexpr (!LAST) line_sync<WL>();
            }
            Stockham<N, E, DIR, NS * R, TWS, OUT_LDS, WL>::run(v, lds, t, tw);
        }
    }
};

// Full transform of one line.  On entry v[e] = x[t + e*T]; on exit
// v[e] = X[t + e*T] and, with OUT_LDS, lds[lpad(k)] = X[k] for all k (after
// a line_sync).  With WL false all threads of the workgroup must call it
// (it contains barriers); WL as for line_sync.
template <int N, int E, int DIR, int TWS = TW_N, bool OUT_LDS = true, bool WL = false>
__device__ __forceinline__ void fft_line(float2 (&v)[E], float2 *lds, int t, const float2 *__restrict__ tw)
{
    if constexpr (N == 1) {
        if constexpr (OUT_LDS) {
            lds[0] = v[0];
            line_sync<WL>();
        }
    } else {
        Stockham<N, E, DIR, 1, TWS, OUT_LDS, WL>::run(v, lds, t, tw);
    }
}

}  // namespace cupss
