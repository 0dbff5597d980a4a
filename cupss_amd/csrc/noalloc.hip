// noalloc.hip -- cupss_c2r_noalloc: the stand-alone inverse transform
// (cupss_c2r) for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// The intermediate half spectrum of the column passes, which cupss_c2r
// takes as a workspace argument, lives here in a file-scope device array of
// NOALLOC_MAX_BINS float2 (16 MiB per device, held until the process ends),
// and the twiddle table, which cupss_c2r allocates on its first call, in a
// second one.  Host code never passes these arrays' names to a kernel (on
// the host they are stand-ins, not device addresses): the first call on
// each device, with that device current, looks their device addresses up
// with hipGetSymbolAddress, uploads the twiddle table with one synchronous
// copy (no allocation), and keeps both addresses per device.  Every call on
// a device shares that device's scratch array, which keeps the last call's
// contents; nothing in it has to start at zero (every entry a pass reads
// was written by the pass before it in the same call).  Two calls that
// overlap on one device (different streams, or host threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "fft_core.hpp"
#include "internal.hpp"
#include "system_impl.hpp"

namespace cupss {
namespace {

constexpr long NOALLOC_MAX_BINS = 1L << 21;  // float2 entries: 16 MiB

static __device__ __align__(256) float2 scratch[NOALLOC_MAX_BINS];
static __device__ __align__(256) float2 scratch_twiddles[TW_N];

bool pow2(int v) { return v >= 1 && (v & (v - 1)) == 0; }

struct DevicePtrs {
    float2 *scratch = nullptr;
    float2 *twiddles = nullptr;
};

std::mutex g_mutex;
std::vector<DevicePtrs> g_ptrs;  // per device, filled at the first call

int device_ptrs(DevicePtrs &out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mutex);
    if ((int)g_ptrs.size() <= dev) g_ptrs.resize(dev + 1);
    DevicePtrs &p = g_ptrs[dev];
    if (!p.scratch) {
        void *sp = nullptr, *tp = nullptr;
        HIP_TRY(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
        HIP_TRY(hipGetSymbolAddress(&tp, HIP_SYMBOL(scratch_twiddles)));
        std::vector<float2> h(TW_N);
        twiddle_table(h.data());
        HIP_TRY(hipMemcpy(tp, h.data(), sizeof(float2) * TW_N, hipMemcpyHostToDevice));
        p.twiddles = (float2 *)tp;
        p.scratch = (float2 *)sp;  // set last: a set scratch pointer means the set-up is done
    }
    out = p;
    return CUPSS_OK;
}

}  // namespace
}  // namespace cupss

using namespace cupss;

extern "C" int cupss_c2r_noalloc(int sx, int sy, int sz, const void *in, int mask_p, float scale, float *out,
                                 void *hip_stream)
{
    if (!pow2(sx) || sx < 2 || sx > 4096 || !pow2(sy) || sy > 4096 || !pow2(sz) || sz > 4096)
        return fail(CUPSS_ERR_UNSUPPORTED, "unsupported grid");
    Geom g{};
    g.sx = sx;
    g.sy = sy;
    g.sz = sz;
    g.nc = sx / 2;
    g.nc1 = sx / 2 + 1;
    g.nh = (long)g.nc1 * sy * sz;
    g.n = (long)sx * sy * sz;
    if (g.nh > NOALLOC_MAX_BINS)
        return fail(CUPSS_ERR_INVALID, "cupss_c2r_noalloc: the half spectrum has " + std::to_string(g.nh) +
                                           " bins; the static scratch holds at most " +
                                           std::to_string(NOALLOC_MAX_BINS));
    if (!in || !out) return fail(CUPSS_ERR_INVALID, "null buffer");
    DevicePtrs p;
    if (int r = device_ptrs(p)) return r;
    HIP_TRY(c2r_tw(g, (const float2 *)in, mask_p, scale, out, p.scratch, (hipStream_t)hip_stream, p.twiddles));
    return CUPSS_OK;
}
