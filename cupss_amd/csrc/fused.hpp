// fused.hpp -- arguments and launchers of the fused two-kernel 2D step
// (fused.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.hpp"

namespace cupss {

constexpr int FUSED_MAX_REAL = 3;    // dealiased real fields read by the products
constexpr int FUSED_MAX_GROUPS = 4;  // groups of product terms with equal prefactors
constexpr int FUSED_MAX_PRODS = 8;   // product terms over all groups
constexpr int FUSED_MAX_FAC = 4;     // factors of one product
constexpr int FUSED_MAX_LPRES = 16;  // prefactors of the derived-field multipliers

struct FusedRowArgs {
    long nrows;
    float scale;                          // 1/N of the inverse transform
    int nreal;
    const float2 *Z[FUSED_MAX_REAL];
    int ngroups;
    int g0[FUSED_MAX_GROUPS + 1];         // products of group g: [g0[g], g0[g+1])
    int nfac[FUSED_MAX_PRODS];
    int fac[FUSED_MAX_PRODS][FUSED_MAX_FAC];  // indices into Z / real fields
    float2 *X[FUSED_MAX_GROUPS];
    int noisy;
    float2 *eta;
    uint32_t key0, key1, step_lo, step_hi, stream;
};

struct FusedColArgs {
    UpdateArgs u;                         // the update of phi; grouped terms are TERM_GROUPED
    int ngroups;
    const float2 *X[FUSED_MAX_GROUPS];
    int gterm[FUSED_MAX_GROUPS];          // u.terms index carrying the group's prefactors
    const float2 *spec_in;
    float2 *spec_out;
    int do_update;
    int nreal;
    int p[FUSED_MAX_REAL];
    int zero[FUSED_MAX_REAL];
    float2 *Z[FUSED_MAX_REAL];
    TermRefDev lmul[FUSED_MAX_REAL];      // multiplier L_f (npres = 0: identity)
    PresDev lpres[FUSED_MAX_LPRES];
};

bool fused_supported(int sx, int sy, int sz);
hipError_t fused_rows(const FusedRowArgs &a, int nc, hipStream_t s);
hipError_t fused_cols(const FusedColArgs &a, int n, hipStream_t s);

}  // namespace cupss
