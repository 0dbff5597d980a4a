// fused.hpp -- the fused two-kernel step for 2D systems with one dynamic
// field phi whose other fields are linear in phi (fused.hip), batched over
// any number of independent instances (batch.hip).
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
constexpr int FUSED_MAX_DERIVED = 4; // derived fields written back after a run

enum NoiseMode : int {
    NOISE_NONE = 0,
    NOISE_REAL = 1,      // xi ~ N(0,1) per real-space cell, then R2C (the reference's way)
    NOISE_SPECTRAL = 2,  // eta_hat drawn directly with the law of R2C(xi)
};

// Device-side descriptor of one instance (one system) of a fused batch.
struct FusedInst {
    Geom g;
    float2 *spec[2];                          // phi's spectrum: [parity] in, [parity^1] out
    float2 *Z[FUSED_MAX_REAL];                // IFFT_y(mask_f L_f phi_hat)
    float2 *X[FUSED_MAX_GROUPS];              // R2C_x(products of group g)
    float2 *eta;                              // R2C_x(xi) rows (NOISE_REAL)
    uint64_t step_base;                       // step counter of the first step of a launch sequence
    // row pass
    float scale;                              // 1/N of the inverse transform
    int nreal;
    int ngroups;
    int g0[FUSED_MAX_GROUPS + 1];             // products of group g: [g0[g], g0[g+1])
    int nfac[FUSED_MAX_PRODS];
    int fac[FUSED_MAX_PRODS][FUSED_MAX_FAC];  // indices into Z / the real fields
    int noise_mode;
    uint32_t key0, key1, stream;
    // column pass
    UpdateArgs u;                             // phi's update; grouped product terms are TERM_GROUPED
    int gterm[FUSED_MAX_GROUPS];              // u.terms index carrying group g's prefactors
    int p[FUSED_MAX_REAL];                    // dealias order of each real field
    int zero[FUSED_MAX_REAL];                 // real field is zero in the prologue (step-0 state)
    TermRefDev lmul[FUSED_MAX_REAL];          // L_f (npres = 0: identity)
    PresDev lpres[FUSED_MAX_LPRES];
    // derived fields psi_d = P(M_d phi_old), written after a run
    int nder;
    float2 *dout[FUSED_MAX_DERIVED];
    TermRefDev dmul[FUSED_MAX_DERIVED];
    PresDev dpres[FUSED_MAX_LPRES];
};

// Work item of one workgroup: instance and first row / column (inst < 0: idle).
struct FusedBlock {
    int inst;
    int line;
};

// Configuration of a size class (all its instances share nc and sy).
int fused_rows_per_block(int nc);
int fused_cols_per_block(int ny);
bool fused_supported(int sx, int sy, int sz);

hipError_t fused_rows_launch(int nc, const FusedInst *inst, const FusedBlock *blocks, int nblocks,
                             uint32_t step_rel, hipStream_t s);
// mode 0: prologue (real fields of the current state), 1: step
hipError_t fused_cols_launch(int ny, const FusedInst *inst, const FusedBlock *blocks, int nblocks, int parity,
                             uint32_t step_rel, int mode, hipStream_t s);
// derived fields of every instance from phi's spectrum buffer [parity]
// (blocks: one spectral row each)
hipError_t fused_derived_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int parity,
                                hipStream_t s);
// elementwise step of instances with no products and no real-space noise
// (blocks: one spectral row each)
hipError_t fused_elem_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int parity,
                             uint32_t step_rel, hipStream_t s);

}  // namespace cupss
