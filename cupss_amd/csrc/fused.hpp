// fused.hpp -- the fused two-kernel step for 2D systems with one dynamic
// field phi whose other fields are linear in phi (fused.hip), batched over
// any number of independent instances (batch.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "internal.hpp"

namespace cupss {

constexpr int FUSED_MAX_REAL = 3;    // dealiased real fields read by the products
constexpr int FUSED_MAX_GROUPS = 4;  // groups of product terms with equal prefactors
constexpr int FUSED_MAX_PRODS = 8;   // product terms over all groups
constexpr int FUSED_MAX_FAC = 4;     // factors of one product
constexpr int FUSED_MAX_LPRES = 16;  // prefactors of the derived-field multipliers
constexpr int FUSED_MAX_DERIVED = 4; // derived fields written back after a run
constexpr int FUSED_TILE = 32;       // side of the transpose tiles
constexpr int FUSED_ELEM_BLOCK = 256;  // bin pairs per workgroup of the elementwise kernel

enum NoiseMode : int {
    NOISE_NONE = 0,
    NOISE_REAL = 1,      // xi ~ N(0,1) per real-space cell, then R2C (the reference's way)
    NOISE_SPECTRAL = 2,  // eta_hat drawn directly with the law of R2C(xi)
};

// Device-side descriptor of one instance (one system) of a fused batch.
//
// Transform instances keep phi's spectrum and every pipeline buffer
// column-major during a run: element (kx, y) at kx * sy + y, so the column
// kernel reads and writes whole contiguous columns.  Elementwise instances
// use the natural half-spectrum layout (ky, kx) of the system buffers.
struct FusedInst {
    Geom g;
    float2 *spec[2];                          // phi's spectrum: [parity] in, [parity^1] out
    float2 *nat;                              // the system's spectrum (natural layout), transform instances
    float2 *Z[FUSED_MAX_REAL];                // IFFT_y(mask_f L_f phi_hat), columns kx <= kmax[f]
    float2 *X[FUSED_MAX_GROUPS];              // R2C_x(products of group g)
    float2 *eta;                              // R2C_x(xi) rows (NOISE_REAL)
    uint64_t step_base;                       // step counter of the first step of a launch sequence
    // row pass
    float scale;                              // 1/N of the inverse transform
    int nreal;
    int ngroups;
    int g0[FUSED_MAX_GROUPS + 1];             // products of group g: [g0[g], g0[g+1])
    int cnt[FUSED_MAX_PRODS][FUSED_MAX_REAL]; // power of each real field in product p
    int noise_mode;
    uint32_t key0, key1, stream;
    // column pass
    UpdateArgs u;                             // phi's update; grouped product terms are TERM_GROUPED
    int gterm[FUSED_MAX_GROUPS];              // u.terms index carrying group g's prefactors
    int p[FUSED_MAX_REAL];                    // dealias order of each real field
    int kmax[FUSED_MAX_REAL];                 // last column kx the dealias mask of field f keeps
    int kymax[FUSED_MAX_REAL];                // largest |ky| the mask keeps
    int zero[FUSED_MAX_REAL];                 // real field is zero in the prologue (step-0 state)
    TermRefDev lmul[FUSED_MAX_REAL];          // L_f (npres = 0: identity)
    PresDev lpres[FUSED_MAX_LPRES];
    // derived fields psi_d = P(M_d phi_old), written after a run
    int nder;
    float2 *dout[FUSED_MAX_DERIVED];
    TermRefDev dmul[FUSED_MAX_DERIVED];
    PresDev dpres[FUSED_MAX_LPRES];
};

// Work item of one workgroup: instance and first row / column / tile / bin
// (inst < 0: idle).
struct FusedBlock {
    int inst;
    int line;
};

bool fused_supported(int sx, int sy, int sz);

// Tile shapes a size class can be launched with: rows per workgroup of the
// row kernel (nc = sx/2) and columns per workgroup of the column kernel
// (ny = sy).  Any choice gives bit-identical results.
std::vector<int> fused_row_shapes(int nc, int ny);
std::vector<int> fused_col_shapes(int ny);

hipError_t fused_rows_launch(int nc, int lpb, const FusedInst *inst, const FusedBlock *blocks, int nblocks,
                             uint32_t step_rel, hipStream_t s);
// mode 0: prologue (real fields of the current state), 1: step
hipError_t fused_cols_launch(int ny, int twd, const FusedInst *inst, const FusedBlock *blocks, int nblocks,
                             int parity, uint32_t step_rel, int mode, hipStream_t s);
// natural spectrum -> column-major spec[0] (blocks: one tile each)
hipError_t fused_transpose_in_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, hipStream_t s);
// column-major spec[fin] -> natural spectrum, derived fields from spec[fin^1]
hipError_t fused_settle_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int fin,
                               hipStream_t s);
// nsteps steps of instances with no products and no real-space noise
// (blocks: FUSED_ELEM_BLOCK bin pairs each); derived fields from the last step's start
hipError_t fused_elem_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int parity,
                             uint32_t step_rel, int nsteps, hipStream_t s);

}  // namespace cupss
