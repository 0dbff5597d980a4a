// fused.hpp -- the fused two-kernel step for 2D systems (fused.hip),
// batched over any number of independent instances (batch.hip), and the
// four-pass fused step of 3D systems with one dynamic field (one at a time).
//
// The first dynamic field is phi; the fields of such a system are
//   derived:  psi = M_psi u for a dynamic field u (one single-field term, no
//             implicit operator, no noise) -- never stored during a run,
//             evaluated from u where needed;
//   stored:   the other dynamic fields (2D); constraints whose products
//             read dynamic fields only and whose other terms are dynamic
//             fields or constants ("fresh": solved at the start of a pass);
//             then any other constraint field whose product
//             terms read only derived or stored constraint fields (or that
//             has no terms: it keeps its value) -- kept as
//             state, updated every pass (the "phase A/B" of the next step,
//             see fused.hip).  No product reads a fresh constraint.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "internal.hpp"

namespace cupss {

constexpr int FUSED_MAX_REAL = 6;     // dealiased real fields read by the products
constexpr int FUSED_MAX_GROUPS = 8;   // groups of product terms (one transform each)
constexpr int FUSED_MAX_PRODS = 12;   // product terms over all groups
constexpr int FUSED_MAX_CONS = 6;     // stored constraint fields
constexpr int FUSED_MAX_TARGETS = 1 + FUSED_MAX_CONS;
constexpr int FUSED_MAX_TTERMS = 8;   // non-product terms of one field
constexpr int FUSED_MAX_POOL = 96;    // prefactors of all multipliers of an instance
constexpr int FUSED_MAX_DERIVED = 4;  // derived fields written back after a run
constexpr int FUSED_TILE = 32;        // side of the transpose tiles
constexpr int FUSED_ELEM_BLOCK = 256; // bin pairs per workgroup of the elementwise kernel
constexpr int FUSED_TS_SLOTS = 64;    // timestamp slots per timed launch
constexpr int FUSED_TS_STRIDE = 16;   // u64 between slots (one 128-byte line each)

enum NoiseMode : int {
    NOISE_NONE = 0,
    NOISE_REAL = 1,      // xi ~ N(0,1) per real-space cell, then R2C (the reference's way)
    NOISE_SPECTRAL = 2,  // eta_hat drawn directly with the law of R2C(xi)
};

// Source of a non-product term (one factor, or none) in a pass that takes
// phi from step m to m+1 and the stored constraints from m to m+1.
enum FusedSrc : int {
    SRC_OLD = 0,      // phi_m
    SRC_NEW = 1,      // phi_{m+1} (terms of constraint fields: phi's pre-step value of step m+1)
    SRC_CONST = 2,    // no factor: N at q = 0
    SRC_DERIVED = 3,  // derived field idx: M_idx phi_m
    SRC_CONS = 4,     // stored field idx at step m (a dynamic one: its value before the step)
    SRC_DYN_NEW = 5,  // stored dynamic field idx after the step (terms of constraint fields)
};

struct FusedTerm {
    int src, idx;
    int mult_i, pres0, npres;  // multiplier in the pool
};

// One field a pass updates: phi (target 0) or stored field c (1 + c)
struct FusedTarget {
    int nterms;
    FusedTerm term[FUSED_MAX_TTERMS];
    int g0, g1;      // its product groups [g0, g1)
    int nimp, imp0;  // stored field: implicit prefactors in the pool (constraint: division for q != 0;
                     // dynamic: D(q) = 1 - dt sum)
    int z;           // stored field: index of its real field, -1 if none
    int dyn;         // stored dynamic field: semi-implicit Euler step of its own value
    int fresh;       // stored constraint solved at the start of the pass from the dynamic
                     // fields before the step (its products read dynamic fields)
    int keep;        // stored constraint without terms: keeps its value (a coefficient field),
                     // divided by its implicit operator's sum every step if it has one
    int noisy;       // stored field: its noise mode (NoiseMode), amplitude as UpdateArgs
    float noise_base;
    int noise_q2n, noise_invq;
    float sdt;       // constraint: 1/sqrt(dt) on the amplitude (setNotDynamic_k)
    int step_off;    // step counter of its noise: the pass's step + step_off (1: a
                     // constraint solved for the next step)
    uint32_t key0, key1, stream;
    float2 *eta;     // NOISE_REAL: R2C_x(xi) rows, column-major (row pass)
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
    float2 *cs[FUSED_MAX_CONS][2];            // stored constraints, column-major, like spec
    float2 *cnat[FUSED_MAX_CONS];             // their system buffers (natural layout)
    float2 *Z[FUSED_MAX_REAL];                // IFFT_y(mask_f * field_f), columns kx <= kmax[f]
    float2 *X[FUSED_MAX_GROUPS];              // R2C_x(sum of the products of group g)
    float2 *W[FUSED_MAX_REAL];                // 3D: IFFT_z(mask_f * L_f * phi), natural layout
    float2 *eta;                              // R2C_x(xi) rows (NOISE_REAL)
    uint64_t step_base;                       // step counter of the first step of a launch sequence
    // row pass
    float scale;                              // 1/N of the inverse transform
    int nreal;
    int ngroups;
    int pg0[FUSED_MAX_GROUPS + 1];            // products of group g: [pg0[g], pg0[g+1])
    int cnt[FUSED_MAX_PRODS][FUSED_MAX_REAL]; // power of each real field in product p
    float coef[FUSED_MAX_PRODS];              // real-space coefficient of product p
    int noise_mode;
    uint32_t key0, key1, stream;
    // column pass
    UpdateArgs u;                             // phi: dt, noise, implicit D(q) (and its terms, elementwise kernel)
    int ncons;
    int nfresh;                               // fresh constraints among the stored fields
    int cont;                                 // cs[.][0] already holds this run's first phase A/B
    int resume;                               // spec[0] and Z_f already hold this run's start (no prologue)
    FusedTarget tgt[FUSED_MAX_TARGETS];
    TermRefDev gmul[FUSED_MAX_GROUPS];        // multiplier of each group (pool)
    PresDev pool[FUSED_MAX_POOL];
    int p[FUSED_MAX_REAL];                    // dealias order of each real field
    int kmax[FUSED_MAX_REAL];                 // last column kx the dealias mask of field f keeps
    int kymax[FUSED_MAX_REAL];                // largest |ky| the mask keeps
    int zero[FUSED_MAX_REAL];                 // real field is zero in the prologue (step-0 state)
    int zsrc[FUSED_MAX_REAL];                 // -1: L_f u (lmul, u = lbase); c >= 0: stored field c
    int lbase[FUSED_MAX_REAL];                // zsrc -1: u is phi (-1) or stored dynamic field c
    TermRefDev lmul[FUSED_MAX_REAL];          // L_f (npres = 0: identity; pool)
    // derived fields psi_d = P(M_d u_old), written after a run
    int nder;
    float2 *dout[FUSED_MAX_DERIVED];
    TermRefDev dmul[FUSED_MAX_DERIVED];       // pool
    int dbase[FUSED_MAX_DERIVED];             // u: phi (-1) or stored dynamic field c
};

// Work item of one workgroup: instance and first row / column / tile / bin
// (inst < 0: idle).
struct FusedBlock {
    int inst;
    int line;
};

bool fused_supported(int sx, int sy, int sz);
// the row and column (2D) or row (3D) kernels exist for this grid with
// nreal real fields (more than 3: the wide row kernel)
bool fused_shapes_exist(int sx, int sy, int sz, int nreal);

// Tile shapes a size class can be launched with: rows per workgroup of the
// row kernel (nc = sx/2; wide = more than 3 real fields) and columns per
// workgroup of the column kernel (ny = sy).  Any choice gives bit-identical
// results.
// row kernel width (real fields held at once) for a system with nreal
int fused_row_width(int nreal);
std::vector<int> fused_row_shapes(int nc, int ny, int width);
std::vector<int> fused_col_shapes(int ny);

hipError_t fused_rows_launch(int nc, int lpb, int width, const FusedInst *inst, const FusedBlock *blocks,
                             int nblocks, uint32_t step_rel, hipStream_t s);
// mode 0: prologue (real fields of the current state), 1: step
// (ts: null, or FUSED_TS_SLOTS pairs {earliest start, latest end} of
// wall-clock ticks, FUSED_TS_STRIDE apart, that the launch's workgroups
// min/max-update; the launch spans min(starts) .. max(ends))
// (stored: 0 = no instance of the launch has stored fields, 1 = some have
// stored constraint fields, 2 = some have stored dynamic fields or fresh
// constraints)
hipError_t fused_cols_launch(int ny, int twd, int stored, const FusedInst *inst, const FusedBlock *blocks, int nblocks,
                             int parity, uint32_t step_rel, int mode, hipStream_t s, unsigned long long *ts = nullptr);
// Size classes whose instances take whole steps in one workgroup each
// (64^2, 128^2): k steps [n0, n0+k) of the instances ids[0..nids) in one
// launch (step n reads spectrum buffer n&1, noise counter offset n - rel0);
// stored and ts as for fused_cols_launch (one slot pair for the launch).
bool fused_small_class(int nc, int ny, int width);
hipError_t fused_small_launch(int nc, int ny, int stored, const FusedInst *inst, const int *ids, int nids, int n0,
                              int k, int rel0, hipStream_t s, unsigned long long *ts = nullptr);
// the slots of n timed launches to {~0, 0} (a kernel, so that graphs can hold it)
hipError_t fused_ts_reset_launch(unsigned long long *ts, int n, hipStream_t s);
// ---- 3D systems (one instance, natural half-spectrum layout; insts[0]).
// A step is four passes: rows (fused3d_rows_launch), FFT_y of X_g
// (lines_y), the z pass (fused3d_z_launch) and IFFT_y of W_f into Z_f
// (lines_y).
// z-line tile widths (kx lines per workgroup) for sz
std::vector<int> fused3d_z_shapes(int nz);
hipError_t fused3d_rows_launch(int nc, int maxr, const FusedInst *inst, long nrows, hipStream_t s);
// mode 0: prologue (W_f of the current state), 1: step, 2: step without W_f
// (timing); ts as for fused_cols_launch
hipError_t fused3d_z_launch(int nz, int twd, const FusedInst *inst, int nc1, int sy, int parity, uint32_t step_rel,
                            int mode, hipStream_t s, unsigned long long *ts = nullptr);
// derived fields from phi's spectrum [parity], natural layout
hipError_t fused3d_derived_launch(const FusedInst *inst, long nh, int parity, hipStream_t s);

// natural spectra -> column-major buffer 0 (phi and the stored constraints;
// blocks: one tile each)
hipError_t fused_transpose_in_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, hipStream_t s);
// column-major phi and stored dynamic fields [fin], stored constraints
// [fin^1] -> natural; derived fields from their base's [fin^1]
hipError_t fused_settle_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int fin,
                               hipStream_t s);
// nsteps steps of instances with no products, no stored constraints and no
// real-space noise (blocks: FUSED_ELEM_BLOCK bin pairs each); derived fields
// from the last step's start
hipError_t fused_elem_launch(const FusedInst *inst, const FusedBlock *blocks, int nblocks, int parity,
                             uint32_t step_rel, int nsteps, hipStream_t s);

}  // namespace cupss
