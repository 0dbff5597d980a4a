// internal.hpp -- structures shared by the host step program and the
// kernels of libcupss (not part of the public C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cupss_hip.h"

namespace cupss {

constexpr int MAX_FACTORS = 8;     // fields in one product term
constexpr int MAX_TERM_REFS = 16;  // terms of one field
constexpr int MAX_PRES = 64;       // prefactors over all terms of one field
constexpr int MAX_IMP = 16;        // implicit (left-hand side) operators of one field

// Grid geometry and the reference q convention (src/field_init.cpp:207-213):
// q_d(i) = (i < (s+1)/2 ? i : i-s) * stepq_d, stepq_d = 2*PI/(d*s) in float.
struct Geom {
    int sx, sy, sz;
    int nc;     // sx/2: complex points per row in the R2C pass
    int nc1;    // sx/2+1: row length of the half spectrum
    long nh;    // nc1*sy*sz: bins of the half spectrum
    long n;     // sx*sy*sz: real points
    float sqx, sqy, sqz;
};

// One prefactor (monomial) of a term with its sign folded in:
// c * q^(2 q2n) * qx^a * qy^b * qz^c * |q|^(-invq)   (S2 of term_init.cpp:135-200)
struct PresDev {
    float c;
    int q2n, a, b, cz, invq;
};

enum TermKind : int {
    TERM_BUFFER = 0,  // X = R2C(product) in a half-spectrum buffer
    TERM_CONST = 1,   // empty product: X = N at q = 0 (term_kernels.cu:58-63)
    TERM_FIELD = 2,   // single field: X = that field's spectrum (copyComp, term.cpp:104-126)
    TERM_GROUPED = 3, // product evaluated by the fused step as part of a group
};

struct TermRefDev {
    const float2 *X;
    int kind;
    int mult_i;   // multiply_by_i_pre (term_init.cpp:143-149)
    int pres0;    // first prefactor in UpdateArgs::pres
    int npres;
};

// Everything the semi-implicit update of one field needs
// (setDynamic_k / setNotDynamic_k, src/field_kernels.cu:130-227).
struct UpdateArgs {
    Geom g;
    float2 *out;
    const float2 *cur;
    const float2 *noise;   // eta_hat = R2C(xi), or null
    int dynamic;
    int nterms;
    int nimp;
    int noisy;
    float dt;
    float sdt;            // 1/sqrt(dt)
    float noise_base;     // sqrt(dt/(dx dy dz) * A)   (field_init.cpp:269)
    int noise_q2n;
    int noise_invq;
    TermRefDev terms[MAX_TERM_REFS];
    PresDev pres[MAX_PRES];
    PresDev imp[MAX_IMP];
};

struct MaskArgs {
    int p;   // aliasing order; 0 = no mask
};

// --- launchers (kernels.hip) ---------------------------------------------

// Twiddle table (device) for the current device, created on first use:
// exp(-2 pi i j / TW_N), j < TW_N.
const float2 *twiddles(hipStream_t s);
// the TW_N-entry table twiddles() uploads, computed on the host
void twiddle_table(float2 *host_out);
// All power-of-two tables: table M (exp(-2 pi i j / M), j < M) at offset M,
// M = 1, 2, 4, ..., TW_N.
const float2 *twiddle_sets(hipStream_t s);

enum LoadKind { LOAD_REAL = 0, LOAD_PRODUCT = 1, LOAD_NOISE = 2 };
struct R2CSource {
    int kind;
    const float *f[MAX_FACTORS];
    int nf;
    uint64_t seed;
    uint64_t step;
    uint32_t stream;
};

// half spectrum out[nh] <- R2C of the real source over the whole grid
hipError_t r2c(const Geom &g, const R2CSource &src, float2 *out, hipStream_t s);
// real out[n] <- scale * C2R(mask_p(in)); scratch[nh] is clobbered (may not alias in)
hipError_t c2r(const Geom &g, const float2 *in, int mask_p, float scale, float *out, float2 *scratch,
               hipStream_t s);
// the same with the caller's twiddle table (TW_N entries, twiddle_table())
hipError_t c2r_tw(const Geom &g, const float2 *in, int mask_p, float scale, float *out, float2 *scratch,
                  hipStream_t s, const float2 *tw);
// the update of one field, then the Hermitian projection of its kx = 0 and
// kx = sx/2 planes (hermitian_planes)
hipError_t update(const UpdateArgs &a, hipStream_t s);
// bins (kx, ky, kz), (kx, -ky, -kz) of the planes kx = 0, sx/2 -> their
// Hermitian mean, in place
hipError_t hermitian_planes(const Geom &g, float2 *spec, hipStream_t s);
// unnormalised C2C transforms (dir -1 forward, +1 inverse) of the lines along
// y of a natural half spectrum with kx < ncols: line (kx, kz) holds
// in[(kz*sy + ky)*nc1 + kx], ky < sy; out may be in
hipError_t lines_y(const Geom &g, int dir, const float2 *in, float2 *out, long ncols, hipStream_t s);
// eta_hat with the law of R2C(xi), xi ~ N(0,1) per cell, from Philox
hipError_t spectral_noise_fill(const Geom &g, float2 *eta, uint64_t seed, uint64_t step, uint32_t stream,
                               hipStream_t s);

}  // namespace cupss
