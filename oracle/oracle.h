/*
 * oracle.h -- CPU restatement of cuPSS's per-timestep pseudospectral update.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (libcupss) never links or calls it.
 *
 * It restates the reference algorithm with the reference's own data layout
 * (full complex C2C spectra of N = sx*sy*sz bins, x fastest,
 * idx = k*sx*sy + j*sx + i) and its own step structure:
 *   - evolver::advanceTime phases A-D          src/evolver.cpp:199-226
 *   - field::updateTerms / createNoise         src/field.cpp:15-23, 300-330
 *   - field::setRHS (update, dealias, IFFT,
 *     normalize, FFT)                          src/field.cpp:25-92, 247-298
 *   - setDynamic_k / setNotDynamic_k (GPU
 *     semantics, not the buggy CPU loops)      src/field_kernels.cu:130-227
 *   - dealias_k (GPU semantics)                src/field_kernels.cu:229-256
 *   - term::update / computeProduct /
 *     copyComp / applyPres_vector              src/term.cpp:12-126, 225-255
 *   - term::prepareDevice/precomputePrefactors src/term_init.cpp:108-200
 *   - field::precalculateImplicit              src/field_init.cpp:237-284
 *   - q grid, PI                               src/field_init.cpp:191-213,
 *                                              inc/cupss/defines.h:54
 * The transforms are the project's own radix-2 FFT (power-of-two lengths)
 * with an O(n^2) DFT fallback for other lengths; twiddles are computed in
 * double precision.  Two builds exist: orc_f32_* computes in float (the
 * reference FFTW-float CPU scheme), orc_f64_* in double (the same scheme in
 * double precision).
 *
 * Noise: the reference draws xi ~ N(0,1) per real-space cell from cuRAND
 * Philox (time seeded, src/field_init.cpp:130-137).  The oracle draws xi from
 * the project's counter-based Philox4x32-10 + Box-Muller stream
 * (orc_philox4x32_10 / orc_normal), keyed by the field's seed and counted by
 * (cell/4, step); the HIP path uses the same stream, so noisy runs are
 * comparable pointwise.  The cuRAND stream itself is not reproduced.
 *
 * Pinning (see tests/test_oracle.py): FFT against an O(n^2) double DFT;
 * spectral operators against the reference's golden files
 * (tests/golden/base_truths, from tests/base_truths of the reference);
 * diffusion against the exact per-mode decay; the cubic product and dealias
 * cutoff against the closed form for cos(k x)^3; noise amplitude against the
 * closed form sqrt(dt*A/dV)*xi; the constraint-field division against hand
 * derived values; Philox against published known-answer vectors.
 */
#ifndef CUPSS_ORACLE_H
#define CUPSS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same layout as the reference's `pres` (inc/cupss/defines.h:31-39). */
typedef struct {
    float preFactor;
    int q2n;
    int iqx;
    int iqy;
    int iqz;
    int invq;
} orc_pres;

/* Philox4x32-10 (Salmon et al., SC'11): ctr[4], key[2] -> out[4]. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* N(0,1) deviate number `cell` of the stream (seed, step): counter
 * (cell/4, step_lo, step_hi, stream), Box-Muller on pairs of uniforms. */
float orc_normal(uint64_t seed, uint64_t step, uint32_t stream, uint32_t cell);
void orc_normal_fill(uint64_t seed, uint64_t step, uint32_t stream, float *out, int n);

/* Field callbacks, as the reference's field::callbackFourier / callback
 * (src/field.cpp:48-58, 70-86): `data` is the field's full spectrum (after
 * the update, before dealiasing) or its real array (after normalisation,
 * then the dealiased copy too), interleaved (re, im) of N complex values in
 * the oracle's precision; `user` is passed through. */
typedef void (*orc_hook)(void *user, void *data, int sx, int sy, int sz);

#define ORC_PLANT_PHASES 1
#define ORC_PLANT_UNDEALIAS 2
void orc_set_plant(int code);

#define ORC_DECL(PFX, REAL)                                                       \
    typedef struct PFX##sys PFX##sys;                                             \
    PFX##sys *PFX##create(int sx, int sy, int sz, float dx, float dy, float dz,   \
                          float dt);                                              \
    void PFX##destroy(PFX##sys *s);                                               \
    int PFX##add_field(PFX##sys *s, int dynamic);                                 \
    void PFX##add_implicit(PFX##sys *s, int f, orc_pres p);                       \
    int PFX##add_term(PFX##sys *s, int f, int nfac, const int *fac, int npres,    \
                      const orc_pres *pres);                                      \
    void PFX##set_noise(PFX##sys *s, int f, orc_pres amp, uint64_t seed);         \
    void PFX##set_real(PFX##sys *s, int f, const float *x);                       \
    void PFX##set_hooks(PFX##sys *s, int f, orc_hook fourier, orc_hook real,      \
                        void *user);                                              \
    void PFX##prepare(PFX##sys *s);                                               \
    void PFX##step(PFX##sys *s);                                                  \
    void PFX##steps(PFX##sys *s, int n);                                          \
    void PFX##get_real(PFX##sys *s, int f, REAL *out);                            \
    void PFX##get_comp(PFX##sys *s, int f, REAL *out);                            \
    void PFX##get_dealiased(PFX##sys *s, int f, REAL *out);                       \
    int PFX##aliasing_order(PFX##sys *s, int f);                                  \
    void PFX##fft(REAL *data, int sx, int sy, int sz, int dir);

ORC_DECL(orc_f32_, float)
ORC_DECL(orc_f64_, double)

/* O(n^2) multidimensional DFT in double: reference for the FFTs. */
void orc_dft_f64(const double *in, double *out, int sx, int sy, int sz, int dir);

/* Switch documenting the reference CPU dealias bug (src/field.cpp:220:
 * |nj| tested against sz/(p+1) instead of |nk|).  Off by default; the
 * oracle follows the GPU kernel dealias_k (src/field_kernels.cu:241-254). */
void orc_set_ref_cpu_dealias_bug(int on);

#ifdef __cplusplus
}
#endif
#endif
