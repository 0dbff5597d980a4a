/*
 * cupss_hip.h -- C-ABI of the MI355X HIP layer of libcupss.
 *
 * Plain C types only (pointers, sizes, ints, floats); every entry point
 * returns 0 (CUPSS_OK) or a negative CUPSS_ERR_* code and never exits; the
 * message of the last failure on the calling thread is cupss_last_error().
 *
 * The reference has no foreign-function interface: its seam between the
 * host classes and the device is the set of extern "C" launchers in
 * inc/cupss/field_kernels.cuh:7-19 and inc/cupss/term_kernels.cuh:7-15,
 * plus the cuFFT/cuRAND/FFTW calls inside field/term (src/field.cpp,
 * src/field_init.cpp, src/term.cpp, src/term_init.cpp).  This header
 * replaces that seam at the granularity of a whole system: the C++ API in
 * include/cupss/ (evolver, field, term, parser) calls these functions, and
 * so can C, ctypes or any other FFI.  Each declaration names what it
 * replaces.
 *
 * Grids are power-of-two sx in [2, 4096], sy and sz in [1, 4096]; real
 * arrays are float[sz][sy][sx] (x fastest, idx = k*sx*sy + j*sx + i);
 * spectra are kept on the device as the Hermitian half
 * float2[sz][sy][sx/2+1] ("half spectrum", nh bins).
 */
#ifndef CUPSS_HIP_H
#define CUPSS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUPSS_ABI_VERSION 1

enum {
    CUPSS_OK = 0,
    CUPSS_ERR_INVALID = -1,     /* bad argument or id */
    CUPSS_ERR_UNSUPPORTED = -2, /* size or system outside what the HIP path implements */
    CUPSS_ERR_STATE = -3,       /* call out of order (e.g. before prepare) */
    CUPSS_ERR_HIP = -4,         /* a HIP runtime call failed */
    CUPSS_ERR_NOMEM = -5,       /* device allocation failed */
    CUPSS_ERR_BUDGET = -6       /* serving: request does not fit the HBM budget */
};

/* Operator descriptor, same layout as the reference `pres`
 * (inc/cupss/defines.h:31-39):
 * preFactor * q^(2*q2n) * (iqx)^iqx (iqy)^iqy (iqz)^iqz * |q|^(-invq). */
typedef struct cupss_pres {
    float preFactor;
    int q2n;
    int iqx;
    int iqy;
    int iqz;
    int invq;
} cupss_pres;

typedef struct cupss_system_t *cupss_system;

/* Real-space hook (replaces the call of field::callback on real_array_d /
 * real_dealiased_d, src/field.cpp:70-86): `real2` is a device float2[N]
 * array holding the field's real values in .x (.y = 0); `which` is 0 for
 * the field, 1 for its dealiased copy.  The hook may change .x. */
typedef void (*cupss_real_hook)(void *user, int field, void *real2, int which);
/* Fourier-space hook (replaces field::callbackFourier on comp_array_d,
 * src/field.cpp:48-57): `full2` is a device float2[N] array holding the
 * full C2C spectrum; the hook may change it (the change is then projected
 * onto real fields, as the reference's IFFT/normalize/FFT does). */
typedef void (*cupss_spec_hook)(void *user, int field, void *full2);
/* Device-side hooks: the fast path of the two hooks above (no counterpart
 * in the reference, whose callbacks always receive float2 arrays; for
 * boundary-condition extensions such as its examples 07 and 08).  They are
 * called with the library's own device buffers and its HIP stream and must
 * enqueue their work on that stream: no float2 copies, no host
 * synchronisation.
 *   real: `real` holds the field's real values, float[sz][sy][sx] (x
 *         fastest); which = 0 for the field (then transformed back to the
 *         spectrum), 1 for its dealiased copy; changed in place.
 *   spectral: `half2` is the field's half spectrum, float2[sz][sy][sx/2+1]
 *         (the R2C layout: bin (kx, ky, kz) at (kz*sy + ky)*(sx/2+1) + kx),
 *         changed in place.
 * They run after the update of the field, in the reference's order
 * (spectral, then real), before the host hooks above when both are set. */
typedef void (*cupss_real_hook_dev)(void *user, int field, float *real, int sx, int sy, int sz, int which,
                                    void *hip_stream);
typedef void (*cupss_spec_hook_dev)(void *user, int field, void *half2, int sx, int sy, int sz, void *hip_stream);

const char *cupss_last_error(void);
int cupss_abi_version(void);

/* evolver constructors + field/term geometry (src/evolver.cpp:13-74,
 * src/field_init.cpp:191-213): q_d = n*2*PI/(d*s), PI = 3.1415926535f. */
int cupss_system_create(int sx, int sy, int sz, float dx, float dy, float dz, float dt, cupss_system *out);
/* evolver::~evolver / field::~field / term::~term (frees every buffer) */
int cupss_system_destroy(cupss_system s);

/* evolver::createField -> field::common_constructor device allocations
 * (src/evolver.cpp:177-196, src/field_init.cpp:15-154). */
int cupss_field_create(cupss_system s, int dynamic, int *field_id);
/* field::implicit push (src/parser.cpp:728-738) */
int cupss_field_add_implicit(cupss_system s, int field, cupss_pres p);
/* field::updateParameter -> precalculateImplicit (src/field.cpp:422-431) */
int cupss_field_set_implicits(cupss_system s, int field, const cupss_pres *p, int n);
/* evolver::createTerm (src/evolver.cpp:326-362): a product of `nfactors`
 * fields (0 = constant term) with summed prefactors. */
int cupss_term_create(cupss_system s, int field, const int *factors, int nfactors, const cupss_pres *pres,
                      int npres, int *term_id);
/* term::precomputePrefactors after a parameter change (src/field.cpp:432-439) */
int cupss_term_set_prefactors(cupss_system s, int term, const cupss_pres *pres, int npres);
/* evolver::addNoise (src/evolver.cpp:162-171) + the cuRAND generator of
 * field_init.cpp:130-137: Gaussian white noise from in-kernel Philox4x32-10
 * keyed by `seed`, counted by (cell, step, field). */
int cupss_field_set_noise(cupss_system s, int field, cupss_pres amplitude, uint64_t seed);
/* field::hasCB/callback and hasCBFourier/callbackFourier (field.h:99-103);
 * null hooks switch them off. */
int cupss_field_set_hooks(cupss_system s, int field, cupss_real_hook real_hook, cupss_spec_hook spec_hook,
                          void *user);
/* The device-side hooks of a field (see cupss_real_hook_dev); null hooks
 * switch them off.  No counterpart in the reference's API. */
int cupss_field_set_device_hooks(cupss_system s, int field, cupss_real_hook_dev real_hook,
                                 cupss_spec_hook_dev spec_hook, void *user);

/* field::copyHostToDevice + toComp (src/field.cpp:332-335, 261-273):
 * `host` holds N floats (the .x of real_array). */
int cupss_field_set_real(cupss_system s, int field, const float *host);
/* field::copyRealDeviceToHost (src/field.cpp:346-348): N floats */
int cupss_field_get_real(cupss_system s, int field, float *host);
/* real_dealiased (src/field.cpp:77-83): N floats */
int cupss_field_get_dealiased(cupss_system s, int field, float *host);
/* comp_array as the half spectrum: 2*nh floats (re, im) */
int cupss_field_get_spectrum(cupss_system s, int field, float *host);
int cupss_field_set_spectrum(cupss_system s, int field, const float *host);
/* number of bins of the half spectrum */
long cupss_system_half_bins(cupss_system s);

/* prepareProblem's device part (src/evolver.cpp:114-125): dealias orders
 * (term::prepareDevice, src/term_init.cpp:108-133) and buffers.  Re-callable
 * after fields are re-initialised; it does not reset the step counter. */
int cupss_system_prepare(cupss_system s);
/* dealias order of a field (0 = not dealiased) */
int cupss_field_aliasing_order(cupss_system s, int field);

/* evolver::advanceTime x nsteps without output (src/evolver.cpp:199-226):
 * phases A-D, each field's update (setDynamic_k / setNotDynamic_k),
 * dealias, transforms and noise, all on the system's stream. */
int cupss_system_advance(cupss_system s, int nsteps);
/* Same, timed with HIP events recorded on the system's stream around the
 * launches; writes the elapsed milliseconds. */
int cupss_system_time_advance(cupss_system s, int nsteps, float *ms);
int cupss_system_synchronize(cupss_system s);
int cupss_system_get_step(cupss_system s, uint64_t *step);
int cupss_system_set_step(cupss_system s, uint64_t step);
/* use the caller's hipStream_t (null = the library's own stream) */
int cupss_system_set_stream(cupss_system s, void *hip_stream);
/* Step implementation: 0 (default) = a fused step when the system
 * qualifies, else the general step program; 1 = always the general step
 * program.  Both run on the GPU and follow the same semantics.  The fused
 * two-kernel 2D step takes any number of dynamic fields with fields
 * linear in one of them, constraint fields of products (of dynamic fields,
 * or of derived and other such constraint fields) and coefficient fields
 * without terms; the fused 3D step takes one dynamic field with fields
 * linear in it and spectral-space noise.  Neither takes field callbacks
 * (either kind) or grids outside 16..4096 x 8..4096 (x 8..1024). */
int cupss_system_set_path(cupss_system s, int mode);
/* Noise draw: 0 (default) = xi ~ N(0,1) per real-space cell, then R2C (the
 * reference's createNoise, src/field.cpp:300-316; the oracle reproduces it
 * pointwise); 1 = eta_hat drawn directly in spectral space with the same
 * law (complex Gaussian, E|eta|^2 = N, Hermitian), which saves the
 * transform.  Not in the reference. */
int cupss_system_set_noise_mode(cupss_system s, int mode);
/* 1 if cupss_system_advance would take the fused 2D step (and
 * cupss_batch_add accept the system) */
int cupss_system_fused_eligible(cupss_system s);
/* 0 = direct launches (default); K (even, >= 2) = runs of at least K steps
 * replay K steps from a hipGraph captured at the first such run.  Applies
 * to the fused path and to the general step program of noiseless systems
 * without callbacks (others keep direct launches).  No counterpart in the
 * reference (src/evolver.cpp:199-226 launches every kernel of every step). */
int cupss_system_set_graph_steps(cupss_system s, int k);
/* device bytes the system holds, including the fused path's buffers it
 * allocates at its first fused run (for admission by memory budget) */
int cupss_system_device_bytes(cupss_system s, size_t *bytes);
/* the same kernel timing for a system's own fused runs */
int cupss_system_profile_kernel(cupss_system s, int on);
int cupss_system_kernel_time(cupss_system s, float *avg_ms, double *points, char *name, int len);
/* evolver::dt change is not part of the reference API; exposed for sweeps */
int cupss_system_set_dt(cupss_system s, float dt);

/* ---- batched path ---------------------------------------------------
 * Many independent systems advanced together: per step, one row and one
 * column launch per size class of systems (and one elementwise launch for
 * systems that need no transform), optionally replayed from a hipGraph.
 * Member systems must be prepared 2D systems that qualify for the fused
 * step (cupss_system_fused_eligible; no callbacks) and must not be advanced
 * on their own while in the batch.  Each keeps its own parameters, step
 * counter and noise stream, so a system advanced in a batch follows the
 * same sequence as when advanced alone.  No counterpart in the reference
 * (which steps one evolver at a time, src/evolver.cpp:199-226). */
typedef struct cupss_batch_t *cupss_batch;
int cupss_batch_create(void *hip_stream, cupss_batch *out);
int cupss_batch_destroy(cupss_batch b);
int cupss_batch_add(cupss_batch b, cupss_system s);
/* take a system out of the batch (it keeps its state and can be advanced
 * on its own again) */
int cupss_batch_remove(cupss_batch b, cupss_system s);
int cupss_batch_size(cupss_batch b);
/* 0 = direct launches; K (even, >= 2) = capture K steps in a hipGraph */
int cupss_batch_set_graph_steps(cupss_batch b, int k);
/* 1 (default) = time the tile shapes and chunk sizes on the batch's own
 * data at its first run and keep the fastest (results are identical for
 * every choice); 0 = fixed default shapes, one chunk per size class */
int cupss_batch_set_autotune(cupss_batch b, int on);
/* text description of the schedule in use (size classes, tile shapes,
 * chunks) into buf (len bytes, NUL-terminated) */
int cupss_batch_schedule(cupss_batch b, char *buf, int len);
int cupss_batch_advance(cupss_batch b, int nsteps);
/* Kernel timing for roofline figures: with `on`, every column launch of the
 * size class with the most points records its earliest workgroup start and
 * latest workgroup end on the device's constant-rate wall clock (inside the
 * replayed graph too); cupss_batch_kernel_time returns the average duration
 * of those launches in the last run, the real points one launch covers and
 * the kernel's name. */
int cupss_batch_profile_kernel(cupss_batch b, int on);
int cupss_batch_kernel_time(cupss_batch b, float *avg_ms, double *points, char *name, int len);
int cupss_batch_time_advance(cupss_batch b, int nsteps, float *ms);
int cupss_batch_synchronize(cupss_batch b);

/* ---- serving -----------------------------------------------------------
 * Jobs (systems with a step count) arrive over time and are served with
 * continuous batching: each round admits the jobs that have arrived, in
 * arrival order, while the head of the queue fits `budget_bytes` of device
 * memory and fewer than `max_active` jobs are resident (an idle device
 * always takes the head); then every resident job advances by up to
 * `quantum` steps -- jobs that qualify for the fused step together in one
 * batch, the others on their own streams -- and finished jobs are handed
 * back and freed.  A job's system is built only at admission, by the
 * caller's `build` callback (which returns a prepared system, e.g. built
 * through the parser); `done` is called with the finished system just
 * before it is destroyed, to read results.  A job's result is the same as
 * advancing its system alone.  No counterpart in the reference. */
typedef struct cupss_server_t *cupss_server;
typedef int (*cupss_job_build)(void *user, int job, cupss_system *out);
typedef void (*cupss_job_done)(void *user, int job, cupss_system s);
int cupss_server_create(size_t budget_bytes, int quantum, int max_active, cupss_server *out);
int cupss_server_destroy(cupss_server srv);
/* arrival: seconds after cupss_server_run starts; bytes_estimate: device
 * memory the job will hold (cupss_system_device_bytes of a like system) */
int cupss_server_submit(cupss_server srv, double arrival, int nsteps, size_t bytes_estimate,
                        cupss_job_build build, cupss_job_done done, void *user, int *job);
/* serve every submitted job to completion (wall-clock arrivals) */
int cupss_server_run(cupss_server srv);
/* seconds after the start of the run; -1 where not reached */
int cupss_server_job_times(cupss_server srv, int job, double *arrival, double *admitted, double *finished);
int cupss_server_stats(cupss_server srv, double *wall_s, double *point_steps, size_t *peak_bytes,
                       int *peak_active);

/* Device buffer access for callers that bring their own kernels (e.g. the
 * reference's user callbacks that launch __global__ functions): device
 * pointer to the current half spectrum of a field. */
int cupss_field_spectrum_device(cupss_system s, int field, void **dev_ptr);

/* Stand-alone transforms on caller device buffers (plain pointers), used by
 * tests and tools: half spectrum out[nh] = R2C(in[N]); real out[N] =
 * scale * C2R(in) with dealias mask order mask_p (0 = none); `scratch` holds
 * nh float2 and must not alias `in`. */
int cupss_r2c(int sx, int sy, int sz, const float *in, void *out, void *hip_stream);
int cupss_c2r(int sx, int sy, int sz, const void *in, int mask_p, float scale, float *out, void *scratch,
              void *hip_stream);
/* cupss_c2r without its workspace argument and without any device
 * allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate (code under HIP graph capture, serving loops, harnesses
 * that forbid hipMalloc).  The workspace is a static per-device array of
 * 2^21 float2 (16 MiB, held until the process ends): half spectra of more
 * bins ((sx/2+1)*sy*sz > 2^21) are rejected with CUPSS_ERR_INVALID.  The
 * first call on a device sets up (a synchronous 64 KiB twiddle upload);
 * calls that overlap on one device share the array and race. */
int cupss_c2r_noalloc(int sx, int sy, int sz, const void *in, int mask_p, float scale, float *out,
                      void *hip_stream);
/* Philox normals of a stream (seed, step, stream) for cells [0, n). */
int cupss_normals(uint64_t seed, uint64_t step, uint32_t stream, float *out, long n, void *hip_stream);

/* Device memory helpers so that callers without a HIP toolchain (ctypes)
 * can stage data. */
int cupss_device_alloc(size_t bytes, void **dev_ptr);
int cupss_device_free(void *dev_ptr);
int cupss_memcpy_h2d(void *dev_dst, const void *host_src, size_t bytes);
int cupss_memcpy_d2h(void *host_dst, const void *dev_src, size_t bytes);
/* device-to-device copy (also the known-size reference kernel that HBM
 * counter readings are scaled by) */
int cupss_memcpy_d2d(void *dev_dst, const void *dev_src, size_t bytes);
int cupss_device_synchronize(void);
int cupss_device_count(int *n);

#ifdef __cplusplus
}
#endif
#endif
