// system_impl.hpp -- internal state of a cupss_system (the device side of
// one evolver) shared by system.hip and batch.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "fused.hpp"
#include "internal.hpp"

namespace cupss {

int fail(int code, const std::string &msg);

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess)                                                                          \
            return ::cupss::fail(CUPSS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
    } while (0)

struct FieldRec {
    int dynamic = 0;
    int noisy = 0;
    cupss_pres noise{};
    uint64_t seed = 0;
    std::vector<cupss_pres> imp;
    std::vector<int> terms;
    int order = 0;  // dealias order, 0 = none
    float2 *spec[2] = {nullptr, nullptr};
    int cur = 0;
    float *dreal = nullptr;
    // dealiased real copy: 0 = zero (never updated, the reference's step-0
    // state), 1 = dreal holds C2R(mask*spec), 2 = same but not yet written
    // to dreal (after fused steps), 3 = stale (spec reset by set_real after
    // updates; the reference keeps the old real_dealiased then)
    int dstate = 0;
    float2 *eta = nullptr;
    cupss_real_hook real_hook = nullptr;
    cupss_spec_hook spec_hook = nullptr;
    void *hook_user = nullptr;
    cupss_real_hook_dev real_dev = nullptr;   // device-side hooks (fast path)
    cupss_spec_hook_dev spec_dev = nullptr;
    void *dev_user = nullptr;
};

struct TermRec {
    int field = 0;
    std::vector<int> fac;
    std::vector<cupss_pres> pres;
    float2 *xhat = nullptr;
};

// Product terms of one field transformed together: M * FFT(sum_t c_t prod_t)
// (terms with equal prefactor lists, c_t = 1, or single-monomial prefactor
// lists that differ only in the coefficient, folded into c_t).
struct FusedGroup {
    int target = -1;                  // field the terms belong to
    std::vector<int> terms;
    std::vector<float> coef;
    std::vector<cupss_pres> pres;     // the group's multiplier
};

struct FusedPlan {
    bool ok = false;
    int phi = -1;
    std::vector<int> real;            // fields read in real space by the products
    std::vector<FusedGroup> groups;   // product groups, phi's first, then by stored field
    std::vector<int> derived;         // constraint fields psi = M_psi u, u dynamic
    std::vector<int> dbase;           // u of each derived field
    std::vector<int> cons;            // stored fields: the other dynamic fields (nxdyn), the fresh
                                      // constraints (nfresh), then the other constraints
    int nxdyn = 0;
    int nfresh = 0;
};

}  // namespace cupss

struct cupss_system_t {
    cupss::Geom g{};
    float dx = 1, dy = 1, dz = 1, dt = 0;
    std::vector<cupss::FieldRec> f;
    std::vector<cupss::TermRec> t;
    float2 *scratch = nullptr;
    float *rtmp = nullptr;
    float2 *full = nullptr;   // float2[N] for hooks
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool prepared = false;
    int path_mode = 0;        // 0 = fused step when eligible, 1 = general step only
    int noise_mode = 0;       // 0 = real-space Philox + R2C, 1 = spectral-space Philox
    cupss::FusedPlan plan;
    float2 *fz[cupss::FUSED_MAX_REAL] = {};
    float2 *fx[cupss::FUSED_MAX_GROUPS] = {};
    float2 *ft[2] = {nullptr, nullptr};  // phi's spectrum, column-major, during fused runs
    float2 *fc[cupss::FUSED_MAX_CONS][2] = {};  // stored constraint fields, column-major
    float2 *fw[cupss::FUSED_MAX_REAL] = {};     // 3D: IFFT_z of the masked real fields (zero off the kept lines)
    // After a fused run, fc[.][cont_buf] holds the stored fields' values of
    // the next step's phase A/B; a run that starts at step cont_step with
    // generation cont_gen continues from them.  Anything that changes the
    // state in between (general steps, set_real, ...) resets cont_step.
    uint64_t cont_step = ~0ull;
    uint64_t cont_gen = 0;
    int cont_buf = 0;
    bool cont_now = false;                        // the current run continues
    // 2D fused path: after a run, ft[cont_ft] holds phi column-major and Z_f
    // the real fields of that state; a run that starts at cont_step with
    // generation cont_gen resumes from them (no transpose in, no prologue)
    int cont_ft = 0;
    bool resume_now = false;
    long fused_steps = 0;     // steps taken by the fused path (diagnostics)
    uint64_t step = 0;
    int device = 0;
    void *batch = nullptr;    // internal one-instance fused batch
    // The fused batch (internal or a cupss_batch's) whose last run left this
    // system's state column-major only: its transpose back to the natural
    // layout is deferred until something reads or writes the state outside
    // such a run (fused_flush), so runs that follow runs pay none.
    void *pending = nullptr;
    // Member of a cupss_batch: the system's work goes to the batch's stream
    // while it is a member (ordered with the batch's runs without per-run
    // synchronisation); its own stream is kept here and restored on removal.
    void *member_of = nullptr;
    hipStream_t home_stream = nullptr;
    uint64_t gen = 1;         // bumped by every change that alters the step program
    uint64_t vgen = 1;        // bumped by prefactor value changes that keep the program's structure
    // general path: K steps captured once in a hipGraph and replayed
    // (noiseless systems without hooks; K even so every field's spectrum
    // returns to its first buffer)
    int graph_steps = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint64_t graph_gen = 0;
    uint64_t graph_vgen = 0;
    hipStream_t graph_stream = nullptr;
    std::vector<int> graph_cur;  // each field's buffer parity the graph was captured at
};

namespace cupss {
int build_update(cupss_system s, int fi, UpdateArgs &a);
int materialize(cupss_system s, int fi);
int general_steps(cupss_system s, int nsteps);
// phase A/B of the reference step (constraint fields) with the general program
int general_constraint_phase(cupss_system s);
// the fused plan's noise can be drawn as the noise mode asks
bool fused_noise_ok(cupss_system s);
// a dynamic field of the fused plan holds a stale dealiased copy
bool fused_stale(cupss_system s);
// fused path (batch.hip)
int fused_steps(cupss_system s, int nsteps);
void fused_release(cupss_system s);
// write back the natural-layout state that a fused run deferred (see
// cupss_system_t::pending); no-op when nothing is pending
int fused_flush(cupss_system s);
int fused_profile(cupss_system s, int on);
int fused_kernel_time(cupss_system s, float *avg_ms, double *points, char *name, int len);
}  // namespace cupss
