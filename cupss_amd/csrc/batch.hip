// batch.hip -- many independent systems advanced together by the fused
// two-kernel step (fused.hip): the batched path of cuPSS workloads such as
// ensembles and parameter sweeps, and (as a batch of one) the fused path of
// a single system.
//
// Instances are grouped into size classes (equal sx/2 and sy).  Per step:
//   for each class: one row launch, one column launch over the rows /
//   columns of all its instances (a block table maps each workgroup to an
//   instance and a first line; consecutive workgroup ids are dealt to the
//   XCDs so that each XCD sweeps a contiguous range of lines);
//   one elementwise launch for all instances whose step needs no transform
//   (no products, noise drawn in spectral space or none).
// Before the first step of a run a prologue column launch writes the real
// fields of the current state.  A run of K steps can be captured once in a
// hipGraph (K even, so the spectra return to their first buffer) and
// replayed; a small kernel inside the graph advances the step counters
// that key the noise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "fft_core.hpp"
#include "fused.hpp"
#include "system_impl.hpp"

using namespace cupss;

namespace cupss {

__global__ void k_bump_steps(FusedInst *inst, int n, uint32_t by)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inst[i].step_base += by;
}

struct FusedBatch {
    std::vector<cupss_system> sys;
    hipStream_t stream = nullptr;
    bool own_stream = false;

    struct Class {
        int nc = 0, ny = 0;
        bool rows = false;      // any instance needs the row pass
        bool real = false;      // any instance has real fields (prologue)
        std::vector<int> insts;
        FusedBlock *rb = nullptr;
        int nrb = 0;
        FusedBlock *cb = nullptr;
        int ncb = 0;
    };
    std::vector<Class> classes;
    std::vector<int> elem;      // elementwise instances
    FusedBlock *eb = nullptr;
    int neb = 0;
    FusedBlock *db = nullptr;   // rows of instances with derived fields
    int ndb = 0;
    hipEvent_t done = nullptr;  // recorded after each run; member streams wait on it

    std::vector<FusedInst> h;
    FusedInst *d = nullptr;
    size_t dcap = 0;
    FusedInst *pinned = nullptr;  // staging of the descriptor upload
    bool built = false;

    int graph_steps = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;

    long steps_run = 0;
    std::vector<uint64_t> gens;   // system generations the tables were built for

    ~FusedBatch() { release(); }

    void release_tables()
    {
        for (auto &c : classes) {
            hipFree(c.rb);
            hipFree(c.cb);
        }
        classes.clear();
        hipFree(eb);
        eb = nullptr;
        neb = 0;
        hipFree(db);
        db = nullptr;
        ndb = 0;
        elem.clear();
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
        built = false;
    }

    void release()
    {
        if (stream) hipStreamSynchronize(stream);
        release_tables();
        hipFree(d);
        d = nullptr;
        hipHostFree(pinned);
        pinned = nullptr;
        dcap = 0;
        if (done) hipEventDestroy(done);
        done = nullptr;
        if (own_stream && stream) hipStreamDestroy(stream);
        stream = nullptr;
    }
};

// device descriptor of system s (its fused plan must be ok)
static int describe(cupss_system s, FusedInst &a)
{
    const FusedPlan &P = s->plan;
    const Geom &g = s->g;
    FieldRec &F = s->f[P.phi];
    for (size_t i = 0; i < P.real.size(); i++)
        if (!s->fz[i] && hipMalloc(&s->fz[i], sizeof(float2) * g.nh) != hipSuccess)
            return fail(CUPSS_ERR_NOMEM, "device allocation failed");
    for (size_t i = 0; i < P.groups.size(); i++)
        if (!s->fx[i] && hipMalloc(&s->fx[i], sizeof(float2) * g.nh) != hipSuccess)
            return fail(CUPSS_ERR_NOMEM, "device allocation failed");
    std::memset(&a, 0, sizeof(a));
    a.g = g;
    a.spec[0] = F.spec[F.cur];
    a.spec[1] = F.spec[F.cur ^ 1];
    a.eta = F.eta;
    a.step_base = s->step;
    build_update(s, P.phi, a.u);
    a.u.noise = F.eta;
    a.ngroups = (int)P.groups.size();
    for (size_t gi = 0; gi < P.groups.size(); gi++) {
        a.X[gi] = s->fx[gi];
        for (size_t k = 0; k < F.terms.size(); k++)
            for (int ti : P.groups[gi])
                if (F.terms[k] == ti) {
                    a.u.terms[k].kind = TERM_GROUPED;
                    if (ti == P.groups[gi][0]) a.gterm[gi] = (int)k;
                }
    }
    a.nreal = (int)P.real.size();
    int lp = 0;
    for (size_t i = 0; i < P.real.size(); i++) {
        const int f = P.real[i];
        a.p[i] = s->f[f].order;
        a.Z[i] = s->fz[i];
        a.zero[i] = (f == P.phi && F.dstate == 0) ? 1 : 0;
        if (f != P.phi) {
            const TermRec &T = s->t[s->f[f].terms[0]];
            const int m = (T.pres[0].iqx + T.pres[0].iqy + T.pres[0].iqz) % 2;
            a.lmul[i].mult_i = m;
            a.lmul[i].pres0 = lp;
            a.lmul[i].npres = (int)T.pres.size();
            for (const cupss_pres &p : T.pres) {
                const int iu = p.iqx + p.iqy + p.iqz;
                PresDev &D = a.lpres[lp++];
                D.c = p.preFactor * (float)(-2 * (((iu - m) / 2) % 2) + 1);
                D.q2n = p.q2n;
                D.a = p.iqx;
                D.b = p.iqy;
                D.cz = p.iqz;
                D.invq = p.invq;
            }
        }
    }
    a.scale = 1.0f / (float)g.n;
    int np = 0;
    for (size_t gi = 0; gi < P.groups.size(); gi++) {
        a.g0[gi] = np;
        for (int ti : P.groups[gi]) {
            const TermRec &T = s->t[ti];
            a.nfac[np] = (int)T.fac.size();
            for (size_t c = 0; c < T.fac.size(); c++)
                a.fac[np][c] = (int)(std::find(P.real.begin(), P.real.end(), T.fac[c]) - P.real.begin());
            np++;
        }
    }
    a.g0[P.groups.size()] = np;
    a.noise_mode = !F.noisy ? NOISE_NONE : (s->noise_mode == 1 ? NOISE_SPECTRAL : NOISE_REAL);
    if ((int)P.derived.size() > FUSED_MAX_DERIVED) return fail(CUPSS_ERR_UNSUPPORTED, "too many derived fields");
    int dp = 0;
    a.nder = (int)P.derived.size();
    for (size_t d = 0; d < P.derived.size(); d++) {
        const FieldRec &D = s->f[P.derived[d]];
        const TermRec &T = s->t[D.terms[0]];
        const int m = (T.pres[0].iqx + T.pres[0].iqy + T.pres[0].iqz) % 2;
        a.dout[d] = D.spec[D.cur];
        a.dmul[d].mult_i = m;
        a.dmul[d].pres0 = dp;
        a.dmul[d].npres = (int)T.pres.size();
        for (const cupss_pres &p : T.pres) {
            if (dp >= FUSED_MAX_LPRES) return fail(CUPSS_ERR_UNSUPPORTED, "too many derived prefactors");
            const int iu = p.iqx + p.iqy + p.iqz;
            PresDev &Q = a.dpres[dp++];
            Q.c = p.preFactor * (float)(-2 * (((iu - m) / 2) % 2) + 1);
            Q.q2n = p.q2n;
            Q.a = p.iqx;
            Q.b = p.iqy;
            Q.cz = p.iqz;
            Q.invq = p.invq;
        }
    }
    a.key0 = (uint32_t)F.seed;
    a.key1 = (uint32_t)(F.seed >> 32);
    a.stream = (uint32_t)P.phi;
    return CUPSS_OK;
}

static bool is_elementwise(const FusedInst &a) { return a.ngroups == 0 && a.nreal == 0 && a.noise_mode != NOISE_REAL; }

// table of (instance, first line) items in XCD-aware order
static int upload_blocks(const std::vector<FusedBlock> &items, FusedBlock **dst, int *n)
{
    const long nb = xcd_blocks((long)items.size());
    std::vector<FusedBlock> t(nb, FusedBlock{-1, 0});
    for (long b = 0; b < nb; b++) {
        const long it = (b & 7) * (nb >> 3) + (b >> 3);
        if (it < (long)items.size()) t[b] = items[it];
    }
    if (hipMalloc(dst, sizeof(FusedBlock) * std::max<long>(nb, 1)) != hipSuccess)
        return fail(CUPSS_ERR_NOMEM, "device allocation failed");
    if (nb > 0 && hipMemcpy(*dst, t.data(), sizeof(FusedBlock) * nb, hipMemcpyHostToDevice) != hipSuccess)
        return fail(CUPSS_ERR_HIP, "hipMemcpy failed");
    *n = (int)nb;
    return CUPSS_OK;
}

static int build(FusedBatch &B)
{
    B.release_tables();
    B.h.resize(B.sys.size());
    for (size_t i = 0; i < B.sys.size(); i++)
        if (int r = describe(B.sys[i], B.h[i])) return r;
    if (B.dcap < B.h.size()) {
        HIP_TRY(hipStreamSynchronize(B.stream));
        hipFree(B.d);
        hipHostFree(B.pinned);
        B.d = nullptr;
        B.pinned = nullptr;
        const size_t n = std::max<size_t>(B.h.size(), 1);
        if (hipMalloc(&B.d, sizeof(FusedInst) * n) != hipSuccess ||
            hipHostMalloc(&B.pinned, sizeof(FusedInst) * n, hipHostMallocDefault) != hipSuccess)
            return fail(CUPSS_ERR_NOMEM, "device allocation failed");
        B.dcap = B.h.size();
    }
    std::map<std::pair<int, int>, int> cls;
    for (size_t i = 0; i < B.h.size(); i++) {
        const FusedInst &a = B.h[i];
        if (is_elementwise(a)) {
            B.elem.push_back((int)i);
            continue;
        }
        auto key = std::make_pair(a.g.nc, a.g.sy);
        if (!cls.count(key)) {
            cls[key] = (int)B.classes.size();
            B.classes.emplace_back();
            B.classes.back().nc = a.g.nc;
            B.classes.back().ny = a.g.sy;
        }
        FusedBatch::Class &c = B.classes[cls[key]];
        c.insts.push_back((int)i);
        c.rows = true;
        c.real = c.real || a.nreal > 0;
    }
    for (auto &c : B.classes) {
        std::vector<FusedBlock> rows, cols;
        const int lpb = fused_rows_per_block(c.nc);
        const int twd = fused_cols_per_block(c.ny);
        for (int i : c.insts) {
            const Geom &g = B.h[i].g;
            for (int r0 = 0; r0 < g.sy * g.sz; r0 += lpb) rows.push_back(FusedBlock{i, r0});
            for (int c0 = 0; c0 < g.nc1; c0 += twd) cols.push_back(FusedBlock{i, c0});
        }
        if (int r = upload_blocks(rows, &c.rb, &c.nrb)) return r;
        if (int r = upload_blocks(cols, &c.cb, &c.ncb)) return r;
    }
    std::vector<FusedBlock> er, dr;
    for (int i : B.elem) {
        const Geom &g = B.h[i].g;
        for (int r0 = 0; r0 < g.sy * g.sz; r0++) er.push_back(FusedBlock{i, r0});
    }
    for (size_t i = 0; i < B.h.size(); i++)
        if (B.h[i].nder > 0)
            for (int r0 = 0; r0 < B.h[i].g.sy * B.h[i].g.sz; r0++) dr.push_back(FusedBlock{(int)i, r0});
    if (int r = upload_blocks(er, &B.eb, &B.neb)) return r;
    if (int r = upload_blocks(dr, &B.db, &B.ndb)) return r;
    B.gens.clear();
    for (cupss_system s : B.sys) B.gens.push_back(s->gen);
    B.built = true;
    return CUPSS_OK;
}

static bool stale_tables(const FusedBatch &B)
{
    if (!B.built || B.gens.size() != B.sys.size()) return true;
    for (size_t i = 0; i < B.sys.size(); i++)
        if (B.sys[i]->gen != B.gens[i]) return true;
    return false;
}

// refresh the per-run parts of the descriptors (buffers, step counters,
// step-0 flags) and upload them; the previous run must be finished with the
// device table and the staging buffer before either is rewritten
static int refresh(FusedBatch &B)
{
    HIP_TRY(hipStreamSynchronize(B.stream));
    for (size_t i = 0; i < B.sys.size(); i++) {
        cupss_system s = B.sys[i];
        FieldRec &F = s->f[s->plan.phi];
        FusedInst &a = B.h[i];
        a.spec[0] = F.spec[F.cur];
        a.spec[1] = F.spec[F.cur ^ 1];
        a.step_base = s->step;
        for (int k = 0; k < a.nreal; k++) a.zero[k] = (s->plan.real[k] == s->plan.phi && F.dstate == 0) ? 1 : 0;
        for (int d = 0; d < a.nder; d++) {
            FieldRec &D = s->f[s->plan.derived[d]];
            a.dout[d] = D.spec[D.cur];
        }
    }
    std::memcpy(B.pinned, B.h.data(), sizeof(FusedInst) * B.h.size());
    HIP_TRY(hipMemcpyAsync(B.d, B.pinned, sizeof(FusedInst) * B.h.size(), hipMemcpyHostToDevice, B.stream));
    return CUPSS_OK;
}

static int launch_prologue(FusedBatch &B)
{
    for (auto &c : B.classes)
        if (c.real) HIP_TRY(fused_cols_launch(c.ny, B.d, c.cb, c.ncb, 0, 0, 0, B.stream));
    return CUPSS_OK;
}

static int launch_step(FusedBatch &B, int parity, uint32_t rel)
{
    for (auto &c : B.classes) {
        if (c.rows) HIP_TRY(fused_rows_launch(c.nc, B.d, c.rb, c.nrb, rel, B.stream));
        HIP_TRY(fused_cols_launch(c.ny, B.d, c.cb, c.ncb, parity, rel, 1, B.stream));
    }
    if (B.neb > 0) HIP_TRY(fused_elem_launch(B.d, B.eb, B.neb, parity, rel, B.stream));
    return CUPSS_OK;
}

// after a run: the general path's view of every system.  Each derived field
// holds P(M_psi phi) of the last step's starting phi, i.e. of the spectrum
// buffer the last step read (parity (nsteps-1) & 1); all of them are
// written by one launch.  Member streams then wait for the batch stream.
static int settle(FusedBatch &B, int nsteps)
{
    if (B.ndb > 0) HIP_TRY(fused_derived_launch(B.d, B.db, B.ndb, (nsteps - 1) & 1, B.stream));
    if (!B.done) HIP_TRY(hipEventCreateWithFlags(&B.done, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(B.done, B.stream));
    for (cupss_system s : B.sys) {
        const FusedPlan &P = s->plan;
        FieldRec &F = s->f[P.phi];
        F.cur ^= (nsteps & 1);
        s->step += (uint64_t)nsteps;
        s->fused_steps += nsteps;
        if (F.order > 0) F.dstate = 2;
        for (int dfi : P.derived)
            if (s->f[dfi].order > 0) s->f[dfi].dstate = 2;
        if (s->stream != B.stream) HIP_TRY(hipStreamWaitEvent(s->stream, B.done, 0));
    }
    return CUPSS_OK;
}

static int run(FusedBatch &B, int nsteps)
{
    if (nsteps <= 0) return CUPSS_OK;
    for (cupss_system s : B.sys)
        if (s->stream != B.stream) HIP_TRY(hipStreamSynchronize(s->stream));
    if (stale_tables(B))
        if (int r = build(B)) return r;
    if (int r = refresh(B)) return r;
    if (!twiddle_sets(B.stream)) return fail(CUPSS_ERR_NOMEM, "twiddle tables");
    if (int r = launch_prologue(B)) return r;
    int done = 0;
    const int K = B.graph_steps;
    if (K >= 2 && nsteps >= K) {
        if (!B.gexec) {
            HIP_TRY(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
            int rc = CUPSS_OK;
            for (int n = 0; n < K && rc == CUPSS_OK; n++) rc = launch_step(B, n & 1, (uint32_t)n);
            if (rc == CUPSS_OK) {
                const int ni = (int)B.h.size();
                hipLaunchKernelGGL(k_bump_steps, dim3((ni + 255) / 256), dim3(256), 0, B.stream, B.d, ni,
                                   (uint32_t)K);
            }
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamEndCapture(B.stream, &g));
            if (rc != CUPSS_OK) {
                hipGraphDestroy(g);
                return rc;
            }
            B.graph = g;
            HIP_TRY(hipGraphInstantiate(&B.gexec, g, nullptr, nullptr, 0));
        }
        while (nsteps - done >= K) {
            HIP_TRY(hipGraphLaunch(B.gexec, B.stream));
            done += K;
        }
        // the remainder continues from the bumped counters
        for (size_t i = 0; i < B.h.size(); i++) B.h[i].step_base += (uint64_t)done;
    }
    for (int n = done; n < nsteps; n++)
        if (int r = launch_step(B, n & 1, (uint32_t)(n - done))) return r;
    B.steps_run += nsteps;
    return settle(B, nsteps);
}

int fused_steps(cupss_system s, int nsteps)
{
    FusedBatch *B = static_cast<FusedBatch *>(s->batch);
    if (!B) {
        B = new FusedBatch();
        B->sys.push_back(s);
        B->stream = s->stream;
        s->batch = B;
    }
    if (B->stream != s->stream) {
        B->stream = s->stream;
        B->built = false;
    }
    return run(*B, nsteps);
}

void fused_release(cupss_system s)
{
    delete static_cast<FusedBatch *>(s->batch);
    s->batch = nullptr;
}

}  // namespace cupss

// ------------------------------------------------------------ batch C-ABI

struct cupss_batch_t {
    FusedBatch b;
};

extern "C" int cupss_batch_create(void *hip_stream, cupss_batch *out)
{
    if (!out) return fail(CUPSS_ERR_INVALID, "null out");
    cupss_batch b = new cupss_batch_t();
    if (hip_stream) {
        b->b.stream = (hipStream_t)hip_stream;
    } else {
        if (hipStreamCreateWithFlags(&b->b.stream, hipStreamNonBlocking) != hipSuccess) {
            delete b;
            return fail(CUPSS_ERR_HIP, "hipStreamCreate failed");
        }
        b->b.own_stream = true;
    }
    *out = b;
    return CUPSS_OK;
}

extern "C" int cupss_batch_destroy(cupss_batch b)
{
    delete b;
    return CUPSS_OK;
}

extern "C" int cupss_batch_add(cupss_batch b, cupss_system s)
{
    if (!b || !s) return fail(CUPSS_ERR_INVALID, "null batch or system");
    if (!s->prepared) return fail(CUPSS_ERR_STATE, "prepare the system before adding it to a batch");
    if (!s->plan.ok) return fail(CUPSS_ERR_UNSUPPORTED, "system does not qualify for the fused batched step");
    for (auto &F : s->f)
        if (F.real_hook || F.spec_hook) return fail(CUPSS_ERR_UNSUPPORTED, "systems with callbacks cannot be batched");
    if (s->f[s->plan.phi].dstate == 3) {
        // consume a stale dealiased copy with one general step, as a lone system would
        if (int r = general_steps(s, 1)) return r;
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    b->b.sys.push_back(s);
    b->b.built = false;
    return CUPSS_OK;
}

extern "C" int cupss_batch_size(cupss_batch b) { return b ? (int)b->b.sys.size() : 0; }

extern "C" int cupss_batch_set_graph_steps(cupss_batch b, int k)
{
    if (!b) return fail(CUPSS_ERR_INVALID, "null batch");
    if (k != 0 && (k < 2 || (k & 1))) return fail(CUPSS_ERR_INVALID, "graph steps must be 0 or an even number >= 2");
    b->b.graph_steps = k;
    b->b.built = false;
    return CUPSS_OK;
}

extern "C" int cupss_batch_advance(cupss_batch b, int nsteps)
{
    if (!b) return fail(CUPSS_ERR_INVALID, "null batch");
    return run(b->b, nsteps);
}

extern "C" int cupss_batch_time_advance(cupss_batch b, int nsteps, float *ms)
{
    if (!b) return fail(CUPSS_ERR_INVALID, "null batch");
    if (stale_tables(b->b))
        if (int r = build(b->b)) return r;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, b->b.stream));
    int r = run(b->b, nsteps);
    HIP_TRY(hipEventRecord(e1, b->b.stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (ms) *ms = t;
    return r;
}

extern "C" int cupss_batch_synchronize(cupss_batch b)
{
    if (!b) return fail(CUPSS_ERR_INVALID, "null batch");
    HIP_TRY(hipStreamSynchronize(b->b.stream));
    return CUPSS_OK;
}
