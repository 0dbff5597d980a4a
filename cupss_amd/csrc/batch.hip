// batch.hip -- many independent systems advanced together by the fused
// two-kernel step (fused.hip): the batched path of cuPSS workloads such as
// ensembles and parameter sweeps, and (as a batch of one) the fused path of
// a single system.
//
// Instances are grouped into size classes (equal sx/2 and sy); those whose
// step needs no transform form the elementwise set.  A run of n steps is
//   prologue:   tiled transpose of each transform instance's spectrum into
//               its column-major buffer, column launch (mode 0) writing the
//               real fields of the current state;
//   steps:      per class and chunk of instances, K x (row launch, column
//               launch); one elementwise launch advancing its instances K
//               steps in registers;
//   settle:     transpose back, derived fields.
// Chunks: the instances of a class are split so that the buffers one chunk
// touches (two spectra, Z_f, X_g, eta) stay resident in the last-level
// cache for the K steps the chunk takes before the next chunk starts.
// Instances never interact, so the order changes no result.  K steps can be
// captured once in a hipGraph (K even, so the spectra return to their first
// buffer) and replayed; a small kernel inside the graph advances the step
// counters that key the noise.
//
// Tile shapes (rows per row workgroup, columns per column workgroup) and
// the chunk size are picked per batch by timing the candidates on the
// batch's own data (autotune); every candidate gives bit-identical results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "fft_core.hpp"
#include "fused.hpp"
#include "system_impl.hpp"

using namespace cupss;

namespace cupss {

// End of a run of `by` steps: the step counters advance, and after an odd
// count each double buffer (phi's spectrum, the stored fields) swaps its
// halves, so that [0] is again the buffer holding the current state.  The
// host applies the same changes to its copies (mirror_run_end), so the next
// run finds the descriptors unchanged and uploads nothing for them.
__global__ void k_bump_steps(FusedInst *inst, int n, uint32_t by)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    FusedInst &a = inst[i];
    a.step_base += by;
    if (by & 1) {
        float2 *t = a.spec[0];
        a.spec[0] = a.spec[1];
        a.spec[1] = t;
        for (int c = 0; c < a.ncons; c++) {
            t = a.cs[c][0];
            a.cs[c][0] = a.cs[c][1];
            a.cs[c][1] = t;
        }
    }
}

// the host side of k_bump_steps on the descriptor copies
static void mirror_run_end(std::vector<FusedInst> &h, uint32_t by)
{
    for (FusedInst &a : h) {
        a.step_base += by;
        if (by & 1) {
            std::swap(a.spec[0], a.spec[1]);
            for (int c = 0; c < a.ncons; c++) std::swap(a.cs[c][0], a.cs[c][1]);
        }
    }
}

struct Tables {
    FusedBlock *rb = nullptr;
    int nrb = 0;
    FusedBlock *cb = nullptr;
    int ncb = 0;
};

struct FusedBatch {
    std::vector<cupss_system> sys;
    hipStream_t stream = nullptr;
    bool own_stream = false;

    struct Class {
        int nc = 0, ny = 0;
        bool real = false;      // any instance has real fields (prologue)
        int width = 3;          // real fields the row kernel holds at once (fused_row_width)
        int cons = 0;           // stored fields of the instances: 1 = constraints, 2 = dynamic fields too
        std::vector<int> insts;
        int lpb = 0, twd = 0;   // tile shapes
        std::vector<Tables> chunks;
        // whole steps of an instance in one workgroup (fused_small_class):
        // one launch per run of k steps over these instance ids
        bool small = false;
        int *ids = nullptr;
    };
    std::vector<Class> classes;
    std::vector<int> elem;      // elementwise instances
    std::vector<char> is_elem;  // per instance
    FusedBlock *eb = nullptr;   // elementwise: FUSED_ELEM_BLOCK pairs of bins per block
    int neb = 0;
    FusedBlock *tb = nullptr;   // transpose tiles of the transform instances
    int ntb = 0;
    hipEvent_t done = nullptr;  // recorded after each run; member streams wait on it
    hipEvent_t staged = nullptr;  // recorded after each descriptor upload

    std::vector<FusedInst> h;
    FusedInst *d = nullptr;
    size_t dcap = 0;
    FusedInst *pinned = nullptr;  // staging of the descriptor upload
    std::vector<char> dirty;      // descriptors that differ from the device's copy
    bool built = false;

    // the last run's transpose back to the natural layout, deferred (see
    // settle and flush)
    bool settle_pending = false;

    bool autotune = true;
    bool tuned = false;
    double chunk_bytes = 0;       // chunk working-set target (0: one chunk per class)

    int graph_steps = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;

    // kernel timing: every column launch of the class with the most points
    // (one chunk) records its earliest workgroup start and latest end on the
    // device wall clock, inside the timed runs themselves (and the replayed
    // graphs); the last run's launches are read back
    bool profile = false;
    int pclass = -1;
    unsigned long long *pts = nullptr;  // timestamp slots of up to PROF_SLOTS timed launches
    int pcount = 0;
    int psteps = 1;  // steps one timed launch takes (a small-instance launch takes a run's k)
    int gpcount = 0, gpsteps = 1;  // the same for one replay of the captured graph
    static constexpr int PROF_SLOTS = 64;

    // independent work (size classes, the elementwise set) runs on up to
    // MAX_LANES streams: the batch stream and side streams forked from it
    // and joined back with events (inside captured graphs as well)
    static constexpr int MAX_LANES = 4;
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t> ljoin;
    hipEvent_t lfork = nullptr;

    long steps_run = 0;
    std::vector<uint64_t> gens;   // system generations the tables were built for
    std::vector<uint64_t> vgens;  // prefactor-value generations the descriptors hold

    // 3D path (the batch of one of a 3D system): h[0] is the descriptor on
    // the device, twd3 the z-line tile width
    bool d3 = false;
    int twd3 = 0;

    ~FusedBatch() { release(); }

    void drop_graph()
    {
        if (gexec) hipGraphExecDestroy(gexec);
        if (graph) hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }

    static void free_tables(std::vector<Tables> &v)
    {
        for (auto &t : v) {
            hipFree(t.rb);
            hipFree(t.cb);
        }
        v.clear();
    }

    void release_tables()
    {
        if (stream) hipStreamSynchronize(stream);  // launches in flight may still read the tables
        for (auto &c : classes) {
            free_tables(c.chunks);
            hipFree(c.ids);
        }
        classes.clear();
        hipFree(eb);
        eb = nullptr;
        neb = 0;
        hipFree(tb);
        tb = nullptr;
        ntb = 0;
        elem.clear();
        is_elem.clear();
        drop_graph();
        built = false;
        tuned = false;
    }

    void release()
    {
        if (stream) hipStreamSynchronize(stream);
        release_tables();
        hipFree(pts);
        pts = nullptr;
        for (hipStream_t l : lanes) hipStreamDestroy(l);
        for (hipEvent_t e : ljoin) hipEventDestroy(e);
        lanes.clear();
        ljoin.clear();
        if (lfork) hipEventDestroy(lfork);
        lfork = nullptr;
        hipFree(d);
        d = nullptr;
        hipHostFree(pinned);
        pinned = nullptr;
        dcap = 0;
        if (done) hipEventDestroy(done);
        done = nullptr;
        if (staged) hipEventDestroy(staged);
        staged = nullptr;
        if (own_stream && stream) hipStreamDestroy(stream);
        stream = nullptr;
    }
};

// no transform needed: no products, no stored fields, no real-space noise,
// and phi's other terms read phi itself or nothing
static int retable(FusedBatch &B);

static bool is_elementwise(const FusedInst &a)
{
    if (a.ngroups != 0 || a.nreal != 0 || a.ncons != 0 || a.noise_mode == NOISE_REAL) return false;
    for (int i = 0; i < a.tgt[0].nterms; i++)
        if (a.tgt[0].term[i].src != SRC_OLD && a.tgt[0].term[i].src != SRC_CONST) return false;
    return true;
}

static int alloc_spec(float2 **p, const Geom &g)
{
    if (!*p && hipMalloc(p, sizeof(float2) * g.nh) != hipSuccess) return fail(CUPSS_ERR_NOMEM, "device allocation failed");
    return CUPSS_OK;
}

// prefactors of one term multiplier into the pool (S2, src/term_init.cpp:
// 135-200: i^m factored out, the sign of each monomial's i^(I-m) folded in)
static int pool_term(FusedInst &a, int &np, const std::vector<cupss_pres> &pres, TermRefDev &R)
{
    const int m = pres.empty() ? 0 : (pres[0].iqx + pres[0].iqy + pres[0].iqz) % 2;
    R.X = nullptr;
    R.kind = 0;
    R.mult_i = m;
    R.pres0 = np;
    R.npres = (int)pres.size();
    for (const cupss_pres &p : pres) {
        if (np >= FUSED_MAX_POOL) return fail(CUPSS_ERR_UNSUPPORTED, "too many prefactors for the fused step");
        const int iu = p.iqx + p.iqy + p.iqz;
        PresDev &D = a.pool[np++];
        D.c = p.preFactor * (float)(-2 * (((iu - m) / 2) % 2) + 1);
        D.q2n = p.q2n;
        D.a = p.iqx;
        D.b = p.iqy;
        D.cz = p.iqz;
        D.invq = p.invq;
    }
    return CUPSS_OK;
}

// field f is a dynamic field of the plan whose dealiased copy was never
// written: it reads as zero in the products (the reference's step 0)
static bool step0_zero(cupss_system s, int f)
{
    const FusedPlan &P = s->plan;
    if (s->f[f].dstate != 0) return false;
    if (f == P.phi) return true;
    for (int k = 0; k < P.nxdyn; k++)
        if (P.cons[k] == f) return true;
    return false;
}

// device descriptor of system s (its fused plan must be ok); natural: the
// 3D path, which needs no column-major spectra
static int describe(cupss_system s, FusedInst &a, bool natural = false)
{
    const FusedPlan &P = s->plan;
    const Geom &g = s->g;
    FieldRec &F = s->f[P.phi];
    for (size_t i = 0; i < P.real.size(); i++)
        if (int r = alloc_spec(&s->fz[i], g)) return r;
    for (size_t i = 0; i < P.groups.size(); i++)
        if (int r = alloc_spec(&s->fx[i], g)) return r;
    std::memset(&a, 0, sizeof(a));
    a.g = g;
    a.eta = F.eta;
    a.step_base = s->step;
    a.scale = 1.0f / (float)g.n;
    build_update(s, P.phi, a.u);
    a.u.noise = F.eta;
    a.noise_mode = !F.noisy ? NOISE_NONE : (s->noise_mode == 1 ? NOISE_SPECTRAL : NOISE_REAL);
    a.key0 = (uint32_t)F.seed;
    a.key1 = (uint32_t)(F.seed >> 32);
    a.stream = (uint32_t)P.phi;
    int np = 0;
    auto index_of = [](const std::vector<int> &v, int x) { return (int)(std::find(v.begin(), v.end(), x) - v.begin()); };
    // derived fields: one multiplier each (also L_f of their real fields)
    if (!fused_noise_ok(s)) return fail(CUPSS_ERR_UNSUPPORTED, "the fused 3D step draws noise in spectral space only");
    a.nder = (int)P.derived.size();
    for (size_t d = 0; d < P.derived.size(); d++) {
        if (int r = pool_term(a, np, s->t[s->f[P.derived[d]].terms[0]].pres, a.dmul[d])) return r;
        a.dbase[d] = P.dbase[d] == P.phi ? -1 : index_of(P.cons, P.dbase[d]);
    }
    // stored constraint fields
    a.ncons = (int)P.cons.size();
    a.nfresh = P.nfresh;
    // targets: phi, then the stored fields; groups are ordered the same way
    std::vector<int> targets = {P.phi};
    targets.insert(targets.end(), P.cons.begin(), P.cons.end());
    a.ngroups = (int)P.groups.size();
    int gi = 0, pi = 0;
    for (size_t k = 0; k < targets.size(); k++) {
        const int f = targets[k];
        FusedTarget &TG = a.tgt[k];
        TG.g0 = gi;
        while (gi < (int)P.groups.size() && P.groups[gi].target == f) {
            const FusedGroup &grp = P.groups[gi];
            a.X[gi] = s->fx[gi];
            a.pg0[gi] = pi;
            for (size_t q = 0; q < grp.terms.size(); q++) {
                for (int c : s->t[grp.terms[q]].fac) a.cnt[pi][index_of(P.real, c)]++;
                a.coef[pi] = grp.coef[q];
                pi++;
            }
            if (int r = pool_term(a, np, grp.pres, a.gmul[gi])) return r;
            gi++;
        }
        TG.g1 = gi;
        TG.nterms = 0;
        for (int ti : s->f[f].terms) {
            const TermRec &T = s->t[ti];
            if (T.fac.size() >= 2) continue;
            FusedTerm &FT = TG.term[TG.nterms++];
            TermRefDev R;
            if (int r = pool_term(a, np, T.pres, R)) return r;
            FT.mult_i = R.mult_i;
            FT.pres0 = R.pres0;
            FT.npres = R.npres;
            FT.idx = 0;
            // a dynamic target (phi, or a stored dynamic field) and a fresh
            // constraint read the values before the step, the other
            // constraints the dynamic fields' after it
            const bool dyn_target = k == 0 || (int)k - 1 < P.nxdyn + P.nfresh;
            if (T.fac.empty()) {
                FT.src = SRC_CONST;
            } else if (T.fac[0] == P.phi) {
                FT.src = dyn_target ? SRC_OLD : SRC_NEW;
            } else if (std::find(P.derived.begin(), P.derived.end(), T.fac[0]) != P.derived.end()) {
                FT.src = SRC_DERIVED;
                FT.idx = index_of(P.derived, T.fac[0]);
            } else {
                FT.idx = index_of(P.cons, T.fac[0]);
                FT.src = (!dyn_target && FT.idx < P.nxdyn) ? SRC_DYN_NEW : SRC_CONS;
            }
        }
        TG.nimp = 0;
        TG.imp0 = np;
        TG.z = -1;
        TG.dyn = 0;
        TG.fresh = (k > 0 && (int)k - 1 >= P.nxdyn && (int)k - 1 < P.nxdyn + P.nfresh) ? 1 : 0;
        TG.keep = (k > 0 && (int)k - 1 >= P.nxdyn && s->f[f].terms.empty() && !s->f[f].noisy) ? 1 : 0;
        TG.noisy = 0;
        if (k > 0 && s->f[f].noisy) {
            // a stored field's noise: its own key, stream and buffer; a
            // constraint solved for the next step draws that step's noise
            UpdateArgs U;
            build_update(s, f, U);
            const FieldRec &G = s->f[f];
            TG.noisy = s->noise_mode == 1 ? NOISE_SPECTRAL : NOISE_REAL;
            TG.eta = G.eta;
            TG.noise_base = U.noise_base;
            TG.noise_q2n = U.noise_q2n;
            TG.noise_invq = U.noise_invq;
            TG.sdt = U.sdt;
            TG.key0 = (uint32_t)G.seed;
            TG.key1 = (uint32_t)(G.seed >> 32);
            TG.stream = (uint32_t)f;
            TG.step_off = (int)k - 1 >= P.nxdyn + P.nfresh ? 1 : 0;
        }
        if (k > 0 && (int)k - 1 < P.nxdyn) TG.dyn = 1;  // a further dynamic field: D(q) from its implicit prefactors
        if (k > 0) {
            for (const cupss_pres &q : s->f[f].imp) {
                if (np >= FUSED_MAX_POOL) return fail(CUPSS_ERR_UNSUPPORTED, "too many prefactors for the fused step");
                PresDev &D = a.pool[np++];
                D.c = q.preFactor;
                D.q2n = q.q2n;
                D.a = q.iqx;
                D.b = q.iqy;
                D.cz = q.iqz;
                D.invq = q.invq;
                TG.nimp++;
            }
            const int c = (int)k - 1;
            if (int r = alloc_spec(&s->fc[c][0], g)) return r;
            if (int r = alloc_spec(&s->fc[c][1], g)) return r;
            a.cs[c][0] = s->fc[c][0];
            a.cs[c][1] = s->fc[c][1];
        }
    }
    a.pg0[gi] = pi;
    // real fields
    a.nreal = (int)P.real.size();
    for (size_t i = 0; i < P.real.size(); i++) {
        const int f = P.real[i];
        a.p[i] = s->f[f].order;
        a.kmax[i] = a.p[i] > 0 ? std::min(g.nc, g.sx / (a.p[i] + 1)) : g.nc;
        a.kymax[i] = a.p[i] > 0 ? g.sy / (a.p[i] + 1) : g.sy;
        a.Z[i] = s->fz[i];
        a.zero[i] = step0_zero(s, f) ? 1 : 0;
        a.zsrc[i] = -1;
        a.lbase[i] = -1;
        a.lmul[i] = TermRefDev{nullptr, 0, 0, 0, 0};
        const int d = index_of(P.derived, f), c = index_of(P.cons, f);
        if (d < (int)P.derived.size()) {
            a.lmul[i] = a.dmul[d];
            a.lbase[i] = a.dbase[d];
        } else if (c < (int)P.cons.size()) {
            a.zsrc[i] = c;
            a.tgt[1 + c].z = (int)i;
        }
    }
    if (!natural && !is_elementwise(a)) {
        if (int r = alloc_spec(&s->ft[0], g)) return r;
        if (int r = alloc_spec(&s->ft[1], g)) return r;
    }
    return CUPSS_OK;
}

// the per-run parts of a descriptor: buffers, step counter, step-0 flags
// the per-run fields of an instance's descriptor; true when any changed
// (step_base: the device copy advances with the steps, see run)
static bool run_view(cupss_system s, bool elementwise, FusedInst &a)
{
    bool changed = false;
    auto set = [&changed](auto &field, auto value) {
        if (field != value) {
            field = value;
            changed = true;
        }
    };
    FieldRec &F = s->f[s->plan.phi];
    if (elementwise) {
        set(a.spec[0], F.spec[F.cur]);
        set(a.spec[1], F.spec[F.cur ^ 1]);
        set(a.nat, (float2 *)nullptr);
    } else {
        const int b = s->resume_now ? s->cont_ft : 0;
        set(a.spec[0], s->ft[b]);
        set(a.spec[1], s->ft[b ^ 1]);
        set(a.nat, F.spec[F.cur]);
    }
    set(a.resume, !elementwise && s->resume_now ? 1 : 0);
    set(a.step_base, (uint64_t)s->step);
    set(a.cont, s->cont_now ? 1 : 0);
    const int b0 = s->cont_now ? s->cont_buf : 0;
    for (int c = 0; c < a.ncons; c++) {
        FieldRec &C = s->f[s->plan.cons[c]];
        set(a.cnat[c], C.spec[C.cur]);
        set(a.cs[c][0], s->fc[c][b0]);
        set(a.cs[c][1], s->fc[c][b0 ^ 1]);
    }
    for (int k = 0; k < a.nreal; k++) set(a.zero[k], step0_zero(s, s->plan.real[k]) ? 1 : 0);
    for (int d = 0; d < a.nder; d++) {
        FieldRec &D = s->f[s->plan.derived[d]];
        set(a.dout[d], D.spec[D.cur]);
    }
    return changed;
}

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

// bytes of the pipeline buffers one transform instance touches per step
static double inst_bytes(const FusedInst &a)
{
    int arrays = 2 + 2 * a.ncons + a.nreal + a.ngroups + (a.noise_mode == NOISE_REAL ? 1 : 0);
    for (int c = 0; c < a.ncons; c++)
        if (a.tgt[1 + c].noisy == NOISE_REAL) arrays++;
    return 8.0 * (double)a.g.nh * arrays;
}

// row and column tables of the chunks of class c for shapes (lpb, twd)
static int class_tables(const FusedBatch &B, const FusedBatch::Class &c, int lpb, int twd, double chunk_bytes,
                        std::vector<Tables> &out)
{
    std::vector<std::vector<int>> parts;
    double acc = 0;
    for (int i : c.insts) {
        const double b = inst_bytes(B.h[i]);
        if (parts.empty() || (chunk_bytes > 0 && acc + b > chunk_bytes && acc > 0)) {
            parts.emplace_back();
            acc = 0;
        }
        parts.back().push_back(i);
        acc += b;
    }
    for (auto &part : parts) {
        std::vector<FusedBlock> rows, cols;
        for (int i : part) {
            const Geom &g = B.h[i].g;
            for (int r0 = 0; r0 < g.sy; r0 += lpb) rows.push_back(FusedBlock{i, r0});
            // column tiles interleaved with stride 8: each XCD takes a
            // contiguous range of the table, and the columns the dealias
            // masks keep (which also transform Z_f) sit at low kx; a stride-8
            // order gives every XCD its share of them
            const int ntile = (g.nc1 + twd - 1) / twd;
            for (int j = 0; j < 8; j++)
                for (int tc = j; tc < ntile; tc += 8) cols.push_back(FusedBlock{i, tc * twd});
        }
        out.push_back(Tables());
        if (int r = upload_blocks(rows, &out.back().rb, &out.back().nrb)) return r;
        if (int r = upload_blocks(cols, &out.back().cb, &out.back().ncb)) return r;
    }
    return CUPSS_OK;
}

// shapes before (or without) tuning: up to 8 rows per row workgroup, up to
// 256 threads per column workgroup
static void default_shapes(FusedBatch::Class &c)
{
    const std::vector<int> rs = fused_row_shapes(c.nc, c.ny, c.width), cs = fused_col_shapes(c.ny);
    c.lpb = rs.empty() ? 0 : rs.front();
    for (int v : rs)
        if (v <= 8) c.lpb = v;
    c.twd = cs.empty() ? 0 : cs.front();
    const int T = c.ny >= 4096 ? c.ny / 16 : (c.ny >= 8 ? c.ny / 8 : 1);
    for (int v : cs)
        if (v * T <= 256) c.twd = v;
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
    B.is_elem.assign(B.h.size(), 0);
    std::vector<FusedBlock> tiles, er;
    for (size_t i = 0; i < B.h.size(); i++) {
        const FusedInst &a = B.h[i];
        if (is_elementwise(a)) {
            B.elem.push_back((int)i);
            B.is_elem[i] = 1;
            for (long h0 = 0; h0 < a.g.nh / 2; h0 += FUSED_ELEM_BLOCK) er.push_back(FusedBlock{(int)i, (int)h0});
            continue;
        }
        int stored = a.ncons > 0 ? 1 : 0;
        for (int k = 0; k < a.ncons; k++)
            if (a.tgt[1 + k].dyn || a.tgt[1 + k].fresh || a.tgt[1 + k].noisy) stored = 2;
        const int width = fused_row_width(a.nreal);
        auto key = std::make_pair(a.g.nc * 16 + (width == 2 ? 0 : width == 3 ? 4 : 8) + stored, a.g.sy);
        if (!cls.count(key)) {
            cls[key] = (int)B.classes.size();
            B.classes.emplace_back();
            B.classes.back().nc = a.g.nc;
            B.classes.back().ny = a.g.sy;
            B.classes.back().width = width;
            B.classes.back().cons = stored;
        }
        FusedBatch::Class &c = B.classes[cls[key]];
        c.insts.push_back((int)i);
        c.real = c.real || a.nreal > 0;
        const int ntx = (a.g.nc1 + FUSED_TILE - 1) / FUSED_TILE, nty = (a.g.sy + FUSED_TILE - 1) / FUSED_TILE;
        for (int q = 0; q < ntx * nty; q++) tiles.push_back(FusedBlock{(int)i, q});
    }
    for (auto &c : B.classes) {
        default_shapes(c);
        if (c.lpb <= 0 || c.twd <= 0) return fail(CUPSS_ERR_UNSUPPORTED, "no tile shape for this size class");
        if (int r = class_tables(B, c, c.lpb, c.twd, B.chunk_bytes, c.chunks)) return r;
        c.small = fused_small_class(c.nc, c.ny, c.width);
        if (c.small) {
            if (hipMalloc(&c.ids, sizeof(int) * c.insts.size()) != hipSuccess)
                return fail(CUPSS_ERR_NOMEM, "device allocation failed");
            HIP_TRY(hipMemcpy(c.ids, c.insts.data(), sizeof(int) * c.insts.size(), hipMemcpyHostToDevice));
        }
    }
    if (int r = upload_blocks(er, &B.eb, &B.neb)) return r;
    if (int r = upload_blocks(tiles, &B.tb, &B.ntb)) return r;
    if (B.profile)
        if (int r = retable(B)) return r;
    B.gens.clear();
    B.vgens.clear();
    for (cupss_system s : B.sys) {
        B.gens.push_back(s->gen);
        B.vgens.push_back(s->vgen);
    }
    B.dirty.assign(B.h.size(), 1);
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

// upload the per-run descriptors that changed (a run that follows a run of
// an even number of steps usually changes none).  The device table is
// rewritten in stream order, after the previous run's kernels; the staging
// buffer may still feed the previous upload, so only that copy is waited
// for -- the host does not wait for the steps.
static int refresh(FusedBatch &B)
{
    const size_t n = B.h.size();
    if (B.dirty.size() != n) B.dirty.assign(n, 1);
    bool any = false;
    for (size_t i = 0; i < n; i++) {
        if (B.vgens[i] != B.sys[i]->vgen) {  // new prefactor values, same structure: new pool
            if (int r = describe(B.sys[i], B.h[i])) return r;
            B.vgens[i] = B.sys[i]->vgen;
            B.dirty[i] = 1;
        }
        if (run_view(B.sys[i], B.is_elem[i], B.h[i])) B.dirty[i] = 1;
        any = any || B.dirty[i];
    }
    if (!any) return CUPSS_OK;
    if (B.staged) HIP_TRY(hipEventSynchronize(B.staged));
    for (size_t i = 0; i < n;) {
        if (!B.dirty[i]) {
            i++;
            continue;
        }
        size_t j = i;
        for (; j < n && B.dirty[j]; j++) {
            std::memcpy(B.pinned + j, &B.h[j], sizeof(FusedInst));
            B.dirty[j] = 0;
        }
        HIP_TRY(hipMemcpyAsync(B.d + i, B.pinned + i, sizeof(FusedInst) * (j - i), hipMemcpyHostToDevice, B.stream));
        i = j;
    }
    if (!B.staged) HIP_TRY(hipEventCreateWithFlags(&B.staged, hipEventDisableTiming));
    HIP_TRY(hipEventRecord(B.staged, B.stream));
    return CUPSS_OK;
}

// Launches are skipped where every instance they serve resumes (its
// workgroups would only return): the one-step-per-call loop then costs the
// steps' own launches.
static int launch_prologue(FusedBatch &B)
{
    bool all = true;
    for (size_t i = 0; i < B.h.size(); i++)
        if (!B.is_elem[i] && !B.h[i].resume) all = false;
    if (B.ntb > 0 && !all) HIP_TRY(fused_transpose_in_launch(B.d, B.tb, B.ntb, B.stream));
    for (auto &c : B.classes) {
        bool resumed = true;
        for (int i : c.insts) resumed = resumed && B.h[i].resume;
        if (c.real && !resumed)
            for (auto &t : c.chunks)
                HIP_TRY(fused_cols_launch(c.ny, c.twd, c.cons, B.d, t.cb, t.ncb, 0, 0, 0, B.stream));
    }
    return CUPSS_OK;
}

// steps [n0, n0+k) of every transform instance, chunk by chunk; step n
// reads spectrum buffer n&1 and uses counter offset n - rel0
// steps [n0, n0+k) of the instances of class ci, chunk by chunk, on stream st
static int launch_class(FusedBatch &B, size_t ci, int n0, int k, int rel0, hipStream_t st)
{
    auto &c = B.classes[ci];
    const bool timed = B.profile && (int)ci == B.pclass && c.chunks.size() == 1;
    if (c.small) {
        unsigned long long *ts = nullptr;
        if (timed && B.pcount < FusedBatch::PROF_SLOTS) {
            ts = B.pts + (size_t)(B.pcount++) * FUSED_TS_SLOTS * FUSED_TS_STRIDE;
            B.psteps = k;
        }
        HIP_TRY(fused_small_launch(c.nc, c.ny, c.cons, B.d, c.ids, (int)c.insts.size(), n0, k, rel0, st, ts));
        return CUPSS_OK;
    }
    for (auto &t : c.chunks)
        for (int n = n0; n < n0 + k; n++) {
            HIP_TRY(fused_rows_launch(c.nc, c.lpb, c.width, B.d, t.rb, t.nrb, (uint32_t)(n - rel0), st));
            unsigned long long *ts = nullptr;
            if (timed && B.pcount < FusedBatch::PROF_SLOTS)
                ts = B.pts + (size_t)(B.pcount++) * FUSED_TS_SLOTS * FUSED_TS_STRIDE;
            HIP_TRY(fused_cols_launch(c.ny, c.twd, c.cons, B.d, t.cb, t.ncb, n & 1, (uint32_t)(n - rel0), 1, st, ts));
        }
    return CUPSS_OK;
}

// steps [n0, n0+k) of every transform instance on the batch stream; step n
// reads spectrum buffer n&1 and uses counter offset n - rel0
static int launch_steps(FusedBatch &B, int n0, int k, int rel0)
{
    if (B.profile) {
        B.pcount = 0;
        HIP_TRY(fused_ts_reset_launch(B.pts, FusedBatch::PROF_SLOTS, B.stream));
    }
    for (size_t ci = 0; ci < B.classes.size(); ci++)
        if (int r = launch_class(B, ci, n0, k, rel0, B.stream)) return r;
    return CUPSS_OK;
}

// steps [n0, n0+k) of every instance, the size classes and the elementwise
// set spread over lanes (longest work first onto the least loaded lane)
static int launch_all(FusedBatch &B, int n0, int k, int rel0)
{
    if (B.profile) {
        B.pcount = 0;
        HIP_TRY(fused_ts_reset_launch(B.pts, FusedBatch::PROF_SLOTS, B.stream));
    }
    std::vector<std::pair<double, int>> items;  // (cost, class index; -1 = elementwise)
    for (size_t ci = 0; ci < B.classes.size(); ci++) {
        double cost = 0;
        for (int i : B.classes[ci].insts) cost += (double)B.h[i].g.n * (4 + B.h[i].nreal + B.h[i].ngroups);
        items.push_back({cost, (int)ci});
    }
    if (B.neb > 0) {
        double cost = 0;
        for (int i : B.elem) cost += (double)B.h[i].g.n;
        items.push_back({cost, -1});
    }
    const int L = std::min<int>(FusedBatch::MAX_LANES, (int)items.size());
    // Work that takes all k steps in one launch on few workgroups (the
    // small-instance classes, the elementwise set) is enqueued first, so it
    // starts beside the per-step launches of the large classes instead of
    // running alone after them; then the rest, longest first.
    auto one_launch = [&B](int ci) { return ci < 0 || B.classes[ci].small; };
    std::sort(items.begin(), items.end(), [&](const std::pair<double, int> &a, const std::pair<double, int> &b) {
        const bool oa = one_launch(a.second), ob = one_launch(b.second);
        if (oa != ob) return oa;
        return a.first > b.first;
    });
    std::vector<double> load(std::max(L, 1), 0.0);
    std::vector<int> lane(items.size(), 0);
    for (size_t it = 0; it < items.size(); it++) {
        const int j = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        lane[it] = j;
        load[j] += items[it].first;
    }
    while ((int)B.lanes.size() < L - 1) {
        hipStream_t st;
        hipEvent_t ev;
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        B.lanes.push_back(st);
        B.ljoin.push_back(ev);
    }
    if (L > 1) {
        if (!B.lfork) HIP_TRY(hipEventCreateWithFlags(&B.lfork, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(B.lfork, B.stream));
        for (int j = 1; j < L; j++) HIP_TRY(hipStreamWaitEvent(B.lanes[j - 1], B.lfork, 0));
    }
    for (size_t it = 0; it < items.size(); it++) {
        hipStream_t st = lane[it] == 0 ? B.stream : B.lanes[lane[it] - 1];
        if (items[it].second >= 0) {
            if (int r = launch_class(B, (size_t)items[it].second, n0, k, rel0, st)) return r;
        } else {
            HIP_TRY(fused_elem_launch(B.d, B.eb, B.neb, n0 & 1, (uint32_t)(n0 - rel0), k, st));
        }
    }
    for (int j = 1; j < L; j++) {
        HIP_TRY(hipEventRecord(B.ljoin[j - 1], B.lanes[j - 1]));
        HIP_TRY(hipStreamWaitEvent(B.stream, B.ljoin[j - 1], 0));
    }
    return CUPSS_OK;
}

// the class the timing events follow: the most real points, of the classes
// that launch a column pass per step if there are any
static void pick_profile_class(FusedBatch &B)
{
    B.pclass = -1;
    B.psteps = 1;
    double best = -1;
    bool per_step = false;
    for (const auto &c : B.classes) per_step = per_step || !c.small;
    for (size_t ci = 0; ci < B.classes.size(); ci++) {
        if (per_step && B.classes[ci].small) continue;
        double pts = 0;
        for (int i : B.classes[ci].insts) pts += (double)B.h[i].g.n;
        if (pts > best) {
            best = pts;
            B.pclass = (int)ci;
        }
    }
}

// ---- autotune

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer()
    {
        if (a) hipEventDestroy(a);
        if (b) hipEventDestroy(b);
    }
};

template <class F>
static int time_ms(FusedBatch &B, Timer &T, int reps, F &&launch, float *ms)
{
    if (!T.a) {
        HIP_TRY(hipEventCreate(&T.a));
        HIP_TRY(hipEventCreate(&T.b));
    }
    if (int r = launch()) return r;  // warm-up
    HIP_TRY(hipEventRecord(T.a, B.stream));
    for (int i = 0; i < reps; i++)
        if (int r = launch()) return r;
    HIP_TRY(hipEventRecord(T.b, B.stream));
    HIP_TRY(hipEventSynchronize(T.b));
    HIP_TRY(hipEventElapsedTime(ms, T.a, T.b));
    return CUPSS_OK;
}

// Time every tile shape of every class on the batch's own data, then the
// chunk sizes, and keep the fastest.  Runs between a prologue and the
// restoring prologue that follows it: the launches only rewrite buffers
// that the prologue and the steps rewrite anyway.
static int tune(FusedBatch &B)
{
    Timer T;
    const int reps = 3;
    for (auto &c : B.classes) {
        if (c.small) continue;  // fixed tiles, one launch per run
        // rows
        float best = 1e30f;
        int pick = c.lpb;
        for (int lpb : fused_row_shapes(c.nc, c.ny, c.width)) {
            std::vector<Tables> t;
            if (int r = class_tables(B, c, lpb, c.twd, 0, t)) return r;
            float ms = 0;
            const int r = time_ms(
                B, T, reps,
                [&]() -> int {
                    HIP_TRY(fused_rows_launch(c.nc, lpb, c.width, B.d, t[0].rb, t[0].nrb, 0, B.stream));
                    return CUPSS_OK;
                },
                &ms);
            HIP_TRY(hipStreamSynchronize(B.stream));
            FusedBatch::free_tables(t);
            if (r) return r;
            if (ms < best) {
                best = ms;
                pick = lpb;
            }
        }
        c.lpb = pick;
        // columns
        best = 1e30f;
        pick = c.twd;
        for (int twd : fused_col_shapes(c.ny)) {
            std::vector<Tables> t;
            if (int r = class_tables(B, c, c.lpb, twd, 0, t)) return r;
            float ms = 0;
            const int r = time_ms(
                B, T, reps,
                [&]() -> int {
                    HIP_TRY(fused_cols_launch(c.ny, twd, c.cons, B.d, t[0].cb, t[0].ncb, 0, 0, 1, B.stream));
                    return CUPSS_OK;
                },
                &ms);
            HIP_TRY(hipStreamSynchronize(B.stream));
            FusedBatch::free_tables(t);
            if (r) return r;
            if (ms < best) {
                best = ms;
                pick = twd;
            }
        }
        c.twd = pick;
    }
    // chunk size: whole classes, or working sets of 128 / 64 / 32 MB
    double total = 0;
    for (auto &c : B.classes)
        for (int i : c.insts) total += inst_bytes(B.h[i]);
    std::vector<double> cand = {0};
    for (double mb : {128.0, 64.0, 32.0})
        if (total > mb * 1048576.0) cand.push_back(mb * 1048576.0);
    float best = 1e30f;
    double pick = 0;
    for (double cb : cand) {
        HIP_TRY(hipStreamSynchronize(B.stream));
        for (auto &c : B.classes) {
            FusedBatch::free_tables(c.chunks);
            if (int r = class_tables(B, c, c.lpb, c.twd, cb, c.chunks)) return r;
        }
        float ms = 0;
        if (int r = time_ms(B, T, 1, [&]() { return launch_steps(B, 0, 4, 0); }, &ms)) return r;
        if (ms < best) {
            best = ms;
            pick = cb;
        }
    }
    HIP_TRY(hipStreamSynchronize(B.stream));
    B.chunk_bytes = pick;
    B.tuned = true;
    return retable(B);
}

// block tables of every class for its shapes and the chunk target (the
// class being timed stays one chunk)
static int retable(FusedBatch &B)
{
    HIP_TRY(hipStreamSynchronize(B.stream));
    if (B.profile) pick_profile_class(B);
    for (size_t ci = 0; ci < B.classes.size(); ci++) {
        auto &c = B.classes[ci];
        FusedBatch::free_tables(c.chunks);
        const double cb = B.profile && (int)ci == B.pclass ? 0.0 : B.chunk_bytes;
        if (int r = class_tables(B, c, c.lpb, c.twd, cb, c.chunks)) return r;
    }
    return CUPSS_OK;
}

// after a run: the general path's view of every system.  Transform
// instances get their spectrum back in natural layout and each derived
// field P(M_psi phi) of the last step's starting phi; elementwise instances
// wrote both in place.  Member streams then wait for the batch stream.
static int settle(FusedBatch &B, int nsteps)
{
    // the transpose back is deferred: a following run resumes from the
    // column-major state; anything else that touches a member flushes it
    B.settle_pending = true;
    bool waits = false;
    for (cupss_system s : B.sys) waits = waits || s->stream != B.stream;
    if (waits) {
        if (!B.done) HIP_TRY(hipEventCreateWithFlags(&B.done, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(B.done, B.stream));
    }
    for (size_t i = 0; i < B.sys.size(); i++) {
        cupss_system s = B.sys[i];
        s->pending = &B;
        const FusedPlan &P = s->plan;
        FieldRec &F = s->f[P.phi];
        if (B.is_elem[i]) F.cur ^= (nsteps & 1);
        s->step += (uint64_t)nsteps;
        s->fused_steps += nsteps;
        if (F.order > 0) F.dstate = 2;
        for (int dfi : P.derived)
            if (s->f[dfi].order > 0) s->f[dfi].dstate = 2;
        for (int cfi : P.cons)
            if (s->f[cfi].order > 0) s->f[cfi].dstate = 2;
        if (!P.cons.empty()) {
            // the last pass left the next step's phase A/B in buffer nsteps&1
            s->cont_buf = (s->cont_now ? s->cont_buf : 0) ^ (nsteps & 1);
        }
        if (!B.is_elem[i]) {
            // phi (column-major) and Z_f of the final state stay for a resume
            s->cont_ft = (s->resume_now ? s->cont_ft : 0) ^ (nsteps & 1);
            s->cont_step = s->step;
            s->cont_gen = s->gen;
        }
        if (s->stream != B.stream) HIP_TRY(hipStreamWaitEvent(s->stream, B.done, 0));
    }
    return CUPSS_OK;
}

// the deferred transpose back of the last run: every transform instance's
// spectrum, stored fields and derived fields in natural layout again
static int flush(FusedBatch &B)
{
    if (!B.settle_pending) return CUPSS_OK;
    B.settle_pending = false;
    // the run's last k_bump_steps left the final state in buffer [0] and
    // the last step's start in [1]
    if (B.ntb > 0) HIP_TRY(fused_settle_launch(B.d, B.tb, B.ntb, 0, B.stream));
    bool waits = false;
    for (cupss_system s : B.sys) {
        if (s->pending == &B) s->pending = nullptr;
        waits = waits || s->stream != B.stream;
    }
    if (waits) {
        if (!B.done) HIP_TRY(hipEventCreateWithFlags(&B.done, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(B.done, B.stream));
        for (cupss_system s : B.sys)
            if (s->stream != B.stream) HIP_TRY(hipStreamWaitEvent(s->stream, B.done, 0));
    }
    return CUPSS_OK;
}

int fused_flush(cupss_system s)
{
    if (!s || !s->pending) return CUPSS_OK;
    return flush(*static_cast<FusedBatch *>(s->pending));
}

static int run(FusedBatch &B, int nsteps)
{
    if (nsteps <= 0) return CUPSS_OK;
    // a member whose state another batch's run left column-major: written
    // back first
    for (cupss_system s : B.sys)
        if (s->pending && s->pending != &B)
            if (int r = fused_flush(s)) return r;
    for (cupss_system s : B.sys)
        if (s->stream != B.stream) HIP_TRY(hipStreamSynchronize(s->stream));
    // this batch's own deferred transpose: needed only where a transform
    // instance cannot resume from the column-major state (tables rebuilt,
    // autotuning, a state changed since), and before the descriptors and
    // tables it reads change
    if (B.settle_pending) {
        bool need = stale_tables(B) || (B.autotune && !B.tuned);
        for (size_t i = 0; i < B.sys.size() && !need; i++)
            need = !B.is_elem[i] && (B.sys[i]->cont_step != B.sys[i]->step || B.sys[i]->cont_gen != B.sys[i]->gen);
        if (need)
            if (int r = flush(B)) return r;
    }
    if (stale_tables(B))
        if (int r = build(B)) return r;
    // Systems with stored constraint fields: phase A/B of the run's first
    // step with the general program (it reads the dealiased copies as left
    // by earlier steps, whatever path made them); each pass then takes phi
    // one step and the stored fields to the next step's phase A/B.
    // A run continues from the previous fused run's values when nothing
    // changed the system since (and no autotuning, which reuses the buffers).
    const bool will_tune = B.autotune && !B.tuned && !B.classes.empty();
    bool general = false;
    for (cupss_system s : B.sys) {
        s->resume_now = !will_tune && s->cont_step == s->step && s->cont_gen == s->gen;
        s->cont_now = !s->plan.cons.empty() && s->resume_now;
        if ((int)s->plan.cons.size() > s->plan.nxdyn + s->plan.nfresh && !s->cont_now) {
            if (int r = general_constraint_phase(s)) return r;
            general = true;
        }
    }
    if (general)
        for (cupss_system s : B.sys)
            if (s->stream != B.stream) HIP_TRY(hipStreamSynchronize(s->stream));
    if (int r = refresh(B)) return r;
    if (!twiddle_sets(B.stream)) return fail(CUPSS_ERR_NOMEM, "twiddle tables");
    if (int r = launch_prologue(B)) return r;
    if (B.autotune && !B.tuned && !B.classes.empty()) {
        B.drop_graph();
        if (int r = tune(B)) return r;
        if (int r = launch_prologue(B)) return r;
    }
    int done = 0;
    const int K = B.graph_steps;
    if (K >= 2 && nsteps >= K) {
        if (!B.gexec) {
            HIP_TRY(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
            int rc = launch_all(B, 0, K, 0);
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
            B.gpcount = B.pcount;
            B.gpsteps = B.psteps;
        }
        while (nsteps - done >= K) {
            HIP_TRY(hipGraphLaunch(B.gexec, B.stream));
            done += K;
        }
        // a replay clears the timestamp slots and fills them as the captured
        // launches did (the host counted those at the capture)
        B.pcount = B.gpcount;
        B.psteps = B.gpsteps;
        // the remainder continues from the bumped counters
        mirror_run_end(B.h, (uint32_t)done);
    }
    if (nsteps > done) {
        if (int r = launch_all(B, done, nsteps - done, done)) return r;
        // the device's step counters and buffer order follow the steps, so
        // that the next run's descriptors need no upload for them
        const int ni = (int)B.h.size();
        hipLaunchKernelGGL(k_bump_steps, dim3((ni + 255) / 256), dim3(256), 0, B.stream, B.d, ni,
                           (uint32_t)(nsteps - done));
        HIP_TRY(hipGetLastError());
        mirror_run_end(B.h, (uint32_t)(nsteps - done));
    }
    B.steps_run += nsteps;
    return settle(B, nsteps);
}

// z-line tile width of a 3D system: every candidate timed on the system's own
// buffers.  A z pass in mode 2 reads phi [par] and X_g and writes only the
// other spectrum buffer, which the next step overwrites: the state is kept.
static int tune3d(FusedBatch &B, cupss_system s, int par, uint32_t rel)
{
    const Geom &g = s->g;
    const std::vector<int> v = fused3d_z_shapes(g.sz);
    if (v.empty()) return fail(CUPSS_ERR_UNSUPPORTED, "no z-line tile for this size");
    hipStream_t st = s->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 0;
    int pick = v[0];
    int rc = CUPSS_OK;
    for (int w : v) {
        float ms = 0;
        hipError_t e = fused3d_z_launch(g.sz, w, B.d, g.nc1, g.sy, par, rel, 2, st);  // warm-up
        if (e == hipSuccess) e = hipEventRecord(e0, st);
        for (int r = 0; r < 3 && e == hipSuccess; r++) e = fused3d_z_launch(g.sz, w, B.d, g.nc1, g.sy, par, rel, 2, st);
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            rc = fail(CUPSS_ERR_HIP, std::string("z-pass tuning: ") + hipGetErrorString(e));
            break;
        }
        if (w == v[0] || ms < best) {
            best = ms;
            pick = w;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc == CUPSS_OK) B.twd3 = pick;
    return rc;
}

// The fused step of a 3D system (fused.hip: k_f3rows, k_f3z; the y lines
// by lines_y), natural layout, no transposes.  A run is
//   prologue:  z pass in mode 0 and IFFT_y: W_f, Z_f of the current state
//              (skipped when the previous run left them and nothing changed
//              the system since: per-step callers pay no prologue);
//   steps:     rows, FFT_y of every group, z pass (update; W_f of the new
//              state), IFFT_y of every real field;
//   end:       the derived fields of the last step's starting phi.
// The descriptor keeps both spectrum buffers of phi in a fixed order (the
// launches pass the parity) and is uploaded again only when it changes.
static int run3d(FusedBatch &B, cupss_system s, int nsteps)
{
    const FusedPlan &P = s->plan;
    FieldRec &F = s->f[P.phi];
    const Geom &g = s->g;
    FusedInst a;
    if (int r = describe(s, a, true)) return r;
    for (int i = 0; i < a.nreal; i++) {
        if (!s->fw[i]) {
            if (int r = alloc_spec(&s->fw[i], g)) return r;
            HIP_TRY(hipMemsetAsync(s->fw[i], 0, sizeof(float2) * g.nh, s->stream));  // lines off the mask stay zero
        }
        a.W[i] = s->fw[i];
    }
    a.spec[0] = F.spec[0];
    a.spec[1] = F.spec[1];
    a.nat = nullptr;
    a.u.cur = nullptr;  // per-run buffers of the general update, unused here (and they
    a.u.out = nullptr;  // would make every run's descriptor differ)
    for (int k = 0; k < a.nreal; k++) a.zero[k] = (P.real[k] == P.phi && F.dstate == 0) ? 1 : 0;
    for (int d = 0; d < a.nder; d++) {
        FieldRec &D = s->f[P.derived[d]];
        a.dout[d] = D.spec[D.cur];
    }
    bool same = B.d3 && B.h.size() == 1;
    if (same) {
        const uint64_t base = B.h[0].step_base;
        a.step_base = base;
        same = std::memcmp(&a, &B.h[0], sizeof(FusedInst)) == 0 && s->step >= base &&
               s->step - base + (uint64_t)nsteps < (1ull << 31);
    }
    if (!same) {
        a.step_base = s->step;
        HIP_TRY(hipStreamSynchronize(s->stream));  // the staging buffer may still feed an earlier copy
        if (B.dcap < 1) {
            hipFree(B.d);
            hipHostFree(B.pinned);
            B.d = nullptr;
            B.pinned = nullptr;
            if (hipMalloc(&B.d, sizeof(FusedInst)) != hipSuccess ||
                hipHostMalloc(&B.pinned, sizeof(FusedInst)) != hipSuccess)
                return fail(CUPSS_ERR_NOMEM, "descriptor allocation failed");
            B.dcap = 1;
        }
        std::memcpy(B.pinned, &a, sizeof(FusedInst));
        HIP_TRY(hipMemcpyAsync(B.d, B.pinned, sizeof(FusedInst), hipMemcpyHostToDevice, s->stream));
        B.h.assign(1, a);
        B.d3 = true;
    }
    if (!twiddle_sets(s->stream) || !twiddles(s->stream)) return fail(CUPSS_ERR_NOMEM, "twiddle tables");
    const uint32_t rel0 = (uint32_t)(s->step - B.h[0].step_base);
    const int par0 = F.cur, maxr = a.nreal > 3 ? FUSED_MAX_REAL : 3;
    const long nrows = (long)g.sy * g.sz;
    hipStream_t st = s->stream;
    if (!B.twd3)
        if (int r = tune3d(B, s, par0, rel0)) return r;
    // Z_f of the current state: left by the previous run when nothing changed
    // the system since (each run ends with them), else the prologue
    const bool cont = s->cont_step == s->step && s->cont_gen == s->gen;
    if (a.nreal > 0 && !cont) {
        HIP_TRY(fused3d_z_launch(g.sz, B.twd3, B.d, g.nc1, g.sy, par0, rel0, 0, st));
        for (int f = 0; f < a.nreal; f++) HIP_TRY(lines_y(g, +1, s->fw[f], s->fz[f], a.kmax[f] + 1, st));
    }
    if (B.profile) {
        B.pclass = -2;
        B.pcount = 0;
        HIP_TRY(fused_ts_reset_launch(B.pts, FusedBatch::PROF_SLOTS, st));
    }
    for (int i = 0; i < nsteps; i++) {
        if (a.ngroups > 0) {
            HIP_TRY(fused3d_rows_launch(g.nc, maxr, B.d, nrows, st));
            for (int gi = 0; gi < a.ngroups; gi++) HIP_TRY(lines_y(g, -1, s->fx[gi], s->fx[gi], g.nc1, st));
        }
        unsigned long long *ts = nullptr;
        if (B.profile && B.pcount < FusedBatch::PROF_SLOTS)
            ts = B.pts + (size_t)(B.pcount++) * FUSED_TS_SLOTS * FUSED_TS_STRIDE;
        HIP_TRY(fused3d_z_launch(g.sz, B.twd3, B.d, g.nc1, g.sy, par0 ^ (i & 1), rel0 + (uint32_t)i, 1, st, ts));
        // the kx = 0, sx/2 planes of the new state: partners are in other z lines
        HIP_TRY(hermitian_planes(g, F.spec[par0 ^ (i & 1) ^ 1], st));
        for (int f = 0; f < a.nreal; f++) HIP_TRY(lines_y(g, +1, s->fw[f], s->fz[f], a.kmax[f] + 1, st));
    }
    if (a.nder > 0) HIP_TRY(fused3d_derived_launch(B.d, g.nh, par0 ^ ((nsteps - 1) & 1), st));
    F.cur = par0 ^ (nsteps & 1);
    s->step += (uint64_t)nsteps;
    s->fused_steps += nsteps;
    s->cont_step = s->step;
    s->cont_gen = s->gen;
    if (F.order > 0) F.dstate = 2;
    for (int dfi : P.derived)
        if (s->f[dfi].order > 0) s->f[dfi].dstate = 2;
    return CUPSS_OK;
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
    if (s->g.sz > 1) {
        B->stream = s->stream;
        return nsteps > 0 ? run3d(*B, s, nsteps) : CUPSS_OK;
    }
    if (B->stream != s->stream) {
        B->stream = s->stream;
        B->built = false;
    }
    if (B->graph_steps != s->graph_steps) {
        B->graph_steps = s->graph_steps;
        B->drop_graph();
    }
    return run(*B, nsteps);
}

void fused_release(cupss_system s)
{
    delete static_cast<FusedBatch *>(s->batch);
    s->batch = nullptr;
}

static std::string schedule_text(const FusedBatch &B)
{
    std::string out;
    char line[256];
    for (const auto &c : B.classes) {
        std::snprintf(line, sizeof line, "class %dx%d%s%s: %zu instances, rows/wg %d, cols/wg %d, chunks %zu\n",
                      2 * c.nc, c.ny, c.width > 3 ? " (wide)" : "", c.cons ? " (stored fields)" : "", c.insts.size(),
                      c.lpb, c.twd, c.chunks.size());
        out += line;
    }
    std::snprintf(line, sizeof line, "elementwise: %zu instances\nchunk target: %.0f MB%s\n", B.elem.size(),
                  B.chunk_bytes / 1048576.0, B.tuned ? " (autotuned)" : "");
    out += line;
    return out;
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

// a member leaves the batch: its state written back, the batch's work on it
// done, its own stream restored
static int leave(cupss_batch b, cupss_system s)
{
    if (int r = flush(b->b)) return r;
    HIP_TRY(hipStreamSynchronize(b->b.stream));
    if (s->member_of == b) {
        s->stream = s->home_stream;
        s->home_stream = nullptr;
        s->member_of = nullptr;
    }
    return CUPSS_OK;
}

extern "C" int cupss_batch_destroy(cupss_batch b)
{
    if (!b) return CUPSS_OK;
    int rc = CUPSS_OK;
    for (cupss_system s : b->b.sys)
        if (int r = leave(b, s)) rc = r;
    b->b.sys.clear();
    delete b;
    return rc;
}

extern "C" int cupss_batch_add(cupss_batch b, cupss_system s)
{
    if (!b || !s) return fail(CUPSS_ERR_INVALID, "null batch or system");
    if (!s->prepared) return fail(CUPSS_ERR_STATE, "prepare the system before adding it to a batch");
    if (!s->plan.ok || s->g.sz > 1) return fail(CUPSS_ERR_UNSUPPORTED, "system does not qualify for the fused batched step");
    for (auto &F : s->f)
        if (F.real_hook || F.spec_hook || F.real_dev || F.spec_dev)
            return fail(CUPSS_ERR_UNSUPPORTED, "systems with callbacks cannot be batched");
    if (s->member_of) return fail(CUPSS_ERR_STATE, "the system is already in a batch");
    if (fused_stale(s)) {
        // consume a stale dealiased copy with one general step, as a lone system would
        if (int r = general_steps(s, 1)) return r;
    }
    // its own queued work done, the member's work then goes to the batch stream
    if (int r = fused_flush(s)) return r;
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->home_stream = s->stream;
    s->stream = b->b.stream;
    s->member_of = b;
    b->b.sys.push_back(s);
    b->b.built = false;
    return CUPSS_OK;
}

extern "C" int cupss_batch_remove(cupss_batch b, cupss_system s)
{
    if (!b || !s) return fail(CUPSS_ERR_INVALID, "null batch or system");
    auto &v = b->b.sys;
    auto it = std::find(v.begin(), v.end(), s);
    if (it == v.end()) return fail(CUPSS_ERR_INVALID, "system is not in the batch");
    if (int r = leave(b, s)) return r;
    v.erase(it);
    b->b.built = false;
    return CUPSS_OK;
}

extern "C" int cupss_batch_size(cupss_batch b) { return b ? (int)b->b.sys.size() : 0; }

extern "C" int cupss_batch_set_graph_steps(cupss_batch b, int k)
{
    if (!b) return fail(CUPSS_ERR_INVALID, "null batch");
    if (k != 0 && (k < 2 || (k & 1))) return fail(CUPSS_ERR_INVALID, "graph steps must be 0 or an even number >= 2");
    b->b.graph_steps = k;
    b->b.drop_graph();
    return CUPSS_OK;
}

extern "C" int cupss_batch_set_autotune(cupss_batch b, int on)
{
    if (!b) return fail(CUPSS_ERR_INVALID, "null batch");
    b->b.autotune = on != 0;
    b->b.built = false;
    return CUPSS_OK;
}

extern "C" int cupss_batch_schedule(cupss_batch b, char *buf, int len)
{
    if (!b || !buf || len <= 0) return fail(CUPSS_ERR_INVALID, "null batch or buffer");
    const std::string t = schedule_text(b->b);
    std::snprintf(buf, (size_t)len, "%s", t.c_str());
    return CUPSS_OK;
}

static int set_profile(FusedBatch &B, int on)
{
    B.profile = on != 0;
    B.drop_graph();
    if (B.profile && !B.pts &&
        hipMalloc(&B.pts, sizeof(unsigned long long) * FusedBatch::PROF_SLOTS * FUSED_TS_SLOTS * FUSED_TS_STRIDE) !=
            hipSuccess)
        return fail(CUPSS_ERR_NOMEM, "device allocation failed");
    if (B.built)
        if (int r = retable(B)) return r;
    return CUPSS_OK;
}

static int kernel_time(FusedBatch &B, float *avg_ms, double *points, char *name, int len)
{
    if (!B.profile || B.pclass == -1 || B.pcount == 0) return fail(CUPSS_ERR_STATE, "no timed launches yet");
    HIP_TRY(hipStreamSynchronize(B.stream));
    const size_t per = (size_t)FUSED_TS_SLOTS * FUSED_TS_STRIDE;
    std::vector<unsigned long long> raw(per * B.pcount), h(2 * B.pcount);
    HIP_TRY(hipMemcpy(raw.data(), B.pts, sizeof(unsigned long long) * raw.size(), hipMemcpyDeviceToHost));
    for (int j = 0; j < B.pcount; j++) {
        unsigned long long lo = ~0ull, hi = 0;
        for (int q = 0; q < FUSED_TS_SLOTS; q++) {
            const unsigned long long *sl = raw.data() + j * per + (size_t)q * FUSED_TS_STRIDE;
            if (sl[1] == 0) continue;  // no workgroup of this launch used the slot
            lo = std::min(lo, sl[0]);
            hi = std::max(hi, sl[1]);
        }
        h[2 * j] = lo;
        h[2 * j + 1] = hi;
    }
    int dev = 0, khz = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz <= 0) return fail(CUPSS_ERR_HIP, "no wall-clock rate");
    double sum = 0;
    for (int j = 0; j < B.pcount; j++) {
        if (h[2 * j + 1] < h[2 * j]) return fail(CUPSS_ERR_STATE, "timed launch did not record");
        sum += (double)(h[2 * j + 1] - h[2 * j]) / (double)khz;
    }
    // per step: a small-instance launch takes psteps steps
    if (avg_ms) *avg_ms = (float)(sum / B.pcount / B.psteps);
    if (B.pclass == -2) {  // 3D: the z pass of the one system
        if (points) *points = (double)B.h[0].g.n;
        if (name && len > 0) std::snprintf(name, (size_t)len, "k_f3z<%d, %d>", B.h[0].g.sz, B.twd3);
        return CUPSS_OK;
    }
    const auto &c = B.classes[B.pclass];
    double pts = 0;
    for (int i : c.insts) pts += (double)B.h[i].g.n;
    if (points) *points = pts;
    if (name && len > 0) {
        if (c.small)
            std::snprintf(name, (size_t)len, "k_fsmall<%d, %d, %s, %s>", c.nc, c.ny, c.cons ? "true" : "false",
                          c.cons == 2 ? "true" : "false");
        else
            std::snprintf(name, (size_t)len, "k_fcols<%d, %d, %s, %s>", c.ny, c.twd, c.cons ? "true" : "false",
                          c.cons == 2 ? "true" : "false");
    }
    return CUPSS_OK;
}

int cupss::fused_profile(cupss_system s, int on)
{
    if (!s->batch) {
        FusedBatch *B = new FusedBatch();
        B->sys.push_back(s);
        B->stream = s->stream;
        s->batch = B;
    }
    return set_profile(*static_cast<FusedBatch *>(s->batch), on);
}

int cupss::fused_kernel_time(cupss_system s, float *avg_ms, double *points, char *name, int len)
{
    if (!s->batch) return fail(CUPSS_ERR_STATE, "the system has not run the fused path");
    return kernel_time(*static_cast<FusedBatch *>(s->batch), avg_ms, points, name, len);
}

extern "C" int cupss_batch_profile_kernel(cupss_batch b, int on)
{
    if (!b) return fail(CUPSS_ERR_INVALID, "null batch");
    return set_profile(b->b, on);
}

extern "C" int cupss_batch_kernel_time(cupss_batch b, float *avg_ms, double *points, char *name, int len)
{
    if (!b) return fail(CUPSS_ERR_INVALID, "null batch");
    return kernel_time(b->b, avg_ms, points, name, len);
}

extern "C" int cupss_batch_advance(cupss_batch b, int nsteps)
{
    if (!b) return fail(CUPSS_ERR_INVALID, "null batch");
    return run(b->b, nsteps);
}

extern "C" int cupss_batch_time_advance(cupss_batch b, int nsteps, float *ms)
{
    if (!b) return fail(CUPSS_ERR_INVALID, "null batch");
    if (stale_tables(b->b)) {  // host set-up outside the timed region
        if (int r = flush(b->b)) return r;
        if (int r = build(b->b)) return r;
    }
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
