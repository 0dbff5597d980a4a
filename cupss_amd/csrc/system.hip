// system.hip -- the C-ABI of include/cupss_hip.h: one system (the device
// side of one evolver) and its step program.
//
// Per step (evolver::advanceTime, src/evolver.cpp:199-226), for the
// constraint fields and then for the dynamic fields:
//   1. noise of every noisy field of the phase:  eta = R2C(Philox xi)
//   2. every product term of the phase:           X_t = R2C(prod of dealiased)
//   3. every field of the phase: update kernel -> the field's other buffer
//      (reads only current buffers, so all terms see the pre-phase state,
//      as the reference's updateTerms-then-setRHS loops do)
//   4. swap the buffers of the phase's fields
//   5. post-update of every field of the phase: hooks (slow path), and
//      dealiased real copy = C2R(mask_p(spectrum)) when the field is a
//      factor of a product (field::setRHS, src/field.cpp:58-89).
// The reference's comp->real->comp round trip (projection onto real fields)
// is implicit in the half-spectrum representation and in the Nyquist rule
// of the update kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "fused.hpp"
#include "internal.hpp"
#include "philox.hpp"
#include "system_impl.hpp"

using namespace cupss;

static thread_local std::string g_last_error;

int cupss::fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

static bool is_pow2(int v) { return v >= 1 && (v & (v - 1)) == 0; }

extern "C" const char *cupss_last_error(void) { return g_last_error.c_str(); }
extern "C" int cupss_abi_version(void) { return CUPSS_ABI_VERSION; }

static int check_sys(cupss_system s)
{
    if (!s) return fail(CUPSS_ERR_INVALID, "null system");
    return CUPSS_OK;
}

static int check_field(cupss_system s, int field)
{
    if (!s) return fail(CUPSS_ERR_INVALID, "null system");
    if (field < 0 || field >= (int)s->f.size()) return fail(CUPSS_ERR_INVALID, "field id out of range");
    return CUPSS_OK;
}

extern "C" int cupss_system_create(int sx, int sy, int sz, float dx, float dy, float dz, float dt,
                                   cupss_system *out)
{
    if (!out) return fail(CUPSS_ERR_INVALID, "null out");
    *out = nullptr;
    if (!is_pow2(sx) || sx < 2 || sx > 4096 || !is_pow2(sy) || sy > 4096 || !is_pow2(sz) || sz > 4096)
        return fail(CUPSS_ERR_UNSUPPORTED,
                    "grid must be powers of two with 2 <= sx <= 4096 and 1 <= sy, sz <= 4096");
    if ((long)sx * sy * sz >= (1L << 31)) return fail(CUPSS_ERR_UNSUPPORTED, "grid too large");
    if (!(dx > 0) || !(dy > 0) || !(dz > 0)) return fail(CUPSS_ERR_INVALID, "cell sizes must be positive");
    cupss_system s = new cupss_system_t();
    s->g.sx = sx;
    s->g.sy = sy;
    s->g.sz = sz;
    s->g.nc = sx / 2;
    s->g.nc1 = sx / 2 + 1;
    s->g.nh = (long)s->g.nc1 * sy * sz;
    s->g.n = (long)sx * sy * sz;
    // src/field_init.cpp:207-213 (createField always uses the 3D constructor)
    const float PI = 3.1415926535f;
    s->g.sqx = 2.0f * PI / (dx * (float)sx);
    s->g.sqy = 2.0f * PI / (dy * (float)sy);
    s->g.sqz = 2.0f * PI / (dz * (float)sz);
    s->dx = dx;
    s->dy = dy;
    s->dz = dz;
    s->dt = dt;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return fail(CUPSS_ERR_HIP, "no HIP device");
    }
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
        delete s;
        return fail(CUPSS_ERR_HIP, "hipStreamCreate failed");
    }
    s->own_stream = true;
    if (hipMalloc(&s->scratch, sizeof(float2) * s->g.nh) != hipSuccess ||
        hipMalloc(&s->rtmp, sizeof(float) * s->g.n) != hipSuccess) {
        cupss_system_destroy(s);
        return fail(CUPSS_ERR_NOMEM, "device allocation failed");
    }
    *out = s;
    return CUPSS_OK;
}

extern "C" int cupss_system_destroy(cupss_system s)
{
    if (!s) return CUPSS_OK;
    if (s->member_of) cupss_batch_remove((cupss_batch)s->member_of, s);
    if (s->stream) hipStreamSynchronize(s->stream);
    s->pending = nullptr;  // only the system's own batch can still hold it
    fused_release(s);
    if (s->gexec) hipGraphExecDestroy(s->gexec);
    if (s->graph) hipGraphDestroy(s->graph);
    for (auto &F : s->f) {
        hipFree(F.spec[0]);
        hipFree(F.spec[1]);
        hipFree(F.dreal);
        hipFree(F.eta);
    }
    for (auto &T : s->t) hipFree(T.xhat);
    for (auto *p : s->fz) hipFree(p);
    for (auto *p : s->fx) hipFree(p);
    for (auto *p : s->ft) hipFree(p);
    for (auto *p : s->fw) hipFree(p);
    for (auto &c : s->fc) {
        hipFree(c[0]);
        hipFree(c[1]);
    }
    hipFree(s->scratch);
    hipFree(s->rtmp);
    hipFree(s->full);
    if (s->own_stream && s->stream) hipStreamDestroy(s->stream);
    delete s;
    return CUPSS_OK;
}

extern "C" int cupss_field_create(cupss_system s, int dynamic, int *field_id)
{
    if (int r = check_sys(s)) return r;
    FieldRec F;
    F.dynamic = dynamic ? 1 : 0;
    if (hipMalloc(&F.spec[0], sizeof(float2) * s->g.nh) != hipSuccess ||
        hipMalloc(&F.spec[1], sizeof(float2) * s->g.nh) != hipSuccess) {
        hipFree(F.spec[0]);
        return fail(CUPSS_ERR_NOMEM, "device allocation failed");
    }
    HIP_TRY(hipMemsetAsync(F.spec[0], 0, sizeof(float2) * s->g.nh, s->stream));
    HIP_TRY(hipMemsetAsync(F.spec[1], 0, sizeof(float2) * s->g.nh, s->stream));
    s->f.push_back(F);
    if (field_id) *field_id = (int)s->f.size() - 1;
    s->prepared = false;
    s->gen++;
    return CUPSS_OK;
}

extern "C" int cupss_field_add_implicit(cupss_system s, int field, cupss_pres p)
{
    if (int r = check_field(s, field)) return r;
    if ((int)s->f[field].imp.size() >= MAX_IMP) return fail(CUPSS_ERR_UNSUPPORTED, "too many implicit terms");
    s->f[field].imp.push_back(p);
    s->gen++;
    return CUPSS_OK;
}

static void plan_fused(cupss_system s);
static void parameters_changed(cupss_system s, bool same_shape);

// same length and the same exponents in every monomial (values may differ)
static bool same_shape(const std::vector<cupss_pres> &a, const cupss_pres *b, int n)
{
    if ((int)a.size() != n) return false;
    for (int i = 0; i < n; i++)
        if (a[i].q2n != b[i].q2n || a[i].iqx != b[i].iqx || a[i].iqy != b[i].iqy || a[i].iqz != b[i].iqz ||
            a[i].invq != b[i].invq)
            return false;
    return true;
}

extern "C" int cupss_field_set_implicits(cupss_system s, int field, const cupss_pres *p, int n)
{
    if (int r = check_field(s, field)) return r;
    if (n < 0 || n > MAX_IMP || (n > 0 && !p)) return fail(CUPSS_ERR_INVALID, "bad implicit list");
    const bool shape = same_shape(s->f[field].imp, p, n);
    s->f[field].imp.assign(p, p + n);
    parameters_changed(s, shape);
    return CUPSS_OK;
}

static int field_pres_count(cupss_system s, int field, int extra)
{
    int n = extra;
    for (int t : s->f[field].terms) n += (int)s->t[t].pres.size();
    return n;
}

extern "C" int cupss_term_create(cupss_system s, int field, const int *factors, int nfactors,
                                 const cupss_pres *pres, int npres, int *term_id)
{
    if (int r = check_field(s, field)) return r;
    if (nfactors < 0 || nfactors > MAX_FACTORS) return fail(CUPSS_ERR_UNSUPPORTED, "too many factors in a product");
    if (npres < 1 || !pres) return fail(CUPSS_ERR_INVALID, "a term needs at least one prefactor");
    if ((int)s->f[field].terms.size() >= MAX_TERM_REFS) return fail(CUPSS_ERR_UNSUPPORTED, "too many terms");
    if (field_pres_count(s, field, npres) > MAX_PRES) return fail(CUPSS_ERR_UNSUPPORTED, "too many prefactors");
    for (int i = 0; i < nfactors; i++)
        if (factors[i] < 0 || factors[i] >= (int)s->f.size()) return fail(CUPSS_ERR_INVALID, "bad factor id");
    TermRec T;
    T.field = field;
    T.fac.assign(factors, factors + nfactors);
    T.pres.assign(pres, pres + npres);
    s->t.push_back(T);
    s->f[field].terms.push_back((int)s->t.size() - 1);
    if (term_id) *term_id = (int)s->t.size() - 1;
    s->prepared = false;
    s->gen++;
    return CUPSS_OK;
}

extern "C" int cupss_term_set_prefactors(cupss_system s, int term, const cupss_pres *pres, int npres)
{
    if (int r = check_sys(s)) return r;
    if (term < 0 || term >= (int)s->t.size()) return fail(CUPSS_ERR_INVALID, "term id out of range");
    if (npres < 1 || !pres) return fail(CUPSS_ERR_INVALID, "a term needs at least one prefactor");
    TermRec &T = s->t[term];
    if (field_pres_count(s, T.field, npres - (int)T.pres.size()) > MAX_PRES)
        return fail(CUPSS_ERR_UNSUPPORTED, "too many prefactors");
    const bool shape = same_shape(T.pres, pres, npres);
    T.pres.assign(pres, pres + npres);
    parameters_changed(s, shape);
    return CUPSS_OK;
}

extern "C" int cupss_field_set_noise(cupss_system s, int field, cupss_pres amplitude, uint64_t seed)
{
    if (int r = check_field(s, field)) return r;
    s->f[field].noisy = 1;
    s->f[field].noise = amplitude;
    s->f[field].seed = seed;
    s->prepared = false;
    s->gen++;
    return CUPSS_OK;
}

extern "C" int cupss_field_set_hooks(cupss_system s, int field, cupss_real_hook real_hook,
                                     cupss_spec_hook spec_hook, void *user)
{
    if (int r = check_field(s, field)) return r;
    FieldRec &F = s->f[field];
    // the C++ evolver re-registers its hooks before every step: only a
    // change alters the step program (and invalidates plans and graphs)
    if (F.real_hook == real_hook && F.spec_hook == spec_hook && F.hook_user == user) return CUPSS_OK;
    F.real_hook = real_hook;
    F.spec_hook = spec_hook;
    F.hook_user = user;
    s->gen++;
    return CUPSS_OK;
}

extern "C" int cupss_field_set_device_hooks(cupss_system s, int field, cupss_real_hook_dev real_hook,
                                            cupss_spec_hook_dev spec_hook, void *user)
{
    if (int r = check_field(s, field)) return r;
    FieldRec &F = s->f[field];
    if (F.real_dev == real_hook && F.spec_dev == spec_hook && F.dev_user == user) return CUPSS_OK;
    F.real_dev = real_hook;
    F.spec_dev = spec_hook;
    F.dev_user = user;
    s->gen++;
    return CUPSS_OK;
}

extern "C" int cupss_field_set_real(cupss_system s, int field, const float *host)
{
    if (int r = check_field(s, field)) return r;
    if (int r = fused_flush(s)) return r;
    s->cont_step = ~0ull;
    if (!host) return fail(CUPSS_ERR_INVALID, "null host pointer");
    // the dealiased copy keeps its old values (prepareProblem re-runs only
    // toComp, src/evolver.cpp:114-117): write them out before the spectrum
    // changes, and mark them stale
    if (int r = materialize(s, field)) return r;
    if (s->f[field].dstate == 1) s->f[field].dstate = 3;
    HIP_TRY(hipMemcpyAsync(s->rtmp, host, sizeof(float) * s->g.n, hipMemcpyHostToDevice, s->stream));
    R2CSource src{};
    src.kind = LOAD_REAL;
    src.f[0] = s->rtmp;
    src.nf = 1;
    FieldRec &F = s->f[field];
    HIP_TRY(r2c(s->g, src, F.spec[F.cur], s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CUPSS_OK;
}

extern "C" int cupss_field_get_real(cupss_system s, int field, float *host)
{
    if (int r = check_field(s, field)) return r;
    if (!host) return fail(CUPSS_ERR_INVALID, "null host pointer");
    if (int r = fused_flush(s)) return r;
    FieldRec &F = s->f[field];
    HIP_TRY(c2r(s->g, F.spec[F.cur], 0, 1.0f / (float)s->g.n, s->rtmp, s->scratch, s->stream));
    HIP_TRY(hipMemcpyAsync(host, s->rtmp, sizeof(float) * s->g.n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CUPSS_OK;
}

extern "C" int cupss_field_get_dealiased(cupss_system s, int field, float *host)
{
    if (int r = check_field(s, field)) return r;
    if (!host) return fail(CUPSS_ERR_INVALID, "null host pointer");
    if (int r = fused_flush(s)) return r;
    FieldRec &F = s->f[field];
    if (!F.dreal) {
        std::memset(host, 0, sizeof(float) * s->g.n);
        return CUPSS_OK;
    }
    if (int r = materialize(s, field)) return r;
    HIP_TRY(hipMemcpyAsync(host, F.dreal, sizeof(float) * s->g.n, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CUPSS_OK;
}

extern "C" int cupss_field_get_spectrum(cupss_system s, int field, float *host)
{
    if (int r = check_field(s, field)) return r;
    if (int r = fused_flush(s)) return r;
    FieldRec &F = s->f[field];
    HIP_TRY(hipMemcpyAsync(host, F.spec[F.cur], sizeof(float2) * s->g.nh, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CUPSS_OK;
}

extern "C" int cupss_field_set_spectrum(cupss_system s, int field, const float *host)
{
    if (int r = check_field(s, field)) return r;
    if (int r = fused_flush(s)) return r;
    s->cont_step = ~0ull;
    FieldRec &F = s->f[field];
    HIP_TRY(hipMemcpyAsync(F.spec[F.cur], host, sizeof(float2) * s->g.nh, hipMemcpyHostToDevice, s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CUPSS_OK;
}

extern "C" long cupss_system_half_bins(cupss_system s) { return s ? s->g.nh : 0; }

extern "C" int cupss_field_spectrum_device(cupss_system s, int field, void **dev_ptr)
{
    if (int r = check_field(s, field)) return r;
    if (int r = fused_flush(s)) return r;
    *dev_ptr = s->f[field].spec[s->f[field].cur];
    s->cont_step = ~0ull;  // the caller may write the spectrum
    return CUPSS_OK;
}

extern "C" int cupss_system_prepare(cupss_system s)
{
    if (int r = check_sys(s)) return r;
    // term::prepareDevice (src/term_init.cpp:108-133): factors of products of
    // two or more fields are dealiased with order max(product sizes)
    for (auto &F : s->f) F.order = 0;
    for (auto &T : s->t)
        if (T.fac.size() >= 2)
            for (int c : T.fac) {
                FieldRec &G = s->f[c];
                int p = (int)T.fac.size();
                if (G.order < p) G.order = p;
            }
    for (auto &F : s->f) {
        if (F.order > 0 && !F.dreal) {
            if (hipMalloc(&F.dreal, sizeof(float) * s->g.n) != hipSuccess)
                return fail(CUPSS_ERR_NOMEM, "device allocation failed");
            // zero until the field's first update (the reference's step-0 state)
            HIP_TRY(hipMemsetAsync(F.dreal, 0, sizeof(float) * s->g.n, s->stream));
        }
        if (F.noisy && !F.eta) {
            if (hipMalloc(&F.eta, sizeof(float2) * s->g.nh) != hipSuccess)
                return fail(CUPSS_ERR_NOMEM, "device allocation failed");
        }
        bool hooks = F.real_hook || F.spec_hook;
        if (hooks && !s->full) {
            if (hipMalloc(&s->full, sizeof(float2) * s->g.n) != hipSuccess)
                return fail(CUPSS_ERR_NOMEM, "device allocation failed");
        }
    }
    for (auto &T : s->t)
        if (T.fac.size() >= 2 && !T.xhat)
            if (hipMalloc(&T.xhat, sizeof(float2) * s->g.nh) != hipSuccess)
                return fail(CUPSS_ERR_NOMEM, "device allocation failed");
    HIP_TRY(hipStreamSynchronize(s->stream));
    plan_fused(s);
    s->gen++;
    s->prepared = true;
    return CUPSS_OK;
}

extern "C" int cupss_field_aliasing_order(cupss_system s, int field)
{
    if (check_field(s, field)) return 0;
    return s->f[field].order;
}

// ------------------------------------------------------------ the step

// host-side precomputation of one field's update descriptor
int cupss::build_update(cupss_system s, int fi, UpdateArgs &a)
{
    FieldRec &F = s->f[fi];
    std::memset(&a, 0, sizeof(a));
    a.g = s->g;
    a.cur = F.spec[F.cur];
    a.out = F.spec[F.cur ^ 1];
    a.noise = F.eta;
    a.dynamic = F.dynamic;
    a.nterms = (int)F.terms.size();
    a.nimp = (int)F.imp.size();
    a.noisy = F.noisy;
    a.dt = s->dt;
    a.sdt = 1.0f / std::sqrt(s->dt);
    if (F.noisy) {
        // src/field_init.cpp:239-241,269: dx recomputed from stepq, float sqrt
        const float PI = 3.1415926535f;
        float dxr = (float)(2.0 * (double)PI / (double)(s->g.sqx * (float)s->g.sx));
        float dyr = (float)(2.0 * (double)PI / (double)(s->g.sqy * (float)s->g.sy));
        float dzr = (float)(2.0 * (double)PI / (double)(s->g.sqz * (float)s->g.sz));
        a.noise_base = std::sqrt(s->dt / (dxr * dyr * dzr) * F.noise.preFactor);
        a.noise_q2n = F.noise.q2n;
        a.noise_invq = F.noise.invq;
    }
    int np = 0;
    for (int k = 0; k < a.nterms; k++) {
        TermRec &T = s->t[F.terms[k]];
        TermRefDev &R = a.terms[k];
        if (T.fac.size() >= 2) {
            R.kind = TERM_BUFFER;
            R.X = T.xhat;
        } else if (T.fac.size() == 1) {
            R.kind = TERM_FIELD;
            FieldRec &G = s->f[T.fac[0]];
            R.X = G.spec[G.cur];
        } else {
            R.kind = TERM_CONST;
            R.X = nullptr;
        }
        // term::precomputePrefactors (src/term_init.cpp:135-200)
        const int m = (T.pres[0].iqx + T.pres[0].iqy + T.pres[0].iqz) % 2;
        R.mult_i = m;
        R.pres0 = np;
        R.npres = (int)T.pres.size();
        for (const cupss_pres &p : T.pres) {
            const int iu = p.iqx + p.iqy + p.iqz;
            if (iu % 2 != m) std::fprintf(stderr, "PANIC: Inconsistent powers of I in prefactors\n");
            const int negate = -2 * (((iu - m) / 2) % 2) + 1;
            PresDev &P = a.pres[np++];
            P.c = p.preFactor * (float)negate;
            P.q2n = p.q2n;
            P.a = p.iqx;
            P.b = p.iqy;
            P.cz = p.iqz;
            P.invq = p.invq;
        }
    }
    for (int k = 0; k < a.nimp; k++) {
        PresDev &P = a.imp[k];
        P.c = F.imp[k].preFactor;
        P.q2n = F.imp[k].q2n;
        P.a = F.imp[k].iqx;
        P.b = F.imp[k].iqy;
        P.cz = F.imp[k].iqz;
        P.invq = F.imp[k].invq;
    }
    return CUPSS_OK;
}

__global__ void k_half_to_full(const float2 *__restrict__ h, float2 *__restrict__ full, Geom g)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.n) return;
    const int i = (int)(idx % g.sx);
    const long r = idx / g.sx;
    const int j = (int)(r % g.sy);
    const int k = (int)(r / g.sy);
    if (i <= g.nc) {
        full[idx] = h[((long)k * g.sy + j) * g.nc1 + i];
    } else {
        const int mi = g.sx - i, mj = (g.sy - j) % g.sy, mk = (g.sz - k) % g.sz;
        const float2 v = h[((long)mk * g.sy + mj) * g.nc1 + mi];
        full[idx] = make_float2(v.x, -v.y);
    }
}

// projection onto real fields: half(k) = (F(k) + conj(F(-k))) / 2
__global__ void k_full_to_half(const float2 *__restrict__ full, float2 *__restrict__ h, Geom g)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= g.nh) return;
    const int i = (int)(idx % g.nc1);
    const long r = idx / g.nc1;
    const int j = (int)(r % g.sy);
    const int k = (int)(r / g.sy);
    const float2 a = full[((long)k * g.sy + j) * g.sx + i];
    const int mi = (g.sx - i) % g.sx, mj = (g.sy - j) % g.sy, mk = (g.sz - k) % g.sz;
    const float2 b = full[((long)mk * g.sy + mj) * g.sx + mi];
    h[idx] = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
}

__global__ void k_real_to_pair(const float *__restrict__ r, float2 *__restrict__ p, long n)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) p[idx] = make_float2(r[idx], 0.0f);
}

__global__ void k_pair_to_real(const float2 *__restrict__ p, float *__restrict__ r, long n)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) r[idx] = p[idx].x;
}

static unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

// field::setRHS after the update (src/field.cpp:48-89)
static int post_update(cupss_system s, int fi)
{
    FieldRec &F = s->f[fi];
    const Geom &g = s->g;
    const float norm = 1.0f / (float)g.n;
    float2 *spec = F.spec[F.cur];
    // device-side hooks: the library's own buffers, stream-ordered
    if (F.spec_dev) F.spec_dev(F.dev_user, fi, spec, g.sx, g.sy, g.sz, (void *)s->stream);
    if (F.real_dev) {
        HIP_TRY(c2r(g, spec, 0, norm, s->rtmp, s->scratch, s->stream));
        F.real_dev(F.dev_user, fi, s->rtmp, g.sx, g.sy, g.sz, 0, (void *)s->stream);
        if (F.order > 0) {
            HIP_TRY(c2r(g, spec, F.order, norm, F.dreal, s->scratch, s->stream));
            F.real_dev(F.dev_user, fi, F.dreal, g.sx, g.sy, g.sz, 1, (void *)s->stream);
        }
        HIP_TRY(hipGetLastError());
        R2CSource src{};
        src.kind = LOAD_REAL;
        src.f[0] = s->rtmp;
        src.nf = 1;
        HIP_TRY(r2c(g, src, spec, s->stream));
        if (F.order > 0) F.dstate = 1;
        if (!F.spec_hook && !F.real_hook) return CUPSS_OK;
    }
    if (F.spec_hook) {
        hipLaunchKernelGGL(k_half_to_full, dim3(nblk(g.n)), dim3(256), 0, s->stream, spec, s->full, g);
        HIP_TRY(hipStreamSynchronize(s->stream));
        F.spec_hook(F.hook_user, fi, s->full);
        HIP_TRY(hipDeviceSynchronize());
        hipLaunchKernelGGL(k_full_to_half, dim3(nblk(g.nh)), dim3(256), 0, s->stream, s->full, spec, g);
    }
    if (F.real_hook) {
        // real copy -> hook -> back to the spectrum
        HIP_TRY(c2r(g, spec, 0, norm, s->rtmp, s->scratch, s->stream));
        hipLaunchKernelGGL(k_real_to_pair, dim3(nblk(g.n)), dim3(256), 0, s->stream, s->rtmp, s->full, g.n);
        HIP_TRY(hipStreamSynchronize(s->stream));
        F.real_hook(F.hook_user, fi, s->full, 0);
        HIP_TRY(hipDeviceSynchronize());
        if (F.order > 0) {
            HIP_TRY(c2r(g, spec, F.order, norm, F.dreal, s->scratch, s->stream));
            // the dealiased copy gets the hook as well; it goes through `full` too
            hipLaunchKernelGGL(k_pair_to_real, dim3(nblk(g.n)), dim3(256), 0, s->stream, s->full, s->rtmp, g.n);
            hipLaunchKernelGGL(k_real_to_pair, dim3(nblk(g.n)), dim3(256), 0, s->stream, F.dreal, s->full, g.n);
            HIP_TRY(hipStreamSynchronize(s->stream));
            F.real_hook(F.hook_user, fi, s->full, 1);
            HIP_TRY(hipDeviceSynchronize());
            hipLaunchKernelGGL(k_pair_to_real, dim3(nblk(g.n)), dim3(256), 0, s->stream, s->full, F.dreal, g.n);
        } else {
            hipLaunchKernelGGL(k_pair_to_real, dim3(nblk(g.n)), dim3(256), 0, s->stream, s->full, s->rtmp, g.n);
        }
        R2CSource src{};
        src.kind = LOAD_REAL;
        src.f[0] = s->rtmp;
        src.nf = 1;
        HIP_TRY(r2c(g, src, spec, s->stream));
        if (F.order > 0) F.dstate = 1;
        return CUPSS_OK;
    }
    if (F.order > 0) {
        HIP_TRY(c2r(g, spec, F.order, norm, F.dreal, s->scratch, s->stream));
        F.dstate = 1;
    }
    return CUPSS_OK;
}

// write the dealiased copy of a field whose copy is current but not stored
int cupss::materialize(cupss_system s, int fi)
{
    if (int r = fused_flush(s)) return r;
    FieldRec &F = s->f[fi];
    if (F.dstate != 2 || !F.dreal || F.order <= 0) return CUPSS_OK;
    HIP_TRY(c2r(s->g, F.spec[F.cur], F.order, 1.0f / (float)s->g.n, F.dreal, s->scratch, s->stream));
    F.dstate = 1;
    return CUPSS_OK;
}

static int run_phase(cupss_system s, int dynamic)
{
    const Geom &g = s->g;
    const int nf = (int)s->f.size();
    // 1. noise (field::createNoise, src/field.cpp:300-316)
    for (int i = 0; i < nf; i++) {
        FieldRec &F = s->f[i];
        if (F.dynamic != dynamic || !F.noisy) continue;
        if (s->noise_mode == 1) {
            HIP_TRY(spectral_noise_fill(g, F.eta, F.seed, s->step, (uint32_t)i, s->stream));
            continue;
        }
        R2CSource src{};
        src.kind = LOAD_NOISE;
        src.seed = F.seed;
        src.step = s->step;
        src.stream = (uint32_t)i;
        HIP_TRY(r2c(g, src, F.eta, s->stream));
    }
    // 2. product terms (term::computeProduct + toComp, src/term.cpp:48-102)
    for (int i = 0; i < nf; i++) {
        FieldRec &F = s->f[i];
        if (F.dynamic != dynamic) continue;
        for (int ti : F.terms) {
            TermRec &T = s->t[ti];
            if (T.fac.size() < 2) continue;
            R2CSource src{};
            src.kind = LOAD_PRODUCT;
            src.nf = (int)T.fac.size();
            for (int c = 0; c < src.nf; c++) src.f[c] = s->f[T.fac[c]].dreal;
            HIP_TRY(r2c(g, src, T.xhat, s->stream));
        }
    }
    // 3. updates (setDynamic_k / setNotDynamic_k)
    for (int i = 0; i < nf; i++) {
        if (s->f[i].dynamic != dynamic) continue;
        UpdateArgs a;
        build_update(s, i, a);
        HIP_TRY(update(a, s->stream));
    }
    // 4. swap
    for (int i = 0; i < nf; i++)
        if (s->f[i].dynamic == dynamic) s->f[i].cur ^= 1;
    // 5. dealias + hooks
    for (int i = 0; i < nf; i++)
        if (s->f[i].dynamic == dynamic)
            if (int r = post_update(s, i)) return r;
    return CUPSS_OK;
}

int cupss::general_constraint_phase(cupss_system s)
{
    if (int r = fused_flush(s)) return r;
    for (int i = 0; i < (int)s->f.size(); i++)
        if (int r = materialize(s, i)) return r;
    return run_phase(s, 0);
}

int cupss::general_steps(cupss_system s, int nsteps)
{
    if (int r = fused_flush(s)) return r;
    if (nsteps > 0) s->cont_step = ~0ull;
    for (int i = 0; i < (int)s->f.size(); i++)
        if (int r = materialize(s, i)) return r;
    for (int n = 0; n < nsteps; n++) {
        if (int r = run_phase(s, 0)) return r;
        if (int r = run_phase(s, 1)) return r;
        s->step += 1;
    }
    return CUPSS_OK;
}

static bool hooks_present(cupss_system s)
{
    for (auto &F : s->f)
        if (F.real_hook || F.spec_hook || F.real_dev || F.spec_dev) return true;
    return false;
}

static void drop_general_graph(cupss_system s)
{
    if (s->gexec) hipGraphExecDestroy(s->gexec);
    if (s->graph) hipGraphDestroy(s->graph);
    s->gexec = nullptr;
    s->graph = nullptr;
}

// The general step program, replayed from a hipGraph of K steps where that
// is valid: the launches of a step depend on host state only through the
// buffer parities (each field's spectrum alternates between two buffers;
// after K even steps they are back where they started) and, for noise,
// through the step counter -- so noisy systems and systems with hooks run
// direct launches, and a replay starts only at the parities of the capture
// (one direct step realigns them).  Capturing records K steps without
// running them (the host state already moves on by K steps); launching the
// graph once then makes the device catch up.
static bool graph_parity(cupss_system s)
{
    if (s->graph_cur.size() != s->f.size()) return false;
    for (size_t i = 0; i < s->f.size(); i++)
        if (s->f[i].cur != s->graph_cur[i]) return false;
    return true;
}

static int general_advance(cupss_system s, int nsteps)
{
    if (int r = fused_flush(s)) return r;
    const int K = s->graph_steps;
    bool noisy = false;
    for (auto &F : s->f) noisy = noisy || F.noisy;
    if (K < 2 || nsteps < K || noisy || hooks_present(s)) return general_steps(s, nsteps);
    for (int i = 0; i < (int)s->f.size(); i++)
        if (int r = materialize(s, i)) return r;
    if (s->gexec && (s->graph_gen != s->gen || s->graph_vgen != s->vgen || s->graph_stream != s->stream))
        drop_general_graph(s);
    int done = 0;
    if (s->gexec && !graph_parity(s)) {
        if (int r = general_steps(s, 1)) return r;
        done = 1;
    }
    if (!s->gexec) {
        s->graph_cur.clear();
        for (auto &F : s->f) s->graph_cur.push_back(F.cur);
        if (!twiddles(s->stream)) return fail(CUPSS_ERR_NOMEM, "twiddle tables");
        HIP_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        const int rc = general_steps(s, K);
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamEndCapture(s->stream, &g));
        if (rc != CUPSS_OK) {
            hipGraphDestroy(g);
            return rc;
        }
        s->graph = g;
        HIP_TRY(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
        s->graph_gen = s->gen;
        s->graph_vgen = s->vgen;
        s->graph_stream = s->stream;
        HIP_TRY(hipGraphLaunch(s->gexec, s->stream));
        done = K;
    }
    while (nsteps - done >= K) {
        HIP_TRY(hipGraphLaunch(s->gexec, s->stream));
        s->step += (uint64_t)K;
        done += K;
    }
    return general_steps(s, nsteps - done);
}


// The fused 3D step draws noise in spectral space only.
bool cupss::fused_noise_ok(cupss_system s)
{
    if (!s->plan.ok || s->noise_mode == 1) return true;
    return !(s->g.sz > 1 && s->f[s->plan.phi].noisy);
}

// a dynamic field of the fused plan holds a stale dealiased copy
bool cupss::fused_stale(cupss_system s)
{
    if (s->f[s->plan.phi].dstate == 3) return true;
    for (int k = 0; k < s->plan.nxdyn; k++)
        if (s->f[s->plan.cons[k]].dstate == 3) return true;
    return false;
}

extern "C" int cupss_system_advance(cupss_system s, int nsteps)
{
    if (int r = check_sys(s)) return r;
    if (!s->prepared) return fail(CUPSS_ERR_STATE, "cupss_system_prepare must be called before advancing");
    if (nsteps <= 0) return CUPSS_OK;
    if (s->path_mode == 0 && s->plan.ok && !hooks_present(s) && fused_noise_ok(s)) {
        // a stale dealiased copy is consumed by one general step first
        if (fused_stale(s)) {
            if (int r = general_steps(s, 1)) return r;
            nsteps -= 1;
        }
        return fused_steps(s, nsteps);
    }
    return general_advance(s, nsteps);
}

extern "C" int cupss_system_set_graph_steps(cupss_system s, int k)
{
    if (int r = check_sys(s)) return r;
    if (k != 0 && (k < 2 || (k & 1))) return fail(CUPSS_ERR_INVALID, "graph steps must be 0 or an even number >= 2");
    s->graph_steps = k;
    drop_general_graph(s);
    return CUPSS_OK;
}

extern "C" int cupss_system_profile_kernel(cupss_system s, int on)
{
    if (int r = check_sys(s)) return r;
    return fused_profile(s, on);
}

extern "C" int cupss_system_kernel_time(cupss_system s, float *avg_ms, double *points, char *name, int len)
{
    if (int r = check_sys(s)) return r;
    return fused_kernel_time(s, avg_ms, points, name, len);
}

extern "C" int cupss_system_device_bytes(cupss_system s, size_t *bytes)
{
    if (int r = check_sys(s)) return r;
    if (!bytes) return fail(CUPSS_ERR_INVALID, "null bytes");
    const size_t spec = sizeof(float2) * (size_t)s->g.nh, real = sizeof(float) * (size_t)s->g.n;
    size_t b = spec + real;  // scratch, rtmp
    if (s->full) b += sizeof(float2) * (size_t)s->g.n;
    for (auto &F : s->f) {
        b += 2 * spec;
        if (F.dreal) b += real;
        if (F.eta) b += spec;
    }
    for (auto &T : s->t)
        if (T.xhat) b += spec;
    if (s->plan.ok && s->g.sz > 1)  // the 3D fused path's buffers (Z_f, W_f, X_g)
        b += spec * (2 * s->plan.real.size() + s->plan.groups.size());
    else if (s->plan.ok)  // the fused path's buffers, allocated at its first run
        b += spec * (2 + 2 * s->plan.cons.size() + s->plan.real.size() + s->plan.groups.size());
    *bytes = b;
    return CUPSS_OK;
}

extern "C" int cupss_system_set_path(cupss_system s, int mode)
{
    if (int r = check_sys(s)) return r;
    if (mode != 0 && mode != 1) return fail(CUPSS_ERR_INVALID, "path mode is 0 (auto) or 1 (general)");
    s->path_mode = mode;
    return CUPSS_OK;
}

extern "C" int cupss_system_set_noise_mode(cupss_system s, int mode)
{
    if (int r = check_sys(s)) return r;
    if (mode != 0 && mode != 1) return fail(CUPSS_ERR_INVALID, "noise mode is 0 (real space) or 1 (spectral)");
    s->noise_mode = mode;
    s->gen++;
    return CUPSS_OK;
}

extern "C" int cupss_system_fused_eligible(cupss_system s)
{
    if (check_sys(s)) return 0;
    return s->plan.ok && s->g.sz == 1 && !hooks_present(s) && fused_noise_ok(s) ? 1 : 0;
}

// ------------------------------------------------------------ fused path

static bool same_pres(const std::vector<cupss_pres> &a, const std::vector<cupss_pres> &b)
{
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); i++)
        if (std::memcmp(&a[i], &b[i], sizeof(cupss_pres)) != 0) return false;
    return true;
}

static bool single_monomial(const std::vector<cupss_pres> &a)
{
    return a.size() == 1;
}

static bool same_monomial(const cupss_pres &a, const cupss_pres &b)
{
    return a.q2n == b.q2n && a.iqx == b.iqx && a.iqy == b.iqy && a.iqz == b.iqz && a.invq == b.invq;
}

// Does the fused step apply, and how?  One dynamic field phi; every other
// field is either derived (psi = M_psi phi: one single-field term of phi,
// no implicit operator, no noise) or stored (a constraint without noise
// whose product terms read only derived or stored fields -- a product with
// phi as a factor would need phi's dealiased copy one step early).  phi's
// own terms may read any field.  3D systems (their own four-pass step, run
// alone) take no stored fields.
static void plan_fused(cupss_system s)
{
    auto &P = s->plan;
    P = FusedPlan();
    if (!fused_supported(s->g.sx, s->g.sy, s->g.sz)) return;
    const int nf = (int)s->f.size();
    // the first dynamic field is phi; in 2D the others are stored fields
    int phi = -1;
    std::vector<char> dyn(nf, 0);
    for (int i = 0; i < nf; i++)
        if (s->f[i].dynamic) {
            dyn[i] = 1;
            if (phi < 0) {
                phi = i;
            } else if (s->g.sz > 1) {
                return;
            }
        }
    if (phi < 0) return;
    std::vector<char> derived(nf, 0), stored(nf, 0);
    std::vector<int> base(nf, -1);
    for (int i = 0; i < nf; i++) {
        if (i == phi) continue;
        if (dyn[i]) {
            stored[i] = 1;
            continue;
        }
        const FieldRec &F = s->f[i];
        if (!F.noisy && F.imp.empty() && F.terms.size() == 1) {
            const TermRec &T = s->t[F.terms[0]];
            if (T.fac.size() == 1 && dyn[T.fac[0]] && (int)T.pres.size() <= FUSED_MAX_POOL / 8) {
                derived[i] = 1;
                base[i] = T.fac[0];
                continue;
            }
        }
        // no terms: it keeps its value (a coefficient field), divided by its
        // implicit operator's sum every step if it has one (setNotDynamic_k)
        if (F.noisy && s->g.sz > 1) return;
        stored[i] = 3;
    }
    // A constraint whose products read a dynamic field is "fresh" (kind 2):
    // the reference's phase A of step m reads the dynamic fields of step m,
    // so it is solved at the start of the pass that advances them, from
    // products of dynamic fields only and single terms of dynamic fields or
    // constants (what that pass holds); nothing reads it in real space.
    // The other constraints (kind 3) are solved for the next step.
    for (int i = 0; i < nf; i++) {
        if (stored[i] != 3) continue;
        bool reads_dyn = false;
        for (int ti : s->f[i].terms)
            if (s->t[ti].fac.size() >= 2)
                for (int c : s->t[ti].fac) reads_dyn = reads_dyn || dyn[c];
        if (!reads_dyn) continue;
        for (int ti : s->f[i].terms)
            for (int c : s->t[ti].fac)
                if (!dyn[c]) return;
        stored[i] = 2;
    }
    // product factors: dynamic fields' products read anything but fresh
    // constraints; the other constraints' products read derived or kind-3
    // constraint fields only
    std::vector<std::vector<int>> prods_of(nf);
    for (int i = 0; i < nf; i++) {
        if (derived[i]) continue;
        int single = 0;
        for (int ti : s->f[i].terms) {
            const TermRec &T = s->t[ti];
            if (T.fac.size() >= 2) {
                if ((int)T.fac.size() > 8) return;
                for (int c : T.fac) {
                    if (stored[c] == 2) return;
                    if (stored[i] == 3 && dyn[c]) return;
                    if (std::find(P.real.begin(), P.real.end(), c) == P.real.end()) P.real.push_back(c);
                }
                prods_of[i].push_back(ti);
            } else {
                single++;
            }
        }
        if (single > FUSED_MAX_TTERMS) return;
    }
    if ((int)P.real.size() > FUSED_MAX_REAL) return;
    if (!fused_shapes_exist(s->g.sx, s->g.sy, s->g.sz, (int)P.real.size())) return;
    for (int kind = 1; kind <= 3; kind++)
        for (int i = 0; i < nf; i++)
            if (stored[i] == kind) P.cons.push_back(i);
    for (int i = 0; i < nf; i++) {
        if (stored[i] == 1) P.nxdyn++;
        if (stored[i] == 2) P.nfresh++;
    }
    if ((int)P.cons.size() > FUSED_MAX_CONS) return;
    if (s->g.sz > 1 && !P.cons.empty()) return;  // the 3D step has no stored fields
    // groups: phi's first, then each stored field's
    std::vector<int> order = {phi};
    order.insert(order.end(), P.cons.begin(), P.cons.end());
    int nprod = 0;
    for (int f : order) {
        std::vector<FusedGroup> gs;
        for (int ti : prods_of[f]) {
            const std::vector<cupss_pres> &pr = s->t[ti].pres;
            bool placed = false;
            for (auto &grp : gs)
                if (same_pres(grp.pres, pr)) {
                    grp.terms.push_back(ti);
                    grp.coef.push_back(1.0f);
                    placed = true;
                    break;
                }
            if (!placed) {
                FusedGroup grp;
                grp.target = f;
                grp.terms.push_back(ti);
                grp.coef.push_back(1.0f);
                grp.pres = pr;
                gs.push_back(grp);
            }
            nprod++;
        }
        // fold groups whose lists are one monomial of the same shape: the
        // coefficient moves into real space, the multiplier keeps c = 1
        for (size_t a = 0; a < gs.size(); a++) {
            if (!single_monomial(gs[a].pres)) continue;
            for (size_t b = a + 1; b < gs.size();) {
                if (single_monomial(gs[b].pres) && same_monomial(gs[a].pres[0], gs[b].pres[0])) {
                    const float ca = gs[a].pres[0].preFactor, cb = gs[b].pres[0].preFactor;
                    if (ca != 1.0f) {
                        for (float &c : gs[a].coef) c *= ca;
                        gs[a].pres[0].preFactor = 1.0f;
                    }
                    for (size_t k = 0; k < gs[b].terms.size(); k++) {
                        gs[a].terms.push_back(gs[b].terms[k]);
                        gs[a].coef.push_back(gs[b].coef[k] * cb);
                    }
                    gs.erase(gs.begin() + b);
                } else {
                    b++;
                }
            }
        }
        P.groups.insert(P.groups.end(), gs.begin(), gs.end());
    }
    if ((int)P.groups.size() > FUSED_MAX_GROUPS || nprod > FUSED_MAX_PRODS) return;
    for (int i = 0; i < nf; i++)
        if (derived[i]) {
            P.derived.push_back(i);
            P.dbase.push_back(base[i]);
        }
    if ((int)P.derived.size() > FUSED_MAX_DERIVED) return;
    // the prefactor pool must hold every multiplier of the instance
    size_t pool = 0;
    for (auto &grp : P.groups) pool += grp.pres.size();
    for (int f : order) {
        pool += s->f[f].imp.size();
        for (int ti : s->f[f].terms)
            if (s->t[ti].fac.size() < 2) pool += s->t[ti].pres.size();
    }
    for (int d : P.derived) pool += s->t[s->f[d].terms[0]].pres.size();
    if (pool > (size_t)FUSED_MAX_POOL) return;
    P.phi = phi;
    P.ok = true;
}

static bool same_plan_shape(const FusedPlan &a, const FusedPlan &b)
{
    if (a.ok != b.ok || a.phi != b.phi || a.real != b.real || a.derived != b.derived || a.dbase != b.dbase ||
        a.cons != b.cons || a.nxdyn != b.nxdyn || a.nfresh != b.nfresh || a.groups.size() != b.groups.size())
        return false;
    for (size_t i = 0; i < a.groups.size(); i++) {
        const FusedGroup &x = a.groups[i], &y = b.groups[i];
        if (x.target != y.target || x.terms != y.terms || x.coef.size() != y.coef.size() ||
            !same_shape(x.pres, y.pres.data(), (int)y.pres.size()))
            return false;
    }
    return true;
}

// evolver::updateParameter (src/evolver.cpp:386-394) through new prefactor
// values: the fused plan is made again (its groups hold the product terms'
// prefactors and folded coefficients).  When the monomials keep their shape
// and the plan its structure, only values changed: vgen moves and the
// batches and graphs refresh their descriptors without re-planning tables,
// re-tuning or re-classifying; otherwise gen moves as for any program change.
static void parameters_changed(cupss_system s, bool same)
{
    s->cont_step = ~0ull;  // real fields left by a run used the old multipliers
    if (!s->prepared) {
        s->gen++;
        return;
    }
    FusedPlan old = s->plan;
    plan_fused(s);
    if (same && same_plan_shape(old, s->plan))
        s->vgen++;
    else
        s->gen++;
}

extern "C" int cupss_system_time_advance(cupss_system s, int nsteps, float *ms)
{
    if (int r = check_sys(s)) return r;
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, s->stream));
    int r = cupss_system_advance(s, nsteps);
    HIP_TRY(hipEventRecord(b, s->stream));
    HIP_TRY(hipEventSynchronize(b));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, a, b));
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (ms) *ms = t;
    return r;
}

extern "C" int cupss_system_synchronize(cupss_system s)
{
    if (int r = check_sys(s)) return r;
    HIP_TRY(hipStreamSynchronize(s->stream));
    return CUPSS_OK;
}

extern "C" int cupss_system_get_step(cupss_system s, uint64_t *step)
{
    if (int r = check_sys(s)) return r;
    *step = s->step;
    return CUPSS_OK;
}

extern "C" int cupss_system_set_step(cupss_system s, uint64_t step)
{
    if (int r = check_sys(s)) return r;
    s->cont_step = ~0ull;
    s->step = step;
    return CUPSS_OK;
}

extern "C" int cupss_system_set_stream(cupss_system s, void *hip_stream)
{
    if (int r = check_sys(s)) return r;
    if (s->member_of) return fail(CUPSS_ERR_STATE, "a batch member runs on the batch's stream: remove it first");
    if (int r = fused_flush(s)) return r;
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->own_stream) hipStreamDestroy(s->stream);
    s->own_stream = false;
    if (hip_stream) {
        s->stream = (hipStream_t)hip_stream;
    } else {
        HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->own_stream = true;
    }
    return CUPSS_OK;
}

extern "C" int cupss_system_set_dt(cupss_system s, float dt)
{
    if (int r = check_sys(s)) return r;
    s->dt = dt;
    s->gen++;
    return CUPSS_OK;
}

// ------------------------------------------------------------ stand-alone

static int make_geom(int sx, int sy, int sz, Geom &g)
{
    if (!is_pow2(sx) || sx < 2 || sx > 4096 || !is_pow2(sy) || sy > 4096 || !is_pow2(sz) || sz > 4096)
        return fail(CUPSS_ERR_UNSUPPORTED, "unsupported grid");
    g.sx = sx;
    g.sy = sy;
    g.sz = sz;
    g.nc = sx / 2;
    g.nc1 = sx / 2 + 1;
    g.nh = (long)g.nc1 * sy * sz;
    g.n = (long)sx * sy * sz;
    return CUPSS_OK;
}

extern "C" int cupss_r2c(int sx, int sy, int sz, const float *in, void *out, void *hip_stream)
{
    Geom g{};
    if (int r = make_geom(sx, sy, sz, g)) return r;
    R2CSource src{};
    src.kind = LOAD_REAL;
    src.f[0] = in;
    src.nf = 1;
    HIP_TRY(r2c(g, src, (float2 *)out, (hipStream_t)hip_stream));
    return CUPSS_OK;
}

extern "C" int cupss_c2r(int sx, int sy, int sz, const void *in, int mask_p, float scale, float *out,
                         void *scratch, void *hip_stream)
{
    Geom g{};
    if (int r = make_geom(sx, sy, sz, g)) return r;
    if (scratch == in) return fail(CUPSS_ERR_INVALID, "scratch must not alias the input");
    HIP_TRY(c2r(g, (const float2 *)in, mask_p, scale, out, (float2 *)scratch, (hipStream_t)hip_stream));
    return CUPSS_OK;
}

__global__ void k_normals(float *out, long n, uint32_t k0, uint32_t k1, uint32_t slo, uint32_t shi, uint32_t st)
{
    const long c = 2 * ((long)blockIdx.x * blockDim.x + threadIdx.x);
    if (c >= n) return;
    const float2 z = philox_normal_pair((uint32_t)c, slo, shi, st, k0, k1);
    out[c] = z.x;
    if (c + 1 < n) out[c + 1] = z.y;
}

extern "C" int cupss_normals(uint64_t seed, uint64_t step, uint32_t stream, float *out, long n, void *hip_stream)
{
    if (n <= 0) return CUPSS_OK;
    hipLaunchKernelGGL(k_normals, dim3(nblk((n + 1) / 2)), dim3(256), 0, (hipStream_t)hip_stream, out, n,
                       (uint32_t)seed, (uint32_t)(seed >> 32), (uint32_t)step, (uint32_t)(step >> 32), stream);
    HIP_TRY(hipGetLastError());
    return CUPSS_OK;
}

extern "C" int cupss_device_alloc(size_t bytes, void **dev_ptr)
{
    if (!dev_ptr) return fail(CUPSS_ERR_INVALID, "null out");
    if (hipMalloc(dev_ptr, bytes) != hipSuccess) return fail(CUPSS_ERR_NOMEM, "hipMalloc failed");
    return CUPSS_OK;
}
extern "C" int cupss_device_free(void *p)
{
    HIP_TRY(hipFree(p));
    return CUPSS_OK;
}
extern "C" int cupss_memcpy_h2d(void *d, const void *h, size_t bytes)
{
    HIP_TRY(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
    return CUPSS_OK;
}
extern "C" int cupss_memcpy_d2h(void *h, const void *d, size_t bytes)
{
    HIP_TRY(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
    return CUPSS_OK;
}
extern "C" int cupss_memcpy_d2d(void *d, const void *s, size_t bytes)
{
    HIP_TRY(hipMemcpy(d, s, bytes, hipMemcpyDeviceToDevice));
    return CUPSS_OK;
}

extern "C" int cupss_device_synchronize(void)
{
    HIP_TRY(hipDeviceSynchronize());
    return CUPSS_OK;
}
extern "C" int cupss_device_count(int *n)
{
    HIP_TRY(hipGetDeviceCount(n));
    return CUPSS_OK;
}
