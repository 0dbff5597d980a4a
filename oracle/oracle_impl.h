/*
 * oracle_impl.h -- body of the oracle, included twice by oracle.c with
 *   REAL = float,  FN(x) = orc_f32_##x     (FFTW-float-like scheme)
 *   REAL = double, FN(x) = orc_f64_##x     (the same scheme in double)
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Every block names the reference lines it restates.
 */

typedef struct {
    REAL re, im;
} CPLX;

typedef struct {
    int field;
    int nfac;
    int fac[ORC_MAXF];
    int npres;
    orc_pres pres[ORC_MAXP];
    CPLX *term_real;
    CPLX *term_comp;
    REAL *precomp;
    int mult_by_i;
} TERM;

typedef struct {
    int dynamic;
    int noisy;
    orc_pres noise_amp;
    uint64_t seed;
    uint32_t stream;
    int nimp;
    orc_pres imp[ORC_MAXP];
    int needsaliasing;
    int aliasing_order;
    int nterms;
    int terms[ORC_MAXT];
    CPLX *real;
    CPLX *comp;
    CPLX *real_dealiased;
    CPLX *comp_dealiased;
    CPLX *noise_comp;
    REAL *precomp_implicit;
    REAL *precomp_noise;
    orc_hook hook_fourier;
    orc_hook hook_real;
    void *hook_user;
} FIELD;

struct FN(sys) {
    int sx, sy, sz, n;
    float dx, dy, dz, dt;
    float stepqx, stepqy, stepqz;
    int nfields, nterms;
    FIELD f[ORC_MAXFIELDS];
    TERM t[ORC_MAXTERMS];
    uint64_t step;
    CPLX *scratch;
};

/* ---------------------------------------------------------------- FFT */

/* In-place 1D FFT of a contiguous line of length n, unnormalised, FFTW sign
 * convention: dir = -1 forward (exp(-2 pi i k m/n)), +1 backward.  Radix-2
 * iterative for powers of two, direct DFT otherwise.  Twiddles in double. */
static void FN(fft1d)(CPLX *a, int n, int dir, CPLX *tmp)
{
    if (n <= 1) return;
    if ((n & (n - 1)) == 0) {
        /* bit reversal */
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) {
                CPLX t = a[i];
                a[i] = a[j];
                a[j] = t;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int half = len >> 1;
            for (int k = 0; k < half; k++) {
                double ang = dir * 2.0 * M_PI * (double)k / (double)len;
                REAL wr = (REAL)cos(ang), wi = (REAL)sin(ang);
                for (int s = 0; s < n; s += len) {
                    CPLX u = a[s + k];
                    CPLX v = a[s + k + half];
                    REAL tr = v.re * wr - v.im * wi;
                    REAL ti = v.re * wi + v.im * wr;
                    a[s + k].re = u.re + tr;
                    a[s + k].im = u.im + ti;
                    a[s + k + half].re = u.re - tr;
                    a[s + k + half].im = u.im - ti;
                }
            }
        }
    } else {
        for (int k = 0; k < n; k++) {
            REAL sr = 0, si = 0;
            for (int m = 0; m < n; m++) {
                long long km = ((long long)k * m) % n;
                double ang = dir * 2.0 * M_PI * (double)km / (double)n;
                REAL wr = (REAL)cos(ang), wi = (REAL)sin(ang);
                sr += a[m].re * wr - a[m].im * wi;
                si += a[m].re * wi + a[m].im * wr;
            }
            tmp[k].re = sr;
            tmp[k].im = si;
        }
        memcpy(a, tmp, sizeof(CPLX) * n);
    }
}

/* Multidimensional C2C over (sz, sy, sx), x fastest: the transform the
 * reference gets from fftwf_plan_dft_{1,2,3}d / cufftPlan{1,2,3}d
 * (src/field_init.cpp:47-120).  out may alias in. */
static void FN(fftnd)(const CPLX *in, CPLX *out, int sx, int sy, int sz, int dir)
{
    int nmax = sx > sy ? sx : sy;
    if (sz > nmax) nmax = sz;
    CPLX *line = (CPLX *)malloc(sizeof(CPLX) * nmax);
    CPLX *tmp = (CPLX *)malloc(sizeof(CPLX) * nmax);
    if (out != in) memcpy(out, in, sizeof(CPLX) * (size_t)sx * sy * sz);
    for (int k = 0; k < sz; k++)
        for (int j = 0; j < sy; j++) {
            CPLX *row = out + (size_t)k * sx * sy + (size_t)j * sx;
            FN(fft1d)(row, sx, dir, tmp);
        }
    if (sy > 1)
        for (int k = 0; k < sz; k++)
            for (int i = 0; i < sx; i++) {
                for (int j = 0; j < sy; j++) line[j] = out[(size_t)k * sx * sy + (size_t)j * sx + i];
                FN(fft1d)(line, sy, dir, tmp);
                for (int j = 0; j < sy; j++) out[(size_t)k * sx * sy + (size_t)j * sx + i] = line[j];
            }
    if (sz > 1)
        for (int j = 0; j < sy; j++)
            for (int i = 0; i < sx; i++) {
                for (int k = 0; k < sz; k++) line[k] = out[(size_t)k * sx * sy + (size_t)j * sx + i];
                FN(fft1d)(line, sz, dir, tmp);
                for (int k = 0; k < sz; k++) out[(size_t)k * sx * sy + (size_t)j * sx + i] = line[k];
            }
    free(line);
    free(tmp);
}

void FN(fft)(REAL *data, int sx, int sy, int sz, int dir)
{
    FN(fftnd)((CPLX *)data, (CPLX *)data, sx, sy, sz, dir);
}

/* ------------------------------------------------------------- set-up */

FN(sys) *FN(create)(int sx, int sy, int sz, float dx, float dy, float dz, float dt)
{
    FN(sys) *s = (FN(sys) *)calloc(1, sizeof(FN(sys)));
    s->sx = sx;
    s->sy = sy;
    s->sz = sz;
    s->n = sx * sy * sz;
    s->dx = dx;
    s->dy = dy;
    s->dz = dz;
    s->dt = dt;
    /* evolver::createField always uses the 3D field constructor
     * (src/evolver.cpp:184, src/field_init.cpp:207-213): stepq = 2*PI/(d*s)
     * in float, with d = 1 and s = 1 on the unused axes of 1D/2D systems. */
    s->stepqx = 2.0f * ORC_PI / (dx * (float)sx);
    s->stepqy = 2.0f * ORC_PI / (dy * (float)sy);
    s->stepqz = 2.0f * ORC_PI / (dz * (float)sz);
    s->scratch = (CPLX *)calloc((size_t)s->n, sizeof(CPLX));
    return s;
}

void FN(destroy)(FN(sys) *s)
{
    if (!s) return;
    for (int i = 0; i < s->nfields; i++) {
        FIELD *f = &s->f[i];
        free(f->real);
        free(f->comp);
        free(f->real_dealiased);
        free(f->comp_dealiased);
        free(f->noise_comp);
        free(f->precomp_implicit);
        free(f->precomp_noise);
    }
    for (int i = 0; i < s->nterms; i++) {
        free(s->t[i].term_real);
        free(s->t[i].term_comp);
        free(s->t[i].precomp);
    }
    free(s->scratch);
    free(s);
}

int FN(add_field)(FN(sys) *s, int dynamic)
{
    if (s->nfields >= ORC_MAXFIELDS) return -1;
    int id = s->nfields++;
    FIELD *f = &s->f[id];
    memset(f, 0, sizeof(*f));
    f->dynamic = dynamic;
    f->aliasing_order = 1; /* src/field_init.cpp:38-39 */
    f->stream = (uint32_t)id;
    size_t n = (size_t)s->n;
    f->real = (CPLX *)calloc(n, sizeof(CPLX));
    f->comp = (CPLX *)calloc(n, sizeof(CPLX));
    f->real_dealiased = (CPLX *)calloc(n, sizeof(CPLX));
    f->comp_dealiased = (CPLX *)calloc(n, sizeof(CPLX));
    f->noise_comp = (CPLX *)calloc(n, sizeof(CPLX));
    f->precomp_implicit = (REAL *)calloc(n, sizeof(REAL));
    f->precomp_noise = (REAL *)calloc(n, sizeof(REAL));
    return id;
}

void FN(add_implicit)(FN(sys) *s, int f, orc_pres p)
{
    FIELD *F = &s->f[f];
    if (F->nimp < ORC_MAXP) F->imp[F->nimp++] = p;
}

int FN(add_term)(FN(sys) *s, int f, int nfac, const int *fac, int npres, const orc_pres *pres)
{
    if (s->nterms >= ORC_MAXTERMS || nfac > ORC_MAXF || npres > ORC_MAXP || npres < 1) return -1;
    int id = s->nterms++;
    TERM *t = &s->t[id];
    memset(t, 0, sizeof(*t));
    t->field = f;
    t->nfac = nfac;
    for (int i = 0; i < nfac; i++) t->fac[i] = fac[i];
    t->npres = npres;
    for (int i = 0; i < npres; i++) t->pres[i] = pres[i];
    size_t n = (size_t)s->n;
    t->term_real = (CPLX *)calloc(n, sizeof(CPLX));
    t->term_comp = (CPLX *)calloc(n, sizeof(CPLX));
    t->precomp = (REAL *)calloc(n, sizeof(REAL));
    FIELD *F = &s->f[f];
    F->terms[F->nterms++] = id;
    return id;
}

void FN(set_noise)(FN(sys) *s, int f, orc_pres amp, uint64_t seed)
{
    s->f[f].noisy = 1;
    s->f[f].noise_amp = amp;
    s->f[f].seed = seed;
}

void FN(set_real)(FN(sys) *s, int f, const float *x)
{
    for (int i = 0; i < s->n; i++) {
        s->f[f].real[i].re = (REAL)x[i];
        s->f[f].real[i].im = 0;
    }
}

void FN(set_hooks)(FN(sys) *s, int f, orc_hook fourier, orc_hook real, void *user)
{
    s->f[f].hook_fourier = fourier;
    s->f[f].hook_real = real;
    s->f[f].hook_user = user;
}

static inline void FN(qvec)(const FN(sys) *s, int i, int j, int k, REAL *qx, REAL *qy, REAL *qz)
{
    /* q_d(i) = (i < (s+1)/2 ? i : i - s) * stepq  (e.g. src/term_init.cpp:157-160) */
    float fi = (i < (s->sx + 1) / 2 ? (float)i : (float)(i - s->sx));
    float fj = (j < (s->sy + 1) / 2 ? (float)j : (float)(j - s->sy));
    float fk = (k < (s->sz + 1) / 2 ? (float)k : (float)(k - s->sz));
    *qx = (REAL)fi * (REAL)s->stepqx;
    *qy = (REAL)fj * (REAL)s->stepqy;
    *qz = (REAL)fk * (REAL)s->stepqz;
}

/* term::precomputePrefactors, src/term_init.cpp:135-200.  The reference's
 * std::pow(float,int) promotes to double; REAL = float rounds back. */
static void FN(precompute_term)(FN(sys) *s, TERM *t)
{
    int m = (t->pres[0].iqx + t->pres[0].iqy + t->pres[0].iqz) % 2;
    for (int p = 0; p < t->npres; p++) {
        int iu = t->pres[p].iqx + t->pres[p].iqy + t->pres[p].iqz;
        if (iu % 2 != m) fprintf(stderr, "oracle: PANIC: Inconsistent powers of I in prefactors\n");
    }
    for (int k = 0; k < s->sz; k++)
        for (int j = 0; j < s->sy; j++)
            for (int i = 0; i < s->sx; i++) {
                int index = k * s->sx * s->sy + j * s->sx + i;
                REAL qx, qy, qz;
                FN(qvec)(s, i, j, k, &qx, &qy, &qz);
                REAL q2 = qx * qx + qy * qy + qz * qz;
                REAL total = 0;
                for (int p = 0; p < t->npres; p++) {
                    const orc_pres *P = &t->pres[p];
                    REAL num = (REAL)P->preFactor;
                    int iu = P->iqx + P->iqy + P->iqz;
                    int negate = -2 * (((iu - m) / 2) % 2) + 1;
                    num *= (REAL)negate;
                    if (P->q2n > 0) num = (REAL)((double)num * pow((double)q2, P->q2n));
                    if (P->iqx > 0) num = (REAL)((double)num * pow((double)qx, P->iqx));
                    if (P->iqy > 0) num = (REAL)((double)num * pow((double)qy, P->iqy));
                    if (P->iqz > 0) num = (REAL)((double)num * pow((double)qz, P->iqz));
                    if (P->invq > 0) {
                        REAL invq = 0;
                        if (index > 0) invq = (REAL)1 / (REAL)sqrt((double)q2);
                        num = (REAL)((double)num * pow((double)invq, P->invq));
                    }
                    total += num;
                }
                t->precomp[index] = total;
            }
    t->mult_by_i = m;
}

/* field::precalculateImplicit, src/field_init.cpp:237-284. */
static void FN(precalculate_implicit)(FN(sys) *s, FIELD *F)
{
    REAL dt = (REAL)s->dt;
    /* dx recomputed from stepq (field_init.cpp:239-241), double math stored to float */
    float dxr = (float)(2.0 * (double)ORC_PI / (double)(s->stepqx * (float)s->sx));
    float dyr = (float)(2.0 * (double)ORC_PI / (double)(s->stepqy * (float)s->sy));
    float dzr = (float)(2.0 * (double)ORC_PI / (double)(s->stepqz * (float)s->sz));
    for (int k = 0; k < s->sz; k++)
        for (int j = 0; j < s->sy; j++)
            for (int i = 0; i < s->sx; i++) {
                int index = k * s->sx * s->sy + j * s->sx + i;
                REAL qx, qy, qz;
                FN(qvec)(s, i, j, k, &qx, &qy, &qz);
                REAL q2 = qx * qx + qy * qy + qz * qz;
                REAL impf = F->dynamic ? (REAL)1 : (REAL)0;
                for (int p = 0; p < F->nimp; p++) {
                    const orc_pres *P = &F->imp[p];
                    REAL th = (REAL)P->preFactor;
                    if (P->q2n != 0) th = (REAL)((double)th * pow((double)q2, P->q2n));
                    if (P->invq != 0) {
                        REAL invq = 0;
                        if (i > 0 || j > 0) invq = (REAL)1 / (REAL)sqrt((double)q2); /* Q8 */
                        th = (REAL)((double)th * pow((double)invq, P->invq));
                    }
                    if (F->dynamic)
                        impf -= dt * th;
                    else
                        impf += th;
                }
                F->precomp_implicit[index] = impf;
                REAL nf = (REAL)sqrt((double)((REAL)s->dt / ((REAL)dxr * (REAL)dyr * (REAL)dzr) *
                                              (REAL)F->noise_amp.preFactor));
                if (F->noise_amp.q2n != 0) nf = (REAL)((double)nf * sqrt(pow((double)q2, F->noise_amp.q2n)));
                if (F->noise_amp.invq != 0) {
                    REAL invq = 0;
                    if (index > 0) invq = (REAL)1 / (REAL)sqrt((double)q2);
                    nf = (REAL)((double)nf * sqrt(pow((double)invq, F->noise_amp.invq)));
                }
                F->precomp_noise[index] = nf;
            }
}

/* evolver::prepareProblem, src/evolver.cpp:81-126: comp = FFT(real) for
 * every field while needsaliasing is still false (so real_dealiased stays
 * zero until the first setRHS -- the step-0 quirk), then prepareDevice
 * (dealias order, term prefactors) and precalculateImplicit. */
void FN(prepare)(FN(sys) *s)
{
    for (int i = 0; i < s->nfields; i++) {
        FIELD *F = &s->f[i];
        FN(fftnd)(F->real, F->comp, s->sx, s->sy, s->sz, -1);
    }
    for (int i = 0; i < s->nfields; i++) {
        FIELD *F = &s->f[i];
        for (int tt = 0; tt < F->nterms; tt++) {
            TERM *t = &s->t[F->terms[tt]];
            /* term::prepareDevice, src/term_init.cpp:108-133 */
            if (t->nfac != 1)
                for (int c = 0; c < t->nfac; c++) {
                    FIELD *G = &s->f[t->fac[c]];
                    G->needsaliasing = 1;
                    if (G->aliasing_order < t->nfac) G->aliasing_order = t->nfac;
                }
            FN(precompute_term)(s, t);
        }
        FN(precalculate_implicit)(s, F);
    }
    /* the step counter (and with it the noise stream) is not reset: the
     * reference's cuRAND generator continues its sequence */
}

/* ------------------------------------------------------------ the step */

/* field::createNoise, src/field.cpp:300-330: xi -> FFT -> noise_comp. */
static void FN(create_noise)(FN(sys) *s, FIELD *F)
{
    CPLX *g = s->scratch;
    for (int c = 0; c < s->n; c++) {
        g[c].re = (REAL)orc_normal(F->seed, s->step, F->stream, (uint32_t)c);
        g[c].im = 0;
    }
    FN(fftnd)(g, F->noise_comp, s->sx, s->sy, s->sz, -1);
}

/* term::update, src/term.cpp:12-25 (+ computeProduct :66-102, copyComp
 * :104-126, applyPres_vector :225-255). */
static void FN(term_update)(FN(sys) *s, TERM *t)
{
    int n = s->n;
    if (t->nfac == 1) {
        memcpy(t->term_comp, s->f[t->fac[0]].comp, sizeof(CPLX) * n);
    } else {
        for (int c = 0; c < n; c++) {
            REAL r = 1;
            for (int p = 0; p < t->nfac; p++) {
                const FIELD *G = &s->f[t->fac[p]];
                r *= (orc_plant_flag == ORC_PLANT_UNDEALIAS ? G->real : G->real_dealiased)[c].re;
            }
            t->term_real[c].re = r;
            t->term_real[c].im = 0;
        }
        FN(fftnd)(t->term_real, t->term_comp, s->sx, s->sy, s->sz, -1);
    }
    for (int c = 0; c < n; c++) {
        t->term_comp[c].re *= t->precomp[c];
        t->term_comp[c].im *= t->precomp[c];
        if (t->mult_by_i) {
            REAL aux = t->term_comp[c].re;
            t->term_comp[c].re = -t->term_comp[c].im;
            t->term_comp[c].im = aux;
        }
    }
}

static void FN(update_terms)(FN(sys) *s, FIELD *F)
{
    if (F->noisy) FN(create_noise)(s, F);
    for (int tt = 0; tt < F->nterms; tt++) FN(term_update)(s, &s->t[F->terms[tt]]);
}

/* field::setRHS, src/field.cpp:25-92 with the GPU kernels setDynamic_k /
 * setNotDynamic_k / dealias_k / normalize_k (src/field_kernels.cu). */
static void FN(set_rhs)(FN(sys) *s, FIELD *F)
{
    int sx = s->sx, sy = s->sy, sz = s->sz, n = s->n;
    REAL dt = (REAL)s->dt;
    if (F->dynamic) {
        /* setDynamic_k, field_kernels.cu:199-227 */
        for (int c = 0; c < n; c++) {
            for (int tt = 0; tt < F->nterms; tt++) {
                CPLX *tc = s->t[F->terms[tt]].term_comp;
                F->comp[c].re += dt * tc[c].re;
                F->comp[c].im += dt * tc[c].im;
            }
            if (F->noisy) {
                F->comp[c].re += F->precomp_noise[c] * F->noise_comp[c].re;
                F->comp[c].im += F->precomp_noise[c] * F->noise_comp[c].im;
            }
            if (F->nimp > 0) {
                F->comp[c].re /= F->precomp_implicit[c];
                F->comp[c].im /= F->precomp_implicit[c];
            }
        }
    } else {
        /* setNotDynamic_k, field_kernels.cu:130-197 */
        REAL sdt = (REAL)1 / (REAL)sqrt((double)dt);
        for (int k = 0; k < sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++) {
                    int c = k * sx * sy + j * sx + i;
                    for (int tt = 0; tt < F->nterms; tt++) {
                        CPLX *tc = s->t[F->terms[tt]].term_comp;
                        if (tt == 0) {
                            F->comp[c] = tc[c];
                        } else {
                            F->comp[c].re += tc[c].re;
                            F->comp[c].im += tc[c].im;
                        }
                    }
                    if (F->noisy) {
                        REAL a = sdt * F->precomp_noise[c];
                        if (F->nterms == 0) {
                            F->comp[c].re = a * F->noise_comp[c].re;
                            F->comp[c].im = a * F->noise_comp[c].im;
                        } else {
                            F->comp[c].re += a * F->noise_comp[c].re;
                            F->comp[c].im += a * F->noise_comp[c].im;
                        }
                    }
                    if (F->nimp > 0 && c > 0) {
                        REAL qx, qy, qz;
                        FN(qvec)(s, i, j, k, &qx, &qy, &qz);
                        REAL q2 = qx * qx + qy * qy + qz * qz;
                        REAL impf = 0;
                        for (int p = 0; p < F->nimp; p++) {
                            REAL th = (REAL)F->imp[p].preFactor;
                            if (F->imp[p].q2n != 0) th *= (REAL)pow((double)q2, F->imp[p].q2n);
                            if (F->imp[p].invq != 0) {
                                REAL invq = (REAL)1 / (REAL)sqrt((double)q2);
                                th *= (REAL)pow((double)invq, F->imp[p].invq);
                            }
                            impf += th;
                        }
                        F->comp[c].re /= impf;
                        F->comp[c].im /= impf;
                    }
                }
    }
    /* Fourier-space callback on the updated spectrum (field.cpp:48-58) */
    if (F->hook_fourier) F->hook_fourier(F->hook_user, F->comp, sx, sy, sz);
    /* dealias (field.cpp:59-61 -> dealias_k, field_kernels.cu:229-256) */
    if (F->needsaliasing) {
        int p1 = F->aliasing_order + 1;
        for (int k = 0; k < sz; k++)
            for (int j = 0; j < sy; j++)
                for (int i = 0; i < sx; i++) {
                    int c = k * sx * sy + j * sx + i;
                    int ni = i > sx / 2 ? i - sx : i;
                    int nj = j > sy / 2 ? j - sy : j;
                    int nk = k > sz / 2 ? k - sz : k;
                    int cut = abs(ni) > sx / p1 || abs(nj) > sy / p1 || abs(nk) > sz / p1;
                    if (orc_ref_cpu_dealias_bug_flag)
                        cut = abs(ni) > sx / p1 || abs(nj) > sy / p1 || abs(nj) > sz / p1;
                    if (cut) {
                        F->comp_dealiased[c].re = 0;
                        F->comp_dealiased[c].im = 0;
                    } else {
                        F->comp_dealiased[c] = F->comp[c];
                    }
                }
    }
    /* toReal + normalize (field.cpp:247-298, normalize_k field_kernels.cu:115-128) */
    FN(fftnd)(F->comp, F->real, sx, sy, sz, +1);
    if (F->needsaliasing) FN(fftnd)(F->comp_dealiased, F->real_dealiased, sx, sy, sz, +1);
    REAL norm = (REAL)1 / (REAL)((float)n);
    for (int c = 0; c < n; c++) {
        F->real[c].re *= norm;
        F->real[c].im = 0;
        if (F->needsaliasing) {
            F->real_dealiased[c].re *= norm;
            F->real_dealiased[c].im = 0;
        }
    }
    /* real-space callback on real, then on real_dealiased (field.cpp:70-86) */
    if (F->hook_real) {
        F->hook_real(F->hook_user, F->real, sx, sy, sz);
        if (F->needsaliasing) F->hook_real(F->hook_user, F->real_dealiased, sx, sy, sz);
    }
    /* toComp (field.cpp:261-273); the forward transform of real_dealiased
     * (:266) is never read and is skipped. */
    FN(fftnd)(F->real, F->comp, sx, sy, sz, -1);
}

/* evolver::advanceTime, src/evolver.cpp:199-226 (output omitted). */
void FN(step)(FN(sys) *s)
{
    if (orc_plant_flag == ORC_PLANT_PHASES) {
        for (int i = 0; i < s->nfields; i++)
            if (s->f[i].dynamic) FN(update_terms)(s, &s->f[i]);
        for (int i = 0; i < s->nfields; i++)
            if (s->f[i].dynamic) FN(set_rhs)(s, &s->f[i]);
        for (int i = 0; i < s->nfields; i++)
            if (!s->f[i].dynamic) FN(update_terms)(s, &s->f[i]);
        for (int i = 0; i < s->nfields; i++)
            if (!s->f[i].dynamic) FN(set_rhs)(s, &s->f[i]);
        s->step += 1;
        return;
    }
    for (int i = 0; i < s->nfields; i++)
        if (!s->f[i].dynamic) FN(update_terms)(s, &s->f[i]);
    for (int i = 0; i < s->nfields; i++)
        if (!s->f[i].dynamic) FN(set_rhs)(s, &s->f[i]);
    for (int i = 0; i < s->nfields; i++)
        if (s->f[i].dynamic) FN(update_terms)(s, &s->f[i]);
    for (int i = 0; i < s->nfields; i++)
        if (s->f[i].dynamic) FN(set_rhs)(s, &s->f[i]);
    s->step += 1;
}

void FN(steps)(FN(sys) *s, int n)
{
    for (int i = 0; i < n; i++) FN(step)(s);
}

void FN(get_real)(FN(sys) *s, int f, REAL *out)
{
    for (int c = 0; c < s->n; c++) out[c] = s->f[f].real[c].re;
}

void FN(get_comp)(FN(sys) *s, int f, REAL *out)
{
    memcpy(out, s->f[f].comp, sizeof(CPLX) * s->n);
}

void FN(get_dealiased)(FN(sys) *s, int f, REAL *out)
{
    for (int c = 0; c < s->n; c++) out[c] = s->f[f].real_dealiased[c].re;
}

int FN(aliasing_order)(FN(sys) *s, int f)
{
    return s->f[f].needsaliasing ? s->f[f].aliasing_order : 0;
}
