/*
 * oracle.c -- CPU restatement of cuPSS's per-step update (see oracle.h).
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg, never by the product.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* inc/cupss/defines.h:54 -- the reference's PI is a float literal. */
#define ORC_PI 3.1415926535f

#define ORC_MAXF 8
#define ORC_MAXP 16
#define ORC_MAXT 16
#define ORC_MAXFIELDS 32
#define ORC_MAXTERMS 128

/* ------------------------------------------------------------- Philox */

/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As
 * Easy as 1, 2, 3", SC'11: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
 * increments 0x9E3779B9 / 0xBB67AE85, ten rounds. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* uniform in (0,1), exactly representable in float: (2*(x>>9)+1) * 2^-24 */
static inline double orc_u01(uint32_t x)
{
    return (double)((float)(2u * (x >> 9) + 1u) * 5.9604644775390625e-08f);
}

float orc_normal(uint64_t seed, uint64_t step, uint32_t stream, uint32_t cell)
{
    uint32_t ctr[4] = {cell >> 2, (uint32_t)step, (uint32_t)(step >> 32), stream};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    orc_philox4x32_10(ctr, key, o);
    int pair = (cell & 3) >> 1;
    double u0 = orc_u01(o[2 * pair]);
    double u1 = orc_u01(o[2 * pair + 1]);
    double r = sqrt(-2.0 * log(u0));
    double th = 2.0 * M_PI * u1;
    return (float)((cell & 1) ? r * sin(th) : r * cos(th));
}

void orc_normal_fill(uint64_t seed, uint64_t step, uint32_t stream, float *out, int n)
{
    for (int c = 0; c < n; c++) out[c] = orc_normal(seed, step, stream, (uint32_t)c);
}

/* --------------------------------------------------------------- DFT */

void orc_dft_f64(const double *in, double *out, int sx, int sy, int sz, int dir)
{
    int n = sx * sy * sz;
    for (int k2 = 0; k2 < sz; k2++)
        for (int k1 = 0; k1 < sy; k1++)
            for (int k0 = 0; k0 < sx; k0++) {
                double sr = 0, si = 0;
                for (int m2 = 0; m2 < sz; m2++)
                    for (int m1 = 0; m1 < sy; m1++)
                        for (int m0 = 0; m0 < sx; m0++) {
                            double ph = (double)((long long)k0 * m0 % sx) / sx +
                                        (double)((long long)k1 * m1 % sy) / sy +
                                        (double)((long long)k2 * m2 % sz) / sz;
                            double ang = dir * 2.0 * M_PI * ph;
                            int m = m2 * sx * sy + m1 * sx + m0;
                            double c = cos(ang), s = sin(ang);
                            sr += in[2 * m] * c - in[2 * m + 1] * s;
                            si += in[2 * m] * s + in[2 * m + 1] * c;
                        }
                int kk = k2 * sx * sy + k1 * sx + k0;
                out[2 * kk] = sr;
                out[2 * kk + 1] = si;
            }
    (void)n;
}

static int orc_ref_cpu_dealias_bug_flag_f32;
static int orc_ref_cpu_dealias_bug_flag_f64;

void orc_set_ref_cpu_dealias_bug(int on)
{
    orc_ref_cpu_dealias_bug_flag_f32 = on;
    orc_ref_cpu_dealias_bug_flag_f64 = on;
}

/* Planted mistakes, for the tests that show a check can fail (never set
 * outside those tests):
 *   ORC_PLANT_PHASES    advanceTime runs the dynamic phases (C, D) before
 *                       the constraint phases (A, B)
 *   ORC_PLANT_UNDEALIAS products read a factor's `real` array instead of
 *                       its dealiased copy */
static int orc_plant_flag;

void orc_set_plant(int code) { orc_plant_flag = code; }

/* ------------------------------------------------ float / double bodies */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)

#define REAL float
#define FN(x) CAT(orc_f32_, x)
#define CPLX cplx_f32
#define TERM term_f32
#define FIELD field_f32
#define orc_ref_cpu_dealias_bug_flag orc_ref_cpu_dealias_bug_flag_f32
#include "oracle_impl.h"
#undef REAL
#undef FN
#undef CPLX
#undef TERM
#undef FIELD
#undef orc_ref_cpu_dealias_bug_flag

#define REAL double
#define FN(x) CAT(orc_f64_, x)
#define CPLX cplx_f64
#define TERM term_f64
#define FIELD field_f64
#define orc_ref_cpu_dealias_bug_flag orc_ref_cpu_dealias_bug_flag_f64
#include "oracle_impl.h"
#undef REAL
#undef FN
#undef CPLX
#undef TERM
#undef FIELD
#undef orc_ref_cpu_dealias_bug_flag
