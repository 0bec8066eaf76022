/*
 * math_sweep.c -- counts, over every float of an argument range (or every
 * STRIDE-th), how often wam_math.h's single-precision functions differ
 *   (a) from the correctly rounded result, taken as the x87 long-double libm
 *       value rounded to float (its error, a few units of 2^-64, can only
 *       matter within 2^-62 of a rounding boundary), and
 *   (b) from the container's glibc float functions (expf, logf, ...), which
 *       is what the reference's flang build calls.
 * Usage: math_sweep STRIDE   (1 = exhaustive).  Prints one JSON object.
 * Test infrastructure: built and run by tests/test_math_cr.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "wam_math.h"

typedef float (*ff)(float);
typedef long double (*fl)(long double);

static float w_exp(float x) { return wam_expf(x); }
static float w_log(float x) { return wam_logf(x); }
static float w_log10(float x) { return wam_log10f(x); }
static float w_tanh(float x) { return wam_tanhf(x); }
static float w_cosh(float x) { return wam_coshf(x); }
static float w_sinh(float x) { return wam_sinhf(x); }
static float w_sin(float x) { return wam_sinf(x); }
static float w_cos(float x) { return wam_cosf(x); }

static int same(float a, float b)
{
    if (a != a && b != b) return 1;
    uint32_t ia, ib;
    __builtin_memcpy(&ia, &a, 4);
    __builtin_memcpy(&ib, &b, 4);
    return ia == ib;
}

/* one signed range [-hi, -lo] U [lo, hi] of float magnitudes (bit patterns) */
static void sweep1(const char *name, ff w, ff g, fl l, uint32_t lo_bits, uint32_t hi_bits, int both_signs,
                   long stride, int first)
{
    long n = 0, bad_cr = 0, bad_glibc = 0, glibc_cr = 0;
    float worst_x = 0.f;
#pragma omp parallel for reduction(+ : n, bad_cr, bad_glibc, glibc_cr) schedule(dynamic, 4096)
    for (long b = (long)lo_bits; b <= (long)hi_bits; b += stride) {
        for (int s = 0; s <= both_signs; s++) {
            uint32_t bits = (uint32_t)b | (s ? 0x80000000u : 0u);
            float x;
            __builtin_memcpy(&x, &bits, 4);
            float ours = w(x), gl = g(x), cr = (float)l((long double)x);
            n++;
            if (!same(ours, cr)) {
                bad_cr++;
                worst_x = x;
            }
            if (!same(ours, gl)) bad_glibc++;
            if (!same(gl, cr)) glibc_cr++;
        }
    }
    printf("%s\"%s\": {\"n\": %ld, \"ne_correctly_rounded\": %ld, \"ne_glibc\": %ld, \"glibc_ne_correctly_rounded\": %ld, "
           "\"example\": %.9g}",
           first ? "" : ", ", name, n, bad_cr, bad_glibc, glibc_cr, (double)worst_x);
    fflush(stdout);
}

static uint32_t fbits(float f)
{
    uint32_t i;
    __builtin_memcpy(&i, &f, 4);
    return i;
}

/* two-argument functions on a seeded sample */
static void sweep2(const char *name, int fn, long count, int first)
{
    long n = 0, bad_cr = 0, bad_glibc = 0, glibc_cr = 0;
#pragma omp parallel for reduction(+ : n, bad_cr, bad_glibc, glibc_cr) schedule(static)
    for (long i = 0; i < count; i++) {
        uint64_t s = 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
        s ^= s >> 29; s *= 0xBF58476D1CE4E5B9ull; s ^= s >> 32;
        double u1 = (double)(s & 0xffffffu) / 16777216.0, u2 = (double)((s >> 24) & 0xffffffu) / 16777216.0;
        int sg = (int)((s >> 48) & 3);
        float x, y, ours, gl, cr;
        if (fn == 0) {        /* pow: x in [1e-4, 1e6], y in [-4, 4] */
            x = (float)pow(10.0, -4.0 + 10.0 * u1);
            y = (float)(-4.0 + 8.0 * u2);
            ours = wam_powf(x, y); gl = powf(x, y); cr = (float)powl((long double)x, (long double)y);
        } else {              /* atan2: magnitudes 1e-6 .. 1e6, all four signs */
            y = (float)pow(10.0, -6.0 + 12.0 * u1) * ((sg & 1) ? -1.f : 1.f);
            x = (float)pow(10.0, -6.0 + 12.0 * u2) * ((sg & 2) ? -1.f : 1.f);
            ours = wam_atan2f(y, x); gl = atan2f(y, x); cr = (float)atan2l((long double)y, (long double)x);
        }
        n++;
        if (!same(ours, cr)) bad_cr++;
        if (!same(ours, gl)) bad_glibc++;
        if (!same(gl, cr)) glibc_cr++;
    }
    printf("%s\"%s\": {\"n\": %ld, \"ne_correctly_rounded\": %ld, \"ne_glibc\": %ld, \"glibc_ne_correctly_rounded\": %ld}",
           first ? "" : ", ", name, n, bad_cr, bad_glibc, glibc_cr);
    fflush(stdout);
}

int main(int argc, char **argv)
{
    const long stride = argc > 1 ? atol(argv[1]) : 1;
    const long n2 = argc > 2 ? atol(argv[2]) : 4000000;
    printf("{\"stride\": %ld, ", stride);
    /* the argument ranges of the reference's uses, whole float ranges where finite */
    sweep1("exp", w_exp, expf, expl, 0, fbits(104.f), 1, stride, 1);
    sweep1("log", w_log, logf, logl, 1, 0x7f7fffffu, 0, stride, 0);
    sweep1("log10", w_log10, log10f, log10l, 1, 0x7f7fffffu, 0, stride, 0);
    sweep1("tanh", w_tanh, tanhf, tanhl, 0, fbits(10.f), 1, stride, 0);
    sweep1("cosh", w_cosh, coshf, coshl, 0, fbits(90.f), 1, stride, 0);
    sweep1("sinh", w_sinh, sinhf, sinhl, 0, fbits(90.f), 1, stride, 0);
    sweep1("sin", w_sin, sinf, sinl, 0, fbits(65536.f) - 1, 1, stride, 0);
    sweep1("cos", w_cos, cosf, cosl, 0, fbits(65536.f) - 1, 1, stride, 0);
    sweep2("pow", 0, n2, 0);
    sweep2("atan2", 1, n2, 0);
    printf("}\n");
    return 0;
}
