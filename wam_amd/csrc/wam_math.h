/*
 * wam_math.h -- deterministic single-precision elementary functions.
 *
 * The WAM source terms call the Fortran runtime's EXP, LOG, LOG10, TANH,
 * COSH, SINH, COS, SIN, ATAN2 and ** (wam_source_module.f90, e.g. SINPUT
 * :2279, TAU_PHI_HF :3946-3958, WAVNU2 :4420-4448, SINPUT_ST6 :4846-4866;
 * FETCH_LAW's ** in wam_jonswap_module.f90 feeds the MIN_ENERGY floor).  An
 * ulp of difference in one of them is carried forward by the implicit
 * scheme, and an ulp in FMEAN or USTAR can move the integer cut-off
 * MIJ = NINT(LOG10(.)/LOG10(CO))+1 (FRCUTINDEX :3639).
 *
 * Each float function here is the correctly rounded float of a double
 * evaluation whose relative error is bounded well below 2^-44 (exp, log,
 * log10, sin, cos, tanh, sinh, cosh, atan2: below 2^-49; pow: below
 * |y log x| * 2^-51 + 2^-51), so it returns the correctly rounded result
 * except where the exact value lies within that bound of a rounding
 * boundary between two floats.  A correctly rounded result is what any
 * accurate libm returns for most arguments, so the reference's values
 * (glibc's float functions behind flang) are met far more often than by a
 * float-only evaluation.
 *
 * Only IEEE-754 double +, -, *, /, fused multiply-add (one rounding), exact
 * conversions, the final double -> float rounding and exact bit
 * manipulation are used, in a fixed order, so the same source gives
 * bit-identical results on the host (gcc; fma is the correctly rounded
 * fused operation) and on gfx950 (hipcc; v_fma_f64, IEEE division), as
 * long as no multiply and add are contracted (-ffp-contract=off).
 *
 * Methods (standard; constants derived in exact rational arithmetic by
 * gen_math_consts.py): Cody-Waite reduction x = k ln2 + r (exp), x = k pi/2
 * + r (sin, cos) with a split constant whose leading part times k is
 * exact; Taylor polynomials on the reduced interval (exp, expm1, sin, cos);
 * log m = 2 atanh((m-1)/(m+1)) on m in [sqrt(1/2), sqrt(2)); atan by the
 * addition formula about j/8 with a table of atan(j/8); tanh, sinh and cosh
 * from expm1 and exp without cancellation.
 */
#ifndef WAM_MATH_H
#define WAM_MATH_H

#include <stdint.h>

/* split and rounded constants (gen_math_consts.py) */
#define WAM_PIO2_1 0x1.921fb54440000p+0      /* pi/2, leading 36 bits */
#define WAM_PIO2_2 0x1.68c234c4c0000p-39     /* next 36 bits */
#define WAM_PIO2_3 0x1.98a2e03707345p-77     /* rest, rounded */
#define WAM_2_OVER_PI 0x1.45f306dc9c883p-1
#define WAM_PI_D 0x1.921fb54442d18p+1
#define WAM_PIO2_D 0x1.921fb54442d18p+0
#define WAM_LN2_HI 0x1.62e42fefa3800p-1      /* ln 2, leading 44 bits */
#define WAM_LN2_LO 0x1.ef35793c76730p-45     /* rest, rounded */
#define WAM_INV_LN2 0x1.71547652b82fep+0
#define WAM_INV_LN10 0x1.bcb7b1526e50ep-2

#if defined(__HIPCC__)
#define WAM_HD __host__ __device__ __attribute__((always_inline)) inline
#else
#define WAM_HD static inline
#endif

WAM_HD int32_t wam_f2i(float f) { int32_t i; __builtin_memcpy(&i, &f, 4); return i; }
WAM_HD float wam_i2f(int32_t i) { float f; __builtin_memcpy(&f, &i, 4); return f; }

WAM_HD float wam_fabsf(float x) { return wam_i2f(wam_f2i(x) & 0x7fffffff); }
WAM_HD float wam_copysignf(float a, float b)
{
    return wam_i2f((wam_f2i(a) & 0x7fffffff) | (wam_f2i(b) & (int32_t)0x80000000));
}
WAM_HD int wam_isnan(float x) { return (wam_f2i(x) & 0x7fffffff) > 0x7f800000; }
WAM_HD float wam_maxf(float a, float b) { return a > b ? a : b; }
WAM_HD float wam_minf(float a, float b) { return a < b ? a : b; }

/* x / 1.1f, correctly rounded, for every float x: the double product of x
   and RN(1/1.1f) (1.1f = 9227469/2^23), rounded once to float.  The exact
   quotient's fraction of an ulp is k/9227469, so it lies at least
   2^-24.1 ulp (2^-48 relative) from a rounding boundary, while the double
   product is within 2^-52 relative of it.  Compared with IEEE division over
   all 2^32 floats by tests/math/div_sweep.c. */
WAM_HD float wam_div_1p1(float x) { return (float)((double)x * 0x1.d1745c6e043b8p-1); }

/* floor/ceil/round on floats without libm; exact. */
WAM_HD float wam_truncf(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_truncf(x);                  /* v_trunc_f32: exact, the same value */
#else
    int32_t i = wam_f2i(x);
    int e = ((i >> 23) & 0xff) - 127;
    if (e >= 23) return x;                       /* integral, inf or nan */
    if (e < 0) return wam_i2f(i & (int32_t)0x80000000);   /* |x| < 1 -> +-0 */
    int32_t m = (int32_t)0x007fffff >> e;
    return wam_i2f(i & ~m);
#endif
}
WAM_HD float wam_floorf(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_floorf(x);                  /* v_floor_f32: exact, the same value */
#else
    float t = wam_truncf(x);
    return (t > x) ? t - 1.0f : t;
#endif
}
WAM_HD float wam_ceilf(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_ceilf(x);                   /* v_ceil_f32: exact, the same value */
#else
    float t = wam_truncf(x);
    return (t < x) ? t + 1.0f : t;
#endif
}
/* an integral float to int, saturating (NaN -> 0), the same on host and
   device (a plain conversion out of range is undefined) */
WAM_HD int wam_f2int(float t)
{
    if (!(t == t)) return 0;
    if (t >= 2147483648.0f) return 2147483647;
    if (t < -2147483648.0f) return (int)0x80000000;
    return (int)t;
}
/* Fortran NINT: nearest integer, halves away from zero. */
WAM_HD int wam_nint(float x)
{
    float a = wam_fabsf(x);
    float t = wam_truncf(a);
    if (a - t >= 0.5f) t = t + 1.0f;     /* a - t is exact for |a| < 2**23 */
    return wam_f2int(x < 0.0f ? -t : t);
}
/* Fortran INT: truncation toward zero. */
WAM_HD int wam_int(float x) { return wam_f2int(wam_truncf(x)); }
/* Fortran SIGN(a,b). */
WAM_HD float wam_signf(float a, float b) { return b >= 0.0f ? wam_fabsf(a) : -wam_fabsf(a); }

/* x**n for integer n >= 0 by square-and-multiply from the low bit. */
WAM_HD float wam_ipowf(float x, int n)
{
    float r = 1.0f, b = x;
    int first = 1;
    while (n > 0) {
        if (n & 1) { r = first ? b : r * b; first = 0; }
        n >>= 1;
        if (n) b = b * b;
    }
    return r;
}


/* ---- double-precision cores ------------------------------------------- */

WAM_HD int64_t wam_d2i(double d) { int64_t i; __builtin_memcpy(&i, &d, 8); return i; }
WAM_HD double wam_i2d(int64_t i) { double d; __builtin_memcpy(&d, &i, 8); return d; }
WAM_HD double wam_fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
/* 2**k as a double, -1022 <= k <= 1023 */
WAM_HD double wam_pow2d(int k) { return wam_i2d((int64_t)(k + 1023) << 52); }
/* nearest integer, ties to even, exactly, for |x| < 2**51 */
WAM_HD double wam_rintd(double x) { return (x + 0x1.8p52) - 0x1.8p52; }

/* exp(x) for -740 < x < 709: x = k ln2 + r with |r| <= ln2/2 (k * LN2_HI is
   exact and so, by Sterbenz's lemma, is x - k * LN2_HI); exp(r) by its
   Taylor polynomial of degree 13 (truncation below 2^-57).  Relative error
   below 2^-51. */
WAM_HD double wam_exp_d(double x)
{
    const double kd = wam_rintd(x * WAM_INV_LN2);
    const double r = (x - kd * WAM_LN2_HI) - kd * WAM_LN2_LO;
    double p = 0x1.6124613a86d09p-33;                 /* 1/13! */
    p = wam_fmad(p, r, 0x1.1eed8eff8d898p-29);        /* 1/12! */
    p = wam_fmad(p, r, 0x1.ae64567f544e4p-26);        /* 1/11! */
    p = wam_fmad(p, r, 0x1.27e4fb7789f5cp-22);        /* 1/10! */
    p = wam_fmad(p, r, 0x1.71de3a556c734p-19);        /* 1/9! */
    p = wam_fmad(p, r, 0x1.a01a01a01a01ap-16);        /* 1/8! */
    p = wam_fmad(p, r, 0x1.a01a01a01a01ap-13);        /* 1/7! */
    p = wam_fmad(p, r, 0x1.6c16c16c16c17p-10);        /* 1/6! */
    p = wam_fmad(p, r, 0x1.1111111111111p-7);         /* 1/5! */
    p = wam_fmad(p, r, 0x1.5555555555555p-5);         /* 1/4! */
    p = wam_fmad(p, r, 0x1.5555555555555p-3);         /* 1/3! */
    p = wam_fmad(p, r, 0.5);
    p = wam_fmad(p, r, 1.0);
    p = wam_fmad(p, r, 1.0);
    const int k = (int)kd;
    if (k < -1000) return (p * wam_pow2d(k + 200)) * wam_pow2d(-200);
    return p * wam_pow2d(k);
}

/* exp(x) - 1 for 0 <= x < 709: below ln2/2 the Taylor polynomial of degree
   14 (no cancellation), above exp(x) - 1 (cancellation at most a factor
   3.5).  Relative error below 2^-49. */
WAM_HD double wam_expm1_d(double x)
{
    if (x > 0.34657359027997264) return wam_exp_d(x) - 1.0;
    double p = 0x1.93974a8c07c9dp-37;                 /* 1/14! */
    p = wam_fmad(p, x, 0x1.6124613a86d09p-33);
    p = wam_fmad(p, x, 0x1.1eed8eff8d898p-29);
    p = wam_fmad(p, x, 0x1.ae64567f544e4p-26);
    p = wam_fmad(p, x, 0x1.27e4fb7789f5cp-22);
    p = wam_fmad(p, x, 0x1.71de3a556c734p-19);
    p = wam_fmad(p, x, 0x1.a01a01a01a01ap-16);
    p = wam_fmad(p, x, 0x1.a01a01a01a01ap-13);
    p = wam_fmad(p, x, 0x1.6c16c16c16c17p-10);
    p = wam_fmad(p, x, 0x1.1111111111111p-7);
    p = wam_fmad(p, x, 0x1.5555555555555p-5);
    p = wam_fmad(p, x, 0x1.5555555555555p-3);
    p = wam_fmad(p, x, 0.5);
    p = wam_fmad(p, x, 1.0);
    return p * x;
}

/* log(x) for a positive finite float x: x = 2^e m with m in [sqrt(1/2),
   sqrt(2)), log m = 2 atanh(s), s = (m-1)/(m+1) (m-1 and m+1 exact, one
   correctly rounded division), |s| < 0.1716, series to s^19 (truncation
   below 2^-55); log x = e LN2_HI + (e LN2_LO + log m).  Relative error
   below 2^-51. */
WAM_HD double wam_log_d(float x)
{
    int32_t i = wam_f2i(x);
    int e = 0;
    if (i < 0x00800000) {                 /* subnormal: scale up by 2**25 */
        i = wam_f2i(x * 33554432.0f);
        e = -25;
    }
    e += ((i >> 23) & 0xff) - 127;
    int32_t mi = (i & 0x007fffff) | 0x3f800000;        /* m in [1, 2) */
    if (mi > 0x3fb504f3) { mi -= 0x00800000; e += 1; } /* m > sqrt(2): m/2 */
    const double m = (double)wam_i2f(mi);
    const double s = (m - 1.0) / (m + 1.0);
    const double z = s * s;
    double p = 0x1.af286bca1af28p-5;                  /* 1/19 */
    p = wam_fmad(p, z, 0x1.e1e1e1e1e1e1ep-5);         /* 1/17 */
    p = wam_fmad(p, z, 0x1.1111111111111p-4);         /* 1/15 */
    p = wam_fmad(p, z, 0x1.3b13b13b13b14p-4);         /* 1/13 */
    p = wam_fmad(p, z, 0x1.745d1745d1746p-4);         /* 1/11 */
    p = wam_fmad(p, z, 0x1.c71c71c71c71cp-4);         /* 1/9 */
    p = wam_fmad(p, z, 0x1.2492492492492p-3);         /* 1/7 */
    p = wam_fmad(p, z, 0x1.999999999999ap-3);         /* 1/5 */
    p = wam_fmad(p, z, 0x1.5555555555555p-2);         /* 1/3 */
    const double s2 = s + s;
    const double lm = wam_fmad(s2 * z, p, s2);
    const double ed = (double)e;
    return ed * WAM_LN2_HI + (ed * WAM_LN2_LO + lm);
}

/* sin(r) and cos(r) for |r| <= pi/4 (+ rounding): Taylor polynomials to r^15
   and r^16 (truncation below 2^-54). */
WAM_HD double wam_sin_poly(double r)
{
    const double z = r * r;
    double p = -0x1.ae7f3e733b81fp-41;                /* -1/15! */
    p = wam_fmad(p, z, 0x1.6124613a86d09p-33);        /* 1/13! */
    p = wam_fmad(p, z, -0x1.ae64567f544e4p-26);       /* -1/11! */
    p = wam_fmad(p, z, 0x1.71de3a556c734p-19);        /* 1/9! */
    p = wam_fmad(p, z, -0x1.a01a01a01a01ap-13);       /* -1/7! */
    p = wam_fmad(p, z, 0x1.1111111111111p-7);         /* 1/5! */
    p = wam_fmad(p, z, -0x1.5555555555555p-3);        /* -1/3! */
    return wam_fmad(r * z, p, r);
}
WAM_HD double wam_cos_poly(double r)
{
    const double z = r * r;
    double p = 0x1.ae7f3e733b81fp-45;                 /* 1/16! */
    p = wam_fmad(p, z, -0x1.93974a8c07c9dp-37);       /* -1/14! */
    p = wam_fmad(p, z, 0x1.1eed8eff8d898p-29);        /* 1/12! */
    p = wam_fmad(p, z, -0x1.27e4fb7789f5cp-22);       /* -1/10! */
    p = wam_fmad(p, z, 0x1.a01a01a01a01ap-16);        /* 1/8! */
    p = wam_fmad(p, z, -0x1.6c16c16c16c17p-10);       /* -1/6! */
    p = wam_fmad(p, z, 0x1.5555555555555p-5);         /* 1/4! */
    p = wam_fmad(p, z, -0.5);
    return wam_fmad(z, p, 1.0);
}

/* x = q pi/2 + r for a float x, |x| < 2^16: r with relative error below
   2^-51 (the nearest multiple of pi/2 to a float of that range is no
   closer than about 2^-26, and x - q PIO2_1 is exact); *q gets q mod 4. */
WAM_HD double wam_reduce_pio2(float x, int *q)
{
    const double xd = (double)x;
    const double qd = wam_rintd(xd * WAM_2_OVER_PI);
    *q = (int)qd & 3;
    return ((xd - qd * WAM_PIO2_1) - qd * WAM_PIO2_2) - qd * WAM_PIO2_3;
}

/* atan(t) for 0 <= t <= 1: atan(t) = atan(c) + atan((t - c)/(1 + t c)) with
   c = j/8 the nearest eighth (t - c exact by Sterbenz's lemma for j >= 1),
   |u| <= 1/16, series to u^13 (truncation below 2^-60). */
WAM_HD double wam_atan01_d(double t)
{
    const int j = (int)wam_rintd(t * 8.0);
    const double c = (double)j * 0.125;
    const double u = (t - c) / wam_fmad(t, c, 1.0);
    const double z = u * u;
    double p = 0x1.3b13b13b13b14p-4;                  /* 1/13 */
    p = wam_fmad(p, z, -0x1.745d1745d1746p-4);        /* -1/11 */
    p = wam_fmad(p, z, 0x1.c71c71c71c71cp-4);         /* 1/9 */
    p = wam_fmad(p, z, -0x1.2492492492492p-3);        /* -1/7 */
    p = wam_fmad(p, z, 0x1.999999999999ap-3);         /* 1/5 */
    p = wam_fmad(p, z, -0x1.5555555555555p-2);        /* -1/3 */
    const double a = wam_fmad(u * z, p, u);
    /* atan(j/8), j = 0..8 */
    const double A = j == 0 ? 0.0 : j == 1 ? 0x1.fd5ba9aac2f6ep-4 : j == 2 ? 0x1.f5b75f92c80ddp-3
                   : j == 3 ? 0x1.6f61941e4def1p-2 : j == 4 ? 0x1.dac670561bb4fp-2 : j == 5 ? 0x1.1e00babdefeb4p-1
                   : j == 6 ? 0x1.4978fa3269ee1p-1 : j == 7 ? 0x1.700a7c5784634p-1 : 0x1.921fb54442d18p-1;
    return A + a;
}

/* ---- the float functions ---------------------------------------------- */

WAM_HD float wam_expf(float x)
{
    if (wam_isnan(x)) return x;
    if (x > 89.0f) return wam_i2f(0x7f800000);
    if (x < -104.0f) return 0.0f;
    return (float)wam_exp_d((double)x);
}

WAM_HD float wam_logf(float x)
{
    if (wam_isnan(x)) return x;
    if (x < 0.0f) return wam_i2f(0x7fc00000);
    if (x == 0.0f) return wam_i2f((int32_t)0xff800000);
    if (wam_f2i(x) == 0x7f800000) return x;
    return (float)wam_log_d(x);
}

WAM_HD float wam_log10f(float x)
{
    if (wam_isnan(x)) return x;
    if (x < 0.0f) return wam_i2f(0x7fc00000);
    if (x == 0.0f) return wam_i2f((int32_t)0xff800000);
    if (wam_f2i(x) == 0x7f800000) return x;
    return (float)(wam_log_d(x) * WAM_INV_LN10);
}

/* tanh x = t / (t + 2), t = expm1(2|x|); |x| > 10 rounds to +-1. */
WAM_HD float wam_tanhf(float x)
{
    if (wam_isnan(x) || x == 0.0f) return x;
    const float a = wam_fabsf(x);
    if (a > 10.0f) return wam_copysignf(1.0f, x);
    const double t = wam_expm1_d(2.0 * (double)a);
    return wam_copysignf((float)(t / (t + 2.0)), x);
}

/* cosh x = (e + 1/e)/2, e = exp|x|. */
WAM_HD float wam_coshf(float x)
{
    if (wam_isnan(x)) return x;
    const float a = wam_fabsf(x);
    if (a > 90.0f) return wam_i2f(0x7f800000);
    const double e = wam_exp_d((double)a);
    return (float)(0.5 * (e + 1.0 / e));
}

/* sinh x = (t + t/(t + 1))/2, t = expm1|x|. */
WAM_HD float wam_sinhf(float x)
{
    if (wam_isnan(x) || x == 0.0f) return x;
    const float a = wam_fabsf(x);
    if (a > 90.0f) return wam_copysignf(wam_i2f(0x7f800000), x);
    const double t = wam_expm1_d((double)a);
    return wam_copysignf((float)(0.5 * (t + t / (t + 1.0))), x);
}

/* sin and cos for |x| < 2^16 (the angles WAM evaluates are a few radians);
   beyond, and for infinities, NaN. */
WAM_HD float wam_sinf(float x)
{
    if (wam_isnan(x) || x == 0.0f) return x;
    if (!(wam_fabsf(x) < 65536.0f)) return wam_i2f(0x7fc00000);
    int q;
    const double r = wam_reduce_pio2(x, &q);
    double v = (q & 1) ? wam_cos_poly(r) : wam_sin_poly(r);
    if (q & 2) v = -v;
    return (float)v;
}
WAM_HD float wam_cosf(float x)
{
    if (wam_isnan(x)) return x;
    if (!(wam_fabsf(x) < 65536.0f)) return wam_i2f(0x7fc00000);
    int q;
    const double r = wam_reduce_pio2(x, &q);
    double v = (q & 1) ? wam_sin_poly(r) : wam_cos_poly(r);
    if (q == 1 || q == 2) v = -v;
    return (float)v;
}

/* x**y for a real exponent (x >= 0; a negative base gives NaN), as
   exp(y log x) in double. */
WAM_HD float wam_powf(float x, float y)
{
    if (wam_isnan(x) || wam_isnan(y)) return wam_i2f(0x7fc00000);
    if (y == 0.0f || x == 1.0f) return 1.0f;
    if (x < 0.0f) return wam_i2f(0x7fc00000);
    if (x == 0.0f) return y > 0.0f ? 0.0f : wam_i2f(0x7f800000);
    if (wam_f2i(x) == 0x7f800000) return y > 0.0f ? x : 0.0f;
    const double t = (double)y * wam_log_d(x);
    if (t > 90.0) return wam_i2f(0x7f800000);
    if (t < -105.0) return 0.0f;
    return (float)wam_exp_d(t);
}

/* atan2(y, x) in [-pi, pi] with the IEEE signed-zero conventions: the ratio
   of the smaller to the larger magnitude (correctly rounded in double),
   atan of it, then the octant. */
WAM_HD float wam_atan2f(float y, float x)
{
    if (wam_isnan(x) || wam_isnan(y)) return wam_i2f(0x7fc00000);
    const double ay = (double)wam_fabsf(y), ax = (double)wam_fabsf(x);
    const int xneg = wam_f2i(x) < 0;
    double a;
    if (ay == 0.0) a = xneg ? WAM_PI_D : 0.0;
    else if (ax == 0.0) a = WAM_PIO2_D;
    else if (ay > ax) {
        a = WAM_PIO2_D - wam_atan01_d(ax / ay);
        if (xneg) a = WAM_PI_D - a;
    } else {
        a = wam_atan01_d(ay / ax);
        if (xneg) a = WAM_PI_D - a;
    }
    return wam_copysignf((float)a, y);
}

WAM_HD float wam_atanf(float x) { return wam_atan2f(x, 1.0f); }

#endif
