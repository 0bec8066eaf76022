/*
 * wam_math.h -- deterministic single-precision elementary functions.
 *
 * The WAM source terms call the Fortran runtime's EXP, LOG, LOG10, TANH,
 * COSH, SINH, COS, SIN and ** (wam_source_module.f90, e.g. SINPUT :2279,
 * TAU_PHI_HF :3946-3958, WAVNU2 :4420-4448, SINPUT_ST6 :4846-4866).  Those
 * results differ by an ulp or two between vendor libraries, and an ulp in
 * FMEAN or USTAR can move the integer cut-off MIJ = NINT(LOG10(.)/LOG10(CO))+1
 * (FRCUTINDEX :3639).  Every routine here uses only IEEE-754 single-precision
 * +, -, *, / and exact bit manipulation, evaluated in a fixed order, so the
 * same source gives bit-identical results on the host (gcc, SSE, no FMA
 * contraction) and on gfx950 (hipcc with -ffp-contract=off).  Truncation,
 * floor and ceiling are exact operations: the device takes them from the
 * hardware instructions, the host from bit manipulation.
 *
 * Algorithms: Cody-Waite range reduction followed by the minimax polynomials
 * published with the Cephes library (S. L. Moshier, "Methods and Programs
 * for Mathematical Functions", 1989): expf, logf, sinf/cosf, tanhf, sinhf,
 * atanf (and atan2f from it).
 * Accuracy (checked against double-precision libm in the CPU test-suite):
 * <= 2 ulp over the argument ranges WAM uses.
 */
#ifndef WAM_MATH_H
#define WAM_MATH_H

#include <stdint.h>

#if defined(__HIPCC__)
#define WAM_HD __host__ __device__ __attribute__((always_inline)) inline
#else
#define WAM_HD static inline
#endif

WAM_HD int32_t wam_f2i(float f) { int32_t i; __builtin_memcpy(&i, &f, 4); return i; }
WAM_HD float wam_i2f(int32_t i) { float f; __builtin_memcpy(&f, &i, 4); return f; }

/* 2**k as a float, exact for -126 <= k <= 127. */
WAM_HD float wam_exp2i(int k) { return wam_i2f((int32_t)(k + 127) << 23); }

/* x * 2**k with a single final IEEE rounding (only in the subnormal range). */
WAM_HD float wam_ldexpf(float x, int k)
{
    if (k > 127) { x = x * wam_exp2i(127); k -= 127; if (k > 127) k = 127; }
    if (k < -126) { x = x * wam_exp2i(-126); k += 126; if (k < -126) k = -126; }
    return x * wam_exp2i(k);
}

WAM_HD float wam_fabsf(float x) { return wam_i2f(wam_f2i(x) & 0x7fffffff); }
WAM_HD float wam_copysignf(float a, float b)
{
    return wam_i2f((wam_f2i(a) & 0x7fffffff) | (wam_f2i(b) & (int32_t)0x80000000));
}
WAM_HD int wam_isnan(float x) { return (wam_f2i(x) & 0x7fffffff) > 0x7f800000; }
WAM_HD float wam_maxf(float a, float b) { return a > b ? a : b; }
WAM_HD float wam_minf(float a, float b) { return a < b ? a : b; }

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
/* Fortran NINT: nearest integer, halves away from zero. */
WAM_HD int wam_nint(float x)
{
    float a = wam_fabsf(x);
    float t = wam_truncf(a);
    if (a - t >= 0.5f) t = t + 1.0f;     /* a - t is exact for |a| < 2**23 */
    int r = (int)t;
    return x < 0.0f ? -r : r;
}
/* Fortran INT: truncation toward zero. */
WAM_HD int wam_int(float x) { return (int)wam_truncf(x); }
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

WAM_HD float wam_expf(float x)
{
    if (wam_isnan(x)) return x;
    if (x > 88.72283935546875f) return wam_i2f(0x7f800000);
    if (x < -103.972084045410f) return 0.0f;
    float fk = wam_floorf(x * 1.44269504088896341f + 0.5f);
    int k = (int)fk;
    float r = x - fk * 0.693359375f;
    r = r - fk * -2.12194440e-4f;
    float z = r * r;
    float y = 1.9875691500E-4f;
    y = y * r + 1.3981999507E-3f;
    y = y * r + 8.3334519073E-3f;
    y = y * r + 4.1665795894E-2f;
    y = y * r + 1.6666665459E-1f;
    y = y * r + 5.0000001201E-1f;
    y = y * z + r;
    y = y + 1.0f;
    return wam_ldexpf(y, k);
}

WAM_HD float wam_logf(float x)
{
    if (wam_isnan(x)) return x;
    if (x < 0.0f) return wam_i2f(0x7fc00000);
    if (x == 0.0f) return wam_i2f((int32_t)0xff800000);
    if (wam_f2i(x) == 0x7f800000) return x;
    int32_t i = wam_f2i(x);
    int e = 0;
    if (i < 0x00800000) {                 /* subnormal: scale up by 2**25 */
        x = x * 33554432.0f;
        i = wam_f2i(x);
        e = -25;
    }
    /* frexp: x = m * 2**e with m in [0.5, 1) */
    e += ((i >> 23) & 0xff) - 126;
    float m = wam_i2f((i & 0x007fffff) | 0x3f000000);
    if (m < 0.707106781186547524f) {
        e -= 1;
        m = m + m - 1.0f;
    } else {
        m = m - 1.0f;
    }
    float z = m * m;
    float y = 7.0376836292E-2f;
    y = y * m - 1.1514610310E-1f;
    y = y * m + 1.1676998740E-1f;
    y = y * m - 1.2420140846E-1f;
    y = y * m + 1.4249322787E-1f;
    y = y * m - 1.6668057665E-1f;
    y = y * m + 2.0000714765E-1f;
    y = y * m - 2.4999993993E-1f;
    y = y * m + 3.3333331174E-1f;
    y = y * m * z;
    float fe = (float)e;
    y = y + fe * -2.12194440e-4f;
    y = y - 0.5f * z;
    float r = m + y;
    r = r + fe * 0.693359375f;
    return r;
}

WAM_HD float wam_log10f(float x)
{
    /* log10(x) = log(x) * log10(e), with log10(e) split hi+lo */
    float l = wam_logf(x);
    return l * 4.3359375E-1f + l * 7.00731903251827651129E-4f;
}

WAM_HD float wam_tanhf(float x)
{
    float a = wam_fabsf(x);
    if (wam_isnan(x)) return x;
    if (a > 9.01f) return x > 0.0f ? 1.0f : -1.0f;
    if (a >= 0.625f) {
        float s = wam_expf(a + a);
        float y = 1.0f - 2.0f / (s + 1.0f);
        return x < 0.0f ? -y : y;
    }
    if (x == 0.0f) return x;
    float z = x * x;
    float y = -5.70498872745E-3f;
    y = y * z + 2.06390887954E-2f;
    y = y * z - 5.37397155531E-2f;
    y = y * z + 1.33314422036E-1f;
    y = y * z - 3.33332819422E-1f;
    y = y * z * x + x;
    return y;
}

WAM_HD float wam_coshf(float x)
{
    float a = wam_fabsf(x);
    if (a > 89.41598629223294f) return wam_i2f(0x7f800000);
    float e = wam_expf(a);
    return 0.5f * e + 0.5f / e;
}

WAM_HD float wam_sinhf(float x)
{
    float a = wam_fabsf(x);
    if (a > 89.41598629223294f) return x > 0.0f ? wam_i2f(0x7f800000) : wam_i2f((int32_t)0xff800000);
    float y;
    if (a > 1.0f) {
        float e = wam_expf(a);
        y = 0.5f * e - 0.5f / e;
    } else {
        float z = a * a;
        y = 2.03721912945E-4f;
        y = y * z + 8.33028376239E-3f;
        y = y * z + 1.66667160211E-1f;
        y = y * z * a + a;
    }
    return x < 0.0f ? -y : y;
}

/* sin and cos: reduction by pi/4 with a three-part constant; fine for |x| < 8192. */
WAM_HD float wam_sincos_core(float x, int want_cos)
{
    int sign = 0;
    float a = x;
    if (a < 0.0f) { a = -a; if (!want_cos) sign = 1; }
    int j = (int)(a * 1.27323954473516f);   /* 4/pi */
    float y = (float)j;
    if (j & 1) { j += 1; y += 1.0f; }
    j &= 7;
    if (j > 3) { j -= 4; sign = !sign; }
    if (want_cos && j > 1) sign = !sign;
    float r = ((a - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
    float z = r * r;
    int use_cos_poly = want_cos ? (j == 0 || j == 3) : (j == 1 || j == 2);
    float p;
    if (use_cos_poly) {
        p = 2.443315711809948E-5f;
        p = p * z - 1.388731625493765E-3f;
        p = p * z + 4.166664568298827E-2f;
        p = p * z * z;
        p = p - 0.5f * z;
        p = p + 1.0f;
    } else {
        p = -1.9515295891E-4f;
        p = p * z + 8.3321608736E-3f;
        p = p * z - 1.6666654611E-1f;
        p = p * z * r;
        p = p + r;
    }
    return sign ? -p : p;
}
WAM_HD float wam_sinf(float x) { return wam_sincos_core(x, 0); }
WAM_HD float wam_cosf(float x) { return wam_sincos_core(x, 1); }

/* x**y for x >= 0 (real exponent), as exp(y*log(x)). */
WAM_HD float wam_powf(float x, float y)
{
    if (x == 0.0f) return y > 0.0f ? 0.0f : (y == 0.0f ? 1.0f : wam_i2f(0x7f800000));
    return wam_expf(y * wam_logf(x));
}

/* atan(x): reduction to |x| <= tan(pi/8) by atan(x) = pi/2 - atan(1/x) and
   pi/4 + atan((x-1)/(x+1)), then Cephes' odd minimax polynomial. */
WAM_HD float wam_atanf(float x)
{
    const float ax = wam_fabsf(x);
    float y = 0.0f, r = ax;
    if (ax > 2.414213562373095f) {           /* tan(3 pi/8) */
        y = 1.5707963267948966f;
        r = -(1.0f / ax);
    } else if (ax > 0.4142135623730950f) {  /* tan(pi/8) */
        y = 0.7853981633974483f;
        r = (ax - 1.0f) / (ax + 1.0f);
    }
    const float z = r * r;
    y = y + ((((8.05374449538e-2f * z - 1.38776856032e-1f) * z + 1.99777106478e-1f) * z - 3.33329491539e-1f) * z *
                 r +
             r);
    return x < 0.0f ? -y : y;
}

/* atan2(y, x) in (-pi, pi], by quadrant from wam_atanf(y/x). */
WAM_HD float wam_atan2f(float y, float x)
{
    const float pi = 3.14159265358979323846f;
    if (x == 0.0f) {
        if (y > 0.0f) return 1.5707963267948966f;
        if (y < 0.0f) return -1.5707963267948966f;
        return 0.0f;
    }
    if (y == 0.0f) return x > 0.0f ? 0.0f : pi;
    const float w = wam_atanf(y / x);
    if (x > 0.0f) return w;
    return y < 0.0f ? w - pi : w + pi;
}

#endif
