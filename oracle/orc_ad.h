/*
 * ORACLE -- test infrastructure only; nothing in the product uses it.
 *
 * orc_ad.h: a number type that records every operation on a tape, so that
 * the CPU restatement (orc_implsch.c, orc_propags.c, compiled with -DORC_AD
 * as C++ in orc_ad.cpp) yields the exact tangent-linear model of its
 * computation (a forward sweep over the tape) and its adjoint (a reverse
 * sweep).  Each tape node holds at most two operands with their partial
 * derivatives (in double).  Branches (MAX, MIN, comparisons, NINT, the
 * limiter's clamp, the frequency cut-off) follow the values at the
 * linearisation point, as a tangent-linear model linearised there does.
 * Node 0 is the constant slot, so zero-filled memory holds constants.
 * Single-threaded use only.
 *
 * The values are FP64 by default.  With -DORC_AD32 they are the REAL*4
 * restatement's own values (float, the elementary functions of wam_math.h),
 * bit for bit, so the linearisation point and every branch are those of the
 * FP32 build; the partial derivatives are still formed in double.
 */
#ifndef ORC_AD_H
#define ORC_AD_H
#ifndef __cplusplus
#error "orc_ad.h is C++"
#endif
#include <math.h>
#include <stdint.h>
#include <vector>
#ifdef ORC_AD32
#include "wam_math.h"
typedef float adv;
#else
typedef double adv;
#endif

struct ad_tape {
    std::vector<int32_t> a, b;
    std::vector<double> da, db;
    void clear()
    {
        a.assign(1, 0); b.assign(1, 0); da.assign(1, 0.0); db.assign(1, 0.0);
    }
    int32_t push(int32_t ia, double pa, int32_t ib, double pb)
    {
        a.push_back(ia); b.push_back(ib); da.push_back(pa); db.push_back(pb);
        return (int32_t)a.size() - 1;
    }
    size_t size() const { return a.size(); }
};
extern ad_tape ad_t;

struct adr {
    adv v;
    int32_t i;            /* tape node, 0 = constant */
    adr() = default;
    constexpr adr(double x) : v((adv)x), i(0) {}
    constexpr adr(float x) : v(x), i(0) {}
    constexpr adr(int x) : v((adv)x), i(0) {}
    constexpr adr(adv x, int32_t n) : v(x), i(n) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator int() const { return (int)v; }
};

/* a new independent variable (an input of the linearised map) */
inline adr ad_input(adv x) { return adr(x, ad_t.push(0, 0.0, 0, 0.0)); }

inline adr ad_node(adv v, const adr &x, double px)
{
    if (!x.i) return adr(v, 0);
    return adr(v, ad_t.push(x.i, px, 0, 0.0));
}
inline adr ad_node2(adv v, const adr &x, double px, const adr &y, double py)
{
    if (!x.i && !y.i) return adr(v, 0);
    if (!y.i) return adr(v, ad_t.push(x.i, px, 0, 0.0));
    if (!x.i) return adr(v, ad_t.push(y.i, py, 0, 0.0));
    return adr(v, ad_t.push(x.i, px, y.i, py));
}

inline adr operator+(const adr &x, const adr &y) { return ad_node2(x.v + y.v, x, 1.0, y, 1.0); }
inline adr operator-(const adr &x, const adr &y) { return ad_node2(x.v - y.v, x, 1.0, y, -1.0); }
inline adr operator*(const adr &x, const adr &y) { return ad_node2(x.v * y.v, x, (double)y.v, y, (double)x.v); }
inline adr operator/(const adr &x, const adr &y)
{
    const adv q = x.v / y.v;
    return ad_node2(q, x, 1.0 / (double)y.v, y, -(double)x.v / ((double)y.v * (double)y.v));
}
inline adr operator-(const adr &x) { return ad_node(-x.v, x, -1.0); }
inline adr &operator+=(adr &x, const adr &y) { x = x + y; return x; }
inline adr &operator-=(adr &x, const adr &y) { x = x - y; return x; }
inline adr &operator*=(adr &x, const adr &y) { x = x * y; return x; }
inline adr &operator/=(adr &x, const adr &y) { x = x / y; return x; }
inline bool operator<(const adr &x, const adr &y) { return x.v < y.v; }
inline bool operator>(const adr &x, const adr &y) { return x.v > y.v; }
inline bool operator<=(const adr &x, const adr &y) { return x.v <= y.v; }
inline bool operator>=(const adr &x, const adr &y) { return x.v >= y.v; }
inline bool operator==(const adr &x, const adr &y) { return x.v == y.v; }
inline bool operator!=(const adr &x, const adr &y) { return x.v != y.v; }

#ifdef ORC_AD32
/* values by wam_math.h (the FP32 build's), derivatives in double at them */
inline adr ad_exp(const adr &x) { const float e = wam_expf(x.v); return ad_node(e, x, (double)e); }
inline adr ad_log(const adr &x) { return ad_node(wam_logf(x.v), x, 1.0 / (double)x.v); }
inline adr ad_log10(const adr &x) { return ad_node(wam_log10f(x.v), x, 1.0 / ((double)x.v * log(10.0))); }
inline adr ad_tanh(const adr &x)
{
    const double t = tanh((double)x.v);
    return ad_node(wam_tanhf(x.v), x, 1.0 - t * t);
}
inline adr ad_cosh(const adr &x) { return ad_node(wam_coshf(x.v), x, sinh((double)x.v)); }
inline adr ad_sinh(const adr &x) { return ad_node(wam_sinhf(x.v), x, cosh((double)x.v)); }
inline adr ad_cos(const adr &x) { return ad_node(wam_cosf(x.v), x, -sin((double)x.v)); }
inline adr ad_sin(const adr &x) { return ad_node(wam_sinf(x.v), x, cos((double)x.v)); }
inline adr ad_sqrt(const adr &x)
{
    const float r = sqrtf(x.v);
    return ad_node(r, x, r > 0 ? 0.5 / (double)r : 0.0);
}
inline adr ad_acos(const adr &x)
{
    return ad_node((float)acos((double)x.v), x, -1.0 / sqrt(1.0 - (double)x.v * x.v));
}
inline adr ad_abs(const adr &x) { return x.v < 0 ? -x : ad_node(wam_fabsf(x.v), x, 1.0); }
inline adr ad_pow(const adr &x, const adr &y)
{
    const float p = wam_powf(x.v, y.v);
    return ad_node2(p, x, x.v != 0.0f ? (double)y.v * p / (double)x.v : 0.0, y,
                    x.v > 0.0f ? (double)p * log((double)x.v) : 0.0);
}
inline adr ad_atan2(const adr &y, const adr &x)
{
    const double r2 = (double)x.v * x.v + (double)y.v * y.v;
    return ad_node2(wam_atan2f(y.v, x.v), y, (double)x.v / r2, x, -(double)y.v / r2);
}
/* x**n, integer n >= 0: wam_ipowf's value */
inline adr ad_ipow(const adr &x, int n)
{
    if (n == 0) return adr(1.0f);
    return ad_node(wam_ipowf(x.v, n), x, (double)n * (double)wam_ipowf(x.v, n - 1));
}
#else
inline adr ad_exp(const adr &x) { const double e = exp(x.v); return ad_node(e, x, e); }
inline adr ad_log(const adr &x) { return ad_node(log(x.v), x, 1.0 / x.v); }
inline adr ad_log10(const adr &x) { return ad_node(log10(x.v), x, 1.0 / (x.v * log(10.0))); }
inline adr ad_tanh(const adr &x) { const double t = tanh(x.v); return ad_node(t, x, 1.0 - t * t); }
inline adr ad_cosh(const adr &x) { return ad_node(cosh(x.v), x, sinh(x.v)); }
inline adr ad_sinh(const adr &x) { return ad_node(sinh(x.v), x, cosh(x.v)); }
inline adr ad_cos(const adr &x) { return ad_node(cos(x.v), x, -sin(x.v)); }
inline adr ad_sin(const adr &x) { return ad_node(sin(x.v), x, cos(x.v)); }
inline adr ad_sqrt(const adr &x) { const double r = sqrt(x.v); return ad_node(r, x, r > 0 ? 0.5 / r : 0.0); }
inline adr ad_acos(const adr &x) { return ad_node(acos(x.v), x, -1.0 / sqrt(1.0 - x.v * x.v)); }
inline adr ad_abs(const adr &x) { return x.v < 0 ? -x : x; }
inline adr ad_pow(const adr &x, const adr &y)
{
    const double p = pow(x.v, y.v);
    return ad_node2(p, x, x.v != 0.0 ? y.v * p / x.v : 0.0, y, x.v > 0.0 ? p * log(x.v) : 0.0);
}
inline adr ad_atan2(const adr &y, const adr &x)
{
    const double r2 = x.v * x.v + y.v * y.v;
    return ad_node2(atan2(y.v, x.v), y, x.v / r2, x, -y.v / r2);
}
/* x**n, integer n >= 0, by the same square-and-multiply as the other builds */
inline adr ad_ipow(const adr &x, int n)
{
    adr r = 1.0, b = x;
    int first = 1;
    while (n > 0) {
        if (n & 1) { r = first ? b : r * b; first = 0; }
        n >>= 1;
        if (n) b = b * b;
    }
    return r;
}
#endif
#endif
