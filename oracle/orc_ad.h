/*
 * ORACLE -- test infrastructure only; nothing in the product uses it.
 *
 * orc_ad.h: an FP64 number type that records every operation on a tape, so
 * that the CPU restatement (orc_implsch.c, orc_propags.c, compiled with
 * -DORC_AD as C++ in orc_ad.cpp) yields the exact tangent-linear model of
 * its computation (a forward sweep over the tape) and its adjoint (a reverse
 * sweep).  Each tape node holds at most two operands with their partial
 * derivatives.  Branches (MAX, MIN, comparisons, NINT, the limiter's clamp,
 * the frequency cut-off) follow the values at the linearisation point, as a
 * tangent-linear model linearised there does.  Node 0 is the constant slot,
 * so zero-filled memory holds constants.  Single-threaded use only.
 */
#ifndef ORC_AD_H
#define ORC_AD_H
#ifndef __cplusplus
#error "orc_ad.h is C++"
#endif
#include <math.h>
#include <stdint.h>
#include <vector>

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
    double v;
    int32_t i;            /* tape node, 0 = constant */
    adr() = default;
    constexpr adr(double x) : v(x), i(0) {}
    constexpr adr(float x) : v(x), i(0) {}
    constexpr adr(int x) : v(x), i(0) {}
    constexpr adr(double x, int32_t n) : v(x), i(n) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator int() const { return (int)v; }
};

/* a new independent variable (an input of the linearised map) */
inline adr ad_input(double x) { return adr(x, ad_t.push(0, 0.0, 0, 0.0)); }

inline adr ad_node(double v, const adr &x, double px)
{
    if (!x.i) return adr(v);
    return adr(v, ad_t.push(x.i, px, 0, 0.0));
}
inline adr ad_node2(double v, const adr &x, double px, const adr &y, double py)
{
    if (!x.i && !y.i) return adr(v);
    if (!y.i) return adr(v, ad_t.push(x.i, px, 0, 0.0));
    if (!x.i) return adr(v, ad_t.push(y.i, py, 0, 0.0));
    return adr(v, ad_t.push(x.i, px, y.i, py));
}

inline adr operator+(const adr &x, const adr &y) { return ad_node2(x.v + y.v, x, 1.0, y, 1.0); }
inline adr operator-(const adr &x, const adr &y) { return ad_node2(x.v - y.v, x, 1.0, y, -1.0); }
inline adr operator*(const adr &x, const adr &y) { return ad_node2(x.v * y.v, x, y.v, y, x.v); }
inline adr operator/(const adr &x, const adr &y)
{
    const double q = x.v / y.v;
    return ad_node2(q, x, 1.0 / y.v, y, -q / y.v);
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
