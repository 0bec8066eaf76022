/*
 * ORACLE -- test and measurement infrastructure only; nothing in the
 * product uses it.
 *
 * orc_opcount.h: a REAL*4 number type that counts the floating-point
 * operations the CPU restatement performs, so the algorithmic work of
 * IMPLSCH per sea point can be stated.  Same size and layout as float, so
 * the set-up tables built by the float restatement are read as they are.
 * Counting rules: add/sub, mul, div and sqrt count one each; comparisons,
 * MAX/MIN, ABS, SIGN and the integer rounding intrinsics count none; each
 * elementary-function call (EXP, LOG, LOG10, TANH, COSH, SINH, COS, SIN,
 * ACOS, real **) counts once as a transcendental; x**n with integer n counts
 * its multiplications.  Single-threaded use only.
 */
#ifndef ORC_OPCOUNT_H
#define ORC_OPCOUNT_H
#ifndef __cplusplus
#error "orc_opcount.h is C++"
#endif
#include <stdint.h>
#include "wam_math.h"

struct oc_counts { uint64_t add, mul, div, sqrt, trans; };
extern oc_counts oc_cnt;

/* Per-stage totals: ORC_STAGE(n) in the restatement closes the running stage
   (its operations since the last marker go to oc_stage_cnt[stage]) and opens
   stage n. */
constexpr int OC_NSTAGE = 3;
extern oc_counts oc_stage_cnt[OC_NSTAGE], oc_mark;
extern int oc_cur_stage;
inline void oc_stage(int n)
{
    oc_counts &s = oc_stage_cnt[oc_cur_stage];
    s.add += oc_cnt.add - oc_mark.add; s.mul += oc_cnt.mul - oc_mark.mul; s.div += oc_cnt.div - oc_mark.div;
    s.sqrt += oc_cnt.sqrt - oc_mark.sqrt; s.trans += oc_cnt.trans - oc_mark.trans;
    oc_mark = oc_cnt;
    oc_cur_stage = n;
}

struct oc_real {
    float v;
    oc_real() = default;
    constexpr oc_real(float x) : v(x) {}
    constexpr oc_real(double x) : v((float)x) {}
    constexpr oc_real(int x) : v((float)x) {}
    explicit operator float() const { return v; }
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
};
static_assert(sizeof(oc_real) == sizeof(float), "oc_real must be laid out as float");

inline oc_real operator+(oc_real a, oc_real b) { oc_cnt.add++; return oc_real(a.v + b.v); }
inline oc_real operator-(oc_real a, oc_real b) { oc_cnt.add++; return oc_real(a.v - b.v); }
inline oc_real operator*(oc_real a, oc_real b) { oc_cnt.mul++; return oc_real(a.v * b.v); }
inline oc_real operator/(oc_real a, oc_real b) { oc_cnt.div++; return oc_real(a.v / b.v); }
inline oc_real operator-(oc_real a) { return oc_real(-a.v); }
inline oc_real &operator+=(oc_real &a, oc_real b) { a = a + b; return a; }
inline oc_real &operator-=(oc_real &a, oc_real b) { a = a - b; return a; }
inline oc_real &operator*=(oc_real &a, oc_real b) { a = a * b; return a; }
inline oc_real &operator/=(oc_real &a, oc_real b) { a = a / b; return a; }
inline bool operator<(oc_real a, oc_real b) { return a.v < b.v; }
inline bool operator>(oc_real a, oc_real b) { return a.v > b.v; }
inline bool operator<=(oc_real a, oc_real b) { return a.v <= b.v; }
inline bool operator>=(oc_real a, oc_real b) { return a.v >= b.v; }
inline bool operator==(oc_real a, oc_real b) { return a.v == b.v; }
inline bool operator!=(oc_real a, oc_real b) { return a.v != b.v; }

inline oc_real oc_t(float r) { oc_cnt.trans++; return oc_real(r); }
inline oc_real oc_sqrt(oc_real x) { oc_cnt.sqrt++; return oc_real(sqrtf(x.v)); }
inline oc_real oc_ipow(oc_real x, int n)
{
    oc_real r = 1.0f, b = x;
    int first = 1;
    while (n > 0) {
        if (n & 1) { r = first ? b : r * b; first = 0; }
        n >>= 1;
        if (n) b = b * b;
    }
    return r;
}
#endif
