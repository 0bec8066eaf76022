// wam_dual.hpp -- forward derivatives of IMPLSCH's per-point scalar chains
// (AIRSEA's Newton iteration, WSIGSTAR, TAU_PHI_HF's nodes) along their two
// inputs, for the adjoint.  The reference has no adjoint; these are the
// chains of wam_source_module.f90 differentiated as evaluated here.
#pragma once
#include "wam_dev.hpp"
#include "wam_math.h"

namespace wam {

// ---------------------------------------------------------------------------
// dual numbers: a value and its derivatives along two input directions.  The
// value part performs the forward kernels' operations in their order, so it
// is bit-identical to them and takes the same branches.
// ---------------------------------------------------------------------------
struct D2 {
    float v, a, b;
};
__device__ __forceinline__ D2 dc(float v) { return {v, 0.f, 0.f}; }
__device__ __forceinline__ D2 operator+(D2 x, D2 y) { return {x.v + y.v, x.a + y.a, x.b + y.b}; }
__device__ __forceinline__ D2 operator-(D2 x, D2 y) { return {x.v - y.v, x.a - y.a, x.b - y.b}; }
__device__ __forceinline__ D2 operator*(D2 x, D2 y)
{
    return {x.v * y.v, x.a * y.v + x.v * y.a, x.b * y.v + x.v * y.b};
}
__device__ __forceinline__ D2 operator/(D2 x, D2 y)
{
    const float q = x.v / y.v;
    return {q, (x.a - q * y.a) / y.v, (x.b - q * y.b) / y.v};
}
__device__ __forceinline__ D2 operator+(D2 x, float c) { return {x.v + c, x.a, x.b}; }
__device__ __forceinline__ D2 operator+(float c, D2 x) { return {c + x.v, x.a, x.b}; }
__device__ __forceinline__ D2 operator-(D2 x, float c) { return {x.v - c, x.a, x.b}; }
__device__ __forceinline__ D2 operator-(float c, D2 x) { return {c - x.v, -x.a, -x.b}; }
__device__ __forceinline__ D2 operator*(D2 x, float c) { return {x.v * c, x.a * c, x.b * c}; }
__device__ __forceinline__ D2 operator*(float c, D2 x) { return {c * x.v, c * x.a, c * x.b}; }
__device__ __forceinline__ D2 operator/(D2 x, float c) { return {x.v / c, x.a / c, x.b / c}; }
__device__ __forceinline__ D2 operator/(float c, D2 y)
{
    const float q = c / y.v;
    return {q, -(q * y.a) / y.v, -(q * y.b) / y.v};
}
__device__ __forceinline__ D2 dsqrt(D2 x)
{
    const float r = sqrtf(x.v), g = r > 0.f ? 0.5f / r : 0.f;
    return {r, g * x.a, g * x.b};
}
__device__ __forceinline__ D2 dlog(D2 x) { return {wam_logf(x.v), x.a / x.v, x.b / x.v}; }
__device__ __forceinline__ D2 dexp(D2 x)
{
    const float e = wam_expf(x.v);
    return {e, e * x.a, e * x.b};
}
__device__ __forceinline__ D2 dpow(D2 x, float y)
{
    const float p = wam_powf(x.v, y), g = x.v != 0.f ? (y * p) / x.v : 0.f;
    return {p, g * x.a, g * x.b};
}
__device__ __forceinline__ D2 dipow(D2 x, int n)
{
    const float g = (float)n * wam_ipowf(x.v, n - 1);
    return {wam_ipowf(x.v, n), g * x.a, g * x.b};
}
// MAX and MIN as the forward kernels take them (ties select the second)
__device__ __forceinline__ D2 dmax(D2 x, D2 y) { return x.v > y.v ? x : y; }
__device__ __forceinline__ D2 dmin(D2 x, D2 y) { return x.v < y.v ? x : y; }

// AIRSEA (wam_source_module.f90:868-971) with derivatives along (UTOP, TAUW);
// the same operations as airsea().  ja = dUST/dUTOP, dUST/dTAUW, dZ0/dUTOP,
// dZ0/dTAUW.
__device__ __forceinline__ void airsea_d(const DevTables &t, float utop_v, float tauw_v, float &ust_v, float &z0_v,
                                         float *ja)
{
    const D2 utop = {utop_v, 1.f, 0.f}, tauw = {tauw_v, 0.f, 1.f};
    const float xlogxl = wam_logf(t.XNLEV), alphaog = t.ALPHA / t.G;
    const D2 xkutop = t.XKAPPA * utop;
    const D2 ustold = utop * dsqrt(8.0E-4f + 8.0E-5f * utop);
    D2 tauold = dmax(ustold * ustold, tauw + t.EPS1);
    D2 u = dsqrt(tauold);
    D2 ustm1 = 1.0f / dmax(u, dc(1.0E-6f));
    D2 z0ch = dc(0.f);
    for (int it = 1; it <= 15; it++) {
        const D2 x = tauw / tauold;
        z0ch = (alphaog * tauold) / dsqrt(dmax(1.0f - x, dc(t.EPS1)));
        const D2 z0vis = t.RNUAIRM * ustm1;
        const D2 z0tot = z0ch + z0vis;
        const D2 xologz0 = 1.0f / (xlogxl - dlog(z0tot));
        const D2 f = u - xkutop * xologz0;
        const D2 zz = (ustm1 * ((z0ch * (2.0f - 3.0f * x)) / (1.0f - x) - z0vis)) / z0tot;
        const D2 delf = 1.0f - (xkutop * (xologz0 * xologz0)) * zz;
        u = u - f / delf;
        const D2 taunew = dmax(u * u, tauw + t.EPS1);
        u = dsqrt(taunew);
        if (taunew.v == tauold.v) break;
        ustm1 = 1.0f / dmax(u, dc(1.0E-6f));
        tauold = taunew;
    }
    ust_v = u.v;
    z0_v = z0ch.v;
    ja[0] = u.a; ja[1] = u.b; ja[2] = z0ch.a; ja[3] = z0ch.b;
}

// WSIGSTAR (:4250-4339) with derivatives along (USTAR, Z0).
__device__ __forceinline__ D2 wsigstar_d(const DevTables &t, float ust_v, float z0_v, float wstar)
{
    const D2 ust = {ust_v, 1.f, 0.f}, z0 = {z0_v, 0.f, 1.f};
    const float p1 = 1.48f, p2 = -0.21f, c1 = 1.03E-3f, c2 = 0.04E-3f;
    D2 u10 = (ust * (1.0f / t.XKAPPA)) * (2.30258509299405f - dlog(z0));
    u10 = dmax(u10, dc(1.0f));
    const D2 u10m1 = 1.0f / u10;
    const D2 c2u = c2 * dpow(u10, p1);
    const D2 u10p2 = dpow(u10, p2);
    const D2 c_d = (c1 + c2u) * u10p2;
    const D2 dcddu = (((p2 * c1) + (p1 + p2) * c2u) * u10p2) * u10m1;
    const D2 sigc = 1.0f + ((0.5f * u10) / c_d) * dcddu;
    const D2 base = 0.0f * dipow(ust, 3) + 0.2f * wam_ipowf(wstar, 3);
    return dmin(dc(0.1f), (sigc * u10m1) * dpow(base, 1.0f / 3.0f));
}

// TAU_PHI_HF's node j (1..19) of the stress integral (:3853-3966, XLEVTAIL =
// 0) with derivatives along (USTAR, Z0), as the parallel form in stresso().
__device__ __forceinline__ D2 tau_node_d(const DevTables &t, const float *fr, int mij, float ust_v, float z0_v, int j)
{
    const D2 ust0 = {ust_v, 1.f, 0.f}, z0 = {z0_v, 0.f, 1.f};
    const float gm1 = 1.0f / t.G;
    const float omegac = t.ZPI * fr[mij - 1];
    const D2 tauw0 = ust0 * ust0;
    const D2 xloggz0 = dlog(t.G * z0);
    const D2 omegacc = dmax(dc(omegac), (t.X0TAUHF * t.G) / ust0);
    const D2 sqz = dsqrt(z0 * gm1), sqg = 1.0f / sqz;
    const D2 zinf = dlog(omegacc * sqz);
    const D2 delz = dmax((0.0f - zinf) / 18.f, dc(0.0f));
    const D2 ust = j == 1 ? ust0 : dsqrt(tauw0);
    const D2 y = dexp(zinf + (float)(j - 1) * delz);
    const D2 cm1 = (y * sqg) * gm1;
    const D2 zx = ust * cm1 + t.ZALP;
    const D2 zarg = t.XKAPPA / zx;
    D2 zlog = (xloggz0 + 2.0f * dlog(cm1)) + zarg;
    zlog = dmin(zlog, dc(0.0f));
    const D2 zbeta = dexp(zlog) * dipow(zlog, 4);
    return ((zbeta * tauw0) * t.WTAUHF[j - 1]) * delz;
}

// dTAU1/d(USTAR, Z0) of TAU_PHI_HF's stress integral: its nodes' derivatives
// added in node order
__device__ __forceinline__ void tau_jac(const DevTables &t, const float *fr, int mij, float ust, float z0, float &ja,
                                        float &jb)
{
    float a = 0.f, b = 0.f;
    for (int j = 1; j <= 19; j++) {
        const D2 h = tau_node_d(t, fr, mij, ust, z0, j);
        a += h.a;
        b += h.b;
    }
    ja = a;
    jb = b;
}

}  // namespace wam
