/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 *
 * orc_ad.cpp: one model step of the CPU restatement, PROPAGS then IMPLSCH
 * (wamodel.f90:332,354), on the tape of orc_ad.h (FP64, or with ORC_AD32
 * the REAL*4 values of the FP32 build), with the wind
 * given as components: U10 = SQRT(U**2 + V**2), UDIR = ATAN2(U, V) (+ 2 pi
 * when negative), wam_wind_module.f90:1157-1167.  It evaluates the step, its
 * tangent-linear model and its adjoint with respect to
 *     x = (FL3, TAUW, Z0, U, V)  ->  y = (FL3', TAUW', Z0')
 * at the given state, exactly (to FP64 round-off) for the linearisation the
 * restatement's branches define there.  The context is the FP64 build's
 * (FP32 with ORC_AD32: orc_api_create of the build linked in the same
 * library).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <cstring>
#include <vector>
#define ORC_AD
#include "orc_common.h"

ad_tape ad_t;

namespace orc_adns {
#include "orc_implsch.c"
#include "orc_propags.c"
}

namespace {
// The step recorded on the tape: returns the tape nodes of x and y.
void record(const void *ctx, const double *F, const double *tauw, const double *z0, const double *u,
            const double *v, const double *roairn, const double *wstar, const double *depth,
            std::vector<int32_t> &xin, std::vector<int32_t> &yout, std::vector<double> &yval)
{
    const orc_ctx *c = static_cast<const orc_ctx *>(ctx);
    const int ns = c->NSEA;
    const size_t nb = (size_t)c->KL * c->ML * ns;
    ad_t.clear();
    std::vector<adr> f(nb), tw(ns), zz(ns), uu(ns), vv(ns), u10(ns), udir(ns), ro(ns), ws(ns), dp(ns), us(ns),
        ph(ns), flux(6 * (size_t)ns);
    std::vector<int> mij(ns);
    xin.clear();
    for (size_t i = 0; i < nb; i++) { f[i] = ad_input(F[i]); xin.push_back(f[i].i); }
    for (int i = 0; i < ns; i++) { tw[i] = ad_input(tauw[i]); xin.push_back(tw[i].i); }
    for (int i = 0; i < ns; i++) { zz[i] = ad_input(z0[i]); xin.push_back(zz[i].i); }
    for (int i = 0; i < ns; i++) { uu[i] = ad_input(u[i]); xin.push_back(uu[i].i); }
    for (int i = 0; i < ns; i++) { vv[i] = ad_input(v[i]); xin.push_back(vv[i].i); }
    const treal zpi = c->ZPI;
    for (int i = 0; i < ns; i++) {
        u10[i] = ad_sqrt(uu[i] * uu[i] + vv[i] * vv[i]);
        udir[i] = u10[i].v != 0.0 ? ad_atan2(uu[i], vv[i]) : adr(0.0);
        if (udir[i].v < 0.0) udir[i] = udir[i] + zpi;
        ro[i] = roairn[i]; ws[i] = wstar[i]; dp[i] = depth[i];
    }
    orc_adns::orc_propags(c, f.data());
    orc_adns::orc_implsch(c, f.data(), u10.data(), udir.data(), tw.data(), us.data(), zz.data(), ro.data(), ws.data(),
                          dp.data(), c->INDEP, 0, ph.data(), flux.data(), mij.data(), 1);
    yout.clear();
    yval.clear();
    for (size_t i = 0; i < nb; i++) { yout.push_back(f[i].i); yval.push_back(f[i].v); }
    for (int i = 0; i < ns; i++) { yout.push_back(tw[i].i); yval.push_back(tw[i].v); }
    for (int i = 0; i < ns; i++) { yout.push_back(zz[i].i); yval.push_back(zz[i].v); }
}
}  // namespace

/* mode 0: y = step(x); 1: also dy = L dx (dx packed as x); 2: also
   xbar = L^T ybar (ybar packed as y).  y has NB + 2 NSEA values, x NB + 4
   NSEA.  Returns the tape length. */
extern "C" long orc_ad_step(const void *ctx, int mode, const double *F, const double *tauw, const double *z0,
                            const double *u, const double *v, const double *roairn, const double *wstar,
                            const double *depth, const double *in, double *y, double *out)
{
    std::vector<int32_t> xin, yout;
    std::vector<double> yval;
    record(ctx, F, tauw, z0, u, v, roairn, wstar, depth, xin, yout, yval);
    std::memcpy(y, yval.data(), yval.size() * sizeof(double));
    const size_t n = ad_t.size();
    if (mode == 1) {
        std::vector<double> t(n, 0.0);
        for (size_t k = 0; k < xin.size(); k++) t[xin[k]] = in[k];
        for (size_t j = 1; j < n; j++) {
            if (ad_t.a[j] == 0) continue;      // inputs (no operands)
            double s = ad_t.da[j] * t[ad_t.a[j]];
            if (ad_t.b[j]) s += ad_t.db[j] * t[ad_t.b[j]];
            t[j] = s;
        }
        for (size_t k = 0; k < yout.size(); k++) out[k] = yout[k] ? t[yout[k]] : 0.0;
    } else if (mode == 2) {
        std::vector<double> a(n, 0.0);
        for (size_t k = 0; k < yout.size(); k++)
            if (yout[k]) a[yout[k]] += in[k];
        for (size_t j = n - 1; j >= 1; j--) {
            const double w = a[j];
            if (w == 0.0 || ad_t.a[j] == 0) continue;
            a[ad_t.a[j]] += ad_t.da[j] * w;
            if (ad_t.b[j]) a[ad_t.b[j]] += ad_t.db[j] * w;
        }
        for (size_t k = 0; k < xin.size(); k++) out[k] = a[xin[k]];
    }
    return (long)n;
}
