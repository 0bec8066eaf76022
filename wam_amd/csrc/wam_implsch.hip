// wam_implsch.hip -- IMPLSCH (wam_source_module.f90:374-700) on gfx950.
//
// One workgroup advances one sea point at a time (persistent loop).  The
// point's spectrum F, the source SL, its diagonal derivative FL and the
// positive input SPOS live in LDS; threads own spectral bins for the
// element-wise stages, one thread per frequency performs each direction sum,
// and thread 0 runs the per-point scalar stages (AIRSEA / W3FLX4, the
// frequency integrals, FRCUTINDEX, TAU_PHI_HF, the LFACTOR stress balance).
// Every reduction is evaluated in the reference's sequential order and no
// multiply-add is contracted, so results are bit-identical to a sequential
// REAL*4 evaluation with the same elementary functions (wam_math.h).
//
// SNONLIN's scatter of each DIA quadruplet into nine bins (:2890-3043) is
// rewritten as a gather: the quadruplet terms are computed in parallel into
// LDS, then every bin sums its contributions from a host-built list that is
// ordered exactly like the reference's sequential scatter.
#include "wam_dev.hpp"
#include "wam_math.h"

namespace wam {

constexpr int NT = 256;          // threads per workgroup
constexpr int MAXF = 64;         // frequencies, directions and 10-Hz bins <= 64

struct PointScalars {
    float u10, udir, tauw, z0, roairn, wstar, depth;
    int indep;
    float ust, emean, fmean, f1mean, akmean, xkmean, emeanws, fmeanws, phiaw, sign;
    int mij, ju;
    // LFACTOR state
    float twx, twy, tauvx, tauvy, tau_tot, rtau, drtau;
    int sign_new, overshot, done;
    float enhfr;
};

// Dynamic LDS layout (floats unless noted): F, SL, FL, SPOS [NB]; SMIN, SLBOT
// [NB] when needed; SCR = max(4*NQ, 4*NB); LLWS [NB] bytes at the end.
struct Lds {
    float *F, *SL, *FL, *SPOS, *SMIN, *SLBOT, *SCR;
    unsigned char *LLWS;
};

__device__ __forceinline__ int tb(const DevTables &t, int m, int indep) { return m * t.NDEPTH + indep - 1; }

// TEMP(M) = SUM over K of F (optionally masked), thread m, natural order.
__device__ void dirsum(const DevTables &t, const float *F, const unsigned char *mask, float *T)
{
    const int KL = t.KL;
    for (int m = threadIdx.x; m < t.ML; m += NT) {
        float s = 0.f;
        bool first = true;
        for (int k = 0; k < KL; k++) {
            const int b = k + KL * m;
            if (mask && !mask[b]) continue;
            s = first ? F[b] : s + F[b];
            first = false;
        }
        T[m] = s;
    }
}

// TOTAL_ENERGY_B and FEMEAN (wam_interface_module.f90:1342-1410, 207-286).
__device__ void energy_fmean(const DevTables &t, const float *T, float &e, float &fm)
{
    const int ML = t.ML;
    float s = t.MO_TAIL * T[ML - 1];
    for (int m = 0; m < ML; m++) s = s + T[m] * t.DFIM[m];
    e = wam_maxf(s, t.EMIN);
    float f = t.MM1_TAIL * T[ML - 1];
    for (int m = 0; m < ML; m++) f = f + T[m] * t.DFIMOFR[m];
    fm = e / wam_maxf(f, t.EMIN);
}

// W3FLX4 (wam_source_module.f90:5372-5443)
__device__ void w3flx4(const DevTables &t, float u10, float &ust, float &z0)
{
    float cd;
    if (u10 >= 50.33f) {
        ust = 2.026f * sqrtf(t.CDFAC);
        cd = wam_ipowf(ust / u10, 2);
    } else {
        cd = (t.CDFAC * ((8.058f + 0.967f * u10) - 0.016f * (u10 * u10))) * 1E-4f;
        ust = u10 * sqrtf(cd);
    }
    z0 = t.XNLEV * wam_expf(-0.4f / sqrtf(cd));
}

// AIRSEA (wam_source_module.f90:868-971)
__device__ void airsea(const DevTables &t, float utop, float tauw, float &ust, float &z0)
{
    const float xlogxl = wam_logf(t.XNLEV), alphaog = t.ALPHA / t.G;
    const float xkutop = t.XKAPPA * utop;
    const float ustold = utop * sqrtf(8.0E-4f + 8.0E-5f * utop);
    float tauold = wam_maxf(ustold * ustold, tauw + t.EPS1);
    float u = sqrtf(tauold);
    float ustm1 = 1.0f / wam_maxf(u, 1.0E-6f);
    float z0ch = 0.f;
    for (int it = 1; it <= 15; it++) {
        const float x = tauw / tauold;
        z0ch = (alphaog * tauold) / sqrtf(wam_maxf(1.0f - x, t.EPS1));
        const float z0vis = t.RNUAIRM * ustm1;
        const float z0tot = z0ch + z0vis;
        const float xologz0 = 1.0f / (xlogxl - wam_logf(z0tot));
        const float f = u - xkutop * xologz0;
        const float zz = (ustm1 * ((z0ch * (2.0f - 3.0f * x)) / (1.0f - x) - z0vis)) / z0tot;
        const float delf = 1.0f - (xkutop * (xologz0 * xologz0)) * zz;
        u = u - f / delf;
        const float taunew = wam_maxf(u * u, tauw + t.EPS1);
        u = sqrtf(taunew);
        if (taunew == tauold) break;
        ustm1 = 1.0f / wam_maxf(u, 1.0E-6f);
        tauold = taunew;
    }
    ust = u;
    z0 = z0ch;
}

// WSIGSTAR (wam_source_module.f90:4250-4339)
__device__ float wsigstar(const DevTables &t, float ust, float z0, float wstar)
{
    const float p1 = 1.48f, p2 = -0.21f, c1 = 1.03E-3f, c2 = 0.04E-3f;
    float u10 = (ust * (1.0f / t.XKAPPA)) * (2.30258509299405f - wam_logf(z0));
    u10 = wam_maxf(u10, 1.0f);
    const float u10m1 = 1.0f / u10;
    const float c2u = c2 * wam_powf(u10, p1);
    const float u10p2 = wam_powf(u10, p2);
    const float c_d = (c1 + c2u) * u10p2;
    const float dcddu = (((p2 * c1) + (p1 + p2) * c2u) * u10p2) * u10m1;
    const float sigc = 1.0f + ((0.5f * u10) / c_d) * dcddu;
    const float base = 0.0f * wam_ipowf(ust, 3) + 0.2f * wam_ipowf(wstar, 3);
    return wam_minf(0.1f, (sigc * u10m1) * wam_powf(base, 1.0f / 3.0f));
}

// FRCUTINDEX (:3569-3647) and FRCUTINDEX_ST6 (:5447-5515)
__device__ int frcut(const DevTables &t, float fm, float fmws, float ust)
{
    if (t.iphys == 2) {
        const int nk = t.ML;
        const float fxfm = 6.f * t.ZPI, fxpm = (4.0f * t.G) / 28.f, signk = t.ZPI * t.FR[nk - 1];
        float fhigh = wam_maxf(fxfm * fm, fxpm / ust);
        const float f1 = 1.f / wam_logf(t.CO);
        const float f2 = 1.f - wam_logf(t.ZPI * t.FR[0]) * f1;
        int nkh = min(nk, wam_int(f2 + f1 * wam_logf(wam_maxf(1.E-7f, fhigh))));
        const int nkh1 = min(nk, nkh + 1);
        fhigh = wam_minf(signk, wam_maxf(fxfm * fm, fxpm / ust));
        return max(2, min(nkh1, wam_int(f2 + f1 * wam_logf(wam_maxf(1.E-7f, fhigh)))));
    }
    const float fpmh = 2.5f / t.FR[0];
    const float fppm = (3.0f * t.G) / ((28.0f * t.ZPI) * t.FR[0]);
    const float fm2 = wam_maxf(fmws, fm) * fpmh;
    const float fpm = fppm / wam_maxf(ust, 1.0E-6f);
    const float fpm4 = wam_maxf(fm2, fpm);
    const int mij = wam_nint(wam_log10f(fpm4) * t.INV_LOG_CO) + 1;
    return min(max(1, mij), t.ML);
}

// TAU_PHI_HF with XLEVTAIL = 0 (:3853-3966)
__device__ void tau_phi_hf(const DevTables &t, int mij, float ust0, float z0, float &tauhf, float &phihf)
{
    const float gm1 = 1.0f / t.G, x0g = t.X0TAUHF * t.G;
    const float omegac = t.ZPI * t.FR[mij - 1];
    const float tauw0 = ust0 * ust0;
    const float xloggz0 = wam_logf(t.G * z0);
    const float omegacc = wam_maxf(omegac, x0g / ust0);
    const float sqz = sqrtf(z0 * gm1), sqg = 1.0f / sqz;
    const float zinf = wam_logf(omegacc * sqz);
    const float delz = wam_maxf((0.0f - zinf) / 18.f, 0.0f);
    float th = 0.f, ph = 0.f, tauw = tauw0, ust = ust0;
    for (int j = 1; j <= 19; j++) {
        const float y = wam_expf(zinf + (float)(j - 1) * delz);
        const float cm1 = (y * sqg) * gm1;
        const float zx = ust * cm1 + t.ZALP;
        const float zarg = t.XKAPPA / zx;
        float zlog = (xloggz0 + 2.0f * wam_logf(cm1)) + zarg;
        zlog = wam_minf(zlog, 0.0f);
        const float zbeta = wam_expf(zlog) * wam_ipowf(zlog, 4);
        const float fnc2 = ((zbeta * tauw) * t.WTAUHF[j - 1]) * delz;
        tauw = wam_maxf(tauw - 0.f * fnc2, 0.0f);
        ust = sqrtf(tauw);
        th = th + fnc2;
        ph = ph + fnc2 / y;
    }
    tauhf = th;
    phihf = sqz * ph;
}

// SINPUT (:2111-2308), NSIN = 2: SL, SPOS, FL, LLWS for every bin.
__device__ void sinput(const DevTables &t, const Lds &s, PointScalars &P, float *cu, float *q)
{
    const int KL = t.KL, ML = t.ML, NB = KL * ML;
    // q[4*m + ..] = CNSN, UCN1, UCN2, XV1D1, XV1D2, ZCN1 (ZCN2 == ZCN1) per frequency
    __shared__ float sig_n;
    if (threadIdx.x == 0) sig_n = wsigstar(t, P.ust, P.z0, P.wstar);
    __syncthreads();
    const float xk2 = t.XKAPPA * t.XKAPPA;
    const float usn1 = P.ust * (1.0f - sig_n), usn2 = P.ust * (1.0f + sig_n);
    for (int m = threadIdx.x; m < ML; m += NT) {
        const float fac = t.ZPI * t.FR[m];
        const float cst = ((fac * t.BETAMAX) / xk2) / t.ROWATER;
        float xk, cm, cnsn;
        if (t.shallow) {
            xk = t.TFAK[tb(t, m, P.indep)];
            cm = xk / fac;
            cnsn = ((cst * wam_ipowf(fac, 2)) / (t.G * xk)) * P.roairn;
        } else {
            xk = wam_ipowf(fac, 2) / t.G;
            cm = fac / t.G;
            cnsn = cst * P.roairn;
        }
        const float zcn = wam_logf(xk * P.z0);
        float *r = q + 6 * m;
        r[0] = cnsn;
        r[1] = usn1 * cm + t.ZALP;
        r[2] = usn2 * cm + t.ZALP;
        r[3] = -1.f / (((usn1 / t.XKAPPA) * zcn) * cm);
        r[4] = -1.f / (((usn2 / t.XKAPPA) * zcn) * cm);
        r[5] = zcn;
    }
    __syncthreads();
    const float const3 = ((2.f * t.XKAPPA) / t.BETAMAX) * xk2;
    for (int b = threadIdx.x; b < NB; b += NT) {
        const int k = b % KL, m = b / KL;
        const float *r = q + 6 * m;
        const float tk = cu[k];
        float ufac1 = 0.f;
        unsigned char ll = 0;
        if (tk > 0.01f) {
#pragma unroll
            for (int is = 0; is < 2; is++) {
                const float x = tk * r[1 + is];
                const float zlog = r[5] + t.XKAPPA / x;
                if (zlog < 0.f) {
                    const float z2x = (zlog * zlog) * x;
                    ufac1 = ufac1 + ((0.5f * wam_expf(zlog)) * z2x) * z2x;
                    ll = 1;
                }
            }
        }
        float ufac2 = (const3 * (tk - r[3])) * wam_ipowf(r[1], 2);
        ufac2 = 0.5f * ufac2 + ((0.5f * const3) * (tk - r[4])) * wam_ipowf(r[2], 2);
        const float sp = r[0] * ufac1;
        const float fl = sp + r[0] * ufac2;
        const float f = s.F[b];
        s.SPOS[b] = sp * f;
        s.FL[b] = fl;
        s.SL[b] = fl * f;
        s.LLWS[b] = ll;
    }
    __syncthreads();
}

// SINPUT_ST6 (:4704-4925) with LFACTOR (:4465-4700) and TAUWINDS (:5555-5587).
// cu[k] holds COS(TH(K)); wn/cg the point's wave numbers and group velocities.
__device__ void sinput_st6(const DevTables &t, const Lds &s, PointScalars &P, float *pf, const float *wn,
                           const float *cg)
{
    const int KL = t.KL, NK = t.ML, NB = KL * NK, NK10 = t.NK10;
    // per-frequency work in pf: SIG, CINV, SQRTBN, LFACT, SD, SX, SY, CINV10, LF10, UCINV  (each MAXF/ 128)
    float *sig = pf, *cinv = pf + MAXF, *sqrtbn = pf + 2 * MAXF, *lfact = pf + 3 * MAXF;
    float *sd = pf + 4 * MAXF, *sx = sd + 128, *sy = sx + 128, *cinv10 = sy + 128, *lf10 = cinv10 + 128,
          *ucinv = lf10 + 128;
    float *A = s.SCR, *D = s.SCR + NB, *W2 = s.SCR + 2 * NB, *SDS = s.SCR + 3 * NB;
    const float ZPI = t.ZPI;
    for (int m = threadIdx.x; m < NK; m += NT) sig[m] = ZPI * t.FR[m];
    __syncthreads();
    for (int b = threadIdx.x; b < NB; b += NT) {
        const int m = b / KL;
        A[b] = (s.F[b] * cg[m]) / (ZPI * sig[m]);
    }
    __syncthreads();
    for (int m = threadIdx.x; m < NK; m += NT) {
        const float *a = A + KL * m;
        float sm = a[0], mx = a[0];
        for (int k = 1; k < KL; k++) { sm = sm + a[k]; mx = wam_maxf(mx, a[k]); }
        const float adens = (sm * sig[m]) * t.DELTH;
        float sn = 0.f;
        for (int k = 0; k < KL; k++) {
            const float v = (mx < 1.0E-34f) ? 1.f : a[k] / mx;
            sn = k == 0 ? v : sn + v;
        }
        const float anar = 1.0f / (sn * t.DELTH);
        sqrtbn[m] = sqrtf((anar * adens) * wam_ipowf(wn[m], 3));
        cinv[m] = wn[m] / sig[m];
    }
    __syncthreads();
    const float cosu = wam_cosf(P.udir), sinu = wam_sinf(P.udir);
    const float uproxy = 32.f * P.ust, rw = P.roairn / t.ROWATER;
    for (int b = threadIdx.x; b < NB; b += NT) {
        const int k = b % KL, m = b / KL;
        const float cosd = t.COSTH[k] * cosu + t.SINTH[k] * sinu;
        const float x = ((uproxy * cinv[m]) * cosd) - 1.f;
        const float w1 = wam_ipowf(wam_maxf(0.f, x), 2);
        W2[b] = wam_ipowf(wam_minf(0.f, x), 2);
        const float d = (((rw * sig[m]) * (2.8f - (1.f + wam_tanhf((10.f * sqrtbn[m]) * w1 - 11.f)))) * sqrtbn[m]) * w1;
        D[b] = d;
        SDS[b] = ((d * A[b]) * sig[m]) / cg[m];
    }
    __syncthreads();
    // LFACTOR: spectral densities of the input (positive parts for the components)
    for (int m = threadIdx.x; m < NK; m += NT) {
        const float *v = SDS + KL * m;
        float a = 0.f, bx = 0.f, by = 0.f;
        for (int k = 0; k < KL; k++) {
            const float vp = wam_maxf(0.f, v[k]);
            if (k == 0) { a = v[k]; bx = vp * t.COSTH[k]; by = vp * t.SINTH[k]; }
            else { a = a + v[k]; bx = bx + vp * t.COSTH[k]; by = by + vp * t.SINTH[k]; }
        }
        sd[m] = a * t.DELTH; sx[m] = bx * t.DELTH; sy[m] = by * t.DELTH;
    }
    __syncthreads();
    const float GRW = t.G * t.ROWATER;
    if (threadIdx.x == 0) {
        for (int i = 0; i < NK; i++) cinv10[i] = cinv[i];
        if (NK < NK10) {
            for (int i = NK; i < NK10; i++) {
                cinv10[i] = t.SIG10[i] * 0.101978f;
                const float r2 = wam_ipowf(t.SIG10[NK - 1] / t.SIG10[i], 2);
                sd[i] = sd[NK - 1] * r2; sx[i] = sx[NK - 1] * r2; sy[i] = sy[NK - 1] * r2;
            }
        }
        const float u10 = P.u10;
        const float tau_tot = (P.ust * P.ust) * P.roairn;
        float tau_vis = (wam_maxf(0.0f, -5.0E-5f * u10 + 1.1E-3f) * (u10 * u10)) * P.roairn;
        tau_vis = wam_minf(0.95f * tau_tot, tau_vis);
        P.tauvx = tau_vis * wam_cosf(P.udir);
        P.tauvy = tau_vis * wam_sinf(P.udir);
        float twx = 0.f, twy = 0.f;
        for (int i = 0; i < NK10; i++) {
            const float vx = (sx[i] * cinv10[i]) * t.DSII10[i], vy = (sy[i] * cinv10[i]) * t.DSII10[i];
            twx = i == 0 ? vx : twx + vx; twy = i == 0 ? vy : twy + vy;
        }
        twx = GRW * twx; twy = GRW * twy;
        const float taux = P.tauvx + twx, tauy = P.tauvy + twy;
        const float tau = sqrtf(taux * taux + tauy * tauy);
        const float err = (tau - tau_tot) / tau_tot;
        P.twx = twx; P.twy = twy; P.tau_tot = tau_tot;
        P.done = !(tau > tau_tot);
        P.overshot = 0;
        P.rtau = err / 90.f;
        P.drtau = 2.0f;
        P.sign_new = err >= 0.f ? 1 : -1;
    }
    for (int i = threadIdx.x; i < NK10; i += NT) {
        lf10[i] = 1.0f;
        ucinv[i] = 1.0f - (32.f * P.ust) * (i < NK ? cinv[i] : t.SIG10[i] * 0.101978f);
    }
    __syncthreads();
    for (int it = 1; it <= 80 && !P.done; it++) {
        const float rtau = P.rtau;
        for (int i = threadIdx.x; i < NK10; i += NT) lf10[i] = wam_minf(1.0f, wam_expf(ucinv[i] * rtau));
        __syncthreads();
        if (threadIdx.x == 0) {
            float twx = 0.f, twy = 0.f;
            for (int i = 0; i < NK10; i++) {
                const float vx = ((sx[i] * lf10[i]) * cinv10[i]) * t.DSII10[i];
                const float vy = ((sy[i] * lf10[i]) * cinv10[i]) * t.DSII10[i];
                twx = i == 0 ? vx : twx + vx; twy = i == 0 ? vy : twy + vy;
            }
            twx = GRW * twx; twy = GRW * twy;
            const float taux = P.tauvx + twx, tauy = P.tauvy + twy;
            const float tau = sqrtf(taux * taux + tauy * tauy);
            const float err = (tau - P.tau_tot) / P.tau_tot;
            const int sold = P.sign_new;
            P.sign_new = err >= 0.f ? 1 : -1;
            if (P.sign_new != sold) P.overshot = 1;
            if (P.overshot) P.drtau = wam_maxf(0.5f * (1.0f + P.drtau), 1.00010f);
            P.rtau = P.sign_new > 0 ? P.rtau * P.drtau : P.rtau * (1.f / P.drtau);
            P.twx = twx; P.twy = twy;
            if (wam_fabsf(err) < 1.54E-4f) P.done = 1;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float ls = lf10[0];
        for (int i = 1; i < NK; i++) ls = ls + lf10[i];
        P.sign = ls < (float)NK ? 1.f : 0.f;     // apply LFACT
        P.tauw = sqrtf(P.twx * P.twx + P.twy * P.twy);
    }
    __syncthreads();
    const bool apply = P.sign != 0.f;
    for (int b = threadIdx.x; b < NB; b += NT) {
        const int m = b / KL;
        float d = D[b];
        if (apply) d = d * lf10[m];
        const float dpos = d;
        const float w2 = W2[b];
        d = d - ((((rw * sig[m]) * 0.09f) * (2.8f - (1.f + wam_tanhf((10.f * sqrtbn[m]) * w2 - 11.f)))) * sqrtbn[m]) * w2;
        const float f = s.F[b];
        s.SL[b] = d * f;
        s.SPOS[b] = dpos * f;
        s.FL[b] = d;
        s.LLWS[b] = s.SPOS[b] > 0.f;
    }
    __syncthreads();
}

// STRESSO (:3164-3356) with SL = SPOS.  Returns through P (tauw only when want_tauw).
__device__ void stresso(const DevTables &t, const Lds &s, PointScalars &P, const float *cu, float *pm,
                        bool want_tauw)
{
    const int KL = t.KL, ML = t.ML, mij = P.mij;
    float *st = pm, *sxa = pm + MAXF, *sya = pm + 2 * MAXF;
    for (int m = threadIdx.x; m < mij; m += NT) {
        const float *v = s.SPOS + KL * m;
        float a = 0.f, x = 0.f, y = 0.f;
        for (int k = 0; k < KL; k++) {
            const float sp = wam_maxf(v[k], 0.f);
            if (k == 0) { a = sp; x = sp * t.SINTH[k]; y = sp * t.COSTH[k]; }
            else { a = a + sp; x = x + sp * t.SINTH[k]; y = y + sp * t.COSTH[k]; }
        }
        st[m] = a; sxa[m] = x; sya[m] = y;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const float gm1 = 1.0f / t.G;
        const float cst = (t.DELTH * t.ZPI4) * gm1;
        const float c1 = (cst * t.FR5[mij - 1]) * gm1;
        const float c2 = (t.ROAIR * cst) * t.FR5[mij - 1];
        float phi = 0.f, xs = 0.f, ys = 0.f;
        for (int m = 0; m < mij; m++) {
            const float cm = t.shallow ? t.TFAK[tb(t, m, P.indep)] * (1.0f / (t.ZPI * t.FR[m]))
                                       : (t.ZPI * t.FR[m]) * gm1;
            float rw = t.RHOWG_DFIM[m];
            if (m == mij - 1 && mij != ML) rw = 0.5f * rw;
            phi = phi + st[m] * rw;
            const float cmrw = cm * rw;
            xs = xs + sxa[m] * cmrw;
            ys = ys + sya[m] * cmrw;
        }
        xs = xs / wam_maxf(t.ROAIR, 1.f);
        ys = ys / wam_maxf(t.ROAIR, 1.f);
        float t1 = 0.f, t2 = 0.f;
        for (int k = 0; k < KL; k++) {
            const float cosw = wam_maxf(cu[k], 0.f);
            const float f = s.F[k + KL * (mij - 1)];
            if (k == 0) { t2 = f * wam_ipowf(cosw, 2); t1 = t2 * cosw; }
            else { t1 = t1 + f * wam_ipowf(cosw, 3); t2 = t2 + f * wam_ipowf(cosw, 2); }
        }
        float tau1, phi1;
        tau_phi_hf(t, mij, P.ust, P.z0, tau1, phi1);
        const float tauhf = (c1 * t1) * tau1, phihf = (c2 * t2) * phi1;
        phi = phi + phihf;
        xs = xs + tauhf * wam_sinf(P.udir);
        ys = ys + tauhf * wam_cosf(P.udir);
        float tw = sqrtf(xs * xs + ys * ys);
        tw = wam_minf(tw, P.ust * P.ust - t.EPS1);
        tw = wam_maxf(tw, 0.f);
        if (want_tauw) P.tauw = tw;
        P.phiaw = phi;
    }
    __syncthreads();
}

// IMPHFTAIL (:3651-3731)
__device__ void imphftail(const DevTables &t, const Lds &s, PointScalars &P, float *t2)
{
    const int KL = t.KL, ML = t.ML, mij = P.mij;
    if (threadIdx.x == 0) {
        for (int m = mij - 1; m < ML; m++) {
            if (t.shallow) {
                const float akm1 = 1.f / t.TFAK[tb(t, m, P.indep)];
                t2[m] = wam_ipowf(akm1, 3) / t.TCGOND[tb(t, m, P.indep)];
            } else t2[m] = t.FRM5[m];
        }
        for (int m = mij; m < ML; m++) t2[m] = t2[m] / t2[mij - 1];
    }
    __syncthreads();
    for (int b = threadIdx.x + KL * mij; b < KL * ML; b += NT) {
        const int k = b % KL, m = b / KL;
        s.F[b] = t2[m] * s.F[k + KL * (mij - 1)];
    }
    __syncthreads();
}

// SNONLIN (:2742-3047): quadruplet terms in parallel, then the ordered gather.
__device__ void snonlin(const DevTables &t, const Lds &s, float ftemp_scale)
{
    const int KL = t.KL, NQ = t.NQ, NB = KL * t.ML;
    float *QAD = s.SCR, *QDD = s.SCR + NQ, *QDP = s.SCR + 2 * NQ, *QDM = s.SCR + 3 * NQ;
    const float *F = s.F;
    for (int q = threadIdx.x; q < NQ; q += NT) {
        const int mc = q / (2 * KL) + 1, r = q % (2 * KL), kh = r / KL, k = r % KL;
        const int *ic = t.INLCOEF + (mc - 1) * 5;
        const float *rc = t.RNLCOEF + (mc - 1) * 25;
        const int k1 = t.K1W[kh * KL + k], k2 = t.K2W[kh * KL + k];
        const int k11 = t.K11W[kh * KL + k], k21 = t.K21W[kh * KL + k];
        const int IC = (ic[0] - 1) * KL, IP = (ic[1] - 1) * KL, IP1 = (ic[2] - 1) * KL;
        const int IM = (ic[3] - 1) * KL, IM1 = (ic[4] - 1) * KL;
        const float sap = rc[1] * F[k1 + IP] + rc[2] * F[k11 + IP] + rc[3] * F[k1 + IP1] + rc[4] * F[k11 + IP1];
        const float sam = rc[13] * F[k2 + IM] + rc[14] * F[k21 + IM] + rc[15] * F[k2 + IM1] + rc[16] * F[k21 + IM1];
        const float fij = F[k + IC] * rc[0];
        const float ftemp = t.AF11[mc - 1] * ftemp_scale;
        float fad1 = fij * (sap + sam);
        const float fad2 = fad1 - (2.f * sap) * sam;
        fad1 = fad1 + fad2;
        const float fcen = ftemp * fij;
        QAD[q] = fad2 * fcen;
        QDD[q] = fad1 * ftemp;
        QDP[q] = ((fij - 2.f * sam) * t.DAL1) * fcen;
        QDM[q] = ((fij - 2.f * sap) * t.DAL2) * fcen;
    }
    __syncthreads();
    // slots in RNLCOEF for the SL / FL weights of update types 1..8
    const int wsl[9] = {0, 20, 19, 17, 18, 8, 7, 5, 6};
    const int wfl[9] = {0, 23, 24, 21, 22, 11, 12, 9, 10};
    for (int b = threadIdx.x; b < NB; b += NT) {
        const int n = t.NCONTRIB[b];
        const int *lst = t.CONTRIB + (size_t)b * t.NCMAX;
        float sl = s.SL[b], fl = s.FL[b];
        for (int j = 0; j < n; j++) {
            const int e = lst[j];
            const int q = e & 0xffff, ty = e >> 16;
            if (ty == 0) {
                sl = sl + (-(2.f * QAD[q]));
                fl = fl + (-(2.f * QDD[q]));
            } else {
                const float *rc = t.RNLCOEF + (q / (2 * KL)) * 25;
                sl = sl + QAD[q] * rc[wsl[ty]];
                fl = fl + (ty <= 4 ? QDM[q] : QDP[q]) * rc[wfl[ty]];
            }
        }
        s.SL[b] = sl;
        s.FL[b] = fl;
    }
    __syncthreads();
}

template <int IPHYS, int SHALLOW>
__global__ void __launch_bounds__(NT) implsch_kernel(DevTables t, ImplArgs a)
{
    extern __shared__ float lds_raw[];
    __shared__ PointScalars P;
    __shared__ float T[MAXF], cu[MAXF], sprd[MAXF], t2[MAXF], pm[3 * MAXF], q[6 * MAXF];
    __shared__ float pf[4 * MAXF + 6 * 128];
    const int KL = t.KL, ML = t.ML, NB = KL * ML, NS = t.NSEA;
    Lds s;
    s.F = lds_raw; s.SL = s.F + NB; s.FL = s.SL + NB; s.SPOS = s.FL + NB;
    float *nxt = s.SPOS + NB;
    s.SMIN = nullptr; s.SLBOT = nullptr;
    if (a.lcflx) { s.SMIN = nxt; nxt += NB; }
    if (SHALLOW) { s.SLBOT = nxt; nxt += NB; }
    s.SCR = nxt;
    const int scr = 4 * (t.NQ > NB ? t.NQ : NB);
    s.LLWS = (unsigned char *)(s.SCR + scr);

    // XCD-aware persistent order: the blocks resident on one XCD walk one
    // contiguous range of sea points side by side, so the bin-major loads of
    // neighbouring points share L2 lines.
    const int nblk = gridDim.x, xcd = blockIdx.x % 8, lb = blockIdx.x / 8;
    const int nb_x = (nblk + 7 - xcd) / 8;
    const int per = (NS + 7) / 8;
    const int p0 = xcd * per, p1 = min(NS, p0 + per);
    const float delt = t.DELT;
    for (int p = p0 + lb; p < p1; p += nb_x) {
        // 0. load the point
        for (int b = threadIdx.x; b < NB; b += NT) s.F[b] = a.fl3[(size_t)b * NS + p];
        if (threadIdx.x == 0) {
            P.u10 = a.u10[p]; P.udir = a.udir[p]; P.tauw = a.tauw[p]; P.z0 = a.z0[p];
            P.roairn = a.roairn[p]; P.wstar = a.wstar[p]; P.depth = t.DEPTH[p];
            P.indep = SHALLOW ? t.INDEP[p] : 1;
            P.ju = min(t.JUMAX, max(wam_nint(P.u10 / t.DELU), 1));
        }
        __syncthreads();
        for (int k = threadIdx.x; k < KL; k += NT) {
            const float c = wam_cosf(t.TH[k] - P.udir);
            cu[k] = c;
            sprd[k] = wam_ipowf(wam_maxf(0.f, c), 2);
        }
        dirsum(t, s.F, nullptr, T);
        __syncthreads();
        // 2. mean parameters, 3. stress
        if (threadIdx.x == 0) {
            float e, fm;
            energy_fmean(t, T, e, fm);
            P.emean = e; P.fmean = fm;
            float tm = t.MP1_TAIL * T[ML - 1];
            for (int m = 0; m < ML; m++) tm = tm + T[m] * t.DFIM_FR[m];
            tm = e > t.EMIN ? e / tm : 1.f;
            P.f1mean = 1.f / tm;
            if (SHALLOW) {
                float d2 = sqrtf(t.G) / t.ZPI;
                float w1 = (t.MM1_TAIL * d2) * T[ML - 1];
                for (int m = 0; m < ML; m++) w1 = w1 + (T[m] / sqrtf(t.TFAK[tb(t, m, P.indep)])) * t.DFIM[m];
                P.akmean = e > t.EMIN ? wam_ipowf(e / w1, 2) : 1.f;
                d2 = t.ZPI / sqrtf(t.G);
                float w2 = (t.MP1_TAIL * d2) * T[ML - 1];
                for (int m = 0; m < ML; m++) w2 = w2 + (T[m] * sqrtf(t.TFAK[tb(t, m, P.indep)])) * t.DFIM[m];
                P.xkmean = e > t.EMIN ? wam_ipowf(w2 / e, 2) : 1.f;
            }
            if (IPHYS == 2) w3flx4(t, P.u10, P.ust, P.z0);
            else airsea(t, P.u10, P.tauw, P.ust, P.z0);
        }
        __syncthreads();
        const float *wn = nullptr, *cg = nullptr;
        float *wnl = pm, *cgl = pm + MAXF;
        if (IPHYS == 2) {
            for (int m = threadIdx.x; m < ML; m += NT) {
                wnl[m] = t.WN[(size_t)m * NS + p];
                cgl[m] = t.CGG[(size_t)m * NS + p];
            }
            __syncthreads();
            // keep wn/cg out of pm (reused by STRESSO): copy to t2/q tail
            for (int m = threadIdx.x; m < ML; m += NT) { q[m] = wnl[m]; q[MAXF + m] = cgl[m]; }
            __syncthreads();
            wn = q; cg = q + MAXF;
            sinput_st6(t, s, P, pf, wn, cg);
        } else {
            sinput(t, s, P, cu, q);
        }
        dirsum(t, s.F, s.LLWS, T);
        __syncthreads();
        if (threadIdx.x == 0) {
            float e, fm;
            energy_fmean(t, T, e, fm);
            P.emeanws = e; P.fmeanws = fm;
            P.mij = frcut(t, P.fmean, P.fmeanws, P.ust);
        }
        __syncthreads();
        stresso(t, s, P, cu, pm, IPHYS != 2);
        if (threadIdx.x == 0) {
            if (IPHYS == 2) w3flx4(t, P.u10, P.ust, P.z0);
            else airsea(t, P.u10, P.tauw, P.ust, P.z0);
        }
        __syncthreads();
        imphftail(t, s, P, t2);
        if (IPHYS != 2) sinput(t, s, P, cu, q);
        if (a.lcflx) {
            for (int b = threadIdx.x; b < NB; b += NT) s.SMIN[b] = s.SL[b] - s.SPOS[b];
        }
        stresso(t, s, P, cu, pm, IPHYS != 2);
        // dissipation
        if (IPHYS == 2) {
            // SDISSIP_ST6 (:4929-5102) and W3SWL6 (:5106-5252); wn, cg in q
            float *t12 = pm, *ddis = pm + MAXF, *aband = pm + 2 * MAXF, *bn = pf;
            for (int m = threadIdx.x; m < ML; m += NT) {
                const float *f = s.F + KL * m;
                float ed = f[0];
                for (int k = 1; k < KL; k++) ed = ed + f[k];
                T[m] = ed * t.DELTH;                              // EDENS
                const float sigm = t.ZPI * t.FR[m];
                const float c0 = t.ZPI * sigm;
                float ab = 0.f, mx = 0.f;
                for (int k = 0; k < KL; k++) {
                    const float v = (f[k] * cg[m]) / c0;
                    ab = k == 0 ? v : ab + v;
                    mx = k == 0 ? v : wam_maxf(mx, v);
                }
                float sn = 0.f;
                for (int k = 0; k < KL; k++) {
                    const float v = (f[k] * cg[m]) / c0;
                    const float kn = (mx < 1.0E-34f) ? 1.f : v / mx;
                    sn = k == 0 ? kn : sn + kn;
                }
                const float anar = 1.0f / (sn * t.DELTH);
                aband[m] = ab;
                bn[m] = (anar * ((ab * sigm) * t.DELTH)) * wam_ipowf(wn[m], 3);
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                const float bnt = 0.035f * 0.035f;
                const float xfac = (1.0f - 1.0f / t.CO) / (t.CO - 1.0f / t.CO);
                float t2s = 0.f;
                for (int ik = 0; ik < ML; ik++) {
                    const float et = (t.ZPI * bnt) / ((1.0f * cg[ik]) * wam_ipowf(wn[ik], 3));
                    const float ex = wam_maxf(0.0f, T[ik] - et);
                    const float nex = ex / et;
                    const float n4 = wam_ipowf(nex, 4);
                    const float t1 = ((4.75E-6f * 1.0f) * t.FR[ik]) * n4;
                    float dfii = t.DF[ik];
                    if (ik > 0 && ik < ML - 1) dfii = dfii * xfac;
                    const float term = (1.0f * n4) * dfii;
                    t2s = ik == 0 ? term : t2s + term;
                    t12[ik] = -1.0f * (wam_maxf(0.0f, t1) + wam_maxf(0.0f, 7.00E-5f * t2s));
                }
                int ikx = 0;
                for (int m = 1; m < ML; m++) if (aband[m] > aband[ikx]) ikx = m;
                float sm = 0.f;
                for (int m = 0; m < ML; m++) {
                    const float sigm = t.ZPI * t.FR[m];
                    const float dden = (t.ZPI * t.DFIM[m]) * sigm;
                    const float v = (aband[m] * dden) / cg[m];
                    sm = m == 0 ? v : sm + v;
                }
                const float b1 = 0.0041f * ((2.f * sqrtf(sm)) * wn[ikx]);
                for (int m = 0; m < ML; m++) {
                    const float sigm = t.ZPI * t.FR[m];
                    ddis[m] = aband[m] > 1.E-30f ? (((-(2.f / 3.f)) * b1) * sigm) * sqrtf(bn[m]) : 0.f;
                }
            }
            __syncthreads();
            for (int b = threadIdx.x; b < NB; b += NT) {
                const int m = b / KL;
                const float f = s.F[b];
                float sl = s.SL[b] + t12[m] * f, fl = s.FL[b] + t12[m];
                s.SL[b] = sl + ddis[m] * f;
                s.FL[b] = fl + ddis[m];
            }
            __syncthreads();
        } else {
            // SDISSIP (:1593-1700) with FMEAN = F1MEAN, AKMEAN = XKMEAN
            for (int m = threadIdx.x; m < ML; m += NT) {
                float sds, x;
                const float e2 = P.emean * P.emean;
                if (SHALLOW) {
                    sds = ((t.CONSS * P.f1mean) * e2) * wam_ipowf(P.xkmean, 4);
                    x = t.TFAK[tb(t, m, P.indep)] / P.xkmean;
                } else {
                    sds = (t.CONSD * e2) * wam_ipowf(P.f1mean, 9);
                    x = wam_ipowf(t.FR[m] / P.f1mean, 2);
                }
                t2[m] = sds * (0.5f * x + 0.5f * wam_ipowf(x, 2));
            }
            __syncthreads();
            for (int b = threadIdx.x; b < NB; b += NT) {
                const int m = b / KL;
                s.SL[b] = s.SL[b] + t2[m] * s.F[b];
                s.FL[b] = s.FL[b] + t2[m];
            }
            __syncthreads();
        }
        // SNONLIN, with the ISNONLIN = 0 shallow-water enhancement
        if (SHALLOW && threadIdx.x == 0) {
            float enh = wam_maxf((0.75f * P.depth) * P.akmean, .5f);
            P.enhfr = 1.f + ((5.5f / enh) * (1.f - .833f * enh)) * wam_expf(-1.25f * enh);
        }
        __syncthreads();
        snonlin(t, s, SHALLOW ? P.enhfr : 1.f);
        if (SHALLOW) {
            // SBOTTOM (:1527-1589) into SL_BOT
            for (int m = threadIdx.x; m < ML; m += NT) {
                const float wav = t.TFAK[tb(t, m, P.indep)];
                const float sbo = wam_minf((2.f * P.depth) * wav, 50.f);
                t2[m] = (((-2.0f * 0.038f) / t.G) * wav) / wam_sinhf(sbo);
            }
            __syncthreads();
            for (int b = threadIdx.x; b < NB; b += NT) {
                const int m = b / KL;
                s.SLBOT[b] = 0.f + t2[m] * s.F[b];
                s.FL[b] = s.FL[b] + t2[m];
            }
            __syncthreads();
        }
        if (a.lcflx) {
            // WNFLUXES (:3970-4054) on SSOURCE = SL - SMIN
            for (int b = threadIdx.x; b < NB; b += NT) s.SMIN[b] = s.SL[b] - s.SMIN[b];
            __syncthreads();
            if (threadIdx.x == 0) {
                const float rog = t.ROWATER * t.G;
                float phioc = P.phiaw, tau = t.ROAIR * (P.ust * P.ust);
                float tx = tau * wam_sinf(P.udir), ty = tau * wam_cosf(P.udir);
                float pb = 0.f, bx = 0.f, by = 0.f;
                const float scdfm = (0.5f * t.DELTH) * (1.f - 1.f / t.CO);
                const int mij = P.mij;
                for (int m = 0; m < mij; m++) {
                    const float cfm = (m < mij - 1) ? rog * t.DFIM[m] : (rog * scdfm) * t.FR[mij - 1];
                    const float cm = SHALLOW ? t.TFAK[tb(t, m, P.indep)] / (t.ZPI * t.FR[m]) : (t.ZPI * t.FR[m]) / t.G;
                    for (int k = 0; k < KL; k++) {
                        const int b = k + KL * m;
                        const float ss = s.SMIN[b];
                        phioc = phioc - ss * cfm;
                        tx = tx - ((cm * ss) * cfm) * t.SINTH[k];
                        ty = ty - ((cm * ss) * cfm) * t.COSTH[k];
                        if (SHALLOW) {
                            const float sb = s.SLBOT[b];
                            pb = pb - sb * cfm;
                            bx = bx - ((cm * sb) * cfm) * t.SINTH[k];
                            by = by - ((cm * sb) * cfm) * t.COSTH[k];
                        }
                    }
                }
                if (a.fluxes) {
                    a.fluxes[0 * (size_t)NS + p] = phioc; a.fluxes[1 * (size_t)NS + p] = tx;
                    a.fluxes[2 * (size_t)NS + p] = ty; a.fluxes[3 * (size_t)NS + p] = pb;
                    a.fluxes[4 * (size_t)NS + p] = bx; a.fluxes[5 * (size_t)NS + p] = by;
                }
            }
        }
        // 4. limiter and implicit update
        for (int m = threadIdx.x; m < ML; m += NT) {
            const float delfl = (t.C5G / wam_ipowf(t.FR[m], 4)) * delt;
            t2[m] = (P.ust * delfl) * wam_maxf(P.fmeanws, P.fmean);
            T[m] = t.FLMINFR[(size_t)m * t.JUMAX + P.ju - 1];
        }
        __syncthreads();
        for (int b = threadIdx.x; b < NB; b += NT) {
            const int k = b % KL, m = b / KL;
            float sl = s.SL[b];
            if (SHALLOW) sl = sl + s.SLBOT[b];
            const float temp1 = (delt * sl) / wam_maxf(1.f, 1.f - delt * s.FL[b]);
            const float temp3 = wam_minf(wam_fabsf(temp1), t2[m]);
            const float f = s.F[b] + wam_signf(temp3, temp1);
            s.F[b] = wam_maxf(f, T[m] * sprd[k]);
        }
        __syncthreads();
        // 5. new mean parameters, cut-off index and tail
        dirsum(t, s.F, nullptr, T);
        __syncthreads();
        if (threadIdx.x == 0) {
            float e, fm;
            energy_fmean(t, T, e, fm);
            P.emean = e; P.fmean = fm;
        }
        __syncthreads();
        dirsum(t, s.F, s.LLWS, T);
        __syncthreads();
        if (threadIdx.x == 0) {
            float e, fm;
            energy_fmean(t, T, e, fm);
            P.emeanws = e; P.fmeanws = fm;
            P.mij = frcut(t, P.fmean, P.fmeanws, P.ust);
        }
        __syncthreads();
        imphftail(t, s, P, t2);
        for (int b = threadIdx.x; b < NB; b += NT) a.fl3[(size_t)b * NS + p] = s.F[b];
        if (threadIdx.x == 0) {
            a.tauw[p] = P.tauw; a.ustar[p] = P.ust; a.z0[p] = P.z0;
            if (a.phiaw) a.phiaw[p] = P.phiaw;
            if (a.mij) a.mij[p] = P.mij;
        }
        __syncthreads();
    }
}

size_t implsch_lds_bytes(const DevTables &t, int lcflx)
{
    const int NB = t.KL * t.ML;
    size_t n = 4 * (size_t)NB + (lcflx ? NB : 0) + (t.shallow ? NB : 0) + 4 * (size_t)(t.NQ > NB ? t.NQ : NB);
    return n * sizeof(float) + (size_t)NB;
}

hipError_t launch_implsch(const DevTables &t, const ImplArgs &a, int nblocks, hipStream_t st)
{
    const size_t lds = implsch_lds_bytes(t, a.lcflx);
    void (*k)(DevTables, ImplArgs) = nullptr;
    if (t.iphys == 2) k = t.shallow ? implsch_kernel<2, 1> : implsch_kernel<2, 0>;
    else k = t.shallow ? implsch_kernel<0, 1> : implsch_kernel<0, 0>;
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(NT), lds, st, t, a);
    return hipGetLastError();
}

}  // namespace wam
