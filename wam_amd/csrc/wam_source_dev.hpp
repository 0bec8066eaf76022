// wam_source_dev.hpp -- the device stages of IMPLSCH (wam_source_module.f90)
// shared by the forward kernels (wam_implsch.hip) and the adjoint kernel
// (wam_adjoint.hip): one sea point per workgroup, its spectrum and source
// terms in LDS, every sum in the reference's order with separate roundings.
#pragma once
#include "wam_dev.hpp"
#include "wam_math.h"

namespace wam {

constexpr int NT = 256;          // threads per workgroup
constexpr int MAXF = 64;         // frequencies and directions <= 64

struct PointScalars {
    float u10, udir, tauw, z0, roairn, wstar, depth;
    int indep, ju;
    float ust, emean, fmean, f1mean, akmean, xkmean, emeanws, fmeanws, phiaw, sig_n, b1, enhfr;
    int mij;
    float s_e, s_fm, s_tm, s_w1, s_w2;                     // partial integrals
    float em, sbr, dsbr;                                   // SDEPTHLIM's factor, SFBRK's rates (WAVE_BREAKING_RUN)
    // LFACTOR state
    float twx, twy, tauvx, tauvy, tau_tot, rtau, drtau;
    int sign_new, overshot, done, apply;
};

struct Smem {
    // point-invariant tables
    float *fr, *dfim, *dfimofr, *dfim_fr, *rhowg, *fr5, *frm5, *th, *costh, *sinth, *rnl, *af11, *sig10, *dsii10;
    int *inl, *kw;                                       // kw: K1W, K2W, K11W, K21W, each [2][KL], 0-based
    int *sprw;                                           // SP_RW (products-form run table)
    // per point
    float *F, *SL, *FL, *SPOS, *SMIN, *SLBOT, *SCR;
    unsigned char *LLWS;
};

__device__ __forceinline__ int tb(const DevTables &t, int m, int indep) { return m * t.NDEPTH + indep - 1; }

// TEMP(M) = SUM over K of F (optionally masked), thread base+m, natural order.
__device__ __forceinline__ void dirsum(const DevTables &t, const float *F, const unsigned char *mask, float *T,
                                       int tid)
{
    const int KL = t.KL;
    if (tid >= 0 && tid < t.ML) {
        const int m = tid;
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

// ---- the off-by-default sources (WAVE_BREAKING_RUN, PHILLIPS_RUN) ----------

constexpr float GAMD_DEV = 0.8f;   // breaker parameter of SDEPTHLIM and SFBRK (wam_source_module.f90:67)

// SDEPTHLIM's factor (:3735-3780): EM = MIN((0.25 GAMD DEPTH)**2 / EMEAN, 1);
// the spectrum and EMEAN are multiplied by it.
__device__ __forceinline__ float sdepthlim_em(float depth, float emean)
{
    const float h = (0.25f * GAMD_DEV) * depth;
    return wam_minf((h * h) / emean, 1.0f);
}

// SFBRK (:1966-2037): SL += SBR * F, FL += DSBR, with the fraction of
// breaking waves QB from CMPQB's Newton-Raphson step (:2041-2107) on
// BB = 8 EMEAN / (GAMD DEPTH)**2.
__device__ __forceinline__ void sfbrk_rates(float emean, float fmean, float depth, float &sbr, float &dsbr)
{
    const float gd = GAMD_DEV * depth;
    const float bb = (8.f * emean) / (gd * gd);
    float q0 = 0.f;
    if (bb >= 0.25f) {
        const float q = 2.f * sqrtf(bb) - 1.f;
        q0 = q * q;
    }
    float qb = 1.f;
    if (bb < 1.f) {
        const float ex = wam_expf((q0 - 1.f) / bb);
        qb = q0 - (bb * (q0 - ex)) / (bb - ex);
    }
    qb = wam_minf(1.f, qb);
    sbr = -(2.f * fmean);                   // -ALPHA*2.*FMEAN with ALPHA = 1
    if (bb <= 1.f) sbr = (sbr * qb) / bb;
    dsbr = 0.f;
    if (bb < 1.f && wam_fabsf(bb - qb) > 0.f) dsbr = (sbr * (1.f - qb)) / (bb - qb);
}

// SOURCE_PHILLIPS (:3051-3160): SL(K,M) += TEMP(M) * TPHOLD(K).  TEMP of
// frequency fr (tfk = TFAK(INDEP, M), shallow water only) ...
template <int SHALLOW>
__device__ __forceinline__ float phillips_temp(const DevTables &t, float fr, float ust, float tfk, float depth)
{
    const float fpm = wam_maxf(0.0000001f, (28.f / (0.13f * t.G)) * ust);
    const float xeps = t.ROAIR / t.ROWATER;
    const float c1 = ((80.f * 16.0f) * (xeps * xeps)) / ((0.5f * 3.0f) * (t.G * t.G));
    const float tp = wam_expf(-(1.f / wam_ipowf(fr * fpm, 4)));     // EXP(-(FR*FPM)**(-4))
    if (SHALLOW) {
        const float kd = wam_minf(tfk * depth, 40.f);
        return (c1 * tp) / (1.0f + (2.0f * kd) / wam_sinhf(2.f * kd));
    }
    return c1 * tp;
}

// ... and TPHOLD of direction th.
__device__ __forceinline__ float phillips_tph(float th, float udir, float ust)
{
    return wam_ipowf(ust * wam_maxf(0.f, wam_cosf(th - udir)), 4);
}

// seeded sequential sum c0*T(ML) + SUM_M T(M)*w(M) (the frequency integrals of
// TOTAL_ENERGY_B / FEMEAN / TM1_TM2_PERIODS_B, wam_interface_module.f90)
__device__ __forceinline__ float integ(int ml, float c0, const float *T, const float *w)
{
    float s = c0 * T[ml - 1];
    for (int m = 0; m < ml; m++) s = s + T[m] * w[m];
    return s;
}

// W3FLX4 (wam_source_module.f90:5372-5443)
__device__ __forceinline__ void w3flx4(const DevTables &t, float u10, float &ust, float &z0)
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
__device__ __forceinline__ void airsea(const DevTables &t, float utop, float tauw, float &ust, float &z0)
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
__device__ __forceinline__ float wsigstar(const DevTables &t, float ust, float z0, float wstar)
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
__device__ __forceinline__ int frcut(const DevTables &t, const float *fr, float fm, float fmws, float ust)
{
    if (t.iphys == 2) {
        const int nk = t.ML;
        const float fxfm = 6.f * t.ZPI, fxpm = (4.0f * t.G) / 28.f, signk = t.ZPI * fr[nk - 1];
        float fhigh = wam_maxf(fxfm * fm, fxpm / ust);
        const float f1 = 1.f / wam_logf(t.CO);
        const float f2 = 1.f - wam_logf(t.ZPI * fr[0]) * f1;
        int nkh = min(nk, wam_int(f2 + f1 * wam_logf(wam_maxf(1.E-7f, fhigh))));
        const int nkh1 = min(nk, nkh + 1);
        fhigh = wam_minf(signk, wam_maxf(fxfm * fm, fxpm / ust));
        return max(2, min(nkh1, wam_int(f2 + f1 * wam_logf(wam_maxf(1.E-7f, fhigh)))));
    }
    const float fpmh = 2.5f / fr[0];
    const float fppm = (3.0f * t.G) / ((28.0f * t.ZPI) * fr[0]);
    const float fm2 = wam_maxf(fmws, fm) * fpmh;
    const float fpm = fppm / wam_maxf(ust, 1.0E-6f);
    const float fpm4 = wam_maxf(fm2, fpm);
    // MIN(ML, MAX(1, NINT(...) + 1)), clamped before the + 1 (a saturated
    // NINT of a non-finite argument must not wrap)
    return min(max(0, wam_nint(wam_log10f(fpm4) * t.INV_LOG_CO)), t.ML - 1) + 1;
}

// SINPUT's factors of one frequency (:2111-2308): CNSN, UCN1, UCN2, the two
// XV1D terms and ZCN.
struct SinFreq {
    float cnsn, ucn1, ucn2, xv1, xv2, zcn;
};

template <int SHALLOW>
__device__ __forceinline__ SinFreq sin_freq(const DevTables &t, float fr_m, float tfk_m, float usn1, float usn2,
                                            float z0, float roairn)
{
    const float xk2 = t.XKAPPA * t.XKAPPA;
    const float fac = t.ZPI * fr_m;
    const float cst = ((fac * t.BETAMAX) / xk2) / t.ROWATER;
    float xk, cm, cnsn;
    if (SHALLOW) {
        xk = tfk_m;
        cm = xk / fac;
        cnsn = ((cst * wam_ipowf(fac, 2)) / (t.G * xk)) * roairn;
    } else {
        xk = wam_ipowf(fac, 2) / t.G;
        cm = fac / t.G;
        cnsn = cst * roairn;
    }
    SinFreq r;
    r.zcn = wam_logf(xk * z0);
    r.cnsn = cnsn;
    r.ucn1 = usn1 * cm + t.ZALP;
    r.ucn2 = usn2 * cm + t.ZALP;
    r.xv1 = -1.f / (((usn1 / t.XKAPPA) * r.zcn) * cm);
    r.xv2 = -1.f / (((usn2 / t.XKAPPA) * r.zcn) * cm);
    return r;
}

// SINPUT of one bin with TEMPD = COS(TH(K) - UDIR) = tk: the positive input
// SPOS/F (sp), the derivative FL = SL/F (fl) and LLWS (ll).
__device__ __forceinline__ void sin_bin(const DevTables &t, const SinFreq &r, float tk, float const3, float &sp,
                                        float &fl, bool &ll)
{
    float ufac1 = 0.f;
    ll = false;
    if (tk > 0.01f) {
#pragma unroll
        for (int is = 0; is < 2; is++) {
            const float x = tk * (is == 0 ? r.ucn1 : r.ucn2);
            const float zlog = r.zcn + t.XKAPPA / x;
            if (zlog < 0.f) {
                const float z2x = (zlog * zlog) * x;
                ufac1 = ufac1 + ((0.5f * wam_expf(zlog)) * z2x) * z2x;
                ll = true;
            }
        }
    }
    float ufac2 = (const3 * (tk - r.xv1)) * wam_ipowf(r.ucn1, 2);
    ufac2 = 0.5f * ufac2 + ((0.5f * const3) * (tk - r.xv2)) * wam_ipowf(r.ucn2, 2);
    sp = r.cnsn * ufac1;
    fl = sp + r.cnsn * ufac2;
}

// SINPUT (:2111-2308), NSIN = 2: SL, SPOS, FL, LLWS for every bin.
template <int SHALLOW>
__device__ __forceinline__ void sinput(const DevTables &t, const Smem &s, PointScalars &P, const float *cu, const float *tfk,
                       SinFreq *q)
{
    const int KL = t.KL, ML = t.ML, NB = KL * ML, tid = threadIdx.x;
    if (tid == 0) P.sig_n = wsigstar(t, P.ust, P.z0, P.wstar);
    __syncthreads();
    const float usn1 = P.ust * (1.0f - P.sig_n), usn2 = P.ust * (1.0f + P.sig_n);
    if (tid < ML) q[tid] = sin_freq<SHALLOW>(t, s.fr[tid], SHALLOW ? tfk[tid] : 0.f, usn1, usn2, P.z0, P.roairn);
    __syncthreads();
    const float const3 = ((2.f * t.XKAPPA) / t.BETAMAX) * (t.XKAPPA * t.XKAPPA);
    for (int b = tid; b < NB; b += NT) {
        const int m = b / KL, k = b - KL * m;
        float sp, fl;
        bool ll;
        sin_bin(t, q[m], cu[k], const3, sp, fl, ll);
        const float f = s.F[b];
        s.SPOS[b] = sp * f;
        s.FL[b] = fl;
        s.SL[b] = fl * f;
        s.LLWS[b] = ll;
    }
    __syncthreads();
}

// value of v in lane i of the wavefront (i uniform)
__device__ __forceinline__ float lane_at(float v, int i)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), i));
}

// Two ordered sums at once: sa = a_0 + a_1 + ... + a_{n-1} and sb likewise,
// added in this order (1 <= n <= 128), where a_i is a0 of lane i for i < 64
// and a1 of lane i - 64 above (b the same).  The first 64 terms are unrolled
// without branches; the whole wavefront must be active.
__device__ __forceinline__ void lane_sum2x(float a0, float a1, float b0, float b1, int n, float &sa, float &sb)
{
    float x = lane_at(a0, 0), y = lane_at(b0, 0);
    const int n0 = n < 64 ? n : 64;
    for (int i0 = 0; i0 < n0; i0 += 8) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int i = i0 + u;
            if (i == 0) continue;
            const float xi = x + lane_at(a0, i), yi = y + lane_at(b0, i);
            x = i < n0 ? xi : x;
            y = i < n0 ? yi : y;
        }
    }
    for (int i = 64; i < n; i++) {
        x = x + lane_at(a1, i - 64);
        y = y + lane_at(b1, i - 64);
    }
    sa = x;
    sb = y;
}

// SINPUT_ST6 (:4704-4925) with LFACTOR (:4465-4700) and TAUWINDS (:5555-5587).
// pf: per-frequency work (>= 5*MAXF floats).
__device__ __forceinline__ void sinput_st6(const DevTables &t, const Smem &s, PointScalars &P, float *pf, const float *wn,
                           const float *cg)
{
    const int KL = t.KL, NK = t.ML, NB = KL * NK, NK10 = t.NK10, tid = threadIdx.x;
    float *sig = pf, *cinv = pf + MAXF, *sqrtbn = pf + 2 * MAXF;
    float *lf10 = pf + 4 * MAXF;
    float *D = s.SCR, *W2 = s.SCR + NB, *SDS = s.SCR + 2 * NB;
    const float ZPI = t.ZPI;
    if (tid < NK) {
        const int m = tid;
        const float sg = ZPI * s.fr[m];
        sig[m] = sg;
        const float c0 = ZPI * sg;
        const float *f = s.F + KL * m;
        float sm = 0.f, mx = 0.f;
        for (int k = 0; k < KL; k++) {
            const float a = (f[k] * cg[m]) / c0;
            sm = k == 0 ? a : sm + a;
            mx = k == 0 ? a : wam_maxf(mx, a);
        }
        const float adens = (sm * sg) * t.DELTH;
        float sn = 0.f;
        for (int k = 0; k < KL; k++) {
            const float a = (f[k] * cg[m]) / c0;
            const float v = (mx < 1.0E-34f) ? 1.f : a / mx;
            sn = k == 0 ? v : sn + v;
        }
        const float anar = 1.0f / (sn * t.DELTH);
        sqrtbn[m] = sqrtf((anar * adens) * wam_ipowf(wn[m], 3));
        cinv[m] = wn[m] / sg;
    }
    __syncthreads();
    const float cosu = wam_cosf(P.udir), sinu = wam_sinf(P.udir);
    const float uproxy = 32.f * P.ust, rw = P.roairn / t.ROWATER;
    for (int b = tid; b < NB; b += NT) {
        const int m = b / KL, k = b - KL * m;
        const float cosd = s.costh[k] * cosu + s.sinth[k] * sinu;
        const float x = ((uproxy * cinv[m]) * cosd) - 1.f;
        const float w1 = wam_ipowf(wam_maxf(0.f, x), 2);
        W2[b] = wam_ipowf(wam_minf(0.f, x), 2);
        const float d = (((rw * sig[m]) * (2.8f - (1.f + wam_tanhf((10.f * sqrtbn[m]) * w1 - 11.f)))) * sqrtbn[m]) * w1;
        D[b] = d;
        const float a = (s.F[b] * cg[m]) / (ZPI * sig[m]);
        SDS[b] = ((d * a) * sig[m]) / cg[m];
    }
    __syncthreads();
    // LFACTOR with TAUWINDS on wave 0 alone: lane l holds the 10-Hz grid
    // points i = l and i = l + 64, and the ordered sums over i read the lanes
    // in turn, so the stress iteration needs no workgroup barrier.
    if (tid < 64) {
        const int l = tid;
        float sxm = 0.f, sym = 0.f;            // input spectral density components at frequency l
        if (l < NK) {
            const float *v = SDS + KL * l;
            float bx = 0.f, by = 0.f;
            for (int k = 0; k < KL; k++) {
                const float vp = wam_maxf(0.f, v[k]);
                if (k == 0) { bx = vp * s.costh[k]; by = vp * s.sinth[k]; }
                else { bx = bx + vp * s.costh[k]; by = by + vp * s.sinth[k]; }
            }
            sxm = bx * t.DELTH; sym = by * t.DELTH;
        }
        const float sxl = lane_at(sxm, NK - 1), syl = lane_at(sym, NK - 1);
        const float ust = P.ust;
        float ci[2], sxv[2], syv[2], ds[2], uc[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int i = l + 64 * h;
            ci[h] = 0.f; sxv[h] = 0.f; syv[h] = 0.f; ds[h] = 0.f;
            if (i < NK) {
                ci[h] = cinv[i]; sxv[h] = sxm; syv[h] = sym;
            } else if (i < NK10) {
                ci[h] = s.sig10[i] * 0.101978f;
                const float r2 = wam_ipowf(s.sig10[NK - 1] / s.sig10[i], 2);
                sxv[h] = sxl * r2; syv[h] = syl * r2;
            }
            if (i < NK10) ds[h] = s.dsii10[i];
            uc[h] = 1.0f - (32.f * ust) * ci[h];
        }
        const float GRW = t.G * t.ROWATER;
        const float u10 = P.u10;
        const float tau_tot = (ust * ust) * P.roairn;
        float tau_vis = (wam_maxf(0.0f, -5.0E-5f * u10 + 1.1E-3f) * (u10 * u10)) * P.roairn;
        tau_vis = wam_minf(0.95f * tau_tot, tau_vis);
        const float tauvx = tau_vis * wam_cosf(P.udir), tauvy = tau_vis * wam_sinf(P.udir);
        float twx, twy;
        lane_sum2x((sxv[0] * ci[0]) * ds[0], (sxv[1] * ci[1]) * ds[1], (syv[0] * ci[0]) * ds[0],
                   (syv[1] * ci[1]) * ds[1], NK10, twx, twy);
        twx = GRW * twx;
        twy = GRW * twy;
        float taux = tauvx + twx, tauy = tauvy + twy;
        float tau = sqrtf(taux * taux + tauy * tauy);
        float err = (tau - tau_tot) / tau_tot;
        float lf0 = 1.0f, lf1 = 1.0f;
        if (tau > tau_tot) {
            bool overshot = false;
            float rtau = err / 90.f, drtau = 2.0f;
            int sign_new = err >= 0.f ? 1 : -1;
            for (int it = 1; it <= 80; it++) {
                lf0 = wam_minf(1.0f, wam_expf(uc[0] * rtau));
                lf1 = wam_minf(1.0f, wam_expf(uc[1] * rtau));
                lane_sum2x(((sxv[0] * lf0) * ci[0]) * ds[0], ((sxv[1] * lf1) * ci[1]) * ds[1],
                           ((syv[0] * lf0) * ci[0]) * ds[0], ((syv[1] * lf1) * ci[1]) * ds[1], NK10, twx, twy);
                twx = GRW * twx;
                twy = GRW * twy;
                taux = tauvx + twx; tauy = tauvy + twy;
                tau = sqrtf(taux * taux + tauy * tauy);
                err = (tau - tau_tot) / tau_tot;
                const int sign_old = sign_new;
                sign_new = err >= 0.f ? 1 : -1;
                if (sign_new != sign_old) overshot = true;
                if (overshot) drtau = wam_maxf(0.5f * (1.0f + drtau), 1.00010f);
                rtau = sign_new > 0 ? rtau * drtau : rtau * (1.f / drtau);
                if (wam_fabsf(err) < 1.54E-4f) break;
            }
        }
        if (l < NK) lf10[l] = lf0;
        float ls, ls_unused;
        lane_sum2x(lf0, lf1, lf0, lf1, NK, ls, ls_unused);
        if (l == 0) {
            P.apply = ls < (float)NK;
            P.tauw = sqrtf(twx * twx + twy * twy);
        }
    }
    __syncthreads();
    const bool apply = P.apply != 0;
    for (int b = tid; b < NB; b += NT) {
        const int m = b / KL;
        float d = D[b];
        if (apply) d = d * lf10[m];
        const float dpos = d;
        const float w2 = W2[b];
        d = d - ((((rw * sig[m]) * 0.09f) * (2.8f - (1.f + wam_tanhf((10.f * sqrtbn[m]) * w2 - 11.f)))) * sqrtbn[m]) * w2;
        const float f = s.F[b];
        const float sp = dpos * f;
        s.SL[b] = d * f;
        s.SPOS[b] = sp;
        s.FL[b] = d;
        s.LLWS[b] = sp > 0.f;
    }
    __syncthreads();
}

// TAU_PHI_HF with XLEVTAIL = 0 (:3853-3966), sequential form.
__device__ __forceinline__ void tau_phi_hf_seq(const DevTables &t, const float *fr, int mij, float ust0, float z0, float &tauhf,
                               float &phihf)
{
    const float gm1 = 1.0f / t.G;
    const float omegac = t.ZPI * fr[mij - 1];
    const float tauw0 = ust0 * ust0;
    const float xloggz0 = wam_logf(t.G * z0);
    const float omegacc = wam_maxf(omegac, (t.X0TAUHF * t.G) / ust0);
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

// STRESSO (:3164-3356) with SL = SPOS and TAU_PHI_HF (:3853-3966, XLEVTAIL = 0).
// The per-frequency products and the 19 quadrature nodes are formed in
// parallel; thread 0 adds them up in the reference's order.  pm >= 3*MAXF + 64.
// keep (the adjoint's recomputation) receives the stress components after
// the high-frequency part, TAUHF, TAU_PHI_HF's stress integral, the
// cos**3-weighted sum of F(:, MIJ), the stress modulus before its clamps and
// CONST1.
template <int SHALLOW>
__device__ __forceinline__ void stresso(const DevTables &t, const Smem &s, PointScalars &P, const float *cu, const float *tfk,
                        float *pm, bool want_tauw, float *keep = nullptr)
{
    const int KL = t.KL, ML = t.ML, mij = P.mij, tid = threadIdx.x;
    float *am = pm, *xm = pm + MAXF, *ym = pm + 2 * MAXF, *hf = pm + 3 * MAXF;   // hf: 19 fnc2, 19 fnc2/y, t1, t2
    const float gm1 = 1.0f / t.G;
    if (tid < mij) {
        const int m = tid;
        const float *v = s.SPOS + KL * m;
        float a = 0.f, x = 0.f, y = 0.f;
        for (int k = 0; k < KL; k++) {
            const float sp = wam_maxf(v[k], 0.f);
            if (k == 0) { a = sp; x = sp * s.sinth[k]; y = sp * s.costh[k]; }
            else { a = a + sp; x = x + sp * s.sinth[k]; y = y + sp * s.costh[k]; }
        }
        const float cm = SHALLOW ? tfk[m] * (1.0f / (t.ZPI * s.fr[m])) : (t.ZPI * s.fr[m]) * gm1;
        float rw = s.rhowg[m];
        if (m == mij - 1 && mij != ML) rw = 0.5f * rw;
        am[m] = a * rw;
        const float cmrw = cm * rw;
        xm[m] = x * cmrw;
        ym[m] = y * cmrw;
    } else if (tid >= 64 && tid < 64 + 19) {
        // TAU_PHI_HF node j; with XLEVTAIL = 0 the wave stress stays at USTAR**2,
        // so node 1 uses USTAR and the later nodes SQRT(USTAR**2)
        const int j = tid - 64 + 1;
        const float ust0 = P.ust, z0 = P.z0;
        const float omegac = t.ZPI * s.fr[mij - 1];
        const float tauw0 = ust0 * ust0;
        const float xloggz0 = wam_logf(t.G * z0);
        const float omegacc = wam_maxf(omegac, (t.X0TAUHF * t.G) / ust0);
        const float sqz = sqrtf(z0 * gm1), sqg = 1.0f / sqz;
        const float zinf = wam_logf(omegacc * sqz);
        const float delz = wam_maxf((0.0f - zinf) / 18.f, 0.0f);
        const float ust = j == 1 ? ust0 : sqrtf(tauw0);
        const float y = wam_expf(zinf + (float)(j - 1) * delz);
        const float cm1 = (y * sqg) * gm1;
        const float zx = ust * cm1 + t.ZALP;
        const float zarg = t.XKAPPA / zx;
        float zlog = (xloggz0 + 2.0f * wam_logf(cm1)) + zarg;
        zlog = wam_minf(zlog, 0.0f);
        const float zbeta = wam_expf(zlog) * wam_ipowf(zlog, 4);
        const float fnc2 = ((zbeta * tauw0) * t.WTAUHF[j - 1]) * delz;
        hf[j - 1] = fnc2;
        hf[19 + j - 1] = fnc2 / y;
        if (j == 1) hf[40] = sqz;
    } else if (tid == 128) {
        float t1 = 0.f, t2 = 0.f;
        for (int k = 0; k < KL; k++) {
            const float cosw = wam_maxf(cu[k], 0.f);
            const float f = s.F[k + KL * (mij - 1)];
            if (k == 0) { t2 = f * wam_ipowf(cosw, 2); t1 = t2 * cosw; }
            else { t1 = t1 + f * wam_ipowf(cosw, 3); t2 = t2 + f * wam_ipowf(cosw, 2); }
        }
        hf[38] = t1; hf[39] = t2;
    }
    __syncthreads();
    if (tid == 0) {
        const float cst = (t.DELTH * t.ZPI4) * gm1;
        const float c1 = (cst * s.fr5[mij - 1]) * gm1;
        const float c2 = (t.ROAIR * cst) * s.fr5[mij - 1];
        float phi = 0.f, xs = 0.f, ys = 0.f;
        for (int m = 0; m < mij; m++) { phi = phi + am[m]; xs = xs + xm[m]; ys = ys + ym[m]; }
        xs = xs / wam_maxf(t.ROAIR, 1.f);
        ys = ys / wam_maxf(t.ROAIR, 1.f);
        float th = 0.f, ph = 0.f;
        bool fin = true;
        for (int j = 0; j < 19; j++) {
            th = th + hf[j];
            ph = ph + hf[19 + j];
            fin = fin && (j == 18 || (wam_f2i(hf[j]) & 0x7f800000) != 0x7f800000);
        }
        float tau1 = th, phi1 = hf[40] * ph;
        // a non-finite node feeds back into the running stress of the later
        // nodes in the sequential form: evaluate that form then
        if (!fin) tau_phi_hf_seq(t, s.fr, mij, P.ust, P.z0, tau1, phi1);
        const float tauhf = (c1 * hf[38]) * tau1, phihf = (c2 * hf[39]) * phi1;
        phi = phi + phihf;
        xs = xs + tauhf * wam_sinf(P.udir);
        ys = ys + tauhf * wam_cosf(P.udir);
        float tw = sqrtf(xs * xs + ys * ys);
        if (keep) {
            keep[0] = xs; keep[1] = ys; keep[2] = tauhf; keep[3] = tau1; keep[4] = hf[38]; keep[5] = tw;
            keep[6] = c1;
        }
        tw = wam_minf(tw, P.ust * P.ust - t.EPS1);
        tw = wam_maxf(tw, 0.f);
        if (want_tauw) P.tauw = tw;
        P.phiaw = phi;
    }
    __syncthreads();
}

// IMPHFTAIL (:3651-3731) on the bins above MIJ, by the threads from tid0 on.
template <int SHALLOW>
__device__ __forceinline__ void imphftail(const DevTables &t, const Smem &s, int mij, const float *tfk, const float *tcg, float *t2,
                          int tid0)
{
    const int KL = t.KL, ML = t.ML, tid = threadIdx.x;
    if (tid >= tid0 && tid - tid0 < ML) {
        const int m = tid - tid0;
        if (m >= mij) {
            float a, b;
            if (SHALLOW) {
                const float akm1 = 1.f / tfk[m], akm0 = 1.f / tfk[mij - 1];
                a = wam_ipowf(akm1, 3) / tcg[m];
                b = wam_ipowf(akm0, 3) / tcg[mij - 1];
            } else { a = s.frm5[m]; b = s.frm5[mij - 1]; }
            t2[m] = a / b;
        }
    }
}

// RNLCOEF slots (0-based) of the SL and FL weights of DIA update types 1..8
// (type 0, the centre bin, has weight -2 for both), and which derivative the
// FL update takes: 0 = DD, 1 = DM (types 1-4), 2 = DP (types 5-8).
static __constant__ int SLOT_SL[9] = {0, 20, 19, 17, 18, 8, 7, 5, 6};
static __constant__ int SLOT_FL[9] = {0, 23, 24, 21, 22, 11, 12, 9, 10};

// The terms of DIA quadruplet (MC = mcm1 + 1, KH*KL + K = r) of SNONLIN
// (:2890-2960): {AD, DD, DM, DP} = the SL update of the centre bin over -2,
// the FL update of the centre bin over -2, and the two FL update factors of
// the MM/MM1 and MP/MP1 bins.
__device__ __forceinline__ float4 dia_terms(const DevTables &t, const Smem &s, int mcm1, int r, float ftemp_scale)
{
    const int KL = t.KL, KL2 = 2 * KL;
    const float *F = s.F;
    const int *ic = s.inl + mcm1 * 5;
    const float *rc = s.rnl + mcm1 * 25;
    const int k1 = s.kw[r], k2 = s.kw[KL2 + r], k11 = s.kw[2 * KL2 + r], k21 = s.kw[3 * KL2 + r];
    const int k = r < KL ? r : r - KL;
    const int IC = (ic[0] - 1) * KL, IP = (ic[1] - 1) * KL, IP1 = (ic[2] - 1) * KL;
    const int IM = (ic[3] - 1) * KL, IM1 = (ic[4] - 1) * KL;
    const float sap = rc[1] * F[k1 + IP] + rc[2] * F[k11 + IP] + rc[3] * F[k1 + IP1] + rc[4] * F[k11 + IP1];
    const float sam = rc[13] * F[k2 + IM] + rc[14] * F[k21 + IM] + rc[15] * F[k2 + IM1] + rc[16] * F[k21 + IM1];
    const float fij = F[k + IC] * rc[0];
    const float ftemp = s.af11[mcm1] * ftemp_scale;
    float fad1 = fij * (sap + sam);
    const float fad2 = fad1 - (2.f * sap) * sam;
    fad1 = fad1 + fad2;
    const float fcen = ftemp * fij;
    return make_float4(fad2 * fcen, fad1 * ftemp, ((fij - 2.f * sap) * t.DAL2) * fcen,
                       ((fij - 2.f * sam) * t.DAL1) * fcen);
}

// SNONLIN (:2742-3047) in chunks of CHUNK frequencies.  The quadruplet terms
// of a chunk are formed in parallel, as float4 {AD, DD, DM, DP} per
// quadruplet (slot 511 is kept zero for the padding entries), with a table
// {SL weight, FL weight, derivative} per (MC, update type); then every bin
// the chunk feeds adds its updates in the reference's order.  Adding
// AD*(-2) is the reference's -(2*AD) exactly.  qslot[i] = MC-offset | r << 8
// of the quadruplet slot threadIdx.x + i*NT within a chunk (r = KH*KL + K).
__device__ __forceinline__ void snonlin(const DevTables &t, const Smem &s, float ftemp_scale, const int (&qslot)[2])
{
    constexpr int G = SN_GROUP_DEV;
    const int KL = t.KL, tid = threadIdx.x, KL2 = 2 * KL;
    float4 *QV = reinterpret_cast<float4 *>(s.SCR);
    float4 *WT = QV + 512;
    const int lane0 = tid & ~63;
    for (int c = 0; c < t.NCHUNK; c++) {
        const int mc0 = c * t.CHUNK;                  // 0-based first MC of the chunk
        const int nmc = min(t.CHUNK, t.MLSTHG - mc0);
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const int mcl = qslot[i] & 255, r = qslot[i] >> 8;
            if (mcl >= nmc) continue;
            QV[tid + i * NT] = dia_terms(t, s, mc0 + mcl, r, ftemp_scale);
        }
        if (tid >= NT - 9 * nmc) {
            const int wi = tid - (NT - 9 * nmc), mcl = wi / 9, ty = wi - 9 * mcl;
            const float *rc = s.rnl + (mc0 + mcl) * 25;
            const float wsl = ty == 0 ? -2.f : rc[SLOT_SL[ty]], wfl = ty == 0 ? -2.f : rc[SLOT_FL[ty]];
            WT[wi] = make_float4(wsl, wfl, __int_as_float(ty == 0 ? 0 : (ty <= 4 ? 1 : 2)), 0.f);
        }
        if (tid == 0) QV[511] = make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        const int *ch = t.SN_CH + 5 * c;
        const int blo = ch[0], bhi = ch[1], nbc = bhi - blo;
        const uint16_t *E = t.SN_E + ch[3];
        const int *ww = t.SN_WW + ch[4];
        for (int b0 = blo + lane0; b0 < bhi; b0 += NT) {
            // one wavefront: bins b0 .. b0 + 63, list length ww of that run
            const int b = b0 + (tid & 63);
            const int W = ww[(b0 - blo) >> 6];
            const bool inb = b < bhi;
            const uint16_t *e = E + (inb ? b - blo : 0);
            float sl = inb ? s.SL[b] : 0.f, fl = inb ? s.FL[b] : 0.f;
            for (int j0 = 0; j0 < W; j0 += G) {
                unsigned v[G];
#pragma unroll
                for (int u = 0; u < G; u++) v[u] = inb ? (unsigned)e[(size_t)(j0 + u) * nbc] : SN_NONE_DEV;
#pragma unroll
                for (int u = 0; u < G; u++) {
                    const float4 q = QV[v[u] & 511u];
                    const float4 w = WT[v[u] >> 9];
                    const int cls = __float_as_int(w.z);
                    const float dq = cls == 0 ? q.y : (cls == 1 ? q.z : q.w);
                    sl = sl + q.x * w.x;
                    fl = fl + dq * w.y;
                }
            }
            if (inb) {
                s.SL[b] = sl;
                s.FL[b] = fl;
            }
        }
        __syncthreads();
    }
}

// WNFLUXES (:3970-4054) on SSOURCE = SL - SMIN (in s.SMIN) and SL_BOT, one
// thread: the six flux components out[c * ns], c = 0..5.
template <int SHALLOW>
__device__ __forceinline__ void wnfluxes(const DevTables &t, const Smem &s, float phiaw, float ust, float sinu,
                                         float cosu, int mij, const float *tfk, float *out, int ns)
{
    const int KL = t.KL;
    const float rog = t.ROWATER * t.G;
    float phioc = phiaw, tau = t.ROAIR * (ust * ust);
    float tx = tau * sinu, ty = tau * cosu;      // SIN(UDIR), COS(UDIR)
    float pb = 0.f, bx = 0.f, by = 0.f;
    const float scdfm = (0.5f * t.DELTH) * (1.f - 1.f / t.CO);
    for (int m = 0; m < mij; m++) {
        const float cfm = (m < mij - 1) ? rog * s.dfim[m] : (rog * scdfm) * s.fr[mij - 1];
        const float cm = SHALLOW ? tfk[m] / (t.ZPI * s.fr[m]) : (t.ZPI * s.fr[m]) / t.G;
        for (int k = 0; k < KL; k++) {
            const int b = k + KL * m;
            const float ss = s.SMIN[b];
            phioc = phioc - ss * cfm;
            tx = tx - ((cm * ss) * cfm) * s.sinth[k];
            ty = ty - ((cm * ss) * cfm) * s.costh[k];
            if (SHALLOW) {
                const float sb = s.SLBOT[b];
                pb = pb - sb * cfm;
                bx = bx - ((cm * sb) * cfm) * s.sinth[k];
                by = by - ((cm * sb) * cfm) * s.costh[k];
            }
        }
    }
    out[0] = phioc; out[(size_t)ns] = tx; out[2 * (size_t)ns] = ty;
    out[3 * (size_t)ns] = pb; out[4 * (size_t)ns] = bx; out[5 * (size_t)ns] = by;
}

// SNONLIN (:2742-3047) in the products form for a workgroup of NTH threads,
// for the bins this thread owns, b = threadIdx.x + NTH j, whose SL and FL
// run in sl[j], fl[j].  Per chunk
// of SP_CHUNK frequencies every quadruplet's thread forms its nine updates,
// U[type*S + ql] = {AD * SL weight, (DD|DM|DP) * FL weight} (type 0 the
// centre bin, weights -2; types 1-4 the MM/MM1 bins with DM, 5-8 the MP/MP1
// bins with DP), each product rounded as in the reference's update; then
// every bin adds its updates from the chunk in the reference's order
// (host-built lists, two entries per 32-bit word).  U[S - 1] = (-0, -0) pads.
// The nine updates of one quadruplet, u[type * S], S the type stride (see
// snonlin_products), from the MC's update weights wa..wd (snl_coef_block).
template <int S>
__device__ __forceinline__ void dia_products(float2 *u, float4 wa, float4 wb, float4 wc, float4 wd, float qx,
                                             float qy, float qz, float qw)
{
    u[0] = make_float2(qx * -2.f, qy * -2.f);
    u[1 * S] = make_float2(qx * wa.x, qz * wa.y);
    u[2 * S] = make_float2(qx * wa.z, qz * wa.w);
    u[3 * S] = make_float2(qx * wb.x, qz * wb.y);
    u[4 * S] = make_float2(qx * wb.z, qz * wb.w);
    u[5 * S] = make_float2(qx * wc.x, qw * wc.y);
    u[6 * S] = make_float2(qx * wc.z, qw * wc.w);
    u[7 * S] = make_float2(qx * wd.x, qw * wd.y);
    u[8 * S] = make_float2(qx * wd.z, qw * wd.w);
}

// The LDS block of one MC for implsch_snl_kernel, 36 words (9 x 16 bytes):
// RNLCOEF(1..5, 14..17) = rc[0..4], rc[13..16]; AF11; the byte offsets
// (INLCOEF - 1)*KL*4 of rows IC, IP, IP1, IM, IM1; then the update weights
// (SL, FL) of types 1..8: rc[20, 23], rc[19, 24], rc[17, 21], rc[18, 22],
// rc[8, 11], rc[7, 12], rc[5, 9], rc[6, 10].
__device__ __forceinline__ float snl_coef_block(const DevTables &t, int mc, int u)
{
    const float *rc = t.RNLCOEF + 25 * mc;
    switch (u) {
    case 0: case 1: case 2: case 3: case 4: return rc[u];
    case 5: case 6: case 7: case 8: return rc[u + 8];
    case 9: return t.AF11[mc];
    case 12: case 13: case 14: case 15: case 16:
        return __int_as_float((t.INLCOEF[5 * mc + u - 12] - 1) * t.KL * 4);
    case 20: return rc[20]; case 21: return rc[23]; case 22: return rc[19]; case 23: return rc[24];
    case 24: return rc[17]; case 25: return rc[21]; case 26: return rc[18]; case 27: return rc[22];
    case 28: return rc[8]; case 29: return rc[11]; case 30: return rc[7]; case 31: return rc[12];
    case 32: return rc[5]; case 33: return rc[9]; case 34: return rc[6]; case 35: return rc[10];
    default: return 0.f;
    }
}

// One gather step of a bin: SP_GROUP_DEV entry pairs (byte offsets into U,
// two per word), all eight LDS reads issued before the adds, which keep the
// entries' order.
__device__ __forceinline__ void gather_group(const char *Ub, const uint32_t (&v)[SP_GROUP_DEV], float &a, float &b)
{
    float2 x0[SP_GROUP_DEV], x1[SP_GROUP_DEV];
#pragma unroll
    for (int u = 0; u < SP_GROUP_DEV; u++) {
        x0[u] = *reinterpret_cast<const float2 *>(Ub + (v[u] & 0xffffu));
        x1[u] = *reinterpret_cast<const float2 *>(Ub + (v[u] >> 16));
    }
#pragma unroll
    for (int u = 0; u < SP_GROUP_DEV; u++) {
        a = a + x0[u].x; b = b + x0[u].y;
        a = a + x1[u].x; b = b + x1[u].y;
    }
}

template <int NTH, int NJ>
__device__ __forceinline__ void snonlin_products(const DevTables &t, const Smem &s, float2 *U, float ftemp_scale,
                                                 const int (&qslot)[256 * SP_QPT_DEV / NTH],
                                                 const unsigned (&qdir)[256 * SP_QPT_DEV / NTH],
                                                 const unsigned (&qdk)[256 * SP_QPT_DEV / NTH], float (&sl)[NJ],
                                                 float (&fl)[NJ])
{
    constexpr int S = 256 * SP_QPT_DEV, QT = S / NTH, NW = NTH / 64, G = SP_GROUP_DEV;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const char *Fb = reinterpret_cast<const char *>(s.F);
    const char *Ub = reinterpret_cast<const char *>(U);
    auto f_at = [&](unsigned off) { return *reinterpret_cast<const float *>(Fb + off); };
    for (int c = 0; c < t.SP_NCHUNK; c++) {
        const int mc0 = c * t.SP_CHUNK;
        const int nmc = min(t.SP_CHUNK, t.MLSTHG - mc0);
        // this wave's runs of the chunk (entry offset, pair count: uniform,
        // from the LDS copy of SP_RW) and the first group of each run's
        // list, loaded now so that the loads overlap the quadruplet phase
        const int2 *rw = reinterpret_cast<const int2 *>(s.sprw) + (size_t)c * t.SP_NRUN;
        int roff[NJ], rnp[NJ];
        uint32_t pv[NJ][G];
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int r = NW * j + wave;
            const int2 v = r < t.SP_NRUN ? rw[r] : make_int2(t.SP_PAD, 0);
            roff[j] = __builtin_amdgcn_readfirstlane(v.x);
            rnp[j] = __builtin_amdgcn_readfirstlane(v.y) >> 1;
            const uint32_t *e = reinterpret_cast<const uint32_t *>(t.SP_E + roff[j]) + lane;
#pragma unroll
            for (int u = 0; u < G; u++) pv[j][u] = e[u * 64];
        }
#pragma unroll
        for (int i = 0; i < QT; i++) {
            const int mcl = qslot[i] & 255;
            if (mcl >= nmc) continue;
            const int ql = tid + i * NTH, mcm1 = mc0 + mcl;
            // the MC's block of 9 float4 (snl_coef_block): the terms'
            // coefficients, the byte offsets into F of the frequency rows
            // IC, IP, IP1, IM, IM1 (s.rnl holds (INLCOEF - 1) * KL * 4), the
            // update weights
            const float4 *cb = reinterpret_cast<const float4 *>(s.rnl) + mcm1 * 9;
            const float4 ca = cb[0], cc = cb[1], ce = cb[2];
            const int4 ic4 = reinterpret_cast<const int4 *>(cb)[3];
            const unsigned IM1 = reinterpret_cast<const int4 *>(cb)[4].x;
            const float4 wa = cb[5], wb = cb[6], wc = cb[7], wd = cb[8];
            const unsigned IC = ic4.x, IP = ic4.y, IP1 = ic4.z, IM = ic4.w;
            // the quadruplet's directions K1, K11, K2, K21 (bytes of qdir)
            // and K, as byte offsets
            const unsigned d = qdir[i];
            const unsigned k1 = d & 255u, k11 = (d >> 8) & 255u, k2 = (d >> 16) & 255u, k21 = d >> 24;
            const unsigned k = qdk[i];
            // the nine spectral values first, then dia_terms' arithmetic
            const float fp0 = f_at(k1 + IP), fp1 = f_at(k11 + IP), fp2 = f_at(k1 + IP1), fp3 = f_at(k11 + IP1);
            const float fm0 = f_at(k2 + IM), fm1 = f_at(k21 + IM), fm2 = f_at(k2 + IM1), fm3 = f_at(k21 + IM1);
            const float fc = f_at(k + IC);
            __builtin_amdgcn_sched_barrier(0);
            const float sap = ca.y * fp0 + ca.z * fp1 + ca.w * fp2 + cc.x * fp3;
            const float sam = cc.y * fm0 + cc.z * fm1 + cc.w * fm2 + ce.x * fm3;
            const float fij = fc * ca.x;
            const float ftemp = ce.y * ftemp_scale;
            float fad1 = fij * (sap + sam);
            const float fad2 = fad1 - (2.f * sap) * sam;
            fad1 = fad1 + fad2;
            const float fcen = ftemp * fij;
            const float qx = fad2 * fcen, qy = fad1 * ftemp, qz = ((fij - 2.f * sap) * t.DAL2) * fcen;
            const float qw = ((fij - 2.f * sam) * t.DAL1) * fcen;
            dia_products<SP_TSTRIDE_DEV>(U + ql, wa, wb, wc, wd, qx, qy, qz, qw);
        }
        __syncthreads();
        // every bin adds its updates from the chunk, in the reference's order
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int np = rnp[j];
            if (np == 0) continue;
            const uint32_t *e = reinterpret_cast<const uint32_t *>(t.SP_E + roff[j]) + lane;
            float a = sl[j], b = fl[j];
            uint32_t cur[G];
#pragma unroll
            for (int u = 0; u < G; u++) cur[u] = pv[j][u];
            for (int i = G; i < np; i += G) {          // np is a whole number of groups
                // the next group's entries are loaded before this group's LDS reads
                uint32_t nxt[G];
#pragma unroll
                for (int u = 0; u < G; u++) nxt[u] = e[(i + u) * 64];
                gather_group(Ub, cur, a, b);
#pragma unroll
                for (int u = 0; u < G; u++) cur[u] = nxt[u];
            }
            gather_group(Ub, cur, a, b);
            sl[j] = a;
            fl[j] = b;
        }
        __syncthreads();
    }
}

// Scratch floats: SNONLIN's quadruplet terms (float4 per quadruplet of a
// chunk) and weight table (float4 per MC and update type), or SINPUT_ST6's
// three per-bin arrays.
__host__ __device__ __forceinline__ int scr_floats(int KL, int ML, int CHUNK, int iphys)
{
    const int snl = 4 * 512 + 4 * 9 * CHUNK, st6 = iphys == 2 ? 3 * KL * ML : 0;
    return snl > st6 ? snl : st6;
}

}  // namespace wam
