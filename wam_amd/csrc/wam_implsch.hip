// wam_implsch.hip -- IMPLSCH (wam_source_module.f90:374-700) on gfx950.
//
// One workgroup (4 waves) advances one sea point at a time in a persistent
// loop.  The point's spectrum F, the source SL, its diagonal derivative FL
// and the positive input SPOS live in LDS next to the point-invariant tables
// (loaded once per workgroup).  Element-wise stages run over the spectral
// bins; the per-frequency direction sums run one thread per frequency; the
// independent scalar stages run on different waves at the same time (AIRSEA
// on wave 1 while wave 0 integrates the spectrum, TAU_PHI_HF's 19 nodes in
// parallel) and only the ordered sums stay serial, reading LDS.  Every sum is
// evaluated in the reference's sequential order and no multiply-add is
// contracted, so results are bit-identical to a sequential REAL*4 evaluation
// with the same elementary functions (wam_math.h).
//
// SNONLIN's scatter of each DIA quadruplet into nine bins (:2890-3043) is a
// gather here: the quadruplet terms of a chunk of frequencies are computed in
// parallel into LDS, then every bin adds its contributions from that chunk in
// the order of the reference's sequential scatter (host-built lists).
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
    const int mij = wam_nint(wam_log10f(fpm4) * t.INV_LOG_CO) + 1;
    return min(max(1, mij), t.ML);
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
template <int SHALLOW>
__device__ __forceinline__ void stresso(const DevTables &t, const Smem &s, PointScalars &P, const float *cu, const float *tfk,
                        float *pm, bool want_tauw)
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
__constant__ int SLOT_SL[9] = {0, 20, 19, 17, 18, 8, 7, 5, 6};
__constant__ int SLOT_FL[9] = {0, 23, 24, 21, 22, 11, 12, 9, 10};

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
            gather_group(Ub, pv[j], a, b);
            for (int i = G; i < np; i += G) {          // np is a whole number of groups
                uint32_t v[G];
#pragma unroll
                for (int u = 0; u < G; u++) v[u] = e[(i + u) * 64];
                gather_group(Ub, v, a, b);
            }
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

template <int IPHYS, int SHALLOW>
__global__ void __launch_bounds__(NT) implsch_kernel(DevTables t, ImplArgs a)
{
    extern __shared__ __align__(16) float lds_raw[];
    __shared__ PointScalars P;
    __shared__ float T[MAXF], T2[MAXF], cu[MAXF], sprd[MAXF], t2[MAXF], flm[MAXF], tfk[MAXF], tcg[MAXF];
    __shared__ float wn[MAXF], cgv[MAXF], pm[3 * MAXF + 64];
    __shared__ SinFreq q[MAXF];
    __shared__ float pf[5 * MAXF];
    const int KL = t.KL, ML = t.ML, NB = KL * ML, NS = t.NSEA, tid = threadIdx.x;
    const int MH = t.MLSTHG;

    // carve the dynamic LDS; the scratch area comes first (16-byte aligned)
    Smem s;
    float *p = lds_raw;
    s.SCR = p;
    p += scr_floats(KL, ML, t.CHUNK, IPHYS);
    s.fr = p; p += ML; s.dfim = p; p += ML; s.dfimofr = p; p += ML; s.dfim_fr = p; p += ML;
    s.rhowg = p; p += ML; s.fr5 = p; p += ML; s.frm5 = p; p += ML;
    s.th = p; p += KL; s.costh = p; p += KL; s.sinth = p; p += KL;
    s.rnl = p; p += 25 * MH; s.af11 = p; p += MH;
    s.sig10 = p; p += (IPHYS == 2 ? t.NK10 : 0); s.dsii10 = p; p += (IPHYS == 2 ? t.NK10 : 0);
    s.inl = (int *)p; p += 5 * MH; s.kw = (int *)p; p += 8 * KL;
    s.F = p; p += NB; s.SL = p; p += NB; s.FL = p; p += NB; s.SPOS = p; p += NB;
    s.SMIN = nullptr; s.SLBOT = nullptr;
    if (a.lcflx) { s.SMIN = p; p += NB; }
    if (SHALLOW) { s.SLBOT = p; p += NB; }
    s.LLWS = (unsigned char *)p;

    // point-invariant tables, once per workgroup
    for (int i = tid; i < ML; i += NT) {
        s.fr[i] = t.FR[i]; s.dfim[i] = t.DFIM[i]; s.dfimofr[i] = t.DFIMOFR[i]; s.dfim_fr[i] = t.DFIM_FR[i];
        s.rhowg[i] = t.RHOWG_DFIM[i]; s.fr5[i] = t.FR5[i]; s.frm5[i] = t.FRM5[i];
    }
    for (int i = tid; i < KL; i += NT) { s.th[i] = t.TH[i]; s.costh[i] = t.COSTH[i]; s.sinth[i] = t.SINTH[i]; }
    for (int i = tid; i < 25 * MH; i += NT) s.rnl[i] = t.RNLCOEF[i];
    for (int i = tid; i < 5 * MH; i += NT) s.inl[i] = t.INLCOEF[i];
    for (int i = tid; i < MH; i += NT) s.af11[i] = t.AF11[i];
    for (int i = tid; i < 2 * KL; i += NT) {
        s.kw[i] = t.K1W[i]; s.kw[2 * KL + i] = t.K2W[i]; s.kw[4 * KL + i] = t.K11W[i]; s.kw[6 * KL + i] = t.K21W[i];
    }
    if (IPHYS == 2)
        for (int i = tid; i < t.NK10; i += NT) { s.sig10[i] = t.SIG10[i]; s.dsii10[i] = t.DSII10[i]; }

    // SNONLIN quadruplet slots of this thread (MC offset, KH*KL + K); MC
    // offset 255 marks a slot past the chunk
    int qslot[2];
    for (int i = 0; i < 2; i++) {
        const int ql = tid + i * NT, mcl = ql / (2 * KL);
        qslot[i] = ql < 511 && mcl < t.CHUNK ? mcl | (ql - mcl * 2 * KL) << 8 : 255;
    }

    // XCD-aware persistent order: the workgroups resident on one XCD walk one
    // contiguous range of sea points side by side, so the bin-major loads of
    // neighbouring points share L2 lines.
    const int nblk = gridDim.x, xcd = blockIdx.x % 8, lb = blockIdx.x / 8;
    const int nb_x = (nblk + 7 - xcd) / 8;
    const int per = (NS + 7) / 8;
    const int p0 = xcd * per, p1 = min(NS, p0 + per);
    const float delt = t.DELT;
    for (int pt = p0 + lb; pt < p1; pt += nb_x) {
        // 0. load the point
        for (int b = tid; b < NB; b += NT) s.F[b] = a.fl3[(size_t)b * NS + pt];
        if (tid == 0) {
            P.u10 = a.u10[pt]; P.udir = a.udir[pt]; P.tauw = a.tauw[pt]; P.z0 = a.z0[pt];
            P.roairn = a.roairn[pt]; P.wstar = a.wstar[pt]; P.depth = t.DEPTH[pt];
            P.indep = SHALLOW ? t.INDEP[pt] : 1;
            P.ju = min(t.JUMAX, max(wam_nint(P.u10 / t.DELU), 1));
        }
        __syncthreads();
        if (tid < KL) {
            const float c = wam_cosf(s.th[tid] - P.udir);
            cu[tid] = c;
            sprd[tid] = wam_ipowf(wam_maxf(0.f, c), 2);
        }
        if (tid >= 64 && tid - 64 < ML) {
            const int m = tid - 64;
            flm[m] = t.FLMINFR[(size_t)m * t.JUMAX + P.ju - 1];
            if (SHALLOW) { tfk[m] = t.TFAK[tb(t, m, P.indep)]; tcg[m] = t.TCGOND[tb(t, m, P.indep)]; }
            if (IPHYS == 2) { wn[m] = t.WN[(size_t)m * NS + pt]; cgv[m] = t.CGG[(size_t)m * NS + pt]; }
        }
        if (tid == 128) {                       // wind stress, overlapped with the direction sums
            if (IPHYS == 2) w3flx4(t, P.u10, P.ust, P.z0);
            else airsea(t, P.u10, P.tauw, P.ust, P.z0);
        }
        dirsum(t, s.F, nullptr, T, tid);
        __syncthreads();
        // 2. mean parameters (the integrals run on separate waves)
        if (tid == 0) P.s_e = integ(ML, t.MO_TAIL, T, s.dfim);
        else if (tid == 64) P.s_fm = integ(ML, t.MM1_TAIL, T, s.dfimofr);
        else if (tid == 192) P.s_tm = integ(ML, t.MP1_TAIL, T, s.dfim_fr);
        else if (SHALLOW && tid == 128) {
            // the two wave-number integrals of the mean-parameter pass, as
            // two independent chains in one loop
            const float d1 = sqrtf(t.G) / t.ZPI, d2 = t.ZPI / sqrtf(t.G);
            float w1 = (t.MM1_TAIL * d1) * T[ML - 1];
            float w2 = (t.MP1_TAIL * d2) * T[ML - 1];
            for (int m = 0; m < ML; m++) {
                const float sk = sqrtf(tfk[m]);
                w1 = w1 + (T[m] / sk) * s.dfim[m];
                w2 = w2 + (T[m] * sk) * s.dfim[m];
            }
            P.s_w1 = w1;
            P.s_w2 = w2;
        }
        __syncthreads();
        if (tid == 0) {
            const float e = wam_maxf(P.s_e, t.EMIN);
            P.emean = e;
            P.fmean = e / wam_maxf(P.s_fm, t.EMIN);
            P.f1mean = 1.f / (e > t.EMIN ? e / P.s_tm : 1.f);
            if (SHALLOW) {
                P.akmean = e > t.EMIN ? wam_ipowf(e / P.s_w1, 2) : 1.f;
                P.xkmean = e > t.EMIN ? wam_ipowf(P.s_w2 / e, 2) : 1.f;
            }
        }
        __syncthreads();
        // 3. input source
        if (IPHYS == 2) sinput_st6(t, s, P, pf, wn, cgv);
        else sinput<SHALLOW>(t, s, P, cu, tfk, q);
        dirsum(t, s.F, s.LLWS, T, tid);
        __syncthreads();
        if (tid == 0) P.s_e = integ(ML, t.MO_TAIL, T, s.dfim);
        else if (tid == 64) P.s_fm = integ(ML, t.MM1_TAIL, T, s.dfimofr);
        __syncthreads();
        if (tid == 0) {
            const float e = wam_maxf(P.s_e, t.EMIN);
            P.emeanws = e;
            P.fmeanws = e / wam_maxf(P.s_fm, t.EMIN);
            P.mij = frcut(t, s.fr, P.fmean, P.fmeanws, P.ust);
        }
        __syncthreads();
        stresso<SHALLOW>(t, s, P, cu, tfk, pm, IPHYS != 2);
        // wind stress update (thread 0) and the diagnostic tail (waves 1-3) in parallel
        if (tid == 0) {
            if (IPHYS == 2) w3flx4(t, P.u10, P.ust, P.z0);
            else airsea(t, P.u10, P.tauw, P.ust, P.z0);
        }
        imphftail<SHALLOW>(t, s, P.mij, tfk, tcg, t2, 64);
        __syncthreads();
        if (tid >= 64) {
            const int mij = P.mij;
            for (int b = tid - 64 + KL * mij; b < NB; b += NT - 64) {
                const int m = b / KL, k = b - KL * m;
                s.F[b] = t2[m] * s.F[k + KL * (mij - 1)];
            }
        }
        __syncthreads();
        if (IPHYS != 2) sinput<SHALLOW>(t, s, P, cu, tfk, q);
        if (a.lcflx)
            for (int b = tid; b < NB; b += NT) s.SMIN[b] = s.SL[b] - s.SPOS[b];
        stresso<SHALLOW>(t, s, P, cu, tfk, pm, IPHYS != 2);
        // dissipation
        if (IPHYS == 2) {
            // SDISSIP_ST6 (:4929-5102) and W3SWL6 (:5106-5252)
            // per frequency on wave 0 (lane m); the cumulative saturation sum,
            // the peak search and the B1 integral read the lanes in order
            float *t12 = pm, *ddis = pm + MAXF;
            if (tid < 64) {
                const int m = tid;
                const bool on = m < ML;
                float ab = 0.f, bnm = 0.f, t2m = 0.f, term = 0.f, t1 = 0.f, sigm = 0.f;
                if (on) {
                    const float *f = s.F + KL * m;
                    float ed = f[0];
                    for (int k = 1; k < KL; k++) ed = ed + f[k];
                    const float edens = ed * t.DELTH;
                    sigm = t.ZPI * s.fr[m];
                    const float c0 = t.ZPI * sigm;
                    float mx = 0.f;
                    for (int k = 0; k < KL; k++) {
                        const float v = (f[k] * cgv[m]) / c0;
                        ab = k == 0 ? v : ab + v;
                        mx = k == 0 ? v : wam_maxf(mx, v);
                    }
                    float sn = 0.f;
                    for (int k = 0; k < KL; k++) {
                        const float v = (f[k] * cgv[m]) / c0;
                        const float kn = (mx < 1.0E-34f) ? 1.f : v / mx;
                        sn = k == 0 ? kn : sn + kn;
                    }
                    const float anar = 1.0f / (sn * t.DELTH);
                    bnm = (anar * ((ab * sigm) * t.DELTH)) * wam_ipowf(wn[m], 3);
                    const float dden = (t.ZPI * s.dfim[m]) * sigm;
                    t2m = (ab * dden) / cgv[m];
                    // W3SWL6 terms of frequency m
                    const float bnt = 0.035f * 0.035f;
                    const float xfac = (1.0f - 1.0f / t.CO) / (t.CO - 1.0f / t.CO);
                    const float et = (t.ZPI * bnt) / ((1.0f * cgv[m]) * wam_ipowf(wn[m], 3));
                    const float ex = wam_maxf(0.0f, edens - et);
                    const float nex = ex / et;
                    const float n4 = wam_ipowf(nex, 4);
                    t1 = ((4.75E-6f * 1.0f) * s.fr[m]) * n4;
                    float dfii = t.DF[m];
                    if (m > 0 && m < ML - 1) dfii = dfii * xfac;
                    term = (1.0f * n4) * dfii;
                }
                // running sum of the terms up to m, the first peak of ABAND, SUM of T2
                float run = lane_at(term, 0), pre = run;
                float cur = lane_at(ab, 0), sm = lane_at(t2m, 0);
                int ikx = 0;
                for (int i0 = 0; i0 < ML; i0 += 8) {
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int i = i0 + u;
                        if (i == 0) continue;
                        const float r = run + lane_at(term, i);
                        run = i < ML ? r : run;
                        pre = m == i ? run : pre;
                        const float v = lane_at(ab, i);
                        const bool take = i < ML && v > cur;
                        cur = take ? v : cur;
                        ikx = take ? i : ikx;
                        const float q = sm + lane_at(t2m, i);
                        sm = i < ML ? q : sm;
                    }
                }
                const float b1 = 0.0041f * ((2.f * sqrtf(sm)) * wn[ikx]);
                if (on) {
                    t12[m] = -1.0f * (wam_maxf(0.0f, t1) + wam_maxf(0.0f, 7.00E-5f * pre));
                    ddis[m] = ab > 1.E-30f ? (((-(2.f / 3.f)) * b1) * sigm) * sqrtf(bnm) : 0.f;
                }
            }
            __syncthreads();
            for (int b = tid; b < NB; b += NT) {
                const int m = b / KL;
                const float f = s.F[b];
                const float sl = s.SL[b] + t12[m] * f, fl = s.FL[b] + t12[m];
                s.SL[b] = sl + ddis[m] * f;
                s.FL[b] = fl + ddis[m];
            }
            __syncthreads();
        } else {
            // SDISSIP (:1593-1700) with FMEAN = F1MEAN, AKMEAN = XKMEAN
            if (tid < ML) {
                const int m = tid;
                float sds, x;
                const float e2 = P.emean * P.emean;
                if (SHALLOW) {
                    sds = ((t.CONSS * P.f1mean) * e2) * wam_ipowf(P.xkmean, 4);
                    x = tfk[m] / P.xkmean;
                } else {
                    sds = (t.CONSD * e2) * wam_ipowf(P.f1mean, 9);
                    x = wam_ipowf(s.fr[m] / P.f1mean, 2);
                }
                t2[m] = sds * (0.5f * x + 0.5f * wam_ipowf(x, 2));
            }
            if (SHALLOW && tid == 64) {
                const float enh = wam_maxf((0.75f * P.depth) * P.akmean, .5f);
                P.enhfr = 1.f + ((5.5f / enh) * (1.f - .833f * enh)) * wam_expf(-1.25f * enh);
            }
            __syncthreads();
            for (int b = tid; b < NB; b += NT) {
                const int m = b / KL;
                s.SL[b] = s.SL[b] + t2[m] * s.F[b];
                s.FL[b] = s.FL[b] + t2[m];
            }
            __syncthreads();
        }
        if (IPHYS == 2 && SHALLOW) {
            if (tid == 0) {
                const float enh = wam_maxf((0.75f * P.depth) * P.akmean, .5f);
                P.enhfr = 1.f + ((5.5f / enh) * (1.f - .833f * enh)) * wam_expf(-1.25f * enh);
            }
            __syncthreads();
        }
        // SNONLIN, with the ISNONLIN = 0 shallow-water enhancement
        snonlin(t, s, SHALLOW ? P.enhfr : 1.f, qslot);
        if (SHALLOW) {
            // SBOTTOM (:1527-1589) into SL_BOT
            if (tid < ML) {
                const float wav = tfk[tid];
                const float sbo = wam_minf((2.f * P.depth) * wav, 50.f);
                t2[tid] = (((-2.0f * 0.038f) / t.G) * wav) / wam_sinhf(sbo);
            }
            __syncthreads();
            for (int b = tid; b < NB; b += NT) {
                const int m = b / KL;
                s.SLBOT[b] = 0.f + t2[m] * s.F[b];
                s.FL[b] = s.FL[b] + t2[m];
            }
            __syncthreads();
        }
        if (a.lcflx) {
            // WNFLUXES (:3970-4054) on SSOURCE = SL - SMIN
            for (int b = tid; b < NB; b += NT) s.SMIN[b] = s.SL[b] - s.SMIN[b];
            __syncthreads();
            if (tid == 0 && a.fluxes)
                wnfluxes<SHALLOW>(t, s, P.phiaw, P.ust, wam_sinf(P.udir), wam_cosf(P.udir), P.mij, tfk, a.fluxes + pt,
                                  a.NMP);
        }
        // 4. limiter and implicit update
        if (tid < ML) {
            const float delfl = (t.C5G / wam_ipowf(s.fr[tid], 4)) * delt;
            t2[tid] = (P.ust * delfl) * wam_maxf(P.fmeanws, P.fmean);
        }
        __syncthreads();
        for (int b = tid; b < NB; b += NT) {
            const int m = b / KL, k = b - KL * m;
            float sl = s.SL[b];
            if (SHALLOW) sl = sl + s.SLBOT[b];
            const float temp1 = (delt * sl) / wam_maxf(1.f, 1.f - delt * s.FL[b]);
            const float temp3 = wam_minf(wam_fabsf(temp1), t2[m]);
            const float f = s.F[b] + wam_signf(temp3, temp1);
            s.F[b] = wam_maxf(f, flm[m] * sprd[k]);
        }
        __syncthreads();
        // 5. new mean parameters, cut-off index and tail (both direction sums at once)
        dirsum(t, s.F, nullptr, T, tid);
        dirsum(t, s.F, s.LLWS, T2, tid - 64);
        __syncthreads();
        if (tid == 0) P.s_e = integ(ML, t.MO_TAIL, T, s.dfim);
        else if (tid == 64) P.s_fm = integ(ML, t.MM1_TAIL, T, s.dfimofr);
        else if (tid == 128) P.s_w1 = integ(ML, t.MO_TAIL, T2, s.dfim);
        else if (tid == 192) P.s_w2 = integ(ML, t.MM1_TAIL, T2, s.dfimofr);
        __syncthreads();
        if (tid == 0) {
            const float e = wam_maxf(P.s_e, t.EMIN);
            P.emean = e;
            P.fmean = e / wam_maxf(P.s_fm, t.EMIN);
            const float ew = wam_maxf(P.s_w1, t.EMIN);
            P.emeanws = ew;
            P.fmeanws = ew / wam_maxf(P.s_w2, t.EMIN);
            P.mij = frcut(t, s.fr, P.fmean, P.fmeanws, P.ust);
        }
        __syncthreads();
        imphftail<SHALLOW>(t, s, P.mij, tfk, tcg, t2, 0);
        __syncthreads();
        {
            const int mij = P.mij;
            for (int b = tid; b < NB; b += NT) {
                const int m = b / KL, k = b - KL * m;
                const float f = m >= mij ? t2[m] * s.F[k + KL * (mij - 1)] : s.F[b];
                a.fl3[(size_t)b * NS + pt] = f;
            }
        }
        if (tid == 0) {
            a.tauw[pt] = P.tauw; a.ustar[pt] = P.ust; a.z0[pt] = P.z0;
            if (a.phiaw) a.phiaw[pt] = P.phiaw;
            if (a.mij) a.mij[pt] = P.mij;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// IMPLSCH for IPHYS = 0 in three kernels.  The scalar chain of a sea point
// (AIRSEA's Newton iteration, the ordered frequency integrals, STRESSO's
// sums, TAU_PHI_HF, FRCUTINDEX) is sequential within the point but
// independent across points, so the two per-lane kernels give every sea
// point its own lane (PW points per wavefront; FL3's point-fastest rows load
// and store coalesced) and run each stage at full width in the reference's
// order.  SNONLIN's DIA couples the bins of one point and stays one
// workgroup per point in between, on point-major copies of F, SL and FL.
//   implsch_pre_kernel:  AIRSEA, mean parameters, SINPUT, FRCUTINDEX,
//                        STRESSO, AIRSEA, IMPHFTAIL, SINPUT, STRESSO, SDISSIP
//   implsch_snl_kernel:  SNONLIN, SBOTTOM, WNFLUXES, limiter and update
//   implsch_post_kernel: mean parameters, FRCUTINDEX, IMPHFTAIL, FL3 out
// ---------------------------------------------------------------------------

constexpr int PW = 256;   // most sea points per workgroup of the per-lane kernels (the launch's width, 64-256)

// STRESSO's closing part (:3164-3356) for one point: the direction sums phi,
// xs, ys over M < MIJ, the tail moments h1, h2 of F(., MIJ) and TAU_PHI_HF in
// its sequential form.  Returns TAUW; PHIAW in phiaw.
__device__ __forceinline__ float stresso_close(const DevTables &t, int mij, float ust, float z0, float udir, float phi,
                                               float xs, float ys, float h1, float h2, float &phiaw)
{
    const float gm1 = 1.0f / t.G;
    const float cst = (t.DELTH * t.ZPI4) * gm1;
    const float c1 = (cst * t.FR5[mij - 1]) * gm1;
    const float c2 = (t.ROAIR * cst) * t.FR5[mij - 1];
    xs = xs / wam_maxf(t.ROAIR, 1.f);
    ys = ys / wam_maxf(t.ROAIR, 1.f);
    float tau1, phi1;
    tau_phi_hf_seq(t, t.FR, mij, ust, z0, tau1, phi1);
    const float tauhf = (c1 * h1) * tau1, phihf = (c2 * h2) * phi1;
    phi = phi + phihf;
    xs = xs + tauhf * wam_sinf(udir);
    ys = ys + tauhf * wam_cosf(udir);
    float tw = sqrtf(xs * xs + ys * ys);
    tw = wam_minf(tw, ust * ust - t.EPS1);
    tw = wam_maxf(tw, 0.f);
    phiaw = phi;
    return tw;
}

// STRESSO's weight of frequency m: RHOWG_DFIM, halved at MIJ below ML.
__device__ __forceinline__ float stresso_rw(const DevTables &t, int m, int mij)
{
    float rw = t.RHOWG_DFIM[m];
    if (m == mij - 1 && mij != t.ML) rw = 0.5f * rw;
    return rw;
}

template <int SHALLOW>
__device__ __forceinline__ float stresso_cm(const DevTables &t, int m, float tfk_m)
{
    return SHALLOW ? tfk_m * (1.0f / (t.ZPI * t.FR[m])) : (t.ZPI * t.FR[m]) * (1.0f / t.G);
}

template <int SHALLOW>
__global__ void __launch_bounds__(PW) __attribute__((amdgpu_waves_per_eu(4))) implsch_pre_kernel(DevTables t, ImplArgs a, ImplScratch x)
// (workgroups of 64, 128 or 256 lanes, one sea point each)
{
    extern __shared__ float cu_lds[];            // [KL][W]: COS(TH(K) - UDIR) of each lane's point
    const int KL = t.KL, ML = t.ML, NB = KL * ML, NS = t.NSEA, NB4 = x.NB4, XS = x.NX, lane = threadIdx.x;
    // lp: the point's row in the work arrays (its place in the chunk), pt: the sea point
    const int W = blockDim.x;
    const int lp0 = blockIdx.x * W + lane;
    const bool on = lp0 < x.NP;
    const int lp = on ? lp0 : x.NP - 1, mp = x.P0 + lp;      // member-point (mp = member * NSEA + sea point)
    const int mem = mp / NS, pt = mp - mem * NS;            // ensemble member, sea point
    float *cu = cu_lds + lane;
    const float *F = a.fl3 + (size_t)mem * NB * NS + pt;   // F(b) = F[b * NS]
    const float u10 = a.u10[mp], udir = a.udir[mp], roairn = a.roairn[mp], wstar = a.wstar[mp];
    const int indep = SHALLOW ? t.INDEP[pt] : 1;
    float tauw = a.tauw[mp], ust = 0.f, z0 = a.z0[mp];
    for (int k = 0; k < KL; k++) {
        const float c = wam_cosf(t.TH[k] - udir);
        cu[k * W] = c;
        // the limiter's floor FLMINFR(JU, M) * SPRD(K) (:494-498) is applied by implsch_snl_kernel
        if (on) x.sprd[(size_t)k * XS + lp] = wam_ipowf(wam_maxf(0.f, c), 2);
    }
    airsea(t, u10, tauw, ust, z0);

    const float const3 = ((2.f * t.XKAPPA) / t.BETAMAX) * (t.XKAPPA * t.XKAPPA);
    const float d1 = sqrtf(t.G) / t.ZPI, d2 = t.ZPI / sqrtf(t.G);
    // 1. one pass over the spectrum: the mean parameters' integrals, the
    //    first SINPUT's LLWS-masked integrals and STRESSO's direction sums.
    //    Frequency ML goes first, since it seeds every integral; its own
    //    terms are added last.  Bins go four at a time (KL % 4 == 0).
    float s_e = 0.f, s_fm = 0.f, s_tm = 0.f, s_w1 = 0.f, s_w2 = 0.f, q_e = 0.f, q_fm = 0.f, Tl = 0.f, Ql = 0.f;
    {
        const float sig_n = wsigstar(t, ust, z0, wstar);
        const float usn1 = ust * (1.0f - sig_n), usn2 = ust * (1.0f + sig_n);
        for (int i = 0; i < ML; i++) {
            const int m = i == 0 ? ML - 1 : i - 1;
            const float tfk = SHALLOW ? t.TFAK[tb(t, m, indep)] : 0.f;
            const SinFreq r = sin_freq<SHALLOW>(t, t.FR[m], tfk, usn1, usn2, z0, roairn);
            float T = 0.f, Q = 0.f, sa = 0.f, sx = 0.f, sy = 0.f;
            bool qfirst = true;
            for (int k0 = 0; k0 < KL; k0 += 4) {
                float f4[4];
#pragma unroll
                for (int u = 0; u < 4; u++) f4[u] = F[(size_t)(k0 + u + KL * m) * NS];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = k0 + u;
                    const float f = f4[u];
                    T = k == 0 ? f : T + f;
                    float sp, fl;
                    bool ll;
                    sin_bin(t, r, cu[k * W], const3, sp, fl, ll);
                    if (ll) {
                        Q = qfirst ? f : Q + f;
                        qfirst = false;
                    }
                    const float spp = wam_maxf(sp * f, 0.f);
                    if (k == 0) { sa = spp; sx = spp * t.SINTH[k]; sy = spp * t.COSTH[k]; }
                    else { sa = sa + spp; sx = sx + spp * t.SINTH[k]; sy = sy + spp * t.COSTH[k]; }
                }
            }
            if (on) {
                float *axy = x.axy + (size_t)m * XS + lp;
                axy[0] = sa;
                axy[(size_t)ML * XS] = sx;
                axy[(size_t)2 * ML * XS] = sy;
            }
            if (i == 0) {
                Tl = T; Ql = Q;
                s_e = t.MO_TAIL * T; s_fm = t.MM1_TAIL * T; s_tm = t.MP1_TAIL * T;
                if (SHALLOW) { s_w1 = (t.MM1_TAIL * d1) * T; s_w2 = (t.MP1_TAIL * d2) * T; }
                q_e = t.MO_TAIL * Q; q_fm = t.MM1_TAIL * Q;
            } else {
                s_e = s_e + T * t.DFIM[m]; s_fm = s_fm + T * t.DFIMOFR[m]; s_tm = s_tm + T * t.DFIM_FR[m];
                if (SHALLOW) {
                    const float sk = sqrtf(tfk);
                    s_w1 = s_w1 + (T / sk) * t.DFIM[m];
                    s_w2 = s_w2 + (T * sk) * t.DFIM[m];
                }
                q_e = q_e + Q * t.DFIM[m]; q_fm = q_fm + Q * t.DFIMOFR[m];
            }
        }
        const int m = ML - 1;
        s_e = s_e + Tl * t.DFIM[m]; s_fm = s_fm + Tl * t.DFIMOFR[m]; s_tm = s_tm + Tl * t.DFIM_FR[m];
        if (SHALLOW) {
            const float sk = sqrtf(t.TFAK[tb(t, m, indep)]);
            s_w1 = s_w1 + (Tl / sk) * t.DFIM[m];
            s_w2 = s_w2 + (Tl * sk) * t.DFIM[m];
        }
        q_e = q_e + Ql * t.DFIM[m]; q_fm = q_fm + Ql * t.DFIMOFR[m];
    }
    const float emean = wam_maxf(s_e, t.EMIN);
    const float fmean = emean / wam_maxf(s_fm, t.EMIN);
    const float f1mean = 1.f / (emean > t.EMIN ? emean / s_tm : 1.f);
    float akmean = 1.f, xkmean = 1.f;
    if (SHALLOW) {
        akmean = emean > t.EMIN ? wam_ipowf(emean / s_w1, 2) : 1.f;
        xkmean = emean > t.EMIN ? wam_ipowf(s_w2 / emean, 2) : 1.f;
    }
    const float ews = wam_maxf(q_e, t.EMIN);
    const float fmeanws = ews / wam_maxf(q_fm, t.EMIN);
    const int mij = frcut(t, t.FR, fmean, fmeanws, ust);

    // 2. STRESSO on the first SINPUT.  The moments of F(., MIJ) serve both
    //    STRESSO calls: IMPHFTAIL changes the frequencies above MIJ only.
    float h1 = 0.f, h2 = 0.f;
    for (int k = 0; k < KL; k++) {
        const float cosw = wam_maxf(cu[k * W], 0.f);
        const float f = F[(size_t)(k + KL * (mij - 1)) * NS];
        if (k == 0) { h2 = f * wam_ipowf(cosw, 2); h1 = h2 * cosw; }
        else { h1 = h1 + f * wam_ipowf(cosw, 3); h2 = h2 + f * wam_ipowf(cosw, 2); }
    }
    {
        float phi = 0.f, xs = 0.f, ys = 0.f;
        for (int m = 0; m < mij; m++) {
            const float *axy = x.axy + (size_t)m * XS + lp;
            const float rw = stresso_rw(t, m, mij);
            const float cmrw = stresso_cm<SHALLOW>(t, m, SHALLOW ? t.TFAK[tb(t, m, indep)] : 0.f) * rw;
            phi = phi + axy[0] * rw;
            xs = xs + axy[(size_t)ML * XS] * cmrw;
            ys = ys + axy[(size_t)2 * ML * XS] * cmrw;
        }
        float phiaw;
        tauw = stresso_close(t, mij, ust, z0, udir, phi, xs, ys, h1, h2, phiaw);
    }
    airsea(t, u10, tauw, ust, z0);

    // 3. IMPHFTAIL, the second SINPUT, STRESSO's sums and SDISSIP in one pass;
    //    F, SL and FL go to the tiled work arrays, LLWS to bits.
    float tden;
    if (SHALLOW) {
        const float akm0 = 1.f / t.TFAK[tb(t, mij - 1, indep)];
        tden = wam_ipowf(akm0, 3) / t.TCGOND[tb(t, mij - 1, indep)];
    } else {
        tden = t.FRM5[mij - 1];
    }
    const float e2 = emean * emean;
    const float sds = SHALLOW ? ((t.CONSS * f1mean) * e2) * wam_ipowf(xkmean, 4) : (t.CONSD * e2) * wam_ipowf(f1mean, 9);
    float phi = 0.f, xs = 0.f, ys = 0.f;
    {
        const float sig_n = wsigstar(t, ust, z0, wstar);
        const float usn1 = ust * (1.0f - sig_n), usn2 = ust * (1.0f + sig_n);
        const size_t q0 = quad_offset(lp, 0, NB4);
        unsigned *LLW = x.llws + lp;
        unsigned word = 0;
        for (int m = 0; m < ML; m++) {
            const float tfk = SHALLOW ? t.TFAK[tb(t, m, indep)] : 0.f;
            const SinFreq r = sin_freq<SHALLOW>(t, t.FR[m], tfk, usn1, usn2, z0, roairn);
            const bool tail = m >= mij;
            float ttail = 0.f;
            if (tail) {
                const float an = SHALLOW ? wam_ipowf(1.f / tfk, 3) / t.TCGOND[tb(t, m, indep)] : t.FRM5[m];
                ttail = an / tden;
            }
            const float xd = SHALLOW ? tfk / xkmean : wam_ipowf(t.FR[m] / f1mean, 2);
            const float t2d = sds * (0.5f * xd + 0.5f * wam_ipowf(xd, 2));
            const int mrow = tail ? mij - 1 : m;
            float sa = 0.f, sx = 0.f, sy = 0.f;
            for (int k0 = 0; k0 < KL; k0 += 4) {
                float f4[4], sl4[4], fl4[4], sm4[4];
#pragma unroll
                for (int u = 0; u < 4; u++) f4[u] = F[(size_t)(k0 + u + KL * mrow) * NS];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = k0 + u, b = k + KL * m;
                    const float f = tail ? ttail * f4[u] : f4[u];
                    float sp, fl;
                    bool ll;
                    sin_bin(t, r, cu[k * W], const3, sp, fl, ll);
                    const float sl = fl * f, spos = sp * f;
                    f4[u] = f;
                    sl4[u] = sl + t2d * f;
                    fl4[u] = fl + t2d;
                    sm4[u] = sl - spos;
                    word |= (unsigned)ll << (b & 31);
                    if ((b & 31) == 31 || b == NB - 1) {
                        if (on) LLW[(size_t)(b >> 5) * XS] = word;
                        word = 0;
                    }
                    const float spp = wam_maxf(spos, 0.f);
                    if (k == 0) { sa = spp; sx = spp * t.SINTH[k]; sy = spp * t.COSTH[k]; }
                    else { sa = sa + spp; sx = sx + spp * t.SINTH[k]; sy = sy + spp * t.COSTH[k]; }
                }
                if (on) {
                    const size_t o = q0 + (size_t)((k0 + KL * m) >> 2) * 256;
                    *reinterpret_cast<float4 *>(x.ft + o) = make_float4(f4[0], f4[1], f4[2], f4[3]);
                    *reinterpret_cast<float4 *>(x.slt + o) = make_float4(sl4[0], sl4[1], sl4[2], sl4[3]);
                    *reinterpret_cast<float4 *>(x.flt + o) = make_float4(fl4[0], fl4[1], fl4[2], fl4[3]);
                    if (a.lcflx)
                        *reinterpret_cast<float4 *>(x.smt + o) = make_float4(sm4[0], sm4[1], sm4[2], sm4[3]);
                }
            }
            if (m < mij) {
                const float rw = stresso_rw(t, m, mij);
                const float cmrw = stresso_cm<SHALLOW>(t, m, tfk) * rw;
                phi = phi + sa * rw;
                xs = xs + sx * cmrw;
                ys = ys + sy * cmrw;
            }
        }
    }
    float phiaw;
    tauw = stresso_close(t, mij, ust, z0, udir, phi, xs, ys, h1, h2, phiaw);
    if (on) {
        a.tauw[mp] = tauw;
        a.ustar[mp] = ust;
        a.z0[mp] = z0;
        if (a.phiaw) a.phiaw[mp] = phiaw;
        float enhfr = 1.f;
        if (SHALLOW) {
            const float enh = wam_maxf((0.75f * t.DEPTH[pt]) * akmean, .5f);
            enhfr = 1.f + ((5.5f / enh) * (1.f - .833f * enh)) * wam_expf(-1.25f * enh);
        }
        x.pp[lp] = enhfr;
        x.pp[(size_t)XS + lp] = wam_maxf(fmeanws, fmean);
        x.pp[(size_t)2 * XS + lp] = (float)mij;
        x.pp[(size_t)3 * XS + lp] = phiaw;
        if (a.lcflx) {   // for WNFLUXES in implsch_snl_kernel
            x.pp[(size_t)4 * XS + lp] = wam_sinf(udir);
            x.pp[(size_t)5 * XS + lp] = wam_cosf(udir);
        }
    }
    // implsch_snl_kernel's point tickets start from zero
    if (blockIdx.x == 0 && lane < 8) x.tick[32 * lane] = 0u;
}

constexpr int SNL_NT = 512;  // threads of implsch_snl_kernel: one quadruplet of a chunk each
constexpr int SNL_NJ = 3;    // bins per thread: KL*ML <= SNL_NT * SNL_NJ
constexpr int SNL_QT = 256 * SP_QPT_DEV / SNL_NT;

// Three workgroups (LDS-limited) of 8 waves per CU: registers for 6 waves per SIMD.
template <int SHALLOW>
__global__ void __launch_bounds__(SNL_NT) __attribute__((amdgpu_waves_per_eu(6)))
implsch_snl_kernel(DevTables t, ImplArgs a, ImplScratch x)
{
    extern __shared__ __align__(16) float lds_raw[];
    __shared__ float sprd[MAXF], flm[MAXF], tfk[MAXF], tlim[MAXF], tbot[MAXF];
    __shared__ float Ps[5];
    __shared__ int Pmij, Pnext;
    constexpr int S = 256 * SP_QPT_DEV;
    const int KL = t.KL, ML = t.ML, NB = KL * ML, NS = t.NSEA, NB4 = x.NB4, XS = x.NX, tid = threadIdx.x;
    const int MH = t.MLSTHG;

    // LDS: the products U [9][S] float2, F, (SMIN, SL_BOT with LCFLX), tables
    Smem s{};
    float *p = lds_raw;
    float2 *U = reinterpret_cast<float2 *>(p); p += 2 * 9 * SP_TSTRIDE_DEV;
    s.F = p; p += NB;
    if (a.lcflx) {
        s.SMIN = p; p += NB;
        if (SHALLOW) { s.SLBOT = p; p += NB; }
    }
    s.rnl = p; p += 36 * MH;      // per MC a 16-byte aligned block (snl_coef_block)
    s.sprw = reinterpret_cast<int *>(p); p += 2 * t.SP_NCHUNK * t.SP_NRUN;
    s.fr = p; p += ML; s.dfim = p; p += ML;
    s.th = p; p += KL; s.costh = p; p += KL; s.sinth = p; p += KL;

    for (int i = tid; i < ML; i += SNL_NT) { s.fr[i] = t.FR[i]; s.dfim[i] = t.DFIM[i]; }
    for (int i = tid; i < KL; i += SNL_NT) { s.th[i] = t.TH[i]; s.costh[i] = t.COSTH[i]; s.sinth[i] = t.SINTH[i]; }
    for (int i = tid; i < 36 * MH; i += SNL_NT) s.rnl[i] = snl_coef_block(t, i / 36, i % 36);
    for (int i = tid; i < 2 * t.SP_NCHUNK * t.SP_NRUN; i += SNL_NT) s.sprw[i] = t.SP_RW[i];
    if (tid == 0) U[S - 1] = make_float2(-0.f, -0.f);
    // this thread's quadruplet slots (MC offset, KH*KL + K); MC offset 255
    // marks a slot past the chunk
    // and its directions as byte offsets: K1W, K11W, K2W, K21W (one byte
    // each of qdir, KL <= 64) and K (qdk)
    int qslot[SNL_QT];
    unsigned qdir[SNL_QT], qdk[SNL_QT];
    for (int i = 0; i < SNL_QT; i++) {
        const int ql = tid + i * SNL_NT, mcl = ql / (2 * KL);
        const int r = ql - mcl * 2 * KL;
        const bool in = ql < S - 1 && mcl < t.SP_CHUNK;
        qslot[i] = in ? mcl | r << 8 : 255;
        qdir[i] = in ? (unsigned)(4 * t.K1W[r]) | (unsigned)(4 * t.K11W[r]) << 8 | (unsigned)(4 * t.K2W[r]) << 16 |
                           (unsigned)(4 * t.K21W[r]) << 24
                     : 0u;
        qdk[i] = in ? (unsigned)(4 * (r < KL ? r : r - KL)) : 0u;
    }

    // Points by tickets: each XCD (blockIdx % 8) owns a contiguous range of
    // sea points and hands them out in order from its counter, so the
    // workgroups of an XCD work on neighbouring points at the same time and
    // share the 128-byte lines of the tiled arrays in its L2.  The ticket for
    // the next point is taken at the start of the current one.
    const int xcd = blockIdx.x % 8;
    const int per = (x.NP + 7) / 8;                // points of the chunk, rows of the work arrays
    const int p0 = xcd * per, p1 = min(x.NP, p0 + per);
    unsigned *tick = x.tick + 32 * xcd;
    const float delt = t.DELT;
    if (tid == 0) Pnext = p0 + (int)atomicAdd(tick, 1u);
    __syncthreads();
    for (int lp = Pnext; lp < p1; lp = Pnext) {
        const int mp = x.P0 + lp, mem = mp / NS, pt = mp - mem * NS;   // member-point, member, sea point
        const size_t q0 = quad_offset(lp, 0, NB4);
        int tnext = 0;
        if (tid == 0) tnext = (int)atomicAdd(tick, 1u);
        {
            float4 *F4 = reinterpret_cast<float4 *>(s.F), *SM4 = reinterpret_cast<float4 *>(s.SMIN);
            for (int q = tid; q < NB4; q += SNL_NT) {
                const size_t o = q0 + (size_t)q * 256;
                F4[q] = *reinterpret_cast<const float4 *>(x.ft + o);
                if (a.lcflx) SM4[q] = *reinterpret_cast<const float4 *>(x.smt + o);
            }
        }
        // SL and FL of the bins tid + SNL_NT j
        float sl[SNL_NJ], fl[SNL_NJ];
#pragma unroll
        for (int j = 0; j < SNL_NJ; j++) {
            const int b = tid + SNL_NT * j;
            sl[j] = 0.f;
            fl[j] = 0.f;
            if (b < NB) {
                const size_t o = q0 + (size_t)(b >> 2) * 256 + (b & 3);
                sl[j] = x.slt[o];
                fl[j] = x.flt[o];
            }
        }
        if (tid == 0) {
            Ps[0] = a.ustar[mp]; Ps[1] = x.pp[lp]; Ps[2] = x.pp[(size_t)XS + lp]; Ps[3] = x.pp[(size_t)3 * XS + lp];
            Pmij = (int)x.pp[(size_t)2 * XS + lp];
        }
        // no elementary functions here: SPRD comes from implsch_pre_kernel and
        // SBOTTOM's factor from PREPARE_PROPAGATION's table (their double-
        // precision evaluation would make this kernel's SNONLIN loop spill)
        if (tid < KL) sprd[tid] = x.sprd[(size_t)tid * XS + lp];
        if (tid >= 64 && tid - 64 < ML) {
            const int m = tid - 64;
            const int ju = min(t.JUMAX, max(wam_nint(a.u10[mp] / t.DELU), 1));
            flm[m] = t.FLMINFR[(size_t)m * t.JUMAX + ju - 1];
            if (SHALLOW) {
                tfk[m] = t.TFAK[tb(t, m, t.INDEP[pt])];
                tbot[m] = t.SBOT[(size_t)m * NS + pt];
            }
        }
        __syncthreads();
        // SNONLIN, with the ISNONLIN = 0 shallow-water enhancement
        snonlin_products<SNL_NT, SNL_NJ>(t, s, U, SHALLOW ? Ps[1] : 1.f, qslot, qdir, qdk, sl, fl);
        // the limiter's factor (:652-690)
        if (tid < ML) {
            const float delfl = (t.C5G / wam_ipowf(s.fr[tid], 4)) * delt;
            tlim[tid] = (Ps[0] * delfl) * Ps[2];
        }
        __syncthreads();
        if (a.lcflx) {
            // WNFLUXES on SSOURCE = SL - SMIN and SL_BOT = 0 + factor * F (SBOTTOM)
#pragma unroll
            for (int j = 0; j < SNL_NJ; j++) {
                const int b = tid + SNL_NT * j;
                if (b < NB) {
                    if (SHALLOW) s.SLBOT[b] = 0.f + tbot[b / KL] * s.F[b];
                    s.SMIN[b] = sl[j] - s.SMIN[b];
                }
            }
            __syncthreads();
            if (tid == 0 && a.fluxes)
                wnfluxes<SHALLOW>(t, s, Ps[3], Ps[0], x.pp[(size_t)4 * XS + lp], x.pp[(size_t)5 * XS + lp], Pmij, tfk,
                                  a.fluxes + mp, a.NMP);
        }
        // SBOTTOM (SL_BOT = 0 + factor * F, FL + factor), the limiter and the
        // implicit update, into F in place (nothing reads F after SNONLIN's
        // last quadruplet pass)
#pragma unroll
        for (int j = 0; j < SNL_NJ; j++) {
            const int b = tid + SNL_NT * j;
            if (b < NB) {
                const int m = b / KL, k = b - KL * m;
                float sv = sl[j], fv = fl[j];
                if (SHALLOW) {
                    const float slb = 0.f + tbot[m] * s.F[b];
                    fv = fv + tbot[m];
                    sv = sv + slb;
                }
                const float temp1 = (delt * sv) / wam_maxf(1.f, 1.f - delt * fv);
                const float temp3 = wam_minf(wam_fabsf(temp1), tlim[m]);
                const float f = s.F[b] + wam_signf(temp3, temp1);
                s.F[b] = wam_maxf(f, flm[m] * sprd[k]);
            }
        }
        __syncthreads();
        {
            const float4 *F4 = reinterpret_cast<const float4 *>(s.F);
            for (int q = tid; q < NB4; q += SNL_NT) *reinterpret_cast<float4 *>(x.ft + q0 + (size_t)q * 256) = F4[q];
        }
        if (tid == 0) Pnext = p0 + tnext;
        __syncthreads();
    }
}

template <int SHALLOW>
__global__ void __launch_bounds__(PW) implsch_post_kernel(DevTables t, ImplArgs a, ImplScratch x)
{
    const int KL = t.KL, ML = t.ML, NS = t.NSEA, NB4 = x.NB4, XS = x.NX, lane = threadIdx.x;
    const int lp0 = blockIdx.x * blockDim.x + lane;   // the point's row in the work arrays
    const bool on = lp0 < x.NP;
    const int lp = on ? lp0 : x.NP - 1, mp = x.P0 + lp;      // member-point
    const int mem = mp / NS, pt = mp - mem * NS;            // ensemble member, sea point
    const float *FQ = x.ft + quad_offset(lp, 0, NB4);   // bins 4q..4q+3 at FQ[q * 256]
    const unsigned *LLW = x.llws + lp;
    const float ust = a.ustar[mp];
    // direction sums (plain and LLWS-masked) and their integrals; frequency ML first
    float s_e = 0.f, s_fm = 0.f, s_w1 = 0.f, s_w2 = 0.f, Tl = 0.f, Ql = 0.f;
    for (int i = 0; i < ML; i++) {
        const int m = i == 0 ? ML - 1 : i - 1;
        float T = 0.f, Q = 0.f;
        bool qfirst = true;
        for (int k0 = 0; k0 < KL; k0 += 4) {
            const int b0 = k0 + KL * m;
            const float4 f4 = *reinterpret_cast<const float4 *>(FQ + (size_t)(b0 >> 2) * 256);
            const unsigned w = LLW[(size_t)(b0 >> 5) * XS] >> (b0 & 31);   // 4 | 32 and b0 % 4 == 0
            const float fv[4] = {f4.x, f4.y, f4.z, f4.w};
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int k = k0 + u;
                const float f = fv[u];
                T = k == 0 ? f : T + f;
                if ((w >> u) & 1u) {
                    Q = qfirst ? f : Q + f;
                    qfirst = false;
                }
            }
        }
        if (i == 0) {
            Tl = T; Ql = Q;
            s_e = t.MO_TAIL * T; s_fm = t.MM1_TAIL * T; s_w1 = t.MO_TAIL * Q; s_w2 = t.MM1_TAIL * Q;
        } else {
            s_e = s_e + T * t.DFIM[m]; s_fm = s_fm + T * t.DFIMOFR[m];
            s_w1 = s_w1 + Q * t.DFIM[m]; s_w2 = s_w2 + Q * t.DFIMOFR[m];
        }
    }
    {
        const int m = ML - 1;
        s_e = s_e + Tl * t.DFIM[m]; s_fm = s_fm + Tl * t.DFIMOFR[m];
        s_w1 = s_w1 + Ql * t.DFIM[m]; s_w2 = s_w2 + Ql * t.DFIMOFR[m];
    }
    const float e = wam_maxf(s_e, t.EMIN);
    const float fmean = e / wam_maxf(s_fm, t.EMIN);
    const float ew = wam_maxf(s_w1, t.EMIN);
    const float fmeanws = ew / wam_maxf(s_w2, t.EMIN);
    const int mij = frcut(t, t.FR, fmean, fmeanws, ust);
    const int indep = SHALLOW ? t.INDEP[pt] : 1;
    float tden;
    if (SHALLOW) {
        const float akm0 = 1.f / t.TFAK[tb(t, mij - 1, indep)];
        tden = wam_ipowf(akm0, 3) / t.TCGOND[tb(t, mij - 1, indep)];
    } else {
        tden = t.FRM5[mij - 1];
    }
    for (int m = 0; m < ML; m++) {
        const bool tail = m >= mij;
        float ttail = 0.f;
        if (tail) {
            const float an = SHALLOW ? wam_ipowf(1.f / t.TFAK[tb(t, m, indep)], 3) / t.TCGOND[tb(t, m, indep)]
                                     : t.FRM5[m];
            ttail = an / tden;
        }
        const int mrow = tail ? mij - 1 : m;
        for (int k0 = 0; k0 < KL; k0 += 4) {
            const float4 f4 = *reinterpret_cast<const float4 *>(FQ + (size_t)((k0 + KL * mrow) >> 2) * 256);
            const float fv[4] = {f4.x, f4.y, f4.z, f4.w};
            if (on) {
#pragma unroll
                for (int u = 0; u < 4; u++)
                    a.fl3[((size_t)mem * KL * ML + k0 + u + KL * m) * NS + pt] = tail ? ttail * fv[u] : fv[u];
            }
        }
    }
    if (on && a.mij) a.mij[mp] = mij;
}

size_t implsch_snl_lds_bytes(const DevTables &t, int lcflx)
{
    const int ML = t.ML, KL = t.KL, MH = t.MLSTHG, NB = KL * ML;
    size_t n = 2 * 9 * (size_t)SP_TSTRIDE_DEV + (size_t)NB;
    if (lcflx) n += (t.shallow ? 2 : 1) * (size_t)NB;
    n += 2 * (size_t)ML + 3 * (size_t)KL + 36 * (size_t)MH + 2 * (size_t)t.SP_NCHUNK * t.SP_NRUN;
    return n * sizeof(float);
}

// Whether the three-kernel form applies: Janssen physics, whole 4-bin quads
// per direction row, at most SNL_NJ bins per thread, one chunk's quadruplets
// of one frequency within the slots.
bool implsch_split_supported(const DevTables &t)
{
    return t.iphys == 0 && t.KL % 4 == 0 && t.KL * t.ML <= SNL_NT * SNL_NJ && 2 * t.KL < 256 * SP_QPT_DEV;
}

// The three kernels of IPHYS = 0 on one stream for the chunk of sea points
// x.P0 .. x.P0 + x.NP - 1 (rows 0 .. NP - 1 of the work arrays); the middle
// one persistent, one workgroup per resident slot.
hipError_t launch_implsch_split(const DevTables &t, const ImplArgs &a, const ImplScratch &x, int ncu, hipStream_t st,
                                const hipEvent_t *ev, const LaunchTune &tu)
{
    const int W = tu.src_wg;
    const int nwg = (x.NP + W - 1) / W;
    auto mark = [&](int i) { if (ev && ev[i]) (void)hipEventRecord(ev[i], st); };
    mark(0);
    const size_t lds1 = (size_t)t.KL * W * sizeof(float);
    if (t.shallow) hipLaunchKernelGGL(implsch_pre_kernel<1>, dim3(nwg), dim3(W), lds1, st, t, a, x);
    else hipLaunchKernelGGL(implsch_pre_kernel<0>, dim3(nwg), dim3(W), lds1, st, t, a, x);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    mark(1);
    void (*k)(DevTables, ImplArgs, ImplScratch) = t.shallow ? implsch_snl_kernel<1> : implsch_snl_kernel<0>;
    const size_t lds = implsch_snl_lds_bytes(t, a.lcflx);
    e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, SNL_NT, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    if (tu.snl_per_cu > 0 && tu.snl_per_cu < per_cu) per_cu = tu.snl_per_cu;
    const long want = (long)per_cu * ncu;
    const int nblocks = (int)(want < x.NP ? want : x.NP);
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(SNL_NT), lds, st, t, a, x);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    mark(2);
    if (t.shallow) hipLaunchKernelGGL(implsch_post_kernel<1>, dim3(nwg), dim3(W), 0, st, t, a, x);
    else hipLaunchKernelGGL(implsch_post_kernel<0>, dim3(nwg), dim3(W), 0, st, t, a, x);
    e = hipGetLastError();
    mark(3);
    return e;
}

size_t implsch_lds_bytes(const DevTables &t, int lcflx)
{
    const int NB = t.KL * t.ML, ML = t.ML, KL = t.KL, MH = t.MLSTHG;
    size_t n = 7 * (size_t)ML + 3 * KL + 26 * (size_t)MH + (t.iphys == 2 ? 2 * (size_t)t.NK10 : 0) + 5 * (size_t)MH +
               8 * (size_t)KL;
    n += 4 * (size_t)NB + (lcflx ? NB : 0) + (t.shallow ? NB : 0);
    n += (size_t)scr_floats(KL, ML, t.CHUNK, t.iphys);
    return n * sizeof(float) + (size_t)NB;
}

// Launches one persistent workgroup per resident slot: occupancy (set by the
// LDS footprint) times the number of CUs, capped at the number of sea points.
hipError_t launch_implsch(const DevTables &t, const ImplArgs &a, int ncu, hipStream_t st)
{
    const size_t lds = implsch_lds_bytes(t, a.lcflx);
    void (*k)(DevTables, ImplArgs) = nullptr;
    if (t.iphys == 2) k = t.shallow ? implsch_kernel<2, 1> : implsch_kernel<2, 0>;
    else k = t.shallow ? implsch_kernel<0, 1> : implsch_kernel<0, 0>;
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, NT, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    const long want = (long)per_cu * ncu;
    const int nblocks = (int)(want < t.NSEA ? want : t.NSEA);
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(NT), lds, st, t, a);
    return hipGetLastError();
}

}  // namespace wam
