/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 *
 * IMPLSCH restated per sea point, wam_source_module.f90:374-700.  The
 * reference vectorises every statement over the block index IJ and no
 * statement couples two sea points (STRESSO's loop to MAXVAL(MIJ) only adds
 * exact zeros beyond a point's own MIJ), so the per-point restatement gives
 * the same arithmetic.  Per-point spectra are held as F[k + KL*m].
 */
#include <stdlib.h>
#include <string.h>
#include "orc_common.h"
#ifdef _OPENMP
#include <omp.h>
#endif

real orc_aki(const orc_ctx *c, real OM, real BETA);

typedef struct {
    real *F, *SL, *FL, *SPOS, *SMIN, *SLBOT, *TEMP, *TMP2;
    char *LLWS;
    real *WN, *CGG, *A, *S, *D, *DPOS, *KN, *W;   /* ST6 work */
} orc_work;

#define TB(t, m) (t)[(size_t)(m) * c->cfg.ndepth + indep - 1]

/* TOTAL_ENERGY_B / FEMEAN / TM1_TM2_PERIODS_B / WM1_WM2_WAVENUMBER_B,
   wam_interface_module.f90:1342-1410, 207-286, 1152-1245, 1591-1716.
   TEMP(M) = SUM(F(K,M), K) in the natural order, masked when mask != 0. */
static void dirsum(const orc_ctx *c, const real *F, const char *mask, real *T)
{
    int KL = c->KL, ML = c->ML;
    for (int m = 0; m < ML; m++) {
        real s = R(0.);
        int first = 1;
        for (int k = 0; k < KL; k++) {
            if (mask && !mask[k + KL * m]) continue;
            s = first ? F[k + KL * m] : s + F[k + KL * m];
            first = 0;
        }
        T[m] = s;
    }
}
static real total_energy(const orc_ctx *c, const real *T)
{
    int ML = c->ML;
    real e = c->MO_TAIL * T[ML - 1];
    for (int m = 0; m < ML; m++) e = e + T[m] * c->DFIM[m];
    return MAXR(e, c->EMIN);
}
static real femean(const orc_ctx *c, const real *T, real emean)
{
    int ML = c->ML;
    real fm = c->MM1_TAIL * T[ML - 1];
    for (int m = 0; m < ML; m++) fm = fm + T[m] * c->DFIMOFR[m];
    return emean / MAXR(fm, c->EMIN);
}
static real tm1(const orc_ctx *c, const real *T, real emean)
{
    int ML = c->ML;
    real t = c->MP1_TAIL * T[ML - 1];
    for (int m = 0; m < ML; m++) t = t + T[m] * c->DFIM_FR[m];
    return emean > c->EMIN ? emean / t : R(1.);
}
static void wm1_wm2(const orc_ctx *c, const real *T, real emean, int indep, real *wm1, real *wm2)
{
    int ML = c->ML;
    real DEL2 = SQRT(c->G) / c->ZPI;
    real w1 = (c->MM1_TAIL * DEL2) * T[ML - 1];
    for (int m = 0; m < ML; m++) w1 = w1 + (T[m] / SQRT(TB(c->TFAK, m))) * c->DFIM[m];
    *wm1 = emean > c->EMIN ? IPOW(emean / w1, 2) : R(1.);
    DEL2 = c->ZPI / SQRT(c->G);
    real w2 = (c->MP1_TAIL * DEL2) * T[ML - 1];
    for (int m = 0; m < ML; m++) w2 = w2 + (T[m] * SQRT(TB(c->TFAK, m))) * c->DFIM[m];
    *wm2 = emean > c->EMIN ? IPOW(w2 / emean, 2) : R(1.);
}

/* WAVNU2, wam_source_module.f90:4343-4461 */
static void wavnu2(const orc_ctx *c, real H, real *WN, real *CGG)
{
    const real EPS = R(1E-7); const int NMAX = 15;
    real G = c->G;
    for (int m = 0; m < c->ML; m++) {
        real SIG = c->ZPI * c->FR[m];
        real KOLD = R(0.), K, W0 = ABS(SIG), CG;
        if (W0 < SQRT(G / H)) K = W0 / SQRT(G * H); else K = (W0 * W0) / G;
        for (int i = 1; i <= NMAX; i++) {
            real DIF = ABS(K - KOLD);
            real RDIF = (K != R(0.)) ? DIF / K : R(0.);
            if (DIF < EPS && RDIF < EPS) break;
            KOLD = K;
            real th = TANH(KOLD * H);
            real F1 = ((G * KOLD) * th) - W0 * W0;
            real FD;
            if (KOLD * H > R(25.)) FD = G * th;
            else { real ch = COSH(KOLD * H); FD = G * th + ((G * KOLD) * H) / (ch * ch); }
            K = KOLD - F1 / FD;
        }
        if ((R(2.) * K) * H > R(25.)) CG = (W0 / K) * R(0.5);
        else CG = ((W0 / K) * R(0.5)) * (R(1.) + ((R(2.) * K) * H) / SINH((R(2.) * K) * H));
        WN[m] = K; CGG[m] = CG;
    }
}

/* W3FLX4, wam_source_module.f90:5372-5443 */
static void w3flx4(const orc_ctx *c, real U10, real *UST, real *Z0)
{
    real CD;
    if (U10 >= R(50.33)) {
        *UST = R(2.026) * SQRT(c->CDFAC);
        CD = IPOW(*UST / U10, 2);
    } else {
        CD = (c->CDFAC * ((R(8.058) + R(0.967) * U10) - R(0.016) * (U10 * U10))) * R(1E-4);
        *UST = U10 * SQRT(CD);
    }
    *Z0 = c->XNLEV * EXP(R(-0.4) / SQRT(CD));
}

/* AIRSEA, wam_source_module.f90:868-971 */
static void airsea(const orc_ctx *c, real UTOP, real TAUW, real *USTAR, real *Z0)
{
    const real TWOXMP1 = R(3.0), EPSUS = R(1.0E-6), ACD = R(8.0E-4), BCD = R(8.0E-5);
    real XLOGXL = LOG(c->XNLEV), ALPHAOG = c->ALPHA / c->G;
    real XKUTOP = c->XKAPPA * UTOP;
    real USTOLD = UTOP * SQRT(ACD + BCD * UTOP);
    real TAUOLD = MAXR(USTOLD * USTOLD, TAUW + c->EPS1);
    real UST = SQRT(TAUOLD);
    real USTM1 = R(1.0) / MAXR(UST, EPSUS);
    real Z0CH = R(0.);
    for (int it = 1; it <= 15; it++) {
        real X = TAUW / TAUOLD;
        Z0CH = (ALPHAOG * TAUOLD) / SQRT(MAXR(R(1.0) - X, c->EPS1));
        real Z0VIS = c->RNUAIRM * USTM1;
        real Z0TOT = Z0CH + Z0VIS;
        real XOLOGZ0 = R(1.0) / (XLOGXL - LOG(Z0TOT));
        real F = UST - XKUTOP * XOLOGZ0;
        real ZZ = (USTM1 * ((Z0CH * (R(2.0) - TWOXMP1 * X)) / (R(1.0) - X) - Z0VIS)) / Z0TOT;
        real DELF = R(1.0) - (XKUTOP * (XOLOGZ0 * XOLOGZ0)) * ZZ;
        UST = UST - F / DELF;
        real TAUNEW = MAXR(UST * UST, TAUW + c->EPS1);
        UST = SQRT(TAUNEW);
        if (TAUNEW == TAUOLD) break;
        USTM1 = R(1.0) / MAXR(UST, EPSUS);
        TAUOLD = TAUNEW;
    }
    *USTAR = UST; *Z0 = Z0CH;
}

/* WSIGSTAR, wam_source_module.f90:4250-4339 */
static real wsigstar(const orc_ctx *c, real USTAR, real Z0, real WSTAR)
{
    const real LOG10C = R(2.30258509299405), ONETHIRD = R(1.0) / R(3.0);
    const real C1 = R(1.03E-3), C2 = R(0.04E-3), P1 = R(1.48), P2 = R(-0.21);
    real XKAPPAD = R(1.0) / c->XKAPPA;
    real U10 = (USTAR * XKAPPAD) * (LOG10C - LOG(Z0));
    U10 = MAXR(U10, R(1.0));
    real U10M1 = R(1.0) / U10;
    real C2U10P1 = C2 * POW(U10, P1);
    real U10P2 = POW(U10, P2);
    real C_D = (C1 + C2U10P1) * U10P2;
    real DC_DDU = (((P2 * C1) + (P1 + P2) * C2U10P1) * U10P2) * U10M1;
    real SIG_CONV = R(1.0) + ((R(0.5) * U10) / C_D) * DC_DDU;
    real base = R(0.0) * IPOW(USTAR, 3) + R(0.2) * IPOW(WSTAR, 3);
    return MINR(R(0.1), (SIG_CONV * U10M1) * POW(base, ONETHIRD));
}

/* SINPUT, wam_source_module.f90:2111-2308 (NSIN = 2). */
static void sinput(const orc_ctx *c, orc_work *w, real USTAR, real UDIR, real Z0,
                   real ROAIRN, real WSTAR, int indep)
{
    int KL = c->KL, ML = c->ML;
    real *TEMP = w->TEMP;                        /* cos(TH-UDIR) per K */
    for (int k = 0; k < KL; k++) TEMP[k] = COS(c->TH[k] - UDIR);
    real XK2 = c->XKAPPA * c->XKAPPA;
    real CONST3 = ((R(2.) * c->XKAPPA) / c->BETAMAX) * XK2;
    real SIG_N = wsigstar(c, USTAR, Z0, WSTAR);
    const real WSIN[2] = {R(0.5), R(0.5)};
    real USN[2] = {USTAR * (R(1.0) - SIG_N), USTAR * (R(1.0) + SIG_N)};
    for (int m = 0; m < ML; m++) {
        real FAC = c->ZPI * c->FR[m];
        real CONST = ((FAC * c->BETAMAX) / XK2) / c->ROWATER;
        real XK, CM, CNSN;
        if (c->cfg.shallow) {
            XK = TB(c->TFAK, m);
            CM = XK / FAC;
            CNSN = ((CONST * IPOW(FAC, 2)) / (c->G * XK)) * ROAIRN;
        } else {
            XK = IPOW(FAC, 2) / c->G;
            CM = FAC / c->G;
            CNSN = CONST * ROAIRN;
        }
        real UCN[2], ZCN[2], XV1D[2];
        for (int is = 0; is < 2; is++) {
            UCN[is] = USN[is] * CM + c->ZALP;
            ZCN[is] = LOG(XK * Z0);
            XV1D[is] = R(-1.) / (((USN[is] / c->XKAPPA) * ZCN[is]) * CM);
        }
        for (int k = 0; k < KL; k++) {
            int b = k + KL * m;
            real UFAC1 = R(0.);
            char llws = 0;
            for (int is = 0; is < 2; is++) {
                if (TEMP[k] > R(0.01)) {
                    real X = TEMP[k] * UCN[is];
                    real ZLOG = ZCN[is] + c->XKAPPA / X;
                    if (ZLOG < R(0.)) {
                        real ZLOG2X = (ZLOG * ZLOG) * X;
                        UFAC1 = UFAC1 + ((WSIN[is] * EXP(ZLOG)) * ZLOG2X) * ZLOG2X;
                        llws = 1;
                    }
                }
            }
            real UFAC2 = (CONST3 * (TEMP[k] - XV1D[0])) * IPOW(UCN[0], 2);
            UFAC2 = WSIN[0] * UFAC2 + ((WSIN[1] * CONST3) * (TEMP[k] - XV1D[1])) * IPOW(UCN[1], 2);
            real sp = CNSN * UFAC1;
            real fl = sp + CNSN * UFAC2;
            w->SPOS[b] = sp * w->F[b];
            w->FL[b] = fl;
            w->SL[b] = fl * w->F[b];
            w->LLWS[b] = llws;
        }
    }
}

static real sum_seq(const real *x, int n) { real s = x[0]; for (int i = 1; i < n; i++) s = s + x[i]; return s; }

/* TAUWINDS, wam_source_module.f90:5555-5587 */
static real tauwinds(const orc_ctx *c, const real *sd, const real *lf, const real *cinv, const real *dsii, int n)
{
    real s = R(0.);
    for (int i = 0; i < n; i++) {
        real v = lf ? sd[i] * lf[i] : sd[i];
        v = (v * cinv[i]) * dsii[i];
        s = (i == 0) ? v : s + v;
    }
    return (c->G * c->ROWATER) * s;
}

/* LFACTOR, wam_source_module.f90:4465-4700.  S is [NK][NTH] (direction fastest). */
static void lfactor(const orc_ctx *c, const real *S, const real *CINV, real U10, real USTAR,
                    real USDIR, real ROAIRN, const real *SIG, const real *DSII, real *LFACT,
                    real *TAUWX, real *TAUWY)
{
    int NTH = c->KL, NK = c->ML;
    const real FRQMAX = R(10.), SIN6WS = R(32.);
    int NK10 = (int)CEILR(LOG(FRQMAX / (SIG[0] / c->ZPI)) / LOG(c->CO)) + 1;
    NK10 = MAXI(NK, NK10);
    real *SIG10 = calloc(NK10, sizeof(real)), *CINV10 = calloc(NK10, sizeof(real));
    real *DSII10 = calloc(NK10, sizeof(real)), *LF10 = calloc(NK10, sizeof(real));
    real *SD = calloc(NK10, sizeof(real)), *SX = calloc(NK10, sizeof(real)), *SY = calloc(NK10, sizeof(real));
    real *UCINV = calloc(NK10, sizeof(real));
    for (int ik = 0; ik < NK; ik++) {
        real s = R(0.), sx = R(0.), sy = R(0.);
        for (int ith = 0; ith < NTH; ith++) {
            real v = S[ith + NTH * ik];
            real vp = MAXR(R(0.), v);
            if (ith == 0) { s = v; sx = vp * c->COSTH[ith]; sy = vp * c->SINTH[ith]; }
            else { s = s + v; sx = sx + vp * c->COSTH[ith]; sy = sy + vp * c->SINTH[ith]; }
        }
        SD[ik] = s * c->DELTH; SX[ik] = sx * c->DELTH; SY[ik] = sy * c->DELTH;
    }
    if (NK < NK10) {
        for (int i = 0; i < NK10; i++) SIG10[i] = SIG[0] * POW(c->CO, (real)(i + 1) - R(1.0));
        for (int i = 0; i < NK; i++) CINV10[i] = CINV[i];
        for (int i = NK; i < NK10; i++) CINV10[i] = SIG10[i] * R(0.101978);
        for (int i = 0; i < NK10; i++) DSII10[i] = (R(0.5) * SIG10[i]) * (c->CO - R(1.0) / c->CO);
        DSII10[0] = (R(0.5) * SIG10[0]) * (c->CO - R(1.0));
        DSII10[NK10 - 1] = ((R(0.5) * SIG10[NK10 - 1]) * (c->CO - R(1.0))) / c->CO;
        for (int i = NK; i < NK10; i++) {
            real r2 = IPOW(SIG10[NK - 1] / SIG10[i], 2);
            SD[i] = SD[NK - 1] * r2; SX[i] = SX[NK - 1] * r2; SY[i] = SY[NK - 1] * r2;
        }
    } else {
        for (int i = 0; i < NK; i++) { SIG10[i] = SIG[i]; CINV10[i] = CINV[i]; DSII10[i] = DSII[i]; }
    }
    real TAU_TOT = (USTAR * USTAR) * ROAIRN;
    real TAU_VIS = (MAXR(R(0.0), R(-5.0E-5) * U10 + R(1.1E-3)) * (U10 * U10)) * ROAIRN;
    TAU_VIS = MINR(R(0.95) * TAU_TOT, TAU_VIS);
    real TAUVX = TAU_VIS * COS(USDIR), TAUVY = TAU_VIS * SIN(USDIR);
    real TWX = tauwinds(c, SX, NULL, CINV10, DSII10, NK10);
    real TWY = tauwinds(c, SY, NULL, CINV10, DSII10, NK10);
    real TAUX = TAUVX + TWX, TAUY = TAUVY + TWY;
    real TAU = SQRT(TAUX * TAUX + TAUY * TAUY);
    real ERR = (TAU - TAU_TOT) / TAU_TOT;
    for (int i = 0; i < NK10; i++) LF10[i] = R(1.0);
    if (TAU > TAU_TOT) {
        int OVERSHOT = 0;
        real RTAU = ERR / R(90.), DRTAU = R(2.0);
        int SIGN_NEW = (ERR >= R(0.)) ? 1 : -1, SIGN_OLD;
        real UPROXY = SIN6WS * USTAR;
        for (int i = 0; i < NK10; i++) UCINV[i] = R(1.0) - UPROXY * CINV10[i];
        for (int it = 1; it <= 80; it++) {
            for (int i = 0; i < NK10; i++) LF10[i] = MINR(R(1.0), EXP(UCINV[i] * RTAU));
            TWX = tauwinds(c, SX, LF10, CINV10, DSII10, NK10);
            TWY = tauwinds(c, SY, LF10, CINV10, DSII10, NK10);
            TAUX = TAUVX + TWX; TAUY = TAUVY + TWY;
            TAU = SQRT(TAUX * TAUX + TAUY * TAUY);
            ERR = (TAU - TAU_TOT) / TAU_TOT;
            SIGN_OLD = SIGN_NEW;
            SIGN_NEW = (ERR >= R(0.)) ? 1 : -1;
            if (SIGN_NEW != SIGN_OLD) OVERSHOT = 1;
            if (OVERSHOT) DRTAU = MAXR(R(0.5) * (R(1.0) + DRTAU), R(1.00010));
            RTAU = (SIGN_NEW > 0) ? RTAU * DRTAU : RTAU * (R(1.) / DRTAU);
            if (ABS(ERR) < R(1.54E-4)) break;
        }
    }
    for (int i = 0; i < NK; i++) LFACT[i] = LF10[i];
    *TAUWX = TWX; *TAUWY = TWY;
    free(SIG10); free(CINV10); free(DSII10); free(LF10); free(SD); free(SX); free(SY); free(UCINV);
}

/* SINPUT_ST6, wam_source_module.f90:4704-4925 (TAUNW is not used by IMPLSCH). */
static void sinput_st6(const orc_ctx *c, orc_work *w, real UABS, real USTAR, real USDIR,
                       real ROAIRN, real *TAUW)
{
    int NTH = c->KL, NK = c->ML;
    const real SIN6WS = R(32.), SIN6A0 = R(0.09);
    real SIG[64], DSII[64], ADENSIG[64], KMAX[64], ANAR[64], SQRTBN[64], CINV[64], LFACT[64];
    for (int m = 0; m < NK; m++) { SIG[m] = c->ZPI * c->FR[m]; DSII[m] = c->ZPI * c->DF[m]; }
    real COSU = COS(USDIR), SINU = SIN(USDIR);
    real UPROXY = SIN6WS * USTAR;
    real *A = w->A, *S = w->S, *D = w->D, *KN = w->KN, *W = w->W;
    for (int m = 0; m < NK; m++)
        for (int k = 0; k < NTH; k++) {
            int b = k + NTH * m;
            A[b] = (w->F[b] * w->CGG[m]) / (c->ZPI * SIG[m]);
            KN[b] = A[b];
        }
    for (int m = 0; m < NK; m++) {
        ADENSIG[m] = (sum_seq(&KN[NTH * m], NTH) * SIG[m]) * c->DELTH;
        real mx = KN[NTH * m];
        for (int k = 1; k < NTH; k++) mx = MAXR(mx, KN[k + NTH * m]);
        KMAX[m] = mx;
        for (int k = 0; k < NTH; k++) {
            int b = k + NTH * m;
            KN[b] = (KMAX[m] < R(1.0E-34)) ? R(1.) : KN[b] / KMAX[m];
        }
        ANAR[m] = R(1.0) / (sum_seq(&KN[NTH * m], NTH) * c->DELTH);
        SQRTBN[m] = SQRT((ANAR[m] * ADENSIG[m]) * IPOW(w->WN[m], 3));
        CINV[m] = w->WN[m] / SIG[m];
    }
    real RW = ROAIRN / c->ROWATER;
    for (int m = 0; m < NK; m++)
        for (int k = 0; k < NTH; k++) {
            int b = k + NTH * m;
            real cosd = c->COSTH[k] * COSU + c->SINTH[k] * SINU;
            real x = ((UPROXY * CINV[m]) * cosd) - R(1.);
            real W1 = IPOW(MAXR(R(0.), x), 2);
            W[b] = IPOW(MINR(R(0.), x), 2);
            D[b] = ((((RW * SIG[m]) * (R(2.8) - (R(1.) + TANH((R(10.) * SQRTBN[m]) * W1 - R(11.))))) * SQRTBN[m]) * W1);
            S[b] = D[b] * A[b];
            S[b] = (S[b] * SIG[m]) / w->CGG[m];            /* SDENSIG */
        }
    real TAUWX, TAUWY;
    lfactor(c, S, CINV, UABS, USTAR, USDIR, ROAIRN, SIG, DSII, LFACT, &TAUWX, &TAUWY);
    real lsum = sum_seq(LFACT, NK);
    if (lsum < (real)NK)
        for (int m = 0; m < NK; m++)
            for (int k = 0; k < NTH; k++) D[k + NTH * m] = D[k + NTH * m] * LFACT[m];
    for (int b = 0; b < NK * NTH; b++) w->DPOS[b] = D[b];
    for (int m = 0; m < NK; m++)
        for (int k = 0; k < NTH; k++) {
            int b = k + NTH * m;
            real W2 = W[b];
            D[b] = D[b] - ((((RW * SIG[m]) * SIN6A0) * (R(2.8) - (R(1.) + TANH((R(10.) * SQRTBN[m]) * W2 - R(11.))))) * SQRTBN[m]) * W2;
        }
    *TAUW = SQRT(TAUWX * TAUWX + TAUWY * TAUWY);
    for (int b = 0; b < NK * NTH; b++) {
        w->SL[b] = D[b] * w->F[b];
        w->SPOS[b] = w->DPOS[b] * w->F[b];
        w->FL[b] = D[b];
        w->LLWS[b] = w->SPOS[b] > R(0.);
    }
}

/* FRCUTINDEX, wam_source_module.f90:3569-3647 */
static int frcutindex(const orc_ctx *c, real FM, real FMWS, real USTAR)
{
    const real EPSUS = R(1.0E-6), FRIC = R(28.0), TAILFACTOR = R(2.5), TAILFACTOR_PM = R(3.0);
    real FPMH = TAILFACTOR / c->FR[0];
    real FPPM = (TAILFACTOR_PM * c->G) / ((FRIC * c->ZPI) * c->FR[0]);
    real FM2 = MAXR(FMWS, FM) * FPMH;
    real FPM = FPPM / MAXR(USTAR, EPSUS);
    real FPM4 = MAXR(FM2, FPM);
    int MIJ = NINT(LOG10(FPM4) * c->INV_LOG_CO) + 1;
    return MINI(MAXI(1, MIJ), c->ML);
}

/* FRCUTINDEX_ST6, wam_source_module.f90:5447-5515 */
static int frcutindex_st6(const orc_ctx *c, real FMEAN, real USTAR)
{
    int NK = c->ML;
    real FXFM = R(6.) * c->ZPI;
    real FXPM = (R(4.0) * c->G) / R(28.);
    real SIGNK = c->ZPI * c->FR[NK - 1];
    real FHIGH = MAXR(FXFM * FMEAN, FXPM / USTAR);
    real FACTI1 = R(1.) / LOG(c->CO);
    real FACTI2 = R(1.) - LOG(c->ZPI * c->FR[0]) * FACTI1;
    int NKH = MINI(NK, INTR(FACTI2 + FACTI1 * LOG(MAXR(R(1.E-7), FHIGH))));
    int NKH1 = MINI(NK, NKH + 1);
    FHIGH = MINR(SIGNK, MAXR(FXFM * FMEAN, FXPM / USTAR));
    NKH = MAXI(2, MINI(NKH1, INTR(FACTI2 + FACTI1 * LOG(MAXR(R(1.E-7), FHIGH)))));
    return NKH;
}

/* TAU_PHI_HF with XLEVTAIL = 0, wam_source_module.f90:3853-3966 */
static void tau_phi_hf(const orc_ctx *c, int MIJ, real USTAR, real Z0, real *TAUHF, real *PHIHF)
{
    real GM1 = R(1.0) / c->G, X0G = c->X0TAUHF * c->G;
    real OMEGAC = c->ZPI * c->FR[MIJ - 1];
    real UST0 = USTAR, TAUW0 = UST0 * UST0;
    real GZ0 = c->G * Z0, XLOGGZ0 = LOG(GZ0);
    real OMEGACC = MAXR(OMEGAC, X0G / UST0);
    real SQRTZ0OG = SQRT(Z0 * GM1), SQRTGZ0 = R(1.0) / SQRTZ0OG;
    real YC = OMEGACC * SQRTZ0OG, ZINF = LOG(YC);
    real DELZ = MAXR((R(0.0) - ZINF) / (real)18, R(0.0));
    real th = R(0.), ph = R(0.), TAUW = TAUW0, UST = UST0;
    for (int J = 1; J <= 19; J++) {
        real Y = EXP(ZINF + (real)(J - 1) * DELZ);
        real OMEGA = Y * SQRTGZ0;
        real CM1 = OMEGA * GM1;
        real ZX = UST * CM1 + c->ZALP;
        real ZARG = c->XKAPPA / ZX;
        real ZLOG = (XLOGGZ0 + R(2.0) * LOG(CM1)) + ZARG;
        ZLOG = MINR(ZLOG, R(0.0));
        real ZBETA = EXP(ZLOG) * IPOW(ZLOG, 4);
        real FNC2 = ((ZBETA * TAUW) * c->WTAUHF[J - 1]) * DELZ;
        TAUW = MAXR(TAUW - R(0.) * FNC2, R(0.0));
        UST = SQRT(TAUW);
        th = th + FNC2;
        ph = ph + FNC2 / Y;
    }
    *TAUHF = th; *PHIHF = SQRTZ0OG * ph;
}

/* STRESSO, wam_source_module.f90:3164-3356 (SL argument = SPOS in IMPLSCH). */
static void stresso(const orc_ctx *c, orc_work *w, const real *SLIN, real USTAR, real UDIR,
                    real Z0, int MIJ, int indep, real *TAUW, real *PHIAW)
{
    int KL = c->KL, ML = c->ML;
    real UST2 = USTAR * USTAR, GM1 = R(1.0) / c->G;
    real CONST = (c->DELTH * (real)pow((double)c->ZPI, 4.0)) * GM1;   /* (ZPI)**4 folded */
    real CONST1 = (CONST * c->FR5[MIJ - 1]) * GM1;
    real CONST2 = (c->ROAIR * CONST) * c->FR5[MIJ - 1];
    real phi = R(0.), xs = R(0.), ys = R(0.);
    for (int m = 0; m < MIJ; m++) {
        real CM;
        if (c->cfg.shallow) CM = TB(c->TFAK, m) * (R(1.0) / (c->ZPI * c->FR[m]));
        else CM = (c->ZPI * c->FR[m]) * GM1;
        real RW = c->RHOWG_DFIM[m];
        if (m == MIJ - 1 && MIJ != ML) RW = R(0.5) * RW;
        real st = R(0.), sx = R(0.), sy = R(0.);
        for (int k = 0; k < KL; k++) {
            real sp = MAXR(SLIN[k + KL * m], R(0.));
            if (k == 0) { st = sp; sx = sp * c->SINTH[k]; sy = sp * c->COSTH[k]; }
            else { st = st + sp; sx = sx + sp * c->SINTH[k]; sy = sy + sp * c->COSTH[k]; }
        }
        phi = phi + st * RW;
        real CMRW = CM * RW;
        xs = xs + sx * CMRW;
        ys = ys + sy * CMRW;
    }
    xs = xs / MAXR(c->ROAIR, R(1.));
    ys = ys / MAXR(c->ROAIR, R(1.));
    real T1 = R(0.), T2 = R(0.);
    for (int k = 0; k < KL; k++) {
        real COSW = MAXR(COS(c->TH[k] - UDIR), R(0.));
        real f = w->F[k + KL * (MIJ - 1)];
        if (k == 0) { T2 = f * IPOW(COSW, 2); T1 = T2 * COSW; }
        else { T1 = T1 + f * IPOW(COSW, 3); T2 = T2 + f * IPOW(COSW, 2); }
    }
    real TAU1, PHI1;
    tau_phi_hf(c, MIJ, USTAR, Z0, &TAU1, &PHI1);
    real TAUHF = (CONST1 * T1) * TAU1;
    real PHIHF = (CONST2 * T2) * PHI1;
    phi = phi + PHIHF;
    xs = xs + TAUHF * SIN(UDIR);
    ys = ys + TAUHF * COS(UDIR);
    real tw = SQRT(xs * xs + ys * ys);
    tw = MINR(tw, UST2 - c->EPS1);
    tw = MAXR(tw, R(0.));
    if (TAUW) *TAUW = tw;
    *PHIAW = phi;
}

/* IMPHFTAIL, wam_source_module.f90:3651-3731 */
static void imphftail(const orc_ctx *c, orc_work *w, int MIJ, int indep)
{
    int KL = c->KL, ML = c->ML;
    real *T2 = w->TMP2;
    for (int m = MIJ - 1; m < ML; m++) {
        if (c->cfg.shallow) {
            real AKM1 = R(1.) / TB(c->TFAK, m);
            T2[m] = IPOW(AKM1, 3) / TB(c->TCGOND, m);
        } else T2[m] = c->FRM5[m];
    }
    for (int m = MIJ; m < ML; m++) T2[m] = T2[m] / T2[MIJ - 1];
    for (int k = 0; k < KL; k++) {
        real TFAC = w->F[k + KL * (MIJ - 1)];
        for (int m = MIJ; m < ML; m++) w->F[k + KL * m] = T2[m] * TFAC;
    }
}

/* SDISSIP, wam_source_module.f90:1593-1700 */
static void sdissip(const orc_ctx *c, orc_work *w, real EMEAN, real FMEAN, real AKMEAN, int indep)
{
    int KL = c->KL, ML = c->ML;
    const real CDIS = R(1.33), DELTA = R(0.5);
    real G4 = (real)pow((double)c->G, 4.0);
    real CONSD = ((-CDIS) * (real)pow((double)c->ZPI, 9.0)) / G4;
    real CONSS = (-CDIS) * c->ZPI;
    real SDS;
    if (c->cfg.shallow) SDS = ((CONSS * FMEAN) * (EMEAN * EMEAN)) * IPOW(AKMEAN, 4);
    else SDS = (CONSD * (EMEAN * EMEAN)) * IPOW(FMEAN, 9);
    for (int m = 0; m < ML; m++) {
        real T1;
        if (c->cfg.shallow) T1 = TB(c->TFAK, m) / AKMEAN;
        else T1 = IPOW(c->FR[m] / FMEAN, 2);
        T1 = SDS * ((R(1.) - DELTA) * T1 + DELTA * IPOW(T1, 2));
        for (int k = 0; k < KL; k++) {
            int b = k + KL * m;
            w->SL[b] = w->SL[b] + T1 * w->F[b];
            w->FL[b] = w->FL[b] + T1;
        }
    }
}

/* SDISSIP_ST6, wam_source_module.f90:4929-5102 */
static void sdissip_st6(const orc_ctx *c, orc_work *w)
{
    int NTH = c->KL, NK = c->ML;
    const real SDS6A1 = R(4.75E-6), SDS6A2 = R(7.00E-5);
    real ANAR = R(1.0), BNT = R(0.035) * R(0.035);
    real XFAC = (R(1.0) - R(1.0) / c->CO) / (c->CO - R(1.0) / c->CO);
    real ADF[64], T12[64], T2s = R(0.);
    for (int ik = 0; ik < NK; ik++) {
        real ET = (c->ZPI * BNT) / ((ANAR * w->CGG[ik]) * IPOW(w->WN[ik], 3));
        real ED = sum_seq(&w->F[NTH * ik], NTH) * c->DELTH;
        real EX = MAXR(R(0.0), ED - ET);
        real NEX = EX / ET;
        real T1 = ((SDS6A1 * ANAR) * c->FR[ik]) * IPOW(NEX, 4);
        ADF[ik] = ANAR * IPOW(NEX, 4);
        real DFII = c->DF[ik];
        if (ik > 0 && ik < NK - 1) DFII = DFII * XFAC;
        real term = ADF[ik] * DFII;
        T2s = (ik == 0) ? term : T2s + term;
        real T2 = SDS6A2 * T2s;
        T12[ik] = R(-1.0) * (MAXR(R(0.0), T1) + MAXR(R(0.0), T2));
    }
    for (int ik = 0; ik < NK; ik++)
        for (int k = 0; k < NTH; k++) {
            int b = k + NTH * ik;
            w->SL[b] = w->SL[b] + T12[ik] * w->F[b];
            w->FL[b] = w->FL[b] + T12[ik];
        }
}

/* W3SWL6, wam_source_module.f90:5106-5252 */
static void w3swl6(const orc_ctx *c, orc_work *w)
{
    int NTH = c->KL, NK = c->ML;
    const real SWL6B1 = R(0.0041);
    real SIG[64], DDEN[64], ABAND[64], BN[64];
    for (int m = 0; m < NK; m++) { SIG[m] = c->ZPI * c->FR[m]; DDEN[m] = (c->ZPI * c->DFIM[m]) * SIG[m]; }
    real *KN = w->KN;
    for (int m = 0; m < NK; m++) {
        for (int k = 0; k < NTH; k++) {
            int b = k + NTH * m;
            KN[b] = (w->F[b] * w->CGG[m]) / (c->ZPI * SIG[m]);
        }
        ABAND[m] = sum_seq(&KN[NTH * m], NTH);
        real mx = KN[NTH * m];
        for (int k = 1; k < NTH; k++) mx = MAXR(mx, KN[k + NTH * m]);
        for (int k = 0; k < NTH; k++) {
            int b = k + NTH * m;
            KN[b] = (mx < R(1.0E-34)) ? R(1.) : KN[b] / mx;
        }
        real ANAR = R(1.0) / (sum_seq(&KN[NTH * m], NTH) * c->DELTH);
        BN[m] = (ANAR * ((ABAND[m] * SIG[m]) * c->DELTH)) * IPOW(w->WN[m], 3);
    }
    int IK = 0;
    for (int m = 1; m < NK; m++) if (ABAND[m] > ABAND[IK]) IK = m;
    real s = R(0.);
    for (int m = 0; m < NK; m++) { real v = (ABAND[m] * DDEN[m]) / w->CGG[m]; s = (m == 0) ? v : s + v; }
    real B1 = SWL6B1 * ((R(2.) * SQRT(s)) * w->WN[IK]);
    for (int m = 0; m < NK; m++) {
        real DDIS = R(0.);
        if (ABAND[m] > R(1.E-30)) DDIS = (((-(R(2.) / R(3.))) * B1) * SIG[m]) * SQRT(BN[m]);
        for (int k = 0; k < NTH; k++) {
            int b = k + NTH * m;
            w->SL[b] = w->SL[b] + DDIS * w->F[b];
            w->FL[b] = w->FL[b] + DDIS;
        }
    }
}

/* SNONLIN, wam_source_module.f90:2742-3047 */
static void snonlin(const orc_ctx *c, orc_work *w, real DEPTH, real AKMEAN)
{
    int KL = c->KL, ML = c->ML, MF = c->MFRSTLW;
    real ENHFR = R(1.);
    if (c->cfg.shallow) {
        ENHFR = MAXR((R(0.75) * DEPTH) * AKMEAN, R(.5));
        ENHFR = R(1.) + ((R(5.5) / ENHFR) * (R(1.) - R(.833) * ENHFR)) * EXP(R(-1.25) * ENHFR);
    }
    int MFR1STFR = -MF + 1;
    int MFRLSTFR = ML - c->KFRH + MFR1STFR;
    real *F = w->F, *SL = w->SL, *FL = w->FL;
#define FF(k, m) F[(k) + KL * ((m) - 1)]
#define SLA(k, m, v) SL[(k) + KL * ((m) - 1)] = SL[(k) + KL * ((m) - 1)] + (v)
#define FLA(k, m, v) FL[(k) + KL * ((m) - 1)] = FL[(k) + KL * ((m) - 1)] + (v)
    for (int MC = 1; MC <= c->MLSTHG; MC++) {
        int i = MC - MF;
        int MP = c->IKP[i], MP1 = c->IKP1[i], MM = c->IKM[i], MM1 = c->IKM1[i];
        const int *ic = &c->INLCOEF[(MC - 1) * 5];
        int IC = ic[0], IP = ic[1], IP1 = ic[2], IM = ic[3], IM1 = ic[4];
        const treal *rc = &c->RNLCOEF[(MC - 1) * 25];
        real FTAIL = rc[0], GW1 = rc[1], GW2 = rc[2], GW3 = rc[3], GW4 = rc[4];
        real FKLAMPA = rc[5], FKLAMPB = rc[6], FKLAMP2 = rc[7], FKLAMP1 = rc[8];
        real FKLAPA2 = rc[9], FKLAPB2 = rc[10], FKLAP12 = rc[11], FKLAP22 = rc[12];
        real GW5 = rc[13], GW6 = rc[14], GW7 = rc[15], GW8 = rc[16];
        real FKLAMMA = rc[17], FKLAMMB = rc[18], FKLAMM2 = rc[19], FKLAMM1 = rc[20];
        real FKLAMA2 = rc[21], FKLAMB2 = rc[22], FKLAM12 = rc[23], FKLAM22 = rc[24];
        real FTEMP = c->cfg.shallow ? c->AF11[i] * ENHFR : c->AF11[i];
        int regime = (MC > MFR1STFR && MC < MFRLSTFR) ? 0 : (MC >= MFRLSTFR ? 1 : 2);
        for (int KH = 0; KH < 2; KH++) {
            for (int K = 0; K < KL; K++) {
                int K1 = c->K1W[KH * KL + K] - 1, K2 = c->K2W[KH * KL + K] - 1;
                int K11 = c->K11W[KH * KL + K] - 1, K21 = c->K21W[KH * KL + K] - 1;
                real SAP = GW1 * FF(K1, IP) + GW2 * FF(K11, IP) + GW3 * FF(K1, IP1) + GW4 * FF(K11, IP1);
                real SAM = GW5 * FF(K2, IM) + GW6 * FF(K21, IM) + GW7 * FF(K2, IM1) + GW8 * FF(K21, IM1);
                real FIJ = regime == 0 ? FF(K, IC) : FF(K, IC) * FTAIL;
                real FAD1 = FIJ * (SAP + SAM);
                real FAD2 = FAD1 - (R(2.) * SAP) * SAM;
                FAD1 = FAD1 + FAD2;
                real FCEN = FTEMP * FIJ;
                real AD = FAD2 * FCEN;
                real DELAD = FAD1 * FTEMP;
                real DELAP = ((FIJ - R(2.) * SAM) * c->DAL1) * FCEN;
                real DELAM = ((FIJ - R(2.) * SAP) * c->DAL2) * FCEN;
                if (regime == 0) {
                    SLA(K, MC, -(R(2.) * AD)); FLA(K, MC, -(R(2.) * DELAD));
                    SLA(K2, MM, AD * FKLAMM1); FLA(K2, MM, DELAM * FKLAM12);
                    SLA(K21, MM, AD * FKLAMM2); FLA(K21, MM, DELAM * FKLAM22);
                    SLA(K2, MM1, AD * FKLAMMA); FLA(K2, MM1, DELAM * FKLAMA2);
                    SLA(K21, MM1, AD * FKLAMMB); FLA(K21, MM1, DELAM * FKLAMB2);
                    SLA(K1, MP, AD * FKLAMP1); FLA(K1, MP, DELAP * FKLAP12);
                    SLA(K11, MP, AD * FKLAMP2); FLA(K11, MP, DELAP * FKLAP22);
                    SLA(K1, MP1, AD * FKLAMPA); FLA(K1, MP1, DELAP * FKLAPA2);
                    SLA(K11, MP1, AD * FKLAMPB); FLA(K11, MP1, DELAP * FKLAPB2);
                } else if (regime == 1) {
                    SLA(K2, MM, AD * FKLAMM1); FLA(K2, MM, DELAM * FKLAM12);
                    SLA(K21, MM, AD * FKLAMM2); FLA(K21, MM, DELAM * FKLAM22);
                    if (MM1 <= ML) {
                        SLA(K2, MM1, AD * FKLAMMA); FLA(K2, MM1, DELAM * FKLAMA2);
                        SLA(K21, MM1, AD * FKLAMMB); FLA(K21, MM1, DELAM * FKLAMB2);
                        if (MC <= ML) {
                            SLA(K, MC, -(R(2.) * AD)); FLA(K, MC, -(R(2.) * DELAD));
                            if (MP <= ML) {
                                SLA(K1, MP, AD * FKLAMP1); FLA(K1, MP, DELAP * FKLAP12);
                                SLA(K11, MP, AD * FKLAMP2); FLA(K11, MP, DELAP * FKLAP22);
                                if (MP1 <= ML) {
                                    SLA(K1, MP1, AD * FKLAMPA); FLA(K1, MP1, DELAP * FKLAPA2);
                                    SLA(K11, MP1, AD * FKLAMPB); FLA(K11, MP1, DELAP * FKLAPB2);
                                }
                            }
                        }
                    }
                } else {
                    if (MM1 >= 1) {
                        SLA(K2, MM1, AD * FKLAMMA); FLA(K2, MM1, DELAM * FKLAMA2);
                        SLA(K21, MM1, AD * FKLAMMB); FLA(K21, MM1, DELAM * FKLAMB2);
                    }
                    SLA(K, MC, -(R(2.) * AD)); FLA(K, MC, -(R(2.) * DELAD));
                    SLA(K1, MP, AD * FKLAMP1); FLA(K1, MP, DELAP * FKLAP12);
                    SLA(K11, MP, AD * FKLAMP2); FLA(K11, MP, DELAP * FKLAP22);
                    SLA(K1, MP1, AD * FKLAMPA); FLA(K1, MP1, DELAP * FKLAPA2);
                    SLA(K11, MP1, AD * FKLAMPB); FLA(K11, MP1, DELAP * FKLAPB2);
                }
            }
        }
    }
#undef FF
#undef SLA
#undef FLA
}

/* SBOTTOM, wam_source_module.f90:1527-1589 (contribution to SL goes to SL_BOT). */
static void sbottom(const orc_ctx *c, orc_work *w, real DEPTH, int indep)
{
    int KL = c->KL, ML = c->ML;
    real CONST = (R(-2.0) * R(0.038)) / c->G;
    for (int m = 0; m < ML; m++) {
        real WAV = TB(c->TFAK, m);
        real SBO = MINR((R(2.) * DEPTH) * WAV, R(50.));
        SBO = (CONST * WAV) / SINH(SBO);
        for (int k = 0; k < KL; k++) {
            int b = k + KL * m;
            w->SLBOT[b] = w->SLBOT[b] + SBO * w->F[b];
            w->FL[b] = w->FL[b] + SBO;
        }
    }
}

/* WNFLUXES, wam_source_module.f90:3970-4054.  out[0..6] = PHIOC, TAUOC_X,
   TAUOC_Y, PHIBOT, TAUBOT_X, TAUBOT_Y (PHIAW is written by the caller). */
static void wnfluxes(const orc_ctx *c, orc_work *w, const real *SS, real USTAR, real UDIR,
                     real PHIAW, int MIJ, int indep, real *out)
{
    int KL = c->KL;
    real ROG = c->ROWATER * c->G;
    real phioc = PHIAW, tau = c->ROAIR * (USTAR * USTAR);
    real tx = tau * SIN(UDIR), ty = tau * COS(UDIR);
    real pb = R(0.), bx = R(0.), by = R(0.);
    real SCDFM = (R(0.5) * c->DELTH) * (R(1.) - R(1.) / c->CO);
    for (int m = 0; m < MIJ; m++) {
        real CONSTFM = (m < MIJ - 1) ? ROG * c->DFIM[m] : (ROG * SCDFM) * c->FR[MIJ - 1];
        real CM = c->cfg.shallow ? TB(c->TFAK, m) / (c->ZPI * c->FR[m]) : (c->ZPI * c->FR[m]) / c->G;
        for (int k = 0; k < KL; k++) {
            int b = k + KL * m;
            phioc = phioc - SS[b] * CONSTFM;
            tx = tx - ((CM * SS[b]) * CONSTFM) * c->SINTH[k];
            ty = ty - ((CM * SS[b]) * CONSTFM) * c->COSTH[k];
            if (c->cfg.shallow) {
                pb = pb - w->SLBOT[b] * CONSTFM;
                bx = bx - ((CM * w->SLBOT[b]) * CONSTFM) * c->SINTH[k];
                by = by - ((CM * w->SLBOT[b]) * CONSTFM) * c->COSTH[k];
            }
        }
    }
    out[0] = phioc; out[1] = tx; out[2] = ty; out[3] = pb; out[4] = bx; out[5] = by;
}

/* GAMD, the parameter of the depth-limited wave height (wam_source_module.f90:67). */
#define GAMD R(0.8)

/* SDEPTHLIM, wam_source_module.f90:3735-3780 (IMPLSCH :513-515 with
   WAVE_BREAKING_RUN): the spectrum scaled by EM = MIN((0.25 GAMD DEPTH)**2 /
   EMEAN, 1), and EMEAN by the same factor. */
static void sdepthlim(const orc_ctx *c, orc_work *w, real DEPTH, real *EMEAN)
{
    int NB = c->KL * c->ML;
    real h = (R(0.25) * GAMD) * DEPTH;
    real EM = MINR((h * h) / *EMEAN, R(1.0));
    for (int b = 0; b < NB; b++) w->F[b] = w->F[b] * EM;
    *EMEAN = EM * *EMEAN;
}

/* SFBRK, wam_source_module.f90:1966-2037, with CMPQB :2041-2107 (Battjes and
   Janssen's fraction of breaking waves by one Newton-Raphson step): SL +=
   SBR * F, FL += DSBR (IMPLSCH :640-642). */
static void sfbrk(const orc_ctx *c, orc_work *w, real EMEAN, real FMEAN, real DEPTH)
{
    int NB = c->KL * c->ML;
    real gd = GAMD * DEPTH;
    real BB = (R(8.) * EMEAN) / (gd * gd);
    /* CMPQB */
    real Q0 = R(0.);
    if (BB >= R(0.25)) {
        real q = R(2.) * SQRT(BB) - R(1.);
        Q0 = q * q;
    }
    real QB;
    if (BB < R(1.)) {
        real e = EXP((Q0 - R(1.)) / BB);
        QB = Q0 - (BB * (Q0 - e)) / (BB - e);
    } else {
        QB = R(1.);
    }
    QB = MINR(R(1.), QB);
    real SBR = -(R(2.) * FMEAN);                    /* -ALPHA*2.*FMEAN, ALPHA = 1 */
    if (BB <= R(1.)) SBR = (SBR * QB) / BB;
    real DSBR = R(0.);
    if (BB < R(1.) && ABS(BB - QB) > R(0.)) DSBR = (SBR * (R(1.) - QB)) / (BB - QB);
    for (int b = 0; b < NB; b++) {
        w->SL[b] = w->SL[b] + SBR * w->F[b];
        w->FL[b] = w->FL[b] + DSBR;
    }
}

/* SOURCE_PHILLIPS, wam_source_module.f90:3051-3160 (IMPLSCH :644-646):
   SL(K,M) += TEMP(M) * TPHOLD(K), the constants folded in REAL*4 as
   written. */
static void source_phillips(const orc_ctx *c, orc_work *w, real USTAR, real UDIR, real DEPTH, int indep)
{
    int KL = c->KL, ML = c->ML;
    real TPH[64];
    real CONST1 = R(28.) / (R(0.13) * c->G);
    real FPM = MAXR(R(0.0000001), CONST1 * USTAR);
    real XEPS = c->ROAIR / c->ROWATER;
    CONST1 = ((R(80.) * R(16.0)) * (XEPS * XEPS)) / ((R(0.5) * R(3.0)) * (c->G * c->G));
    for (int k = 0; k < KL; k++) TPH[k] = IPOW(USTAR * MAXR(R(0.), COS(c->TH[k] - UDIR)), 4);
    for (int m = 0; m < ML; m++) {
        real x = c->FR[m] * FPM;
        real TEMP = EXP(-(R(1.) / IPOW(x, 4)));     /* (FR*FPM)**(-4) */
        if (c->cfg.shallow) {
            real KD = MINR(TB(c->TFAK, m) * DEPTH, R(40.));
            TEMP = (CONST1 * TEMP) / (R(1.0) + (R(2.0) * KD) / SINH(R(2.) * KD));
        } else {
            TEMP = CONST1 * TEMP;
        }
        for (int k = 0; k < KL; k++) w->SL[k + KL * m] = w->SL[k + KL * m] + TEMP * TPH[k];
    }
}

/* One sea point through IMPLSCH (wam_source_module.f90:374-700). */
static void implsch_point(const orc_ctx *c, orc_work *w, real U10, real UDIR, real *TAUW,
                          real *USTAR, real *Z0, real ROAIRN, real WSTAR, real DEPTH, int indep,
                          int lcflx, real *phiaw_out, real *flux_out, int *mij_out)
{
    int KL = c->KL, ML = c->ML, NB = KL * ML, iphys = c->cfg.iphys;
    real SPRD[64], T[64];
    ORC_STAGE(0);   /* op-count stages: 0 up to SDISSIP, 1 SNONLIN to the update, 2 the new mean parameters and tail */
    for (int k = 0; k < KL; k++) SPRD[k] = IPOW(MAXR(R(0.), COS(c->TH[k] - UDIR)), 2);
    int JU = MINI(c->JUMAX, MAXI(NINT(U10 / c->DELU), 1));

    dirsum(c, w->F, NULL, T);
    real EMEAN = total_energy(c, T);
    if (c->wave_breaking) {
        sdepthlim(c, w, DEPTH, &EMEAN);
        dirsum(c, w->F, NULL, T);
    }
    real FMEAN = femean(c, T, EMEAN);
    real F1MEAN = R(1.) / tm1(c, T, EMEAN);
    real AKMEAN = R(0.), XKMEAN = R(0.);
    if (c->cfg.shallow) wm1_wm2(c, T, EMEAN, indep, &AKMEAN, &XKMEAN);
    if (iphys == 2) wavnu2(c, DEPTH, w->WN, w->CGG);

    real ust, z0 = *Z0, tauw = *TAUW, phiaw = R(0.);
    if (iphys == 2) w3flx4(c, U10, &ust, &z0);
    else airsea(c, U10, tauw, &ust, &z0);

    if (iphys == 2) sinput_st6(c, w, U10, ust, UDIR, ROAIRN, &tauw);
    else sinput(c, w, ust, UDIR, z0, ROAIRN, WSTAR, indep);

    dirsum(c, w->F, w->LLWS, T);
    real EMEANWS = total_energy(c, T);
    real FMEANWS = femean(c, T, EMEANWS);
    int MIJ = iphys == 2 ? frcutindex_st6(c, FMEAN, ust) : frcutindex(c, FMEAN, FMEANWS, ust);

    if (iphys == 2) {
        stresso(c, w, w->SPOS, ust, UDIR, z0, MIJ, indep, NULL, &phiaw);
        w3flx4(c, U10, &ust, &z0);
    } else {
        stresso(c, w, w->SPOS, ust, UDIR, z0, MIJ, indep, &tauw, &phiaw);
        airsea(c, U10, tauw, &ust, &z0);
    }
    imphftail(c, w, MIJ, indep);
    if (iphys != 2) sinput(c, w, ust, UDIR, z0, ROAIRN, WSTAR, indep);
    if (lcflx) for (int b = 0; b < NB; b++) w->SMIN[b] = w->SL[b] - w->SPOS[b];
    if (iphys == 2) stresso(c, w, w->SPOS, ust, UDIR, z0, MIJ, indep, NULL, &phiaw);
    else stresso(c, w, w->SPOS, ust, UDIR, z0, MIJ, indep, &tauw, &phiaw);

    if (iphys == 2) { sdissip_st6(c, w); w3swl6(c, w); }
    else sdissip(c, w, EMEAN, F1MEAN, XKMEAN, indep);

    ORC_STAGE(1);
    snonlin(c, w, DEPTH, AKMEAN);

    if (c->cfg.shallow) {
        for (int b = 0; b < NB; b++) w->SLBOT[b] = R(0.);
        sbottom(c, w, DEPTH, indep);
    }
    if (lcflx) {
        for (int b = 0; b < NB; b++) w->SMIN[b] = w->SL[b] - w->SMIN[b];   /* SSOURCE */
        wnfluxes(c, w, w->SMIN, ust, UDIR, phiaw, MIJ, indep, flux_out);
    }
    if (c->cfg.shallow) for (int b = 0; b < NB; b++) w->SL[b] = w->SL[b] + w->SLBOT[b];
    if (c->wave_breaking) sfbrk(c, w, EMEAN, FMEAN, DEPTH);
    if (c->phillips) source_phillips(c, w, ust, UDIR, DEPTH, indep);

    /* 4. limiter and implicit update */
    real DELT = (real)c->cfg.idelt;
    real c5g = R(5.0E-07) * c->G;
    real fmx = MAXR(FMEANWS, FMEAN);
    const treal *FLM = &c->FLMINFR[JU - 1];
    for (int m = 0; m < ML; m++) {
        real DELFL = (c5g / IPOW(c->FR[m], 4)) * DELT;
        real TEMP2 = (ust * DELFL) * fmx;
        real flmin = FLM[(size_t)m * c->JUMAX];
        for (int k = 0; k < KL; k++) {
            int b = k + KL * m;
            real TEMP1 = (DELT * w->SL[b]) / MAXR(R(1.), R(1.) - DELT * w->FL[b]);
            real TEMP3 = MINR(ABS(TEMP1), TEMP2);
            real f = w->F[b] + SIGNR(TEMP3, TEMP1);
            w->F[b] = MAXR(f, flmin * SPRD[k]);
        }
    }
    /* 5. new mean parameters, cut-off and tail */
    ORC_STAGE(2);
    dirsum(c, w->F, NULL, T);
    EMEAN = total_energy(c, T);
    FMEAN = femean(c, T, EMEAN);
    dirsum(c, w->F, w->LLWS, T);
    EMEANWS = total_energy(c, T);
    FMEANWS = femean(c, T, EMEANWS);
    MIJ = iphys == 2 ? frcutindex_st6(c, FMEAN, ust) : frcutindex(c, FMEAN, FMEANWS, ust);
    imphftail(c, w, MIJ, indep);

    *TAUW = tauw; *USTAR = ust; *Z0 = z0;
    *phiaw_out = phiaw;
    *mij_out = MIJ;
    ORC_STAGE(0);
}

/* Block driver: fl3[(m*KL+k)*NSEA + ij].  out_flux: [6][NSEA] if lcflx. */
void orc_implsch(const orc_ctx *c, real *fl3, const real *u10, const real *udir, real *tauw,
                 real *ustar, real *z0, const real *roairn, const real *wstar, const real *depth,
                 const int *indep, int lcflx, real *phiaw, real *out_flux, int *mij, int nthreads)
{
    int NSEA = c->NSEA, KL = c->KL, ML = c->ML, NB = KL * ML;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
    {
        orc_work w;
        real *buf = calloc((size_t)NB * 13 + 4 * 64, sizeof(real));
        w.F = buf; w.SL = buf + NB; w.FL = buf + 2 * NB; w.SPOS = buf + 3 * NB; w.SMIN = buf + 4 * NB;
        w.SLBOT = buf + 5 * NB; w.A = buf + 6 * NB; w.S = buf + 7 * NB; w.D = buf + 8 * NB;
        w.DPOS = buf + 9 * NB; w.KN = buf + 10 * NB; w.W = buf + 11 * NB; w.TEMP = buf + 12 * NB;
        w.TMP2 = buf + 12 * NB + 64; w.WN = buf + 12 * NB + 128; w.CGG = buf + 12 * NB + 192;
        w.LLWS = calloc(NB, 1);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int ij = 0; ij < NSEA; ij++) {
            for (int m = 0; m < ML; m++)
                for (int k = 0; k < KL; k++) w.F[k + KL * m] = fl3[((size_t)m * KL + k) * NSEA + ij];
            real fl[6] = {0};
            int id = (c->cfg.shallow && indep) ? indep[ij] : 1;
            implsch_point(c, &w, u10[ij], udir[ij], &tauw[ij], &ustar[ij], &z0[ij], roairn[ij],
                          wstar[ij], depth[ij], id, lcflx, &phiaw[ij], fl, &mij[ij]);
            if (lcflx && out_flux) for (int q = 0; q < 6; q++) out_flux[(size_t)q * NSEA + ij] = fl[q];
            for (int m = 0; m < ML; m++)
                for (int k = 0; k < KL; k++) fl3[((size_t)m * KL + k) * NSEA + ij] = w.F[k + KL * m];
        }
        free(buf); free(w.LLWS);
    }
}

/* WAVNU2 for every sea point (used by the tests to compare the cached
   device table): wn, cgg are [ML][NSEA]. */
void orc_wavnu2_all(const orc_ctx *c, const real *depth, real *wn, real *cgg)
{
    real WN[64], CG[64];
    for (int ij = 0; ij < c->NSEA; ij++) {
        wavnu2(c, depth[ij], WN, CG);
        for (int m = 0; m < c->ML; m++) { wn[(size_t)m * c->NSEA + ij] = WN[m]; cgg[(size_t)m * c->NSEA + ij] = CG[m]; }
    }
}

/* SNONLIN alone on one point's spectrum F[k + KL*m] (SL, FL accumulate). */
void orc_snonlin_point(const orc_ctx *c, real *F, real *SL, real *FL, real depth, real akmean)
{
    orc_work w;
    memset(&w, 0, sizeof(w));
    w.F = F; w.SL = SL; w.FL = FL;
    snonlin(c, &w, depth, akmean);
}
