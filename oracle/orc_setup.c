/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 *
 * Set-up part of the CPU restatement: constants, frequency/direction grid,
 * minimum-energy and shallow-water tables, DIA interaction tables, the
 * regular lat-lon sea-point grid with its neighbour tables, and the
 * propagation weights with the CFL check.  Each function names the reference
 * routine (file:line) whose statements it restates, in the same order and
 * with the same association of floating-point operations.
 */
#include <stdlib.h>
#include <string.h>
#include "orc_common.h"

#define M_DEGREE 360000   /* wam_coordinate_module.f90:41-43 */

static void *xcalloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s); if (!p) abort(); return p; }

/* Correctly rounded x**n of a compile-time constant (Fortran constant folding). */
static real fold_ipow(real x, int n) { return (real)pow((double)x, (double)n); }

/* wam_general_module.f90:41-85 and INITMDL :160 */
static void orc_constants(orc_ctx *c)
{
    c->G = R(9.806); c->PI = R(3.1415927); c->ZPI = R(2.) * c->PI;
    c->DEG = R(180.) / c->PI; c->RAD = c->PI / R(180.);
    c->CIRC = R(40000000.); c->RADIUS = c->CIRC / c->ZPI;
    c->ROAIR = R(1.225); c->RNUAIR = R(1.5E-5); c->RNUAIRM = R(0.11) * c->RNUAIR;
    c->ROWATER = R(1000.); c->XKAPPA = R(0.40); c->XNLEV = R(10.0); c->DKMAX = R(40.);
    c->BETAMAX = (real)c->cfg.betamax; c->ZALP = (real)c->cfg.zalp;
    c->ALPHA = (real)c->cfg.alpha; c->CDFAC = (real)c->cfg.cdfac;
    if (c->ML <= 30) c->ALPHA = R(0.0075);               /* initmdl.f90:160 */
    c->EMIN = R(1.0E-12);                                /* wam_interface_module.f90:44 */
    c->EPS1 = R(0.00001);                                /* wam_tables_module.f90:66 */
    c->CO = R(1.1);
}

/* SET_FRE_DIR + MAKE_FRE_DIR, wam_fre_dir_module.f90:246-251,297-461 */
static void orc_fre_dir(orc_ctx *c)
{
    int KL = c->KL, ML = c->ML, M, K;
    c->FR = xcalloc(ML, sizeof(real)); c->DF = xcalloc(ML, sizeof(real));
    c->DF_FR = xcalloc(ML, sizeof(real)); c->DF_FR2 = xcalloc(ML, sizeof(real));
    c->DFIM = xcalloc(ML, sizeof(real)); c->DFIMOFR = xcalloc(ML, sizeof(real));
    c->DFIM_FR = xcalloc(ML, sizeof(real)); c->DFIM_FR2 = xcalloc(ML, sizeof(real));
    c->FR5 = xcalloc(ML, sizeof(real)); c->FRM5 = xcalloc(ML, sizeof(real));
    c->RHOWG_DFIM = xcalloc(ML, sizeof(real)); c->GOM = xcalloc(ML, sizeof(real));
    c->C = xcalloc(ML, sizeof(real)); c->TH = xcalloc(KL, sizeof(real));
    c->COSTH = xcalloc(KL, sizeof(real)); c->SINTH = xcalloc(KL, sizeof(real));
    c->MPM = xcalloc(3 * ML, sizeof(int)); c->KPM = xcalloc(3 * KL, sizeof(int));
    c->JXO = xcalloc(2 * KL, sizeof(int)); c->JYO = xcalloc(2 * KL, sizeof(int));

    real CO = c->CO;
    int n = 1 - c->cfg.iref;
    real cof = n >= 0 ? IPOW(CO, n) : R(1.) / IPOW(CO, -n);
    c->FR[0] = cof * (real)c->cfg.fr1;
    c->INV_LOG_CO = R(1.) / LOG10(CO);
    for (M = 1; M < ML; M++) c->FR[M] = CO * c->FR[M - 1];
    for (M = 0; M < ML; M++) { c->FR5[M] = IPOW(c->FR[M], 5); c->FRM5[M] = R(1.) / c->FR5[M]; }
    for (M = 0; M < ML; M++) c->GOM[M] = c->G / ((R(2.) * c->ZPI) * c->FR[M]);
    for (M = 0; M < ML; M++) c->C[M] = c->G / (c->ZPI * c->FR[M]);
    c->DELTH = c->ZPI / (real)KL;
    c->DELTR = c->DELTH * c->RADIUS;
    for (K = 0; K < KL; K++) c->TH[K] = (real)K * c->DELTH + R(0.5) * c->DELTH;
    for (K = 0; K < KL; K++) { c->COSTH[K] = COS(c->TH[K]); c->SINTH[K] = SIN(c->TH[K]); }
    real CO1 = R(0.5) * (CO - R(1.));
    c->DF[0] = CO1 * c->FR[0];
    for (M = 1; M < ML - 1; M++) c->DF[M] = CO1 * (c->FR[M] + c->FR[M - 1]);
    c->DF[ML - 1] = CO1 * c->FR[ML - 2];
    for (M = 0; M < ML; M++) {
        c->DF_FR[M] = c->DF[M] * c->FR[M];
        c->DF_FR2[M] = c->DF_FR[M] * c->FR[M];
        c->DFIM[M] = c->DF[M] * c->DELTH;
        c->DFIMOFR[M] = c->DFIM[M] / c->FR[M];
        c->DFIM_FR[M] = c->DF_FR[M] * c->DELTH;
        c->DFIM_FR2[M] = c->DF_FR2[M] * c->DELTH;
        c->RHOWG_DFIM[M] = ((c->ROWATER * c->G) * c->DELTH * LOG(CO)) * c->FR[M];
    }
    c->RHOWG_DFIM[0] = R(0.5) * c->RHOWG_DFIM[0];
    c->RHOWG_DFIM[ML - 1] = R(0.5) * c->RHOWG_DFIM[ML - 1];
    real FRML = c->FR[ML - 1];
    c->MO_TAIL = -((c->DELTH / R(-4.)) * FRML);
    c->MM1_TAIL = -(c->DELTH / R(-5.));
    c->MP1_TAIL = -((c->DELTH / R(-3.)) * IPOW(FRML, 2));
    c->MP2_TAIL = -((c->DELTH / R(-2.)) * IPOW(FRML, 3));
    c->FMIN = fold_ipow(R(0.07), 2) / ((R(16.) * (FRML - c->FR[0])) * c->ZPI);
    for (M = 1; M <= ML; M++) {
        c->MPM[(M - 1) * 3 + 0] = MAXI(1, M - 1);
        c->MPM[(M - 1) * 3 + 1] = M;
        c->MPM[(M - 1) * 3 + 2] = MINI(ML, M + 1);
    }
    for (K = 1; K <= KL; K++) {
        int km = K - 1; if (km < 1) km = KL;
        int kp = K + 1; if (kp > KL) kp = 1;
        c->KPM[(K - 1) * 3 + 0] = km; c->KPM[(K - 1) * 3 + 1] = K; c->KPM[(K - 1) * 3 + 2] = kp;
        if (c->COSTH[K - 1] >= R(0.)) { c->JYO[(K - 1) * 2] = 1; c->JYO[(K - 1) * 2 + 1] = 2; }
        else { c->JYO[(K - 1) * 2] = 2; c->JYO[(K - 1) * 2 + 1] = 1; }
        if (c->SINTH[K - 1] >= R(0.)) { c->JXO[(K - 1) * 2] = 1; c->JXO[(K - 1) * 2 + 1] = 2; }
        else { c->JXO[(K - 1) * 2] = 2; c->JXO[(K - 1) * 2 + 1] = 1; }
    }
}

/* AKI, wam_general_module.f90:301-364 */
real orc_aki(const orc_ctx *c, real OM, real BETA)
{
    real G = c->G;
    real AKI = MAXR((OM * OM) / (R(4.) * G), OM / (R(2.) * SQRT(G * BETA)));
    real AKP = R(10000.);
    while (ABS(AKP - AKI) > R(0.0001) * AKI) {
        real BO = BETA * AKI;
        if (BO > c->DKMAX) { AKI = (OM * OM) / G; break; }
        AKP = AKI;
        real TH = (G * AKI) * TANH(BO);
        real STH = SQRT(TH);
        real ch = COSH(BO);
        AKI = AKI + (((OM - STH) * STH) * R(2.)) / (TH / AKI + (G * BO) / (ch * ch));
    }
    return AKI;
}

/* MIN_ENERGY + FETCH_LAW + JONSWAP, wam_tables_module.f90:315-385,
   wam_jonswap_module.f90:63-213 */
static void orc_min_energy(orc_ctx *c)
{
    int JUMAX = 300, J, M, ML = c->ML;
    c->JUMAX = JUMAX; c->UMAX = R(75.); c->FLMIN = R(0.000001);
    c->DELU = c->UMAX / (real)JUMAX;
    real G = c->G;
    real FETCH = MINR(R(0.5) * c->DELPHI, R(25000.));
    real FPMAX = c->FR[ML - 1];
    real *U10 = xcalloc(JUMAX, sizeof(real)), *FP = xcalloc(JUMAX, sizeof(real)),
         *AL = xcalloc(JUMAX, sizeof(real));
    const real A = R(2.84), D = -(R(3.) / R(10.)), B = R(0.033), E = R(2.) / R(3.);
    for (J = 0; J < JUMAX; J++) {
        U10[J] = (real)(J + 1) * c->DELU;
        if (U10[J] > R(0.1E-08)) {
            real UG = G / U10[J];
            real fp = MAXR(R(0.13), A * POW((G * FETCH) / (U10[J] * U10[J]), D));
            fp = MINR(fp, FPMAX / UG);
            AL[J] = MAXR(R(0.0081), B * POW(fp, E));
            FP[J] = fp * UG;
        } else { AL[J] = R(0.0081); FP[J] = FPMAX; }
        AL[J] = R(0.01) * AL[J];
    }
    c->FLMINFR = xcalloc((size_t)JUMAX * ML, sizeof(real));
    const real GAMMA = R(3.0), SA = R(7.0E-02), SB = R(9.0E-02);
    real G2ZPI4 = (G * G) / fold_ipow(c->ZPI, 4);
    real LOG_GAMMA = LOG(GAMMA);
    real SAA = R(0.5) / (SA * SA), SBB = R(0.5) / (SB * SB);
    for (M = 0; M < ML; M++) {
        real FRH = c->FR[M];
        real g2 = G2ZPI4 * (R(1.) / IPOW(FRH, 5));
        for (J = 0; J < JUMAX; J++) {
            real ET;
            real ARG = R(1.25) * IPOW(FP[J] / FRH, 4);
            if (ARG < R(50.)) {
                ET = (AL[J] * g2) * EXP(-ARG);
                if (FRH > FP[J]) ARG = SBB * IPOW(FRH / FP[J] - R(1.), 2);
                else ARG = SAA * IPOW(FRH / FP[J] - R(1.), 2);
                if (ARG < R(99.)) ET = ET * EXP(LOG_GAMMA * EXP(-ARG));
            } else ET = R(0.);
            c->FLMINFR[(size_t)M * JUMAX + J] = MAXR(ET, c->FLMIN);
        }
    }
    free(U10); free(FP); free(AL);
}

/* MAKE_SHALLOW_TABLES, wam_tables_module.f90:388-466 */
static void orc_shallow_tables(orc_ctx *c)
{
    int ND = c->cfg.ndepth, ML = c->ML, M, JD;
    c->DEPTHA = (real)c->cfg.deptha; c->DEPTHD = (real)c->cfg.depthd;
    size_t n = (size_t)ND * ML;
    c->TFAK = xcalloc(n, sizeof(real)); c->TCGOND = xcalloc(n, sizeof(real));
    c->TSIHKD = xcalloc(n, sizeof(real)); c->T_TAIL = xcalloc(n, sizeof(real));
    real G = c->G, GH = G / (R(4.) * c->PI);
    for (M = 0; M < ML; M++) {
        real OM = c->ZPI * c->FR[M];
        for (JD = 1; JD <= ND; JD++) {
            size_t i = (size_t)M * ND + (JD - 1);
            real AD = c->DEPTHA * IPOW(c->DEPTHD, JD - 1);
            real AK = orc_aki(c, OM, AD);
            c->TFAK[i] = AK;
            real AKD = AK * AD;
            if (AKD <= R(10.0)) {
                c->TCGOND[i] = (R(0.5) * SQRT((G * TANH(AKD)) / AK)) * (R(1.0) + (R(2.0) * AKD) / SINH(R(2.) * AKD));
                c->TSIHKD[i] = OM / SINH(R(2.) * AKD);
            } else {
                c->TCGOND[i] = GH / c->FR[M];
                c->TSIHKD[i] = R(0.);
            }
            c->T_TAIL[i] = R(1.) / (IPOW(c->TFAK[i], 3) * c->TCGOND[i]);
        }
    }
}

/* JAFU, src/chief/jafu.f90 */
static int jafu(real CL, int J, int IAN)
{
    int r = J + INTR(CL);
    if (r <= 0) r += IAN;
    if (r > IAN) r -= IAN;
    return r;
}

/* NLWEIGT, wam_source_module.f90:1056-1287 */
static void orc_nlweigt(orc_ctx *c)
{
    int KL = c->KL, ML = c->ML;
    const real ALAMD = R(0.25), CON = R(3000.);
    real CO = c->CO;
    real F1P1 = LOG10(CO);
    c->ISP = INTR(LOG10(R(1.) + ALAMD) / F1P1 + R(.000001));
    c->ISM = (int)FLOORR(LOG10(R(1.) - ALAMD) / F1P1 + R(.0000001));
    c->MFRSTLW = 1 + c->ISM;
    c->MLSTHG = ML - c->ISM;
    c->KFRH = -c->ISM + c->ISP + 2;
    int MF = c->MFRSTLW, MLH = c->MLSTHG, NM = MLH - MF + 1;
    c->IKP = xcalloc(NM, sizeof(int)); c->IKP1 = xcalloc(NM, sizeof(int));
    c->IKM = xcalloc(NM, sizeof(int)); c->IKM1 = xcalloc(NM, sizeof(int));
    c->AF11 = xcalloc(NM, sizeof(real)); c->FKLAP = xcalloc(NM, sizeof(real));
    c->FKLAP1 = xcalloc(NM, sizeof(real)); c->FKLAM = xcalloc(NM, sizeof(real));
    c->FKLAM1 = xcalloc(NM, sizeof(real)); c->FRH = xcalloc(c->KFRH, sizeof(real));
    c->K1W = xcalloc(2 * KL, sizeof(int)); c->K2W = xcalloc(2 * KL, sizeof(int));
    c->K11W = xcalloc(2 * KL, sizeof(int)); c->K21W = xcalloc(2 * KL, sizeof(int));
    int *JA1 = xcalloc(2 * KL, sizeof(int)), *JA2 = xcalloc(2 * KL, sizeof(int));

    /* 1. angular grid; constant expressions folded with correct rounding */
    real t = (R(1.) + ALAMD) / (R(1.) - ALAMD);
    real XF = fold_ipow(t, 4);
    real num = (R(1.) + R(2.) * ALAMD) + R(2.) * fold_ipow(ALAMD, 3);
    real COSTH3 = num / fold_ipow(R(1.) + ALAMD, 2);
    real DELPHI1 = (-(R(180.) / c->PI)) * ACOS(COSTH3);
    real COSTH4 = SQRT((R(1.) - XF) + XF * IPOW(COSTH3, 2));
    real DELPHI2 = (R(180.) / c->PI) * ACOS(COSTH4);
    real DELTHA = c->DELTH * c->DEG;
    real CL1 = DELPHI1 / DELTHA, CL2 = DELPHI2 / DELTHA;

    int KLP1 = KL + 1, IC = 1, KH, K;
    for (KH = 1; KH <= 2; KH++) {
        int KLH = (KH == 2) ? KLP1 : KL;
        for (K = 1; K <= KLH; K++) {
            int KS = K;
            if (KH > 1) KS = KLP1 - K + 1;
            if (KS > KL) continue;
            JA1[(KH - 1) * KL + KS - 1] = jafu((real)IC * CL1, K, KL);
            JA2[(KH - 1) * KL + KS - 1] = jafu((real)IC * CL2, K, KL);
        }
        IC = -1;
    }
    CL1 = CL1 - (real)INTR(CL1);
    CL2 = CL2 - (real)INTR(CL2);
    c->ACL1 = ABS(CL1); c->ACL2 = ABS(CL2);
    c->CL11 = R(1.) - c->ACL1; c->CL21 = R(1.) - c->ACL2;
    real AL11 = fold_ipow(R(1.) + ALAMD, 4), AL12 = fold_ipow(R(1.) - ALAMD, 4);
    c->DAL1 = R(1.) / AL11; c->DAL2 = R(1.) / AL12;

    int ISG = 1;
    for (KH = 1; KH <= 2; KH++) {
        real CL1H = (real)ISG * CL1, CL2H = (real)ISG * CL2;
        for (K = 1; K <= KL; K++) {
            int KS = K;
            if (KH == 2) KS = KL - K + 2;
            if (K == 1) KS = 1;
            int K1 = JA1[(KH - 1) * KL + K - 1], K11, K21;
            c->K1W[(KH - 1) * KL + KS - 1] = K1;
            if (CL1H < R(0.)) { K11 = K1 - 1; if (K11 < 1) K11 = KL; }
            else { K11 = K1 + 1; if (K11 > KL) K11 = 1; }
            c->K11W[(KH - 1) * KL + KS - 1] = K11;
            int K2 = JA2[(KH - 1) * KL + K - 1];
            c->K2W[(KH - 1) * KL + KS - 1] = K2;
            if (CL2H < R(0.)) { K21 = K2 - 1; if (K21 < 1) K21 = KL; }
            else { K21 = K2 + 1; if (K21 > KL) K21 = 1; }
            c->K21W[(KH - 1) * KL + KS - 1] = K21;
        }
        ISG = -1;
    }

    /* 2. frequency grid, FRLON(MFRSTLW:ML+KFRH) */
    int NLON = ML + c->KFRH - MF + 1, M;
    real *FRLON = xcalloc(NLON, sizeof(real));
#define FRL(m) FRLON[(m) - MF]
    for (M = 1; M <= ML; M++) FRL(M) = c->FR[M - 1];
    for (M = 0; M >= MF; M--) FRL(M) = FRL(M + 1) / CO;
    for (M = ML + 1; M <= ML + c->KFRH; M++) FRL(M) = CO * FRL(M - 1);
    for (M = MF; M <= MLH; M++) {
        int i = M - MF;
        real FRG = FRL(M);
        c->AF11[i] = CON * IPOW(FRG, 11);
        real FLP = FRG * (R(1.) + ALAMD);
        real FLM = FRG * (R(1.) - ALAMD);
        int IKN = M + c->ISP;
        c->IKP[i] = IKN;
        real FKP = FRL(c->IKP[i]);
        c->IKP1[i] = c->IKP[i] + 1;
        c->FKLAP[i] = (FLP - FKP) / (FRL(c->IKP1[i]) - FKP);
        c->FKLAP1[i] = R(1.) - c->FKLAP[i];
        if (FRL(MF) >= FLM) {
            c->IKM[i] = 1; c->IKM1[i] = 1; c->FKLAM[i] = R(0.); c->FKLAM1[i] = R(0.);
        } else {
            IKN = M + c->ISM;
            c->IKM[i] = IKN;
            real FKM = FRL(c->IKM[i]);
            c->IKM1[i] = c->IKM[i] + 1;
            c->FKLAM[i] = (FLM - FKM) / (FRL(c->IKM1[i]) - FKM);
            c->FKLAM1[i] = R(1.) - c->FKLAM[i];
            if (IKN < MF) { c->IKM[i] = 1; c->FKLAM1[i] = R(0.); }
        }
    }
    /* 3. tail frequency ratios */
    for (int I = 1; I <= c->KFRH; I++) {
        M = ML + I - 1;
        c->FRH[I - 1] = IPOW(FRL(ML) / FRL(M), 5);
    }
#undef FRL
    free(FRLON); free(JA1); free(JA2);
}

/* INIT_SNONLIN, wam_source_module.f90:1291-1523 */
static void orc_init_snonlin(orc_ctx *c)
{
    int MF = c->MFRSTLW, MLH = c->MLSTHG, ML = c->ML, MC;
    c->FTRF = xcalloc(2 - MF, sizeof(real));          /* FTRF(MFRSTLW:1) */
#define FTR(m) c->FTRF[(m) - MF]
#define EPMMA(X) (EXP(-MINR(R(1.25) * IPOW(X, 4), R(50.))) * IPOW(X, 5))
    real ALPH = R(1.) / EPMMA(R(1.));
    real FRR = R(1.);
    for (MC = 1; MC >= MF; MC--) { FTR(MC) = ALPH * EPMMA(FRR); FRR = FRR * c->CO; }
    c->INLCOEF = xcalloc(5 * MLH, sizeof(int));
    c->RNLCOEF = xcalloc(25 * MLH, sizeof(real));
    for (MC = 1; MC <= MLH; MC++) {
        int i = MC - MF;
        int MP = c->IKP[i], MP1 = c->IKP1[i], MM = c->IKM[i], MM1 = c->IKM1[i];
        real FFACP = R(1.), FFACP1 = R(1.), FFACM = R(1.), FFACM1 = R(1.), FTAIL = R(1.);
        int IC = MC, IP = MP, IP1 = MP1, IM = MM, IM1 = MM1;
        if (IM < 1) {
            FFACM = FTR(IM); IM = 1;
            if (IM1 < 1) { FFACM1 = FTR(IM1); IM1 = 1; }
        }
        if (IP1 > ML) {
            int ITEMP = IP1 - ML + 1;
            if (ITEMP > c->KFRH) ITEMP = c->KFRH;
            FFACP1 = c->FRH[ITEMP - 1];
            IP1 = ML;
            if (IP > ML) {
                FFACP = c->FRH[IP - ML + 1 - 1]; IP = ML;
                if (IC > ML) {
                    FTAIL = c->FRH[IC - ML + 1 - 1]; IC = ML;
                    if (IM1 > ML) { FFACM1 = c->FRH[IM1 - ML + 1 - 1]; IM1 = ML; }
                }
            }
        }
        int *ic = &c->INLCOEF[(MC - 1) * 5];
        ic[0] = IC; ic[1] = IP; ic[2] = IP1; ic[3] = IM; ic[4] = IM1;
        real *rc = &c->RNLCOEF[(MC - 1) * 25];
        real FKLAMP = c->FKLAP[i], FKLAMP1 = c->FKLAP1[i];
        real GW2 = (FKLAMP1 * FFACP) * c->DAL1;
        real GW1 = GW2 * c->CL11;
        GW2 = GW2 * c->ACL1;
        real GW4 = (FKLAMP * FFACP1) * c->DAL1;
        real GW3 = GW4 * c->CL11;
        GW4 = GW4 * c->ACL1;
        real FKLAMPA = FKLAMP * c->CL11, FKLAMPB = FKLAMP * c->ACL1;
        real FKLAMP2 = FKLAMP1 * c->ACL1;
        FKLAMP1 = FKLAMP1 * c->CL11;
        rc[0] = FTAIL; rc[1] = GW1; rc[2] = GW2; rc[3] = GW3; rc[4] = GW4;
        rc[5] = FKLAMPA; rc[6] = FKLAMPB; rc[7] = FKLAMP2; rc[8] = FKLAMP1;
        rc[9] = FKLAMPA * FKLAMPA; rc[10] = FKLAMPB * FKLAMPB;
        rc[11] = FKLAMP1 * FKLAMP1; rc[12] = FKLAMP2 * FKLAMP2;
        real FKLAMM = c->FKLAM[i], FKLAMM1 = c->FKLAM1[i];
        real GW6 = (FKLAMM1 * FFACM) * c->DAL2;
        real GW5 = GW6 * c->CL21;
        GW6 = GW6 * c->ACL2;
        real GW8 = (FKLAMM * FFACM1) * c->DAL2;
        real GW7 = GW8 * c->CL21;
        GW8 = GW8 * c->ACL2;
        real FKLAMMA = FKLAMM * c->CL21, FKLAMMB = FKLAMM * c->ACL2;
        real FKLAMM2 = FKLAMM1 * c->ACL2;
        FKLAMM1 = FKLAMM1 * c->CL21;
        rc[13] = GW5; rc[14] = GW6; rc[15] = GW7; rc[16] = GW8;
        rc[17] = FKLAMMA; rc[18] = FKLAMMB; rc[19] = FKLAMM2; rc[20] = FKLAMM1;
        rc[21] = FKLAMMA * FKLAMMA; rc[22] = FKLAMMB * FKLAMMB;
        rc[23] = FKLAMM1 * FKLAMM1; rc[24] = FKLAMM2 * FKLAMM2;
    }
#undef FTR
#undef EPMMA
}

/* INIT_X0TAUHF, wam_source_module.f90:3784-3849 */
static void orc_init_x0tauhf(orc_ctx *c)
{
    real X0 = R(0.005);
    for (int J = 1; J <= 30; J++) {
        real FF = EXP(c->XKAPPA / (X0 + c->ZALP));
        real F = (c->ALPHA * (X0 * X0)) * FF - R(1.0);
        if (F == R(0.0)) break;
        real DF = (c->ALPHA * FF) * (R(2.0) * X0 - c->XKAPPA * IPOW(X0 / (X0 + c->ZALP), 2));
        X0 = X0 - F / DF;
    }
    c->X0TAUHF = X0;
    real CONST1 = (c->BETAMAX / (c->XKAPPA * c->XKAPPA)) / R(3.0);
    c->WTAUHF[0] = CONST1;
    for (int J = 2; J <= 18; J += 2) { c->WTAUHF[J - 1] = R(4.0) * CONST1; c->WTAUHF[J] = R(2.0) * CONST1; }
    c->WTAUHF[18] = CONST1;
}

/* PREPARE_GRID (regular grid part) + CLOSEST_GP_LAT, preproc_module.f90:2043-2381,
   1318-1488; DCO/DPSN wam_initial_module.f90:205-214.  Sea points are numbered
   row-major from the south-west corner; 0 marks a land neighbour. */
static int orc_grid(orc_ctx *c, const uint8_t *mask)
{
    const orc_config *g = &c->cfg;
    int NX = g->nx, NY = g->ny, I, K, IP;
    c->SINPH = xcalloc(NY, sizeof(real)); c->COSPH = xcalloc(NY, sizeof(real));
    c->DELLAM = xcalloc(NY, sizeof(real)); c->ZDELLO = xcalloc(NY, sizeof(int));
    for (K = 1; K <= NY; K++) {
        real XLAT = ((real)(g->amosop + (int64_t)(K - 1) * g->xdella) / (real)M_DEGREE) * c->RAD;
        c->SINPH[K - 1] = SIN(XLAT); c->COSPH[K - 1] = COS(XLAT);
    }
    c->DELPHI = (((real)g->xdella / (real)M_DEGREE) * c->CIRC) / R(360.);
    for (K = 0; K < NY; K++) {
        if (g->iper) c->ZDELLO[K] = NINT((real)(g->amoeap - g->amowep + g->xdello) / (real)NX);
        else c->ZDELLO[K] = NINT((real)(g->amoeap - g->amowep) / (real)(NX - 1));
        c->DELLAM[K] = (((real)c->ZDELLO[K] / (real)M_DEGREE) * c->CIRC) / R(360.);
    }
    int NSEA = 0;
    for (K = 0; K < NX * NY; K++) NSEA += mask[K] ? 1 : 0;
    c->NSEA = NSEA;
    if (NSEA <= 0) return 1;
    c->IXLG = xcalloc(NSEA, sizeof(int)); c->KXLT = xcalloc(NSEA, sizeof(int));
    IP = 0;
    for (K = 1; K <= NY; K++)
        for (I = 1; I <= NX; I++)
            if (mask[(K - 1) * NX + (I - 1)]) { c->IXLG[IP] = I; c->KXLT[IP] = K; IP++; }
#define MASK(i, k) (mask[((k) - 1) * NX + ((i) - 1)] != 0)
    c->KLAT = xcalloc(4 * (size_t)NSEA, sizeof(int));
    c->KLON = xcalloc(2 * (size_t)NSEA, sizeof(int));
    c->WLAT = xcalloc(2 * (size_t)NSEA, sizeof(real));
#define KLAT_(ij, ic, icl) c->KLAT[((size_t)((icl) - 1) * 2 + ((ic) - 1)) * NSEA + (ij) - 1]
#define KLON_(ij, ic) c->KLON[(size_t)((ic) - 1) * NSEA + (ij) - 1]
    for (IP = 1; IP <= NSEA; IP++) {
        I = c->IXLG[IP - 1]; K = c->KXLT[IP - 1];
        if (I > 1) { if (MASK(I - 1, K)) KLON_(IP, 1) = IP - 1; }
        else if (g->iper && MASK(NX, K)) {
            int cnt = 0; for (int ii = 2; ii <= NX; ii++) cnt += MASK(ii, K);
            KLON_(IP, 1) = IP + cnt;
        }
        if (I < NX) { if (MASK(I + 1, K)) KLON_(IP, 2) = IP + 1; }
        else if (g->iper && MASK(1, K)) {
            int cnt = 0; for (int ii = 1; ii <= NX; ii++) cnt += MASK(ii, K);
            KLON_(IP, 2) = IP + 1 - cnt;
        }
        for (int side = 1; side <= 2; side++) {
            int KH = side == 1 ? K - 1 : K + 1;
            int IP1 = 0;
            if (KH >= 1 && KH <= NY) {
                /* regular grid: same longitude index on the neighbouring row */
                int IPH1 = I;
                if (MASK(IPH1, KH)) {
                    int ICS = side == 1 ? -1 : 1, ICE = side == 1 ? 1 : NSEA, IC;
                    for (IC = IP; IC != ICE + ICS; IC += ICS)
                        if (c->IXLG[IC - 1] == IPH1 && c->KXLT[IC - 1] == KH) break;
                    IP1 = IC;
                }
            }
            KLAT_(IP, side, 1) = IP1; KLAT_(IP, side, 2) = IP1;
            c->WLAT[(size_t)(side - 1) * NSEA + IP - 1] = R(1.);
        }
    }
    /* DCO(0) is the land slot (never used as a divisor for a sea point) */
    c->DCO = xcalloc((size_t)NSEA + 1, sizeof(real));
    c->DPSN = xcalloc(2 * (size_t)NSEA, sizeof(real));
    for (IP = 1; IP <= NSEA; IP++) c->DCO[IP] = R(1.) / c->COSPH[c->KXLT[IP - 1] - 1];
    for (IP = 1; IP <= NSEA; IP++) {
        c->DPSN[IP - 1] = R(1.); c->DPSN[(size_t)NSEA + IP - 1] = R(1.);
        if (g->spherical) {
            if (KLAT_(IP, 1, 1) != 0) c->DPSN[IP - 1] = c->DCO[IP] / c->DCO[KLAT_(IP, 1, 1)];
            if (KLAT_(IP, 2, 1) != 0) c->DPSN[(size_t)NSEA + IP - 1] = c->DCO[IP] / c->DCO[KLAT_(IP, 2, 1)];
        }
    }
#undef MASK
    return 0;
}

/* GRADIENT (regular grid), wam_propagation_module.f90:182-275.  f[0] = land slot. */
static void orc_gradient(const orc_ctx *c, const real *f, real *dlat, real *dlon)
{
    int NSEA = c->NSEA;
    for (int IJ = 1; IJ <= NSEA; IJ++) {
        int IP = KLAT_(IJ, 2, 1), IM = KLAT_(IJ, 1, 1);
        dlat[IJ - 1] = (IP != 0 && IM != 0) ? (f[IP] - f[IM]) / (R(2.) * c->DELPHI) : R(0.0);
        IP = KLON_(IJ, 2); IM = KLON_(IJ, 1);
        int KX = c->KXLT[IJ - 1];
        dlon[IJ - 1] = (IP != 0 && IM != 0) ? (f[IP] - f[IM]) / (R(2.) * c->DELLAM[KX - 1]) : R(0.0);
    }
}

#define COEFP(p, ij, k, m) c->COEF[(((size_t)(p) * ML + (m)) * KL + (k)) * NSEA + (ij)]

/* PREPARE_PROPAGATION + PREPARE_PROPAGATION_COEF (PREP_SPHER_DEEP :1248-1327,
   PREP_SPHER_SHALLOW :1336-1485, PREP_SPHER_DEEP_CURR_REG :1591-1676,
   PREP_SPHER_SHALLOW_CURR_REG :1808-1911) + CHECK_CFL :1917-2138.
   depth/u/v are per sea point (1..NSEA at index 0..NSEA-1).  The land slot
   carries U = V = 0 (the reference reads U(NINF-1) outside its allocation,
   wam_initial_module.f90:187 vs wam_propagation_module.f90:1855). */
int orc_prepare_propagation(orc_ctx *c, const real *depth, const real *u, const real *v)
{
    int NSEA = c->NSEA, KL = c->KL, ML = c->ML, ND = c->cfg.ndepth, IJ, K, M;
    const orc_config *g = &c->cfg;
    if (!g->spherical) return 2;                      /* Cartesian variants not restated */
    real *D0 = xcalloc((size_t)NSEA + 1, sizeof(real));
    real *U0 = xcalloc((size_t)NSEA + 1, sizeof(real));
    real *V0 = xcalloc((size_t)NSEA + 1, sizeof(real));
    for (IJ = 1; IJ <= NSEA; IJ++) { D0[IJ] = depth[IJ - 1]; U0[IJ] = u ? u[IJ - 1] : R(0.); V0[IJ] = v ? v[IJ - 1] : R(0.); }
    free(c->INDEP); c->INDEP = xcalloc(NSEA, sizeof(int));
    real *CGOND = NULL;
#define TAB(t, jd, m) (t)[(size_t)(m) * ND + (jd) - 1]
    if (g->shallow) {
        for (IJ = 1; IJ <= NSEA; IJ++) {
            int id = NINT(LOG(D0[IJ] / c->DEPTHA) / LOG(c->DEPTHD) + R(1.));
            c->INDEP[IJ - 1] = MINI(MAXI(id, 1), ND);
        }
        CGOND = xcalloc((size_t)(NSEA + 1) * ML, sizeof(real));   /* [M][0..NSEA] */
        for (M = 0; M < ML; M++) {
            CGOND[(size_t)M * (NSEA + 1)] = TAB(c->TCGOND, ND, M);
            for (IJ = 1; IJ <= NSEA; IJ++) CGOND[(size_t)M * (NSEA + 1) + IJ] = TAB(c->TCGOND, c->INDEP[IJ - 1], M);
        }
    }
#define CG(ij, m) CGOND[(size_t)(m) * (NSEA + 1) + (ij)]
    int refr = g->refr_d || g->refr_c;
    real *DDPHI = NULL, *DDLAM = NULL, *DUPHI = NULL, *DULAM = NULL, *DVPHI = NULL, *DVLAM = NULL;
    int *INDEP_G = NULL;
    if (refr && g->shallow) {
        DDPHI = xcalloc(NSEA, sizeof(real)); DDLAM = xcalloc(NSEA, sizeof(real));
        orc_gradient(c, D0, DDPHI, DDLAM);
        for (IJ = 0; IJ < NSEA; IJ++) DDLAM[IJ] = DDLAM[IJ] * c->DCO[IJ + 1];
    }
    if (g->refr_c) {
        DUPHI = xcalloc(NSEA, sizeof(real)); DULAM = xcalloc(NSEA, sizeof(real));
        DVPHI = xcalloc(NSEA, sizeof(real)); DVLAM = xcalloc(NSEA, sizeof(real));
        orc_gradient(c, U0, DUPHI, DULAM);
        orc_gradient(c, V0, DVPHI, DVLAM);
        for (IJ = 0; IJ < NSEA; IJ++) { DULAM[IJ] = DULAM[IJ] * c->DCO[IJ + 1]; DVLAM[IJ] = DVLAM[IJ] * c->DCO[IJ + 1]; }
    }
    if (refr && g->shallow) {
        INDEP_G = xcalloc(NSEA, sizeof(int));
        for (IJ = 1; IJ <= NSEA; IJ++) {
            real D_MEAN = R(4.) * D0[IJ]; int IZ = 4;
            int IP = KLAT_(IJ, 2, 1), IM = KLAT_(IJ, 1, 1);
            if (IP != 0 && IM != 0) { D_MEAN = D_MEAN + D0[IP] + D0[IM]; IZ += 2; }
            IP = KLON_(IJ, 2); IM = KLON_(IJ, 1);
            if (IP != 0 && IM != 0) { D_MEAN = D_MEAN + D0[IP] + D0[IM]; IZ += 2; }
            D_MEAN = D_MEAN / (real)IZ;
            IZ = NINT(LOG(D_MEAN / c->DEPTHA) / LOG(c->DEPTHD) + R(1.));
            IZ = MAXI(IZ, 1);
            INDEP_G[IJ - 1] = MINI(IZ, ND);
        }
    }
    /* THDD / SIDC: [K][M][NSEA] (only the M=ML plane when deep) */
    real *THDD = NULL, *SIDC = NULL;
    if (refr) {
        THDD = xcalloc((size_t)NSEA * KL * ML, sizeof(real));
        SIDC = xcalloc((size_t)NSEA * KL * ML, sizeof(real));
    }
#define THD(ij, k, m) THDD[(((size_t)(m) * KL) + (k)) * NSEA + (ij)]
#define SID(ij, k, m) SIDC[(((size_t)(m) * KL) + (k)) * NSEA + (ij)]
    real *TEMP = xcalloc(NSEA, sizeof(real));
    if (refr && g->shallow) {
        for (K = 0; K < KL; K++) {
            for (IJ = 0; IJ < NSEA; IJ++) TEMP[IJ] = c->SINTH[K] * DDPHI[IJ] - c->COSTH[K] * DDLAM[IJ];
            for (M = 0; M < ML; M++)
                for (IJ = 0; IJ < NSEA; IJ++) THD(IJ, K, M) = TEMP[IJ] * TAB(c->TSIHKD, INDEP_G[IJ], M);
        }
    }
    if (g->refr_c && g->shallow)
        for (IJ = 0; IJ < NSEA; IJ++) TEMP[IJ] = V0[IJ + 1] * DDPHI[IJ] + U0[IJ + 1] * DDLAM[IJ];
    if (g->refr_c) {
        for (K = 0; K < KL; K++) {
            real SS = c->SINTH[K] * c->SINTH[K], SC = c->SINTH[K] * c->COSTH[K], CC = c->COSTH[K] * c->COSTH[K];
            for (IJ = 0; IJ < NSEA; IJ++) {
                real THDC = ((SS * DUPHI[IJ] + SC * DVPHI[IJ]) - SC * DULAM[IJ]) - CC * DVLAM[IJ];
                if (g->shallow) { for (M = 0; M < ML; M++) THD(IJ, K, M) = THD(IJ, K, M) + THDC; }
                else THD(IJ, K, ML - 1) = THDC;
            }
        }
        for (K = 0; K < KL; K++) {
            real SD = c->SINTH[K], CD = c->COSTH[K];
            real SS = SD * SD, SC = SD * CD, CC = CD * CD;
            for (IJ = 0; IJ < NSEA; IJ++)
                SID(IJ, K, ML - 1) = (((-(SC * DUPHI[IJ])) - CC * DVPHI[IJ]) - SS * DULAM[IJ]) - SC * DVLAM[IJ];
        }
        if (g->shallow) {
            for (M = 0; M < ML; M++)
                for (K = 0; K < KL; K++)
                    for (IJ = 0; IJ < NSEA; IJ++)
                        SID(IJ, K, M) = (SID(IJ, K, ML - 1) * CG(IJ + 1, M) + TEMP[IJ] * TAB(c->TSIHKD, INDEP_G[IJ], M))
                                        * TAB(c->TFAK, INDEP_G[IJ], M);
        }
    }

    /* coefficients */
    free(c->COEF);
    c->COEF = xcalloc((size_t)C_NPLANES * ML * KL * NSEA, sizeof(real));
    real DELPRO = (real)g->idelpro;
    real DELTH0 = (R(0.5) * DELPRO) / c->DELTR;
    real DELPH0 = (R(0.5) * DELPRO) / c->DELPHI;
    real DELTHR = (R(0.5) * DELPRO) / c->DELTH;
    real *DELLA0 = xcalloc(NSEA, sizeof(real)), *DRGM = xcalloc(NSEA, sizeof(real));
    int deep_nocur = !g->refr_c && !g->shallow;
    for (IJ = 1; IJ <= NSEA; IJ++) {
        int KX = c->KXLT[IJ - 1];
        if (deep_nocur) DELLA0[IJ - 1] = (DELPRO * c->DCO[IJ]) / c->DELLAM[KX - 1];
        else DELLA0[IJ - 1] = ((R(0.5) * DELPRO) * c->DCO[IJ]) / c->DELLAM[KX - 1];
        DRGM[IJ - 1] = c->SINPH[KX - 1] * c->DCO[IJ];
    }
#define KP_(k) (c->KPM[(k) * 3 + 2] - 1)
#define KM_(k) (c->KPM[(k) * 3 + 0] - 1)
    if (!g->refr_c && !g->shallow) {               /* PREP_SPHER_DEEP */
        for (K = 0; K < KL; K++) {
            real SD = c->SINTH[K], CD = c->COSTH[K];
            real SDA = ABS(SD), CDA = ABS(CD * DELPH0);
            int KP = KP_(K), KM = KM_(K);
            int IC = c->JYO[K * 2], ID = c->JYO[K * 2 + 1];
            real SP = DELTH0 * (c->SINTH[K] + c->SINTH[KP]);
            real SM = DELTH0 * (c->SINTH[K] + c->SINTH[KM]);
            for (M = 0; M < ML; M++) {
                real GOM = c->GOM[M];
                for (IJ = 0; IJ < NSEA; IJ++) {
                    real DPN = CDA * (c->DPSN[(size_t)(IC - 1) * NSEA + IJ] + R(1.));
                    real DLE = SDA * DELLA0[IJ];
                    real DTC = CDA * (c->DPSN[(size_t)(ID - 1) * NSEA + IJ] + R(1.)) + DLE;
                    real DTP = DRGM[IJ] * SP, DTM = DRGM[IJ] * SM;
                    DTC = (DTC + MAXR(R(0.), DTP)) - MINR(R(0.), DTM);
                    COEFP(C_LAT11, IJ, K, M) = GOM * DPN;
                    COEFP(C_LON1, IJ, K, M) = GOM * DLE;
                    COEFP(C_TH1, IJ, K, M) = -(GOM * MINR(R(0.), DTP));
                    COEFP(C_TH2, IJ, K, M) = GOM * MAXR(R(0.), DTM);
                    COEFP(C_SUM, IJ, K, M) = R(1.) - GOM * DTC;
                }
            }
        }
        c->ncoef = 5;
    } else if (!g->refr_c && g->shallow) {          /* PREP_SPHER_SHALLOW */
        real *CGKLAT = xcalloc(2 * (size_t)NSEA * ML, sizeof(real));
#define CGK(ij, m, ic) CGKLAT[(((size_t)(ic) - 1) * ML + (m)) * NSEA + (ij)]
        for (int IC = 1; IC <= 2; IC++)
            for (M = 0; M < ML; M++)
                for (IJ = 0; IJ < NSEA; IJ++)
                    CGK(IJ, M, IC) = CG(IJ + 1, M) + c->DPSN[(size_t)(IC - 1) * NSEA + IJ] * CG(KLAT_(IJ + 1, IC, 1), M);
        for (K = 0; K < KL; K++) {
            real SD = c->SINTH[K], CD = c->COSTH[K];
            real SDA = ABS(SD), CDA = ABS(CD * DELPH0);
            int KP = KP_(K), KM = KM_(K);
            real SP = DELTH0 * (c->SINTH[K] + c->SINTH[KP]);
            real SM = DELTH0 * (c->SINTH[K] + c->SINTH[KM]);
            int IS = c->JXO[K * 2], IT = c->JXO[K * 2 + 1], IC = c->JYO[K * 2], ID = c->JYO[K * 2 + 1];
            for (M = 0; M < ML; M++) {
                for (IJ = 0; IJ < NSEA; IJ++) {
                    real DTC = R(1.) - CDA * CGK(IJ, M, ID);
                    COEFP(C_LAT11, IJ, K, M) = CDA * CGK(IJ, M, IC);
                    DTC = DTC - (SDA * (CG(KLON_(IJ + 1, IT), M) + CG(IJ + 1, M))) * DELLA0[IJ];
                    COEFP(C_LON1, IJ, K, M) = (SDA * (CG(KLON_(IJ + 1, IS), M) + CG(IJ + 1, M))) * DELLA0[IJ];
                    real DTP = DRGM[IJ] * CG(IJ + 1, M), DTM;
                    if (g->refr_d) {
                        DTM = DTP * SM + (THD(IJ, K, M) + THD(IJ, KM, M)) * DELTHR;
                        DTP = DTP * SP + (THD(IJ, K, M) + THD(IJ, KP, M)) * DELTHR;
                    } else { DTM = DTP * SM; DTP = DTP * SP; }
                    DTC = (DTC - MAXR(R(0.), DTP)) + MINR(R(0.), DTM);
                    COEFP(C_TH1, IJ, K, M) = -MINR(R(0.), DTP);
                    COEFP(C_TH2, IJ, K, M) = MAXR(R(0.), DTM);
                    COEFP(C_SUM, IJ, K, M) = DTC;
                }
            }
        }
#undef CGK
        free(CGKLAT);
        c->ncoef = 5;
    } else {                                       /* current refraction, regular grid */
        real *WOK1 = xcalloc((size_t)NSEA + 1, sizeof(real)), *WOK2 = xcalloc((size_t)NSEA + 1, sizeof(real));
        real DELFR0 = g->shallow ? (R(0.5) * DELPRO) / (R(0.1) * c->ZPI)
                                 : ((R(0.5) * DELPRO) * R(2.1)) / R(0.2);
        for (M = 0; M < ML; M++) {
            int MM = c->MPM[M * 3 + 0] - 1, MP = c->MPM[M * 3 + 2] - 1;
            real DFP = DELFR0 / c->FR[M], DFM = DELFR0 / c->FR[MM];
            for (K = 0; K < KL; K++) {
                real SD = c->SINTH[K], CD = c->COSTH[K];
                int KP = KP_(K), KM = KM_(K);
                real SP, SM;
                if (g->shallow) {
                    SP = DELTH0 * (c->SINTH[K] + c->SINTH[KP]);
                    SM = DELTH0 * (c->SINTH[K] + c->SINTH[KM]);
                    for (IJ = 0; IJ <= NSEA; IJ++) {
                        WOK1[IJ] = U0[IJ] + SD * CG(IJ, M);
                        WOK2[IJ] = (V0[IJ] + CD * CG(IJ, M)) * DELPH0;
                    }
                } else {
                    SP = (DELTH0 * (c->SINTH[K] + c->SINTH[KP])) * c->GOM[M];
                    SM = (DELTH0 * (c->SINTH[K] + c->SINTH[KM])) * c->GOM[M];
                    real CGS = c->GOM[M] * SD, CGC = c->GOM[M] * CD;
                    for (IJ = 0; IJ <= NSEA; IJ++) {
                        WOK1[IJ] = U0[IJ] + CGS;
                        WOK2[IJ] = (V0[IJ] + CGC) * DELPH0;
                    }
                }
                for (IJ = 0; IJ < NSEA; IJ++) {
                    int ij1 = IJ + 1;
                    real DLW = (WOK1[ij1] + WOK1[KLON_(ij1, 1)]) * DELLA0[IJ];
                    real DLE = (WOK1[ij1] + WOK1[KLON_(ij1, 2)]) * DELLA0[IJ];
                    real DTC = (R(1.) - MAXR(R(0.), DLE)) + MINR(R(0.), DLW);
                    COEFP(C_LON1, IJ, K, M) = MAXR(R(0.), DLW);
                    COEFP(C_LON2, IJ, K, M) = -MINR(R(0.), DLE);
                    real DPS = WOK2[ij1] + c->DPSN[IJ] * WOK2[KLAT_(ij1, 1, 1)];
                    real DPN = WOK2[ij1] + c->DPSN[(size_t)NSEA + IJ] * WOK2[KLAT_(ij1, 2, 1)];
                    if (!g->shallow) {
                        DPS = WOK2[ij1] + WOK2[KLAT_(ij1, 1, 1)] * c->DPSN[IJ];
                        DPN = WOK2[ij1] + WOK2[KLAT_(ij1, 2, 1)] * c->DPSN[(size_t)NSEA + IJ];
                    }
                    DTC = (DTC - MAXR(R(0.), DPN)) + MINR(R(0.), DPS);
                    COEFP(C_LAT11, IJ, K, M) = MAXR(R(0.), DPS);
                    COEFP(C_LAT21, IJ, K, M) = -MINR(R(0.), DPN);
                    real DTP, DTM;
                    if (g->shallow) {
                        DTP = ((SP * DRGM[IJ]) * CG(ij1, M)) + (THD(IJ, K, M) + THD(IJ, KP, M)) * DELTHR;
                        DTM = ((SM * DRGM[IJ]) * CG(ij1, M)) + (THD(IJ, K, M) + THD(IJ, KM, M)) * DELTHR;
                    } else {
                        DTP = SP * DRGM[IJ] + (THD(IJ, K, ML - 1) + THD(IJ, KP, ML - 1)) * DELTHR;
                        DTM = SM * DRGM[IJ] + (THD(IJ, K, ML - 1) + THD(IJ, KM, ML - 1)) * DELTHR;
                    }
                    DTC = (DTC - MAXR(R(0.), DTP)) + MINR(R(0.), DTM);
                    COEFP(C_TH1, IJ, K, M) = -MINR(R(0.), DTP);
                    COEFP(C_TH2, IJ, K, M) = MAXR(R(0.), DTM);
                    if (g->shallow) {
                        real DOP = (SID(IJ, K, M) + SID(IJ, K, MP)) * DFP;
                        real DOM = (SID(IJ, K, M) + SID(IJ, K, MM)) * DFM;
                        DTC = (DTC - MAXR(R(0.), DOP)) + MINR(R(0.), DOM);
                        COEFP(C_SIG1, IJ, K, M) = -MINR(R(0.), DOP) / R(1.1);
                        COEFP(C_SIG2, IJ, K, M) = MAXR(R(0.), DOM) * R(1.1);
                    } else {
                        real DOM = SID(IJ, K, ML - 1) * DELFR0;
                        DTC = DTC - ABS(DOM);
                        COEFP(C_SIG1, IJ, K, M) = -MINR(R(0.), DOM) / R(1.1);
                        COEFP(C_SIG2, IJ, K, M) = MAXR(R(0.), DOM) * R(1.1);
                    }
                    COEFP(C_SUM, IJ, K, M) = DTC;
                }
            }
        }
        free(WOK1); free(WOK2);
        c->ncoef = 9;
    }
    free(DELLA0); free(DRGM); free(TEMP); free(THDD); free(SIDC); free(INDEP_G);
    free(DDPHI); free(DDLAM); free(DUPHI); free(DULAM); free(DVPHI); free(DVLAM);
    free(CGOND); free(D0); free(U0); free(V0);

    /* CHECK_CFL */
    size_t NB = (size_t)ML * KL * NSEA;
#define PMAX(p, out) do { real mx = c->COEF[(size_t)(p) * NB]; for (size_t q = 0; q < NB; q++) mx = MAXR(mx, c->COEF[(size_t)(p) * NB + q]); out = mx; } while (0)
#define PMIN(p, out) do { real mn = c->COEF[(size_t)(p) * NB]; for (size_t q = 0; q < NB; q++) mn = MINR(mn, c->COEF[(size_t)(p) * NB + q]); out = mn; } while (0)
    real lat11mx, lat11mn, lat21mx = R(0.), lat21mn = R(0.), lon1mx, lon1mn, lon2mx = R(0.1), lon2mn = R(0.1);
    real th1mx, th1mn, th2mx, th2mn, s1mx = R(0.1), s1mn = R(0.1), s2mx = R(0.1), s2mn = R(0.1), summx, summn;
    PMAX(C_LAT11, lat11mx); PMIN(C_LAT11, lat11mn);
    if (g->refr_c) {                  /* COEF_LAT(:,:,:,1:2,1): both IC planes */
        PMAX(C_LAT21, lat21mx); PMIN(C_LAT21, lat21mn);
        lat11mx = MAXR(lat11mx, lat21mx); lat11mn = MINR(lat11mn, lat21mn);
        PMAX(C_LON2, lon2mx); PMIN(C_LON2, lon2mn);
        PMAX(C_SIG1, s1mx); PMIN(C_SIG1, s1mn); PMAX(C_SIG2, s2mx); PMIN(C_SIG2, s2mn);
    }
    PMAX(C_LON1, lon1mx); PMIN(C_LON1, lon1mn);
    PMAX(C_TH1, th1mx); PMIN(C_TH1, th1mn); PMAX(C_TH2, th2mx); PMIN(C_TH2, th2mn);
    real sm_lo, sm_hi; PMIN(C_SUM, sm_lo); PMAX(C_SUM, sm_hi);
    summx = R(1.) - sm_lo; summn = R(1.) - sm_hi;
    real lat12 = R(0.1);
    real cmax = lat11mx;
    cmax = MAXR(cmax, lat12); cmax = MAXR(cmax, lon1mx); cmax = MAXR(cmax, lon2mx);
    cmax = MAXR(cmax, th1mx); cmax = MAXR(cmax, th2mx); cmax = MAXR(cmax, s1mx);
    cmax = MAXR(cmax, s2mx); cmax = MAXR(cmax, summx);
    /* the reference lists COEF_LON_1_2_MIN twice and omits COEF_LON_1_1_MIN (:2003-2006) */
    real cmin = lat11mn;
    cmin = MINR(cmin, lat12); cmin = MINR(cmin, lon2mn); cmin = MINR(cmin, lon2mn);
    cmin = MINR(cmin, th1mn); cmin = MINR(cmin, th2mn); cmin = MINR(cmin, s1mn);
    cmin = MINR(cmin, s2mn); cmin = MINR(cmin, summn);
    (void)lon1mn;
    c->cfl_max = cmax; c->cfl_min = cmin;
    if (cmin < R(0.)) return 3;                      /* ABORT1 in the reference */
    /* a window of a larger grid takes that grid's COUNTER (force_counter) */
    if (c->force_counter > 0 ? c->force_counter > 1 : cmax > R(0.999)) {
        c->counter = c->force_counter > 0 ? c->force_counter : (int)CEILR(cmax);
        real rc = (real)c->counter;
        for (int p = 1; p < C_NPLANES; p++)
            for (size_t q = 0; q < NB; q++) c->COEF[(size_t)p * NB + q] = c->COEF[(size_t)p * NB + q] / rc;
        for (size_t q = 0; q < NB; q++) c->COEF[q] = R(1.) - (R(1.) - c->COEF[q]) / rc;
    } else c->counter = 1;
    return 0;
#undef PMAX
#undef PMIN
#undef TAB
#undef CG
#undef THD
#undef SID
}

orc_ctx *orc_create(const orc_config *cfg, const uint8_t *mask, int *status)
{
    orc_ctx *c = xcalloc(1, sizeof(orc_ctx));
    c->cfg = *cfg;
    c->KL = cfg->kl; c->ML = cfg->ml;
    *status = 0;
    if (c->KL < 4 || c->ML < 6) { *status = 1; return c; }
    orc_constants(c);
    orc_fre_dir(c);
    if (orc_grid(c, mask)) { *status = 1; return c; }
    orc_min_energy(c);
    orc_shallow_tables(c);
    orc_nlweigt(c);
    orc_init_snonlin(c);
    orc_init_x0tauhf(c);
    c->counter = 1;
    return c;
}

void orc_free(orc_ctx *c)
{
    if (!c) return;
    void *ptrs[] = {c->FR, c->DF, c->DF_FR, c->DF_FR2, c->DFIM, c->DFIMOFR, c->DFIM_FR, c->DFIM_FR2,
                    c->FR5, c->FRM5, c->RHOWG_DFIM, c->GOM, c->C, c->TH, c->COSTH, c->SINTH, c->MPM,
                    c->KPM, c->JXO, c->JYO, c->FLMINFR, c->TFAK, c->TCGOND, c->TSIHKD, c->T_TAIL,
                    c->IKP, c->IKP1, c->IKM, c->IKM1, c->K1W, c->K2W, c->K11W, c->K21W, c->AF11,
                    c->FKLAP, c->FKLAP1, c->FKLAM, c->FKLAM1, c->FRH, c->FTRF, c->INLCOEF, c->RNLCOEF,
                    c->SINPH, c->COSPH, c->DELLAM, c->ZDELLO, c->IXLG, c->KXLT, c->KLAT, c->KLON,
                    c->WLAT, c->DCO, c->DPSN, c->COEF, c->INDEP, c->WN, c->CGG};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(c);
}
