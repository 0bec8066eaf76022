/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 *
 * orc_nest.c: the nest boundary exchange of WAM Cycle 7.
 *  * BOUNDARY_OUTPUT (wam_boundary_module.f90:291-397): the coarse grid's
 *    spectra at the fine grid's boundary points with their TOTAL_ENERGY,
 *    MEAN_DIRECTION (THQ) and FEMEAN (FM);
 *  * BOUNDARY_INPUT (:168-287): time interpolation of two sets of boundary
 *    spectra, then the space interpolation into FL3 at the fine boundary
 *    points, both by INTSPEC (wam_interface_module.f90:290-408) with ROTSPEC
 *    (:891-941) and STRSPEC (:945-1006).
 * A spectrum here is F(KL, ML), f[m*KL + k].
 *
 * What pins it: golden vectors of the reference's own TOTAL_ENERGY, FEMEAN,
 * MEAN_DIRECTION, INTSPEC, ROTSPEC and STRSPEC built with flang
 * (tests/test_ref_golden_nest.py: EMEAN, FMEAN and ROTSPEC bit-identical,
 * THQ, STRSPEC and INTSPEC within the libm boundary); and (tests/test_nest.py)
 * ROTSPEC by whole direction steps is a
 * cyclic shift and by zero the identity; STRSPEC by CO**n is a frequency
 * shift, by one the identity, by a factor outside the range zero; INTSPEC at
 * the end points returns either spectrum, is symmetric under exchanging the
 * two, and between two copies of one spectrum returns it; the mean
 * parameters against orc_output_parameters (Hs, mean period, mean direction)
 * on the same spectra.  Parity unpinned: REAL**(-n), taken as 1/REAL**n (the
 * GNU runtime's power), where the reference's compiler may differ in the last
 * bit.
 */
#include "orc_common.h"

#ifdef ORC_FP64
#define ATAN2(y, x) atan2(y, x)
#define EPS_R 2.220446049250313e-16
#else
#define ATAN2(y, x) wam_atan2f(y, x)
#define EPS_R 1.1920929e-7f
#endif

/* REAL**INTEGER, negative powers as 1/x**n. */
static real powi_r(real x, int n)
{
    if (n >= 0) return n == 0 ? R(1.) : IPOW(x, n);
    return R(1.) / IPOW(x, -n);
}

/* ROTSPEC: F_OUT = BDIF*CSHIFT(F_IN, INC) + ADIF*CSHIFT(F_IN, INC-1), the
   shift along the directions. */
static void rotspec(const orc_ctx *c, const real *fin, real *fout, real rthet)
{
    const int KL = c->KL, ML = c->ML;
    real adif = rthet * (real)KL / c->ZPI;
    const int inc = -(int)FLOORR(adif);
    adif = adif + (real)inc;
    const real bdif = R(1.) - adif;
    for (int m = 0; m < ML; m++)
        for (int k = 0; k < KL; k++) {
            const int k0 = ((k + inc) % KL + KL) % KL, k1 = ((k + inc - 1) % KL + KL) % KL;
            fout[m * KL + k] = bdif * fin[m * KL + k0] + adif * fin[m * KL + k1];
        }
}

/* STRSPEC: the spectrum stretched in frequency by GAMMA. */
static void strspec(const orc_ctx *c, const real *fin, real *fout, real gamma)
{
    const int KL = c->KL, ML = c->ML;
    if (gamma == R(1.)) {
        for (int i = 0; i < KL * ML; i++) fout[i] = fin[i];
        return;
    }
    for (int i = 0; i < KL * ML; i++) fout[i] = R(0.);
    const int inc = (int)FLOORR(LOG10(gamma) / LOG10(c->CO));
    if ((inc < 0 ? -inc : inc) >= ML - 1) return;
    const real adif = (c->CO - gamma * powi_r(c->CO, -inc)) / (c->CO - R(1.));
    const real bdif = R(1.) - adif;
    if (inc >= 0) {
        /* F_OUT(:,1:ML-INC-1) = ADIF*F_IN(:,1+INC:ML-1) + BDIF*F_IN(:,2+INC:ML); F_OUT(:,ML-INC) = ADIF*F_IN(:,ML) */
        for (int m = 0; m < ML - inc - 1; m++)
            for (int k = 0; k < KL; k++)
                fout[m * KL + k] = adif * fin[(m + inc) * KL + k] + bdif * fin[(m + inc + 1) * KL + k];
        for (int k = 0; k < KL; k++) fout[(ML - inc - 1) * KL + k] = adif * fin[(ML - 1) * KL + k];
    } else {
        /* F_OUT(:,1-INC:ML) = ADIF*F_IN(:,1:ML+INC) + BDIF*F_IN(:,2:ML+INC+1); F_OUT(:,-INC) = BDIF*F_IN(:,1) */
        for (int m = -inc; m < ML; m++)
            for (int k = 0; k < KL; k++)
                fout[m * KL + k] = adif * fin[(m + inc) * KL + k] + bdif * fin[(m + inc + 1) * KL + k];
        for (int k = 0; k < KL; k++) fout[(-inc - 1) * KL + k] = bdif * fin[k];
    }
}

/* INTSPEC: the spectrum at DEL1L between F1 (at 0) and F2 (at DEL12), with
   its mean frequency, energy and direction. */
void orc_intspec(const orc_ctx *c, real del12, real del1l, const real *f1, real fmean1, real emean1, real thetm1,
                 const real *f2, real fmean2, real emean2, real thetm2, real *fl, real *fmean, real *emean,
                 real *thetm)
{
    const int NB = c->KL * c->ML;
    real gw2 = del1l / del12, gw1 = R(1.) - gw2;
    if (ABS(gw1) < EPS_R || emean1 < EPS_R) {
        for (int i = 0; i < NB; i++) fl[i] = gw2 * f2[i];
        *emean = gw2 * emean2; *fmean = gw2 * fmean2; *thetm = gw2 * thetm2;
        return;
    }
    if (ABS(gw2) < EPS_R || emean2 < EPS_R) {
        for (int i = 0; i < NB; i++) fl[i] = gw1 * f1[i];
        *emean = gw1 * emean1; *fmean = gw1 * fmean1; *thetm = gw1 * thetm1;
        return;
    }
    const real em = gw1 * emean1 + gw2 * emean2;
    const real fm = gw1 * fmean1 + gw2 * fmean2;
    const real th = ATAN2(gw1 * SIN(thetm1) + gw2 * SIN(thetm2), gw1 * COS(thetm1) + gw2 * COS(thetm2));
    real rot[64 * 60], fL[64 * 60], fR[64 * 60];
    rotspec(c, f1, rot, th - thetm1);
    strspec(c, rot, fL, fmean1 / fm);
    gw1 = (gw1 * em) / emean1;
    rotspec(c, f2, rot, th - thetm2);
    strspec(c, rot, fR, fmean2 / fm);
    gw2 = (gw2 * em) / emean2;
    for (int i = 0; i < NB; i++) fl[i] = gw1 * fL[i] + gw2 * fR[i];
    *emean = em; *fmean = fm; *thetm = th;
}

/* BOUNDARY_OUTPUT: fbc[(ib*ML + m)*KL + k] = FL3(IJARC(ib), k, m), and per
   point TOTAL_ENERGY (wam_interface_module.f90:1342-1410), MEAN_DIRECTION's
   THQ (:412-533) and FEMEAN's FM (:207-286): mean[ib], mean[nb + ib],
   mean[2 nb + ib] = EMEAN, THQ, FMEAN.  ijarc is 1-based. */
void orc_boundary_output(const orc_ctx *c, const real *fl3, int nb, const int *ijarc, real *fbc, real *mean)
{
    const int KL = c->KL, ML = c->ML;
    const long NS = c->NSEA;
    for (int ib = 0; ib < nb; ib++) {
        const long ij = ijarc[ib] - 1;
        real *f = fbc + (size_t)ib * KL * ML;
        for (int m = 0; m < ML; m++)
            for (int k = 0; k < KL; k++) f[m * KL + k] = fl3[((size_t)m * KL + k) * NS + ij];
        real T[64] = {R(0.)};
        for (int m = 0; m < ML; m++) {
            real s = R(0.);
            for (int k = 0; k < KL; k++) s = k == 0 ? f[m * KL + k] : s + f[m * KL + k];
            T[m] = s;
        }
        real em = c->MO_TAIL * T[ML - 1];
        for (int m = 0; m < ML; m++) em = em + T[m] * c->DFIM[m];
        em = MAXR(em, c->EMIN);
        real fm = c->MM1_TAIL * T[ML - 1];
        for (int m = 0; m < ML; m++) fm = fm + T[m] * c->DFIMOFR[m];
        fm = em / MAXR(fm, c->EMIN);
        real si = R(0.), ci = R(0.);
        for (int k = 0; k < KL; k++) {
            real te = f[k] * c->DFIM[0];
            for (int m = 1; m < ML; m++) te = te + f[m * KL + k] * c->DFIM[m];
            si = si + c->SINTH[k] * te;
            ci = ci + c->COSTH[k] * te;
        }
        if (ci == R(0.)) ci = R(0.1E-30);
        real thq = ATAN2(si, ci);
        if (thq < R(0.)) thq = thq + c->ZPI;
        if (thq > c->ZPI - R(0.001)) thq = R(0.);
        mean[ib] = em;
        mean[nb + ib] = thq;
        mean[2 * nb + ib] = fm;
    }
}

/* BOUNDARY_INPUT.  Time step: which = 1 takes the first set, 2 the second,
   0 interpolates at DEL1L of DEL12 (INTSPEC per input point); fi, mi get the
   set at the model time.  Space step: FL3(IJARF(i),:,:) = INTSPEC(1, BFW(i),
   FI(IBFL(i)), FI(IBFR(i))) in the order of i (all indices 1-based).  m1, m2,
   mi hold EMEAN, THQ, FMEAN by [3][nbinp], the order of the boundary file. */
void orc_boundary_input(const orc_ctx *c, real *fl3, int nbinp, const real *f1, const real *m1, const real *f2,
                        const real *m2, int which, real del12, real del1l, int nbounf, const int *ijarf,
                        const int *ibfl, const int *ibfr, const real *bfw, real *fi, real *mi)
{
    const int KL = c->KL, ML = c->ML, NB = KL * ML;
    const long NS = c->NSEA;
    for (int p = 0; p < nbinp; p++) {
        real *o = fi + (size_t)p * NB;
        if (which == 1 || which == 2) {
            const real *s = (which == 1 ? f1 : f2) + (size_t)p * NB;
            const real *sm = which == 1 ? m1 : m2;
            for (int i = 0; i < NB; i++) o[i] = s[i];
            mi[p] = sm[p]; mi[nbinp + p] = sm[nbinp + p]; mi[2 * nbinp + p] = sm[2 * nbinp + p];
        } else {
            orc_intspec(c, del12, del1l, f1 + (size_t)p * NB, m1[2 * nbinp + p], m1[p], m1[nbinp + p],
                        f2 + (size_t)p * NB, m2[2 * nbinp + p], m2[p], m2[nbinp + p], o, &mi[2 * nbinp + p],
                        &mi[p], &mi[nbinp + p]);
        }
    }
    real out[64 * 60];
    for (int i = 0; i < nbounf; i++) {
        const int l = ibfl[i] - 1, r = ibfr[i] - 1;
        const long ij = ijarf[i] - 1;
        real fmean, emean, thq;
        orc_intspec(c, R(1.), bfw[i], fi + (size_t)l * NB, mi[2 * nbinp + l], mi[l], mi[nbinp + l],
                    fi + (size_t)r * NB, mi[2 * nbinp + r], mi[r], mi[nbinp + r], out, &fmean, &emean, &thq);
        for (int m = 0; m < ML; m++)
            for (int k = 0; k < KL; k++) fl3[((size_t)m * KL + k) * NS + ij] = out[m * KL + k];
    }
}

/* ctypes entry points: scalars as double, arrays in the build's REAL. */
void orc_api_intspec(const orc_ctx *c, double del12, double del1l, const real *f1, const real *m1, const real *f2,
                     const real *m2, real *fl, real *mo)
{
    /* m1, m2, mo: EMEAN, THQ, FMEAN */
    orc_intspec(c, (real)del12, (real)del1l, f1, m1[2], m1[0], m1[1], f2, m2[2], m2[0], m2[1], fl, &mo[2], &mo[0],
                &mo[1]);
}

void orc_api_boundary_output(const orc_ctx *c, const real *fl3, int nb, const int *ijarc, real *fbc, real *mean)
{
    orc_boundary_output(c, fl3, nb, ijarc, fbc, mean);
}

void orc_api_boundary_input(const orc_ctx *c, real *fl3, int nbinp, const real *f1, const real *m1, const real *f2,
                            const real *m2, int which, double del12, double del1l, int nbounf, const int *ijarf,
                            const int *ibfl, const int *ibfr, const real *bfw, real *fi, real *mi)
{
    orc_boundary_input(c, fl3, nbinp, f1, m1, f2, m2, which, (real)del12, (real)del1l, nbounf, ijarf, ibfl, ibfr,
                       bfw, fi, mi);
}

void orc_api_rotspec(const orc_ctx *c, const real *fin, real *fout, double rthet) { rotspec(c, fin, fout, (real)rthet); }

void orc_api_strspec(const orc_ctx *c, const real *fin, real *fout, double gamma) { strspec(c, fin, fout, (real)gamma); }
