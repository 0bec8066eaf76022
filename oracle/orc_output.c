/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 *
 * orc_output.c: the integrated parameters of COMPUTE_OUTPUT_PARAMETER
 * (wam_output_module.f90:481-601) for every sea point, from the spectra and
 * the wind fields: parameters 1-16 and 32 of the reference's BLOCK, in the
 * order of the WAM_OUT_* indices of include/wam_gpu.h.
 */
#include "orc_common.h"

/* MOD(A, 360.) for 0 <= A < 720 (exact). */
static real mod360(real a) { return a >= R(360.) ? a - R(360.) : a; }

/* a direction in radians as output: DEG*a ("coming from") or
   MOD(DEG*a + 180., 360.) ("going to"), wam_output_module.f90:531-535 */
static real out_dir(const orc_ctx *c, real a, int orient)
{
    return orient ? c->DEG * a : mod360(c->DEG * a + R(180.));
}

#ifdef ORC_FP64
#define ATAN2(y, x) atan2(y, x)
#define TINY_R 2.2250738585072014e-308
#else
#define ATAN2(y, x) wam_atan2f(y, x)
#define TINY_R 1.17549435e-38f
#endif

void orc_output_parameters(const orc_ctx *c, const real *fl3, const real *u10, const real *udir, const real *ustar,
                           const real *tauw, const real *z0, const real *depth, const real *u, const real *v,
                           int orient, real *out)
{
    const int KL = c->KL, ML = c->ML, NS = c->NSEA;
#pragma omp parallel for schedule(static)
    for (int ij = 0; ij < NS; ij++) {
        real T[64] = {R(0.)};
        /* SUM(F, DIM=2): direction sums per frequency */
        for (int m = 0; m < ML; m++) {
            real s = R(0.);
            for (int k = 0; k < KL; k++) {
                const real f = fl3[((size_t)m * KL + k) * NS + ij];
                s = k == 0 ? f : s + f;
            }
            T[m] = s;
        }
        /* TOTAL_ENERGY_B (wam_interface_module.f90:1342-1410) */
        real em = c->MO_TAIL * T[ML - 1];
        for (int m = 0; m < ML; m++) em = em + T[m] * c->DFIM[m];
        em = MAXR(em, c->EMIN);
        /* FEMEAN (:207-286); the output is the mean period 1/FM */
        real fm = c->MM1_TAIL * T[ML - 1];
        for (int m = 0; m < ML; m++) fm = fm + T[m] * c->DFIMOFR[m];
        fm = em / MAXR(fm, c->EMIN);
        /* TM1_TM2_PERIODS_B (:1152-1245) */
        real tm1 = c->MP1_TAIL * T[ML - 1], tm2 = c->MP2_TAIL * T[ML - 1];
        for (int m = 0; m < ML; m++) tm1 = tm1 + T[m] * c->DFIM_FR[m];
        for (int m = 0; m < ML; m++) tm2 = tm2 + T[m] * c->DFIM_FR2[m];
        tm1 = em > c->EMIN ? em / tm1 : R(1.);
        tm2 = em > c->EMIN ? SQRT(em / tm2) : R(1.);
        /* PEAK_PERIOD_B (:741-813): MAXLOC takes the first maximum */
        int ip = 0;
        for (int m = 1; m < ML; m++)
            if (T[m] > T[ip]) ip = m;
        const real peakp = ip == 0 ? R(1.) : R(1.) / c->FR[ip];
        /* MEAN_DIRECTION_B (:412-533) */
        real si = R(0.), ci = R(0.), sp = R(0.);
        for (int k = 0; k < KL; k++) {
            real te = fl3[(size_t)k * NS + ij] * c->DFIM[0];
            for (int m = 1; m < ML; m++) te = te + fl3[((size_t)m * KL + k) * NS + ij] * c->DFIM[m];
            si = si + c->SINTH[k] * te;
            ci = ci + c->COSTH[k] * te;
            sp = sp + te;
        }
        if (ci == R(0.)) ci = R(0.1E-30);
        real thq = ATAN2(si, ci);
        if (thq < R(0.)) thq = thq + c->ZPI;
        if (thq > c->ZPI - R(0.001)) thq = R(0.);
        if (ABS(ci) < R(0.1E-15)) ci = SIGNR(R(0.1E-15), ci);
        if (ABS(si) < R(0.1E-15)) si = SIGNR(R(0.1E-15), si);
        if (ABS(sp) < R(0.1E-15)) sp = SIGNR(R(0.1E-15), sp);
        sp = R(2.) * (R(1.) - SQRT(si * si + ci * ci) / sp);
        sp = sp <= R(0.) ? (real)TINY_R : SQRT(sp);
        /* the wind and current parameters (wam_output_module.f90:525-563) */
        const real us = ustar[ij], tau = us * us + R(0.0001);
        const real uu = u ? u[ij] : R(0.), vv = v ? v[ij] : R(0.);
        real cdir = ATAN2(uu, vv) * c->DEG;
        if (!orient) cdir = mod360(cdir + R(180.));
        if (cdir < R(0.)) cdir = cdir + R(360.);
        real *o = out + ij;
        o[0 * (size_t)NS] = u10[ij];
        o[1 * (size_t)NS] = out_dir(c, udir[ij], orient);
        o[2 * (size_t)NS] = us;
        o[3 * (size_t)NS] = tau / (u10[ij] * u10[ij] + R(0.01));
        o[4 * (size_t)NS] = (c->G * z0[ij]) / MAXR(us * us, R(1.0E-6));   /* CHARNOCK_PAR (:159-203) */
        o[5 * (size_t)NS] = c->cfg.shallow ? depth[ij] : R(999.);
        o[6 * (size_t)NS] = SQRT(uu * uu + vv * vv);
        o[7 * (size_t)NS] = cdir;
        o[8 * (size_t)NS] = R(4.) * SQRT(em);
        o[9 * (size_t)NS] = peakp;
        o[10 * (size_t)NS] = R(1.) / fm;
        o[11 * (size_t)NS] = tm1;
        o[12 * (size_t)NS] = tm2;
        o[13 * (size_t)NS] = out_dir(c, thq, orient);
        o[14 * (size_t)NS] = sp * c->DEG;
        o[15 * (size_t)NS] = tauw[ij] / (tau * c->ROAIR);
        o[16 * (size_t)NS] = z0[ij];
    }
}
