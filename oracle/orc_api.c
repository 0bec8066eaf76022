/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 * Flat C entry points for ctypes: create/free, named table access, the
 * propagation set-up, PROPAGS and IMPLSCH.  Built twice: liborc32.so (REAL*4)
 * and liborc64.so (-DORC_FP64, REAL*8).
 */
#include <string.h>
#include "orc_common.h"

orc_ctx *orc_create(const orc_config *cfg, const uint8_t *mask, int *status);
void orc_free(orc_ctx *c);
int orc_prepare_propagation(orc_ctx *c, const real *depth, const real *u, const real *v);
void orc_propags(const orc_ctx *c, real *F3);
void orc_implsch(const orc_ctx *c, real *fl3, const real *u10, const real *udir, real *tauw,
                 real *ustar, real *z0, const real *roairn, const real *wstar, const real *depth,
                 const int *indep, int lcflx, real *phiaw, real *out_flux, int *mij, int nthreads);
void orc_wavnu2_all(const orc_ctx *c, const real *depth, real *wn, real *cgg);
real orc_aki(const orc_ctx *c, real OM, real BETA);

int orc_real_size(void) { return (int)sizeof(real); }

orc_ctx *orc_api_create(const orc_config *cfg, const uint8_t *mask, int *status)
{
    return orc_create(cfg, mask, status);
}
void orc_api_free(orc_ctx *c) { orc_free(c); }

/* type: 0 = real, 1 = int.  Returns NULL for an unknown name. */
const void *orc_api_get(const orc_ctx *c, const char *name, long *n, int *type)
{
    int KL = c->KL, ML = c->ML, NM = c->MLSTHG - c->MFRSTLW + 1;
    long NS = c->NSEA;
    struct { const char *nm; const void *p; long n; int t; } tab[] = {
        {"FR", c->FR, ML, 0}, {"DF", c->DF, ML, 0}, {"DFIM", c->DFIM, ML, 0},
        {"DFIMOFR", c->DFIMOFR, ML, 0}, {"DFIM_FR", c->DFIM_FR, ML, 0}, {"DFIM_FR2", c->DFIM_FR2, ML, 0},
        {"FR5", c->FR5, ML, 0}, {"FRM5", c->FRM5, ML, 0}, {"RHOWG_DFIM", c->RHOWG_DFIM, ML, 0},
        {"GOM", c->GOM, ML, 0}, {"C", c->C, ML, 0}, {"TH", c->TH, KL, 0},
        {"COSTH", c->COSTH, KL, 0}, {"SINTH", c->SINTH, KL, 0},
        {"MPM", c->MPM, 3L * ML, 1}, {"KPM", c->KPM, 3L * KL, 1}, {"JXO", c->JXO, 2L * KL, 1}, {"JYO", c->JYO, 2L * KL, 1},
        {"FLMINFR", c->FLMINFR, (long)c->JUMAX * ML, 0},
        {"TFAK", c->TFAK, (long)c->cfg.ndepth * ML, 0}, {"TCGOND", c->TCGOND, (long)c->cfg.ndepth * ML, 0},
        {"TSIHKD", c->TSIHKD, (long)c->cfg.ndepth * ML, 0}, {"T_TAIL", c->T_TAIL, (long)c->cfg.ndepth * ML, 0},
        {"IKP", c->IKP, NM, 1}, {"IKP1", c->IKP1, NM, 1}, {"IKM", c->IKM, NM, 1}, {"IKM1", c->IKM1, NM, 1},
        {"K1W", c->K1W, 2L * KL, 1}, {"K2W", c->K2W, 2L * KL, 1}, {"K11W", c->K11W, 2L * KL, 1}, {"K21W", c->K21W, 2L * KL, 1},
        {"AF11", c->AF11, NM, 0}, {"FKLAP", c->FKLAP, NM, 0}, {"FKLAP1", c->FKLAP1, NM, 0},
        {"FKLAM", c->FKLAM, NM, 0}, {"FKLAM1", c->FKLAM1, NM, 0}, {"FRH", c->FRH, c->KFRH, 0},
        {"FTRF", c->FTRF, 2 - c->MFRSTLW, 0},
        {"INLCOEF", c->INLCOEF, 5L * c->MLSTHG, 1}, {"RNLCOEF", c->RNLCOEF, 25L * c->MLSTHG, 0},
        {"WTAUHF", c->WTAUHF, 19, 0},
        {"SINPH", c->SINPH, c->cfg.ny, 0}, {"COSPH", c->COSPH, c->cfg.ny, 0}, {"DELLAM", c->DELLAM, c->cfg.ny, 0},
        {"ZDELLO", c->ZDELLO, c->cfg.ny, 1}, {"IXLG", c->IXLG, NS, 1}, {"KXLT", c->KXLT, NS, 1},
        {"KLAT", c->KLAT, 4 * NS, 1}, {"KLON", c->KLON, 2 * NS, 1}, {"WLAT", c->WLAT, 2 * NS, 0},
        {"DCO", c->DCO, NS + 1, 0}, {"DPSN", c->DPSN, 2 * NS, 0},
        {"COEF", c->COEF, c->COEF ? (long)C_NPLANES * ML * KL * NS : 0, 0},
        {"INDEP", c->INDEP, c->INDEP ? NS : 0, 1},
    };
    for (size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); i++)
        if (!strcmp(tab[i].nm, name)) { *n = tab[i].n; *type = tab[i].t; return tab[i].p; }
    return NULL;
}

double orc_api_scalar(const orc_ctx *c, const char *name)
{
    struct { const char *nm; double v; } tab[] = {
        {"G", c->G}, {"ZPI", c->ZPI}, {"DELTH", c->DELTH}, {"DELTR", c->DELTR}, {"INV_LOG_CO", c->INV_LOG_CO},
        {"MO_TAIL", c->MO_TAIL}, {"MM1_TAIL", c->MM1_TAIL}, {"MP1_TAIL", c->MP1_TAIL}, {"MP2_TAIL", c->MP2_TAIL},
        {"ACL1", c->ACL1}, {"ACL2", c->ACL2}, {"CL11", c->CL11}, {"CL21", c->CL21}, {"DAL1", c->DAL1}, {"DAL2", c->DAL2},
        {"ISP", c->ISP}, {"ISM", c->ISM}, {"MFRSTLW", c->MFRSTLW}, {"MLSTHG", c->MLSTHG}, {"KFRH", c->KFRH},
        {"X0TAUHF", c->X0TAUHF}, {"DELPHI", c->DELPHI}, {"NSEA", c->NSEA}, {"DELU", c->DELU},
        {"COUNTER", c->counter}, {"CFL_MAX", c->cfl_max}, {"CFL_MIN", c->cfl_min}, {"NCOEF", c->ncoef},
        {"ALPHA", c->ALPHA}, {"JUMAX", c->JUMAX}, {"CO", c->CO}, {"EMIN", c->EMIN},
    };
    for (size_t i = 0; i < sizeof(tab) / sizeof(tab[0]); i++)
        if (!strcmp(tab[i].nm, name)) return tab[i].v;
    return -1e300;
}

int orc_api_prepare_propagation(orc_ctx *c, const real *depth, const real *u, const real *v)
{
    return orc_prepare_propagation(c, depth, u, v);
}
void orc_api_propags(const orc_ctx *c, real *fl3) { orc_propags(c, fl3); }
/* CHECK_CFL's sub-step count of the grid this context is a window of (0:
   its own); set before orc_api_prepare_propagation */
void orc_api_force_counter(orc_ctx *c, int counter) { c->force_counter = counter; }
/* SET_MODEL_OPTION's WAVE_BREAKING and PHILLIPS (wam_timopt_module.f90:176-177):
   breaking needs a shallow-water run and is switched off otherwise (:213-227). */
void orc_api_source_options(orc_ctx *c, int wave_breaking, int phillips)
{
    c->wave_breaking = wave_breaking && c->cfg.shallow;
    c->phillips = phillips != 0;
}
void orc_api_implsch(const orc_ctx *c, real *fl3, const real *u10, const real *udir, real *tauw,
                     real *ustar, real *z0, const real *roairn, const real *wstar, const real *depth,
                     int lcflx, real *phiaw, real *out_flux, int *mij, int nthreads)
{
    orc_implsch(c, fl3, u10, udir, tauw, ustar, z0, roairn, wstar, depth, c->INDEP, lcflx,
                phiaw, out_flux, mij, nthreads);
}
void orc_api_wavnu2(const orc_ctx *c, const real *depth, real *wn, real *cgg) { orc_wavnu2_all(c, depth, wn, cgg); }
void orc_output_parameters(const orc_ctx *c, const real *fl3, const real *u10, const real *udir, const real *ustar,
                           const real *tauw, const real *z0, const real *depth, const real *u, const real *v,
                           int orient, real *out);
void orc_api_output_parameters(const orc_ctx *c, const real *fl3, const real *u10, const real *udir,
                               const real *ustar, const real *tauw, const real *z0, const real *depth, const real *u,
                               const real *v, int orient, real *out)
{
    orc_output_parameters(c, fl3, u10, udir, ustar, tauw, z0, depth, u, v, orient, out);
}
double orc_api_aki(const orc_ctx *c, double om, double beta) { return orc_aki(c, (real)om, (real)beta); }

void orc_snonlin_point(const orc_ctx *c, real *F, real *SL, real *FL, real depth, real akmean);
void orc_api_snonlin(const orc_ctx *c, real *F, real *SL, real *FL, double depth, double akmean)
{
    orc_snonlin_point(c, F, SL, FL, (real)depth, (real)akmean);
}

/* The shared single-precision elementary functions (wam_math.h), for the
   accuracy tests: fn 0 exp, 1 log, 2 log10, 3 tanh, 4 cosh, 5 sinh, 6 cos,
   7 sin, 8 pow(x, y), 9 nint, 10 ipow(x, (int)y), 11 atan2(x, y). */
#include "wam_math.h"
void orc_api_math(int fn, const float *x, const float *y, float *out, long n)
{
    for (long i = 0; i < n; i++) {
        float a = x[i], b = y ? y[i] : 0.f, r = 0.f;
        switch (fn) {
        case 0: r = wam_expf(a); break;
        case 1: r = wam_logf(a); break;
        case 2: r = wam_log10f(a); break;
        case 3: r = wam_tanhf(a); break;
        case 4: r = wam_coshf(a); break;
        case 5: r = wam_sinhf(a); break;
        case 6: r = wam_cosf(a); break;
        case 7: r = wam_sinf(a); break;
        case 8: r = wam_powf(a, b); break;
        case 9: r = (float)wam_nint(a); break;
        case 10: r = wam_ipowf(a, (int)b); break;
        case 11: r = wam_atan2f(a, b); break;
        case 12: r = wam_div_1p1(a); break;
        }
        out[i] = r;
    }
}
