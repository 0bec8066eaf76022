/*
 * ORACLE -- test infrastructure only (see orc_common.h).
 *
 * orc_propags.c: PROPAGS of the CPU restatement
 * (wam_propagation_module.f90:635-829), on the weights PREPARE_PROPAGATION
 * stored in the context (orc_setup.c).  A file of its own so that the
 * differentiating build (orc_ad.cpp) compiles it with active spectra.
 */
#include <stdlib.h>
#include <string.h>
#include "orc_common.h"

static void *propags_calloc(size_t n, size_t s) { void *p = calloc(n ? n : 1, s); if (!p) abort(); return p; }
#define xcalloc propags_calloc

#define KLAT_(ij, ic, icl) c->KLAT[((size_t)((icl) - 1) * 2 + ((ic) - 1)) * NSEA + (ij) - 1]
#define KLON_(ij, ic) c->KLON[(size_t)((ic) - 1) * NSEA + (ij) - 1]
#define COEFP(p, ij, k, m) c->COEF[(((size_t)(p) * ML + (m)) * KL + (k)) * NSEA + (ij)]
#define KP_(k) (c->KPM[(k) * 3 + 2] - 1)
#define KM_(k) (c->KPM[(k) * 3 + 0] - 1)

/* PROPAGS (PROP_NO_CURR :750-785, PROP_CURR :794-825), COUNTER sub-steps. */
void orc_propags(const orc_ctx *c, real *F3)
{
    int NSEA = c->NSEA, KL = c->KL, ML = c->ML, K, M, IJ;
    size_t NB = (size_t)KL * ML;
    real *F1 = xcalloc(NB * (NSEA + 1), sizeof(real));      /* [M][K][0..NSEA], slot 0 land */
#define F1_(ij, k, m) F1[(((size_t)(m) * KL) + (k)) * (NSEA + 1) + (ij)]
#define F3_(ij, k, m) F3[(((size_t)(m) * KL) + (k)) * NSEA + (ij) - 1]
    for (int it = 0; it < c->counter; it++) {
        for (M = 0; M < ML; M++)
            for (K = 0; K < KL; K++) {
                F1_(0, K, M) = R(0.);
                for (IJ = 1; IJ <= NSEA; IJ++) F1_(IJ, K, M) = F3_(IJ, K, M);
            }
        if (c->cfg.refr_c) {
            for (M = 0; M < ML; M++) {
                int MM = c->MPM[M * 3 + 0] - 1, MP = c->MPM[M * 3 + 2] - 1;
                for (K = 0; K < KL; K++) {
                    int KP = KP_(K), KM = KM_(K);
                    for (IJ = 1; IJ <= NSEA; IJ++) {
                        int i0 = IJ - 1;
                        real s = COEFP(C_SUM, i0, K, M) * F1_(IJ, K, M);
                        s = s + COEFP(C_LAT21, i0, K, M) * F1_(KLAT_(IJ, 2, 1), K, M);
                        s = s + COEFP(C_LAT11, i0, K, M) * F1_(KLAT_(IJ, 1, 1), K, M);
                        s = s + COEFP(C_LON2, i0, K, M) * F1_(KLON_(IJ, 2), K, M);
                        s = s + COEFP(C_LON1, i0, K, M) * F1_(KLON_(IJ, 1), K, M);
                        s = s + COEFP(C_TH1, i0, K, M) * F1_(IJ, KP, M);
                        s = s + COEFP(C_TH2, i0, K, M) * F1_(IJ, KM, M);
                        s = s + COEFP(C_SIG1, i0, K, M) * F1_(IJ, K, MP);
                        s = s + COEFP(C_SIG2, i0, K, M) * F1_(IJ, K, MM);
                        F3_(IJ, K, M) = s;
                    }
                }
            }
        } else {
            for (K = 0; K < KL; K++) {
                int KP = KP_(K), KM = KM_(K);
                int IS = c->JXO[K * 2], IC = c->JYO[K * 2];
                for (M = 0; M < ML; M++)
                    for (IJ = 1; IJ <= NSEA; IJ++) {
                        int i0 = IJ - 1;
                        real s = COEFP(C_SUM, i0, K, M) * F1_(IJ, K, M)
                               + COEFP(C_LAT11, i0, K, M) * F1_(KLAT_(IJ, IC, 1), K, M);
                        s = s + COEFP(C_LON1, i0, K, M) * F1_(KLON_(IJ, IS), K, M);
                        if (c->cfg.refr_d || c->cfg.spherical) {
                            s = s + COEFP(C_TH1, i0, K, M) * F1_(IJ, KP, M);
                            s = s + COEFP(C_TH2, i0, K, M) * F1_(IJ, KM, M);
                        }
                        F3_(IJ, K, M) = s;
                    }
            }
        }
    }
    free(F1);
#undef F1_
#undef F3_
}

