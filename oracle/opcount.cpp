/*
 * ORACLE -- measurement infrastructure only (see orc_common.h).
 *
 * opcount.cpp: IMPLSCH of the CPU restatement (orc_implsch.c) compiled with
 * the counting REAL*4 type of orc_opcount.h, run on one thread, to state the
 * algorithmic floating-point work per sea point.  The set-up tables come
 * from the plain float build (same layout).
 */
#include <math.h>
#include <omp.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#define ORC_OPCOUNT
#include "orc_common.h"

oc_counts oc_cnt;
oc_counts oc_stage_cnt[OC_NSTAGE], oc_mark;
int oc_cur_stage;

namespace orc_counted {
#include "orc_implsch.c"
}

/* Runs IMPLSCH once over every sea point (arguments as orc_api_implsch) and
   returns the operation totals: add, mul, div, sqrt, transcendental; and the
   same five per stage in stage_counts[3][5] (stage 0: up to SDISSIP, 1:
   SNONLIN, SBOTTOM, WNFLUXES and the limited update, 2: the new mean
   parameters, cut-off and tail). */
extern "C" void orc_opcount_implsch(const void *ctx, float *fl3, const float *u10, const float *udir, float *tauw,
                                    float *ustar, float *z0, const float *roairn, const float *wstar,
                                    const float *depth, int lcflx, float *phiaw, float *out_flux, int *mij,
                                    uint64_t *counts, uint64_t *stage_counts)
{
    const orc_ctx *c = static_cast<const orc_ctx *>(ctx);
    auto R_ = [](float *p) { return reinterpret_cast<oc_real *>(p); };
    auto C_ = [](const float *p) { return reinterpret_cast<const oc_real *>(p); };
    oc_cnt = oc_counts{};
    oc_mark = oc_counts{};
    for (auto &x : oc_stage_cnt) x = oc_counts{};
    oc_cur_stage = 0;
    orc_counted::orc_implsch(c, R_(fl3), C_(u10), C_(udir), R_(tauw), R_(ustar), R_(z0), C_(roairn), C_(wstar),
                             C_(depth), c->INDEP, lcflx, R_(phiaw), R_(out_flux), mij, 1);
    counts[0] = oc_cnt.add; counts[1] = oc_cnt.mul; counts[2] = oc_cnt.div;
    counts[3] = oc_cnt.sqrt; counts[4] = oc_cnt.trans;
    for (int i = 0; i < OC_NSTAGE; i++) {
        const oc_counts &x = oc_stage_cnt[i];
        uint64_t *o = stage_counts + 5 * i;
        o[0] = x.add; o[1] = x.mul; o[2] = x.div; o[3] = x.sqrt; o[4] = x.trans;
    }
}
