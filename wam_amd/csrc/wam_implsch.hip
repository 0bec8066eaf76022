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
#include "wam_source_dev.hpp"
#include "wam_dual.hpp"

namespace wam {

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
        if (t.brk) {
            // SDEPTHLIM (:513-515): the spectrum and EMEAN scaled by EM; the
            // other mean parameters come from the scaled spectrum
            if (tid == 0) {
                const float e0 = wam_maxf(integ(ML, t.MO_TAIL, T, s.dfim), t.EMIN);
                const float em = sdepthlim_em(P.depth, e0);
                P.em = em;
                P.emean = em * e0;
            }
            __syncthreads();
            const float em = P.em;
            for (int b = tid; b < NB; b += NT) s.F[b] = s.F[b] * em;
            __syncthreads();
            dirsum(t, s.F, nullptr, T, tid);
            __syncthreads();
        }
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
            const float e = t.brk ? P.emean : wam_maxf(P.s_e, t.EMIN);
            P.emean = e;
            P.fmean = e / wam_maxf(P.s_fm, t.EMIN);
            P.f1mean = 1.f / (e > t.EMIN ? e / P.s_tm : 1.f);
            if (SHALLOW) {
                P.akmean = e > t.EMIN ? wam_ipowf(e / P.s_w1, 2) : 1.f;
                P.xkmean = e > t.EMIN ? wam_ipowf(P.s_w2 / e, 2) : 1.f;
            }
            if (t.brk) sfbrk_rates(e, P.fmean, P.depth, P.sbr, P.dsbr);
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
        // 4. limiter and implicit update; SOURCE_PHILLIPS' TEMP(M) and TPHOLD(K)
        //    into pm (free after the dissipation)
        if (tid < ML) {
            const float delfl = (t.C5G / wam_ipowf(s.fr[tid], 4)) * delt;
            t2[tid] = (P.ust * delfl) * wam_maxf(P.fmeanws, P.fmean);
            if (t.phil) pm[tid] = phillips_temp<SHALLOW>(t, s.fr[tid], P.ust, SHALLOW ? tfk[tid] : 0.f, P.depth);
        }
        if (t.phil && tid >= 64 && tid - 64 < KL) pm[MAXF + tid - 64] = phillips_tph(s.th[tid - 64], P.udir, P.ust);
        __syncthreads();
        for (int b = tid; b < NB; b += NT) {
            const int m = b / KL, k = b - KL * m;
            float sl = s.SL[b], flv = s.FL[b];
            if (SHALLOW) sl = sl + s.SLBOT[b];
            if (t.brk) {                          // SFBRK (:640-642)
                sl = sl + P.sbr * s.F[b];
                flv = flv + P.dsbr;
            }
            if (t.phil) sl = sl + pm[m] * pm[MAXF + k];   // SOURCE_PHILLIPS (:644-646)
            const float temp1 = (delt * sl) / wam_maxf(1.f, 1.f - delt * flv);
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
// its sequential form.  Returns TAUW; PHIAW in phiaw.  keep (the adjoint's
// recomputation) receives the stress components after the high-frequency
// part, TAUHF, TAU_PHI_HF's stress integral, h1, the modulus before its
// clamps and CONST1.
__device__ __forceinline__ float stresso_close(const DevTables &t, int mij, float ust, float z0, float udir, float phi,
                                               float xs, float ys, float h1, float h2, float &phiaw,
                                               float *keep = nullptr)
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
    if (keep) {
        keep[0] = xs; keep[1] = ys; keep[2] = tauhf; keep[3] = tau1; keep[4] = h1; keep[5] = tw; keep[6] = c1;
    }
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

// ADJ = 1 is the adjoint's recomputation of the same stages (identical
// values): it keeps LLWS of the first SINPUT, COS(TH - UDIR), the scalar
// chains' derivatives and STRESSO's values in e, and leaves the caller's
// TAUW, USTAR, Z0 and PHIAW as they are.  BRK = 1 (shallow water with
// WAVE_BREAKING_RUN) scales the spectrum by SDEPTHLIM's factor as it is read
// and leaves SFBRK's rates for implsch_snl_kernel.
template <int SHALLOW, int ADJ = 0, int BRK = 0>
__global__ void __launch_bounds__(PW) __attribute__((amdgpu_waves_per_eu(4))) implsch_pre_kernel(DevTables t, ImplArgs a, ImplScratch x, AdjExtra e)
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
        if (ADJ && on) e.cu[(size_t)k * XS + lp] = c;
    }
    float ja[4];
    auto put = [&](int slot, float v) {
        if (ADJ && on) e.sc[(size_t)slot * XS + lp] = v;
    };
    if (ADJ) airsea_d(t, u10, tauw, ust, z0, ja);
    else airsea(t, u10, tauw, ust, z0);

    const float const3 = ((2.f * t.XKAPPA) / t.BETAMAX) * (t.XKAPPA * t.XKAPPA);
    const float d1 = sqrtf(t.G) / t.ZPI, d2 = t.ZPI / sqrtf(t.G);
    // 0. SDEPTHLIM (:513-515, :3735-3780): TOTAL_ENERGY of the spectrum as
    //    given, then the factor EM that every later read of F applies
    float em = 1.f, emean_lim = 0.f;
    if (BRK) {
        float se = 0.f, Tl = 0.f;
        for (int i = 0; i < ML; i++) {
            const int m = i == 0 ? ML - 1 : i - 1;
            float T = 0.f;
            for (int k = 0; k < KL; k++) {
                const float f = F[(size_t)(k + KL * m) * NS];
                T = k == 0 ? f : T + f;
            }
            if (i == 0) { Tl = T; se = t.MO_TAIL * T; }
            else se = se + T * t.DFIM[m];
        }
        se = se + Tl * t.DFIM[ML - 1];
        const float e0 = wam_maxf(se, t.EMIN);
        em = sdepthlim_em(t.DEPTH[pt], e0);
        emean_lim = em * e0;
    }
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
                for (int u = 0; u < 4; u++) {
                    f4[u] = F[(size_t)(k0 + u + KL * m) * NS];
                    if (BRK) f4[u] = f4[u] * em;
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = k0 + u;
                    const float f = f4[u];
                    T = k == 0 ? f : T + f;
                    float sp, fl;
                    bool ll;
                    sin_bin(t, r, cu[k * W], const3, sp, fl, ll);
                    if (ADJ && on) e.llws1[(size_t)(k + KL * m) * XS + lp] = ll;
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
    const float emean = BRK ? emean_lim : wam_maxf(s_e, t.EMIN);
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
        const float f = BRK ? F[(size_t)(k + KL * (mij - 1)) * NS] * em : F[(size_t)(k + KL * (mij - 1)) * NS];
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
        float phiaw, keep[7];
        tauw = stresso_close(t, mij, ust, z0, udir, phi, xs, ys, h1, h2, phiaw, ADJ ? keep : nullptr);
        if (ADJ) {
            put(AS_UST1, ust); put(AS_Z01, z0);
            for (int i = 0; i < 4; i++) put(AS_JA1 + i, ja[i]);
            const D2 sg = wsigstar_d(t, ust, z0, wstar);
            put(AS_SIG1, sg.v); put(AS_JS1, sg.a); put(AS_JS1 + 1, sg.b);
            float j0, j1;
            tau_jac(t, t.FR, mij, ust, z0, j0, j1);
            put(AS_JT1, j0); put(AS_JT1 + 1, j1);
            for (int i = 0; i < 7; i++) put(AS_K1 + i, keep[i]);
            put(AS_SE, s_e); put(AS_SFM, s_fm); put(AS_STM, s_tm);
            put(AS_EMEAN, emean); put(AS_FMEAN, fmean); put(AS_F1MEAN, f1mean);
            put(AS_SEWS, q_e); put(AS_SFMWS, q_fm); put(AS_EMEANWS, ews); put(AS_FMEANWS, fmeanws);
            put(AS_MIJ, (float)mij); put(AS_TW1, tauw);
        }
    }
    if (ADJ) airsea_d(t, u10, tauw, ust, z0, ja);
    else airsea(t, u10, tauw, ust, z0);

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
                for (int u = 0; u < 4; u++) {
                    f4[u] = F[(size_t)(k0 + u + KL * mrow) * NS];
                    if (BRK) f4[u] = f4[u] * em;
                }
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
    float phiaw, keep[7];
    tauw = stresso_close(t, mij, ust, z0, udir, phi, xs, ys, h1, h2, phiaw, ADJ ? keep : nullptr);
    if (ADJ) {
        put(AS_UST2, ust); put(AS_Z02, z0);
        for (int i = 0; i < 4; i++) put(AS_JA2 + i, ja[i]);
        const D2 sg = wsigstar_d(t, ust, z0, wstar);
        put(AS_SIG2, sg.v); put(AS_JS2, sg.a); put(AS_JS2 + 1, sg.b);
        float j0, j1;
        tau_jac(t, t.FR, mij, ust, z0, j0, j1);
        put(AS_JT2, j0); put(AS_JT2 + 1, j1);
        for (int i = 0; i < 7; i++) put(AS_K2 + i, keep[i]);
    }
    if (on) {
        if (!ADJ) {
            a.tauw[mp] = tauw;
            a.ustar[mp] = ust;
            a.z0[mp] = z0;
            if (a.phiaw) a.phiaw[mp] = phiaw;
        }
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
        if (BRK) {       // SFBRK's rates on EMEAN after SDEPTHLIM and FMEAN of 2.3
            float sbr, dsbr;
            sfbrk_rates(emean, fmean, t.DEPTH[pt], sbr, dsbr);
            x.pp[(size_t)6 * XS + lp] = sbr;
            x.pp[(size_t)7 * XS + lp] = dsbr;
        }
    }
    // implsch_snl_kernel's point tickets start from zero
    if (blockIdx.x == 0 && lane < 8) x.tick[32 * lane] = 0u;
}

// IMPLSCH up to the dissipation for IPHYS = 2 (ST6), one lane per sea point,
// feeding the same SNONLIN and post kernels as IPHYS = 0.  The stages and
// their operation order are the single kernel's (implsch_kernel<2, *>):
// W3FLX4, the mean parameters, SINPUT_ST6 with LFACTOR and TAUWINDS
// (:4704-4925, :4465-4700, :5555-5587), FRCUTINDEX_ST6, STRESSO (PHIAW only),
// IMPHFTAIL, SDISSIP_ST6 and W3SWL6 (:4929-5252).  FRCUTINDEX_ST6 needs FMEAN
// and USTAR only, so MIJ and the tail are known before SINPUT_ST6's last pass,
// and W3FLX4 depends on U10 alone (its second call gives the same USTAR, Z0).
// Three passes over the spectrum:
//   A: the mean-parameter integrals and, per frequency, SINPUT_ST6's
//      saturation factor SQRT(BN), C^-1 and the LFACTOR stress components;
//   C: SDISSIP_ST6's and W3SWL6's per-frequency factors on the spectrum after
//      IMPHFTAIL (rows from MIJ up replaced by the tail);
//   B: SINPUT_ST6 per bin with the stress factor (its tanh terms evaluated
//      again, the same values as in pass A), the dissipation, STRESSO's
//      direction sums, the LLWS-masked integrals, the tiled outputs.
// LFACTOR's iteration runs per lane between A and C.  x.st6 holds the per-
// frequency values: SQRT(BN), C^-1, the two stress components, T12 and
// SQRT(BNM) ([6][ML][NX]).  BRK = 1 applies SDEPTHLIM as in implsch_pre_kernel.
template <int SHALLOW, int BRK = 0>
__global__ void __launch_bounds__(PW) __attribute__((amdgpu_waves_per_eu(4)))
implsch_pre_st6_kernel(DevTables t, ImplArgs a, ImplScratch x)
{
    extern __shared__ float cu_lds[];            // [KL][W]: COS(TH(K) - UDIR) of each lane's point
    const int KL = t.KL, ML = t.ML, NB = KL * ML, NS = t.NSEA, NB4 = x.NB4, XS = x.NX, lane = threadIdx.x;
    const int W = blockDim.x;
    const int lp0 = blockIdx.x * W + lane;
    const bool on = lp0 < x.NP;
    const int lp = on ? lp0 : x.NP - 1, mp = x.P0 + lp;      // member-point
    const int mem = mp / NS, pt = mp - mem * NS;            // ensemble member, sea point
    float *cu = cu_lds + lane;
    const float *F = a.fl3 + (size_t)mem * NB * NS + pt;   // F(b) = F[b * NS]
    const float u10 = a.u10[mp], udir = a.udir[mp], roairn = a.roairn[mp];
    const int indep = SHALLOW ? t.INDEP[pt] : 1;
    const float depth = t.DEPTH[pt];
    float *SQB = x.st6 + lp, *CIV = SQB + (size_t)ML * XS, *SXM = CIV + (size_t)ML * XS, *SYM = SXM + (size_t)ML * XS;
    float *T12 = SYM + (size_t)ML * XS, *SBN = T12 + (size_t)ML * XS;    // row m at [m * XS]
    for (int k = 0; k < KL; k++) {
        const float c = wam_cosf(t.TH[k] - udir);
        cu[k * W] = c;
        if (on) x.sprd[(size_t)k * XS + lp] = wam_ipowf(wam_maxf(0.f, c), 2);
    }
    float ust, z0;
    w3flx4(t, u10, ust, z0);
    const float cosu = wam_cosf(udir), sinu = wam_sinf(udir);
    const float uproxy = 32.f * ust, rwr = roairn / t.ROWATER;
    const float ZPI = t.ZPI;
    const float d1 = sqrtf(t.G) / ZPI, d2 = ZPI / sqrtf(t.G);
    // SDEPTHLIM (:513-515): TOTAL_ENERGY of the spectrum as given first
    float em = 1.f, emean_lim = 0.f;
    if (BRK) {
        float se = 0.f, Tl = 0.f;
        for (int i = 0; i < ML; i++) {
            const int m = i == 0 ? ML - 1 : i - 1;
            float T = 0.f;
            for (int k = 0; k < KL; k++) {
                const float f = F[(size_t)(k + KL * m) * NS];
                T = k == 0 ? f : T + f;
            }
            if (i == 0) { Tl = T; se = t.MO_TAIL * T; }
            else se = se + T * t.DFIM[m];
        }
        se = se + Tl * t.DFIM[ML - 1];
        const float e0 = wam_maxf(se, t.EMIN);
        em = sdepthlim_em(depth, e0);
        emean_lim = em * e0;
    }
    auto fread = [&](int b) { return BRK ? F[(size_t)b * NS] * em : F[(size_t)b * NS]; };
    // SINPUT_ST6's input term D of one bin (and its W2), as the single kernel forms it
    auto dterm = [&](int k, float sg, float ci, float sqb, float &w2) {
        const float cosd = t.COSTH[k] * cosu + t.SINTH[k] * sinu;
        const float xv = ((uproxy * ci) * cosd) - 1.f;
        const float w1 = wam_ipowf(wam_maxf(0.f, xv), 2);
        w2 = wam_ipowf(wam_minf(0.f, xv), 2);
        return (((rwr * sg) * (2.8f - (1.f + wam_tanhf((10.f * sqb) * w1 - 11.f)))) * sqb) * w1;
    };

    // pass A
    float s_e = 0.f, s_fm = 0.f, s_tm = 0.f, s_w1 = 0.f, s_w2 = 0.f, Tl = 0.f;
    for (int i = 0; i < ML; i++) {
        const int m = i == 0 ? ML - 1 : i - 1;
        const float wnm = t.WN[(size_t)m * NS + pt], cgm = t.CGG[(size_t)m * NS + pt];
        const float sg = ZPI * t.FR[m], c0 = ZPI * sg;
        float T = 0.f, sm = 0.f, mx = 0.f;
        for (int k = 0; k < KL; k++) {
            const float f = fread(k + KL * m);
            T = k == 0 ? f : T + f;
            const float av = (f * cgm) / c0;
            sm = k == 0 ? av : sm + av;
            mx = k == 0 ? av : wam_maxf(mx, av);
        }
        float sn = 0.f;
        for (int k = 0; k < KL; k++) {
            const float av = (fread(k + KL * m) * cgm) / c0;
            const float v = (mx < 1.0E-34f) ? 1.f : av / mx;
            sn = k == 0 ? v : sn + v;
        }
        const float anar = 1.0f / (sn * t.DELTH);
        const float adens = (sm * sg) * t.DELTH;
        const float sqb = sqrtf((anar * adens) * wam_ipowf(wnm, 3));
        const float ci = wnm / sg;
        float bx = 0.f, by = 0.f;
        for (int k = 0; k < KL; k++) {
            float w2;
            const float d = dterm(k, sg, ci, sqb, w2);
            const float av = (fread(k + KL * m) * cgm) / c0;
            const float sds = ((d * av) * sg) / cgm;
            const float vp = wam_maxf(0.f, sds);
            if (k == 0) { bx = vp * t.COSTH[k]; by = vp * t.SINTH[k]; }
            else { bx = bx + vp * t.COSTH[k]; by = by + vp * t.SINTH[k]; }
        }
        if (on) {
            SQB[(size_t)m * XS] = sqb;
            CIV[(size_t)m * XS] = ci;
            SXM[(size_t)m * XS] = bx * t.DELTH;
            SYM[(size_t)m * XS] = by * t.DELTH;
        }
        if (i == 0) {
            Tl = T;
            s_e = t.MO_TAIL * T; s_fm = t.MM1_TAIL * T; s_tm = t.MP1_TAIL * T;
            if (SHALLOW) { s_w1 = (t.MM1_TAIL * d1) * T; s_w2 = (t.MP1_TAIL * d2) * T; }
        } else {
            s_e = s_e + T * t.DFIM[m]; s_fm = s_fm + T * t.DFIMOFR[m]; s_tm = s_tm + T * t.DFIM_FR[m];
            if (SHALLOW) {
                const float sk = sqrtf(t.TFAK[tb(t, m, indep)]);
                s_w1 = s_w1 + (T / sk) * t.DFIM[m];
                s_w2 = s_w2 + (T * sk) * t.DFIM[m];
            }
        }
    }
    {
        const int m = ML - 1;
        s_e = s_e + Tl * t.DFIM[m]; s_fm = s_fm + Tl * t.DFIMOFR[m]; s_tm = s_tm + Tl * t.DFIM_FR[m];
        if (SHALLOW) {
            const float sk = sqrtf(t.TFAK[tb(t, m, indep)]);
            s_w1 = s_w1 + (Tl / sk) * t.DFIM[m];
            s_w2 = s_w2 + (Tl * sk) * t.DFIM[m];
        }
    }
    const float emean = BRK ? emean_lim : wam_maxf(s_e, t.EMIN);
    const float fmean = emean / wam_maxf(s_fm, t.EMIN);
    float akmean = 1.f;
    if (SHALLOW) akmean = emean > t.EMIN ? wam_ipowf(emean / s_w1, 2) : 1.f;
    const int mij = frcut(t, t.FR, fmean, fmean, ust);      // FRCUTINDEX_ST6: FMEAN and USTAR

    // LFACTOR (:4465-4700) with TAUWINDS: the stress supported by the waves
    // over the frequencies and their 10-Hz extension, NK10 terms in order
    const int NK = ML, NK10 = t.NK10;
    const float GRW = t.G * t.ROWATER;
    const float sxl = SXM[(size_t)(NK - 1) * XS], syl = SYM[(size_t)(NK - 1) * XS];
    auto term = [&](int i, bool with_lf, float rtau, float &tx, float &ty) {
        float ci, sx, sy;
        if (i < NK) {
            ci = CIV[(size_t)i * XS]; sx = SXM[(size_t)i * XS]; sy = SYM[(size_t)i * XS];
        } else {
            ci = t.SIG10[i] * 0.101978f;
            const float r2 = wam_ipowf(t.SIG10[NK - 1] / t.SIG10[i], 2);
            sx = sxl * r2; sy = syl * r2;
        }
        const float ds = t.DSII10[i];
        if (with_lf) {
            const float lf = wam_minf(1.0f, wam_expf((1.0f - uproxy * ci) * rtau));
            tx = ((sx * lf) * ci) * ds;
            ty = ((sy * lf) * ci) * ds;
        } else {
            tx = (sx * ci) * ds;
            ty = (sy * ci) * ds;
        }
    };
    auto stress = [&](bool with_lf, float rtau, float &twx, float &twy) {
        term(0, with_lf, rtau, twx, twy);
        for (int i = 1; i < NK10; i++) {
            float tx, ty;
            term(i, with_lf, rtau, tx, ty);
            twx = twx + tx;
            twy = twy + ty;
        }
        twx = GRW * twx;
        twy = GRW * twy;
    };
    const float tau_tot = (ust * ust) * roairn;
    float tau_vis = (wam_maxf(0.0f, -5.0E-5f * u10 + 1.1E-3f) * (u10 * u10)) * roairn;
    tau_vis = wam_minf(0.95f * tau_tot, tau_vis);
    const float tauvx = tau_vis * cosu, tauvy = tau_vis * sinu;
    float twx, twy;
    stress(false, 0.f, twx, twy);
    float taux = tauvx + twx, tauy = tauvy + twy;
    float tau = sqrtf(taux * taux + tauy * tauy);
    float err = (tau - tau_tot) / tau_tot;
    bool iterated = false;
    float rtau_lf = 0.f;                   // the RTAU of the last stress factors evaluated
    if (tau > tau_tot) {
        bool overshot = false;
        float rtau = err / 90.f, drtau = 2.0f;
        int sign_new = err >= 0.f ? 1 : -1;
        for (int it = 1; it <= 80; it++) {
            iterated = true;
            rtau_lf = rtau;
            stress(true, rtau, twx, twy);
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
    auto lfac = [&](float ci) { return iterated ? wam_minf(1.0f, wam_expf((1.0f - uproxy * ci) * rtau_lf)) : 1.0f; };
    bool apply;
    {
        float ls = lfac(CIV[0]);
        for (int m = 1; m < NK; m++) ls = ls + lfac(CIV[(size_t)m * XS]);
        apply = ls < (float)NK;
    }
    const float tauw = sqrtf(twx * twx + twy * twy);

    // IMPHFTAIL's factors: rows from MIJ up are TTAIL(M) * F(., MIJ - 1)
    float tden;
    if (SHALLOW) {
        const float akm0 = 1.f / t.TFAK[tb(t, mij - 1, indep)];
        tden = wam_ipowf(akm0, 3) / t.TCGOND[tb(t, mij - 1, indep)];
    } else {
        tden = t.FRM5[mij - 1];
    }
    auto ttail = [&](int m) {
        const float an = SHALLOW ? wam_ipowf(1.f / t.TFAK[tb(t, m, indep)], 3) / t.TCGOND[tb(t, m, indep)] : t.FRM5[m];
        return an / tden;
    };

    // pass C: SDISSIP_ST6 (:4929-5102) and W3SWL6 (:5106-5252) per frequency
    uint64_t abpos = 0;
    float b1;
    {
        float run = 0.f, cur = 0.f, smt = 0.f;
        int ikx = 0;
        const float xfac = (1.0f - 1.0f / t.CO) / (t.CO - 1.0f / t.CO);
        const float bnt = 0.035f * 0.035f;
        for (int m = 0; m < ML; m++) {
            const bool tl = m >= mij;
            const float tt = tl ? ttail(m) : 1.f;
            const int row = tl ? mij - 1 : m;
            auto fp = [&](int k) { return tl ? tt * fread(k + KL * row) : fread(k + KL * m); };
            const float wnm = t.WN[(size_t)m * NS + pt], cgm = t.CGG[(size_t)m * NS + pt];
            float ed = fp(0);
            for (int k = 1; k < KL; k++) ed = ed + fp(k);
            const float edens = ed * t.DELTH;
            const float sigm = ZPI * t.FR[m], c0 = ZPI * sigm;
            float ab = 0.f, mx = 0.f;
            for (int k = 0; k < KL; k++) {
                const float v = (fp(k) * cgm) / c0;
                ab = k == 0 ? v : ab + v;
                mx = k == 0 ? v : wam_maxf(mx, v);
            }
            float sn = 0.f;
            for (int k = 0; k < KL; k++) {
                const float v = (fp(k) * cgm) / c0;
                const float kn = (mx < 1.0E-34f) ? 1.f : v / mx;
                sn = k == 0 ? kn : sn + kn;
            }
            const float anar = 1.0f / (sn * t.DELTH);
            const float bnm = (anar * ((ab * sigm) * t.DELTH)) * wam_ipowf(wnm, 3);
            const float dden = (ZPI * t.DFIM[m]) * sigm;
            const float t2m = (ab * dden) / cgm;
            const float et = (ZPI * bnt) / ((1.0f * cgm) * wam_ipowf(wnm, 3));
            const float ex = wam_maxf(0.0f, edens - et);
            const float nex = ex / et;
            const float n4 = wam_ipowf(nex, 4);
            const float t1 = ((4.75E-6f * 1.0f) * t.FR[m]) * n4;
            float dfii = t.DF[m];
            if (m > 0 && m < ML - 1) dfii = dfii * xfac;
            const float termv = (1.0f * n4) * dfii;
            run = m == 0 ? termv : run + termv;
            if (m == 0 || ab > cur) { cur = ab; ikx = m; }
            smt = m == 0 ? t2m : smt + t2m;
            if (on) {
                T12[(size_t)m * XS] = -1.0f * (wam_maxf(0.0f, t1) + wam_maxf(0.0f, 7.00E-5f * run));
                SBN[(size_t)m * XS] = sqrtf(bnm);
            }
            if (ab > 1.E-30f) abpos |= 1ull << m;
        }
        b1 = 0.0041f * ((2.f * sqrtf(smt)) * t.WN[(size_t)ikx * NS + pt]);
    }

    // pass B: SINPUT_ST6 per bin, the dissipation, STRESSO's sums, the
    // LLWS-masked integrals; F after the tail, SL, FL (SL - SPOS) and LLWS out
    float q_e = 0.f, q_fm = 0.f, Ql = 0.f;
    float phi = 0.f, xs = 0.f, ys = 0.f, sal = 0.f, sxl2 = 0.f, syl2 = 0.f;
    {
        const size_t q0 = quad_offset(lp, 0, NB4);
        unsigned *LLW = x.llws + lp;
        for (int i = 0; i < ML; i++) {
            const int m = i == 0 ? ML - 1 : i - 1;
            const bool tl = m >= mij;
            const float tt = tl ? ttail(m) : 1.f;
            const float cgm = t.CGG[(size_t)m * NS + pt];
            const float sg = ZPI * t.FR[m];
            const float ci = CIV[(size_t)m * XS], sqb = SQB[(size_t)m * XS];
            const float lf = lfac(ci);
            const float t12 = T12[(size_t)m * XS];
            const float ddis = ((abpos >> m) & 1ull) ? (((-(2.f / 3.f)) * b1) * sg) * SBN[(size_t)m * XS] : 0.f;
            float Q = 0.f, sa = 0.f, sx = 0.f, sy = 0.f;
            bool qfirst = true;
            unsigned word = 0;
            for (int k0 = 0; k0 < KL; k0 += 4) {
                float f4[4], sl4[4], fl4[4], sm4[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = k0 + u, b = k + KL * m;
                    const float f = fread(b);
                    const float fpv = tl ? tt * fread(k + KL * (mij - 1)) : f;
                    float w2;
                    float d = dterm(k, sg, ci, sqb, w2);
                    if (apply) d = d * lf;
                    const float dpos = d;
                    d = d - ((((rwr * sg) * 0.09f) * (2.8f - (1.f + wam_tanhf((10.f * sqb) * w2 - 11.f)))) * sqb) * w2;
                    const float sl0 = d * f, sp = dpos * f;
                    const bool ll = sp > 0.f;
                    sm4[u] = sl0 - sp;
                    const float sl1 = sl0 + t12 * fpv, fl1 = d + t12;
                    sl4[u] = sl1 + ddis * fpv;
                    fl4[u] = fl1 + ddis;
                    f4[u] = fpv;
                    word |= (unsigned)ll << (b & 31);
                    if ((b & 31) == 31 || b == NB - 1) {
                        if (on) LLW[(size_t)(b >> 5) * XS] = word;
                        word = 0;
                    }
                    if (ll) {
                        Q = qfirst ? f : Q + f;
                        qfirst = false;
                    }
                    const float spp = wam_maxf(sp, 0.f);
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
            if (i == 0) {
                Ql = Q;
                q_e = t.MO_TAIL * Q; q_fm = t.MM1_TAIL * Q;
                sal = sa; sxl2 = sx; syl2 = sy;          // frequency ML's STRESSO terms go last
            } else {
                q_e = q_e + Q * t.DFIM[m]; q_fm = q_fm + Q * t.DFIMOFR[m];
                if (m < mij) {
                    const float rw = stresso_rw(t, m, mij);
                    const float cmrw = stresso_cm<SHALLOW>(t, m, SHALLOW ? t.TFAK[tb(t, m, indep)] : 0.f) * rw;
                    phi = phi + sa * rw;
                    xs = xs + sx * cmrw;
                    ys = ys + sy * cmrw;
                }
            }
        }
        const int m = ML - 1;
        q_e = q_e + Ql * t.DFIM[m]; q_fm = q_fm + Ql * t.DFIMOFR[m];
        if (m < mij) {
            const float rw = stresso_rw(t, m, mij);
            const float cmrw = stresso_cm<SHALLOW>(t, m, SHALLOW ? t.TFAK[tb(t, m, indep)] : 0.f) * rw;
            phi = phi + sal * rw;
            xs = xs + sxl2 * cmrw;
            ys = ys + syl2 * cmrw;
        }
    }
    const float ews = wam_maxf(q_e, t.EMIN);
    const float fmeanws = ews / wam_maxf(q_fm, t.EMIN);
    // STRESSO (TAUW not kept for ST6) on the moments of F(., MIJ)
    float h1 = 0.f, h2 = 0.f;
    for (int k = 0; k < KL; k++) {
        const float cosw = wam_maxf(cu[k * W], 0.f);
        const float f = fread(k + KL * (mij - 1));
        if (k == 0) { h2 = f * wam_ipowf(cosw, 2); h1 = h2 * cosw; }
        else { h1 = h1 + f * wam_ipowf(cosw, 3); h2 = h2 + f * wam_ipowf(cosw, 2); }
    }
    float phiaw;
    (void)stresso_close(t, mij, ust, z0, udir, phi, xs, ys, h1, h2, phiaw);
    if (on) {
        a.tauw[mp] = tauw;
        a.ustar[mp] = ust;
        a.z0[mp] = z0;
        if (a.phiaw) a.phiaw[mp] = phiaw;
        float enhfr = 1.f;
        if (SHALLOW) {
            const float enh = wam_maxf((0.75f * depth) * akmean, .5f);
            enhfr = 1.f + ((5.5f / enh) * (1.f - .833f * enh)) * wam_expf(-1.25f * enh);
        }
        x.pp[lp] = enhfr;
        x.pp[(size_t)XS + lp] = wam_maxf(fmeanws, fmean);
        x.pp[(size_t)2 * XS + lp] = (float)mij;
        x.pp[(size_t)3 * XS + lp] = phiaw;
        if (a.lcflx) {
            x.pp[(size_t)4 * XS + lp] = sinu;
            x.pp[(size_t)5 * XS + lp] = cosu;
        }
        if (BRK) {
            float sbr, dsbr;
            sfbrk_rates(emean, fmean, depth, sbr, dsbr);
            x.pp[(size_t)6 * XS + lp] = sbr;
            x.pp[(size_t)7 * XS + lp] = dsbr;
        }
    }
    if (blockIdx.x == 0 && lane < 8) x.tick[32 * lane] = 0u;
}

constexpr int SNL_NT = 512;  // threads of implsch_snl_kernel: one quadruplet of a chunk each
constexpr int SNL_NJ = 3;    // bins per thread: KL*ML <= SNL_NT * SNL_NJ
constexpr int SNL_QT = 256 * SP_QPT_DEV / SNL_NT;

// Three workgroups (LDS-limited) of 8 waves per CU: registers for 6 waves per SIMD.
// ADJ = 1 (the adjoint's recomputation) reads USTAR from e.sc and writes
// the final SL and FL over the work arrays' and the limiter's output to e.fn,
// leaving F after the tail in x.ft.  OPT = 1 adds the off-by-default sources
// the context switches on: SFBRK (rates from implsch_pre_kernel) and
// SOURCE_PHILLIPS, after SBOTTOM as IMPLSCH adds them (:639-646).
template <int SHALLOW, int ADJ = 0, int OPT = 0>
__global__ void __launch_bounds__(SNL_NT) __attribute__((amdgpu_waves_per_eu(6)))
implsch_snl_kernel(DevTables t, ImplArgs a, ImplScratch x, AdjExtra e)
{
    extern __shared__ __align__(16) float lds_raw[];
    __shared__ float sprd[MAXF], flm[MAXF], tfk[MAXF], tlim[MAXF], tbot[MAXF];
    __shared__ float Ps[OPT ? 7 : 5];
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
    float *ptemp = p, *ptph = p + ML;     // OPT: SOURCE_PHILLIPS' TEMP(M), TPHOLD(K)
    if (OPT) p += ML + KL;

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
            Ps[0] = ADJ ? e.sc[(size_t)AS_UST2 * XS + lp] : a.ustar[mp];
            Ps[1] = x.pp[lp]; Ps[2] = x.pp[(size_t)XS + lp]; Ps[3] = x.pp[(size_t)3 * XS + lp];
            Pmij = (int)x.pp[(size_t)2 * XS + lp];
            if (OPT && t.brk) { Ps[5] = x.pp[(size_t)6 * XS + lp]; Ps[6] = x.pp[(size_t)7 * XS + lp]; }
        }
        if (OPT && t.phil) {
            // SOURCE_PHILLIPS on USTAR of the second AIRSEA
            const float ust = a.ustar[mp];
            if (tid < ML) {
                const float tk = SHALLOW ? t.TFAK[tb(t, tid, t.INDEP[pt])] : 0.f;
                ptemp[tid] = phillips_temp<SHALLOW>(t, t.FR[tid], ust, tk, t.DEPTH[pt]);
            }
            if (tid >= 256 && tid - 256 < KL) ptph[tid - 256] = phillips_tph(t.TH[tid - 256], a.udir[mp], ust);
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
                if (OPT && t.brk) {
                    sv = sv + Ps[5] * s.F[b];
                    fv = fv + Ps[6];
                }
                if (OPT && t.phil) sv = sv + ptemp[m] * ptph[k];
                const float temp1 = (delt * sv) / wam_maxf(1.f, 1.f - delt * fv);
                const float temp3 = wam_minf(wam_fabsf(temp1), tlim[m]);
                const float f = s.F[b] + wam_signf(temp3, temp1);
                s.F[b] = wam_maxf(f, flm[m] * sprd[k]);
            }
        }
        __syncthreads();
        {
            const float4 *F4 = reinterpret_cast<const float4 *>(s.F);
            float *dst = ADJ ? e.fn : x.ft;
            for (int q = tid; q < NB4; q += SNL_NT) *reinterpret_cast<float4 *>(dst + q0 + (size_t)q * 256) = F4[q];
        }
        if (ADJ) {
#pragma unroll
            for (int j = 0; j < SNL_NJ; j++) {
                const int b = tid + SNL_NT * j;
                if (b < NB) {
                    const size_t o = q0 + (size_t)(b >> 2) * 256 + (b & 3);
                    x.slt[o] = sl[j];
                    x.flt[o] = fl[j];
                }
            }
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
    float *out = a.fl3 + (size_t)mem * KL * ML * NS + pt;   // FL3(pt, k, m) at out[(k + KL*m) * NS]
    // direction sums (plain and LLWS-masked) and their integrals, frequency ML
    // first; every bin is written to FL3 as it is read (the tail rows are
    // written again below, once MIJ is known)
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
                if (on) out[(size_t)(b0 + u) * NS] = f;
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
    // IMPHFTAIL: the rows from MIJ up, from row MIJ - 1
    for (int m = mij; m < ML; m++) {
        const float an = SHALLOW ? wam_ipowf(1.f / t.TFAK[tb(t, m, indep)], 3) / t.TCGOND[tb(t, m, indep)]
                                 : t.FRM5[m];
        const float ttail = an / tden;
        for (int k0 = 0; k0 < KL; k0 += 4) {
            const float4 f4 = *reinterpret_cast<const float4 *>(FQ + (size_t)((k0 + KL * (mij - 1)) >> 2) * 256);
            if (on) {
                const float fv[4] = {f4.x, f4.y, f4.z, f4.w};
#pragma unroll
                for (int u = 0; u < 4; u++) out[(size_t)(k0 + u + KL * m) * NS] = ttail * fv[u];
            }
        }
    }
    if (on && a.mij) a.mij[mp] = mij;
}

size_t implsch_snl_lds_bytes(const DevTables &t, int lcflx, int opt)
{
    const int ML = t.ML, KL = t.KL, MH = t.MLSTHG, NB = KL * ML;
    size_t n = 2 * 9 * (size_t)SP_TSTRIDE_DEV + (size_t)NB;
    if (lcflx) n += (t.shallow ? 2 : 1) * (size_t)NB;
    n += 2 * (size_t)ML + 3 * (size_t)KL + 36 * (size_t)MH + 2 * (size_t)t.SP_NCHUNK * t.SP_NRUN;
    if (opt) n += (size_t)ML + KL;
    return n * sizeof(float);
}

// Whether the three-kernel form applies: Janssen or ST6 physics, whole 4-bin
// quads per direction row, at most SNL_NJ bins per thread, one chunk's
// quadruplets of one frequency within the slots (ST6: at most 64 frequencies,
// as the single kernel assumes).
bool implsch_split_supported(const DevTables &t)
{
    return t.split && (t.iphys == 0 || (t.iphys == 2 && t.ML <= 64)) && t.KL % 4 == 0 &&
           t.KL * t.ML <= SNL_NT * SNL_NJ && 2 * t.KL < 256 * SP_QPT_DEV;
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
    const bool opt = t.brk || t.phil;
    if (t.iphys == 2) {
        if (t.shallow && t.brk)
            hipLaunchKernelGGL((implsch_pre_st6_kernel<1, 1>), dim3(nwg), dim3(W), lds1, st, t, a, x);
        else if (t.shallow) hipLaunchKernelGGL(implsch_pre_st6_kernel<1>, dim3(nwg), dim3(W), lds1, st, t, a, x);
        else hipLaunchKernelGGL(implsch_pre_st6_kernel<0>, dim3(nwg), dim3(W), lds1, st, t, a, x);
    } else if (t.shallow && t.brk)
        hipLaunchKernelGGL((implsch_pre_kernel<1, 0, 1>), dim3(nwg), dim3(W), lds1, st, t, a, x, AdjExtra{});
    else if (t.shallow) hipLaunchKernelGGL(implsch_pre_kernel<1>, dim3(nwg), dim3(W), lds1, st, t, a, x, AdjExtra{});
    else hipLaunchKernelGGL(implsch_pre_kernel<0>, dim3(nwg), dim3(W), lds1, st, t, a, x, AdjExtra{});
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    mark(1);
    void (*k)(DevTables, ImplArgs, ImplScratch, AdjExtra) =
        opt ? (t.shallow ? implsch_snl_kernel<1, 0, 1> : implsch_snl_kernel<0, 0, 1>)
            : (t.shallow ? implsch_snl_kernel<1> : implsch_snl_kernel<0>);
    const size_t lds = implsch_snl_lds_bytes(t, a.lcflx, opt);
    e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, SNL_NT, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    if (tu.snl_per_cu > 0 && tu.snl_per_cu < per_cu) per_cu = tu.snl_per_cu;
    const long want = (long)per_cu * ncu;
    const int nblocks = (int)(want < x.NP ? want : x.NP);
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(SNL_NT), lds, st, t, a, x, AdjExtra{});
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    mark(2);
    if (t.shallow) hipLaunchKernelGGL(implsch_post_kernel<1>, dim3(nwg), dim3(W), 0, st, t, a, x);
    else hipLaunchKernelGGL(implsch_post_kernel<0>, dim3(nwg), dim3(W), 0, st, t, a, x);
    e = hipGetLastError();
    mark(3);
    return e;
}

// The adjoint's recomputation of IMPLSCH up to the limiter (deep water):
// implsch_pre_kernel and implsch_snl_kernel with ADJ = 1 on the chunk of x.
hipError_t launch_implsch_adj_forward(const DevTables &t, const ImplArgs &a, const ImplScratch &x,
                                      const AdjExtra &ex, int ncu, hipStream_t st, const LaunchTune &tu)
{
    const int W = tu.src_wg;
    const int nwg = (x.NP + W - 1) / W;
    const size_t lds1 = (size_t)t.KL * W * sizeof(float);
    hipLaunchKernelGGL((implsch_pre_kernel<0, 1>), dim3(nwg), dim3(W), lds1, st, t, a, x, ex);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    void (*k)(DevTables, ImplArgs, ImplScratch, AdjExtra) = implsch_snl_kernel<0, 1>;
    const size_t lds = implsch_snl_lds_bytes(t, 0, 0);
    e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, SNL_NT, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    if (tu.snl_per_cu > 0 && tu.snl_per_cu < per_cu) per_cu = tu.snl_per_cu;
    const long want = (long)per_cu * ncu;
    const int nblocks = (int)(want < x.NP ? want : x.NP);
    hipLaunchKernelGGL(k, dim3(nblocks), dim3(SNL_NT), lds, st, t, a, x, ex);
    return hipGetLastError();
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
