// wam_adjoint.hip -- the adjoint of one model step, PROPAGS then IMPLSCH
// (wamodel.f90:332,354), on gfx950.
//
// The reference has no tangent-linear or adjoint model; these kernels are
// the transpose of the linearisation of this library's forward step at a
// given state, with every branch (MAX/MIN, the frequency cut-off MIJ, the
// wind-sea mask LLWS, the limiter's clamp and floor, AIRSEA's iteration
// count) taken as the forward step takes it there.
//
//  * implsch_adj_kernel: IMPLSCH for IPHYS = 0 in deep water.  One workgroup
//    per (member-)point recomputes the forward step in LDS with the stages
//    of the forward kernels (bit-identical values, so the same branches),
//    keeping what the reverse sweep needs, then runs the reverse sweep:
//    cut-off tail, limiter, SNONLIN, SDISSIP, STRESSO + SINPUT (second
//    call), AIRSEA, tail, STRESSO + SINPUT (first call), mean parameters,
//    AIRSEA.  The per-point scalar chains (AIRSEA's Newton iteration,
//    WSIGSTAR, TAU_PHI_HF's 19 nodes) carry forward derivatives along their
//    two inputs (dual numbers), whose local Jacobians the reverse sweep
//    applies.  SNONLIN's quadruplet terms are reversed per quadruplet and
//    their derivatives gathered per bin over host-built reader lists, so
//    every sum has a fixed order and the result is deterministic.
//  * propags_adj_kernel: the transposed stencil of PROP_NO_CURR / PROP_CURR
//    (wam_propagation_module.f90:750-825) over the stored weight planes,
//    as a gather: on a regular grid the neighbour relation is symmetric
//    (KLAT_S(i) = j exactly when KLAT_N(j) = i; the host checks it), so the
//    points whose stencil reads point j are j's own opposite neighbours.
//  * wind_uv_kernel / wind_uv_adj_kernel: U10 = SQRT(U**2 + V**2), UDIR =
//    ATAN2(U, V) (+ ZPI when negative), wam_wind_module.f90:1157-1167, and
//    its adjoint.
#include "wam_source_dev.hpp"
#include "wam_dual.hpp"

namespace wam {

// SINPUT of one bin (sin_bin's operations) with what its reverse needs: the
// wind-input terms' ZLOG, X and EXP(ZLOG) of each active NSIN term.
struct SinLin {
    float sp, fl;
    float z[2], x[2], e[2];
    bool on[2];
};

__device__ __forceinline__ SinLin sin_bin_lin(const DevTables &t, const SinFreq &r, float tk, float const3)
{
    SinLin L;
    float ufac1 = 0.f;
    L.on[0] = L.on[1] = false;
    if (tk > 0.01f) {
#pragma unroll
        for (int is = 0; is < 2; is++) {
            const float x = tk * (is == 0 ? r.ucn1 : r.ucn2);
            const float zlog = r.zcn + t.XKAPPA / x;
            if (zlog < 0.f) {
                const float z2x = (zlog * zlog) * x;
                const float e = wam_expf(zlog);
                ufac1 = ufac1 + ((0.5f * e) * z2x) * z2x;
                L.on[is] = true; L.z[is] = zlog; L.x[is] = x; L.e[is] = e;
            }
        }
    }
    float ufac2 = (const3 * (tk - r.xv1)) * wam_ipowf(r.ucn1, 2);
    ufac2 = 0.5f * ufac2 + ((0.5f * const3) * (tk - r.xv2)) * wam_ipowf(r.ucn2, 2);
    L.sp = r.cnsn * ufac1;
    L.fl = L.sp + r.cnsn * ufac2;
    return L;
}

// Reverse of one bin's SINPUT given the bars of SPOS/F (spb) and FL (flb):
// adds to the bar of COS(TH - UDIR) (dtk) and of the frequency's factors
// dm = {ZCN, UCN1, UCN2, XV1, XV2}.  A term 0.5 EXP(Z) Z**4 X**2 with
// Z = ZCN + XKAPPA/X has dterm/dZCN = 0.5 EXP(Z) X**2 Z**3 (Z + 4) and
// dterm/dX = EXP(Z) Z**3 (Z X - 0.5 XKAPPA (Z + 4)).
__device__ __forceinline__ void sin_bin_rev(const DevTables &t, const SinFreq &r, const SinLin &L, float tk,
                                            float const3, float spb, float flb, float &dtk, float (&dm)[5])
{
    const float u1b = r.cnsn * (spb + flb), u2b = r.cnsn * flb;
#pragma unroll
    for (int is = 0; is < 2; is++) {
        if (!L.on[is]) continue;
        const float z = L.z[is], x = L.x[is], e = L.e[is], z3 = (z * z) * z;
        const float dz = (((0.5f * e) * (x * x)) * z3) * (z + 4.f);
        const float dx = (e * z3) * (z * x - (0.5f * t.XKAPPA) * (z + 4.f));
        dm[0] += u1b * dz;
        const float xb = u1b * dx;
        dtk += xb * (is == 0 ? r.ucn1 : r.ucn2);
        dm[1 + is] += xb * tk;
    }
    const float h = 0.5f * const3;
    const float a1 = r.ucn1 * r.ucn1, a2 = r.ucn2 * r.ucn2;
    dtk += u2b * (h * (a1 + a2));
    dm[1] += u2b * ((const3 * (tk - r.xv1)) * r.ucn1);
    dm[2] += u2b * ((const3 * (tk - r.xv2)) * r.ucn2);
    dm[3] -= u2b * (h * a1);
    dm[4] -= u2b * (h * a2);
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Per-point scalars of the adjoint: the forward values it linearises at,
// the local Jacobians of the scalar chains and the running bars.
struct AdjScalars {
    float ja1[4], ja2[4];          // AIRSEA: d(UST, Z0)/d(U10, TAUW), first and second call
    float js1[2], js2[2];          // WSIGSTAR: dSIG_N/d(UST, Z0)
    float jt1[2], jt2[2];          // TAU_PHI_HF: dTAU1/d(UST, Z0)
    float ust1, z01, sig1, ust2, z02, sig2;
    float k1[8], k2[8];            // STRESSO's kept values, first and second call
    float s_e, s_fm, s_tm, s_ews, s_fmws, emean, fmean, f1mean, emeanws, fmeanws;
    int mij, mij2;
    // bars
    float ustb, z0b, twb, u10b, udirb;
    float fmeanb, fmeanwsb, emeanb, f1meanb;
    float xsb, ysb, t1b;
    float seb, sfmb, stmb, sewsb, sfmwsb;
};

// STRESSO's stress modulus reversed to its direction sums (per bin later),
// the cos**3 sum of F(:, MIJ) and the scalars: given the bar twb of the
// clamped TAUW; k the kept values (stresso's keep), jt dTAU1/d(UST, Z0).
__device__ __forceinline__ void stresso_rev_scalar(const DevTables &t, const float *k, float udir, float ust,
                                                   float twb, const float *jt, AdjScalars &A)
{
    const float xs = k[0], ys = k[1], tauhf = k[2], tau1 = k[3], t1 = k[4], tw = k[5], c1 = k[6];
    const float cap = ust * ust - t.EPS1;
    const float tmin = wam_minf(tw, cap);
    float tb = 0.f;
    if (tmin > 0.f) {
        if (tw < cap) tb = twb;
        else A.ustb += twb * (2.f * ust);
    }
    float xb = 0.f, yb = 0.f;
    if (tw > 0.f) { xb = tb * (xs / tw); yb = tb * (ys / tw); }
    const float su = wam_sinf(udir), co = wam_cosf(udir);
    const float hb = xb * su + yb * co;                  // bar of TAUHF = (CONST1 * T1) * TAU1
    A.udirb += tauhf * (xb * co - yb * su);
    A.t1b = hb * (c1 * tau1);
    const float tau1b = hb * (c1 * t1);
    A.ustb += tau1b * jt[0];
    A.z0b += tau1b * jt[1];
    const float ra = wam_maxf(t.ROAIR, 1.f);
    A.xsb = xb / ra;
    A.ysb = yb / ra;
}

// The frequencies' factors of SINPUT (sin_freq) reversed to USN1, USN2 and
// Z0, then USTAR and SIG_N: thread m < ML writes its three bars over dm's
// first rows; thread 0 then adds them up (after a barrier, by the caller).
__device__ __forceinline__ void sin_freq_rev(const DevTables &t, const SinFreq &r, float fr_m, float ust, float sig,
                                             float z0, float (&mb)[5][MAXF], int m)
{
    const float usn1 = ust * (1.0f - sig), usn2 = ust * (1.0f + sig);
    const float cm = (t.ZPI * fr_m) / t.G;
    const float d1 = ((usn1 / t.XKAPPA) * r.zcn) * cm, d2 = ((usn2 / t.XKAPPA) * r.zcn) * cm;
    const float g1 = mb[3][m] / (d1 * d1), g2 = mb[4][m] / (d2 * d2);
    const float zc = (r.zcn * cm) / t.XKAPPA;
    const float u1 = mb[1][m] * cm + g1 * zc;
    const float u2 = mb[2][m] * cm + g2 * zc;
    const float zb = (mb[0][m] + g1 * ((usn1 / t.XKAPPA) * cm)) + g2 * ((usn2 / t.XKAPPA) * cm);
    mb[0][m] = u1;
    mb[1][m] = u2;
    mb[2][m] = zb / z0;
}

__device__ __forceinline__ void sin_scalar_rev(const DevTables &t, const float (&mb)[5][MAXF], float ust, float sig,
                                               const float *js, AdjScalars &A)
{
    float u1 = 0.f, u2 = 0.f, z = 0.f;
    for (int m = 0; m < t.ML; m++) { u1 += mb[0][m]; u2 += mb[1][m]; z += mb[2][m]; }
    const float sigb = ust * (u2 - u1);
    A.ustb += (u1 * (1.0f - sig) + u2 * (1.0f + sig)) + sigb * js[0];
    A.z0b += z + sigb * js[1];
}

__host__ __device__ __forceinline__ size_t adj_lds_floats(int KL, int ML, int MH, int CHUNK)
{
    return (size_t)scr_floats(KL, ML, CHUNK, 0) + 7 * (size_t)ML + 3 * (size_t)KL + 26 * (size_t)MH +
           5 * (size_t)MH + 8 * (size_t)KL + 8 * (size_t)KL + 6 * (size_t)KL * ML;
}

__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2))) implsch_adj_kernel(DevTables t, AdjArgs a)
{
    extern __shared__ __align__(16) float lds_raw[];
    __shared__ PointScalars P;
    __shared__ AdjScalars A;
    __shared__ float T[MAXF], T2[MAXF], cu[MAXF], sprd[MAXF], t2[MAXF], t2b[MAXF], flm[MAXF], tmp2[MAXF], sdd[MAXF];
    __shared__ float pm[3 * MAXF + 64];
    __shared__ SinFreq q1[MAXF], q2[MAXF];
    __shared__ float mb[5][MAXF], kp[4][MAXF], cub[MAXF];
    __shared__ float hd[2][19][2];
    const int KL = t.KL, ML = t.ML, NB = KL * ML, NS = t.NSEA, tid = threadIdx.x, MH = t.MLSTHG;
    const int wv = tid >> 6, lane = tid & 63;

    Smem s;
    float *p = lds_raw;
    s.SCR = p;
    p += scr_floats(KL, ML, t.CHUNK, 0);
    s.fr = p; p += ML; s.dfim = p; p += ML; s.dfimofr = p; p += ML; s.dfim_fr = p; p += ML;
    s.rhowg = p; p += ML; s.fr5 = p; p += ML; s.frm5 = p; p += ML;
    s.th = p; p += KL; s.costh = p; p += KL; s.sinth = p; p += KL;
    s.rnl = p; p += 25 * MH; s.af11 = p; p += MH;
    s.sig10 = nullptr; s.dsii10 = nullptr;
    s.inl = (int *)p; p += 5 * MH; s.kw = (int *)p; p += 8 * KL;
    int *kinv = (int *)p; p += 8 * KL;
    s.F = p; p += NB; s.SL = p; p += NB; s.FL = p; p += NB; s.SPOS = p; p += NB;
    float *F0 = p; p += NB;
    float *FB = p; p += NB;
    s.SMIN = nullptr; s.SLBOT = nullptr;
    s.LLWS = (unsigned char *)p;
    unsigned char *LL1 = s.LLWS + NB;

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
    for (int i = tid; i < 8 * KL; i += NT) kinv[i] = t.AKINV[i];
    int qslot[2];
    for (int i = 0; i < 2; i++) {
        const int ql = tid + i * NT, mcl = ql / (2 * KL);
        qslot[i] = ql < 511 && mcl < t.CHUNK ? mcl | (ql - mcl * 2 * KL) << 8 : 255;
    }
    const float delt = t.DELT, gm1 = 1.0f / t.G;
    const float const3 = ((2.f * t.XKAPPA) / t.BETAMAX) * (t.XKAPPA * t.XKAPPA);

    const int NMP = a.NMP;
    const int nblk = gridDim.x, xcd = blockIdx.x % 8, lb = blockIdx.x / 8;
    const int nb_x = (nblk + 7 - xcd) / 8;
    const int per = (NMP + 7) / 8;
    const int p0 = xcd * per, p1 = min(NMP, p0 + per);
    for (int mp = p0 + lb; mp < p1; mp += nb_x) {
        const int mem = mp / NS, pt = mp - mem * NS;
        const size_t fo = (size_t)mem * NB * NS + pt;
        // ---------------- forward recomputation ----------------
        for (int b = tid; b < NB; b += NT) {
            const float f = a.fl3[fo + (size_t)b * NS];
            s.F[b] = f;
            F0[b] = f;
            FB[b] = a.fl3_bar[fo + (size_t)b * NS];
        }
        if (tid == 0) {
            P.u10 = a.u10[mp]; P.udir = a.udir[mp]; P.tauw = a.tauw[mp]; P.z0 = 0.f;
            P.roairn = a.roairn[mp]; P.wstar = a.wstar[mp]; P.depth = t.DEPTH[pt];
            P.indep = 1;
            P.ju = min(t.JUMAX, max(wam_nint(P.u10 / t.DELU), 1));
            A.ustb = a.ustar_bar ? a.ustar_bar[mp] : 0.f;
            A.z0b = a.z0_bar ? a.z0_bar[mp] : 0.f;
            A.twb = a.tauw_bar ? a.tauw_bar[mp] : 0.f;
            A.u10b = 0.f; A.udirb = 0.f;
        }
        __syncthreads();
        if (tid < KL) {
            const float c = wam_cosf(s.th[tid] - P.udir);
            cu[tid] = c;
            sprd[tid] = wam_ipowf(wam_maxf(0.f, c), 2);
        }
        if (tid >= 64 && tid - 64 < ML) flm[tid - 64] = t.FLMINFR[(size_t)(tid - 64) * t.JUMAX + P.ju - 1];
        if (tid == 128) {
            airsea_d(t, P.u10, P.tauw, P.ust, P.z0, A.ja1);
            A.ust1 = P.ust; A.z01 = P.z0;
        }
        dirsum(t, s.F, nullptr, T, tid);
        __syncthreads();
        if (tid == 0) P.s_e = integ(ML, t.MO_TAIL, T, s.dfim);
        else if (tid == 64) P.s_fm = integ(ML, t.MM1_TAIL, T, s.dfimofr);
        else if (tid == 192) P.s_tm = integ(ML, t.MP1_TAIL, T, s.dfim_fr);
        __syncthreads();
        if (tid == 0) {
            const float e = wam_maxf(P.s_e, t.EMIN);
            P.emean = e;
            P.fmean = e / wam_maxf(P.s_fm, t.EMIN);
            P.f1mean = 1.f / (e > t.EMIN ? e / P.s_tm : 1.f);
            A.s_e = P.s_e; A.s_fm = P.s_fm; A.s_tm = P.s_tm;
            A.emean = P.emean; A.fmean = P.fmean; A.f1mean = P.f1mean;
        }
        __syncthreads();
        sinput<0>(t, s, P, cu, nullptr, q1);
        dirsum(t, s.F, s.LLWS, T, tid);
        for (int b = tid; b < NB; b += NT) LL1[b] = s.LLWS[b];
        __syncthreads();
        if (tid == 0) P.s_e = integ(ML, t.MO_TAIL, T, s.dfim);
        else if (tid == 64) P.s_fm = integ(ML, t.MM1_TAIL, T, s.dfimofr);
        else if (tid == 128) {
            const D2 sg = wsigstar_d(t, A.ust1, A.z01, P.wstar);
            A.sig1 = P.sig_n; A.js1[0] = sg.a; A.js1[1] = sg.b;
        }
        __syncthreads();
        if (tid == 0) {
            const float e = wam_maxf(P.s_e, t.EMIN);
            P.emeanws = e;
            P.fmeanws = e / wam_maxf(P.s_fm, t.EMIN);
            P.mij = frcut(t, s.fr, P.fmean, P.fmeanws, P.ust);
            A.s_ews = P.s_e; A.s_fmws = P.s_fm; A.emeanws = P.emeanws; A.fmeanws = P.fmeanws; A.mij = P.mij;
        }
        __syncthreads();
        stresso<0>(t, s, P, cu, nullptr, pm, true, A.k1);
        if (tid == 0) airsea_d(t, P.u10, P.tauw, P.ust, P.z0, A.ja2);
        imphftail<0>(t, s, P.mij, nullptr, nullptr, t2, 64);
        if (tid >= 128 && tid < 128 + 19) {
            const D2 h = tau_node_d(t, s.fr, A.mij, A.ust1, A.z01, tid - 128 + 1);
            hd[0][tid - 128][0] = h.a; hd[0][tid - 128][1] = h.b;
        }
        __syncthreads();
        if (tid >= 64) {
            const int mij = P.mij;
            for (int b = tid - 64 + KL * mij; b < NB; b += NT - 64) {
                const int m = b / KL, k = b - KL * m;
                s.F[b] = t2[m] * s.F[k + KL * (mij - 1)];
            }
        }
        if (tid == 0) { A.ust2 = P.ust; A.z02 = P.z0; }
        __syncthreads();
        sinput<0>(t, s, P, cu, nullptr, q2);
        stresso<0>(t, s, P, cu, nullptr, pm, true, A.k2);
        // SDISSIP (:1593-1700) with FMEAN = F1MEAN, AKMEAN = XKMEAN (deep water)
        if (tid < ML) {
            const int m = tid;
            const float e2 = P.emean * P.emean;
            const float sds = (t.CONSD * e2) * wam_ipowf(P.f1mean, 9);
            const float x = wam_ipowf(s.fr[m] / P.f1mean, 2);
            sdd[m] = sds * (0.5f * x + 0.5f * wam_ipowf(x, 2));
        } else if (tid == 64) {
            const D2 sg = wsigstar_d(t, A.ust2, A.z02, P.wstar);
            A.sig2 = P.sig_n; A.js2[0] = sg.a; A.js2[1] = sg.b;
        } else if (tid >= 128 && tid < 128 + 19) {
            const D2 h = tau_node_d(t, s.fr, A.mij, A.ust2, A.z02, tid - 128 + 1);
            hd[1][tid - 128][0] = h.a; hd[1][tid - 128][1] = h.b;
        }
        __syncthreads();
        for (int b = tid; b < NB; b += NT) {
            const int m = b / KL;
            s.SL[b] = s.SL[b] + sdd[m] * s.F[b];
            s.FL[b] = s.FL[b] + sdd[m];
        }
        if (tid == 64) {
            float a0 = 0.f, a1 = 0.f, c0 = 0.f, c1 = 0.f;
            for (int j = 0; j < 19; j++) { a0 += hd[0][j][0]; a1 += hd[0][j][1]; c0 += hd[1][j][0]; c1 += hd[1][j][1]; }
            A.jt1[0] = a0; A.jt1[1] = a1; A.jt2[0] = c0; A.jt2[1] = c1;
        }
        __syncthreads();
        snonlin(t, s, 1.f, qslot);
        if (tid < ML) {
            const float delfl = (t.C5G / wam_ipowf(s.fr[tid], 4)) * delt;
            tmp2[tid] = (P.ust * delfl) * wam_maxf(P.fmeanws, P.fmean);
        }
        __syncthreads();
        for (int b = tid; b < NB; b += NT) {
            const int m = b / KL, k = b - KL * m;
            const float temp1 = (delt * s.SL[b]) / wam_maxf(1.f, 1.f - delt * s.FL[b]);
            const float temp3 = wam_minf(wam_fabsf(temp1), tmp2[m]);
            const float f = s.F[b] + wam_signf(temp3, temp1);
            s.SPOS[b] = wam_maxf(f, flm[m] * sprd[k]);
        }
        __syncthreads();
        dirsum(t, s.SPOS, nullptr, T, tid);
        dirsum(t, s.SPOS, s.LLWS, T2, tid - 64);
        __syncthreads();
        if (tid == 0) P.s_e = integ(ML, t.MO_TAIL, T, s.dfim);
        else if (tid == 64) P.s_fm = integ(ML, t.MM1_TAIL, T, s.dfimofr);
        else if (tid == 128) P.s_w1 = integ(ML, t.MO_TAIL, T2, s.dfim);
        else if (tid == 192) P.s_w2 = integ(ML, t.MM1_TAIL, T2, s.dfimofr);
        __syncthreads();
        if (tid == 0) {
            const float e = wam_maxf(P.s_e, t.EMIN);
            const float fm = e / wam_maxf(P.s_fm, t.EMIN);
            const float ew = wam_maxf(P.s_w1, t.EMIN);
            const float fmw = ew / wam_maxf(P.s_w2, t.EMIN);
            A.mij2 = frcut(t, s.fr, fm, fmw, P.ust);
        }
        __syncthreads();
        imphftail<0>(t, s, A.mij2, nullptr, nullptr, t2b, 0);
        __syncthreads();

        // ---------------- reverse sweep ----------------
        // cut-off tail of the new spectrum (MIJ of the new mean parameters)
        if (tid < KL) {
            const int k = tid, mij2 = A.mij2;
            float acc = 0.f;
            for (int m = mij2; m < ML; m++) {
                acc = acc + t2b[m] * FB[k + KL * m];
                FB[k + KL * m] = 0.f;
            }
            FB[k + KL * (mij2 - 1)] += acc;
        }
        __syncthreads();
        // limiter and implicit update: FB <- bar of F, SL <- bar of SL, FL <- bar of FL
        {
            float kacc = 0.f;
            for (int m = wv; m < ML; m += 4) {
                float t2s = 0.f;
                if (lane < KL) {
                    const int k = lane, b = k + KL * m;
                    const float fnb = FB[b];
                    const float y = 1.f - delt * s.FL[b];
                    const float den = wam_maxf(1.f, y);
                    const float temp1 = (delt * s.SL[b]) / den;
                    const float a1 = wam_fabsf(temp1);
                    const float temp3 = wam_minf(a1, tmp2[m]);
                    const float f = s.F[b] + wam_signf(temp3, temp1);
                    float t1b = 0.f, ftb = 0.f;
                    if (f > flm[m] * sprd[k]) {
                        ftb = fnb;
                        if (a1 < tmp2[m]) t1b = fnb;
                        else t2s = temp1 >= 0.f ? fnb : -fnb;
                    } else if (cu[k] > 0.f) {
                        kacc += fnb * (flm[m] * (2.f * cu[k]));
                    }
                    FB[b] = ftb;
                    s.SL[b] = (t1b * delt) / den;
                    s.FL[b] = !(1.f > y) ? ((t1b * temp1) * delt) / den : 0.f;
                }
                t2s = wave_sum(t2s);
                if (lane == 0) mb[0][m] = t2s;
            }
            if (lane < KL) kp[wv][lane] = kacc;
        }
        __syncthreads();
        if (tid < KL) cub[tid] = ((kp[0][tid] + kp[1][tid]) + kp[2][tid]) + kp[3][tid];
        if (tid == 64) {
            const float fmx = wam_maxf(A.fmeanws, A.fmean);
            float ub = 0.f, fb = 0.f;
            for (int m = 0; m < ML; m++) {
                const float delfl = (t.C5G / wam_ipowf(s.fr[m], 4)) * delt;
                ub += mb[0][m] * (delfl * fmx);
                fb += mb[0][m] * (A.ust2 * delfl);
            }
            A.ustb += ub;
            const bool ws = A.fmeanws > A.fmean;
            A.fmeanwsb = ws ? fb : 0.f;
            A.fmeanb = ws ? 0.f : fb;
        }
        __syncthreads();
        // SNONLIN: per chunk of frequencies, every quadruplet's bars of FIJ,
        // SAP and SAM from the bars of its nine updates, then every bin
        // gathers them over the quadruplets that read it
        {
            float *QF = s.SCR, *QP = s.SCR + 512, *QM = s.SCR + 1024;
            const float dal1 = t.DAL1, dal2 = t.DAL2;
            for (int c = 0; c < t.ANCH; c++) {
                const int mc0 = c * t.ACH, nmc = min(t.ACH, MH - mc0), nq = nmc * 2 * KL;
                for (int ql = tid; ql < nq; ql += NT) {
                    const int mcl = ql / (2 * KL), r = ql - mcl * 2 * KL, mcm1 = mc0 + mcl;
                    const int K = r < KL ? r : r - KL;
                    const int *ic = s.inl + mcm1 * 5;
                    const float *rc = s.rnl + mcm1 * 25;
                    const int k1 = s.kw[r], k2 = s.kw[2 * KL + r], k11 = s.kw[4 * KL + r], k21 = s.kw[6 * KL + r];
                    const int IC = (ic[0] - 1) * KL, IP = (ic[1] - 1) * KL, IP1 = (ic[2] - 1) * KL;
                    const int IM = (ic[3] - 1) * KL, IM1 = (ic[4] - 1) * KL;
                    const float *F = s.F;
                    const float sap = rc[1] * F[k1 + IP] + rc[2] * F[k11 + IP] + rc[3] * F[k1 + IP1] + rc[4] * F[k11 + IP1];
                    const float sam = rc[13] * F[k2 + IM] + rc[14] * F[k21 + IM] + rc[15] * F[k2 + IM1] + rc[16] * F[k21 + IM1];
                    const float fij = F[K + IC] * rc[0];
                    const float ftemp = s.af11[mcm1];
                    const float fcen = ftemp * fij;
                    const float fad1 = fij * (sap + sam), fad2 = fad1 - (2.f * sap) * sam;
                    const int8_t *tr = t.AQ_ROW + 9 * mcm1;
                    float adb = 0.f, ddb = 0.f, dmb = 0.f, dpb = 0.f;
                    if (tr[0] >= 0) {
                        const int b = K + KL * tr[0];
                        adb -= 2.f * s.SL[b];
                        ddb -= 2.f * s.FL[b];
                    }
#pragma unroll
                    for (int ty = 1; ty <= 8; ty++) {
                        if (tr[ty] < 0) continue;
                        const int kk = ty <= 4 ? ((ty & 1) ? k2 : k21) : ((ty & 1) ? k1 : k11);
                        const int b = kk + KL * tr[ty];
                        adb += rc[SLOT_SL[ty]] * s.SL[b];
                        const float d = rc[SLOT_FL[ty]] * s.FL[b];
                        if (ty <= 4) dmb += d;
                        else dpb += d;
                    }
                    // AD = FAD2*FCEN, DELAD = (FAD1 + FAD2)*FTEMP,
                    // DELAP = ((FIJ - 2 SAM)*DAL1)*FCEN, DELAM = ((FIJ - 2 SAP)*DAL2)*FCEN
                    float fad2b = adb * fcen, fcenb = adb * fad2;
                    const float fad1pb = ddb * ftemp;
                    fcenb += dpb * ((fij - 2.f * sam) * dal1) + dmb * ((fij - 2.f * sap) * dal2);
                    float fijb = dpb * (dal1 * fcen) + dmb * (dal2 * fcen);
                    float samb = -2.f * (dpb * (dal1 * fcen)), sapb = -2.f * (dmb * (dal2 * fcen));
                    fijb += fcenb * ftemp;
                    float fad1b = fad1pb;
                    fad2b += fad1pb;
                    fad1b += fad2b;
                    sapb -= (2.f * sam) * fad2b;
                    samb -= (2.f * sap) * fad2b;
                    fijb += fad1b * (sap + sam);
                    sapb += fad1b * fij;
                    samb += fad1b * fij;
                    QF[ql] = fijb; QP[ql] = sapb; QM[ql] = samb;
                }
                __syncthreads();
                for (int m = wv; m < ML; m += 4) {
                    const int e0 = t.AR_OFF[c * ML + m], e1 = t.AR_OFF[c * ML + m + 1];
                    if (lane < KL && e1 > e0) {
                        float acc = 0.f;
                        for (int e = e0; e < e1; e++) {
                            const int v = t.AR_E[e], mcm1 = v >> 4, sl = v & 15;
                            const float w = s.rnl[mcm1 * 25 + (sl <= 4 ? sl : sl + 8)];
                            const float *X = sl == 0 ? QF : (sl <= 4 ? QP : QM);
                            const int mpk = sl <= 4 ? ((sl - 1) & 1) : 2 + ((sl - 5) & 1);
                            const int base = (mcm1 - mc0) * 2 * KL;
#pragma unroll
                            for (int kh = 0; kh < 2; kh++) {
                                const int K = sl == 0 ? lane : kinv[(mpk * 2 + kh) * KL + lane];
                                acc = acc + w * X[base + kh * KL + K];
                            }
                        }
                        FB[lane + KL * m] += acc;
                    }
                }
                __syncthreads();
            }
        }
        // SDISSIP: bar of its per-frequency factor, then EMEAN and F1MEAN
        for (int m = wv; m < ML; m += 4) {
            float d = 0.f;
            if (lane < KL) {
                const int b = lane + KL * m;
                FB[b] += s.SL[b] * sdd[m];
                d = s.SL[b] * s.F[b] + s.FL[b];
            }
            d = wave_sum(d);
            if (lane == 0) T2[m] = d;
        }
        if (tid == 64) stresso_rev_scalar(t, A.k2, P.udir, A.ust2, A.twb, A.jt2, A);
        __syncthreads();
        if (tid == 0) {
            const float e = A.emean, f1 = A.f1mean;
            const float sds = (t.CONSD * (e * e)) * wam_ipowf(f1, 9);
            float sdsb = 0.f, f1b = 0.f;
            for (int m = 0; m < ML; m++) {
                const float x = wam_ipowf(s.fr[m] / f1, 2);
                sdsb += T2[m] * (0.5f * x + 0.5f * (x * x));
                const float xb = T2[m] * (sds * (0.5f + x));
                f1b += xb * ((-2.f * x) / f1);
            }
            A.emeanb = sdsb * ((t.CONSD * wam_ipowf(f1, 9)) * (2.f * e));
            A.f1meanb = f1b + sdsb * ((t.CONSD * (e * e)) * (9.f * wam_ipowf(f1, 8)));
        }
        // SINPUT and STRESSO, second call: bins
        {
            const int mij = A.mij;
            float kacc = 0.f;
            for (int m = wv; m < ML; m += 4) {
                float dm[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
                if (lane < KL) {
                    const int k = lane, b = k + KL * m;
                    const float ft = s.F[b], slb = s.SL[b], flb = s.FL[b];
                    const SinLin L = sin_bin_lin(t, q2[m], cu[k], const3);
                    float sposb = 0.f;
                    if (m < mij && L.sp * ft > 0.f) {
                        float rw = s.rhowg[m];
                        if (m == mij - 1 && mij != ML) rw = 0.5f * rw;
                        const float cmrw = ((t.ZPI * s.fr[m]) * gm1) * rw;
                        sposb = (A.xsb * cmrw) * s.sinth[k] + (A.ysb * cmrw) * s.costh[k];
                    }
                    float fbadd = slb * L.fl + sposb * L.sp;
                    float dtk = 0.f;
                    sin_bin_rev(t, q2[m], L, cu[k], const3, sposb * ft, slb * ft + flb, dtk, dm);
                    if (m == mij - 1 && cu[k] > 0.f) {
                        const float c = cu[k];
                        fbadd += A.t1b * ((c * c) * c);
                        dtk += (A.t1b * ft) * (3.f * (c * c));
                    }
                    FB[b] += fbadd;
                    kacc += dtk;
                }
#pragma unroll
                for (int i = 0; i < 5; i++) {
                    const float v = wave_sum(dm[i]);
                    if (lane == 0) mb[i][m] = v;
                }
            }
            if (lane < KL) kp[wv][lane] = kacc;
        }
        __syncthreads();
        if (tid < KL) cub[tid] += ((kp[0][tid] + kp[1][tid]) + kp[2][tid]) + kp[3][tid];
        if (tid >= 64 && tid - 64 < ML)
            sin_freq_rev(t, q2[tid - 64], s.fr[tid - 64], A.ust2, A.sig2, A.z02, mb, tid - 64);
        __syncthreads();
        if (tid == 0) {
            sin_scalar_rev(t, mb, A.ust2, A.sig2, A.js2, A);
            // AIRSEA, second call: (UST, Z0) from (U10, TAUW of the first STRESSO)
            A.u10b += A.ustb * A.ja2[0] + A.z0b * A.ja2[2];
            A.twb = A.ustb * A.ja2[1] + A.z0b * A.ja2[3];
            A.ustb = 0.f;
            A.z0b = 0.f;
        }
        // cut-off tail of the first pass (MIJ)
        if (tid >= 64 && tid - 64 < KL) {
            const int k = tid - 64, mij = A.mij;
            float acc = 0.f;
            for (int m = mij; m < ML; m++) {
                acc = acc + t2[m] * FB[k + KL * m];
                FB[k + KL * m] = 0.f;
            }
            FB[k + KL * (mij - 1)] += acc;
        }
        __syncthreads();
        if (tid == 0) stresso_rev_scalar(t, A.k1, P.udir, A.ust1, A.twb, A.jt1, A);
        if (tid == 64) {
            // mean parameters of the input spectrum: EMEAN, FMEAN (through
            // MAX(FMEANWS, FMEAN)), F1MEAN; and FMEANWS of the first LLWS
            const float e = A.emean;
            const float dfm = wam_maxf(A.s_fm, t.EMIN);
            float eb = A.emeanb + A.fmeanb / dfm, stmb = 0.f;
            if (e > t.EMIN) {
                const float q = e / A.s_tm;
                const float qb = -A.f1meanb / (q * q);
                eb += qb / A.s_tm;
                stmb = -(qb * q) / A.s_tm;
            }
            A.sfmb = A.s_fm > t.EMIN ? -(A.fmeanb * A.fmean) / dfm : 0.f;
            A.seb = A.s_e > t.EMIN ? eb : 0.f;
            A.stmb = stmb;
            const float dws = wam_maxf(A.s_fmws, t.EMIN);
            A.sewsb = A.s_ews > t.EMIN ? A.fmeanwsb / dws : 0.f;
            A.sfmwsb = A.s_fmws > t.EMIN ? -(A.fmeanwsb * A.fmeanws) / dws : 0.f;
        }
        __syncthreads();
        // SINPUT and STRESSO, first call, and the mean parameters: bins
        {
            const int mij = A.mij;
            float kacc = 0.f;
            for (int m = wv; m < ML; m += 4) {
                float dm[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
                if (lane < KL) {
                    const int k = lane, b = k + KL * m;
                    const float f0 = F0[b];
                    const SinLin L = sin_bin_lin(t, q1[m], cu[k], const3);
                    float sposb = 0.f;
                    if (m < mij && L.sp * f0 > 0.f) {
                        float rw = s.rhowg[m];
                        if (m == mij - 1 && mij != ML) rw = 0.5f * rw;
                        const float cmrw = ((t.ZPI * s.fr[m]) * gm1) * rw;
                        sposb = (A.xsb * cmrw) * s.sinth[k] + (A.ysb * cmrw) * s.costh[k];
                    }
                    const bool last = m == ML - 1;
                    const float tb = (A.seb * (s.dfim[m] + (last ? t.MO_TAIL : 0.f)) +
                                      A.sfmb * (s.dfimofr[m] + (last ? t.MM1_TAIL : 0.f))) +
                                     A.stmb * (s.dfim_fr[m] + (last ? t.MP1_TAIL : 0.f));
                    float fbadd = sposb * L.sp + tb;
                    if (LL1[b])
                        fbadd += A.sewsb * (s.dfim[m] + (last ? t.MO_TAIL : 0.f)) +
                                 A.sfmwsb * (s.dfimofr[m] + (last ? t.MM1_TAIL : 0.f));
                    float dtk = 0.f;
                    sin_bin_rev(t, q1[m], L, cu[k], const3, sposb * f0, 0.f, dtk, dm);
                    if (m == mij - 1 && cu[k] > 0.f) {
                        const float c = cu[k];
                        fbadd += A.t1b * ((c * c) * c);
                        dtk += (A.t1b * f0) * (3.f * (c * c));
                    }
                    FB[b] += fbadd;
                    kacc += dtk;
                }
#pragma unroll
                for (int i = 0; i < 5; i++) {
                    const float v = wave_sum(dm[i]);
                    if (lane == 0) mb[i][m] = v;
                }
            }
            if (lane < KL) kp[wv][lane] = kacc;
        }
        __syncthreads();
        if (tid < KL) cub[tid] += ((kp[0][tid] + kp[1][tid]) + kp[2][tid]) + kp[3][tid];
        if (tid >= 64 && tid - 64 < ML)
            sin_freq_rev(t, q1[tid - 64], s.fr[tid - 64], A.ust1, A.sig1, A.z01, mb, tid - 64);
        __syncthreads();
        if (tid == 0) {
            sin_scalar_rev(t, mb, A.ust1, A.sig1, A.js1, A);
            // AIRSEA, first call: (UST, Z0) from (U10, TAUW at the step's start)
            A.u10b += A.ustb * A.ja1[0] + A.z0b * A.ja1[2];
            const float tw0b = A.ustb * A.ja1[1] + A.z0b * A.ja1[3];
            // UDIR through COS(TH(K) - UDIR): d/dUDIR = SIN(TH(K) - UDIR)
            float ub = A.udirb;
            for (int k = 0; k < KL; k++) ub += cub[k] * wam_sinf(s.th[k] - P.udir);
            a.u10_bar[mp] = A.u10b;
            a.udir_bar[mp] = ub;
            a.tauw_bar_out[mp] = tw0b;
        }
        for (int b = tid; b < NB; b += NT) a.fl3_bar_out[fo + (size_t)b * NS] = FB[b];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// IMPLSCH's adjoint in kernels of the forward split's shape (IPHYS = 0, deep
// water, KL % 4 == 0): the per-point scalar chains run one sea point per
// lane, SNONLIN one point per workgroup, on the tiled work arrays.
//   implsch_pre_kernel<0, 1>, implsch_snl_kernel<0, 1> (wam_implsch.hip): the
//       forward recomputation up to the limiter, keeping what the reverse
//       needs (x: F after the tail, the final SL and FL, LLWS; e: the
//       limiter's output, LLWS of the first SINPUT, COS(TH - UDIR), scalars)
//   adj_post_kernel: the new cut-off index, then the bar of the limiter's
//       output (IMPHFTAIL transposed)
//   adj_snl_kernel: the limiter and SNONLIN reversed (per point)
//   adj_rev_kernel: SDISSIP, the second SINPUT and STRESSO, AIRSEA, the
//       tail, the first SINPUT and STRESSO, the mean parameters and AIRSEA
//       reversed (per lane), the input spectrum's bar out
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) adj_post_kernel(DevTables t, AdjArgs a, ImplScratch x, AdjExtra e)
{
    const int KL = t.KL, ML = t.ML, NS = t.NSEA, NB4 = x.NB4, XS = x.NX, lane = threadIdx.x;
    const int lp0 = blockIdx.x * blockDim.x + lane;
    const bool on = lp0 < x.NP;
    const int lp = on ? lp0 : x.NP - 1, mp = x.P0 + lp;
    const int mem = mp / NS, pt = mp - mem * NS;
    float *FQ = e.fn + quad_offset(lp, 0, NB4);
    const unsigned *LLW = x.llws + lp;
    const float ust = e.sc[(size_t)AS_UST2 * XS + lp];
    // the new mean parameters and the cut-off index, as implsch_post_kernel
    float s_e = 0.f, s_fm = 0.f, s_w1 = 0.f, s_w2 = 0.f, Tl = 0.f, Ql = 0.f;
    for (int i = 0; i < ML; i++) {
        const int m = i == 0 ? ML - 1 : i - 1;
        float T = 0.f, Q = 0.f;
        bool qfirst = true;
        for (int k0 = 0; k0 < KL; k0 += 4) {
            const int b0 = k0 + KL * m;
            const float4 f4 = *reinterpret_cast<const float4 *>(FQ + (size_t)(b0 >> 2) * 256);
            const unsigned w = LLW[(size_t)(b0 >> 5) * XS] >> (b0 & 31);
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
    const float ev = wam_maxf(s_e, t.EMIN);
    const float fmean = ev / wam_maxf(s_fm, t.EMIN);
    const float ew = wam_maxf(s_w1, t.EMIN);
    const float fmeanws = ew / wam_maxf(s_w2, t.EMIN);
    const int mij = frcut(t, t.FR, fmean, fmeanws, ust);
    // IMPHFTAIL transposed: F(K, M >= MIJ) = T2(M) F(K, MIJ) sends T2(M) times
    // their bars to row MIJ, the rows above get none
    const float tden = t.FRM5[mij - 1];
    const float *G = a.fl3_bar + (size_t)mem * KL * ML * NS + pt;
    for (int k0 = 0; k0 < KL; k0 += 4) {
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
        for (int m = mij; m < ML; m++) {
            const float t2 = t.FRM5[m] / tden;
#pragma unroll
            for (int u = 0; u < 4; u++) acc[u] = acc[u] + t2 * G[(size_t)(k0 + u + KL * m) * NS];
        }
        for (int m = 0; m < ML; m++) {
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const float g = m < mij ? G[(size_t)(k0 + u + KL * m) * NS] : 0.f;
                v[u] = m == mij - 1 ? g + acc[u] : g;
            }
            if (on)
                *reinterpret_cast<float4 *>(FQ + (size_t)((k0 + KL * m) >> 2) * 256) = make_float4(v[0], v[1], v[2], v[3]);
        }
    }
    if (on) e.sc[(size_t)AS_MIJ2 * XS + lp] = (float)mij;
}

// LDS words of adj_snl_kernel besides its static arrays: the quadruplet
// bars, the DIA tables, the reader lists and their offsets, the update rows,
// and the point's F, SL, FL and bar of F
__host__ __device__ __forceinline__ size_t adj_snl_lds_floats(int KL, int ML, int MH, int NAR, int ANCH)
{
    return 3 * 512 + 25 * (size_t)MH + (size_t)MH + 5 * (size_t)MH + 16 * (size_t)KL + (size_t)NAR +
           ((size_t)ANCH * ML + 1) + (9 * (size_t)MH + 3) / 4 + 4 * (size_t)KL * ML;
}

__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2)))
adj_snl_kernel(DevTables t, AdjArgs a, ImplScratch x, AdjExtra e)
{
    extern __shared__ __align__(16) float lds_raw[];
    __shared__ float cu[MAXF], sprd[MAXF], flm[MAXF], tmp2[MAXF], mb0[MAXF], kp[4][MAXF];
    __shared__ float Pv[3];
    const int KL = t.KL, ML = t.ML, NB = KL * ML, NB4 = x.NB4, XS = x.NX, tid = threadIdx.x, MH = t.MLSTHG;
    const int wv = tid >> 6, lane = tid & 63;
    float *QF = lds_raw, *QP = lds_raw + 512, *QM = lds_raw + 1024;
    float *p = lds_raw + 1536;
    float *rnl = p; p += 25 * MH;
    float *af11 = p; p += MH;
    int *inl = (int *)p; p += 5 * MH;
    int *kw = (int *)p; p += 8 * KL;
    int *kinv = (int *)p; p += 8 * KL;
    int *are = (int *)p; p += t.NAR;
    int *aro = (int *)p; p += t.ANCH * ML + 1;
    signed char *aqr = (signed char *)p; p += (9 * MH + 3) / 4;
    float *F = p; p += NB;
    float *SL = p; p += NB;
    float *FL = p; p += NB;
    float *FB = p; p += NB;
    for (int i = tid; i < t.NAR; i += NT) are[i] = t.AR_E[i];
    for (int i = tid; i < t.ANCH * ML + 1; i += NT) aro[i] = t.AR_OFF[i];
    for (int i = tid; i < 9 * MH; i += NT) aqr[i] = t.AQ_ROW[i];
    for (int i = tid; i < 25 * MH; i += NT) rnl[i] = t.RNLCOEF[i];
    for (int i = tid; i < 5 * MH; i += NT) inl[i] = t.INLCOEF[i];
    for (int i = tid; i < MH; i += NT) af11[i] = t.AF11[i];
    for (int i = tid; i < 2 * KL; i += NT) {
        kw[i] = t.K1W[i]; kw[2 * KL + i] = t.K2W[i]; kw[4 * KL + i] = t.K11W[i]; kw[6 * KL + i] = t.K21W[i];
    }
    for (int i = tid; i < 8 * KL; i += NT) kinv[i] = t.AKINV[i];
    const float delt = t.DELT;
    const int nblk = gridDim.x, xcd = blockIdx.x % 8, lb = blockIdx.x / 8;
    const int nb_x = (nblk + 7 - xcd) / 8;
    const int per = (x.NP + 7) / 8;
    const int p0 = xcd * per, p1 = min(x.NP, p0 + per);
    for (int lp = p0 + lb; lp < p1; lp += nb_x) {
        const int mp = x.P0 + lp;
        const size_t q0 = quad_offset(lp, 0, NB4);
        for (int q = tid; q < NB4; q += NT) {
            const size_t o = q0 + (size_t)q * 256;
            reinterpret_cast<float4 *>(F)[q] = *reinterpret_cast<const float4 *>(x.ft + o);
            reinterpret_cast<float4 *>(SL)[q] = *reinterpret_cast<const float4 *>(x.slt + o);
            reinterpret_cast<float4 *>(FL)[q] = *reinterpret_cast<const float4 *>(x.flt + o);
            reinterpret_cast<float4 *>(FB)[q] = *reinterpret_cast<const float4 *>(e.fn + o);
        }
        if (tid < KL) {
            const float c = e.cu[(size_t)tid * XS + lp];
            cu[tid] = c;
            sprd[tid] = wam_ipowf(wam_maxf(0.f, c), 2);
        }
        if (tid >= 64 && tid - 64 < ML) {
            const int m = tid - 64;
            const int ju = min(t.JUMAX, max(wam_nint(a.u10[mp] / t.DELU), 1));
            flm[m] = t.FLMINFR[(size_t)m * t.JUMAX + ju - 1];
        }
        if (tid == 128) {
            Pv[0] = e.sc[(size_t)AS_UST2 * XS + lp];
            Pv[1] = e.sc[(size_t)AS_FMEANWS * XS + lp];
            Pv[2] = e.sc[(size_t)AS_FMEAN * XS + lp];
        }
        __syncthreads();
        if (tid < ML) {
            const float delfl = (t.C5G / wam_ipowf(t.FR[tid], 4)) * delt;
            tmp2[tid] = (Pv[0] * delfl) * wam_maxf(Pv[1], Pv[2]);
        }
        __syncthreads();
        // the limiter and implicit update reversed: FB <- the bar of F after
        // the tail, SL and FL <- their bars
        {
            float kacc = 0.f;
            for (int m = wv; m < ML; m += 4) {
                float t2s = 0.f;
                if (lane < KL) {
                    const int k = lane, b = k + KL * m;
                    const float fnb = FB[b];
                    const float y = 1.f - delt * FL[b];
                    const float den = wam_maxf(1.f, y);
                    const float temp1 = (delt * SL[b]) / den;
                    const float a1 = wam_fabsf(temp1);
                    const float temp3 = wam_minf(a1, tmp2[m]);
                    const float f = F[b] + wam_signf(temp3, temp1);
                    float t1b = 0.f, ftb = 0.f;
                    if (f > flm[m] * sprd[k]) {
                        ftb = fnb;
                        if (a1 < tmp2[m]) t1b = fnb;
                        else t2s = temp1 >= 0.f ? fnb : -fnb;
                    } else if (cu[k] > 0.f) {
                        kacc += fnb * (flm[m] * (2.f * cu[k]));
                    }
                    FB[b] = ftb;
                    SL[b] = (t1b * delt) / den;
                    FL[b] = !(1.f > y) ? ((t1b * temp1) * delt) / den : 0.f;
                }
                t2s = wave_sum(t2s);
                if (lane == 0) mb0[m] = t2s;
            }
            if (lane < KL) kp[wv][lane] = kacc;
        }
        __syncthreads();
        if (tid < KL) e.cub[(size_t)tid * XS + lp] = ((kp[0][tid] + kp[1][tid]) + kp[2][tid]) + kp[3][tid];
        if (tid == 64) {
            const float fmx = wam_maxf(Pv[1], Pv[2]);
            float ub = 0.f, fb = 0.f;
            for (int m = 0; m < ML; m++) {
                const float delfl = (t.C5G / wam_ipowf(t.FR[m], 4)) * delt;
                ub += mb0[m] * (delfl * fmx);
                fb += mb0[m] * (Pv[0] * delfl);
            }
            e.sc[(size_t)AS_UST2B * XS + lp] = ub;
            e.sc[(size_t)AS_FMXB * XS + lp] = fb;
        }
        // SNONLIN reversed, chunk by chunk of frequencies
        {
            const float dal1 = t.DAL1, dal2 = t.DAL2;
            for (int c = 0; c < t.ANCH; c++) {
                const int mc0 = c * t.ACH, nmc = min(t.ACH, MH - mc0), nq = nmc * 2 * KL;
                for (int ql = tid; ql < nq; ql += NT) {
                    const int mcl = ql / (2 * KL), r = ql - mcl * 2 * KL, mcm1 = mc0 + mcl;
                    const int K = r < KL ? r : r - KL;
                    const int *ic = inl + mcm1 * 5;
                    const float *rc = rnl + mcm1 * 25;
                    const int k1 = kw[r], k2 = kw[2 * KL + r], k11 = kw[4 * KL + r], k21 = kw[6 * KL + r];
                    const int IC = (ic[0] - 1) * KL, IP = (ic[1] - 1) * KL, IP1 = (ic[2] - 1) * KL;
                    const int IM = (ic[3] - 1) * KL, IM1 = (ic[4] - 1) * KL;
                    const float sap = rc[1] * F[k1 + IP] + rc[2] * F[k11 + IP] + rc[3] * F[k1 + IP1] + rc[4] * F[k11 + IP1];
                    const float sam = rc[13] * F[k2 + IM] + rc[14] * F[k21 + IM] + rc[15] * F[k2 + IM1] + rc[16] * F[k21 + IM1];
                    const float fij = F[K + IC] * rc[0];
                    const float ftemp = af11[mcm1];
                    const float fcen = ftemp * fij;
                    const float fad1 = fij * (sap + sam), fad2 = fad1 - (2.f * sap) * sam;
                    const signed char *tr = aqr + 9 * mcm1;
                    float adb = 0.f, ddb = 0.f, dmb = 0.f, dpb = 0.f;
                    if (tr[0] >= 0) {
                        const int b = K + KL * tr[0];
                        adb -= 2.f * SL[b];
                        ddb -= 2.f * FL[b];
                    }
                    // RNLCOEF slots of the SL and FL weights of update types 1..8
                    constexpr int wsl[9] = {0, 20, 19, 17, 18, 8, 7, 5, 6};
                    constexpr int wfl[9] = {0, 23, 24, 21, 22, 11, 12, 9, 10};
#pragma unroll
                    for (int ty = 1; ty <= 8; ty++) {
                        if (tr[ty] < 0) continue;
                        const int kk = ty <= 4 ? ((ty & 1) ? k2 : k21) : ((ty & 1) ? k1 : k11);
                        const int b = kk + KL * tr[ty];
                        adb += rc[wsl[ty]] * SL[b];
                        const float d = rc[wfl[ty]] * FL[b];
                        if (ty <= 4) dmb += d;
                        else dpb += d;
                    }
                    float fad2b = adb * fcen, fcenb = adb * fad2;
                    const float fad1pb = ddb * ftemp;
                    fcenb += dpb * ((fij - 2.f * sam) * dal1) + dmb * ((fij - 2.f * sap) * dal2);
                    float fijb = dpb * (dal1 * fcen) + dmb * (dal2 * fcen);
                    float samb = -2.f * (dpb * (dal1 * fcen)), sapb = -2.f * (dmb * (dal2 * fcen));
                    fijb += fcenb * ftemp;
                    float fad1b = fad1pb;
                    fad2b += fad1pb;
                    fad1b += fad2b;
                    sapb -= (2.f * sam) * fad2b;
                    samb -= (2.f * sap) * fad2b;
                    fijb += fad1b * (sap + sam);
                    sapb += fad1b * fij;
                    samb += fad1b * fij;
                    QF[ql] = fijb; QP[ql] = sapb; QM[ql] = samb;
                }
                __syncthreads();
                // one bin per thread (every lane busy), its row's reader list
                for (int b = tid; b < NB; b += NT) {
                    const int m = b / KL, k = b - KL * m;
                    const int e0 = aro[c * ML + m], e1 = aro[c * ML + m + 1];
                    if (e1 > e0) {
                        float acc = 0.f;
                        for (int en = e0; en < e1; en++) {
                            const int v = are[en], mcm1 = v >> 4, sl = v & 15;
                            const float w = rnl[mcm1 * 25 + (sl <= 4 ? sl : sl + 8)];
                            const float *X = sl == 0 ? QF : (sl <= 4 ? QP : QM);
                            const int mpk = sl <= 4 ? ((sl - 1) & 1) : 2 + ((sl - 5) & 1);
                            const int base = (mcm1 - mc0) * 2 * KL;
#pragma unroll
                            for (int kh = 0; kh < 2; kh++) {
                                const int K = sl == 0 ? k : kinv[(mpk * 2 + kh) * KL + k];
                                acc = acc + w * X[base + kh * KL + K];
                            }
                        }
                        FB[b] += acc;
                    }
                }
                __syncthreads();
            }
        }
        for (int q = tid; q < NB4; q += NT) {
            const size_t o = q0 + (size_t)q * 256;
            *reinterpret_cast<float4 *>(e.fn + o) = reinterpret_cast<const float4 *>(FB)[q];
            *reinterpret_cast<float4 *>(x.slt + o) = reinterpret_cast<const float4 *>(SL)[q];
            *reinterpret_cast<float4 *>(x.flt + o) = reinterpret_cast<const float4 *>(FL)[q];
        }
        __syncthreads();
    }
}

// STRESSO's modulus reversed (one lane): as stresso_rev_scalar, on plain values
__device__ __forceinline__ void stresso_rev_lane(const DevTables &t, const float *k, float udir, float ust, float twb,
                                                 float jt0, float jt1, float &ustb, float &z0b, float &udirb,
                                                 float &xsb, float &ysb, float &t1b)
{
    const float xs = k[0], ys = k[1], tauhf = k[2], tau1 = k[3], t1 = k[4], tw = k[5], c1 = k[6];
    const float cap = ust * ust - t.EPS1;
    const float tmin = wam_minf(tw, cap);
    float tb = 0.f;
    if (tmin > 0.f) {
        if (tw < cap) tb = twb;
        else ustb += twb * (2.f * ust);
    }
    float xb = 0.f, yb = 0.f;
    if (tw > 0.f) { xb = tb * (xs / tw); yb = tb * (ys / tw); }
    const float su = wam_sinf(udir), co = wam_cosf(udir);
    const float hb = xb * su + yb * co;
    udirb += tauhf * (xb * co - yb * su);
    t1b = hb * (c1 * tau1);
    const float tau1b = hb * (c1 * t1);
    ustb += tau1b * jt0;
    z0b += tau1b * jt1;
    const float ra = wam_maxf(t.ROAIR, 1.f);
    xsb = xb / ra;
    ysb = yb / ra;
}

// sin_freq reversed for one frequency (one lane): adds the bars of USN1,
// USN2 and Z0 from the frequency's factor bars dm = {ZCN, UCN1, UCN2, XV1, XV2}
__device__ __forceinline__ void sin_freq_rev_lane(const DevTables &t, const SinFreq &r, float fr_m, float ust,
                                                  float sig, float z0, const float (&dm)[5], float &u1, float &u2,
                                                  float &zs)
{
    const float usn1 = ust * (1.0f - sig), usn2 = ust * (1.0f + sig);
    const float cm = (t.ZPI * fr_m) / t.G;
    const float d1 = ((usn1 / t.XKAPPA) * r.zcn) * cm, d2 = ((usn2 / t.XKAPPA) * r.zcn) * cm;
    const float g1 = dm[3] / (d1 * d1), g2 = dm[4] / (d2 * d2);
    const float zc = (r.zcn * cm) / t.XKAPPA;
    u1 += dm[1] * cm + g1 * zc;
    u2 += dm[2] * cm + g2 * zc;
    zs += ((dm[0] + g1 * ((usn1 / t.XKAPPA) * cm)) + g2 * ((usn2 / t.XKAPPA) * cm)) / z0;
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) adj_rev_kernel(DevTables t, AdjArgs a, ImplScratch x, AdjExtra e)
{
    extern __shared__ float rev_lds[];            // [3][KL][W]: COS(TH - UDIR), its bar, the tail's sums
    const int KL = t.KL, ML = t.ML, NB = KL * ML, NS = t.NSEA, NB4 = x.NB4, XS = x.NX, lane = threadIdx.x;
    const int W = blockDim.x;
    const int lp0 = blockIdx.x * W + lane;
    const bool on = lp0 < x.NP;
    const int lp = on ? lp0 : x.NP - 1, mp = x.P0 + lp;
    const int mem = mp / NS, pt = mp - mem * NS;
    float *cu = rev_lds + lane, *cub = rev_lds + (size_t)KL * W + lane, *acc = rev_lds + (size_t)2 * KL * W + lane;
    for (int k = 0; k < KL; k++) {
        cu[k * W] = e.cu[(size_t)k * XS + lp];
        cub[k * W] = e.cub[(size_t)k * XS + lp];
        acc[k * W] = 0.f;
    }
    auto S = [&](int slot) { return e.sc[(size_t)slot * XS + lp]; };
    const float udir = a.udir[mp], roairn = a.roairn[mp];
    const int mij = (int)S(AS_MIJ);
    const float ust1 = S(AS_UST1), z01 = S(AS_Z01), sig1 = S(AS_SIG1);
    const float ust2 = S(AS_UST2), z02 = S(AS_Z02), sig2 = S(AS_SIG2);
    const float emean = S(AS_EMEAN), f1mean = S(AS_F1MEAN), fmean = S(AS_FMEAN), fmeanws = S(AS_FMEANWS);
    const float const3 = ((2.f * t.XKAPPA) / t.BETAMAX) * (t.XKAPPA * t.XKAPPA);
    const float gm1 = 1.0f / t.G;
    float ustb = S(AS_UST2B) + (a.ustar_bar ? a.ustar_bar[mp] : 0.f);
    float z0b = a.z0_bar ? a.z0_bar[mp] : 0.f;
    float udirb = 0.f, u10b = 0.f;
    const float fmxb = S(AS_FMXB);
    const bool ws = fmeanws > fmean;
    const float fmeanwsb = ws ? fmxb : 0.f, fmeanb = ws ? 0.f : fmxb;
    // STRESSO, second call: the modulus's bar back to its sums
    float xsb, ysb, t1b;
    {
        float k2[7];
        for (int i = 0; i < 7; i++) k2[i] = S(AS_K2 + i);
        stresso_rev_lane(t, k2, udir, ust2, a.tauw_bar ? a.tauw_bar[mp] : 0.f, S(AS_JT2), S(AS_JT2 + 1), ustb, z0b,
                         udirb, xsb, ysb, t1b);
    }
    // pass I: SDISSIP and the second SINPUT/STRESSO reversed; rows at and
    // above MIJ add their final bars to the tail's sums
    const float e2 = emean * emean;
    const float sds = (t.CONSD * e2) * wam_ipowf(f1mean, 9);
    const float tden = t.FRM5[mij - 1];
    float sdsb = 0.f, f1b = 0.f, U1 = 0.f, U2 = 0.f, Z = 0.f;
    const size_t q0 = quad_offset(lp, 0, NB4);
    {
        const float usn1 = ust2 * (1.0f - sig2), usn2 = ust2 * (1.0f + sig2);
        for (int m = 0; m < ML; m++) {
            const SinFreq r = sin_freq<0>(t, t.FR[m], 0.f, usn1, usn2, z02, roairn);
            const float xd = wam_ipowf(t.FR[m] / f1mean, 2);
            const float t2d = sds * (0.5f * xd + 0.5f * wam_ipowf(xd, 2));
            float cmrw = 0.f;
            if (m < mij) {
                float rw = t.RHOWG_DFIM[m];
                if (m == mij - 1 && mij != ML) rw = 0.5f * rw;
                cmrw = ((t.ZPI * t.FR[m]) * gm1) * rw;
            }
            const float t2m = m >= mij ? t.FRM5[m] / tden : 0.f;
            float dsum = 0.f;
            float dm[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < KL; k0 += 4) {
                const size_t o = q0 + (size_t)((k0 + KL * m) >> 2) * 256;
                const float4 f4 = *reinterpret_cast<const float4 *>(x.ft + o);
                const float4 s4 = *reinterpret_cast<const float4 *>(x.slt + o);
                const float4 l4 = *reinterpret_cast<const float4 *>(x.flt + o);
                const float4 b4 = *reinterpret_cast<const float4 *>(e.fn + o);
                const float fv[4] = {f4.x, f4.y, f4.z, f4.w}, sv[4] = {s4.x, s4.y, s4.z, s4.w};
                const float lv[4] = {l4.x, l4.y, l4.z, l4.w};
                float bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = k0 + u;
                    const float ft = fv[u], slb = sv[u], flb = lv[u], ck = cu[k * W];
                    float fb = bv[u] + slb * t2d;
                    dsum = dsum + (slb * ft + flb);
                    const SinLin L = sin_bin_lin(t, r, ck, const3);
                    float sposb = 0.f;
                    if (m < mij && L.sp * ft > 0.f) sposb = (xsb * cmrw) * t.SINTH[k] + (ysb * cmrw) * t.COSTH[k];
                    float fbadd = slb * L.fl + sposb * L.sp;
                    float dtk = 0.f;
                    sin_bin_rev(t, r, L, ck, const3, sposb * ft, slb * ft + flb, dtk, dm);
                    if (m == mij - 1 && ck > 0.f) {
                        fbadd += t1b * ((ck * ck) * ck);
                        dtk += (t1b * ft) * (3.f * (ck * ck));
                    }
                    fb = fb + fbadd;
                    cub[k * W] += dtk;
                    if (m >= mij) acc[k * W] += t2m * fb;
                    bv[u] = fb;
                }
                if (on && m < mij) *reinterpret_cast<float4 *>(e.fn + o) = make_float4(bv[0], bv[1], bv[2], bv[3]);
            }
            sdsb += dsum * (0.5f * xd + 0.5f * (xd * xd));
            const float xb = dsum * (sds * (0.5f + xd));
            f1b += xb * ((-2.f * xd) / f1mean);
            sin_freq_rev_lane(t, r, t.FR[m], ust2, sig2, z02, dm, U1, U2, Z);
        }
    }
    const float emeanb = sdsb * ((t.CONSD * wam_ipowf(f1mean, 9)) * (2.f * emean));
    const float f1meanb = f1b + sdsb * ((t.CONSD * e2) * (9.f * wam_ipowf(f1mean, 8)));
    {
        const float sigb = ust2 * (U2 - U1);
        ustb += (U1 * (1.0f - sig2) + U2 * (1.0f + sig2)) + sigb * S(AS_JS2);
        z0b += Z + sigb * S(AS_JS2 + 1);
    }
    // AIRSEA, second call: (UST, Z0) from (U10, TAUW of the first STRESSO)
    u10b += ustb * S(AS_JA2) + z0b * S(AS_JA2 + 2);
    const float tw1b = ustb * S(AS_JA2 + 1) + z0b * S(AS_JA2 + 3);
    ustb = 0.f;
    z0b = 0.f;
    // STRESSO, first call
    {
        float k1[7];
        for (int i = 0; i < 7; i++) k1[i] = S(AS_K1 + i);
        stresso_rev_lane(t, k1, udir, ust1, tw1b, S(AS_JT1), S(AS_JT1 + 1), ustb, z0b, udirb, xsb, ysb, t1b);
    }
    // the mean parameters' integrals (EMEAN, FMEAN, F1MEAN; FMEANWS of the first LLWS)
    float seb, sfmb, stmb = 0.f, sewsb, sfmwsb;
    {
        const float s_fm = S(AS_SFM), s_tm = S(AS_STM);
        const float dfm = wam_maxf(s_fm, t.EMIN);
        float eb = emeanb + fmeanb / dfm;
        if (emean > t.EMIN) {
            const float q = emean / s_tm;
            const float qb = -f1meanb / (q * q);
            eb += qb / s_tm;
            stmb = -(qb * q) / s_tm;
        }
        sfmb = s_fm > t.EMIN ? -(fmeanb * fmean) / dfm : 0.f;
        seb = S(AS_SE) > t.EMIN ? eb : 0.f;
        const float s_fmws = S(AS_SFMWS);
        const float dws = wam_maxf(s_fmws, t.EMIN);
        sewsb = S(AS_SEWS) > t.EMIN ? fmeanwsb / dws : 0.f;
        sfmwsb = s_fmws > t.EMIN ? -(fmeanwsb * fmeanws) / dws : 0.f;
    }
    // pass II: the tail transposed, the first SINPUT/STRESSO and the mean
    // parameters reversed, the input spectrum's bar out
    U1 = 0.f; U2 = 0.f; Z = 0.f;
    {
        const float usn1 = ust1 * (1.0f - sig1), usn2 = ust1 * (1.0f + sig1);
        const float *F0 = a.fl3 + (size_t)mem * NB * NS + pt;
        float *OUT = a.fl3_bar_out + (size_t)mem * NB * NS + pt;
        const unsigned char *L1 = e.llws1 + lp;
        for (int m = 0; m < ML; m++) {
            const SinFreq r = sin_freq<0>(t, t.FR[m], 0.f, usn1, usn2, z01, roairn);
            float cmrw = 0.f;
            if (m < mij) {
                float rw = t.RHOWG_DFIM[m];
                if (m == mij - 1 && mij != ML) rw = 0.5f * rw;
                cmrw = ((t.ZPI * t.FR[m]) * gm1) * rw;
            }
            const bool last = m == ML - 1;
            const float tb = (seb * (t.DFIM[m] + (last ? t.MO_TAIL : 0.f)) +
                              sfmb * (t.DFIMOFR[m] + (last ? t.MM1_TAIL : 0.f))) +
                             stmb * (t.DFIM_FR[m] + (last ? t.MP1_TAIL : 0.f));
            const float twsb = sewsb * (t.DFIM[m] + (last ? t.MO_TAIL : 0.f)) +
                               sfmwsb * (t.DFIMOFR[m] + (last ? t.MM1_TAIL : 0.f));
            float dm[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < KL; k0 += 4) {
                float bv[4] = {0.f, 0.f, 0.f, 0.f};
                if (m < mij) {
                    const float4 b4 = *reinterpret_cast<const float4 *>(e.fn + q0 + (size_t)((k0 + KL * m) >> 2) * 256);
                    bv[0] = b4.x; bv[1] = b4.y; bv[2] = b4.z; bv[3] = b4.w;
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int k = k0 + u, b = k + KL * m;
                    const float f0 = F0[(size_t)b * NS], ck = cu[k * W];
                    float base = bv[u];
                    if (m == mij - 1) base = base + acc[k * W];
                    const SinLin L = sin_bin_lin(t, r, ck, const3);
                    float sposb = 0.f;
                    if (m < mij && L.sp * f0 > 0.f) sposb = (xsb * cmrw) * t.SINTH[k] + (ysb * cmrw) * t.COSTH[k];
                    float fbadd = sposb * L.sp + tb;
                    if (L1[(size_t)b * XS]) fbadd += twsb;
                    float dtk = 0.f;
                    sin_bin_rev(t, r, L, ck, const3, sposb * f0, 0.f, dtk, dm);
                    if (m == mij - 1 && ck > 0.f) {
                        fbadd += t1b * ((ck * ck) * ck);
                        dtk += (t1b * f0) * (3.f * (ck * ck));
                    }
                    cub[k * W] += dtk;
                    if (on) OUT[(size_t)b * NS] = base + fbadd;
                }
            }
            sin_freq_rev_lane(t, r, t.FR[m], ust1, sig1, z01, dm, U1, U2, Z);
        }
    }
    {
        const float sigb = ust1 * (U2 - U1);
        ustb += (U1 * (1.0f - sig1) + U2 * (1.0f + sig1)) + sigb * S(AS_JS1);
        z0b += Z + sigb * S(AS_JS1 + 1);
    }
    // AIRSEA, first call: (UST, Z0) from (U10, TAUW at the step's start)
    u10b += ustb * S(AS_JA1) + z0b * S(AS_JA1 + 2);
    const float tw0b = ustb * S(AS_JA1 + 1) + z0b * S(AS_JA1 + 3);
    float ub = udirb;
    for (int k = 0; k < KL; k++) ub += cub[k * W] * wam_sinf(t.TH[k] - udir);
    if (on) {
        a.u10_bar[mp] = u10b;
        a.udir_bar[mp] = ub;
        a.tauw_bar_out[mp] = tw0b;
    }
}

// ---------------------------------------------------------------------------
// PROPAGS transposed.  Input j of row (M, K) receives, from every output i
// whose stencil reads it, that output's weight times its bar.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int adj_chunk(int nchunks)
{
    const int per = (nchunks + 7) / 8;
    return (blockIdx.x % 8) * per + blockIdx.x / 8;
}

template <bool CURR>
__global__ void __launch_bounds__(256) propags_adj_kernel(DevTables t, const float *__restrict__ coef,
                                                           const float *__restrict__ gin, float *__restrict__ gout)
{
    const size_t moff = (size_t)blockIdx.y * t.KL * t.ML * t.NSEA;
    gin += moff;
    gout += moff;
    const int wg = blockDim.x, nch = (t.NSEA + wg - 1) / wg;
    const int ch = adj_chunk(nch);
    if (ch >= nch) return;
    const int j = ch * wg + threadIdx.x;
    if (j >= t.NSEA) return;
    const int KL = t.KL, ML = t.ML, NS = t.NSEA;
    const size_t NB = (size_t)KL * ML * NS;
    const int ns = t.KLAT_S[j], nn = t.KLAT_N[j], nw = t.KLON_W[j], ne = t.KLON_E[j];
    // every weight is read once: non-temporal loads keep them from evicting
    // the bars that neighbouring points read again
    auto term = [&](int plane, size_t row, int i) -> float {
        return i >= 0 ? __builtin_nontemporal_load(coef + plane * NB + row + i) * gin[row + i] : 0.f;
    };
    for (int m = 0; m < ML; m++) {
        for (int k = 0; k < KL; k++) {
            const size_t row = ((size_t)m * KL + k) * NS;
            const size_t rkm = ((size_t)m * KL + (t.KPM_M[k] - 1)) * NS;   // output K-1 reads K through TH1
            const size_t rkp = ((size_t)m * KL + (t.KPM_P[k] - 1)) * NS;   // output K+1 reads K through TH2
            float g = __builtin_nontemporal_load(coef + P_SUM * NB + row + j) * gin[row + j];
            if (!CURR) {
                // PROP_NO_CURR reads KLAT_S when JYO1 = 1 (else KLAT_N) and
                // KLON_W when JXO1 = 1 (else KLON_E)
                g = g + term(1, row, t.JYO1[k] == 1 ? nn : ns);
                g = g + term(2, row, t.JXO1[k] == 1 ? ne : nw);
                g = g + term(3, rkm, j);
                g = g + term(4, rkp, j);
            } else {
                g = g + term(P_LAT2, row, ns);     // output KLAT_S(j) reads its north neighbour j
                g = g + term(P_LAT1, row, nn);
                g = g + term(P_LON2, row, nw);
                g = g + term(P_LON1, row, ne);
                g = g + term(P_TH1, rkm, j);
                g = g + term(P_TH2, rkp, j);
                // SIG1 of output M' reads MIN(M'+1, ML); SIG2 reads MAX(M'-1, 1)
                if (m > 0) g = g + term(P_SIG1, ((size_t)(m - 1) * KL + k) * NS, j);
                if (m == ML - 1) g = g + term(P_SIG1, row, j);
                if (m < ML - 1) g = g + term(P_SIG2, ((size_t)(m + 1) * KL + k) * NS, j);
                if (m == 0) g = g + term(P_SIG2, row, j);
            }
            __builtin_nontemporal_store(g, gout + row + j);
        }
    }
}

// The no-current weights as the transposed stencil reads them: plane p at
// point j holds the weight that multiplies the bar gathered from j's
// neighbour in the transposed sum (SUM of j, LAT and LON of the points that
// read j, TH1 of row K-1 and TH2 of row K+1 at j; 0 for a land neighbour).
__global__ void transpose_weights_kernel(DevTables t, const float *__restrict__ coef, float *__restrict__ tw)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= t.NSEA) return;
    const int KL = t.KL, NS = t.NSEA;
    const int bin = blockIdx.y, m = bin / KL, k = bin - KL * m;
    const size_t NB = (size_t)KL * t.ML * NS, row = ((size_t)m * KL + k) * NS;
    const int il = t.JYO1[k] == 1 ? t.KLAT_N[j] : t.KLAT_S[j];
    const int io = t.JXO1[k] == 1 ? t.KLON_E[j] : t.KLON_W[j];
    tw[row + j] = coef[row + j];
    tw[NB + row + j] = il >= 0 ? coef[NB + row + il] : 0.f;
    tw[2 * NB + row + j] = io >= 0 ? coef[2 * NB + row + io] : 0.f;
    tw[3 * NB + row + j] = coef[3 * NB + ((size_t)m * KL + (t.KPM_M[k] - 1)) * NS + j];
    tw[4 * NB + row + j] = coef[4 * NB + ((size_t)m * KL + (t.KPM_P[k] - 1)) * NS + j];
}

// PROP_NO_CURR transposed on the transposed weights: the same products in
// the same order as propags_adj_kernel<false>, every weight load at the
// point's own place.
__global__ void __launch_bounds__(256) propags_adjt_kernel(DevTables t, const float *__restrict__ tw,
                                                            const float *__restrict__ gin, float *__restrict__ gout)
{
    const size_t moff = (size_t)blockIdx.y * t.KL * t.ML * t.NSEA;
    gin += moff;
    gout += moff;
    const int wg = blockDim.x, nch = (t.NSEA + wg - 1) / wg;
    const int ch = adj_chunk(nch);
    if (ch >= nch) return;
    const int j = ch * wg + threadIdx.x;
    if (j >= t.NSEA) return;
    const int KL = t.KL, ML = t.ML, NS = t.NSEA;
    const size_t NB = (size_t)KL * ML * NS;
    const int ns = t.KLAT_S[j], nn = t.KLAT_N[j], nw = t.KLON_W[j], ne = t.KLON_E[j];
    for (int m = 0; m < ML; m++) {
        for (int k = 0; k < KL; k++) {
            const size_t row = ((size_t)m * KL + k) * NS, at = row + j;
            const size_t rkm = ((size_t)m * KL + (t.KPM_M[k] - 1)) * NS;
            const size_t rkp = ((size_t)m * KL + (t.KPM_P[k] - 1)) * NS;
            const int il = t.JYO1[k] == 1 ? nn : ns, io = t.JXO1[k] == 1 ? ne : nw;
            // every load issued up front (a land neighbour reads the point itself, unused)
            const float w0 = __builtin_nontemporal_load(tw + at), w1 = __builtin_nontemporal_load(tw + NB + at);
            const float w2 = __builtin_nontemporal_load(tw + 2 * NB + at);
            const float w3 = __builtin_nontemporal_load(tw + 3 * NB + at);
            const float w4 = __builtin_nontemporal_load(tw + 4 * NB + at);
            const float g0 = gin[at], gl = gin[row + (il >= 0 ? il : j)], go = gin[row + (io >= 0 ? io : j)];
            const float g3 = gin[rkm + j], g4 = gin[rkp + j];
            float g = w0 * g0;
            g = g + (il >= 0 ? w1 * gl : 0.f);
            g = g + (io >= 0 ? w2 * go : 0.f);
            g = g + w3 * g3;
            g = g + w4 * g4;
            __builtin_nontemporal_store(g, gout + at);
        }
    }
}

__global__ void wind_uv_kernel(float zpi, const float *__restrict__ u, const float *__restrict__ v,
                               float *__restrict__ u10, float *__restrict__ udir, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float uu = u[i], vv = v[i];
    const float s = sqrtf(uu * uu + vv * vv);
    float d = s != 0.f ? wam_atan2f(uu, vv) : 0.f;
    if (d < 0.f) d = d + zpi;
    u10[i] = s;
    udir[i] = d;
}

__global__ void wind_uv_adj_kernel(const float *__restrict__ u, const float *__restrict__ v,
                                   const float *__restrict__ u10b, const float *__restrict__ udirb,
                                   float *__restrict__ ub, float *__restrict__ vb, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float uu = u[i], vv = v[i];
    const float r2 = uu * uu + vv * vv, s = sqrtf(r2);
    float gu = 0.f, gv = 0.f;
    if (s != 0.f) {
        // dU10/du = u/U10, dUDIR/du = v/(u**2 + v**2), dUDIR/dv = -u/(u**2 + v**2)
        const float a = u10b[i], d = udirb[i];
        gu = a * (uu / s) + d * (vv / r2);
        gv = a * (vv / s) - d * (uu / r2);
    }
    ub[i] = gu;
    vb[i] = gv;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
size_t implsch_adj_lds_bytes(const DevTables &t)
{
    return adj_lds_floats(t.KL, t.ML, t.MLSTHG, t.CHUNK) * sizeof(float) + 2 * (size_t)t.KL * t.ML;
}

hipError_t launch_implsch_adjoint(const DevTables &t, const AdjArgs &a, int ncu, hipStream_t st)
{
    const size_t lds = implsch_adj_lds_bytes(t);
    hipError_t e = hipFuncSetAttribute((const void *)implsch_adj_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, implsch_adj_kernel, NT, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    const long nblocks = std::min<long>((long)per_cu * ncu, a.NMP);
    hipLaunchKernelGGL(implsch_adj_kernel, dim3((unsigned)nblocks), dim3(NT), lds, st, t, a);
    return hipGetLastError();
}

hipError_t launch_implsch_adj_forward(const DevTables &t, const ImplArgs &a, const ImplScratch &x,
                                      const AdjExtra &ex, int ncu, hipStream_t st, const LaunchTune &tu);

// IMPLSCH's adjoint in the split form on the chunk of member-points of x.
hipError_t launch_implsch_adjoint_split(const DevTables &t, const AdjArgs &a, const ImplScratch &x, const AdjExtra &ex,
                                        int ncu, hipStream_t st, const LaunchTune &tu)
{
    ImplArgs ia{};
    ia.fl3 = const_cast<float *>(a.fl3);
    ia.u10 = a.u10; ia.udir = a.udir; ia.roairn = a.roairn; ia.wstar = a.wstar;
    // read only by the recomputation (TAUW at the step's start; Z0 in is not used by AIRSEA)
    ia.tauw = const_cast<float *>(a.tauw);
    ia.z0 = const_cast<float *>(a.tauw);
    ia.NMP = a.NMP;
    hipError_t e = launch_implsch_adj_forward(t, ia, x, ex, ncu, st, tu);
    if (e != hipSuccess) return e;
    constexpr int W = 64;
    const int nwg = (x.NP + W - 1) / W;
    hipLaunchKernelGGL(adj_post_kernel, dim3(nwg), dim3(W), 0, st, t, a, x, ex);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t lds = adj_snl_lds_floats(t.KL, t.ML, t.MLSTHG, t.NAR, t.ANCH) * sizeof(float);
    e = hipFuncSetAttribute((const void *)adj_snl_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, adj_snl_kernel, NT, lds);
    if (e != hipSuccess) return e;
    if (per_cu < 1) return hipErrorInvalidConfiguration;
    const long nblocks = std::min<long>((long)per_cu * ncu, x.NP);
    hipLaunchKernelGGL(adj_snl_kernel, dim3((unsigned)nblocks), dim3(NT), lds, st, t, a, x, ex);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(adj_rev_kernel, dim3(nwg), dim3(W), (size_t)3 * t.KL * W * sizeof(float), st, t, a, x, ex);
    return hipGetLastError();
}

// tw: the transposed no-current weights (transpose_weights_kernel), else NULL
hipError_t launch_propags_adjoint(const DevTables &t, const float *coef, const float *tw, int nmem, const float *gin,
                                  float *gout, hipStream_t st)
{
    const int wg = 256, nch = (t.NSEA + wg - 1) / wg;
    const dim3 grid(((nch + 7) / 8) * 8, nmem), block(wg);
    if (t.refr_c) hipLaunchKernelGGL(propags_adj_kernel<true>, grid, block, 0, st, t, coef, gin, gout);
    else if (tw) hipLaunchKernelGGL(propags_adjt_kernel, grid, block, 0, st, t, tw, gin, gout);
    else hipLaunchKernelGGL(propags_adj_kernel<false>, grid, block, 0, st, t, coef, gin, gout);
    return hipGetLastError();
}

hipError_t launch_transpose_weights(const DevTables &t, const float *coef, float *tw, hipStream_t st)
{
    hipLaunchKernelGGL(transpose_weights_kernel, dim3((t.NSEA + 255) / 256, t.KL * t.ML), dim3(256), 0, st, t, coef,
                       tw);
    return hipGetLastError();
}

hipError_t launch_wind_uv(float zpi, const float *u, const float *v, float *u10, float *udir, int64_t n,
                          hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(wind_uv_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, zpi, u, v, u10, udir, n);
    return hipGetLastError();
}

hipError_t launch_wind_uv_adjoint(const float *u, const float *v, const float *u10b, const float *udirb, float *ub,
                                  float *vb, int64_t n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(wind_uv_adj_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, u, v, u10b, udirb,
                       ub, vb, n);
    return hipGetLastError();
}

}  // namespace wam
