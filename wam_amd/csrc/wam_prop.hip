// wam_prop.hip -- PROPAGS and its weights on gfx950.
//
//  * prep_coef_kernel: PREPARE_PROPAGATION_COEF (wam_propagation_module.f90:
//    PREP_SPHER_DEEP :1248-1327, PREP_SPHER_SHALLOW :1336-1485,
//    PREP_SPHER_DEEP_CURR_REG :1591-1676, PREP_SPHER_SHALLOW_CURR_REG
//    :1808-1911) generated in HBM, one thread per (sea point, direction,
//    frequency); THDD/SIDC (:400-560) are recomputed inline from per-point
//    gradients instead of being stored.
//  * minmax_kernel / rescale_kernel: CHECK_CFL (:1917-2138).
//  * propags_*_kernel: PROP_NO_CURR (:750-785) and PROP_CURR (:794-825).
//    A thread owns one sea point and walks all (M, K) planes, so the
//    neighbour indices are read once and every plane access is coalesced;
//    blocks are laid out so each XCD owns a contiguous sea-point range and
//    the north/south neighbour rows are served from that XCD's L2.
//  * wavnu2_kernel: WAVNU2 (wam_source_module.f90:4343-4461), which depends
//    on depth only and is tabulated once per depth field.
#include "wam_dev.hpp"
#include "wam_math.h"

namespace wam {

__device__ __forceinline__ float cg_at(const DevTables &t, int ij, int m)
{
    const int nd = t.NDEPTH;
    return ij < 0 ? t.TCGOND[m * nd + nd - 1] : t.TCGOND[m * nd + t.INDEP[ij] - 1];
}

// Depth part of theta-dot times TSIHKD plus the current part (THDD).
__device__ __forceinline__ float thdd(const DevTables &t, int ij, int k, int m)
{
    float th = 0.f;
    if (t.shallow) {
        float temp = t.SINTH[k] * t.DDPHI[ij] - t.COSTH[k] * t.DDLAM[ij];
        th = temp * t.TSIHKD[m * t.NDEPTH + t.INDEP_G[ij] - 1];
    }
    if (t.refr_c) {
        float s = t.SINTH[k], c = t.COSTH[k];
        float ss = s * s, sc = s * c, cc = c * c;
        float thdc = ((ss * t.DUPHI[ij] + sc * t.DVPHI[ij]) - sc * t.DULAM[ij]) - cc * t.DVLAM[ij];
        th = t.shallow ? th + thdc : thdc;
    }
    return th;
}

// Sigma-dot (SIDC); deep water uses the M-independent current part only.
__device__ __forceinline__ float sidc(const DevTables &t, int ij, int k, int m)
{
    float s = t.SINTH[k], c = t.COSTH[k];
    float ss = s * s, sc = s * c, cc = c * c;
    float s0 = (((-(sc * t.DUPHI[ij])) - cc * t.DVPHI[ij]) - ss * t.DULAM[ij]) - sc * t.DVLAM[ij];
    if (!t.shallow) return s0;
    float tempd = t.V[ij] * t.DDPHI[ij] + t.U[ij] * t.DDLAM[ij];
    const int ig = m * t.NDEPTH + t.INDEP_G[ij] - 1;
    return (s0 * cg_at(t, ij, m) + tempd * t.TSIHKD[ig]) * t.TFAK[ig];
}

__global__ void prep_coef_kernel(DevTables t, float *__restrict__ coef)
{
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= t.NSEA) return;
    const int KL = t.KL, ML = t.ML;
    const int bin = blockIdx.y, m = bin / KL, k = bin % KL;
    const size_t NB = (size_t)KL * ML * t.NSEA;
    const size_t at = (size_t)bin * t.NSEA + ij;
    const int kp = t.KPM_P[k] - 1, km = t.KPM_M[k] - 1;
    const float sd = t.SINTH[k], cd = t.COSTH[k];
    if (!t.refr_c && !t.shallow) {                         // PREP_SPHER_DEEP
        const float sda = wam_fabsf(sd), cda = wam_fabsf(cd * t.DELPH0);
        const int ic = t.JYO1[k], id = t.JYO2[k];
        const float sp = t.DELTH0 * (t.SINTH[k] + t.SINTH[kp]);
        const float sm = t.DELTH0 * (t.SINTH[k] + t.SINTH[km]);
        const float gom = t.GOM[m];
        const float dpn = cda * ((ic == 1 ? t.DPSN_S[ij] : t.DPSN_N[ij]) + 1.f);
        const float dle = sda * t.DELLA0[ij];
        float dtc = cda * ((id == 1 ? t.DPSN_S[ij] : t.DPSN_N[ij]) + 1.f) + dle;
        const float dtp = t.DRGM[ij] * sp, dtm = t.DRGM[ij] * sm;
        dtc = (dtc + wam_maxf(0.f, dtp)) - wam_minf(0.f, dtm);
        coef[0 * NB + at] = 1.f - gom * dtc;
        coef[1 * NB + at] = gom * dpn;
        coef[2 * NB + at] = gom * dle;
        coef[3 * NB + at] = -(gom * wam_minf(0.f, dtp));
        coef[4 * NB + at] = gom * wam_maxf(0.f, dtm);
    } else if (!t.refr_c) {                                // PREP_SPHER_SHALLOW
        const float sda = wam_fabsf(sd), cda = wam_fabsf(cd * t.DELPH0);
        const float sp = t.DELTH0 * (t.SINTH[k] + t.SINTH[kp]);
        const float sm = t.DELTH0 * (t.SINTH[k] + t.SINTH[km]);
        const int is = t.JXO1[k], it = t.JXO2[k], ic = t.JYO1[k], id = t.JYO2[k];
        const float cg = cg_at(t, ij, m);
        const float cgk_s = cg + t.DPSN_S[ij] * cg_at(t, t.KLAT_S[ij], m);
        const float cgk_n = cg + t.DPSN_N[ij] * cg_at(t, t.KLAT_N[ij], m);
        float dtc = 1.f - cda * (id == 1 ? cgk_s : cgk_n);
        const float lat = cda * (ic == 1 ? cgk_s : cgk_n);
        const int nt = it == 1 ? t.KLON_W[ij] : t.KLON_E[ij];
        const int ns = is == 1 ? t.KLON_W[ij] : t.KLON_E[ij];
        dtc = dtc - (sda * (cg_at(t, nt, m) + cg)) * t.DELLA0[ij];
        const float lon = (sda * (cg_at(t, ns, m) + cg)) * t.DELLA0[ij];
        float dtp = t.DRGM[ij] * cg, dtm;
        if (t.refr_d) {
            const float thk = thdd(t, ij, k, m);
            dtm = dtp * sm + (thk + thdd(t, ij, km, m)) * t.DELTHR;
            dtp = dtp * sp + (thk + thdd(t, ij, kp, m)) * t.DELTHR;
        } else { dtm = dtp * sm; dtp = dtp * sp; }
        dtc = (dtc - wam_maxf(0.f, dtp)) + wam_minf(0.f, dtm);
        coef[0 * NB + at] = dtc;
        coef[1 * NB + at] = lat;
        coef[2 * NB + at] = lon;
        coef[3 * NB + at] = -wam_minf(0.f, dtp);
        coef[4 * NB + at] = wam_maxf(0.f, dtm);
    } else {                                               // current refraction, regular grid
        const int mm = (m > 0 ? m - 1 : 0), mp = (m < ML - 1 ? m + 1 : ML - 1);
        const float dfp = t.DELFR0 / t.FR[m], dfm = t.DELFR0 / t.FR[mm];
        float sp = t.DELTH0 * (t.SINTH[k] + t.SINTH[kp]);
        float sm = t.DELTH0 * (t.SINTH[k] + t.SINTH[km]);
        if (!t.shallow) { sp = sp * t.GOM[m]; sm = sm * t.GOM[m]; }
        const float cgs = t.GOM[m] * sd, cgc = t.GOM[m] * cd;
        auto wok1 = [&](int j) -> float {
            const float u = j < 0 ? 0.f : t.U[j];
            return t.shallow ? u + sd * cg_at(t, j, m) : u + cgs;
        };
        auto wok2 = [&](int j) -> float {
            const float v = j < 0 ? 0.f : t.V[j];
            return t.shallow ? (v + cd * cg_at(t, j, m)) * t.DELPH0 : (v + cgc) * t.DELPH0;
        };
        const float w1 = wok1(ij), w2 = wok2(ij);
        const float dlw = (w1 + wok1(t.KLON_W[ij])) * t.DELLA0[ij];
        const float dle = (w1 + wok1(t.KLON_E[ij])) * t.DELLA0[ij];
        float dtc = (1.f - wam_maxf(0.f, dle)) + wam_minf(0.f, dlw);
        const float dps = w2 + t.DPSN_S[ij] * wok2(t.KLAT_S[ij]);
        const float dpn = w2 + t.DPSN_N[ij] * wok2(t.KLAT_N[ij]);
        dtc = (dtc - wam_maxf(0.f, dpn)) + wam_minf(0.f, dps);
        float dtp, dtm;
        if (t.shallow) {
            const float cg = cg_at(t, ij, m), thk = thdd(t, ij, k, m);
            dtp = ((sp * t.DRGM[ij]) * cg) + (thk + thdd(t, ij, kp, m)) * t.DELTHR;
            dtm = ((sm * t.DRGM[ij]) * cg) + (thk + thdd(t, ij, km, m)) * t.DELTHR;
        } else {
            const float thk = thdd(t, ij, k, ML - 1);
            dtp = sp * t.DRGM[ij] + (thk + thdd(t, ij, kp, ML - 1)) * t.DELTHR;
            dtm = sm * t.DRGM[ij] + (thk + thdd(t, ij, km, ML - 1)) * t.DELTHR;
        }
        dtc = (dtc - wam_maxf(0.f, dtp)) + wam_minf(0.f, dtm);
        float s1, s2;
        if (t.shallow) {
            const float sk = sidc(t, ij, k, m);
            const float dop = (sk + sidc(t, ij, k, mp)) * dfp;
            const float dom = (sk + sidc(t, ij, k, mm)) * dfm;
            dtc = (dtc - wam_maxf(0.f, dop)) + wam_minf(0.f, dom);
            s1 = -wam_minf(0.f, dop) / 1.1f;
            s2 = wam_maxf(0.f, dom) * 1.1f;
        } else {
            const float dom = sidc(t, ij, k, ML - 1) * t.DELFR0;
            dtc = dtc - wam_fabsf(dom);
            s1 = -wam_minf(0.f, dom) / 1.1f;
            s2 = wam_maxf(0.f, dom) * 1.1f;
        }
        coef[P_SUM * NB + at] = dtc;
        coef[P_LAT1 * NB + at] = wam_maxf(0.f, dps);
        coef[P_LAT2 * NB + at] = -wam_minf(0.f, dpn);
        coef[P_LON1 * NB + at] = wam_maxf(0.f, dlw);
        coef[P_LON2 * NB + at] = -wam_minf(0.f, dle);
        coef[P_TH1 * NB + at] = -wam_minf(0.f, dtp);
        coef[P_TH2 * NB + at] = wam_maxf(0.f, dtm);
        coef[P_SIG1 * NB + at] = s1;
        coef[P_SIG2 * NB + at] = s2;
    }
}

// Per-block min and max of one coefficient plane (order-free, exact).
__global__ void minmax_kernel(const float *__restrict__ x, size_t n, float *__restrict__ out)
{
    __shared__ float smin[256], smax[256];
    float lo = 3.4e38f, hi = -3.4e38f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = x[i];
        lo = wam_minf(lo, v); hi = wam_maxf(hi, v);
    }
    smin[threadIdx.x] = lo; smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = wam_minf(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = wam_maxf(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[2 * blockIdx.x] = smin[0]; out[2 * blockIdx.x + 1] = smax[0]; }
}

// COUNTER > 1: weights divided by COUNTER, centre weight 1 - (1 - SUM)/COUNTER.
__global__ void rescale_kernel(float *__restrict__ coef, size_t nb, int nplanes, float rc)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb * nplanes;
         i += (size_t)gridDim.x * blockDim.x) {
        const float v = coef[i];
        coef[i] = (i < nb) ? 1.f - (1.f - v) / rc : v / rc;
    }
}

// Chunk of 256 sea points handled by this block: each XCD (blocks are
// dispatched round-robin over the 8 XCDs) owns one contiguous range.
__device__ __forceinline__ int xcd_chunk(int nchunks)
{
    const int per = (nchunks + 7) / 8;
    return (blockIdx.x % 8) * per + blockIdx.x / 8;
}

// Weights are read exactly once per sub-step: non-temporal loads keep them
// from evicting the spectrum lines that neighbouring points re-read from L2.
__device__ __forceinline__ float ld_stream(const float *__restrict__ p, size_t i)
{
    return __builtin_nontemporal_load(p + i);
}

__device__ __forceinline__ float ld_or0(const float *__restrict__ p, int j)
{
    return j >= 0 ? p[j] : 0.f;
}

// PROP_NO_CURR: F3 = SUM*F + LAT*F(KLAT(IC)) + LON*F(KLON(IS)) + TH1*F(K+1) + TH2*F(K-1)
__global__ void __launch_bounds__(256) propags_nocurr_kernel(DevTables t, const float *__restrict__ coef,
                                                              const float *__restrict__ fin,
                                                              float *__restrict__ fout)
{
    const int nch = (t.NSEA + 255) / 256;
    const int ch = xcd_chunk(nch);
    if (ch >= nch) return;
    const int ij = ch * 256 + threadIdx.x;
    if (ij >= t.NSEA) return;
    const int KL = t.KL, ML = t.ML, NS = t.NSEA;
    const size_t NB = (size_t)KL * ML * NS;
    const int ns = t.KLAT_S[ij], nn = t.KLAT_N[ij], nw = t.KLON_W[ij], ne = t.KLON_E[ij];
    for (int m = 0; m < ML; m++) {
        for (int k = 0; k < KL; k++) {
            const size_t row = ((size_t)m * KL + k) * NS;
            const float *f = fin + row;
            const int nlat = t.JYO1[k] == 1 ? ns : nn;
            const int nlon = t.JXO1[k] == 1 ? nw : ne;
            const float *fkp = fin + ((size_t)m * KL + (t.KPM_P[k] - 1)) * NS;
            const float *fkm = fin + ((size_t)m * KL + (t.KPM_M[k] - 1)) * NS;
            const size_t at = row + ij;
            float s = ld_stream(coef, P_SUM * NB + at) * f[ij] + ld_stream(coef, 1 * NB + at) * ld_or0(f, nlat);
            s = s + ld_stream(coef, 2 * NB + at) * ld_or0(f, nlon);
            s = s + ld_stream(coef, 3 * NB + at) * fkp[ij];
            s = s + ld_stream(coef, 4 * NB + at) * fkm[ij];
            __builtin_nontemporal_store(s, &fout[at]);
        }
    }
}

// PROP_CURR: nine-point stencil in the reference's summation order.
__global__ void __launch_bounds__(256) propags_curr_kernel(DevTables t, const float *__restrict__ coef,
                                                            const float *__restrict__ fin,
                                                            float *__restrict__ fout)
{
    const int nch = (t.NSEA + 255) / 256;
    const int ch = xcd_chunk(nch);
    if (ch >= nch) return;
    const int ij = ch * 256 + threadIdx.x;
    if (ij >= t.NSEA) return;
    const int KL = t.KL, ML = t.ML, NS = t.NSEA;
    const size_t NB = (size_t)KL * ML * NS;
    const int ns = t.KLAT_S[ij], nn = t.KLAT_N[ij], nw = t.KLON_W[ij], ne = t.KLON_E[ij];
    for (int m = 0; m < ML; m++) {
        const int mm = m > 0 ? m - 1 : 0, mp = m < ML - 1 ? m + 1 : ML - 1;
        for (int k = 0; k < KL; k++) {
            const size_t row = ((size_t)m * KL + k) * NS;
            const float *f = fin + row;
            const size_t at = row + ij;
            const float *fkp = fin + ((size_t)m * KL + (t.KPM_P[k] - 1)) * NS;
            const float *fkm = fin + ((size_t)m * KL + (t.KPM_M[k] - 1)) * NS;
            const float *fmp = fin + ((size_t)mp * KL + k) * NS;
            const float *fmm = fin + ((size_t)mm * KL + k) * NS;
            float s = ld_stream(coef, P_SUM * NB + at) * f[ij];
            s = s + ld_stream(coef, P_LAT2 * NB + at) * ld_or0(f, nn);
            s = s + ld_stream(coef, P_LAT1 * NB + at) * ld_or0(f, ns);
            s = s + ld_stream(coef, P_LON2 * NB + at) * ld_or0(f, ne);
            s = s + ld_stream(coef, P_LON1 * NB + at) * ld_or0(f, nw);
            s = s + ld_stream(coef, P_TH1 * NB + at) * fkp[ij];
            s = s + ld_stream(coef, P_TH2 * NB + at) * fkm[ij];
            s = s + ld_stream(coef, P_SIG1 * NB + at) * fmp[ij];
            s = s + ld_stream(coef, P_SIG2 * NB + at) * fmm[ij];
            __builtin_nontemporal_store(s, &fout[at]);
        }
    }
}

// PROP_CURR for KL = KLT directions: the point's spectrum rows M-1, M and
// M+1 (all directions) are held in registers, so the direction and
// frequency neighbours (cyclic K+1, K-1; M+1, M-1 clamped at the ends) cost
// no memory access; only the four spatial neighbours and the weights are
// loaded.  Same sums in the same order as propags_curr_kernel.
template <int KLT>
__global__ void __launch_bounds__(256) propags_curr_rows_kernel(DevTables t, const float *__restrict__ coef,
                                                                 const float *__restrict__ fin,
                                                                 float *__restrict__ fout)
{
    const int nch = (t.NSEA + 255) / 256;
    const int ch = xcd_chunk(nch);
    if (ch >= nch) return;
    const int ij = ch * 256 + threadIdx.x;
    if (ij >= t.NSEA) return;
    const int ML = t.ML, NS = t.NSEA;
    const size_t NB = (size_t)KLT * ML * NS;
    const int ns = t.KLAT_S[ij], nn = t.KLAT_N[ij], nw = t.KLON_W[ij], ne = t.KLON_E[ij];
    float prev[KLT], cur[KLT], nxt[KLT];
#pragma unroll
    for (int k = 0; k < KLT; k++) cur[k] = fin[(size_t)k * NS + ij];
#pragma unroll
    for (int k = 0; k < KLT; k++) prev[k] = cur[k];
    for (int m = 0; m < ML; m++) {
        if (m < ML - 1) {
#pragma unroll
            for (int k = 0; k < KLT; k++) nxt[k] = fin[((size_t)(m + 1) * KLT + k) * NS + ij];
        } else {
#pragma unroll
            for (int k = 0; k < KLT; k++) nxt[k] = cur[k];
        }
#pragma unroll
        for (int k = 0; k < KLT; k++) {
            const int kp = k + 1 < KLT ? k + 1 : 0, km = k > 0 ? k - 1 : KLT - 1;
            const size_t row = ((size_t)m * KLT + k) * NS;
            const float *f = fin + row;
            const size_t at = row + ij;
            float s = ld_stream(coef, P_SUM * NB + at) * cur[k];
            s = s + ld_stream(coef, P_LAT2 * NB + at) * ld_or0(f, nn);
            s = s + ld_stream(coef, P_LAT1 * NB + at) * ld_or0(f, ns);
            s = s + ld_stream(coef, P_LON2 * NB + at) * ld_or0(f, ne);
            s = s + ld_stream(coef, P_LON1 * NB + at) * ld_or0(f, nw);
            s = s + ld_stream(coef, P_TH1 * NB + at) * cur[kp];
            s = s + ld_stream(coef, P_TH2 * NB + at) * cur[km];
            s = s + ld_stream(coef, P_SIG1 * NB + at) * nxt[k];
            s = s + ld_stream(coef, P_SIG2 * NB + at) * prev[k];
            __builtin_nontemporal_store(s, &fout[at]);
        }
#pragma unroll
        for (int k = 0; k < KLT; k++) { prev[k] = cur[k]; cur[k] = nxt[k]; }
    }
}

template __global__ void propags_curr_rows_kernel<24>(DevTables, const float *, const float *, float *);
template __global__ void propags_curr_rows_kernel<36>(DevTables, const float *, const float *, float *);

// WAVNU2 for every (sea point, frequency): WN and CGG as [ML][NSEA].
__global__ void wavnu2_kernel(DevTables t, float *__restrict__ wn, float *__restrict__ cgg)
{
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= t.NSEA) return;
    const int m = blockIdx.y;
    const float G = t.G, H = t.DEPTH[ij], EPS = 1E-7f;
    const float sig = t.ZPI * t.FR[m];
    float kold = 0.f, k, w0 = wam_fabsf(sig), cg;
    if (w0 < sqrtf(G / H)) k = w0 / sqrtf(G * H); else k = (w0 * w0) / G;
    for (int i = 1; i <= 15; i++) {
        const float dif = wam_fabsf(k - kold);
        const float rdif = (k != 0.f) ? dif / k : 0.f;
        if (dif < EPS && rdif < EPS) break;
        kold = k;
        const float th = wam_tanhf(kold * H);
        const float f1 = ((G * kold) * th) - w0 * w0;
        float fd;
        if (kold * H > 25.f) fd = G * th;
        else { const float ch = wam_coshf(kold * H); fd = G * th + ((G * kold) * H) / (ch * ch); }
        k = kold - f1 / fd;
    }
    if ((2.f * k) * H > 25.f) cg = (w0 / k) * 0.5f;
    else cg = ((w0 / k) * 0.5f) * (1.f + ((2.f * k) * H) / wam_sinhf((2.f * k) * H));
    wn[(size_t)m * t.NSEA + ij] = k;
    cgg[(size_t)m * t.NSEA + ij] = cg;
}

}  // namespace wam
