// wam_prop.hip -- PROPAGS and its weights on gfx950.
//
//  * prep_coef_kernel: PREPARE_PROPAGATION_COEF (wam_propagation_module.f90:
//    PREP_SPHER_DEEP :1248-1327, PREP_SPHER_SHALLOW :1336-1485,
//    PREP_SPHER_DEEP_CURR_REG :1591-1676, PREP_SPHER_SHALLOW_CURR_REG
//    :1808-1911) generated in HBM, one thread per (sea point, direction,
//    frequency); THDD/SIDC (:400-560) are recomputed inline from per-point
//    gradients instead of being stored.
//  * minmax_kernel / rescale_kernel: CHECK_CFL (:1917-2138).
//  * propags_fused_sc_kernel / propags_fused_kernel: PROPAGS with the
//    weights computed inside the stencil (no weight planes in HBM), with
//    cfl_minmax_kernel as their CHECK_CFL.
//  * propags_*_kernel: PROP_NO_CURR (:750-785) and PROP_CURR (:794-825).
//    A thread owns one sea point and walks all (M, K) planes, so the
//    neighbour indices are read once and every plane access is coalesced;
//    blocks are laid out so each XCD owns a contiguous sea-point range and
//    the north/south neighbour rows are served from that XCD's L2.
//  * wavnu2_kernel: WAVNU2 (wam_source_module.f90:4343-4461), which depends
//    on depth only and is tabulated once per depth field.
#include <cfloat>

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

// The propagation weights of one (sea point, direction, frequency), in the
// plane order of wam_dev.hpp (without currents: SUM, LAT1, LON1, TH1, TH2).
// CURR selects the current-refraction formulas (t.refr_c) at compile time.
template <bool CURR>
__device__ __forceinline__ void coef_bin(const DevTables &t, int ij, int m, int k, float *__restrict__ c)
{
    const int ML = t.ML;
    const int kp = t.KPM_P[k] - 1, km = t.KPM_M[k] - 1;
    const float sd = t.SINTH[k], cd = t.COSTH[k];
    if (!CURR && !t.shallow) {                             // PREP_SPHER_DEEP
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
        c[0] = 1.f - gom * dtc;
        c[1] = gom * dpn;
        c[2] = gom * dle;
        c[3] = -(gom * wam_minf(0.f, dtp));
        c[4] = gom * wam_maxf(0.f, dtm);
    } else if (!CURR) {                                    // PREP_SPHER_SHALLOW
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
        c[0] = dtc;
        c[1] = lat;
        c[2] = lon;
        c[3] = -wam_minf(0.f, dtp);
        c[4] = wam_maxf(0.f, dtm);
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
        c[P_SUM] = dtc;
        c[P_LAT1] = wam_maxf(0.f, dps);
        c[P_LAT2] = -wam_minf(0.f, dpn);
        c[P_LON1] = wam_maxf(0.f, dlw);
        c[P_LON2] = -wam_minf(0.f, dle);
        c[P_TH1] = -wam_minf(0.f, dtp);
        c[P_TH2] = wam_maxf(0.f, dtm);
        c[P_SIG1] = s1;
        c[P_SIG2] = s2;
    }
}

// The per-point part of PREPARE_PROPAGATION (wam_propagation_module.f90:
// 279-598): INDEP (depth table index), GRADIENT's (:182-275) depth and
// current gradients (centred differences over the N/S and E/W neighbours,
// zero where one is land; the longitude derivative times DCO), INDEP_G (the
// index of the depth averaged with the four neighbours), DELLA0 and DRGM.
// The same operations, in the same order, as the host's prepare_points.
__global__ void __launch_bounds__(256) prep_points_kernel(DevTables t, PrepPoints p)
{
    const int ij = blockIdx.x * 256 + threadIdx.x;
    if (ij >= t.NSEA) return;
    const int ND = t.NDEPTH;
    const float d0 = p.depth[ij];
    if (!(d0 > 0.f && d0 <= FLT_MAX)) p.bad[0] = 1u;
    if (t.refr_c && !(fabsf(p.u[ij]) <= FLT_MAX && fabsf(p.v[ij]) <= FLT_MAX)) p.bad[1] = 1u;
    auto index = [&](float dep) {
        const int id = wam_nint(wam_logf(dep / p.deptha) / p.lndepthd + 1.f);
        return id < 1 ? 1 : (id > ND ? ND : id);
    };
    p.indep[ij] = t.shallow ? index(d0) : 1;
    const int n = t.KLAT_N[ij], s = t.KLAT_S[ij], e = t.KLON_E[ij], w = t.KLON_W[ij];   // 0-based, -1 = land
    const int kx = t.KXLT_PT[ij] - 1;
    const float dco = t.DCO_PT[ij], dlam = t.DELLAM_ROW[kx];
    auto grad = [&](const float *f, float &dlat, float &dlon) {
        dlat = (n >= 0 && s >= 0) ? (f[n] - f[s]) / (2.f * t.DELPHI) : 0.f;
        dlon = (e >= 0 && w >= 0) ? (f[e] - f[w]) / (2.f * dlam) : 0.f;
    };
    float ddphi = 0.f, ddlam = 0.f, duphi = 0.f, dulam = 0.f, dvphi = 0.f, dvlam = 0.f;
    int indep_g = 1;
    if ((t.refr_d || t.refr_c) && t.shallow) {
        grad(p.depth, ddphi, ddlam);
        ddlam = ddlam * dco;
        float dm = 4.f * d0;
        int iz = 4;
        if (n >= 0 && s >= 0) { dm = dm + p.depth[n] + p.depth[s]; iz += 2; }
        if (e >= 0 && w >= 0) { dm = dm + p.depth[e] + p.depth[w]; iz += 2; }
        dm = dm / (float)iz;
        indep_g = index(dm);
    }
    if (t.refr_c) {
        grad(p.u, duphi, dulam);
        grad(p.v, dvphi, dvlam);
        dulam = dulam * dco;
        dvlam = dvlam * dco;
    }
    p.indep_g[ij] = indep_g;
    p.ddphi[ij] = ddphi; p.ddlam[ij] = ddlam;
    p.duphi[ij] = duphi; p.dulam[ij] = dulam; p.dvphi[ij] = dvphi; p.dvlam[ij] = dvlam;
    const bool deep_nocur = !t.refr_c && !t.shallow;
    p.della0[ij] = deep_nocur ? (p.delpro * dco) / dlam : ((0.5f * p.delpro) * dco) / dlam;
    p.drgm[ij] = t.SINPH_ROW[kx] * dco;
}

__global__ void prep_coef_kernel(DevTables t, float *__restrict__ coef)
{
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= t.NSEA) return;
    const int KL = t.KL;
    const int bin = blockIdx.y, m = bin / KL, k = bin % KL;
    const size_t NB = (size_t)KL * t.ML * t.NSEA;
    const size_t at = (size_t)bin * t.NSEA + ij;
    float c[P_NPLANES];
    if (t.refr_c) {
        coef_bin<true>(t, ij, m, k, c);
#pragma unroll
        for (int p = 0; p < P_NPLANES; p++) coef[p * NB + at] = c[p];
    } else {
        coef_bin<false>(t, ij, m, k, c);
#pragma unroll
        for (int p = 0; p < 5; p++) coef[p * NB + at] = c[p];
    }
}

// min and max that return NaN when either argument is NaN (CHECK_CFL refuses
// NaN weights), the reference's MIN/MAX otherwise
__device__ __forceinline__ float nan_min(float a, float b) { return a != a ? a : (b != b ? b : wam_minf(a, b)); }
__device__ __forceinline__ float nan_max(float a, float b) { return a != a ? a : (b != b ? b : wam_maxf(a, b)); }

// Per-block min and max of one coefficient plane (order-free, exact; NaN if
// the plane holds one).
__global__ void minmax_kernel(const float *__restrict__ x, size_t n, float *__restrict__ out)
{
    __shared__ float smin[256], smax[256];
    float lo = 3.4e38f, hi = -3.4e38f;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = x[i];
        lo = nan_min(lo, v); hi = nan_max(hi, v);
    }
    smin[threadIdx.x] = lo; smax[threadIdx.x] = hi;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = nan_min(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = nan_max(smax[threadIdx.x], smax[threadIdx.x + s]);
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

// Chunk of blockDim.x sea points handled by this block (the grid a multiple
// of 8): each XCD (blocks are dispatched round-robin over the 8 XCDs,
// blockIdx.x % 8) owns one contiguous eighth of the chunks.
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

// Ensemble members share the grid, its weights and per-point fields; member
// blockIdx.y owns the spectra at this offset of the input and output.
__device__ __forceinline__ size_t member_offset(const DevTables &t)
{
    return (size_t)blockIdx.y * t.KL * t.ML * t.NSEA;
}

// PROP_NO_CURR: F3 = SUM*F + LAT*F(KLAT(IC)) + LON*F(KLON(IS)) + TH1*F(K+1) + TH2*F(K-1)
__global__ void __launch_bounds__(256) propags_nocurr_kernel(DevTables t, const float *__restrict__ coef,
                                                              const float *__restrict__ fin,
                                                              float *__restrict__ fout)
{
    fin += member_offset(t);
    fout += member_offset(t);
    const int wg = blockDim.x, nch = (t.NSEA + wg - 1) / wg;
    const int ch = xcd_chunk(nch);
    if (ch >= nch) return;
    const int ij = ch * wg + threadIdx.x;
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
    fin += member_offset(t);
    fout += member_offset(t);
    const int wg = blockDim.x, nch = (t.NSEA + wg - 1) / wg;
    const int ch = xcd_chunk(nch);
    if (ch >= nch) return;
    const int ij = ch * wg + threadIdx.x;
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
    fin += member_offset(t);
    fout += member_offset(t);
    const int wg = blockDim.x, nch = (t.NSEA + wg - 1) / wg;
    const int ch = xcd_chunk(nch);
    if (ch >= nch) return;
    const int ij = ch * wg + threadIdx.x;
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

// CHECK_CFL's minima and maxima of every weight plane without storing the
// planes: a thread computes its sea point's weights bin by bin, the workgroup
// reduces, out[blk][plane][2] = {min, max} (order-free, exact).
template <bool CURR>
__global__ void __launch_bounds__(256) cfl_minmax_kernel(DevTables t, float *__restrict__ out)
{
    __shared__ float red[2 * P_NPLANES][256];
    const int ij = blockIdx.x * 256 + threadIdx.x;
    constexpr int np = CURR ? P_NPLANES : 5;
    float lo[P_NPLANES], hi[P_NPLANES];
#pragma unroll
    for (int p = 0; p < P_NPLANES; p++) { lo[p] = 3.4e38f; hi[p] = -3.4e38f; }
    if (ij < t.NSEA) {
        for (int m = 0; m < t.ML; m++)
            for (int k = 0; k < t.KL; k++) {
                float c[P_NPLANES];
                coef_bin<CURR>(t, ij, m, k, c);
#pragma unroll
                for (int p = 0; p < np; p++) { lo[p] = nan_min(lo[p], c[p]); hi[p] = nan_max(hi[p], c[p]); }
            }
    }
#pragma unroll
    for (int p = 0; p < P_NPLANES; p++) { red[2 * p][threadIdx.x] = lo[p]; red[2 * p + 1][threadIdx.x] = hi[p]; }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int p = 0; p < P_NPLANES; p++) {
                red[2 * p][threadIdx.x] = nan_min(red[2 * p][threadIdx.x], red[2 * p][threadIdx.x + w]);
                red[2 * p + 1][threadIdx.x] = nan_max(red[2 * p + 1][threadIdx.x], red[2 * p + 1][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 2 * P_NPLANES) out[(size_t)blockIdx.x * 2 * P_NPLANES + threadIdx.x] = red[threadIdx.x][0];
}

// PROP_CURR / PROP_NO_CURR with the weights computed in place (coef_bin, then
// CHECK_CFL's rescale for COUNTER > 1 with the same expressions as
// rescale_kernel), so no weight planes are read: the stencil's HBM traffic is
// the spectrum alone.  Same sums in the same order as the stored-weight kernels.
template <bool CURR>
__global__ void __launch_bounds__(256) propags_fused_kernel(DevTables t, int counter, const float *__restrict__ fin,
                                                             float *__restrict__ fout)
{
    fin += member_offset(t);
    fout += member_offset(t);
    const int wg = blockDim.x, nch = (t.NSEA + wg - 1) / wg;
    const int ch = xcd_chunk(nch);
    if (ch >= nch) return;
    const int ij = ch * wg + threadIdx.x;
    if (ij >= t.NSEA) return;
    const int KL = t.KL, ML = t.ML, NS = t.NSEA;
    const float rc = (float)counter;
    const int ns = t.KLAT_S[ij], nn = t.KLAT_N[ij], nw = t.KLON_W[ij], ne = t.KLON_E[ij];
    for (int m = 0; m < ML; m++) {
        const int mm = m > 0 ? m - 1 : 0, mp = m < ML - 1 ? m + 1 : ML - 1;
        for (int k = 0; k < KL; k++) {
            float c[P_NPLANES];
            coef_bin<CURR>(t, ij, m, k, c);
            if (counter > 1) {
                c[0] = 1.f - (1.f - c[0]) / rc;
#pragma unroll
                for (int p = 1; p < (CURR ? P_NPLANES : 5); p++) c[p] = c[p] / rc;
            }
            const size_t row = ((size_t)m * KL + k) * NS;
            const float *f = fin + row;
            const float *fkp = fin + ((size_t)m * KL + (t.KPM_P[k] - 1)) * NS;
            const float *fkm = fin + ((size_t)m * KL + (t.KPM_M[k] - 1)) * NS;
            float s;
            if (CURR) {
                const float *fmp = fin + ((size_t)mp * KL + k) * NS;
                const float *fmm = fin + ((size_t)mm * KL + k) * NS;
                s = c[P_SUM] * f[ij];
                s = s + c[P_LAT2] * ld_or0(f, nn);
                s = s + c[P_LAT1] * ld_or0(f, ns);
                s = s + c[P_LON2] * ld_or0(f, ne);
                s = s + c[P_LON1] * ld_or0(f, nw);
                s = s + c[P_TH1] * fkp[ij];
                s = s + c[P_TH2] * fkm[ij];
                s = s + c[P_SIG1] * fmp[ij];
                s = s + c[P_SIG2] * fmm[ij];
            } else {
                const int nlat = t.JYO1[k] == 1 ? ns : nn;
                const int nlon = t.JXO1[k] == 1 ? nw : ne;
                s = c[0] * f[ij] + c[1] * ld_or0(f, nlat);
                s = s + c[2] * ld_or0(f, nlon);
                s = s + c[3] * fkp[ij];
                s = s + c[4] * fkm[ij];
            }
            __builtin_nontemporal_store(s, &fout[row + ij]);
        }
    }
}

// PROP_CURR with PREP_SPHER_SHALLOW_CURR_REG's weights computed in place, for
// shallow water with current refraction.  The same values as coef_bin<true>
// (every weight is the same sequence of operations), with the work hoisted:
// per-point fields once, the per-frequency group velocities and depth-table
// factors carried across the M loop, the per-direction refraction terms in a
// window carried across the (cyclic) K loop, the per-direction constants from
// KDIR (host-computed with the same operations).  Rows of F are addressed as
// a uniform row pointer plus a per-lane 32-bit byte offset; a land neighbour's
// offset lies past the buffer's end, where the load returns +0 (the
// reference's land slot).  MIN(0, x) is v_min_f32 (the same value as the
// reference's for every x that is not NaN, signed zeros included); MAX(0, x)
// in the running centre weight is v_max_f32, which differs only in the sign of
// a zero result, and a zero of either sign leaves that sum unchanged (it
// starts at 1 - MAX(0, DLE) and never becomes -0).  CHECK_CFL refuses NaN
// weights, so no argument here is NaN.  RESCALE is CHECK_CFL's division of the
// weights by COUNTER: 0 none (COUNTER = 1), 1 a power of two (the exact
// multiplication by 1/COUNTER), 2 a true division.
template <int RESCALE>
__global__ void __launch_bounds__(256) propags_fused_sc_kernel(DevTables t, int counter,
                                                                const float *__restrict__ fin,
                                                                float *__restrict__ fout)
{
    fin += member_offset(t);
    fout += member_offset(t);
    // KDIR in LDS (KL <= 64): uniform reads of it are broadcasts
    __shared__ float4 kdir4[2 * 64];
    const int KL = t.KL, ML = t.ML, NS = t.NSEA, ND = t.NDEPTH;
    for (int i = threadIdx.x; i < 2 * KL; i += blockDim.x) kdir4[i] = reinterpret_cast<const float4 *>(t.KDIR)[i];
    __syncthreads();
    const float *kdir = reinterpret_cast<const float *>(kdir4);
    const int wg = blockDim.x, nch = (NS + wg - 1) / wg;
    const int ch = xcd_chunk(nch);
    if (ch >= nch) return;
    const int ij = ch * wg + threadIdx.x;
    if (ij >= NS) return;
    const int ns = t.KLAT_S[ij], nn = t.KLAT_N[ij], nw = t.KLON_W[ij], ne = t.KLON_E[ij];
    const bool hs = ns >= 0, hn = nn >= 0, hw = nw >= 0, he = ne >= 0;
    // the buffer spans three frequency planes (3*planeb bytes); a land
    // neighbour's offset is that length, out of range whatever the row's
    // scalar offset (the host keeps 6*planeb below 2**32, so no sum wraps)
    const unsigned rowb = 4u * NS, planeb = rowb * KL, oob = 3u * planeb;
    const unsigned o_c = 4u * ij, o_s = hs ? 4u * ns : oob, o_n = hn ? 4u * nn : oob;
    const unsigned o_w = hw ? 4u * nw : oob, o_e = he ? 4u * ne : oob;
    // per point
    const float u = t.U[ij], v = t.V[ij];
    const float uw = hw ? t.U[nw] : 0.f, ue = he ? t.U[ne] : 0.f;
    const float vs = hs ? t.V[ns] : 0.f, vn = hn ? t.V[nn] : 0.f;
    const int i0 = t.INDEP[ij] - 1;
    const int iw = hw ? t.INDEP[nw] - 1 : ND - 1, ie = he ? t.INDEP[ne] - 1 : ND - 1;
    const int is = hs ? t.INDEP[ns] - 1 : ND - 1, in = hn ? t.INDEP[nn] - 1 : ND - 1;
    const int ig = t.INDEP_G[ij] - 1;
    const float ddphi = t.DDPHI[ij], ddlam = t.DDLAM[ij];
    const float duphi = t.DUPHI[ij], dulam = t.DULAM[ij], dvphi = t.DVPHI[ij], dvlam = t.DVLAM[ij];
    const float drgm = t.DRGM[ij], della0 = t.DELLA0[ij], dpsn_s = t.DPSN_S[ij], dpsn_n = t.DPSN_N[ij];
    const float tempd = v * ddphi + u * ddlam;
    const float rc = (float)counter, rinv = 1.0f / rc;
    // F through one buffer resource per frequency M, based at plane M-1
    // (clamped): rows (M', K) at scalar byte offsets (M' - M-1)*planeb +
    // K*rowb, the sea point (or neighbour) as the lane's byte offset
    auto rsrc = [&](const float *base, int m, unsigned bytes) {
        // the base as a wave-uniform value (kept in scalar registers)
        const uint64_t a = reinterpret_cast<uint64_t>(base + (size_t)m * KL * NS);
        const uint64_t u = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a) |
                           (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32)) << 32;
        return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(u), (short)0, (int)bytes, 0x00020000);
    };
    auto ld = [&](__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
    };
    // group velocity, TSIHKD and TFAK of the point's depth rows at frequency m
    auto cgv = [&](int idx, int m) { return t.TCGOND[m * ND + idx]; };
    float cg_m = cgv(i0, 0), ts_m = t.TSIHKD[ig], tf_m = t.TFAK[ig];
    float cg_mm = cg_m, ts_mm = ts_m, tf_mm = tf_m;                 // M-1 clamps to 1
    for (int m = 0; m < ML; m++) {
        const int mp = m < ML - 1 ? m + 1 : ML - 1;
        const float cg_mp = cgv(i0, mp), ts_mp = t.TSIHKD[mp * ND + ig], tf_mp = t.TFAK[mp * ND + ig];
        const int mm = m > 0 ? m - 1 : 0;
        const float dfp = t.DELFR0 / t.FR[m], dfm = t.DELFR0 / t.FR[mm];
        const float cgw = cgv(iw, m), cge = cgv(ie, m), cgs = cgv(is, m), cgn = cgv(in, m);
        const __amdgpu_buffer_rsrc_t pf = rsrc(fin, mm, 3 * planeb), po = rsrc(fout, m, planeb);
        const unsigned sm0 = (m - mm) * planeb, sp0 = (mp - mm) * planeb;   // planes M, M+1 from plane M-1
        // THDD(K, M) = temp(K)*TSIHKD(M) + THDC(K); window over K-1, K, K+1
        auto thdd_k = [&](int k) {
            const float *kd = kdir + 8 * k;
            const float s = kd[0], c = kd[1];
            const float temp = s * ddphi - c * ddlam;
            const float ss = kd[4], sc = kd[5], cc = kd[6];
            const float thdc = ((ss * duphi + sc * dvphi) - sc * dulam) - cc * dvlam;
            return temp * ts_m + thdc;
        };
        float th_km = thdd_k(KL - 1), th_k = thdd_k(0), th_kp = thdd_k(KL > 1 ? 1 : 0);
        // the point's rows K-1, K, K+1 of plane M rotate through registers; the
        // loads of the next direction are issued before this one's arithmetic
        const int k1 = KL > 1 ? 1 : 0;
        float r_km = ld(pf, o_c, sm0 + (KL - 1) * rowb), r_c = ld(pf, o_c, sm0), r_kp = ld(pf, o_c, sm0 + k1 * rowb);
        float q_n = ld(pf, o_n, sm0), q_s = ld(pf, o_s, sm0), q_e = ld(pf, o_e, sm0), q_w = ld(pf, o_w, sm0);
        float q_mp = ld(pf, o_c, sp0), q_mm = ld(pf, o_c, 0);
        for (int k = 0; k < KL; k++) {
            const int kp = k + 1 < KL ? k + 1 : 0, km = k > 0 ? k - 1 : KL - 1;
            const float *kd = kdir + 8 * k;
            const float sd = kd[0], cd = kd[1], sp = kd[2], sm = kd[3];
            // the nine spectral values of the stencil
            const float f_c = r_c, f_kp = r_kp, f_km = r_km;
            const float l_n = q_n, l_s = q_s, l_e = q_e, l_w = q_w, f_mp = q_mp, f_mm = q_mm;
            {
                // the next direction's loads (at the last direction, rows of direction 0 again: unused)
                const int kn = kp, kn2 = kn + 1 < KL ? kn + 1 : 0;
                const unsigned rn = sm0 + kn * rowb;
                r_km = r_c; r_c = r_kp;
                r_kp = ld(pf, o_c, sm0 + kn2 * rowb);
                q_n = ld(pf, o_n, rn); q_s = ld(pf, o_s, rn); q_e = ld(pf, o_e, rn); q_w = ld(pf, o_w, rn);
                q_mp = ld(pf, o_c, sp0 + kn * rowb); q_mm = ld(pf, o_c, kn * rowb);
            }
            const float w1 = u + sd * cg_m;
            const float w2 = (v + cd * cg_m) * t.DELPH0;
            const float dlw = (w1 + (uw + sd * cgw)) * della0;
            const float dle = (w1 + (ue + sd * cge)) * della0;
            const float mn_dlw = __builtin_fminf(0.f, dlw), mn_dle = __builtin_fminf(0.f, dle);
            float dtc = (1.f - __builtin_fmaxf(0.f, dle)) + mn_dlw;
            const float dps = w2 + dpsn_s * ((vs + cd * cgs) * t.DELPH0);
            const float dpn = w2 + dpsn_n * ((vn + cd * cgn) * t.DELPH0);
            const float mn_dps = __builtin_fminf(0.f, dps), mn_dpn = __builtin_fminf(0.f, dpn);
            dtc = (dtc - __builtin_fmaxf(0.f, dpn)) + mn_dps;
            const float dtp = ((sp * drgm) * cg_m) + (th_k + th_kp) * t.DELTHR;
            const float dtm = ((sm * drgm) * cg_m) + (th_k + th_km) * t.DELTHR;
            const float mn_dtp = __builtin_fminf(0.f, dtp), mn_dtm = __builtin_fminf(0.f, dtm);
            dtc = (dtc - __builtin_fmaxf(0.f, dtp)) + mn_dtm;
            // SIDC(K, M') = (S0(K)*CG(M') + TEMPD*TSIHKD(M'))*TFAK(M')
            const float ss = kd[4], sc = kd[5], cc = kd[6];
            const float s0 = (((-(sc * duphi)) - cc * dvphi) - ss * dulam) - sc * dvlam;
            const float sk = (s0 * cg_m + tempd * ts_m) * tf_m;
            const float dop = (sk + (s0 * cg_mp + tempd * ts_mp) * tf_mp) * dfp;
            const float dom = (sk + (s0 * cg_mm + tempd * ts_mm) * tf_mm) * dfm;
            const float mn_dop = __builtin_fminf(0.f, dop);
            dtc = (dtc - __builtin_fmaxf(0.f, dop)) + __builtin_fminf(0.f, dom);
            float c[P_NPLANES];
            c[P_SUM] = dtc;
            c[P_LAT1] = wam_maxf(0.f, dps);
            c[P_LAT2] = -mn_dpn;
            c[P_LON1] = wam_maxf(0.f, dlw);
            c[P_LON2] = -mn_dle;
            c[P_TH1] = -mn_dtp;
            c[P_TH2] = wam_maxf(0.f, dtm);
            c[P_SIG1] = wam_div_1p1(-mn_dop);
            c[P_SIG2] = wam_maxf(0.f, dom) * 1.1f;
            if (RESCALE == 1) {
                c[0] = 1.f - (1.f - c[0]) * rinv;
#pragma unroll
                for (int p = 1; p < P_NPLANES; p++) c[p] = c[p] * rinv;
            } else if (RESCALE == 2) {
                c[0] = 1.f - (1.f - c[0]) / rc;
#pragma unroll
                for (int p = 1; p < P_NPLANES; p++) c[p] = c[p] / rc;
            }
            float sacc = c[P_SUM] * f_c;
            sacc = sacc + c[P_LAT2] * l_n;
            sacc = sacc + c[P_LAT1] * l_s;
            sacc = sacc + c[P_LON2] * l_e;
            sacc = sacc + c[P_LON1] * l_w;
            sacc = sacc + c[P_TH1] * f_kp;
            sacc = sacc + c[P_TH2] * f_km;
            sacc = sacc + c[P_SIG1] * f_mp;
            sacc = sacc + c[P_SIG2] * f_mm;
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(sacc), po, o_c, k * rowb, 2);   // nt
            th_km = th_k;
            th_k = th_kp;
            th_kp = thdd_k(k + 2 < KL ? k + 2 : k + 2 - KL);
        }
        cg_mm = cg_m; ts_mm = ts_m; tf_mm = tf_m;
        cg_m = cg_mp; ts_m = ts_mp; tf_m = tf_mp;
    }
}

// PROP_CURR with the weights computed in place (as propags_fused_sc_kernel,
// the same operations), in register blocks of KB directions with the
// frequencies inside.  For its block K0 .. K0+KB-1 a thread keeps the point's
// rows K0-1 .. K0+KB of planes M-1, M and M+1 in registers, so that each row
// of F is read once per block as plane M+1; only the two rows beside the
// block are read again by the neighbouring blocks.  What depends on the point
// and K alone (THDD's temp and THDC, SIDC's S0, SINTH*DRGM) is formed once per
// block, SIDC(K, M') is carried across M, and DELFR0/FR(M), DELFR0/FR(M-1)
// come from a per-workgroup table.  Land neighbours, MIN/MAX and the division
// by 1.1 as in propags_fused_sc_kernel.  KL must be a multiple of KB.
template <int RESCALE, int KB>
__global__ void __launch_bounds__(256) propags_blk_kernel(DevTables t, int counter, const float *__restrict__ fin,
                                                          float *__restrict__ fout)
{
    fin += member_offset(t);
    fout += member_offset(t);
    __shared__ float4 kdir4[2 * 64];
    __shared__ float2 dfr[64];
    const int KL = t.KL, ML = t.ML, NS = t.NSEA, ND = t.NDEPTH;
    for (int i = threadIdx.x; i < 2 * KL; i += blockDim.x) kdir4[i] = reinterpret_cast<const float4 *>(t.KDIR)[i];
    for (int m = threadIdx.x; m < ML; m += blockDim.x)
        dfr[m] = make_float2(t.DELFR0 / t.FR[m], t.DELFR0 / t.FR[m > 0 ? m - 1 : 0]);
    __syncthreads();
    const float *kdir = reinterpret_cast<const float *>(kdir4);
    const int wg = blockDim.x, nch = (NS + wg - 1) / wg;
    const int ch = xcd_chunk(nch);
    if (ch >= nch) return;
    const int ij = ch * wg + threadIdx.x;
    if (ij >= NS) return;
    const int ns = t.KLAT_S[ij], nn = t.KLAT_N[ij], nw = t.KLON_W[ij], ne = t.KLON_E[ij];
    const bool hs = ns >= 0, hn = nn >= 0, hw = nw >= 0, he = ne >= 0;
    // F through one buffer resource per frequency M, based at plane M-1
    // (clamped), spanning three planes; a land neighbour's offset is the
    // span's length, out of range, where the load returns +0
    const unsigned rowb = 4u * NS, planeb = rowb * KL, oob = 3u * planeb;
    const unsigned o_c = 4u * ij, o_s = hs ? 4u * ns : oob, o_n = hn ? 4u * nn : oob;
    const unsigned o_w = hw ? 4u * nw : oob, o_e = he ? 4u * ne : oob;
    const float u = t.U[ij], v = t.V[ij];
    const float uw = hw ? t.U[nw] : 0.f, ue = he ? t.U[ne] : 0.f;
    const float vs = hs ? t.V[ns] : 0.f, vn = hn ? t.V[nn] : 0.f;
    const int i0 = t.INDEP[ij] - 1;
    const int iw = hw ? t.INDEP[nw] - 1 : ND - 1, ie = he ? t.INDEP[ne] - 1 : ND - 1;
    const int is = hs ? t.INDEP[ns] - 1 : ND - 1, in = hn ? t.INDEP[nn] - 1 : ND - 1;
    const int ig = t.INDEP_G[ij] - 1;
    const float ddphi = t.DDPHI[ij], ddlam = t.DDLAM[ij];
    const float duphi = t.DUPHI[ij], dulam = t.DULAM[ij], dvphi = t.DVPHI[ij], dvlam = t.DVLAM[ij];
    const float drgm = t.DRGM[ij], della0 = t.DELLA0[ij], dpsn_s = t.DPSN_S[ij], dpsn_n = t.DPSN_N[ij];
    const float tempd = v * ddphi + u * ddlam;
    const float rc = (float)counter, rinv = 1.0f / rc;
    auto rsrc = [&](const float *base, int m, unsigned bytes) {
        const uint64_t a = reinterpret_cast<uint64_t>(base + (size_t)m * KL * NS);
        // the base as a wave-uniform value (kept in scalar registers)
        const uint64_t uu = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a) |
                            (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32)) << 32;
        return __builtin_amdgcn_make_buffer_rsrc(reinterpret_cast<void *>(uu), (short)0, (int)bytes, 0x00020000);
    };
    auto ld = [&](__amdgpu_buffer_rsrc_t r, unsigned off, unsigned soff) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0));
    };
    auto cgv = [&](int idx, int m) { return t.TCGOND[m * ND + idx]; };
    for (int kb = 0; kb < KL; kb += KB) {
        // THDD's direction terms of K0-1 .. K0+KB (cyclic)
        int kk[KB + 2];
        float tp[KB + 2], tc[KB + 2];
#pragma unroll
        for (int d = 0; d < KB + 2; d++) {
            const int k = (kb - 1 + d + KL) % KL;
            kk[d] = k;
            const float *kd = kdir + 8 * k;
            tp[d] = kd[0] * ddphi - kd[1] * ddlam;
            tc[d] = ((kd[4] * duphi + kd[5] * dvphi) - kd[5] * dulam) - kd[6] * dvlam;
        }
        float s0[KB], spd[KB], smd[KB], sdk[KB], cdk[KB];
#pragma unroll
        for (int j = 0; j < KB; j++) {
            const float *kd = kdir + 8 * (kb + j);
            sdk[j] = kd[0]; cdk[j] = kd[1];
            spd[j] = kd[2] * drgm; smd[j] = kd[3] * drgm;
            s0[j] = (((-(kd[5] * duphi)) - kd[6] * dvphi) - kd[4] * dulam) - kd[5] * dvlam;
        }
        // the rows K0-1 .. K0+KB of planes M-1, M, M+1 (M-1 clamps to 1)
        float P_mm[KB + 2], P_c[KB + 2], P_mp[KB + 2];
        {
            const __amdgpu_buffer_rsrc_t p0 = rsrc(fin, 0, planeb);
#pragma unroll
            for (int d = 0; d < KB + 2; d++) { P_c[d] = ld(p0, o_c, kk[d] * rowb); P_mm[d] = P_c[d]; }
        }
        float cg_m = cgv(i0, 0), ts_m = t.TSIHKD[ig], tf_m = t.TFAK[ig];
        // SIDC(K, M) and SIDC(K, M-1) of the block's directions
        float S_c[KB], S_mm[KB];
#pragma unroll
        for (int j = 0; j < KB; j++) { S_c[j] = (s0[j] * cg_m + tempd * ts_m) * tf_m; S_mm[j] = S_c[j]; }
        for (int m = 0; m < ML; m++) {
            const int mp = m < ML - 1 ? m + 1 : ML - 1, mm = m > 0 ? m - 1 : 0;
            const float cg_mp = cgv(i0, mp), ts_mp = t.TSIHKD[mp * ND + ig], tf_mp = t.TFAK[mp * ND + ig];
            const float cgw = cgv(iw, m), cge = cgv(ie, m), cgs = cgv(is, m), cgn = cgv(in, m);
            const float2 df = dfr[m];
            const __amdgpu_buffer_rsrc_t pf = rsrc(fin, mm, 3 * planeb), po = rsrc(fout, m, planeb);
            const unsigned sm0 = (m - mm) * planeb, sp0 = (mp - mm) * planeb;
            // plane M+1 (at M = ML the plane itself again, the clamp)
#pragma unroll
            for (int d = 0; d < KB + 2; d++) P_mp[d] = ld(pf, o_c, sp0 + kk[d] * rowb);
            float th[KB + 2];
#pragma unroll
            for (int d = 0; d < KB + 2; d++) th[d] = tp[d] * ts_m + tc[d];
#pragma unroll
            for (int j = 0; j < KB; j++) {
                const int k = kb + j;
                const unsigned rk = sm0 + k * rowb;
                const float l_n = ld(pf, o_n, rk), l_s = ld(pf, o_s, rk), l_e = ld(pf, o_e, rk), l_w = ld(pf, o_w, rk);
                const float f_c = P_c[j + 1], f_kp = P_c[j + 2], f_km = P_c[j], f_mp = P_mp[j + 1], f_mm = P_mm[j + 1];
                const float sd = sdk[j], cd = cdk[j];
                const float w1 = u + sd * cg_m;
                const float w2 = (v + cd * cg_m) * t.DELPH0;
                const float dlw = (w1 + (uw + sd * cgw)) * della0;
                const float dle = (w1 + (ue + sd * cge)) * della0;
                const float mn_dlw = __builtin_fminf(0.f, dlw), mn_dle = __builtin_fminf(0.f, dle);
                float dtc = (1.f - __builtin_fmaxf(0.f, dle)) + mn_dlw;
                const float dps = w2 + dpsn_s * ((vs + cd * cgs) * t.DELPH0);
                const float dpn = w2 + dpsn_n * ((vn + cd * cgn) * t.DELPH0);
                const float mn_dps = __builtin_fminf(0.f, dps), mn_dpn = __builtin_fminf(0.f, dpn);
                dtc = (dtc - __builtin_fmaxf(0.f, dpn)) + mn_dps;
                const float dtp = (spd[j] * cg_m) + (th[j + 1] + th[j + 2]) * t.DELTHR;
                const float dtm = (smd[j] * cg_m) + (th[j + 1] + th[j]) * t.DELTHR;
                const float mn_dtp = __builtin_fminf(0.f, dtp), mn_dtm = __builtin_fminf(0.f, dtm);
                dtc = (dtc - __builtin_fmaxf(0.f, dtp)) + mn_dtm;
                const float s_mp = (s0[j] * cg_mp + tempd * ts_mp) * tf_mp;
                const float dop = (S_c[j] + s_mp) * df.x;
                const float dom = (S_c[j] + S_mm[j]) * df.y;
                const float mn_dop = __builtin_fminf(0.f, dop);
                dtc = (dtc - __builtin_fmaxf(0.f, dop)) + __builtin_fminf(0.f, dom);
                float c[P_NPLANES];
                c[P_SUM] = dtc;
                c[P_LAT1] = wam_maxf(0.f, dps);
                c[P_LAT2] = -mn_dpn;
                c[P_LON1] = wam_maxf(0.f, dlw);
                c[P_LON2] = -mn_dle;
                c[P_TH1] = -mn_dtp;
                c[P_TH2] = wam_maxf(0.f, dtm);
                c[P_SIG1] = wam_div_1p1(-mn_dop);
                c[P_SIG2] = wam_maxf(0.f, dom) * 1.1f;
                if (RESCALE == 1) {
                    c[0] = 1.f - (1.f - c[0]) * rinv;
#pragma unroll
                    for (int p = 1; p < P_NPLANES; p++) c[p] = c[p] * rinv;
                } else if (RESCALE == 2) {
                    c[0] = 1.f - (1.f - c[0]) / rc;
#pragma unroll
                    for (int p = 1; p < P_NPLANES; p++) c[p] = c[p] / rc;
                }
                float sacc = c[P_SUM] * f_c;
                sacc = sacc + c[P_LAT2] * l_n;
                sacc = sacc + c[P_LAT1] * l_s;
                sacc = sacc + c[P_LON2] * l_e;
                sacc = sacc + c[P_LON1] * l_w;
                sacc = sacc + c[P_TH1] * f_kp;
                sacc = sacc + c[P_TH2] * f_km;
                sacc = sacc + c[P_SIG1] * f_mp;
                sacc = sacc + c[P_SIG2] * f_mm;
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(sacc), po, o_c, k * rowb, 2);   // nt
                S_mm[j] = S_c[j];
                S_c[j] = s_mp;
            }
#pragma unroll
            for (int d = 0; d < KB + 2; d++) { P_mm[d] = P_c[d]; P_c[d] = P_mp[d]; }
            cg_m = cg_mp; ts_m = ts_mp; tf_m = tf_mp;
        }
    }
}
template __global__ void propags_blk_kernel<0, 3>(DevTables, int, const float *, float *);
template __global__ void propags_blk_kernel<1, 3>(DevTables, int, const float *, float *);
template __global__ void propags_blk_kernel<2, 3>(DevTables, int, const float *, float *);

template __global__ void propags_fused_sc_kernel<0>(DevTables, int, const float *, float *);
template __global__ void propags_fused_sc_kernel<1>(DevTables, int, const float *, float *);
template __global__ void propags_fused_sc_kernel<2>(DevTables, int, const float *, float *);
template __global__ void cfl_minmax_kernel<true>(DevTables, float *);
template __global__ void cfl_minmax_kernel<false>(DevTables, float *);
template __global__ void propags_fused_kernel<true>(DevTables, int, const float *, float *);
template __global__ void propags_fused_kernel<false>(DevTables, int, const float *, float *);

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

// SBOTTOM's factor of every (sea point, frequency), [ML][NSEA]:
// (-2*0.038/G * XK) / SINH(MIN(2 XK DEPTH, 50)) with XK = TFAK(INDEP, M)
// (wam_source_module.f90:1527-1589), the operations of IMPLSCH's own
// evaluation; it depends on the depth field only.
__global__ void sbottom_factor_kernel(DevTables t, float *__restrict__ sbot)
{
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;
    if (ij >= t.NSEA) return;
    const int m = blockIdx.y;
    const float wav = t.TFAK[m * t.NDEPTH + t.INDEP[ij] - 1];
    const float sbo = wam_minf((2.f * t.DEPTH[ij]) * wav, 50.f);
    sbot[(size_t)m * t.NSEA + ij] = (((-2.0f * 0.038f) / t.G) * wav) / wam_sinhf(sbo);
}

}  // namespace wam
