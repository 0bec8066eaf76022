// wam_output.hip -- the integrated parameters of COMPUTE_OUTPUT_PARAMETER
// (wam_output_module.f90:481-601) on gfx950: wind, drag, Charnock, depth,
// current, Hs (TOTAL_ENERGY), peak period (PEAK_PERIOD), mean period
// (FEMEAN), TM1/TM2 (TM1_TM2_PERIODS), mean direction and spread
// (MEAN_DIRECTION), normalised wave stress and Z0, for every sea point.
//
// One lane per sea point (64 per wavefront): the point-fastest rows of FL3
// load coalesced, and every sum runs in the reference's order within its
// lane.  The per-direction sums of MEAN_DIRECTION (over frequencies, inner
// loop in the reference) are carried across the frequency loop in LDS, one
// column per lane.
#include "wam_dev.hpp"
#include "wam_math.h"

namespace wam {

constexpr int OUT_PW = 64;                // sea points per workgroup
constexpr int NOUT = 17;                  // parameters per sea point

// MOD(A, 360.) for 0 <= A < 720 (exact)
__device__ __forceinline__ float mod360(float a) { return a >= 360.f ? a - 360.f : a; }

__device__ __forceinline__ float out_dir(const DevTables &t, float a, int orient)
{
    return orient ? t.DEG * a : mod360(t.DEG * a + 180.f);
}

__global__ void __launch_bounds__(OUT_PW) output_params_kernel(DevTables t, const float *__restrict__ fl3,
                                                               const float *__restrict__ u10,
                                                               const float *__restrict__ udir,
                                                               const float *__restrict__ ustar,
                                                               const float *__restrict__ tauw,
                                                               const float *__restrict__ z0, int orient,
                                                               float *__restrict__ out)
{
    extern __shared__ float dk_lds[];             // [KL][OUT_PW]: MEAN_DIRECTION's TEMP(K) of each lane
    const int KL = t.KL, ML = t.ML, NS = t.NSEA, lane = threadIdx.x;
    const int pt0 = blockIdx.x * OUT_PW + lane;
    const bool on = pt0 < NS;
    const int pt = on ? pt0 : NS - 1;
    const float *F = fl3 + pt;                    // F(k, m) = F[(m*KL + k) * NS]
    float *dk = dk_lds + lane;
    auto dirsum = [&](int m) {
        float s = 0.f;
        for (int k = 0; k < KL; k++) {
            const float f = F[(size_t)(k + KL * m) * NS];
            s = k == 0 ? f : s + f;
        }
        return s;
    };
    // the highest frequency seeds the integrals (the tail terms)
    const float tl = dirsum(ML - 1);
    float em = t.MO_TAIL * tl, fm = t.MM1_TAIL * tl, tm1 = t.MP1_TAIL * tl, tm2 = t.MP2_TAIL * tl;
    float tmax = 0.f;
    int ip = 0;
    for (int m = 0; m < ML; m++) {
        float s = 0.f;
        const float wm = t.DFIM[m];
        for (int k = 0; k < KL; k++) {
            const float f = F[(size_t)(k + KL * m) * NS];
            s = k == 0 ? f : s + f;
            dk[k * OUT_PW] = m == 0 ? f * wm : dk[k * OUT_PW] + f * wm;
        }
        em = em + s * wm;
        fm = fm + s * t.DFIMOFR[m];
        tm1 = tm1 + s * t.DFIM_FR[m];
        tm2 = tm2 + s * t.DFIM_FR2[m];
        if (m == 0 || s > tmax) {              // MAXLOC: the first maximum
            tmax = s;
            ip = m;
        }
    }
    em = wam_maxf(em, t.EMIN);
    fm = em / wam_maxf(fm, t.EMIN);
    tm1 = em > t.EMIN ? em / tm1 : 1.f;
    tm2 = em > t.EMIN ? sqrtf(em / tm2) : 1.f;
    const float peakp = ip == 0 ? 1.f : 1.f / t.FR[ip];
    // MEAN_DIRECTION_B (wam_interface_module.f90:412-533)
    float si = 0.f, ci = 0.f, sp = 0.f;
    for (int k = 0; k < KL; k++) {
        const float te = dk[k * OUT_PW];
        si = si + t.SINTH[k] * te;
        ci = ci + t.COSTH[k] * te;
        sp = sp + te;
    }
    if (ci == 0.f) ci = 0.1E-30f;
    float thq = wam_atan2f(si, ci);
    if (thq < 0.f) thq = thq + t.ZPI;
    if (thq > t.ZPI - 0.001f) thq = 0.f;
    if (wam_fabsf(ci) < 0.1E-15f) ci = wam_signf(0.1E-15f, ci);
    if (wam_fabsf(si) < 0.1E-15f) si = wam_signf(0.1E-15f, si);
    if (wam_fabsf(sp) < 0.1E-15f) sp = wam_signf(0.1E-15f, sp);
    sp = 2.f * (1.f - sqrtf(si * si + ci * ci) / sp);
    sp = sp <= 0.f ? 1.17549435e-38f : sqrtf(sp);    // TINY(1.)
    // the wind and current parameters (wam_output_module.f90:525-563)
    const float us = ustar[pt], tau = us * us + 0.0001f, w = u10[pt];
    const float uu = t.U ? t.U[pt] : 0.f, vv = t.V ? t.V[pt] : 0.f;
    float cdir = wam_atan2f(uu, vv) * t.DEG;
    if (!orient) cdir = mod360(cdir + 180.f);
    if (cdir < 0.f) cdir = cdir + 360.f;
    if (!on) return;
    float *o = out + pt;
    o[0 * (size_t)NS] = w;
    o[1 * (size_t)NS] = out_dir(t, udir[pt], orient);
    o[2 * (size_t)NS] = us;
    o[3 * (size_t)NS] = tau / (w * w + 0.01f);
    o[4 * (size_t)NS] = (t.G * z0[pt]) / wam_maxf(us * us, 1.0E-6f);   // CHARNOCK_PAR (:159-203)
    o[5 * (size_t)NS] = t.shallow ? t.DEPTH[pt] : 999.f;
    o[6 * (size_t)NS] = sqrtf(uu * uu + vv * vv);
    o[7 * (size_t)NS] = cdir;
    o[8 * (size_t)NS] = 4.f * sqrtf(em);
    o[9 * (size_t)NS] = peakp;
    o[10 * (size_t)NS] = 1.f / fm;
    o[11 * (size_t)NS] = tm1;
    o[12 * (size_t)NS] = tm2;
    o[13 * (size_t)NS] = out_dir(t, thq, orient);
    o[14 * (size_t)NS] = sp * t.DEG;
    o[15 * (size_t)NS] = tauw[pt] / (tau * t.ROAIR);
    o[16 * (size_t)NS] = z0[pt];
}

hipError_t launch_output_params(const DevTables &t, const float *fl3, const float *u10, const float *udir,
                                const float *ustar, const float *tauw, const float *z0, int orient, float *out,
                                hipStream_t st)
{
    const int nwg = (t.NSEA + OUT_PW - 1) / OUT_PW;
    hipLaunchKernelGGL(output_params_kernel, dim3(nwg), dim3(OUT_PW), (size_t)t.KL * OUT_PW * sizeof(float), st, t,
                       fl3, u10, udir, ustar, tauw, z0, orient, out);
    return hipGetLastError();
}

// WRITE_INT_PAR_OUTPUT's GRID(IXLG(I), KXLT(I)) = BLOCK(I), ZMISS on land
// (wam_output_module.f90:866-868) for the rows picked by `rows`: one lane per
// grid cell, reading its sea point through the cell -> point map (-1 land);
// the sea-point order is the cell order, so the reads are near-coalesced.
struct MapRows {
    int p[NOUT];
};

__global__ void __launch_bounds__(256) map_grid_kernel(const float *__restrict__ out, int nsea,
                                                       const int *__restrict__ cell_pt, int ncell, MapRows rows,
                                                       float *__restrict__ grids)
{
    const int c = blockIdx.x * 256 + threadIdx.x, r = blockIdx.y;
    if (c >= ncell) return;
    const int ij = cell_pt[c];
    grids[(size_t)r * ncell + c] = ij >= 0 ? out[(size_t)rows.p[r] * nsea + ij] : -9999999.f;
}

hipError_t launch_map_grids(const float *out, int nsea, const int *cell_pt, int ncell, const int *rows, int nrows,
                            float *grids, hipStream_t st)
{
    if (nrows <= 0) return hipSuccess;
    MapRows r{};
    for (int i = 0; i < nrows && i < NOUT; i++) r.p[i] = rows[i];
    hipLaunchKernelGGL(map_grid_kernel, dim3((ncell + 255) / 256, nrows), dim3(256), 0, st, out, nsea, cell_pt,
                       ncell, r, grids);
    return hipGetLastError();
}

}  // namespace wam
