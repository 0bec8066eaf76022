// wam_nest.hip -- the nest boundary exchange on gfx950.
//
//  * BOUNDARY_OUTPUT (wam_boundary_module.f90:291-397): the spectra of the
//    coarse grid at the nest's boundary points (gathered from FL3 in HBM)
//    with TOTAL_ENERGY, MEAN_DIRECTION's THQ and FEMEAN's FM of each;
//  * BOUNDARY_INPUT (:168-287): the time interpolation of two sets of
//    boundary spectra and the space interpolation into FL3 at the fine grid's
//    boundary points, both by INTSPEC (wam_interface_module.f90:290-408) with
//    ROTSPEC (:891-941) and STRSPEC (:945-1006).
//
// INTSPEC's scalars (weights, mean values, the shift and stretch of either
// spectrum) depend on the spectrum pair only; each output value is then an
// independent function of at most eight input values (two directions x two
// frequencies x two spectra), so lanes split the KL*ML values with no
// exchange.  Work on FL3 (the gather and the insertion) takes one lane per
// boundary point, so that boundary points adjacent in the sea-point order
// (a nest's southern and northern edges) touch FL3's point-fastest rows
// coalesced; the time interpolation of the boundary sets, contiguous
// spectra, takes one workgroup per spectrum.  Every product and sum is
// rounded as the reference writes it (hipcc -ffp-contract=off), and the sums
// of the mean parameters run serially in the reference's order.  A spectrum
// is F(K, M) at f[M*KL + K].
#include "wam_dev.hpp"
#include "wam_math.h"

namespace wam {

constexpr int NEST_WG = 256;
constexpr int NEST_PG = 4;                       // boundary points per workgroup on FL3 (16-byte runs per row)
constexpr int NEST_SL = NEST_WG / NEST_PG;       // lanes per boundary point
constexpr float EPSILON_1 = 1.1920929e-7f;       // EPSILON(1.)

// ROTSPEC's rotation and STRSPEC's stretch of one spectrum, and its weight
struct Adjust {
    int sh_r, inc_s, mode;       // sh_r: ROTSPEC's INC modulo KL; mode 0: GAMMA = 1, 1: shifted out of range, 2: stretch
    float adif_r, bdif_r, adif_s, bdif_s, w;
};

__device__ Adjust make_adjust(const DevTables &t, float rthet, float gamma, float gw, float em, float emk)
{
    Adjust a;
    float ad = rthet * (float)t.KL / t.ZPI;
    const int inc = -(int)wam_floorf(ad);
    ad = ad + (float)inc;
    a.sh_r = inc % t.KL;
    if (a.sh_r < 0) a.sh_r += t.KL;
    a.adif_r = ad;
    a.bdif_r = 1.f - ad;
    a.inc_s = 0;
    a.adif_s = a.bdif_s = 0.f;
    if (gamma == 1.0f) {
        a.mode = 0;
    } else {
        a.inc_s = (int)wam_floorf(wam_log10f(gamma) / wam_log10f(t.CO));
        if (abs(a.inc_s) >= t.ML - 1) {
            a.mode = 1;
        } else {
            a.mode = 2;
            // CO**(-INC); a negative power as the reciprocal of the positive one
            const float p = a.inc_s <= 0 ? wam_ipowf(t.CO, -a.inc_s) : 1.f / wam_ipowf(t.CO, a.inc_s);
            a.adif_s = (t.CO - gamma * p) / (t.CO - 1.f);
            a.bdif_s = 1.f - a.adif_s;
        }
    }
    a.w = (gw * em) / emk;
    return a;
}

// BDIF*CSHIFT(F, INC) + ADIF*CSHIFT(F, INC-1) at (k, m)
__device__ __forceinline__ float rot_at(const float *f, const Adjust &a, int KL, int k, int m)
{
    int k0 = k + a.sh_r;
    if (k0 >= KL) k0 -= KL;
    const int k1 = k0 == 0 ? KL - 1 : k0 - 1;
    return a.bdif_r * f[m * KL + k0] + a.adif_r * f[m * KL + k1];
}

// STRSPEC(ROTSPEC(F)) at (k, m)
__device__ float adjusted(const float *f, const Adjust &a, int KL, int ML, int k, int m)
{
    if (a.mode == 0) return rot_at(f, a, KL, k, m);
    if (a.mode == 1) return 0.f;
    const int s = a.inc_s;
    if (s >= 0) {
        if (m < ML - s - 1) return a.adif_s * rot_at(f, a, KL, k, m + s) + a.bdif_s * rot_at(f, a, KL, k, m + s + 1);
        if (m == ML - s - 1) return a.adif_s * rot_at(f, a, KL, k, ML - 1);
        return 0.f;
    }
    if (m >= -s) return a.adif_s * rot_at(f, a, KL, k, m + s) + a.bdif_s * rot_at(f, a, KL, k, m + s + 1);
    if (m == -s - 1) return a.bdif_s * rot_at(f, a, KL, k, 0);
    return 0.f;
}

// INTSPEC: out[i*os] for i = M*KL + K = i0, i0 + is, ... (the caller's lanes
// cover all i); the mean values (EMEAN, THQ, FMEAN) to mo[0], mo[ms], mo[2*ms]
// from the lane with i0 == 0 when mo is given.
__device__ void intspec_part(const DevTables &t, float del12, float del1l, const float *f1, float em1, float th1,
                             float fm1, const float *f2, float em2, float th2, float fm2, float *out, size_t os,
                             float *mo, int ms, int i0, int is)
{
    const int KL = t.KL, ML = t.ML, NB = KL * ML, l = i0;
    const float gw2 = del1l / del12, gw1 = 1.f - gw2;
    if (wam_fabsf(gw1) < EPSILON_1 || em1 < EPSILON_1) {
        for (int i = i0; i < NB; i += is) out[i * os] = gw2 * f2[i];
        if (mo && l == 0) { mo[0] = gw2 * em2; mo[ms] = gw2 * th2; mo[2 * ms] = gw2 * fm2; }
        return;
    }
    if (wam_fabsf(gw2) < EPSILON_1 || em2 < EPSILON_1) {
        for (int i = i0; i < NB; i += is) out[i * os] = gw1 * f1[i];
        if (mo && l == 0) { mo[0] = gw1 * em1; mo[ms] = gw1 * th1; mo[2 * ms] = gw1 * fm1; }
        return;
    }
    const float em = gw1 * em1 + gw2 * em2;
    const float fm = gw1 * fm1 + gw2 * fm2;
    const float th = wam_atan2f(gw1 * wam_sinf(th1) + gw2 * wam_sinf(th2), gw1 * wam_cosf(th1) + gw2 * wam_cosf(th2));
    const Adjust a = make_adjust(t, th - th1, fm1 / fm, gw1, em, em1);
    const Adjust b = make_adjust(t, th - th2, fm2 / fm, gw2, em, em2);
    for (int i = i0; i < NB; i += is) {
        const int m = i / KL, k = i - m * KL;
        out[i * os] = a.w * adjusted(f1, a, KL, ML, k, m) + b.w * adjusted(f2, b, KL, ML, k, m);
    }
    if (mo && l == 0) { mo[0] = em; mo[ms] = th; mo[2 * ms] = fm; }
}

// BOUNDARY_OUTPUT: FBC(ib) = FL3(IJARC(ib), :, :) and its mean parameters
// mean[ib], mean[nb + ib], mean[2 nb + ib] = EMEAN, THQ, FMEAN.  NEST_PG
// boundary points per workgroup, consecutive lanes on consecutive points, so
// that points adjacent in the sea-point order read FL3 in 16-byte runs; the
// NEST_SL lanes of a point share its sums: lane s forms SUM(F, DIM=2) of
// frequencies s, s + NEST_SL, ... (serial over directions) and
// MEAN_DIRECTION's TEMP(K) of directions s, s + NEST_SL, ... (serial over
// frequencies), and lane 0 of each point finishes the reference's serial
// sums from LDS.
__global__ void __launch_bounds__(NEST_WG) boundary_gather_kernel(DevTables t, const float *__restrict__ fl3,
                                                                  const int *__restrict__ ijarc, int nb,
                                                                  float *__restrict__ fbc, float *__restrict__ mean)
{
    extern __shared__ float sums[];               // [ML][NEST_PG] direction sums, [KL][NEST_PG] frequency integrals
    const int KL = t.KL, ML = t.ML, NB = KL * ML;
    const int lane = threadIdx.x % NEST_PG, w = threadIdx.x / NEST_PG;
    const int ib0 = blockIdx.x * NEST_PG + lane;
    const bool on = ib0 < nb;
    const int ib = on ? ib0 : nb - 1;
    const size_t NS = t.NSEA;
    const float *F = fl3 + (ijarc[ib] - 1);
    float *o = fbc + (size_t)ib * NB;
    float *T = sums, *te = sums + ML * NEST_PG;
    for (int m = w; m < ML; m += NEST_SL) {
        float s = 0.f;
        for (int k = 0; k < KL; k++) {
            const float f = F[(size_t)(m * KL + k) * NS];
            s = k == 0 ? f : s + f;
            if (on) o[m * KL + k] = f;
        }
        T[m * NEST_PG + lane] = s;
    }
    for (int k = w; k < KL; k += NEST_SL) {
        float s = F[(size_t)k * NS] * t.DFIM[0];
        for (int m = 1; m < ML; m++) s = s + F[(size_t)(m * KL + k) * NS] * t.DFIM[m];
        te[k * NEST_PG + lane] = s;
    }
    __syncthreads();
    if (w != 0 || !on) return;
    float em = t.MO_TAIL * T[(ML - 1) * NEST_PG + lane];
    for (int m = 0; m < ML; m++) em = em + T[m * NEST_PG + lane] * t.DFIM[m];
    em = wam_maxf(em, t.EMIN);
    float fm = t.MM1_TAIL * T[(ML - 1) * NEST_PG + lane];
    for (int m = 0; m < ML; m++) fm = fm + T[m * NEST_PG + lane] * t.DFIMOFR[m];
    fm = em / wam_maxf(fm, t.EMIN);
    float si = 0.f, ci = 0.f;
    for (int k = 0; k < KL; k++) {
        si = si + t.SINTH[k] * te[k * NEST_PG + lane];
        ci = ci + t.COSTH[k] * te[k * NEST_PG + lane];
    }
    if (ci == 0.f) ci = 0.1E-30f;
    float thq = wam_atan2f(si, ci);
    if (thq < 0.f) thq = thq + t.ZPI;
    if (thq > t.ZPI - 0.001f) thq = 0.f;
    mean[ib] = em;
    mean[nb + ib] = thq;
    mean[2 * nb + ib] = fm;
}

// BOUNDARY_INPUT's time step: FI(p), MI(p) from the first set (which = 1),
// the second (2) or INTSPEC between them at DEL1L of DEL12 (0).
__global__ void __launch_bounds__(NEST_WG) boundary_time_kernel(DevTables t, int nbinp, const float *__restrict__ f1,
                                                                const float *__restrict__ m1,
                                                                const float *__restrict__ f2,
                                                                const float *__restrict__ m2, int which, float del12,
                                                                float del1l, float *__restrict__ fi,
                                                                float *__restrict__ mi)
{
    const int NB = t.KL * t.ML, p = blockIdx.x, l = threadIdx.x;
    if (which != 0) {
        const float *s = (which == 1 ? f1 : f2) + (size_t)p * NB, *sm = which == 1 ? m1 : m2;
        for (int i = l; i < NB; i += NEST_WG) fi[(size_t)p * NB + i] = s[i];
        if (l < 3) mi[l * nbinp + p] = sm[l * nbinp + p];
        return;
    }
    intspec_part(t, del12, del1l, f1 + (size_t)p * NB, m1[p], m1[nbinp + p], m1[2 * nbinp + p], f2 + (size_t)p * NB,
                 m2[p], m2[nbinp + p], m2[2 * nbinp + p], fi + (size_t)p * NB, 1, mi + p, nbinp, l, NEST_WG);
}

// BOUNDARY_INPUT's space step: FL3(IJARF(i), :, :) = INTSPEC(1, BFW(i),
// FI(IBFL(i)), FI(IBFR(i))).  NEST_PG fine boundary points per workgroup,
// consecutive lanes on consecutive points and NEST_SL lanes splitting the
// KL*ML values of each, so that points adjacent in the sea-point order write
// FL3 in 16-byte runs.  The host passes each IJARF once.
__global__ void __launch_bounds__(NEST_WG) boundary_space_kernel(DevTables t, int nbinp, int nbounf,
                                                                 const float *__restrict__ fi,
                                                                 const float *__restrict__ mi,
                                                                 const int *__restrict__ ijarf,
                                                                 const int *__restrict__ ibfl,
                                                                 const int *__restrict__ ibfr,
                                                                 const float *__restrict__ bfw, float *fl3)
{
    const int NB = t.KL * t.ML, i = blockIdx.x * NEST_PG + threadIdx.x % NEST_PG, w = threadIdx.x / NEST_PG;
    if (i >= nbounf) return;
    const int lft = ibfl[i] - 1, rgt = ibfr[i] - 1;
    const size_t ij = (size_t)(ijarf[i] - 1);
    intspec_part(t, 1.f, bfw[i], fi + (size_t)lft * NB, mi[lft], mi[nbinp + lft], mi[2 * nbinp + lft],
                 fi + (size_t)rgt * NB, mi[rgt], mi[nbinp + rgt], mi[2 * nbinp + rgt], fl3 + ij, (size_t)t.NSEA,
                 nullptr, 0, w, NEST_SL);
}

hipError_t launch_boundary_gather(const DevTables &t, const float *fl3, const int *ijarc, int nb, float *fbc,
                                  float *mean, hipStream_t st)
{
    if (nb <= 0) return hipSuccess;
    const size_t lds = (size_t)(t.ML + t.KL) * NEST_PG * sizeof(float);
    hipLaunchKernelGGL(boundary_gather_kernel, dim3((nb + NEST_PG - 1) / NEST_PG), dim3(NEST_WG), lds, st, t, fl3,
                       ijarc, nb, fbc, mean);
    return hipGetLastError();
}

hipError_t launch_boundary_time(const DevTables &t, int nbinp, const float *f1, const float *m1, const float *f2,
                                const float *m2, int which, float del12, float del1l, float *fi, float *mi,
                                hipStream_t st)
{
    if (nbinp <= 0) return hipSuccess;
    hipLaunchKernelGGL(boundary_time_kernel, dim3(nbinp), dim3(NEST_WG), 0, st, t, nbinp, f1, m1, f2, m2, which,
                       del12, del1l, fi, mi);
    return hipGetLastError();
}

hipError_t launch_boundary_space(const DevTables &t, int nbinp, const float *fi, const float *mi, int nbounf,
                                 const int *ijarf, const int *ibfl, const int *ibfr, const float *bfw, float *fl3,
                                 hipStream_t st)
{
    if (nbounf <= 0) return hipSuccess;
    hipLaunchKernelGGL(boundary_space_kernel, dim3((nbounf + NEST_PG - 1) / NEST_PG), dim3(NEST_WG), 0, st, t, nbinp,
                       nbounf, fi, mi, ijarf, ibfl, ibfr, bfw, fl3);
    return hipGetLastError();
}

}  // namespace wam
