// wam_setup.cpp -- host-side set-up of the WAM hot path (see wam_setup.hpp).
#include "wam_setup.hpp"

#include <cfloat>
#include <cmath>
#include <cstring>

#include "wam_math.h"

namespace wam {

static const int M_DEGREE = 360000;   // wam_coordinate_module.f90:41-43

// Fortran folds constant expressions with one correct rounding per operation;
// x**n of a constant is folded as a correctly rounded power.
static float fold_pow(float x, int n) { return (float)std::pow((double)x, (double)n); }
static float facos(float x) { return (float)std::acos((double)x); }
static inline float fmaxr(float a, float b) { return a > b ? a : b; }
static inline float fminr(float a, float b) { return a < b ? a : b; }

float aki(const Tables &t, float OM, float BETA)   // wam_general_module.f90:301-364
{
    const float G = t.k.G;
    float a = fmaxr((OM * OM) / (4.f * G), OM / (2.f * sqrtf(G * BETA)));
    float akp = 10000.f;
    while (wam_fabsf(akp - a) > 0.0001f * a) {
        float bo = BETA * a;
        if (bo > t.k.DKMAX) { a = (OM * OM) / G; break; }
        akp = a;
        float th = (G * a) * wam_tanhf(bo);
        float sth = sqrtf(th);
        float ch = wam_coshf(bo);
        a = a + (((OM - sth) * sth) * 2.f) / (th / a + (G * bo) / (ch * ch));
    }
    return a;
}

static void set_consts(Tables &t)
{
    Consts &c = t.k;
    c.G = 9.806f; c.PI = 3.1415927f; c.ZPI = 2.f * c.PI; c.DEG = 180.f / c.PI; c.RAD = c.PI / 180.f;
    c.CIRC = 40000000.f; c.RADIUS = c.CIRC / c.ZPI; c.ROAIR = 1.225f; c.RNUAIR = 1.5E-5f;
    c.RNUAIRM = 0.11f * c.RNUAIR; c.ROWATER = 1000.f; c.XKAPPA = 0.40f; c.XNLEV = 10.0f; c.DKMAX = 40.f;
    c.BETAMAX = (float)t.cfg.betamax; c.ZALP = (float)t.cfg.zalp; c.ALPHA = (float)t.cfg.alpha;
    c.CDFAC = (float)t.cfg.cdfac;
    if (t.ML <= 30) c.ALPHA = 0.0075f;    // initmdl.f90:160
    c.EMIN = 1.0E-12f; c.EPS1 = 0.00001f; c.CO = 1.1f;
}

// SET_FRE_DIR + MAKE_FRE_DIR, wam_fre_dir_module.f90:246-251,297-461
static void fre_dir(Tables &t)
{
    const int KL = t.KL, ML = t.ML;
    const Consts &c = t.k;
    const float CO = c.CO;
    auto z = [&](std::vector<float> &v, int n) { v.assign(n, 0.f); };
    z(t.FR, ML); z(t.DF, ML); z(t.DFIM, ML); z(t.DFIMOFR, ML); z(t.DFIM_FR, ML); z(t.DFIM_FR2, ML); z(t.FR5, ML);
    z(t.FRM5, ML); z(t.RHOWG_DFIM, ML); z(t.GOM, ML); z(t.TH, KL); z(t.COSTH, KL); z(t.SINTH, KL);
    int n = 1 - t.cfg.iref;
    float cof = n >= 0 ? wam_ipowf(CO, n) : 1.f / wam_ipowf(CO, -n);
    t.FR[0] = cof * (float)t.cfg.fr1;
    t.INV_LOG_CO = 1.f / wam_log10f(CO);
    for (int m = 1; m < ML; m++) t.FR[m] = CO * t.FR[m - 1];
    for (int m = 0; m < ML; m++) {
        t.FR5[m] = wam_ipowf(t.FR[m], 5);
        t.FRM5[m] = 1.f / t.FR5[m];
        t.GOM[m] = c.G / ((2.f * c.ZPI) * t.FR[m]);
    }
    t.DELTH = c.ZPI / (float)KL;
    t.DELTR = t.DELTH * c.RADIUS;
    for (int k = 0; k < KL; k++) {
        t.TH[k] = (float)k * t.DELTH + 0.5f * t.DELTH;
        t.COSTH[k] = wam_cosf(t.TH[k]);
        t.SINTH[k] = wam_sinf(t.TH[k]);
    }
    const float CO1 = 0.5f * (CO - 1.f);
    t.DF[0] = CO1 * t.FR[0];
    for (int m = 1; m < ML - 1; m++) t.DF[m] = CO1 * (t.FR[m] + t.FR[m - 1]);
    t.DF[ML - 1] = CO1 * t.FR[ML - 2];
    const float lco = wam_logf(CO);
    for (int m = 0; m < ML; m++) {
        float dffr = t.DF[m] * t.FR[m];
        t.DFIM[m] = t.DF[m] * t.DELTH;
        t.DFIMOFR[m] = t.DFIM[m] / t.FR[m];
        t.DFIM_FR[m] = dffr * t.DELTH;
        t.DFIM_FR2[m] = (dffr * t.FR[m]) * t.DELTH;
        t.RHOWG_DFIM[m] = (((c.ROWATER * c.G) * t.DELTH) * lco) * t.FR[m];
    }
    t.RHOWG_DFIM[0] = 0.5f * t.RHOWG_DFIM[0];
    t.RHOWG_DFIM[ML - 1] = 0.5f * t.RHOWG_DFIM[ML - 1];
    const float frl = t.FR[ML - 1];
    t.MO_TAIL = -((t.DELTH / -4.f) * frl);
    t.MM1_TAIL = -(t.DELTH / -5.f);
    t.MP1_TAIL = -((t.DELTH / -3.f) * wam_ipowf(frl, 2));
    t.MP2_TAIL = -((t.DELTH / -2.f) * wam_ipowf(frl, 3));
    t.MPM_m.resize(ML); t.MPM_p.resize(ML);
    for (int m = 1; m <= ML; m++) { t.MPM_m[m - 1] = m > 1 ? m - 1 : 1; t.MPM_p[m - 1] = m < ML ? m + 1 : ML; }
    t.KPM_m.resize(KL); t.KPM_p.resize(KL); t.JXO1.resize(KL); t.JXO2.resize(KL); t.JYO1.resize(KL); t.JYO2.resize(KL);
    for (int k = 1; k <= KL; k++) {
        t.KPM_m[k - 1] = k > 1 ? k - 1 : KL;
        t.KPM_p[k - 1] = k < KL ? k + 1 : 1;
        bool north = t.COSTH[k - 1] >= 0.f, east = t.SINTH[k - 1] >= 0.f;
        t.JYO1[k - 1] = north ? 1 : 2; t.JYO2[k - 1] = north ? 2 : 1;
        t.JXO1[k - 1] = east ? 1 : 2;  t.JXO2[k - 1] = east ? 2 : 1;
    }
}

// MIN_ENERGY + FETCH_LAW + JONSWAP, wam_tables_module.f90:315-385, wam_jonswap_module.f90:63-213
static void min_energy(Tables &t)
{
    const int JU = t.JUMAX, ML = t.ML;
    const float G = t.k.G;
    t.DELU = 75.f / (float)JU;
    const float FETCH = fminr(0.5f * t.DELPHI, 25000.f);
    const float FPMAX = t.FR[ML - 1];
    const float A = 2.84f, D = -(3.f / 10.f), B = 0.033f, E = 2.f / 3.f;
    std::vector<float> fp(JU), al(JU);
    for (int j = 0; j < JU; j++) {
        float u = (float)(j + 1) * t.DELU;
        if (u > 0.1E-08f) {
            float ug = G / u;
            float f = fmaxr(0.13f, A * wam_powf((G * FETCH) / (u * u), D));
            f = fminr(f, FPMAX / ug);
            al[j] = fmaxr(0.0081f, B * wam_powf(f, E));
            fp[j] = f * ug;
        } else { al[j] = 0.0081f; fp[j] = FPMAX; }
        al[j] = 0.01f * al[j];
    }
    t.FLMINFR.assign((size_t)JU * ML, 0.f);
    const float gz4 = (G * G) / fold_pow(t.k.ZPI, 4);
    const float lgam = wam_logf(3.0f);
    const float saa = 0.5f / (0.07f * 0.07f), sbb = 0.5f / (0.09f * 0.09f);
    for (int m = 0; m < ML; m++) {
        const float f = t.FR[m];
        const float g2 = gz4 * (1.f / wam_ipowf(f, 5));
        for (int j = 0; j < JU; j++) {
            float et = 0.f;
            float arg = 1.25f * wam_ipowf(fp[j] / f, 4);
            if (arg < 50.f) {
                et = (al[j] * g2) * wam_expf(-arg);
                float r = f / fp[j] - 1.f;
                arg = (f > fp[j] ? sbb : saa) * wam_ipowf(r, 2);
                if (arg < 99.f) et = et * wam_expf(lgam * wam_expf(-arg));
            }
            t.FLMINFR[(size_t)m * JU + j] = fmaxr(et, 0.000001f);
        }
    }
}

// MAKE_SHALLOW_TABLES, wam_tables_module.f90:388-466
static void shallow_tables(Tables &t)
{
    const int ND = t.cfg.ndepth, ML = t.ML;
    const float G = t.k.G, GH = G / (4.f * t.k.PI);
    const float da = (float)t.cfg.deptha, dd = (float)t.cfg.depthd;
    t.TFAK.assign((size_t)ND * ML, 0.f); t.TCGOND.assign((size_t)ND * ML, 0.f); t.TSIHKD.assign((size_t)ND * ML, 0.f);
    for (int m = 0; m < ML; m++) {
        const float om = t.k.ZPI * t.FR[m];
        for (int jd = 0; jd < ND; jd++) {
            size_t i = (size_t)m * ND + jd;
            float ad = da * wam_ipowf(dd, jd);
            float ak = aki(t, om, ad);
            float akd = ak * ad;
            t.TFAK[i] = ak;
            if (akd <= 10.0f) {
                t.TCGOND[i] = (0.5f * sqrtf((G * wam_tanhf(akd)) / ak)) * (1.0f + (2.0f * akd) / wam_sinhf(2.f * akd));
                t.TSIHKD[i] = om / wam_sinhf(2.f * akd);
            } else {
                t.TCGOND[i] = GH / t.FR[m];
                t.TSIHKD[i] = 0.f;
            }
        }
    }
}

static int jafu(float cl, int j, int ian)   // src/chief/jafu.f90
{
    int r = j + wam_int(cl);
    if (r <= 0) r += ian;
    if (r > ian) r -= ian;
    return r;
}

// NLWEIGT (wam_source_module.f90:1056-1287) and INIT_SNONLIN (:1291-1523)
static void dia_tables(Tables &t)
{
    const int KL = t.KL, ML = t.ML;
    const float CO = t.k.CO, AL = 0.25f, CON = 3000.f;
    const float f1p1 = wam_log10f(CO);
    t.ISP = wam_int(wam_log10f(1.f + AL) / f1p1 + .000001f);
    t.ISM = (int)wam_floorf(wam_log10f(1.f - AL) / f1p1 + .0000001f);
    t.MFRSTLW = 1 + t.ISM;
    t.MLSTHG = ML - t.ISM;
    t.KFRH = -t.ISM + t.ISP + 2;
    const int MF = t.MFRSTLW, MH = t.MLSTHG, NM = MH - MF + 1;

    // angular part
    const float xf = fold_pow((1.f + AL) / (1.f - AL), 4);
    const float costh3 = ((1.f + 2.f * AL) + 2.f * fold_pow(AL, 3)) / fold_pow(1.f + AL, 2);
    const float dphi1 = (-(180.f / t.k.PI)) * facos(costh3);
    const float costh4 = sqrtf((1.f - xf) + xf * wam_ipowf(costh3, 2));
    const float dphi2 = (180.f / t.k.PI) * facos(costh4);
    const float deltha = t.DELTH * t.k.DEG;
    float cl1 = dphi1 / deltha, cl2 = dphi2 / deltha;
    std::vector<int> ja1(2 * KL), ja2(2 * KL);
    for (int kh = 1, ic = 1; kh <= 2; kh++, ic = -1) {
        const int klh = kh == 2 ? KL + 1 : KL;
        for (int k = 1; k <= klh; k++) {
            int ks = kh > 1 ? KL + 2 - k : k;
            if (ks > KL) continue;
            ja1[(kh - 1) * KL + ks - 1] = jafu((float)ic * cl1, k, KL);
            ja2[(kh - 1) * KL + ks - 1] = jafu((float)ic * cl2, k, KL);
        }
    }
    cl1 = cl1 - (float)wam_int(cl1);
    cl2 = cl2 - (float)wam_int(cl2);
    t.ACL1 = wam_fabsf(cl1); t.ACL2 = wam_fabsf(cl2);
    t.CL11 = 1.f - t.ACL1; t.CL21 = 1.f - t.ACL2;
    t.DAL1 = 1.f / fold_pow(1.f + AL, 4);
    t.DAL2 = 1.f / fold_pow(1.f - AL, 4);
    t.K1W.assign(2 * KL, 0); t.K2W.assign(2 * KL, 0); t.K11W.assign(2 * KL, 0); t.K21W.assign(2 * KL, 0);
    for (int kh = 1, isg = 1; kh <= 2; kh++, isg = -1) {
        const float c1h = (float)isg * cl1, c2h = (float)isg * cl2;
        for (int k = 1; k <= KL; k++) {
            int ks = (k == 1) ? 1 : (kh == 2 ? KL - k + 2 : k);
            int o = (kh - 1) * KL + ks - 1;
            int k1 = ja1[(kh - 1) * KL + k - 1], k2 = ja2[(kh - 1) * KL + k - 1];
            int k11 = c1h < 0.f ? (k1 - 1 < 1 ? KL : k1 - 1) : (k1 + 1 > KL ? 1 : k1 + 1);
            int k21 = c2h < 0.f ? (k2 - 1 < 1 ? KL : k2 - 1) : (k2 + 1 > KL ? 1 : k2 + 1);
            t.K1W[o] = k1; t.K11W[o] = k11; t.K2W[o] = k2; t.K21W[o] = k21;
        }
    }
    // frequency part on the extended grid FRLON(MFRSTLW:ML+KFRH)
    std::vector<float> frlon(ML + t.KFRH - MF + 1);
    auto FL = [&](int m) -> float & { return frlon[m - MF]; };
    for (int m = 1; m <= ML; m++) FL(m) = t.FR[m - 1];
    for (int m = 0; m >= MF; m--) FL(m) = FL(m + 1) / CO;
    for (int m = ML + 1; m <= ML + t.KFRH; m++) FL(m) = CO * FL(m - 1);
    t.IKP.assign(NM, 0); t.IKP1.assign(NM, 0); t.IKM.assign(NM, 0); t.IKM1.assign(NM, 0);
    t.AF11.assign(NM, 0.f); t.FKLAP.assign(NM, 0.f); t.FKLAP1.assign(NM, 0.f);
    t.FKLAM.assign(NM, 0.f); t.FKLAM1.assign(NM, 0.f);
    for (int m = MF; m <= MH; m++) {
        const int i = m - MF;
        const float frg = FL(m);
        t.AF11[i] = CON * wam_ipowf(frg, 11);
        const float flp = frg * (1.f + AL), flm = frg * (1.f - AL);
        t.IKP[i] = m + t.ISP;
        t.IKP1[i] = t.IKP[i] + 1;
        const float fkp = FL(t.IKP[i]);
        t.FKLAP[i] = (flp - fkp) / (FL(t.IKP1[i]) - fkp);
        t.FKLAP1[i] = 1.f - t.FKLAP[i];
        if (FL(MF) >= flm) {
            t.IKM[i] = 1; t.IKM1[i] = 1; t.FKLAM[i] = 0.f; t.FKLAM1[i] = 0.f;
        } else {
            const int ikn = m + t.ISM;
            t.IKM[i] = ikn; t.IKM1[i] = ikn + 1;
            const float fkm = FL(ikn);
            t.FKLAM[i] = (flm - fkm) / (FL(ikn + 1) - fkm);
            t.FKLAM1[i] = 1.f - t.FKLAM[i];
            if (ikn < MF) { t.IKM[i] = 1; t.FKLAM1[i] = 0.f; }
        }
    }
    t.FRH.assign(t.KFRH, 0.f);
    for (int i = 1; i <= t.KFRH; i++) t.FRH[i - 1] = wam_ipowf(FL(ML) / FL(ML + i - 1), 5);

    // INIT_SNONLIN
    auto epmma = [](float x) { return wam_expf(-fminr(1.25f * wam_ipowf(x, 4), 50.f)) * wam_ipowf(x, 5); };
    t.FTRF.assign(2 - MF, 0.f);
    {
        const float alph = 1.f / epmma(1.f);
        float frr = 1.f;
        for (int mc = 1; mc >= MF; mc--) { t.FTRF[mc - MF] = alph * epmma(frr); frr = frr * CO; }
    }
    t.INLCOEF.assign(5 * MH, 0);
    t.RNLCOEF.assign(25 * MH, 0.f);
    for (int mc = 1; mc <= MH; mc++) {
        const int i = mc - MF;
        float ffp = 1.f, ffp1 = 1.f, ffm = 1.f, ffm1 = 1.f, ftail = 1.f;
        int ic = mc, ip = t.IKP[i], ip1 = t.IKP1[i], im = t.IKM[i], im1 = t.IKM1[i];
        if (im < 1) {
            ffm = t.FTRF[im - MF]; im = 1;
            if (im1 < 1) { ffm1 = t.FTRF[im1 - MF]; im1 = 1; }
        }
        if (ip1 > ML) {
            int it = ip1 - ML + 1;
            if (it > t.KFRH) it = t.KFRH;
            ffp1 = t.FRH[it - 1]; ip1 = ML;
            if (ip > ML) {
                ffp = t.FRH[ip - ML]; ip = ML;
                if (ic > ML) {
                    ftail = t.FRH[ic - ML]; ic = ML;
                    if (im1 > ML) { ffm1 = t.FRH[im1 - ML]; im1 = ML; }
                }
            }
        }
        int *ib = &t.INLCOEF[(mc - 1) * 5];
        ib[0] = ic; ib[1] = ip; ib[2] = ip1; ib[3] = im; ib[4] = im1;
        float *r = &t.RNLCOEF[(mc - 1) * 25];
        float p = t.FKLAP[i], p1 = t.FKLAP1[i];
        float gw2 = (p1 * ffp) * t.DAL1, gw1 = gw2 * t.CL11; gw2 = gw2 * t.ACL1;
        float gw4 = (p * ffp1) * t.DAL1, gw3 = gw4 * t.CL11; gw4 = gw4 * t.ACL1;
        float pa = p * t.CL11, pb = p * t.ACL1, p2 = p1 * t.ACL1; p1 = p1 * t.CL11;
        r[0] = ftail; r[1] = gw1; r[2] = gw2; r[3] = gw3; r[4] = gw4;
        r[5] = pa; r[6] = pb; r[7] = p2; r[8] = p1;
        r[9] = pa * pa; r[10] = pb * pb; r[11] = p1 * p1; r[12] = p2 * p2;
        float q = t.FKLAM[i], q1 = t.FKLAM1[i];
        float gw6 = (q1 * ffm) * t.DAL2, gw5 = gw6 * t.CL21; gw6 = gw6 * t.ACL2;
        float gw8 = (q * ffm1) * t.DAL2, gw7 = gw8 * t.CL21; gw8 = gw8 * t.ACL2;
        float qa = q * t.CL21, qb = q * t.ACL2, q2 = q1 * t.ACL2; q1 = q1 * t.CL21;
        r[13] = gw5; r[14] = gw6; r[15] = gw7; r[16] = gw8;
        r[17] = qa; r[18] = qb; r[19] = q2; r[20] = q1;
        r[21] = qa * qa; r[22] = qb * qb; r[23] = q1 * q1; r[24] = q2 * q2;
    }
}

// Ordered contribution lists for the gather form of SNONLIN's scatter
// (wam_source_module.f90:2890-3043): the statements of each regime are
// replayed in program order and every update is appended to its target
// bin's list, so a bin accumulates exactly in the reference's order.
// Entry = q | type << 16 with q = ((MC-1)*2 + KH)*KL + K (0-based KH, K).
static int snonlin_lists(Tables &t, std::string &err)
{
    const int KL = t.KL, ML = t.ML, MF = t.MFRSTLW, NB = KL * ML;
    const int m1st = -MF + 1, mlst = ML - t.KFRH + m1st;
    std::vector<std::vector<int32_t>> lists(NB);
    bool bad = false;
    t.NQ = 2 * KL * t.MLSTHG;
    for (int mc = 1; mc <= t.MLSTHG; mc++) {
        const int i = mc - MF;
        const int mp = t.IKP[i], mp1 = t.IKP1[i], mm = t.IKM[i], mm1 = t.IKM1[i];
        const int regime = (mc > m1st && mc < mlst) ? 0 : (mc >= mlst ? 1 : 2);
        for (int kh = 0; kh < 2; kh++)
            for (int k = 0; k < KL; k++) {
                const int q = ((mc - 1) * 2 + kh) * KL + k;
                const int k1 = t.K1W[kh * KL + k] - 1, k2 = t.K2W[kh * KL + k] - 1;
                const int k11 = t.K11W[kh * KL + k] - 1, k21 = t.K21W[kh * KL + k] - 1;
                auto add = [&](int type, int kk, int m) {
                    if (m < 1 || m > ML) { bad = true; return; }
                    lists[kk + KL * (m - 1)].push_back(q | (type << 16));
                };
                auto mmset = [&]() { add(1, k2, mm); add(2, k21, mm); };
                auto mm1set = [&]() { add(3, k2, mm1); add(4, k21, mm1); };
                auto mpset = [&]() { add(5, k1, mp); add(6, k11, mp); };
                auto mp1set = [&]() { add(7, k1, mp1); add(8, k11, mp1); };
                if (regime == 0) {
                    add(0, k, mc); mmset(); mm1set(); mpset(); mp1set();
                } else if (regime == 1) {
                    mmset();
                    if (mm1 <= ML) {
                        mm1set();
                        if (mc <= ML) {
                            add(0, k, mc);
                            if (mp <= ML) { mpset(); if (mp1 <= ML) mp1set(); }
                        }
                    }
                } else {
                    if (mm1 >= 1) mm1set();
                    add(0, k, mc); mpset(); mp1set();
                }
            }
    }
    if (bad) { err = "SNONLIN target frequency outside 1..ML"; return WAM_E_ARG; }
    int nc = 0;
    for (auto &l : lists) nc = l.size() > (size_t)nc ? (int)l.size() : nc;
    t.NCMAX = nc;
    t.CONTRIB.assign((size_t)NB * nc, 0);
    t.NCONTRIB.assign(NB, 0);
    for (int b = 0; b < NB; b++) {
        t.NCONTRIB[b] = (int)lists[b].size();
        for (size_t j = 0; j < lists[b].size(); j++) t.CONTRIB[(size_t)b * nc + j] = lists[b][j];
    }
    // Chunks of at most 512 quadruplets (two passes of a 256-thread
    // workgroup).  The update order inside each bin's list is the
    // reference's, and every list is ordered by MC, so a chunked gather adds
    // in the same order.
    t.CHUNK = std::max(1, std::min(14, 511 / (2 * KL)));
    t.NCHUNK = (t.MLSTHG + t.CHUNK - 1) / t.CHUNK;
    if (2 * KL > 511) { err = "SNONLIN packing needs KL <= 255"; return WAM_E_ARG; }
    t.SN_E.clear();
    t.SN_WW.clear();
    t.SN_CH.assign((size_t)5 * t.NCHUNK, 0);
    for (int b = 0; b < NB; b++) {
        int prev_mc = 0;
        for (size_t j = 0; j < lists[b].size(); j++) {
            const int mc = (lists[b][j] & 0xffff) / (2 * KL) + 1;
            if (mc < prev_mc) { err = "SNONLIN list not ordered by frequency"; return WAM_E_ARG; }
            prev_mc = mc;
        }
    }
    for (int c = 0; c < t.NCHUNK; c++) {
        const int mc_lo = 1 + c * t.CHUNK, mc_hi = std::min(t.MLSTHG, mc_lo + t.CHUNK - 1);
        std::vector<std::vector<uint16_t>> per(NB);
        int blo = NB, bhi = 0, w = 0;
        for (int b = 0; b < NB; b++) {
            for (int e : lists[b]) {
                const int q = e & 0xffff, ty = e >> 16, mc = q / (2 * KL) + 1;
                if (mc < mc_lo || mc > mc_hi) continue;
                const int ql = q - (mc_lo - 1) * 2 * KL;            // quadruplet index within the chunk
                per[b].push_back((uint16_t)(ql | (((mc - mc_lo) * 9 + ty) << 9)));
            }
            if (!per[b].empty()) { blo = std::min(blo, b); bhi = b + 1; w = std::max(w, (int)per[b].size()); }
        }
        if (bhi <= blo) { blo = 0; bhi = 0; w = 0; }
        w = (w + SN_GROUP - 1) / SN_GROUP * SN_GROUP;                // whole groups of entries
        const int nbc = bhi - blo, off = (int)t.SN_E.size(), wwoff = (int)t.SN_WW.size();
        t.SN_CH[5 * c] = blo; t.SN_CH[5 * c + 1] = bhi; t.SN_CH[5 * c + 2] = w; t.SN_CH[5 * c + 3] = off;
        t.SN_CH[5 * c + 4] = wwoff;
        t.SN_E.resize((size_t)off + (size_t)w * nbc, SN_NONE);
        for (int b = blo; b < bhi; b++)
            for (size_t j = 0; j < per[b].size(); j++) t.SN_E[(size_t)off + j * nbc + (b - blo)] = per[b][j];
        // list length each run of 64 consecutive bins (one wavefront) needs
        for (int g0 = blo; g0 < bhi; g0 += 64) {
            int wg = 0;
            for (int b = g0; b < std::min(bhi, g0 + 64); b++) wg = std::max(wg, (int)per[b].size());
            t.SN_WW.push_back((wg + SN_GROUP - 1) / SN_GROUP * SN_GROUP);
        }
    }
    // products form: chunks of at most 255 quadruplets, per run of 64 bins
    t.SP_CHUNK = std::max(1, (256 * SP_QPT - 1) / (2 * KL));
    t.SP_NCHUNK = (t.MLSTHG + t.SP_CHUNK - 1) / t.SP_CHUNK;
    t.SP_NRUN = (NB + 63) / 64;
    if (2 * KL > 256 * SP_QPT - 1) { err = "SNONLIN products form needs more quadruplet slots"; return WAM_E_ARG; }
    t.SP_E.clear();
    t.SP_RW.assign((size_t)2 * t.SP_NCHUNK * t.SP_NRUN, 0);
    for (int c = 0; c < t.SP_NCHUNK; c++) {
        const int mc_lo = 1 + c * t.SP_CHUNK, mc_hi = std::min(t.MLSTHG, mc_lo + t.SP_CHUNK - 1);
        for (int r = 0; r < t.SP_NRUN; r++) {
            std::vector<std::vector<uint16_t>> per(64);
            size_t w = 0;
            for (int l = 0; l < 64 && 64 * r + l < NB; l++) {
                for (int e : lists[64 * r + l]) {
                    const int q = e & 0xffff, ty = e >> 16, mc = q / (2 * KL) + 1;
                    if (mc < mc_lo || mc > mc_hi) continue;
                    per[l].push_back((uint16_t)((ty * SP_TSTRIDE + (q - (mc_lo - 1) * 2 * KL)) * 8));
                }
                w = std::max(w, per[l].size());
            }
            w = (w + 2 * SP_GROUP - 1) / (2 * SP_GROUP) * (2 * SP_GROUP);   // whole groups of pairs
            const size_t off = t.SP_E.size();
            t.SP_RW[2 * ((size_t)c * t.SP_NRUN + r)] = (int32_t)off;
            t.SP_RW[2 * ((size_t)c * t.SP_NRUN + r) + 1] = (int32_t)w;
            t.SP_E.resize(off + w * 64, SP_NONE);
            // pairs of entries lane-interleaved: entries 2i, 2i+1 of lane l at (i*64 + l)*2 + {0, 1}
            for (int l = 0; l < 64; l++)
                for (size_t i = 0; i < per[l].size(); i++) t.SP_E[off + ((i / 2) * 64 + l) * 2 + (i % 2)] = per[l][i];
        }
    }
    // one group of padding pairs for the runs with no entries
    t.SP_PAD = (int)t.SP_E.size();
    t.SP_E.resize(t.SP_E.size() + (size_t)SP_GROUP * 64 * 2, SP_NONE);
    for (size_t i = 0; i < t.SP_RW.size(); i += 2)
        if (t.SP_RW[i + 1] == 0) t.SP_RW[i] = t.SP_PAD;
    return WAM_OK;
}

// Tables of SNONLIN's adjoint (the transpose of the DIA quadruplet terms):
// the rows each quadruplet updates (the regimes of :2890-3043, as in
// snonlin_lists), the quadruplets reading each frequency row, and the
// inverses of the direction maps.  A map that is not a permutation of the
// directions leaves adj_err set (the adjoint then refuses to run).
static void snonlin_adjoint_tables(Tables &t)
{
    const int KL = t.KL, ML = t.ML, MF = t.MFRSTLW;
    const int m1st = -MF + 1, mlst = ML - t.KFRH + m1st;
    t.adj_err.clear();
    t.AQ_ROW.assign((size_t)9 * t.MLSTHG, -1);
    for (int mc = 1; mc <= t.MLSTHG; mc++) {
        const int i = mc - MF;
        const int mp = t.IKP[i], mp1 = t.IKP1[i], mm = t.IKM[i], mm1 = t.IKM1[i];
        const int regime = (mc > m1st && mc < mlst) ? 0 : (mc >= mlst ? 1 : 2);
        int8_t *r = &t.AQ_ROW[(size_t)9 * (mc - 1)];
        auto set = [&](int ty, int m) { r[ty] = (int8_t)(m - 1); };
        auto pair = [&](int ty, int m) { set(ty, m); set(ty + 1, m); };
        if (regime == 0) {
            set(0, mc); pair(1, mm); pair(3, mm1); pair(5, mp); pair(7, mp1);
        } else if (regime == 1) {
            pair(1, mm);
            if (mm1 <= ML) {
                pair(3, mm1);
                if (mc <= ML) {
                    set(0, mc);
                    if (mp <= ML) { pair(5, mp); if (mp1 <= ML) pair(7, mp1); }
                }
            }
        } else {
            if (mm1 >= 1) pair(3, mm1);
            set(0, mc); pair(5, mp); pair(7, mp1);
        }
    }
    t.ACH = std::max(1, 512 / (2 * KL));
    t.ANCH = (t.MLSTHG + t.ACH - 1) / t.ACH;
    static const int rslot[9] = {0, 1, 1, 2, 2, 3, 3, 4, 4};          // INLCOEF column of the read row
    static const int wslot[9] = {0, 1, 2, 3, 4, 13, 14, 15, 16};      // RNLCOEF slot of the read weight
    t.AR_E.clear();
    t.AR_OFF.assign((size_t)t.ANCH * ML + 1, 0);
    for (int c = 0; c < t.ANCH; c++) {
        const int mc_lo = 1 + c * t.ACH, mc_hi = std::min(t.MLSTHG, mc_lo + t.ACH - 1);
        for (int m = 0; m < ML; m++) {
            t.AR_OFF[(size_t)c * ML + m] = (int32_t)t.AR_E.size();
            for (int mc = mc_lo; mc <= mc_hi; mc++)
                for (int sl = 0; sl < 9; sl++) {
                    const int row = t.INLCOEF[(size_t)(mc - 1) * 5 + rslot[sl]] - 1;
                    if (row != m || t.RNLCOEF[(size_t)(mc - 1) * 25 + wslot[sl]] == 0.f) continue;
                    t.AR_E.push_back((uint16_t)((mc - 1) << 4 | sl));
                }
        }
    }
    t.AR_OFF[(size_t)t.ANCH * ML] = (int32_t)t.AR_E.size();
    t.AKINV.assign((size_t)8 * KL, -1);
    const std::vector<int> *maps[4] = {&t.K1W, &t.K11W, &t.K2W, &t.K21W};
    for (int mp = 0; mp < 4; mp++)
        for (int kh = 0; kh < 2; kh++)
            for (int k = 0; k < KL; k++) {
                const int to = (*maps[mp])[(size_t)kh * KL + k] - 1;
                int32_t &inv = t.AKINV[((size_t)mp * 2 + kh) * KL + to];
                if (to < 0 || to >= KL || inv >= 0) t.adj_err = "a DIA direction map is not a permutation";
                else inv = k;
            }
    for (int i = 0; i < t.MLSTHG * 5; i++)
        if (t.INLCOEF[i] < 1 || t.INLCOEF[i] > ML) t.adj_err = "a DIA read row outside 1..ML";
}

// INIT_X0TAUHF, wam_source_module.f90:3784-3849
static void x0tauhf(Tables &t)
{
    const Consts &c = t.k;
    float x0 = 0.005f;
    for (int j = 1; j <= 30; j++) {
        float ff = wam_expf(c.XKAPPA / (x0 + c.ZALP));
        float f = (c.ALPHA * (x0 * x0)) * ff - 1.0f;
        if (f == 0.0f) break;
        float df = (c.ALPHA * ff) * (2.0f * x0 - c.XKAPPA * wam_ipowf(x0 / (x0 + c.ZALP), 2));
        x0 = x0 - f / df;
    }
    t.X0TAUHF = x0;
    const float c1 = (c.BETAMAX / (c.XKAPPA * c.XKAPPA)) / 3.0f;
    t.WTAUHF[0] = c1;
    for (int j = 2; j <= 18; j += 2) { t.WTAUHF[j - 1] = 4.0f * c1; t.WTAUHF[j] = 2.0f * c1; }
    t.WTAUHF[18] = c1;
}

// LFACTOR's 10-Hz extension grid, wam_source_module.f90:4585-4605 (depends on FR(1) only)
static void lfactor_grid(Tables &t)
{
    const float ZPI = t.k.ZPI, CO = t.k.CO;
    const float sig1 = ZPI * t.FR[0];
    int nk10 = (int)wam_ceilf(wam_logf(10.f / (sig1 / ZPI)) / wam_logf(CO)) + 1;
    if (nk10 < t.ML) nk10 = t.ML;
    t.NK10 = nk10;
    t.SIG10.assign(nk10, 0.f); t.DSII10.assign(nk10, 0.f);
    if (t.ML < nk10) {
        for (int i = 0; i < nk10; i++) t.SIG10[i] = sig1 * wam_powf(CO, (float)(i + 1) - 1.0f);
        for (int i = 0; i < nk10; i++) t.DSII10[i] = (0.5f * t.SIG10[i]) * (CO - 1.0f / CO);
        t.DSII10[0] = (0.5f * t.SIG10[0]) * (CO - 1.0f);
        t.DSII10[nk10 - 1] = ((0.5f * t.SIG10[nk10 - 1]) * (CO - 1.0f)) / CO;
    } else {
        for (int i = 0; i < nk10; i++) { t.SIG10[i] = ZPI * t.FR[i]; t.DSII10[i] = ZPI * t.DF[i]; }
    }
}

// PREPARE_GRID (regular grid) + DCO/DPSN (wam_initial_module.f90:205-214).
// Regular grid: CLOSEST_GP_LAT reduces to the same longitude index on the
// neighbouring row with weight 1 (preproc_module.f90:1396-1399).
static int grid(Tables &t, const uint8_t *mask, std::string &err)
{
    const wam_config &g = t.cfg;
    const int NX = g.nx, NY = g.ny;
    if (NX < 2 || NY < 1) { err = "grid needs nx >= 2, ny >= 1"; return WAM_E_GRID; }
    t.SINPH.resize(NY); t.COSPH.resize(NY); t.DELLAM.resize(NY);
    for (int k = 0; k < NY; k++) {
        float xlat = ((float)(g.amosop + (int64_t)k * g.xdella) / (float)M_DEGREE) * t.k.RAD;
        t.SINPH[k] = wam_sinf(xlat); t.COSPH[k] = wam_cosf(xlat);
    }
    t.DELPHI = (((float)g.xdella / (float)M_DEGREE) * t.k.CIRC) / 360.f;
    int zd = g.iper ? wam_nint((float)(g.amoeap - g.amowep + g.xdello) / (float)NX)
                    : wam_nint((float)(g.amoeap - g.amowep) / (float)(NX - 1));
    for (int k = 0; k < NY; k++) t.DELLAM[k] = (((float)zd / (float)M_DEGREE) * t.k.CIRC) / 360.f;
    t.ZDELLO = zd;
    // sea-point numbering and a cell -> sea point map
    std::vector<int> num((size_t)NX * NY, 0);
    int ns = 0;
    for (size_t c = 0; c < (size_t)NX * NY; c++) if (mask[c]) num[c] = ++ns;
    if (ns == 0) { err = "sea mask has no sea point"; return WAM_E_GRID; }
    t.NSEA = ns;
    t.IXLG.resize(ns); t.KXLT.resize(ns);
    t.KLAT_S.assign(ns, 0); t.KLAT_N.assign(ns, 0); t.KLON_W.assign(ns, 0); t.KLON_E.assign(ns, 0);
    for (int k = 1; k <= NY; k++)
        for (int i = 1; i <= NX; i++) {
            int ip = num[(size_t)(k - 1) * NX + i - 1];
            if (!ip) continue;
            t.IXLG[ip - 1] = i; t.KXLT[ip - 1] = k;
            auto at = [&](int ii, int kk) { return num[(size_t)(kk - 1) * NX + ii - 1]; };
            if (i > 1) t.KLON_W[ip - 1] = at(i - 1, k);
            else if (g.iper) t.KLON_W[ip - 1] = at(NX, k);
            if (i < NX) t.KLON_E[ip - 1] = at(i + 1, k);
            else if (g.iper) t.KLON_E[ip - 1] = at(1, k);
            if (k > 1) t.KLAT_S[ip - 1] = at(i, k - 1);
            if (k < NY) t.KLAT_N[ip - 1] = at(i, k + 1);
        }
    t.DCO.resize(ns); t.DPSN_S.assign(ns, 1.f); t.DPSN_N.assign(ns, 1.f);
    for (int ij = 0; ij < ns; ij++) t.DCO[ij] = 1.f / t.COSPH[t.KXLT[ij] - 1];
    if (g.spherical)
        for (int ij = 0; ij < ns; ij++) {
            if (t.KLAT_S[ij]) t.DPSN_S[ij] = t.DCO[ij] / t.DCO[t.KLAT_S[ij] - 1];
            if (t.KLAT_N[ij]) t.DPSN_N[ij] = t.DCO[ij] / t.DCO[t.KLAT_N[ij] - 1];
        }
    return WAM_OK;
}

int build_tables(Tables &t, const wam_config &cfg, const uint8_t *mask, std::string &err)
{
    t = Tables();
    t.cfg = cfg;
    t.KL = cfg.kl; t.ML = cfg.ml;
    if (cfg.kl < 4 || cfg.kl > 64 || cfg.ml < 8 || cfg.ml > 60) { err = "need 4 <= KL <= 64 and 8 <= ML <= 60"; return WAM_E_ARG; }
    if (cfg.iphys != 0 && cfg.iphys != 2) { err = "IPHYS must be 0 or 2"; return WAM_E_ARG; }
    if (cfg.isnonlin != 0) { err = "only ISNONLIN = 0 is supported"; return WAM_E_ARG; }
    if (!cfg.spherical) { err = "only spherical grids are supported"; return WAM_E_ARG; }
    if (cfg.idelpro <= 0 || cfg.idelt <= 0 || cfg.ndepth < 1 || cfg.fr1 <= 0) { err = "bad time step, depth table or FR1"; return WAM_E_ARG; }
    set_consts(t);
    fre_dir(t);
    int rc = grid(t, mask, err);
    if (rc) return rc;
    min_energy(t);
    shallow_tables(t);
    dia_tables(t);
    rc = snonlin_lists(t, err);
    if (rc) return rc;
    snonlin_adjoint_tables(t);
    x0tauhf(t);
    lfactor_grid(t);
    if (cfg.iphys == 2 && t.NK10 > 128) { err = "LFACTOR 10-Hz grid longer than 128 frequencies (FR(1) too low)"; return WAM_E_ARG; }
    return WAM_OK;
}

// PREPARE_PROPAGATION per-point part, wam_propagation_module.f90:279-470
// (regular grid).  Land neighbours are index 0 and are skipped as in GRADIENT.
int prepare_points(Tables &t, const float *depth, const float *u, const float *v, std::string &err)
{
    const int ns = t.NSEA, ND = t.cfg.ndepth;
    const float da = (float)t.cfg.deptha, ldd = wam_logf((float)t.cfg.depthd);
    for (int ij = 0; ij < ns; ij++)
        if (!(depth[ij] > 0.f && depth[ij] <= FLT_MAX)) {
            err = "depth must be positive and finite at every sea point";
            return WAM_E_ARG;
        }
    if (t.cfg.refr_c && u && v)
        for (int ij = 0; ij < ns; ij++)
            if (!(std::fabs(u[ij]) <= FLT_MAX && std::fabs(v[ij]) <= FLT_MAX)) {
                err = "currents must be finite at every sea point";
                return WAM_E_ARG;
            }
    t.INDEP.assign(ns, 1);
    if (t.cfg.shallow)
        for (int ij = 0; ij < ns; ij++) {
            int id = wam_nint(wam_logf(depth[ij] / da) / ldd + 1.f);
            t.INDEP[ij] = id < 1 ? 1 : (id > ND ? ND : id);
        }
    const bool refr = t.cfg.refr_d || t.cfg.refr_c;
    auto gradient = [&](const float *f, std::vector<float> &dlat, std::vector<float> &dlon) {
        dlat.assign(ns, 0.f); dlon.assign(ns, 0.f);
        for (int ij = 0; ij < ns; ij++) {
            int n = t.KLAT_N[ij], s = t.KLAT_S[ij];
            if (n && s) dlat[ij] = (f[n - 1] - f[s - 1]) / (2.f * t.DELPHI);
            int e = t.KLON_E[ij], w = t.KLON_W[ij];
            if (e && w) dlon[ij] = (f[e - 1] - f[w - 1]) / (2.f * t.DELLAM[t.KXLT[ij] - 1]);
        }
    };
    t.DDPHI.assign(ns, 0.f); t.DDLAM.assign(ns, 0.f);
    t.DUPHI.assign(ns, 0.f); t.DULAM.assign(ns, 0.f); t.DVPHI.assign(ns, 0.f); t.DVLAM.assign(ns, 0.f);
    t.INDEP_G.assign(ns, 1);
    if (refr && t.cfg.shallow) {
        gradient(depth, t.DDPHI, t.DDLAM);
        for (int ij = 0; ij < ns; ij++) t.DDLAM[ij] = t.DDLAM[ij] * t.DCO[ij];
        for (int ij = 0; ij < ns; ij++) {
            float dm = 4.f * depth[ij];
            int iz = 4;
            int n = t.KLAT_N[ij], s = t.KLAT_S[ij];
            if (n && s) { dm = dm + depth[n - 1] + depth[s - 1]; iz += 2; }
            int e = t.KLON_E[ij], w = t.KLON_W[ij];
            if (e && w) { dm = dm + depth[e - 1] + depth[w - 1]; iz += 2; }
            dm = dm / (float)iz;
            iz = wam_nint(wam_logf(dm / da) / ldd + 1.f);
            t.INDEP_G[ij] = iz < 1 ? 1 : (iz > ND ? ND : iz);
        }
    }
    if (t.cfg.refr_c) {
        if (!u || !v) { err = "current refraction needs u and v"; return WAM_E_ARG; }
        gradient(u, t.DUPHI, t.DULAM);
        gradient(v, t.DVPHI, t.DVLAM);
        for (int ij = 0; ij < ns; ij++) { t.DULAM[ij] = t.DULAM[ij] * t.DCO[ij]; t.DVLAM[ij] = t.DVLAM[ij] * t.DCO[ij]; }
    }
    const float delpro = (float)t.cfg.idelpro;
    const bool deep_nocur = !t.cfg.refr_c && !t.cfg.shallow;
    t.DELLA0.resize(ns); t.DRGM.resize(ns);
    for (int ij = 0; ij < ns; ij++) {
        const int kx = t.KXLT[ij] - 1;
        t.DELLA0[ij] = deep_nocur ? (delpro * t.DCO[ij]) / t.DELLAM[kx]
                                  : ((0.5f * delpro) * t.DCO[ij]) / t.DELLAM[kx];
        t.DRGM[ij] = t.SINPH[kx] * t.DCO[ij];
    }
    return WAM_OK;
}

}  // namespace wam
