// wam_dev.hpp -- device-side views of the WAM tables and kernel entry points.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace wam {

constexpr int SN_GROUP_DEV = 6;   // SNONLIN list entries a thread loads together
constexpr unsigned SN_NONE_DEV = 0x01FF;   // padding entry: zero quadruplet slot 511, weight -2
constexpr int SP_QPT_DEV = 2;              // products form: quadruplets per thread of a chunk
constexpr int SP_TSTRIDE_DEV = 256 * SP_QPT_DEV + 4;           // products form: update slots per type
constexpr unsigned SP_NONE_DEV = (256 * SP_QPT_DEV - 1) * 8;   // products form: byte offset of the (-0, -0) slot
constexpr int SP_GROUP_DEV = 4;            // products form: entry pairs per gather step (lists are whole groups)

// Coefficient planes of the propagation weights, [plane][M][K][NSEA]:
// COEF_SUM, COEF_LAT(.,1,1), COEF_LAT(.,2,1), COEF_LON(.,1), COEF_LON(.,2),
// COEF_THETA(.,1), COEF_THETA(.,2), COEF_SIGD(.,1), COEF_SIGD(.,2)
// (wam_propagation_module.f90:75-110).  Without current refraction only the
// first five are stored, in the order SUM, LAT1, LON1, TH1, TH2.
enum { P_SUM = 0, P_LAT1, P_LAT2, P_LON1, P_LON2, P_TH1, P_TH2, P_SIG1, P_SIG2, P_NPLANES };

// Everything IMPLSCH and PROPAGS read from module globals in the reference,
// as device pointers and scalars (passed by value to every kernel).
struct DevTables {
    int KL, ML, NSEA, NDEPTH, JUMAX, MFRSTLW, MLSTHG, NQ, NCMAX, NK10;
    int iphys, shallow, refr_d, refr_c;
    int brk, phil;                        // WAVE_BREAKING_RUN (SDEPTHLIM, SFBRK), PHILLIPS_RUN (SOURCE_PHILLIPS)
    int split;                            // IMPLSCH in three kernels (1) or the single kernel (0), where both apply
    float G, ZPI, ROAIR, RNUAIRM, ROWATER, XKAPPA, XNLEV, BETAMAX, ZALP, ALPHA, CDFAC, EMIN, EPS1, CO;
    float DELTH, INV_LOG_CO, MO_TAIL, MM1_TAIL, MP1_TAIL, MP2_TAIL, DELU, DAL1, DAL2, X0TAUHF, DELT;
    float DEG;                            // 180/PI
    float ZPI4, CONSD, CONSS, C5G;        // folded constants of STRESSO, SDISSIP and the limiter
    // per frequency / direction
    const float *FR, *DF, *DFIM, *DFIMOFR, *DFIM_FR, *DFIM_FR2, *FR5, *FRM5, *RHOWG_DFIM, *GOM, *TH, *COSTH, *SINTH;
    const float *FLMINFR;                 // [ML][JUMAX]
    const float *TFAK, *TCGOND, *TSIHKD;  // [ML][NDEPTH]
    const float *SIG10, *DSII10;          // [NK10]
    const float *WTAUHF;                  // [19] Simpson weights of TAU_PHI_HF
    // SNONLIN
    const int *INLCOEF;                   // [MLSTHG][5]
    const float *RNLCOEF;                 // [MLSTHG][25]
    const float *AF11;                    // [MLSTHG] (index MC-1)
    const int *K1W, *K2W, *K11W, *K21W;   // [2][KL], 0-based
    // SNONLIN gather lists by chunks of CHUNK frequencies of quadruplets:
    // per chunk the first and last+1 bin it feeds, the list width (a multiple
    // of SN_GROUP_DEV), the offset into SN_E and the offset into SN_WW (list
    // width per run of 64 bins); entries ql | (MC-offset*9 + update type) << 9,
    // padded with SN_NONE_DEV
    int CHUNK, NCHUNK;
    const uint16_t *SN_E;
    const int *SN_CH, *SN_WW;
    // the products form of the lists (entries the byte offsets of the
    // update slots, (type*SP_TSTRIDE_DEV + ql)*8; per run of 64 bins a block of entry
    // pairs at the SP_RW offset, whole groups of SP_GROUP_DEV pairs; runs
    // with no entries point at the padding group SP_PAD)
    int SP_CHUNK, SP_NCHUNK, SP_NRUN, SP_PAD;
    const uint16_t *SP_E;
    const int *SP_RW;
    // SNONLIN's adjoint by chunks of ACH frequencies of quadruplets: update
    // rows per (MC, update type), -1 = skipped; per (chunk, read row) the
    // reading quadruplets (MC - 1) << 4 | read slot; inverse direction maps
    int ACH, ANCH, NAR;             // NAR: entries of AR_E
    const int8_t *AQ_ROW;
    const uint16_t *AR_E;
    const int *AR_OFF, *AKINV;
    // per sea point (PREPARE_PROPAGATION), 0-based neighbour indices, -1 = land
    const int *KLAT_S, *KLAT_N, *KLON_W, *KLON_E;
    const int *INDEP, *INDEP_G;
    const float *DEPTH, *DPSN_S, *DPSN_N, *DELLA0, *DRGM, *U, *V;
    const float *DDPHI, *DDLAM, *DUPHI, *DULAM, *DVPHI, *DVLAM;
    const float *WN, *CGG;                // [ML][NSEA] (IPHYS = 2)
    // SBOTTOM's factor (-2*0.038/G * k)/SINH(MIN(2 k D, 50)) with k = TFAK(INDEP, M)
    // (wam_source_module.f90:1527-1589): [ML][NSEA], shallow water only; it
    // depends on the depth field alone, so PREPARE_PROPAGATION tabulates it
    const float *SBOT;
    const int *JXO1, *JYO1, *JXO2, *JYO2, *KPM_P, *KPM_M;   // 1-based
    float DELTH0, DELPH0, DELTHR, DELFR0;
    // per direction K (cyclic K+1, K-1): SINTH, COSTH, DELTH0*(SINTH(K)+SINTH(K+1)),
    // DELTH0*(SINTH(K)+SINTH(K-1)), SINTH**2, SINTH*COSTH, COSTH**2, 0
    const float *KDIR;
    // grid geometry for the per-point part of PREPARE_PROPAGATION on the
    // device: DCO and KXLT (1-based latitude row) per sea point, DELLAM and
    // SINPH per latitude row, the latitude increment [rad]
    const float *DCO_PT, *DELLAM_ROW, *SINPH_ROW;
    const int *KXLT_PT;
    float DELPHI;
};

// IMPLSCH's arguments for NMP member-points: an ensemble of NMP / NSEA
// members on one grid, member-point mp = member * NSEA + sea point.  fl3 is
// [member][ML][KL][NSEA] (each member a plain FL3), the per-point arrays
// [member][NSEA], fluxes [6][NMP]; a single run is one member.
struct ImplArgs {
    float *fl3;
    const float *u10, *udir, *roairn, *wstar;
    float *tauw, *ustar, *z0, *phiaw, *fluxes;
    int *mij;
    int lcflx;
    int NMP;
};

// The adjoint of IMPLSCH for NMP member-points (laid out as ImplArgs): fl3
// is IMPLSCH's input spectrum (after PROPAGS) and fl3_bar the bar of its
// output; tauw_bar, ustar_bar, z0_bar may be NULL (zero bars).
struct AdjArgs {
    const float *fl3, *u10, *udir, *tauw, *roairn, *wstar;
    const float *fl3_bar, *tauw_bar, *ustar_bar, *z0_bar;
    float *fl3_bar_out, *u10_bar, *udir_bar, *tauw_bar_out;
    int NMP;
};

// Device work arrays of the three-kernel IMPLSCH (IPHYS = 0).  Spectra are
// in point-quad tiles: 64 sea points x 4 consecutive bins of one direction
// row, each point's 4 bins contiguous (16 bytes), so the per-lane kernels
// move 1 KB per wavefront access and the per-point kernel reads a point's
// bins as float4 (KL % 4 == 0).
// Inputs and outputs of prep_points_kernel (the per-point part of
// PREPARE_PROPAGATION); u, v NULL = no current.
struct PrepPoints {
    const float *depth, *u, *v;
    int *indep, *indep_g;
    float *ddphi, *ddlam, *duphi, *dulam, *dvphi, *dvlam, *della0, *drgm;
    unsigned *bad;                 // [0] set where DEPTH is not positive and finite, [1] where U or V is not finite
    float deptha, lndepthd, delpro;
};

struct ImplScratch {
    int NB4;                       // KL*ML/4
    // the chunk of sea points one launch sequence advances: points P0 ..
    // P0 + NP - 1, in rows 0 .. NP - 1 of the work arrays below, whose
    // per-point rows are NX apart (NX >= NP; one chunk of NSEA points with
    // NX = NSEA when the arrays are sized for the whole grid)
    int P0, NP, NX;
    float *ft, *slt, *flt, *smt;   // F, SL, FL, SL - SPOS of the second SINPUT (LCFLX only)
    float *axy;                    // [3][ML][NX]: STRESSO's direction sums of the first SINPUT
    float *pp;                     // [8][NX]: ENHFR, MAX(FMEANWS, FMEAN), MIJ, PHIAW, SIN(UDIR), COS(UDIR), SFBRK's SBR, DSBR
    float *sprd;                   // [KL][NX]: SPRD(K) = MAX(0, COS(TH(K) - UDIR))**2 of the call
    unsigned *llws;                // [(KL*ML+31)/32][NX]: LLWS bits of the second SINPUT
    unsigned *tick;                // [8][32]: point counters of implsch_snl_kernel, one per XCD (own 128-byte line)
    float *st6;                    // IPHYS = 2: [6][ML][NX] per-frequency values of implsch_pre_st6_kernel
};

// Per-point scalars the adjoint's forward recomputation keeps (rows of
// AdjExtra::sc): AIRSEA's values and Jacobians d(UST, Z0)/d(U10, TAUW) of
// both calls, WSIGSTAR's and TAU_PHI_HF's derivatives along (UST, Z0),
// STRESSO's kept values (stresso_close), the mean parameters and their
// integrals, the cut-off indices and the limiter's bars.
enum {
    AS_UST1 = 0, AS_Z01, AS_JA1, AS_SIG1 = AS_JA1 + 4, AS_JS1, AS_JT1 = AS_JS1 + 2, AS_K1 = AS_JT1 + 2,
    AS_SE = AS_K1 + 7, AS_SFM, AS_STM, AS_EMEAN, AS_FMEAN, AS_F1MEAN, AS_SEWS, AS_SFMWS, AS_EMEANWS, AS_FMEANWS,
    AS_MIJ, AS_TW1, AS_UST2, AS_Z02, AS_JA2, AS_SIG2 = AS_JA2 + 4, AS_JS2, AS_JT2 = AS_JS2 + 2, AS_K2 = AS_JT2 + 2,
    AS_MIJ2 = AS_K2 + 7, AS_UST2B, AS_FMXB, AS_NSC
};

// The adjoint's work arrays besides ImplScratch's (rows NX apart, as there).
struct AdjExtra {
    float *fn;              // tiled: the limiter's output, then its bar, then the bar of F after the tail
    unsigned char *llws1;   // [KL*ML][NX]: LLWS of the first SINPUT
    float *cu;              // [KL][NX]: COS(TH(K) - UDIR)
    float *cub;             // [KL][NX]: the bar of cu from the limiter's floor
    float *sc;              // [AS_NSC][NX]
};

// Launch shapes, chosen by the autotuner (wam_gpu_autotune); every choice
// gives the same results, bit for bit.
struct LaunchTune {
    int prop_wg = 256;        // sea points per PROPAGS workgroup: 64, 128 or 256
    int src_wg = 64;          // sea points per workgroup of implsch_pre/post_kernel: 64, 128 or 256
    int snl_per_cu = 0;       // implsch_snl_kernel workgroups per CU (0: as many as fit)
    long chunk = 0;           // member-points per IMPLSCH launch sequence of an ensemble (0: the default)
};

// offset of the quad of bins 4q..4q+3 of sea point pt in the tiled arrays
__host__ __device__ __forceinline__ size_t quad_offset(int pt, int q, int nb4)
{
    return ((size_t)(pt >> 6) * nb4 + q) * 256 + (size_t)(pt & 63) * 4;
}

}  // namespace wam
