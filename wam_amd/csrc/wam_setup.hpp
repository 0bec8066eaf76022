// wam_setup.hpp -- host-side set-up tables of the WAM hot path.
//
// These are the module tables that IMPLSCH and PROPAGS read through USE
// statements in the reference; a Fortran host that already holds them can
// equally pass its own.  Each builder names the reference routine it follows;
// the floating-point operations are evaluated in the reference's order with
// separate roundings, so the tables are bit-identical to the Fortran REAL*4
// path wherever the elementary functions agree.
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "../../include/wam_gpu.h"

namespace wam {

constexpr int SN_GROUP = 6;    // SNONLIN list entries a thread loads together
// Padding entry of the SNONLIN lists: quadruplet slot 511 (kept zero) with the
// centre weight -2, so it adds (+0)*(-2) = -0, which leaves every value as it is.
constexpr uint16_t SN_NONE = 0x01FF;
constexpr int SP_QPT = 2;               // products form: quadruplets per thread of a chunk
constexpr int SP_TSTRIDE = 256 * SP_QPT + 4;    // products form: update slots per type (padded against LDS bank conflicts)
constexpr uint16_t SP_NONE = (256 * SP_QPT - 1) * 8;   // products form: byte offset of the unused last slot of type 0
constexpr int SP_GROUP = 4;             // products form: entry pairs a gather step takes per bin

struct Consts {           // wam_general_module.f90:41-85
    float G, PI, ZPI, DEG, RAD, CIRC, RADIUS, ROAIR, RNUAIR, RNUAIRM, ROWATER, XKAPPA, XNLEV,
        DKMAX, BETAMAX, ZALP, ALPHA, CDFAC, EMIN, EPS1, CO;
};

struct Tables {
    wam_config cfg{};
    int KL = 0, ML = 0;
    Consts k{};
    // frequency / direction grid (MAKE_FRE_DIR)
    float INV_LOG_CO = 0, DELTH = 0, DELTR = 0, MO_TAIL = 0, MM1_TAIL = 0, MP1_TAIL = 0, MP2_TAIL = 0;
    std::vector<float> FR, DF, DFIM, DFIMOFR, DFIM_FR, DFIM_FR2, FR5, FRM5, RHOWG_DFIM, GOM, TH, COSTH, SINTH;
    std::vector<int> KPM_m, KPM_p, MPM_m, MPM_p, JXO1, JXO2, JYO1, JYO2;   // 1-based values
    // MIN_ENERGY / MAKE_SHALLOW_TABLES
    int JUMAX = 300;
    float DELU = 0;
    std::vector<float> FLMINFR;                          // [ML][JUMAX]
    std::vector<float> TFAK, TCGOND, TSIHKD;             // [ML][NDEPTH]
    // DIA (NLWEIGT, INIT_SNONLIN)
    int ISP = 0, ISM = 0, MFRSTLW = 0, MLSTHG = 0, KFRH = 0;
    float ACL1 = 0, ACL2 = 0, CL11 = 0, CL21 = 0, DAL1 = 0, DAL2 = 0;
    std::vector<int> IKP, IKP1, IKM, IKM1;               // index MC - MFRSTLW
    std::vector<int> K1W, K2W, K11W, K21W;               // [2][KL], 1-based
    std::vector<float> AF11, FKLAP, FKLAP1, FKLAM, FKLAM1, FRH, FTRF;
    std::vector<int> INLCOEF;                            // [MLSTHG][5]
    std::vector<float> RNLCOEF;                          // [MLSTHG][25]
    // TAU_PHI_HF
    float X0TAUHF = 0;
    float WTAUHF[19] = {};
    // LFACTOR 10-Hz extension (depends on FR(1) only)
    int NK10 = 0;
    std::vector<float> SIG10, DSII10;
    // grid (PREPARE_GRID, regular) -- sea points numbered row-major from SW
    int NSEA = 0;
    float DELPHI = 0;
    std::vector<float> SINPH, COSPH, DELLAM;             // [NY]
    std::vector<int> IXLG, KXLT;                         // 1-based
    int ZDELLO = 0;                                      // longitude increment of every row [M_SEC]
    std::vector<int> KLAT_S, KLAT_N, KLON_W, KLON_E;     // 1-based sea point, 0 = land
    std::vector<float> DCO, DPSN_S, DPSN_N;              // DCO[NSEA]
    // propagation per-point data (PREPARE_PROPAGATION)
    std::vector<int> INDEP, INDEP_G;
    std::vector<float> DDPHI, DDLAM, DUPHI, DULAM, DVPHI, DVLAM, DELLA0, DRGM;
    // SNONLIN gather lists: per bin, contributions in the reference's order
    int NQ = 0, NCMAX = 0;
    std::vector<int32_t> CONTRIB;                        // [NB][NCMAX]
    std::vector<int32_t> NCONTRIB;                       // [NB]
    // device form, for a gather done CHUNK frequencies of quadruplets at a
    // time.  Chunk c feeds the bins [SN_CH[5c], SN_CH[5c+1]); its entries are
    // SN_E[SN_CH[5c+3] + j*(bins of the chunk) + (b - first bin)], j <
    // SN_CH[5c+2] (a multiple of SN_GROUP), each ql | (MC-offset*9 + update
    // type) << 9 with ql the quadruplet's index within the chunk (< 511), and
    // SN_NONE after the last entry of a bin.  SN_WW[SN_CH[5c+4] + g] is the
    // list length the bins blo + 64g .. blo + 64g + 63 need.
    int CHUNK = 0, NCHUNK = 0;
    std::vector<uint16_t> SN_E;
    std::vector<int32_t> SN_WW;
    std::vector<int32_t> SN_CH;                          // [NCHUNK][5]
    // products form, for a gather in which thread t of a 256-thread
    // workgroup owns the bins t + 256 j.  Chunk c covers SP_CHUNK
    // frequencies of quadruplets (fewer than S = 256*SP_QPT, SP_QPT per
    // thread); its updates sit at type*SP_TSTRIDE + ql (type 0..8, ql the
    // quadruplet's index within the chunk), float2 each, and slot S - 1 of
    // type 0 holds a zero.  Entries are byte offsets of these slots,
    // (type*SP_TSTRIDE + ql)*8, and SP_NONE = (S - 1)*8.  For run r of 64 bins (64 r .. 64 r + 63) chunk c has W =
    // SP_RW[2(c*SP_NRUN + r) + 1] entries per bin (a multiple of 2*SP_GROUP,
    // 0 when the run gets nothing from the chunk), padded with SP_NONE;
    // entries 2i and 2i+1 of bin 64 r + l sit at
    // SP_E[SP_RW[2(c*SP_NRUN + r)] + (i*64 + l)*2 + {0,1}].  A run with no
    // entries points at SP_PAD, a block of SP_GROUP pairs per lane of SP_NONE,
    // so that a group can always be loaded.
    int SP_CHUNK = 0, SP_NCHUNK = 0, SP_NRUN = 0, SP_PAD = 0;
    std::vector<uint16_t> SP_E;
    std::vector<int32_t> SP_RW;                          // [SP_NCHUNK][SP_NRUN][2]
    // the adjoint of SNONLIN, by chunks of ACH frequencies of quadruplets
    // (2*KL*ACH <= 512).  AQ_ROW[9*(MC-1) + type] is the 0-based frequency
    // row of the quadruplet's update of that type (the scatter targets of
    // snonlin_lists), -1 where the regime skips it.  A quadruplet reads the
    // nine bins (K, IC), (K1, IP), (K11, IP), (K1, IP1), (K11, IP1), (K2, IM),
    // (K21, IM), (K2, IM1), (K21, IM1) (read slots 0..8); AR_E[AR_OFF[c*ML +
    // m] .. AR_OFF[c*ML + m + 1]) lists the (MC - 1) << 4 | slot of chunk c
    // whose read row is m (nonzero weights only).  AKINV[(map*2 + KH)*KL + k]
    // inverts the direction maps K -> K1W, K11W, K2W, K21W (map 0..3).
    // adj_err is empty when every direction map is a permutation.
    int ACH = 0, ANCH = 0;
    std::vector<int8_t> AQ_ROW;
    std::vector<uint16_t> AR_E;
    std::vector<int32_t> AR_OFF;
    std::vector<int32_t> AKINV;
    std::string adj_err;
};

// Builds constants, fre_dir, tables, DIA, X0TAUHF and the grid.  Returns a
// WAM_E_* code.
int build_tables(Tables &t, const wam_config &cfg, const uint8_t *mask, std::string &err);
// PREPARE_PROPAGATION per-point part (INDEP, gradients, INDEP_G, DELLA0, DRGM).
int prepare_points(Tables &t, const float *depth, const float *u, const float *v, std::string &err);
float aki(const Tables &t, float OM, float BETA);

}  // namespace wam
