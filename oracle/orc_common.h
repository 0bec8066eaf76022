/*
 * ORACLE -- test infrastructure only.  Nothing in the product (wam_amd/) may
 * include, link or call anything in this directory; tests/, the smoke check
 * and bench.py's cpu_baseline leg use it as the checker.
 *
 * orc_common.h: precision switch and the data structures of the CPU
 * restatement of WAM Cycle 7's IMPLSCH + PROPAGS path and its set-up.
 * Arrays follow the Fortran layout of the reference: FL3(IJ,K,M) is
 * fl3[(m*KL + k)*NIJ + ij] (0-based), sea point fastest.
 *
 * What pins it (tests/): IMPLSCH (every routine it calls, both physics) and
 * PREPARE_PROPAGATION + PROPAGS against golden vectors written by running the
 * reference's own Fortran (oracle/Makefile.ref, ref_*_driver.f90); the DIA
 * tables against the reference's logged tables; grids, shallow tables,
 * FLMINFR and neighbour tables against closed forms, the dispersion relation
 * and brute force.  Parity unpinned: the elementary functions themselves
 * (wam_math.h vs the reference's libm), beyond the end-to-end bars of those
 * tests.
 *
 * Build FP32 (default, the reference's REAL) or FP64 (-DORC_FP64: the
 * reference built with 8-byte reals); -DORC_OPCOUNT (C++) is the FP32
 * arithmetic with every operation counted (orc_opcount.h), per stage at the
 * ORC_STAGE markers; -DORC_AD (C++) is the FP64 arithmetic recorded on a tape
 * for its tangent-linear model and adjoint (orc_ad.h).  The FP32 build evaluates the Fortran
 * intrinsics with wam_math.h (deterministic single precision); the FP64 build
 * uses libm double.
 */
#ifndef ORC_COMMON_H
#define ORC_COMMON_H

#include <stdint.h>
#include <stddef.h>
#include <math.h>

#ifdef ORC_FP64
typedef double real;
#define R(x) (x)
#define EXP(x) exp(x)
#define LOG(x) log(x)
#define LOG10(x) log10(x)
#define TANH(x) tanh(x)
#define COSH(x) cosh(x)
#define SINH(x) sinh(x)
#define COS(x) cos(x)
#define SIN(x) sin(x)
#define SQRT(x) sqrt(x)
#define ACOS(x) acos(x)
#define POW(x, y) pow(x, y)
#define ABS(x) fabs(x)
static inline real ipow_r(real x, int n)
{
    real r = 1.0, b = x; int first = 1;
    while (n > 0) { if (n & 1) { r = first ? b : r * b; first = 0; } n >>= 1; if (n) b = b * b; }
    return r;
}
#define IPOW(x, n) ipow_r(x, n)
static inline int NINT(real x) { return (int)lround(x); }
static inline int INTR(real x) { return (int)trunc(x); }
static inline real FLOORR(real x) { return floor(x); }
static inline real CEILR(real x) { return ceil(x); }
static inline real SIGNR(real a, real b) { return b >= 0 ? fabs(a) : -fabs(a); }
#elif defined(ORC_AD)
/* the FP64 arithmetic on the tape of orc_ad.h, the context's tables the
   FP64 build's (passive); with ORC_AD32 the REAL*4 arithmetic on the tape,
   the tables the FP32 build's */
#include "orc_ad.h"
typedef adr real;
#ifdef ORC_AD32
typedef float treal;
#define R(x) (x##f)
#else
typedef double treal;
#define R(x) (x)
#endif
#define EXP(x) ad_exp(real(x))
#define LOG(x) ad_log(real(x))
#define LOG10(x) ad_log10(real(x))
#define TANH(x) ad_tanh(real(x))
#define COSH(x) ad_cosh(real(x))
#define SINH(x) ad_sinh(real(x))
#define COS(x) ad_cos(real(x))
#define SIN(x) ad_sin(real(x))
#define SQRT(x) ad_sqrt(real(x))
#define ACOS(x) ad_acos(real(x))
#define POW(x, y) ad_pow(real(x), real(y))
#define ABS(x) ad_abs(real(x))
#define IPOW(x, n) ad_ipow(real(x), n)
#ifdef ORC_AD32
static inline int NINT(real x) { return wam_nint(x.v); }
static inline int INTR(real x) { return wam_int(x.v); }
static inline real FLOORR(real x) { return wam_floorf(x.v); }
static inline real CEILR(real x) { return wam_ceilf(x.v); }
#else
static inline int NINT(real x) { return (int)lround(x.v); }
static inline int INTR(real x) { return (int)trunc(x.v); }
static inline real FLOORR(real x) { return floor(x.v); }
static inline real CEILR(real x) { return ceil(x.v); }
#endif
static inline real SIGNR(real a, real b) { return b.v >= 0 ? ad_abs(a) : -ad_abs(a); }
#elif defined(ORC_OPCOUNT)
#include "orc_opcount.h"
typedef oc_real real;
#define R(x) (x##f)
#define EXP(x) oc_t(wam_expf((x).v))
#define LOG(x) oc_t(wam_logf((x).v))
#define LOG10(x) oc_t(wam_log10f((x).v))
#define TANH(x) oc_t(wam_tanhf((x).v))
#define COSH(x) oc_t(wam_coshf((x).v))
#define SINH(x) oc_t(wam_sinhf((x).v))
#define COS(x) oc_t(wam_cosf(oc_real(x).v))
#define SIN(x) oc_t(wam_sinf(oc_real(x).v))
#define SQRT(x) oc_sqrt(x)
#define ACOS(x) oc_t((float)acos((double)(x)))
#define POW(x, y) oc_t(wam_powf(oc_real(x).v, oc_real(y).v))
#define ABS(x) oc_real(wam_fabsf(oc_real(x).v))
#define IPOW(x, n) oc_ipow(x, n)
static inline int NINT(real x) { return wam_nint(x.v); }
static inline int INTR(real x) { return wam_int(x.v); }
static inline real FLOORR(real x) { return wam_floorf(x.v); }
static inline real CEILR(real x) { return wam_ceilf(x.v); }
static inline real SIGNR(real a, real b) { return wam_signf(a.v, b.v); }
#else
#include "wam_math.h"
typedef float real;
#define R(x) (x##f)
#define EXP(x) wam_expf(x)
#define LOG(x) wam_logf(x)
#define LOG10(x) wam_log10f(x)
#define TANH(x) wam_tanhf(x)
#define COSH(x) wam_coshf(x)
#define SINH(x) wam_sinhf(x)
#define COS(x) wam_cosf(x)
#define SIN(x) wam_sinf(x)
#define SQRT(x) sqrtf(x)
#define ACOS(x) ((float)acos((double)(x)))
#define POW(x, y) wam_powf(x, y)
#define ABS(x) wam_fabsf(x)
#define IPOW(x, n) wam_ipowf(x, n)
static inline int NINT(real x) { return wam_nint(x); }
static inline int INTR(real x) { return wam_int(x); }
static inline real FLOORR(real x) { return wam_floorf(x); }
static inline real CEILR(real x) { return wam_ceilf(x); }
static inline real SIGNR(real a, real b) { return wam_signf(a, b); }
#endif

/* The type of the context's stored tables and constants: the arithmetic type
   in every build except the differentiating one (-DORC_AD, C++), where the
   per-point work is active and the set-up (built in FP64) is passive. */
#ifndef ORC_AD
typedef real treal;
#endif

/* Stage markers of the op-counting build (no-ops otherwise). */
#ifdef ORC_OPCOUNT
#define ORC_STAGE(n) oc_stage(n)
#else
#define ORC_STAGE(n) ((void)0)
#endif

static inline real MAXR(real a, real b) { return a > b ? a : b; }
static inline real MINR(real a, real b) { return a < b ? a : b; }
static inline int MAXI(int a, int b) { return a > b ? a : b; }
static inline int MINI(int a, int b) { return a < b ? a : b; }

/* Options and dimensions (wam_timopt_module.f90:47-68, wam_fre_dir_module.f90). */
typedef struct {
    int kl, ml;             /* directions, frequencies */
    double fr1;             /* reference frequency [Hz] */
    int iref;               /* bin of the reference frequency */
    int iphys;              /* 0 Janssen, 2 ST6 */
    int shallow, spherical, refr_d, refr_c, iper;
    int isnonlin;
    int idelpro, idelt;     /* propagation / source time step [s] */
    int ndepth;             /* shallow-water table length */
    double deptha, depthd;  /* first depth, depth ratio */
    double betamax, zalp, alpha, cdfac;
    /* regular lat-lon grid in M_SEC units (1 degree = 360000) */
    int nx, ny;
    int64_t amosop, amonop, amowep, amoeap, xdella, xdello;
} orc_config;

/* Everything the path reads from module globals in the reference. */
typedef struct {
    orc_config cfg;
    int KL, ML;
    /* constants (wam_general_module.f90:41-85) */
    treal G, PI, ZPI, DEG, RAD, CIRC, RADIUS, ROAIR, RNUAIR, RNUAIRM, ROWATER,
         XKAPPA, XNLEV, DKMAX, BETAMAX, ZALP, ALPHA, CDFAC, EMIN, EPS1;
    /* fre_dir (wam_fre_dir_module.f90:297-461) */
    treal CO, INV_LOG_CO, DELTH, DELTR, MO_TAIL, MM1_TAIL, MP1_TAIL, MP2_TAIL, FMIN;
    treal *FR, *DF, *DF_FR, *DF_FR2, *DFIM, *DFIMOFR, *DFIM_FR, *DFIM_FR2, *FR5, *FRM5,
         *RHOWG_DFIM, *GOM, *C, *TH, *COSTH, *SINTH;
    int *MPM, *KPM, *JXO, *JYO;            /* [ML][3], [KL][3], [KL][2], [KL][2] */
    /* tables (wam_tables_module.f90) */
    int JUMAX; treal UMAX, DELU, FLMIN;
    treal *FLMINFR;                          /* [ML][JUMAX]: FLMINFR(JU,M) */
    treal DEPTHA, DEPTHD;
    treal *TFAK, *TCGOND, *TSIHKD, *T_TAIL;  /* [ML][NDEPTH] */
    /* DIA (wam_source_module.f90:1056-1523) */
    int ISP, ISM, MFRSTLW, MLSTHG, KFRH;
    int *IKP, *IKP1, *IKM, *IKM1;           /* index MFRSTLW..MLSTHG -> [M-MFRSTLW] */
    int *K1W, *K2W, *K11W, *K21W;           /* [2][KL] (KH, K) */
    treal *AF11, *FKLAP, *FKLAP1, *FKLAM, *FKLAM1, *FRH, *FTRF;
    treal ACL1, ACL2, CL11, CL21, DAL1, DAL2;
    int *INLCOEF;                           /* [MLSTHG][5] */
    treal *RNLCOEF;                          /* [MLSTHG][25] */
    /* TAU_PHI_HF (wam_source_module.f90:3784-3849) */
    treal X0TAUHF, WTAUHF[19];
    /* grid (preproc_module.f90:2043-2381) */
    int NSEA;
    treal DELPHI;
    treal *SINPH, *COSPH, *DELLAM;           /* [NY] */
    int *ZDELLO, *IXLG, *KXLT;              /* 1-based grid indices per sea point */
    int *KLAT;                              /* [2 (ICL)][2 (IC)][NSEA] : KLAT(IJ,IC,ICL) 1-based, 0 = land */
    int *KLON;                              /* [2][NSEA] */
    treal *WLAT;                             /* [2][NSEA] */
    treal *DCO;                              /* [NSEA+1], index 0 = land slot */
    treal *DPSN;                             /* [2][NSEA] */
    /* propagation */
    int counter; treal cfl_max, cfl_min;
    int force_counter;   /* > 0: COUNTER of the grid this one is a window of (CHECK_CFL's rescale by it) */
    /* WAVE_BREAKING_RUN and PHILLIPS_RUN (wam_timopt_module.f90:60-63, set by
       SET_MODEL_OPTION :176-177; breaking is switched off in deep water :213-227) */
    int wave_breaking, phillips;
    int ncoef;                              /* coefficient planes present */
    treal *COEF;                             /* [ncoef][ML][KL][NSEA] */
    int *INDEP;                             /* [NSEA] */
    treal *WN, *CGG;                         /* WAVNU2 results [ML][NSEA] (IPHYS=2) */
} orc_ctx;

/* coefficient plane order in COEF: */
enum { C_SUM = 0, C_LAT11, C_LAT21, C_LON1, C_LON2, C_TH1, C_TH2, C_SIG1, C_SIG2, C_NPLANES };

#endif
