/*
 * wam_gpu.h -- C-ABI of the MI355X implementation of WAM Cycle 7's per-step
 * hot path: IMPLSCH (implicit source-term integration) and PROPAGS
 * (first-order upwind propagation over the sea-point neighbour tables).
 *
 * Reference interfaces replaced (myWAveModel/WAM):
 *   IMPLSCH  src/mod/wam_source_module.f90:167-170 (generic), :374,422-431
 *            IMPLSCH(FL3, U10, UDIR, TAUW, USTAR, Z0, ROAIRN, WSTAR, DEPTH, INDEP)
 *            called from src/chief/wamodel.f90:354
 *   PROPAGS  src/mod/wam_propagation_module.f90:151-154, :635,687
 *            PROPAGS(F3(NIJS:NIJL,KL,ML)), called from src/chief/wamodel.f90:332
 *   PREPARE_SOURCE       wam_source_module.f90:704-794   (initmdl.f90:212)
 *   PREPARE_PROPAGATION  wam_propagation_module.f90:279-598 (initmdl.f90:204, wamodel.f90:326)
 *
 * Layout: spectra are the Fortran array FL3(IJ,K,M) seen from C, i.e.
 * fl3[(m*KL + k)*NSEA + ij] with the sea point fastest (bin-major SoA).
 * Per-point fields are float[NSEA].  All entry points return 0 on success and
 * a positive WAM_E_* code otherwise (the reference aborts via ABORT1); the
 * Fortran shim turns a non-zero code into the host's abort.
 * One context is used from one host thread; calls are ordered on the
 * context's HIP stream.  Host-pointer entry points are synchronous.
 */
#ifndef WAM_GPU_H
#define WAM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    WAM_OK = 0,
    WAM_E_ARG = 1,        /* bad argument, size mismatch or unsupported option */
    WAM_E_HIP = 2,        /* HIP runtime error (allocation, launch) */
    WAM_E_CFL = 3,        /* negative propagation weight (CHECK_CFL abort), a NaN weight, or a CFL number not finite
                             or above 2^20 */
    WAM_E_STATE = 4,      /* call order violated (e.g. IMPLSCH before PREPARE_PROPAGATION) */
    WAM_E_GRID = 5,       /* empty or invalid sea mask */
    WAM_E_IO = 6          /* restart or boundary file: cannot be opened, read or written, or inconsistent with the grid */
};

/* Model options and grid definition.  Mirrors the module variables the
   reference reads through USE statements: wam_timopt_module.f90:47-68
   (IPHYS, SHALLOW_RUN, SPHERICAL_RUN, REFRACTION_D/C_RUN, ISNONLIN, IDELPRO,
   IDELT), wam_fre_dir_module.f90 (KL, ML, FR1, IREF), wam_tables_module.f90
   (NDEPTH, DEPTHA, DEPTHD), wam_general_module.f90:60-63 (BETAMAX, CDFAC,
   ZALP, ALPHA) and the regular grid of preproc_module.f90:2043-2381 with
   coordinates in the reference's M_SEC units (1 degree = 360000). */
typedef struct wam_config {
    int32_t kl, ml;
    double fr1;
    int32_t iref;
    int32_t iphys;                 /* 0: Janssen/Komen (default), 2: ST6 */
    int32_t shallow, spherical, refr_d, refr_c, iper;
    int32_t isnonlin;              /* only 0 (the default) is supported */
    int32_t idelpro, idelt;        /* propagation / source time step [s] */
    int32_t ndepth;
    double deptha, depthd;
    double betamax, zalp, alpha, cdfac;
    int32_t nx, ny;
    int64_t amosop, amonop, amowep, amoeap, xdella, xdello;
} wam_config;

typedef struct wam_ctx wam_ctx;

/* INITMDL's grid and table part: PREPARE_GRID neighbour tables from the sea
   mask (nx*ny bytes, row-major from the south-west corner, non-zero = sea),
   MAKE_FRE_DIR, MIN_ENERGY, MAKE_SHALLOW_TABLES and PREPARE_SOURCE (NLWEIGT,
   INIT_SNONLIN, INIT_X0TAUHF); uploads the tables to the device. */
int wam_gpu_init(wam_ctx **ctx, const wam_config *cfg, const uint8_t *sea_mask);

/* Host-only form of wam_gpu_init: builds the same tables without touching a
   device (for hosts that only need the tables, and for CPU-side checks).
   Such a context serves wam_gpu_get_table / wam_gpu_get_scalar /
   wam_gpu_finalize only. */
int wam_gpu_init_host(wam_ctx **ctx, const wam_config *cfg, const uint8_t *sea_mask);

/* PREPARE_PROPAGATION (+ PREPARE_PROPAGATION_COEF and CHECK_CFL): depth and
   currents per sea point (host arrays of NSEA; u, v may be NULL = no
   current).  Computes INDEP, the propagation weights on the device, the CFL
   sub-step count, and the per-point wave numbers for ST6 (WAVNU2 depends on
   depth only).  Must precede IMPLSCH and PROPAGS.  A depth that is not
   positive and finite, or a current that is not finite, returns WAM_E_ARG. */
int wam_gpu_prepare_propagation(wam_ctx *ctx, const float *depth, const float *u, const float *v);

/* Device-resident PREPARE_PROPAGATION (wam_propagation_module.f90:279-598,
   GRADIENT :182-275, PREPARE_PROPAGATION_COEF and CHECK_CFL): depth and
   currents are HBM arrays of NSEA (d_u, d_v may be NULL without current
   refraction), as when they come from a coupled model that lives on the
   device; the reference re-runs PREPARE_PROPAGATION whenever they change
   (wamodel.f90:317-326).  INDEP, INDEP_G, the depth and current gradients,
   DELLA0 and DRGM are computed on the device (bit-identical to
   wam_gpu_prepare_propagation) and copied back for wam_gpu_get_table; the
   fields are copied, so the caller's arrays may change afterwards.  Returns
   after the sub-step count is known (one synchronisation). */
int wam_gpu_prepare_propagation_dev(wam_ctx *ctx, const float *d_depth, const float *d_u, const float *d_v);

/* Where PROPAGS's weights live.  No counterpart in the reference, which
   always stores its COEF_* arrays (wam_propagation_module.f90:75-110):
   stored = 1 keeps the weight planes in HBM (5 or 9 planes of NSEA*KL*ML
   floats) and streams them in every sub-step; stored = 0 computes them
   inside the stencil from the per-point fields; stored = -1, the default,
   computes them for shallow water with current refraction and stores them
   otherwise.  Results are bit-identical in every mode.  Takes effect at the
   next wam_gpu_prepare_propagation. */
int wam_gpu_set_propagation_weights(wam_ctx *ctx, int32_t stored);

/* IMPLSCH's form on the device (no counterpart in the reference): 1 = three
   kernels (the per-lane stages up to the dissipation, SNONLIN with one
   workgroup per sea point, the per-lane new mean parameters and tail), 0 =
   one kernel with one workgroup per sea point, -1 = automatic (the default:
   three kernels for IPHYS = 0, and for ST6 from 8192 sea points on).  Where
   the three-kernel form does not apply (KL not a multiple of 4, more than
   1536 bins) the single kernel runs whatever the setting.  Results are
   bit-identical in both forms.  Takes effect at the next IMPLSCH call. */
int wam_gpu_set_implsch_form(wam_ctx *ctx, int32_t form);

/* SET_MODEL_OPTION's WAVE_BREAKING and PHILLIPS switches
   (wam_timopt_module.f90:176-177, the module variables WAVE_BREAKING_RUN and
   PHILLIPS_RUN read by IMPLSCH, wam_source_module.f90:513-515,639-646): wave
   breaking adds SDEPTHLIM (:3735-3780) and SFBRK with CMPQB (:1966-2107),
   PHILLIPS adds SOURCE_PHILLIPS (:3051-3160).  Both are off by default.  As
   in SET_MODEL_OPTION (:213-227), breaking stays off in deep water.  Takes
   effect at the next IMPLSCH call; the step adjoint refuses PHILLIPS. */
int wam_gpu_set_source_options(wam_ctx *ctx, int32_t wave_breaking, int32_t phillips);
/* The switches as they stand (after the deep-water rule); either pointer may be NULL. */
int wam_gpu_get_source_options(wam_ctx *ctx, int32_t *wave_breaking, int32_t *phillips);

/* Drop-in PROPAGS: F3(NIJS:NIJL,KL,ML) in host memory, updated in place. */
int wam_gpu_propags(wam_ctx *ctx, int32_t nijs, int32_t nijl, int32_t kl, int32_t ml, float *f3);

/* Drop-in IMPLSCH with the reference's argument list (host memory, NSEA
   points).  DEPTH and INDEP (either may be NULL) must be the fields of the
   last PREPARE_PROPAGATION, as the reference's module arrays are; a DEPTH
   that differs, or an INDEP that differs in shallow water (it is undefined
   in deep water), returns WAM_E_STATE.  lcflx != 0 also computes the WNFLUXES side outputs.  phiaw (may be
   NULL) receives PHIAW; fluxes (may be NULL) receives [6][NSEA] = PHIOC,
   TAUOC_X, TAUOC_Y, PHIBOT, TAUBOT_X, TAUBOT_Y (wam_flux_module.f90:15-21). */
int wam_gpu_implsch(wam_ctx *ctx, int32_t nsea, int32_t kl, int32_t ml, float *fl3,
                    const float *u10, const float *udir, float *tauw, float *ustar, float *z0,
                    const float *roairn, const float *wstar, const float *depth,
                    const int32_t *indep, int32_t lcflx, float *phiaw, float *fluxes);

/* Device-resident forms (pointers into HBM, ordered on the context stream,
   asynchronous).  propags_dev reads d_in and writes d_out (distinct buffers of
   NSEA*KL*ML floats).  implsch_dev updates d_fl3 in place; d_phiaw, d_fluxes
   and d_mij may be NULL. */
int wam_gpu_propags_dev(wam_ctx *ctx, const float *d_in, float *d_out);
int wam_gpu_implsch_dev(wam_ctx *ctx, float *d_fl3, const float *d_u10, const float *d_udir,
                        float *d_tauw, float *d_ustar, float *d_z0, const float *d_roairn,
                        const float *d_wstar, int32_t lcflx, float *d_phiaw, float *d_fluxes,
                        int32_t *d_mij);

/* IMPLSCH on device memory without allocating anything: the arguments and
   results of wam_gpu_implsch_dev, bit-identical to it.  No counterpart in the
   reference; it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, hosts that forbid it).  Its work
   arrays are a static device array of 256 MiB per GPU (held until the process
   ends), so the sea points are advanced in chunks of the points whose work
   arrays fit (about 12 000 at 36 x 36); spectra whose 64-point work arrays
   exceed 256 MiB are refused with WAM_E_ARG.  Every call on one GPU uses that
   GPU's array: calls that overlap on one GPU (other streams, other contexts)
   race and must be ordered by the caller.  Calls no hipMalloc or hipFree;
   asynchronous on the context stream; no per-kernel timing. */
int wam_gpu_implsch_dev_noalloc(wam_ctx *ctx, float *d_fl3, const float *d_u10, const float *d_udir, float *d_tauw,
                                float *d_ustar, float *d_z0, const float *d_roairn, const float *d_wstar,
                                int32_t lcflx, float *d_phiaw, float *d_fluxes, int32_t *d_mij);

/* Ensembles: NMEM members on the context's grid, sharing its depth, currents
   and propagation weights (wam_gpu_prepare_propagation), each with its own
   spectrum, forcing and state.  No counterpart in the reference, where an
   ensemble is NMEM separate runs of WAMODEL (wamodel.f90:332,354 per member).
   Member arrays are concatenated: spectra [NMEM][ML][KL][NSEA] (each member a
   plain FL3), per-point fields [NMEM][NSEA], fluxes [6][NMEM*NSEA].  Every
   member is advanced exactly as a plain run of it (bit for bit); each launch
   covers all members (PROPAGS with members on the grid's second dimension,
   IMPLSCH over chunks of member-points, whose size wam_gpu_autotune picks).
   NMEM * NSEA < 2^31.  HBM pointers, asynchronous on the context stream. */
int wam_gpu_propags_ensemble_dev(wam_ctx *ctx, int32_t nmem, const float *d_in, float *d_out);
int wam_gpu_implsch_ensemble_dev(wam_ctx *ctx, int32_t nmem, float *d_fl3, const float *d_u10, const float *d_udir,
                                 float *d_tauw, float *d_ustar, float *d_z0, const float *d_roairn,
                                 const float *d_wstar, int32_t lcflx, float *d_phiaw, float *d_fluxes,
                                 int32_t *d_mij);

/* NSTEPS model steps of an ensemble, as WAMODEL's time loop does them
   (wamodel.f90:332,354): PROPAGS d_fa -> d_fb, IMPLSCH on d_fb, then the two
   swap roles; the state after the last step is in d_fa if NSTEPS is even,
   else in d_fb.  The forcing stays the arrays given.  graph != 0 runs the
   whole loop as one HIP graph, captured at the first call and replayed while
   the arguments, the launch shapes and the prepared fields stay the same
   (graph = 0: the same launches issued one by one).  No LCFLX outputs. */
int wam_gpu_ensemble_run_dev(wam_ctx *ctx, int32_t nmem, int32_t nsteps, float *d_fa, float *d_fb,
                             const float *d_u10, const float *d_udir, float *d_tauw, float *d_ustar, float *d_z0,
                             const float *d_roairn, const float *d_wstar, int32_t graph);

/* Autotuner of the launch shapes (no counterpart in the reference): times, on
   the given ensemble state with HIP events, the points per PROPAGS workgroup
   (64, 128, 256), the member-points per IMPLSCH launch sequence (powers of
   two from 2^18, and all; those whose work arrays fit half the free memory),
   the sea points per workgroup of the per-lane IMPLSCH kernels (64, 128, 256)
   and the workgroups per CU of the SNONLIN kernel, and keeps the fastest (a
   coordinate search, best of reps runs each).  Every shape gives the same
   results bit for bit.  d_fa is read; d_fb is overwritten (it is the
   ensemble's second buffer); TAUW and Z0 are not changed.  The shapes apply to
   every later call of the context, plain and ensemble. */
int wam_gpu_autotune(wam_ctx *ctx, int32_t nmem, const float *d_fa, float *d_fb, const float *d_u10,
                     const float *d_udir, const float *d_tauw, const float *d_z0, const float *d_roairn,
                     const float *d_wstar, int32_t reps);
/* The launch shapes: shapes[4] = PROPAGS workgroup width, per-lane IMPLSCH
   workgroup width, SNONLIN workgroups per CU (0 = as many as fit),
   member-points per IMPLSCH launch sequence (0 = the default, 2^20). */
int wam_gpu_get_tuning(wam_ctx *ctx, int64_t *shapes);
int wam_gpu_set_tuning(wam_ctx *ctx, const int64_t *shapes);
/* The autotuner's measurements, one "parameter value milliseconds" line per
   candidate, into buf (n bytes, NUL-terminated); returns the full length. */
int64_t wam_gpu_get_tuning_log(wam_ctx *ctx, char *buf, int64_t n);

/* Several grids advanced together, with nests (no counterpart in the
   reference, where every grid is its own run of WAMODEL coupled through
   boundary value files): every step runs PROPAGS and IMPLSCH on each grid,
   then each nest's exchange, BOUNDARY_OUTPUT's gather on the coarse grid
   (wam_boundary_module.f90:350-362) and BOUNDARY_INPUT's space
   interpolation into the fine grid (:226-285), as WAMODEL inserts boundary
   values after the sources (wamodel.f90:332,354,370), all in HBM.  A grid's
   state is d_fa/d_fb (the step writes the other one; wam_gpu_multigrid_state
   tells which holds it), its forcing and TAUW, USTAR, Z0 (device, NSEA each).
   wam_gpu_multigrid_run with fused = 0 issues the steps grid after grid on
   one stream; fused = 1 runs them as one HIP graph in which the grids of a
   step are concurrent branches (captured at the first call, replayed while
   nsteps and the state parity stay the same); the results are the same bit
   for bit, and equal to plain runs with the nest entry points in between.
   A nest needs the same directions and frequencies on both grids
   (READ_BOUNDARY_INPUT's check, read_boundary_input.f90:136-156), else
   WAM_E_ARG.  Point lists are 1-based, host arrays, copied. */
typedef struct wam_grid_state {
    wam_ctx *ctx;
    float *d_fa, *d_fb;
    const float *d_u10, *d_udir, *d_roairn, *d_wstar;
    float *d_tauw, *d_ustar, *d_z0;
} wam_grid_state;
typedef struct wam_multigrid wam_multigrid;
int wam_gpu_multigrid_create(wam_multigrid **mg, int32_t ngrid, const wam_grid_state *grids);
int wam_gpu_multigrid_add_nest(wam_multigrid *mg, int32_t coarse, int32_t fine, int32_t nb, const int32_t *ijarc,
                               int32_t nbounf, const int32_t *ijarf, const int32_t *ibfl, const int32_t *ibfr,
                               const float *bfw);
int wam_gpu_multigrid_run(wam_multigrid *mg, int32_t nsteps, int32_t fused);
int wam_gpu_multigrid_state(wam_multigrid *mg, int32_t grid, float **d_cur);
const char *wam_gpu_multigrid_last_error(wam_multigrid *mg);
void wam_gpu_multigrid_free(wam_multigrid *mg);

/* The adjoint of one model step, PROPAGS then IMPLSCH (wamodel.f90:332,354),
   for NMEM members laid out as the ensemble entry points (no counterpart in
   the reference, which has no adjoint model).  It is the transpose of this
   library's forward step linearised at the given state, every branch (MIJ,
   LLWS, the limiter, AIRSEA's iteration count) taken as the forward step
   takes it there: given the step's input state d_fl3 (before PROPAGS) and
   d_tauw, its forcing and the bars of its outputs (d_fl3_bar; d_tauw_bar,
   d_ustar_bar, d_z0_bar may be NULL for zero), it writes the bars of its
   inputs: d_fl3_bar_out (may be d_fl3_bar), d_u10_bar, d_udir_bar and
   d_tauw_bar_out (Z0 and USTAR in are not read by IPHYS = 0, their bars are
   zero).  IPHYS = 0 in deep water, on a grid whose neighbour relation is
   symmetric (regular grids), else WAM_E_ARG.  HBM pointers, asynchronous on
   the context stream. */
int wam_gpu_step_adjoint_dev(wam_ctx *ctx, int32_t nmem, const float *d_fl3, const float *d_u10,
                             const float *d_udir, const float *d_tauw, const float *d_roairn, const float *d_wstar,
                             const float *d_fl3_bar, const float *d_tauw_bar, const float *d_ustar_bar,
                             const float *d_z0_bar, float *d_fl3_bar_out, float *d_u10_bar, float *d_udir_bar,
                             float *d_tauw_bar_out);

/* The adjoint of PROPAGS alone (the transposed stencil of PROP_NO_CURR /
   PROP_CURR, wam_propagation_module.f90:750-825, COUNTER sub-steps): from the
   bar of its output d_in_bar into the bar of its input d_out_bar, NMEM
   members.  Any configuration, on a grid whose neighbour relation is
   symmetric. */
int wam_gpu_propags_adjoint_dev(wam_ctx *ctx, int32_t nmem, const float *d_in_bar, float *d_out_bar);

/* Wind components to speed and direction, as WAM_WIND does
   (wam_wind_module.f90:1157-1167): U10 = SQRT(U**2 + V**2), UDIR =
   ATAN2(U, V) (0 where U10 = 0) plus ZPI where negative; and the adjoint of
   that map (zero where U10 = 0).  n values, HBM pointers, asynchronous. */
int wam_gpu_wind_uv_dev(wam_ctx *ctx, const float *d_u, const float *d_v, float *d_u10, float *d_udir, int64_t n);
int wam_gpu_wind_uv_adjoint_dev(wam_ctx *ctx, const float *d_u, const float *d_v, const float *d_u10_bar,
                                const float *d_udir_bar, float *d_u_bar, float *d_v_bar, int64_t n);

/* Checkpointing schedule for the adjoint of NSTEPS model steps with NSNAP
   checkpoint slots besides slot 0, which holds the initial state (no
   counterpart in the reference).  The schedule recomputes the fewest forward
   steps (the binomial schedule, by dynamic programming); it is a list of
   actions {op, a, b}: WAM_CK_ADVANCE runs steps a .. b-1 from the current
   state (at step a); WAM_CK_STORE copies the current state (step a) into
   slot b; WAM_CK_RESTORE makes slot b (step a) the current state;
   WAM_CK_ADJOINT runs the adjoint of step a at the current state (step a),
   after which the current state is undefined.  The first action stores step
   0 in slot 0, the adjoint steps come in the order NSTEPS-1 .. 0.  Writes up
   to cap actions (3 int32 each) and returns how many the schedule has, or
   -WAM_E_ARG (NSTEPS outside 1..20000, or a table too long to build).
   wam_checkpoint_recomputed_steps returns the forward steps ADVANCE runs. */
enum { WAM_CK_ADVANCE = 0, WAM_CK_STORE = 1, WAM_CK_RESTORE = 2, WAM_CK_ADJOINT = 3 };
int64_t wam_checkpoint_schedule(int32_t nsteps, int32_t nsnap, int32_t *actions, int64_t cap);
int64_t wam_checkpoint_recomputed_steps(int32_t nsteps, int32_t nsnap);

/* Stream used by the context (a hipStream_t; NULL = the context's own).
   No counterpart in the reference (a host-side CPU model). */
int wam_gpu_set_stream(wam_ctx *ctx, void *stream);
/* The stream the context's device work is ordered on, for callers that order
   their own streams against it. */
int wam_gpu_get_stream(wam_ctx *ctx, void **stream);

/* Integrated output parameters (COMPUTE_OUTPUT_PARAMETER,
   wam_output_module.f90:481-601, with TOTAL_ENERGY, FEMEAN,
   TM1_TM2_PERIODS, PEAK_PERIOD, MEAN_DIRECTION and CHARNOCK_PAR of
   wam_interface_module.f90) for every sea point: out[p * NSEA + ij], p one
   of the WAM_OUT_* indices below, which follow the reference's parameter
   numbers 1-16 and 32.  orient = 1 gives directions as DEG * angle
   ("coming from"), orient = 0 as MOD(DEG * angle + 180, 360) ("going to"),
   the reference's orientation_of_directions.  USTAR is taken as given (the
   reference evaluates AIRSEA first when every USTAR is zero).  Depth and
   currents are the fields of wam_gpu_prepare_propagation. */
enum {
    WAM_OUT_U10 = 0,        /*  1 wind speed [m/s] */
    WAM_OUT_UDIR,           /*  2 wind direction [deg] */
    WAM_OUT_USTAR,          /*  3 friction velocity [m/s] */
    WAM_OUT_CD,             /*  4 drag coefficient */
    WAM_OUT_CHARNOCK,       /*  5 Charnock parameter */
    WAM_OUT_DEPTH,          /*  6 water depth [m] (999 in deep water) */
    WAM_OUT_CURRENT_SPEED,  /*  7 current speed [m/s] */
    WAM_OUT_CURRENT_DIR,    /*  8 current direction [deg] */
    WAM_OUT_HS,             /*  9 significant wave height [m] */
    WAM_OUT_TP,             /* 10 peak period [s] */
    WAM_OUT_TM,             /* 11 mean period [s] */
    WAM_OUT_TM1,            /* 12 TM1 period [s] */
    WAM_OUT_TM2,            /* 13 TM2 period [s] */
    WAM_OUT_MDIR,           /* 14 mean wave direction [deg] */
    WAM_OUT_SPREAD,         /* 15 directional spread [deg] */
    WAM_OUT_TAUWN,          /* 16 normalised wave stress */
    WAM_OUT_Z0,             /* 32 roughness length [m] */
    WAM_NOUT
};

/* Device-resident form: HBM pointers, asynchronous on the context stream;
   d_out holds WAM_NOUT * NSEA floats. */
int wam_gpu_output_parameters_dev(wam_ctx *ctx, const float *d_fl3, const float *d_u10, const float *d_udir,
                                  const float *d_ustar, const float *d_tauw, const float *d_z0, int32_t orient,
                                  float *d_out);

/* Host-memory form: FL3(NSEA,KL,ML) and the per-point fields in host
   memory, out (WAM_NOUT * NSEA floats) filled on return. */
int wam_gpu_output_parameters(wam_ctx *ctx, int32_t nsea, int32_t kl, int32_t ml, const float *fl3, const float *u10,
                              const float *udir, const float *ustar, const float *tauw, const float *z0,
                              int32_t orient, float *out);

/* Gridded integrated parameters: WRITE_INT_PAR_OUTPUT
   (wam_output_module.f90:786-911), the MAP file.  fflag[WAM_MAP_NOUT] is
   FFLAG_P, non-zero for a parameter IP = index + 1 to be written; only the
   computed parameters 1-16 and 32 may be set (else WAM_E_ARG).

   wam_gpu_output_grids_dev: from d_out (WAM_NOUT * NSEA, the layout of
   wam_gpu_output_parameters_dev) to d_grids, one GRID(NX, NY) per flagged
   parameter in the order of IP: GRID(IXLG(I), KXLT(I)) = the point's value,
   ZMISS = -9999999 on land.  Asynchronous on the context stream. */
#define WAM_MAP_NOUT 70
int wam_gpu_output_grids_dev(wam_ctx *ctx, const int32_t *fflag, const float *d_out, float *d_grids);

/* One output time of the MAP file: the grids of d_out as above, then the
   records of wam_map_write with the context's grid (NX, NY, corners, NLON_RG =
   NX and the longitude increment of every row).  append = 0 starts a new file.
   cdtpro, cdatea: 14 characters each (CDTPRO, and CDATEA, the run's start). */
int wam_gpu_write_map(wam_ctx *ctx, const char *path, int32_t append, const char *cdtpro, const char *cdatea,
                      const int32_t *fflag, const float *d_out);

/* Host form of the MAP file's records for one output time (no context):
     CDTPRO, REAL(NX), REAL(NY), AMOWEP, AMOSOP, AMOEAP, AMONOP, CDATEA;
     NLON_RG(1:NY), ZDELLO(1:NY); FFLAG_P(1:70) (LOGICAL*4, .TRUE. written as 1);
     GRID(1:NX, 1:NY) per flagged parameter (grids[k][NY][NX] in order of IP);
   corners = AMOWEP, AMOSOP, AMOEAP, AMONOP [M_SEC]. */
int wam_map_write(const char *path, int32_t append, const char *cdtpro, const char *cdatea, int32_t nx, int32_t ny,
                  const int32_t *corners, const int32_t *nlon_rg, const int32_t *zdello, const int32_t *fflag,
                  const float *grids);

/* The restart file, binary form: SAVE_RESTART_FILE
   (wam_restart_module.f90:431-600) writes and CONNECT_RESTART (:109-327)
   reads Fortran sequential unformatted records, 4-byte little-endian counts
   around each record, records longer than 2 147 483 639 bytes split into
   subrecords as by the Intel and GNU Fortran runtimes:
     NSEA, CDTPRO, CDTSOU, CDA, CDTA, CDCA (INTEGER*4, 5 x CHARACTER*14);
     U10; UDIR; TAUW (NSEA each); FL3 (NSEA, KL, ML);
     DEPTH (NSEA) if CDTA is not blank; U, V (one record) if CDCA is not blank;
   sea points in the order of one undivided domain.  dates holds the five
   dates, 70 characters, blank-padded, no terminator.  A file whose NSEA
   differs from the grid's is refused with WAM_E_IO (the reference aborts).
   On read, *present gets bit 0 for a depth record, bit 1 for currents;
   depth, u, v may be NULL to skip those records.

   Device-resident forms: FL3, U10, UDIR, TAUW in HBM, staged through pinned
   host buffers; the depth and currents written are the context's fields of
   PREPARE_PROPAGATION. */
int wam_gpu_save_restart(wam_ctx *ctx, const char *path, const char *dates, const float *d_fl3, const float *d_u10,
                         const float *d_udir, const float *d_tauw);
int wam_gpu_connect_restart(wam_ctx *ctx, const char *path, char *dates, float *d_fl3, float *d_u10, float *d_udir,
                            float *d_tauw, float *d_depth, float *d_u, float *d_v, int32_t *present);

/* Host-memory forms (no context, no device). */
int wam_restart_write(const char *path, int32_t nsea, int32_t kl, int32_t ml, const char *dates, const float *fl3,
                      const float *u10, const float *udir, const float *tauw, const float *depth, const float *u,
                      const float *v);
int wam_restart_read(const char *path, int32_t nsea, int32_t kl, int32_t ml, char *dates, float *fl3, float *u10,
                     float *udir, float *tauw, float *depth, float *u, float *v, int32_t *present);

/* Nest boundary exchange.  A spectrum is F(1:KL, 1:ML), KL fastest; mean
   values are [3][n]: EMEAN, THQ [rad], FMEAN [Hz], the order of the boundary
   file.  Indices are 1-based, as in the reference.  KL, ML <= 64.

   BOUNDARY_OUTPUT (wam_boundary_module.f90:291-397, one undivided domain):
   fbc[nb][ML][KL] = FL3(IJARC(i), :, :) and mean[3][nb] from TOTAL_ENERGY,
   MEAN_DIRECTION (THQ) and FEMEAN (FM) of each, into host memory.  IJARC
   outside 1..NSEA is refused (WAM_E_ARG). */
int wam_gpu_boundary_output_dev(wam_ctx *ctx, int32_t nb, const int32_t *ijarc, const float *d_fl3, float *fbc,
                                float *mean);
int wam_gpu_boundary_output(wam_ctx *ctx, int32_t nsea, int32_t kl, int32_t ml, const float *fl3, int32_t nb,
                            const int32_t *ijarc, float *fbc, float *mean);

/* BOUNDARY_INPUT (wam_boundary_module.f90:168-287) after its reads: the time
   step takes the first set f1, m1 (which = 1: CDTPRO = CDATE1), the second
   f2, m2 (which = 2: CDTPRO = CDATE2 or CDATE1 blank) or INTSPEC
   (wam_interface_module.f90:290-408) between them at DEL1L seconds of DEL12
   (which = 0); then FL3(IJARF(i), :, :) = INTSPEC(1, BFW(i), FI(IBFL(i)),
   FI(IBFR(i))) for i = 1..NBOUNF in FL3 in HBM (the last entry of a sea point
   listed twice stays).  f1, f2 are [nbinp][ML][KL], m1, m2 [3][nbinp]; the
   set not taken may be NULL.  fi [nbinp][ML][KL] and mi [3][nbinp], the set at
   the model time, are returned when not NULL.  Indices outside 1..NSEA or
   1..NBINP are refused (WAM_E_ARG).  Synchronous. */
int wam_gpu_boundary_input_dev(wam_ctx *ctx, int32_t nbinp, const float *f1, const float *m1, const float *f2,
                               const float *m2, int32_t which, float del12, float del1l, int32_t nbounf,
                               const int32_t *ijarf, const int32_t *ibfl, const int32_t *ibfr, const float *bfw,
                               float *d_fl3, float *fi, float *mi);
int wam_gpu_boundary_input(wam_ctx *ctx, int32_t nsea, int32_t kl, int32_t ml, float *fl3, int32_t nbinp,
                           const float *f1, const float *m1, const float *f2, const float *m2, int32_t which,
                           float del12, float del1l, int32_t nbounf, const int32_t *ijarf, const int32_t *ibfl,
                           const int32_t *ibfr, const float *bfw, float *fi, float *mi);

/* The boundary value file, binary form (no context, no device).
   SAVE_BOUNDARY_FILE (wam_boundary_module.f90:709-808) writes the header
   record hdr[8] = KL, ML, TH(1), FR(1), CO, NBOUNC, IDEL_B_OUT, IDEL_BO_FILE;
   BOUNDARY_OUTPUT (:364-378) appends per point the record BLNGC, BLATC
   (INTEGER, M_SEC), CDTPRO (CHARACTER*14), EMEAN, THQ, FMEAN, DEPTH and the
   record FBC(1:KL, 1:ML); READ_BOUNDARY_INPUT (chief/read_boundary_input.f90
   :66-195) reads them.  Reading carries the byte offset *pos from one call
   to the next; dates gets 14 characters per point. */
int wam_boundary_write_header(const char *path, const float *hdr);
int wam_boundary_append(const char *path, int32_t nb, int32_t kl, int32_t ml, const int32_t *lon, const int32_t *lat,
                        const char *cdtpro, const float *mean, const float *depth, const float *fbc);
int wam_boundary_read_header(const char *path, float *hdr, int64_t *pos);
int wam_boundary_read(const char *path, int64_t *pos, int32_t nb, int32_t kl, int32_t ml, int32_t *lon, int32_t *lat,
                      char *dates, float *mean, float *depth, float *fbc);

/* Per-kernel timing on the context stream.  No counterpart in the
   reference (its timing is host wall-clock around whole calls).  on = 1
   records a HIP event pair around every kernel launch of PROPAGS, IMPLSCH and
   the nest boundary exchange from now on and drops earlier records; on = 0
   stops recording. */
int wam_gpu_set_timing(wam_ctx *ctx, int32_t on);

/* Total GPU milliseconds and number of launches of one kernel since timing
   was switched on: name is "propags" (one PROPAGS sub-step), "implsch_pre",
   "implsch_snl", "implsch_post" (the three IMPLSCH kernels of IPHYS = 0),
   "implsch" (the single IMPLSCH kernel of the other configurations),
   "boundary_gather" (BOUNDARY_OUTPUT), "boundary_time" or "boundary_space"
   (BOUNDARY_INPUT's two steps).  Waits for the context stream. */
int wam_gpu_get_timing(wam_ctx *ctx, const char *name, double *total_ms, int64_t *launches);
int wam_gpu_sync(wam_ctx *ctx);

/* Introspection for parity checks: copies a named table into out (n elements
   of the table's type: int32 for index tables, float otherwise); returns the
   table length, or -1 for an unknown name.  out may be NULL to query length. */
int64_t wam_gpu_get_table(wam_ctx *ctx, const char *name, void *out, int64_t n);
double wam_gpu_get_scalar(wam_ctx *ctx, const char *name);
const char *wam_gpu_last_error(wam_ctx *ctx);

/* The shared elementary functions evaluated on the device, for parity
   checks of the device build against the host build (no counterpart in the
   reference, which calls its runtime's intrinsics): out[i] = f(x[i], y[i])
   for fn 0 EXP, 1 LOG, 2 LOG10, 3 TANH, 4 COSH, 5 SINH, 6 COS, 7 SIN,
   8 x**y (real y), 9 NINT, 10 x**INT(y), 11 ATAN2(x, y), 12 x/1.1 (the FMA-corrected
   quotient PROPAGS uses for its SIG1 weight); host arrays of n
   (y may be NULL for one-argument functions), device 0, synchronous. */
int wam_gpu_eval_math(int32_t fn, const float *x, const float *y, float *out, int64_t n);

void wam_gpu_finalize(wam_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
