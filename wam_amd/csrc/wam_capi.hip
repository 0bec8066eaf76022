// wam_capi.hip -- context management and the extern "C" entry points of
// include/wam_gpu.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "../../include/wam_gpu.h"
#include "wam_dev.hpp"
#include "wam_math.h"
#include "wam_restart.hpp"
#include "wam_setup.hpp"

namespace wam {
__global__ void prep_coef_kernel(DevTables t, float *coef);
__global__ void prep_points_kernel(DevTables t, PrepPoints p);
__global__ void minmax_kernel(const float *x, size_t n, float *out);
__global__ void rescale_kernel(float *coef, size_t nb, int nplanes, float rc);
__global__ void propags_nocurr_kernel(DevTables t, const float *coef, const float *fin, float *fout);
__global__ void propags_curr_kernel(DevTables t, const float *coef, const float *fin, float *fout);
template <int KLT>
__global__ void propags_curr_rows_kernel(DevTables t, const float *coef, const float *fin, float *fout);
template <bool CURR>
__global__ void cfl_minmax_kernel(DevTables t, float *out);
template <bool CURR>
__global__ void propags_fused_kernel(DevTables t, int counter, const float *fin, float *fout);
template <int RESCALE>
__global__ void propags_fused_sc_kernel(DevTables t, int counter, const float *fin, float *fout);
template <int RESCALE, int KB>
__global__ void propags_blk_kernel(DevTables t, int counter, const float *fin, float *fout);
__global__ void wavnu2_kernel(DevTables t, float *wn, float *cgg);
__global__ void sbottom_factor_kernel(DevTables t, float *sbot);
hipError_t launch_implsch(const DevTables &t, const ImplArgs &a, int ncu, hipStream_t st);
hipError_t launch_implsch_split(const DevTables &t, const ImplArgs &a, const ImplScratch &x, int ncu, hipStream_t st,
                                const hipEvent_t *ev, const LaunchTune &tu);
bool implsch_split_supported(const DevTables &t);
hipError_t launch_implsch_adjoint(const DevTables &t, const AdjArgs &a, int ncu, hipStream_t st);
hipError_t launch_implsch_adjoint_split(const DevTables &t, const AdjArgs &a, const ImplScratch &x, const AdjExtra &ex,
                                        int ncu, hipStream_t st, const LaunchTune &tu);
hipError_t launch_propags_adjoint(const DevTables &t, const float *coef, const float *tw, int nmem, const float *gin,
                                  float *gout, hipStream_t st);
hipError_t launch_transpose_weights(const DevTables &t, const float *coef, float *tw, hipStream_t st);
hipError_t launch_wind_uv(float zpi, const float *u, const float *v, float *u10, float *udir, int64_t n,
                          hipStream_t st);
hipError_t launch_wind_uv_adjoint(const float *u, const float *v, const float *u10b, const float *udirb, float *ub,
                                  float *vb, int64_t n, hipStream_t st);
hipError_t launch_output_params(const DevTables &t, const float *fl3, const float *u10, const float *udir,
                                const float *ustar, const float *tauw, const float *z0, int orient, float *out,
                                hipStream_t st);
hipError_t launch_map_grids(const float *out, int nsea, const int *cell_pt, int ncell, const int *rows, int nrows,
                            float *grids, hipStream_t st);
hipError_t launch_boundary_gather(const DevTables &t, const float *fl3, const int *ijarc, int nb, float *fbc,
                                  float *mean, hipStream_t st);
hipError_t launch_boundary_time(const DevTables &t, int nbinp, const float *f1, const float *m1, const float *f2,
                                const float *m2, int which, float del12, float del1l, float *fi, float *mi,
                                hipStream_t st);
hipError_t launch_boundary_space(const DevTables &t, int nbinp, const float *fi, const float *mi, int nbounf,
                                 const int *ijarf, const int *ibfl, const int *ibfr, const float *bfw, float *fl3,
                                 hipStream_t st);
}  // namespace wam

using namespace wam;

// The work arrays of wam_gpu_implsch_dev_noalloc: a file-scope device array,
// one copy per GPU (held from the library's first use on that GPU until the
// process ends).  Host code never passes its name to a kernel: the device
// address comes from hipGetSymbolAddress, per device, at wam_gpu_init.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
static void *scratch_dev[64];             // device address of scratch on each device (nullptr: not yet looked up)
static std::mutex scratch_mu;

namespace {
__global__ void eval_math_kernel(int fn, const float *x, const float *y, float *out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float a = x[i], b = y ? y[i] : 0.f;
    float r = 0.f;
    switch (fn) {
    case 0: r = wam_expf(a); break;
    case 1: r = wam_logf(a); break;
    case 2: r = wam_log10f(a); break;
    case 3: r = wam_tanhf(a); break;
    case 4: r = wam_coshf(a); break;
    case 5: r = wam_sinhf(a); break;
    case 6: r = wam_cosf(a); break;
    case 7: r = wam_sinf(a); break;
    case 8: r = wam_powf(a, b); break;
    case 9: r = (float)wam_nint(a); break;
    case 10: r = wam_ipowf(a, (int)b); break;
    case 11: r = wam_atan2f(a, b); break;
    case 12: r = wam_div_1p1(a); break;
    }
    out[i] = r;
}
}  // namespace

struct wam_ctx {
    Tables tab;
    DevTables dt{};
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    std::vector<void *> dev_allocs;       // static tables (freed at finalize)
    std::vector<void *> point_allocs;     // per depth field
    float *coef = nullptr;          // stored weight planes (NULL: computed in the stencil)
    int ncoef = 0;
    int weights_mode = -1;          // wam_gpu_set_propagation_weights: -1 auto, 0 computed, 1 stored
    int implsch_form = -1;          // wam_gpu_set_implsch_form: -1 auto, 0 single kernel, 1 three kernels
    int counter = 1;
    float cfl_max = 0.f, cfl_min = 0.f;
    float *wn = nullptr, *cgg = nullptr;
    float *fscratch = nullptr;            // COUNTER > 1 ping-pong
    ImplScratch isx{};                    // IMPLSCH work arrays (IPHYS = 0), allocated at the first call
    ImplScratch esx{};                    // IMPLSCH work arrays of ensembles (a chunk of member-points)
    float *fscratch_e = nullptr;          // ensembles: COUNTER > 1 ping-pong
    size_t fscratch_e_n = 0;
    LaunchTune tune;                      // launch shapes (wam_gpu_autotune, wam_gpu_set_tuning)
    struct TuneEntry { std::string what; long value; float ms; };
    std::vector<TuneEntry> tune_log;      // the last autotuner's measurements
    // the captured ensemble loop (wam_gpu_ensemble_run_dev): its stream, the
    // two ordering events, the graph and the arguments it was built for
    hipStream_t gstream = nullptr;
    hipEvent_t gev[2] = {nullptr, nullptr};
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<uintptr_t> gkey;
    // per-kernel timing (wam_gpu_set_timing): an event pair per launch
    bool timing = false;
    struct Timed { std::string name; hipEvent_t a, b; };
    std::vector<Timed> timed;
    float *stage = nullptr;               // host drop-in staging: FL3 + per-point fields
    size_t stage_n = 0;
    float *nest = nullptr;                // nest boundary spectra, mean values and indices
    size_t nest_n = 0;
    int *cell_pt = nullptr;               // grid cell -> sea point (0-based, -1 land), for gridded output
    float *grids = nullptr;               // gridded parameter fields
    size_t grids_n = 0;
    std::vector<float> depth_h;           // the DEPTH of the last PREPARE_PROPAGATION (host copy)
    bool prepared = false;
    void *noalloc = nullptr;              // this device's copy of scratch (wam_gpu_implsch_dev_noalloc)
    // the step adjoint (wam_gpu_step_adjoint_dev): IMPLSCH's input spectra
    // and their bars for adj_n values, the weight planes where PROPAGS
    // computes its weights in the stencil (a point allocation), and whether
    // the grid's neighbour relation is symmetric (-1: not yet checked)
    float *adj_fp = nullptr, *adj_fb = nullptr;
    size_t adj_n = 0;
    float *adj_coef = nullptr;
    float *adj_tw = nullptr;              // the no-current weights transposed (a point allocation)
    int adj_sym = -1;
    // the split IMPLSCH adjoint's work arrays for a chunk of member-points
    ImplScratch asx{};
    AdjExtra aex{};
    int ncu = 256;                 // compute units of the device
    std::string err;
};

// A failed HIP call also leaves its code as the thread's last error, which
// a later launch check (hipGetLastError) would report again: it is cleared
// with the failure, so the next call starts clean.
#define HIPCHK(x)                                                                  \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            (void)hipGetLastError();                                               \
            ctx->err = std::string(#x) + ": " + hipGetErrorString(e_);             \
            return WAM_E_HIP;                                                      \
        }                                                                          \
    } while (0)

// n events recorded around launches when timing is on (ev[i] = nullptr otherwise)
static int timing_events(wam_ctx *ctx, hipEvent_t *ev, int n)
{
    for (int i = 0; i < n; i++) {
        ev[i] = nullptr;
        if (ctx->timing) HIPCHK(hipEventCreate(&ev[i]));
    }
    return WAM_OK;
}

static void timing_record(wam_ctx *ctx, hipEvent_t e)
{
    if (e) (void)hipEventRecord(e, ctx->stream);
}

static void timing_keep(wam_ctx *ctx, const char *name, hipEvent_t a, hipEvent_t b)
{
    if (a && b) ctx->timed.push_back({name, a, b});
}

static void timing_clear(wam_ctx *ctx)
{
    for (auto &r : ctx->timed) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    ctx->timed.clear();
}

template <class T>
static int upload(wam_ctx *ctx, std::vector<void *> &owner, const T *host, size_t n, const T **dev)
{
    void *p = nullptr;
    HIPCHK(hipMalloc(&p, (n ? n : 1) * sizeof(T)));
    owner.push_back(p);
    if (n) HIPCHK(hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice));
    *dev = (const T *)p;
    return WAM_OK;
}

static void free_list(std::vector<void *> &v)
{
    for (void *p : v) (void)hipFree(p);
    v.clear();
}

#define UP(owner, vec, field)                                                        \
    do {                                                                           \
        int r_ = upload(ctx, owner, (vec).data(), (vec).size(), &ctx->dt.field);   \
        if (r_) return r_;                                                         \
    } while (0)

static_assert(SN_GROUP == SN_GROUP_DEV && SN_NONE == SN_NONE_DEV, "host and device SNONLIN list layouts differ");
static_assert(SP_QPT == SP_QPT_DEV && SP_NONE == SP_NONE_DEV && SP_TSTRIDE == SP_TSTRIDE_DEV && SP_GROUP == SP_GROUP_DEV,
              "host and device SNONLIN products layouts differ");

// IMPLSCH's form: three kernels (per-lane stages, SNONLIN per workgroup)
// unless the single kernel is asked for; automatic: always for IPHYS = 0,
// for ST6 from 8192 sea points on (its per-lane LFACTOR chain is long, and
// below that the single kernel's one workgroup per point keeps the GPU
// busier).
static int implsch_form_split(int form, int iphys, int nsea)
{
    if (form >= 0) return form;
    return iphys == 0 || nsea >= 8192 ? 1 : 0;
}

static int upload_static(wam_ctx *ctx)
{
    Tables &t = ctx->tab;
    DevTables &d = ctx->dt;
    d.KL = t.KL; d.ML = t.ML; d.NSEA = t.NSEA; d.NDEPTH = t.cfg.ndepth; d.JUMAX = t.JUMAX;
    d.MFRSTLW = t.MFRSTLW; d.MLSTHG = t.MLSTHG; d.NQ = t.NQ; d.NCMAX = t.NCMAX; d.NK10 = t.NK10;
    d.CHUNK = t.CHUNK; d.NCHUNK = t.NCHUNK;
    d.iphys = t.cfg.iphys; d.shallow = t.cfg.shallow; d.refr_d = t.cfg.refr_d; d.refr_c = t.cfg.refr_c;
    d.split = implsch_form_split(ctx->implsch_form, t.cfg.iphys, t.NSEA);
    const Consts &k = t.k;
    d.G = k.G; d.ZPI = k.ZPI; d.ROAIR = k.ROAIR; d.RNUAIRM = k.RNUAIRM; d.ROWATER = k.ROWATER;
    d.XKAPPA = k.XKAPPA; d.XNLEV = k.XNLEV; d.BETAMAX = k.BETAMAX; d.ZALP = k.ZALP; d.ALPHA = k.ALPHA;
    d.CDFAC = k.CDFAC; d.EMIN = k.EMIN; d.EPS1 = k.EPS1; d.CO = k.CO;
    d.DELTH = t.DELTH; d.INV_LOG_CO = t.INV_LOG_CO; d.MO_TAIL = t.MO_TAIL; d.MM1_TAIL = t.MM1_TAIL;
    d.MP1_TAIL = t.MP1_TAIL; d.MP2_TAIL = t.MP2_TAIL; d.DELU = t.DELU; d.DAL1 = t.DAL1; d.DAL2 = t.DAL2; d.X0TAUHF = t.X0TAUHF;
    d.DELT = (float)t.cfg.idelt;
    d.DEG = k.DEG;
    // constant expressions the Fortran compiler folds: one correct rounding each
    d.ZPI4 = (float)std::pow((double)k.ZPI, 4.0);
    const float g4 = (float)std::pow((double)k.G, 4.0);
    d.CONSD = ((-1.33f) * (float)std::pow((double)k.ZPI, 9.0)) / g4;
    d.CONSS = (-1.33f) * k.ZPI;
    d.C5G = 5.0E-07f * k.G;
    const float delpro = (float)t.cfg.idelpro;
    d.DELTH0 = (0.5f * delpro) / t.DELTR;
    d.DELPH0 = (0.5f * delpro) / t.DELPHI;
    d.DELTHR = (0.5f * delpro) / t.DELTH;
    d.DELFR0 = t.cfg.shallow ? (0.5f * delpro) / (0.1f * k.ZPI) : ((0.5f * delpro) * 2.1f) / 0.2f;

    auto &o = ctx->dev_allocs;
    UP(o, t.FR, FR); UP(o, t.DF, DF); UP(o, t.DFIM, DFIM); UP(o, t.DFIMOFR, DFIMOFR);
    UP(o, t.DFIM_FR, DFIM_FR); UP(o, t.DFIM_FR2, DFIM_FR2); UP(o, t.FR5, FR5); UP(o, t.FRM5, FRM5); UP(o, t.RHOWG_DFIM, RHOWG_DFIM);
    UP(o, t.GOM, GOM); UP(o, t.TH, TH); UP(o, t.COSTH, COSTH); UP(o, t.SINTH, SINTH);
    UP(o, t.FLMINFR, FLMINFR); UP(o, t.TFAK, TFAK); UP(o, t.TCGOND, TCGOND); UP(o, t.TSIHKD, TSIHKD);
    UP(o, t.SIG10, SIG10); UP(o, t.DSII10, DSII10);
    std::vector<float> wtauhf(t.WTAUHF, t.WTAUHF + 19);
    UP(o, wtauhf, WTAUHF);
    UP(o, t.INLCOEF, INLCOEF); UP(o, t.RNLCOEF, RNLCOEF);
    std::vector<float> af(t.MLSTHG);
    for (int mc = 1; mc <= t.MLSTHG; mc++) af[mc - 1] = t.AF11[mc - t.MFRSTLW];
    UP(o, af, AF11);
    auto zero_based = [](const std::vector<int> &v) { std::vector<int> r(v); for (auto &x : r) x -= 1; return r; };
    std::vector<int> k1 = zero_based(t.K1W), k2 = zero_based(t.K2W), k11 = zero_based(t.K11W), k21 = zero_based(t.K21W);
    UP(o, k1, K1W); UP(o, k2, K2W); UP(o, k11, K11W); UP(o, k21, K21W);
    UP(o, t.SN_E, SN_E); UP(o, t.SN_CH, SN_CH); UP(o, t.SN_WW, SN_WW);
    d.SP_CHUNK = t.SP_CHUNK; d.SP_NCHUNK = t.SP_NCHUNK; d.SP_NRUN = t.SP_NRUN; d.SP_PAD = t.SP_PAD;
    UP(o, t.SP_E, SP_E); UP(o, t.SP_RW, SP_RW);
    d.ACH = t.ACH; d.ANCH = t.ANCH; d.NAR = (int)t.AR_E.size();
    UP(o, t.AQ_ROW, AQ_ROW); UP(o, t.AR_E, AR_E); UP(o, t.AR_OFF, AR_OFF); UP(o, t.AKINV, AKINV);
    std::vector<int> ks = zero_based(t.KLAT_S), kn = zero_based(t.KLAT_N), kw = zero_based(t.KLON_W), ke = zero_based(t.KLON_E);
    UP(o, ks, KLAT_S); UP(o, kn, KLAT_N); UP(o, kw, KLON_W); UP(o, ke, KLON_E);
    UP(o, t.DPSN_S, DPSN_S); UP(o, t.DPSN_N, DPSN_N);
    UP(o, t.JXO1, JXO1); UP(o, t.JXO2, JXO2); UP(o, t.JYO1, JYO1); UP(o, t.JYO2, JYO2);
    UP(o, t.KPM_p, KPM_P); UP(o, t.KPM_m, KPM_M);
    UP(o, t.DCO, DCO_PT); UP(o, t.DELLAM, DELLAM_ROW); UP(o, t.SINPH, SINPH_ROW); UP(o, t.KXLT, KXLT_PT);
    d.DELPHI = t.DELPHI;
    std::vector<float> kdir((size_t)8 * t.KL, 0.f);
    for (int k = 0; k < t.KL; k++) {
        const int kp = k + 1 < t.KL ? k + 1 : 0, km = k > 0 ? k - 1 : t.KL - 1;
        const float s = t.SINTH[k], c = t.COSTH[k];
        float *kd = kdir.data() + 8 * k;
        kd[0] = s; kd[1] = c;
        kd[2] = d.DELTH0 * (s + t.SINTH[kp]);
        kd[3] = d.DELTH0 * (s + t.SINTH[km]);
        kd[4] = s * s; kd[5] = s * c; kd[6] = c * c;
    }
    UP(o, kdir, KDIR);
    return WAM_OK;
}

extern "C" {

int wam_gpu_init_host(wam_ctx **out, const wam_config *cfg, const uint8_t *sea_mask)
{
    if (!out || !cfg || !sea_mask) return WAM_E_ARG;
    wam_ctx *ctx = new wam_ctx();
    *out = ctx;
    return build_tables(ctx->tab, *cfg, sea_mask, ctx->err);
}

int wam_gpu_init(wam_ctx **out, const wam_config *cfg, const uint8_t *sea_mask)
{
    if (!out || !cfg || !sea_mask) return WAM_E_ARG;
    *out = nullptr;
    int rc = wam_gpu_init_host(out, cfg, sea_mask);
    wam_ctx *ctx = *out;
    if (rc) return rc;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) { ctx->err = "no HIP device"; *out = ctx; return WAM_E_HIP; }
    if (hipStreamCreate(&ctx->own_stream) != hipSuccess) { ctx->err = "hipStreamCreate failed"; *out = ctx; return WAM_E_HIP; }
    ctx->stream = ctx->own_stream;
    // the device address of the no-allocation scratch array, once per device
    // (with that device current, as it is here)
    if (dev >= 0 && dev < 64) {
        std::lock_guard<std::mutex> lock(scratch_mu);
        if (!scratch_dev[dev]) {
            // a failure leaves the no-allocation entry point unusable
            // (it reports WAM_E_STATE) and every other one as it is
            void *ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess && ptr) scratch_dev[dev] = ptr;
            else (void)hipGetLastError();
        }
        ctx->noalloc = scratch_dev[dev];
    }
    rc = upload_static(ctx);
    if (!rc) {
        int ncu = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) == hipSuccess) ncu = prop.multiProcessorCount;
        ctx->ncu = ncu;
    }
    *out = ctx;
    return rc;
}

// PREPARE_PROPAGATION after the per-point fields are in ctx->dt: the
// weights (stored planes or CHECK_CFL over computed ones), COUNTER, WAVNU2.
static int finish_prepare(wam_ctx *ctx)
{
    Tables &t = ctx->tab;
    const int ns = t.NSEA;
    auto &o = ctx->point_allocs;
    void *p = nullptr;
    const size_t nb = (size_t)t.KL * t.ML * ns;
    ctx->ncoef = t.cfg.refr_c ? 9 : 5;
    hipStream_t st = ctx->stream;
    std::vector<float> lo(ctx->ncoef, 3.4e38f), hi(ctx->ncoef, -3.4e38f);
    bool nan_weight = false;
    // auto: computed in the stencil where the hoisted shallow-water + current
    // kernel applies (it is faster than streaming nine planes), else stored
    const bool stored = ctx->weights_mode == 1 ||
                        (ctx->weights_mode == -1 && !(t.cfg.shallow && t.cfg.refr_c));
    if (stored) {
        // the reference's COEF_* planes in HBM, then CHECK_CFL over them
        HIPCHK(hipMalloc(&p, nb * ctx->ncoef * sizeof(float)));
        o.push_back(p);
        ctx->coef = (float *)p;
        hipLaunchKernelGGL(prep_coef_kernel, dim3((ns + 255) / 256, t.KL * t.ML), dim3(256), 0, st, ctx->dt, ctx->coef);
        HIPCHK(hipGetLastError());
        const int nblk = 1024;
        float *mm = nullptr;
        HIPCHK(hipMalloc(&mm, 2 * nblk * sizeof(float) * ctx->ncoef));
        for (int pl = 0; pl < ctx->ncoef; pl++)
            hipLaunchKernelGGL(minmax_kernel, dim3(nblk), dim3(256), 0, st, ctx->coef + pl * nb, nb, mm + 2 * nblk * pl);
        std::vector<float> h(2 * nblk * ctx->ncoef);
        HIPCHK(hipMemcpyAsync(h.data(), mm, h.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        (void)hipFree(mm);
        for (int pl = 0; pl < ctx->ncoef; pl++)
            for (int b = 0; b < nblk; b++) {
                const float a = h[2 * nblk * pl + 2 * b], z = h[2 * nblk * pl + 2 * b + 1];
                if (a != a || z != z) nan_weight = true;
                lo[pl] = fminf(lo[pl], a);
                hi[pl] = fmaxf(hi[pl], z);
            }
    } else {
        // CHECK_CFL over weights computed on the fly; PROPAGS recomputes them
        const int nblk = (ns + 255) / 256;
        float *mm = nullptr;
        HIPCHK(hipMalloc(&mm, (size_t)nblk * 2 * P_NPLANES * sizeof(float)));
        if (t.cfg.refr_c) hipLaunchKernelGGL(cfl_minmax_kernel<true>, dim3(nblk), dim3(256), 0, st, ctx->dt, mm);
        else hipLaunchKernelGGL(cfl_minmax_kernel<false>, dim3(nblk), dim3(256), 0, st, ctx->dt, mm);
        HIPCHK(hipGetLastError());
        std::vector<float> h((size_t)nblk * 2 * P_NPLANES);
        HIPCHK(hipMemcpyAsync(h.data(), mm, h.size() * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        (void)hipFree(mm);
        for (int b = 0; b < nblk; b++)
            for (int pl = 0; pl < ctx->ncoef; pl++) {
                const float a = h[((size_t)b * P_NPLANES + pl) * 2], z = h[((size_t)b * P_NPLANES + pl) * 2 + 1];
                if (a != a || z != z) nan_weight = true;
                lo[pl] = fminf(lo[pl], a);
                hi[pl] = fmaxf(hi[pl], z);
            }
    }
    // plane order: no current SUM LAT1 LON1 TH1 TH2; current SUM LAT1 LAT2 LON1 LON2 TH1 TH2 SIG1 SIG2
    float lat11mx, lat11mn, lon2mx = 0.1f, lon2mn = 0.1f, s1mx = 0.1f, s1mn = 0.1f, s2mx = 0.1f, s2mn = 0.1f;
    float lon1mx, th1mx, th1mn, th2mx, th2mn;
    if (t.cfg.refr_c) {
        lat11mx = fmaxf(hi[P_LAT1], hi[P_LAT2]); lat11mn = fminf(lo[P_LAT1], lo[P_LAT2]);
        lon1mx = hi[P_LON1]; lon2mx = hi[P_LON2]; lon2mn = lo[P_LON2];
        th1mx = hi[P_TH1]; th1mn = lo[P_TH1]; th2mx = hi[P_TH2]; th2mn = lo[P_TH2];
        s1mx = hi[P_SIG1]; s1mn = lo[P_SIG1]; s2mx = hi[P_SIG2]; s2mn = lo[P_SIG2];
    } else {
        lat11mx = hi[1]; lat11mn = lo[1]; lon1mx = hi[2];
        th1mx = hi[3]; th1mn = lo[3]; th2mx = hi[4]; th2mn = lo[4];
    }
    const float summx = 1.f - lo[0], summn = 1.f - hi[0];
    float cmax = lat11mx;
    for (float x : {0.1f, lon1mx, lon2mx, th1mx, th2mx, s1mx, s2mx, summx}) cmax = fmaxf(cmax, x);
    // the reference's minimum omits COEF_LON_1_1_MIN and lists COEF_LON_1_2_MIN twice
    float cmin = lat11mn;
    for (float x : {0.1f, lon2mn, lon2mn, th1mn, th2mn, s1mn, s2mn, summn}) cmin = fminf(cmin, x);
    ctx->cfl_max = cmax; ctx->cfl_min = cmin;
    // a NaN weight (from finite but extreme depths or currents) is refused:
    // the propagation kernels take MIN(0, x) and MAX(0, x) of arguments that
    // are then never NaN
    if (nan_weight) { ctx->err = "a propagation weight is not a number (CHECK_CFL)"; return WAM_E_CFL; }
    if (cmin < 0.f) { ctx->err = "negative propagation weight (CHECK_CFL)"; return WAM_E_CFL; }
    // COUNTER = CEILING(CFL) sub-steps: a CFL number that is not finite, or
    // one that would need more than 2^20 sub-steps, is refused
    if (!(cmax <= 1048576.f)) { ctx->err = "CFL maximum not finite or above 2^20 (CHECK_CFL)"; return WAM_E_CFL; }
    ctx->counter = 1;
    if (cmax > 0.999f) {
        ctx->counter = (int)std::ceil(cmax);
        if (ctx->coef) {
            hipLaunchKernelGGL(rescale_kernel, dim3(4096), dim3(256), 0, st, ctx->coef, nb, ctx->ncoef,
                               (float)ctx->counter);
            HIPCHK(hipGetLastError());
        }
    }
    if (t.cfg.shallow) {
        HIPCHK(hipMalloc(&p, (size_t)t.ML * ns * sizeof(float)));
        o.push_back(p);
        ctx->dt.SBOT = (const float *)p;
        hipLaunchKernelGGL(sbottom_factor_kernel, dim3((ns + 255) / 256, t.ML), dim3(256), 0, st, ctx->dt, (float *)p);
        HIPCHK(hipGetLastError());
    }
    if (t.cfg.iphys == 2) {
        HIPCHK(hipMalloc(&p, 2 * (size_t)t.ML * ns * sizeof(float)));
        o.push_back(p);
        ctx->wn = (float *)p; ctx->cgg = ctx->wn + (size_t)t.ML * ns;
        hipLaunchKernelGGL(wavnu2_kernel, dim3((ns + 255) / 256, t.ML), dim3(256), 0, st, ctx->dt, ctx->wn, ctx->cgg);
        HIPCHK(hipGetLastError());
        ctx->dt.WN = ctx->wn; ctx->dt.CGG = ctx->cgg;
    }
    HIPCHK(hipStreamSynchronize(st));
    ctx->prepared = true;
    return WAM_OK;
}

int wam_gpu_prepare_propagation(wam_ctx *ctx, const float *depth, const float *u, const float *v)
{
    if (!ctx || !depth) return WAM_E_ARG;
    Tables &t = ctx->tab;
    ctx->prepared = false;
    int rc = prepare_points(t, depth, u, v, ctx->err);
    if (rc) return rc;
    free_list(ctx->point_allocs);
    ctx->coef = nullptr; ctx->wn = nullptr; ctx->cgg = nullptr; ctx->dt.SBOT = nullptr;
    ctx->adj_coef = nullptr; ctx->adj_tw = nullptr;
    const int ns = t.NSEA;
    auto &o = ctx->point_allocs;
    std::vector<float> d(depth, depth + ns), uu(ns, 0.f), vv(ns, 0.f);
    ctx->depth_h = d;
    if (u) uu.assign(u, u + ns);
    if (v) vv.assign(v, v + ns);
    UP(o, d, DEPTH); UP(o, uu, U); UP(o, vv, V);
    UP(o, t.INDEP, INDEP); UP(o, t.INDEP_G, INDEP_G); UP(o, t.DELLA0, DELLA0); UP(o, t.DRGM, DRGM);
    UP(o, t.DDPHI, DDPHI); UP(o, t.DDLAM, DDLAM); UP(o, t.DUPHI, DUPHI); UP(o, t.DULAM, DULAM);
    UP(o, t.DVPHI, DVPHI); UP(o, t.DVLAM, DVLAM);
    return finish_prepare(ctx);
}

int wam_gpu_prepare_propagation_dev(wam_ctx *ctx, const float *d_depth, const float *d_u, const float *d_v)
{
    if (!ctx || !d_depth) return WAM_E_ARG;
    Tables &t = ctx->tab;
    ctx->prepared = false;
    if (t.cfg.refr_c && (!d_u || !d_v)) { ctx->err = "current refraction needs u and v"; return WAM_E_ARG; }
    free_list(ctx->point_allocs);
    ctx->coef = nullptr; ctx->wn = nullptr; ctx->cgg = nullptr; ctx->dt.SBOT = nullptr;
    ctx->adj_coef = nullptr; ctx->adj_tw = nullptr;
    const int ns = t.NSEA;
    auto &o = ctx->point_allocs;
    hipStream_t st = ctx->stream;
    // the context's own copies of the fields, then the per-point tables
    auto dev = [&](size_t bytes, void **q) -> int {
        HIPCHK(hipMalloc(q, bytes ? bytes : 4));
        o.push_back(*q);
        return WAM_OK;
    };
    float *dd = nullptr, *du = nullptr, *dv = nullptr, *f[8] = {};
    int *ind = nullptr, *indg = nullptr;
    unsigned *bad = nullptr;
    int r = dev(ns * 4, (void **)&dd) | dev(ns * 4, (void **)&du) | dev(ns * 4, (void **)&dv) |
            dev(ns * 4, (void **)&ind) | dev(ns * 4, (void **)&indg) | dev(8, (void **)&bad);
    for (int i = 0; i < 8; i++) r |= dev(ns * 4, (void **)&f[i]);
    if (r) return r;
    HIPCHK(hipMemcpyAsync(dd, d_depth, ns * 4, hipMemcpyDeviceToDevice, st));
    if (d_u) HIPCHK(hipMemcpyAsync(du, d_u, ns * 4, hipMemcpyDeviceToDevice, st));
    else HIPCHK(hipMemsetAsync(du, 0, ns * 4, st));
    if (d_v) HIPCHK(hipMemcpyAsync(dv, d_v, ns * 4, hipMemcpyDeviceToDevice, st));
    else HIPCHK(hipMemsetAsync(dv, 0, ns * 4, st));
    HIPCHK(hipMemsetAsync(bad, 0, 8, st));
    PrepPoints pp{dd, du, dv, ind, indg, f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], bad,
                  (float)t.cfg.deptha, wam_logf((float)t.cfg.depthd), (float)t.cfg.idelpro};
    DevTables &d = ctx->dt;
    d.DEPTH = dd; d.U = du; d.V = dv; d.INDEP = ind; d.INDEP_G = indg;
    d.DDPHI = f[0]; d.DDLAM = f[1]; d.DUPHI = f[2]; d.DULAM = f[3]; d.DVPHI = f[4]; d.DVLAM = f[5];
    d.DELLA0 = f[6]; d.DRGM = f[7];
    hipLaunchKernelGGL(prep_points_kernel, dim3((ns + 255) / 256), dim3(256), 0, st, d, pp);
    HIPCHK(hipGetLastError());
    // the host copies of the per-point tables (INDEP is IMPLSCH's argument on
    // the host path), and the depth check
    unsigned hbad[2] = {0u, 0u};
    t.INDEP.resize(ns); t.INDEP_G.resize(ns);
    std::vector<float> *hf[8] = {&t.DDPHI, &t.DDLAM, &t.DUPHI, &t.DULAM, &t.DVPHI, &t.DVLAM, &t.DELLA0, &t.DRGM};
    ctx->depth_h.resize(ns);
    HIPCHK(hipMemcpyAsync(ctx->depth_h.data(), dd, ns * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hbad, bad, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(t.INDEP.data(), ind, ns * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(t.INDEP_G.data(), indg, ns * 4, hipMemcpyDeviceToHost, st));
    for (int i = 0; i < 8; i++) {
        hf[i]->resize(ns);
        HIPCHK(hipMemcpyAsync(hf[i]->data(), f[i], ns * 4, hipMemcpyDeviceToHost, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    if (hbad[0]) { ctx->err = "depth must be positive and finite at every sea point"; return WAM_E_ARG; }
    if (hbad[1]) { ctx->err = "currents must be finite at every sea point"; return WAM_E_ARG; }
    return finish_prepare(ctx);
}

// Restart records from and to device memory, through two pinned host buffers:
// the copy of one piece overlaps the file transfer of the other.
namespace {
struct Staging {
    static constexpr size_t CHUNK = (size_t)64 << 20;
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~Staging()
    {
        for (int i = 0; i < 2; i++) {
            if (ev[i]) (void)hipEventDestroy(ev[i]);
            if (buf[i]) (void)hipHostFree(buf[i]);
        }
    }
    int init()
    {
        for (int i = 0; i < 2; i++) {
            if (hipHostMalloc(&buf[i], CHUNK) != hipSuccess) return WAM_E_HIP;
            if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) return WAM_E_HIP;
        }
        return WAM_OK;
    }
};

int put_dev(Staging &sg, hipStream_t st, wam::RecordWriter &w, const void *src, uint64_t n)
{
    const char *s = static_cast<const char *>(src);
    const size_t C = Staging::CHUNK;
    const uint64_t nch = (n + C - 1) / C;
    if (nch == 0) return WAM_OK;
    auto issue = [&](uint64_t i) {
        const size_t k = (size_t)std::min<uint64_t>(C, n - i * C);
        if (hipMemcpyAsync(sg.buf[i % 2], s + i * C, k, hipMemcpyDefault, st) != hipSuccess) return false;
        return hipEventRecord(sg.ev[i % 2], st) == hipSuccess;
    };
    if (!issue(0)) return WAM_E_HIP;
    for (uint64_t i = 0; i < nch; i++) {
        if (i + 1 < nch && !issue(i + 1)) return WAM_E_HIP;
        if (hipEventSynchronize(sg.ev[i % 2]) != hipSuccess) return WAM_E_HIP;
        const size_t k = (size_t)std::min<uint64_t>(C, n - i * C);
        if (!w.put(sg.buf[i % 2], k)) return WAM_E_IO;
    }
    return WAM_OK;
}

int get_dev(Staging &sg, hipStream_t st, wam::RecordReader &r, void *dst, uint64_t n)
{
    char *d = static_cast<char *>(dst);
    const size_t C = Staging::CHUNK;
    for (uint64_t i = 0; i * C < n; i++) {
        const size_t k = (size_t)std::min<uint64_t>(C, n - i * C);
        // the buffer's previous copy must be done before it is refilled
        if (i >= 2 && hipEventSynchronize(sg.ev[i % 2]) != hipSuccess) return WAM_E_HIP;
        if (!r.get(sg.buf[i % 2], k)) return WAM_E_IO;
        if (hipMemcpyAsync(d + i * C, sg.buf[i % 2], k, hipMemcpyDefault, st) != hipSuccess) return WAM_E_HIP;
        if (hipEventRecord(sg.ev[i % 2], st) != hipSuccess) return WAM_E_HIP;
    }
    return hipStreamSynchronize(st) == hipSuccess ? WAM_OK : WAM_E_HIP;
}

struct FileCloser {
    FILE *f = nullptr;
    ~FileCloser() { if (f) std::fclose(f); }
};
}  // namespace

int wam_gpu_save_restart(wam_ctx *ctx, const char *path, const char *dates, const float *d_fl3, const float *d_u10,
                         const float *d_udir, const float *d_tauw)
{
    if (!ctx || !path || !dates || !d_fl3 || !d_u10 || !d_udir || !d_tauw) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    const bool with_depth = !wam::date_blank(dates + 42), with_cur = !wam::date_blank(dates + 56);
    if ((with_depth || with_cur) && !ctx->prepared) {
        ctx->err = "depth or currents requested before PREPARE_PROPAGATION";
        return WAM_E_STATE;
    }
    Staging sg;
    int rc = sg.init();
    if (rc) { ctx->err = "pinned staging buffers"; return rc; }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    FileCloser fc;
    fc.f = std::fopen(path, "wb");
    if (!fc.f) { ctx->err = std::string("cannot open ") + path; return WAM_E_IO; }
    wam::RecordWriter w(fc.f);
    const int32_t nsea = t.NSEA;
    char head[74];
    std::memcpy(head, &nsea, 4);
    std::memcpy(head + 4, dates, 70);
    const uint64_t n = (uint64_t)nsea * 4, nb = n * (uint64_t)t.KL * (uint64_t)t.ML;
    hipStream_t st = ctx->stream;
    auto rec = [&](const void *p, uint64_t bytes) {
        if (!w.begin(bytes)) return (int)WAM_E_IO;
        int r = put_dev(sg, st, w, p, bytes);
        if (r) return r;
        return w.end() ? (int)WAM_OK : (int)WAM_E_IO;
    };
    if (!(w.begin(74) && w.put(head, 74) && w.end())) rc = WAM_E_IO;
    if (!rc) rc = rec(d_u10, n);
    if (!rc) rc = rec(d_udir, n);
    if (!rc) rc = rec(d_tauw, n);
    if (!rc) rc = rec(d_fl3, nb);
    if (!rc && with_depth) rc = rec(ctx->dt.DEPTH, n);
    if (!rc && with_cur) {
        if (!w.begin(2 * n)) rc = WAM_E_IO;
        if (!rc) rc = put_dev(sg, st, w, ctx->dt.U, n);
        if (!rc) rc = put_dev(sg, st, w, ctx->dt.V, n);
        if (!rc && !w.end()) rc = WAM_E_IO;
    }
    if (!rc) {
        const int c = std::fclose(fc.f);
        fc.f = nullptr;
        if (c != 0) rc = WAM_E_IO;
    }
    if (rc) ctx->err = rc == WAM_E_IO ? std::string("cannot write ") + path : "HIP copy to the staging buffers";
    return rc;
}

int wam_gpu_connect_restart(wam_ctx *ctx, const char *path, char *dates, float *d_fl3, float *d_u10, float *d_udir,
                            float *d_tauw, float *d_depth, float *d_u, float *d_v, int32_t *present)
{
    if (!ctx || !path || !dates || !d_fl3 || !d_u10 || !d_udir || !d_tauw) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    Staging sg;
    int rc = sg.init();
    if (rc) { ctx->err = "pinned staging buffers"; return rc; }
    FileCloser fc;
    fc.f = std::fopen(path, "rb");
    if (!fc.f) { ctx->err = std::string("cannot open ") + path; return WAM_E_IO; }
    wam::RecordReader r(fc.f);
    char head[74];
    if (!(r.begin(74) && r.get(head, 74) && r.end())) { ctx->err = "restart header record"; return WAM_E_IO; }
    int32_t nsea_r;
    std::memcpy(&nsea_r, head, 4);
    if (nsea_r != t.NSEA) {
        ctx->err = "restart fields are inconsistent with the model grid (NSEA)";
        return WAM_E_IO;
    }
    std::memcpy(dates, head + 4, 70);
    const bool with_depth = !wam::date_blank(dates + 42), with_cur = !wam::date_blank(dates + 56);
    const uint64_t n = (uint64_t)t.NSEA * 4, nb = n * (uint64_t)t.KL * (uint64_t)t.ML;
    hipStream_t st = ctx->stream;
    auto rec = [&](void *p, uint64_t bytes) {
        if (!r.begin(bytes)) return (int)WAM_E_IO;
        int e = p ? get_dev(sg, st, r, p, bytes) : (r.skip(bytes) ? (int)WAM_OK : (int)WAM_E_IO);
        if (e) return e;
        return r.end() ? (int)WAM_OK : (int)WAM_E_IO;
    };
    rc = rec(d_u10, n);
    if (!rc) rc = rec(d_udir, n);
    if (!rc) rc = rec(d_tauw, n);
    if (!rc) rc = rec(d_fl3, nb);
    if (!rc && with_depth) rc = rec(d_depth, n);
    if (!rc && with_cur) {
        if (!r.begin(2 * n)) rc = WAM_E_IO;
        if (!rc) rc = d_u && d_v ? get_dev(sg, st, r, d_u, n) : (r.skip(n) ? WAM_OK : WAM_E_IO);
        if (!rc) rc = d_u && d_v ? get_dev(sg, st, r, d_v, n) : (r.skip(n) ? WAM_OK : WAM_E_IO);
        if (!rc && !r.end()) rc = WAM_E_IO;
    }
    if (rc) {
        ctx->err = rc == WAM_E_IO ? std::string("cannot read ") + path : "HIP copy from the staging buffers";
        return rc;
    }
    if (present) *present = (with_depth ? 1 : 0) | (with_cur ? 2 : 0);
    return WAM_OK;
}

// PROPAGS for nmem members (members on blockIdx.y, the weights and per-point
// fields shared), COUNTER sub-steps alternating between d_out and scratch so
// that the last one lands in d_out.
static int propags_members(wam_ctx *ctx, int nmem, const float *d_in, float *d_out, float *scratch, bool timed)
{
    const Tables &t = ctx->tab;
    const int wg = ctx->tune.prop_wg;
    const int nch = (t.NSEA + wg - 1) / wg;
    const dim3 grid(((nch + 7) / 8) * 8, nmem), block(wg);
    // KPM is cyclic (K+1, K-1 modulo KL), so for the common direction counts
    // the register-row form of PROP_CURR applies
    auto kern = propags_nocurr_kernel;
    if (t.cfg.refr_c)
        kern = t.KL == 36 ? propags_curr_rows_kernel<36> : t.KL == 24 ? propags_curr_rows_kernel<24> : propags_curr_kernel;
    const float *src = d_in;
    for (int i = 0; i < ctx->counter; i++) {
        const int left = ctx->counter - 1 - i;
        float *dst = (left % 2 == 0) ? d_out : scratch;
        hipEvent_t ev[2] = {nullptr, nullptr};
        if (timed) {
            const int rc = timing_events(ctx, ev, 2);
            if (rc) return rc;
        }
        timing_record(ctx, ev[0]);
        if (ctx->coef)
            hipLaunchKernelGGL(kern, grid, block, 0, ctx->stream, ctx->dt, ctx->coef, src, dst);
        else if (t.cfg.refr_c && t.cfg.shallow && 6ull * 4 * t.KL * t.NSEA < (1ull << 32)) {
            // (32-bit buffer offsets over three frequency planes, and land
            // neighbours' offsets past them)
            // in register blocks of three directions where KL allows, else a direction at a time
            const int c = ctx->counter;
            const int rs = c == 1 ? 0 : (c & (c - 1)) == 0 ? 1 : 2;
            auto k = t.KL % 3 == 0 ? (rs == 0 ? propags_blk_kernel<0, 3> : rs == 1 ? propags_blk_kernel<1, 3>
                                                                              : propags_blk_kernel<2, 3>)
                                   : (rs == 0 ? propags_fused_sc_kernel<0> : rs == 1 ? propags_fused_sc_kernel<1>
                                                                                : propags_fused_sc_kernel<2>);
            hipLaunchKernelGGL(k, grid, block, 0, ctx->stream, ctx->dt, c, src, dst);
        }
        else if (t.cfg.refr_c)
            hipLaunchKernelGGL(propags_fused_kernel<true>, grid, block, 0, ctx->stream, ctx->dt, ctx->counter, src, dst);
        else
            hipLaunchKernelGGL(propags_fused_kernel<false>, grid, block, 0, ctx->stream, ctx->dt, ctx->counter, src, dst);
        HIPCHK(hipGetLastError());
        timing_record(ctx, ev[1]);
        timing_keep(ctx, "propags", ev[0], ev[1]);
        src = dst;
    }
    return WAM_OK;
}

int wam_gpu_propags_dev(wam_ctx *ctx, const float *d_in, float *d_out)
{
    if (!ctx || !d_in || !d_out || d_in == d_out) return WAM_E_ARG;
    if (!ctx->prepared) { ctx->err = "PREPARE_PROPAGATION not done"; return WAM_E_STATE; }
    const size_t nb = (size_t)ctx->tab.KL * ctx->tab.ML * ctx->tab.NSEA;
    if (ctx->counter > 1 && !ctx->fscratch) HIPCHK(hipMalloc(&ctx->fscratch, nb * sizeof(float)));
    return propags_members(ctx, 1, d_in, d_out, ctx->fscratch, true);
}

// Work arrays of the three-kernel IMPLSCH for cap points (member-points) per
// launch sequence: tiled F, SL, FL (and SL - SPOS with LCFLX), STRESSO's
// direction sums, per-point scalars, SPRD, LLWS bits, ticket counters.
// Reallocated when the capacity changes.
static void free_impl_scratch(ImplScratch &x)
{
    for (void *p : {(void *)x.ft, (void *)x.slt, (void *)x.flt, (void *)x.smt, (void *)x.axy, (void *)x.pp,
                    (void *)x.sprd, (void *)x.llws, (void *)x.tick, (void *)x.st6})
        if (p) (void)hipFree(p);
    x = ImplScratch{};
}

static int ensure_impl_scratch(wam_ctx *ctx, ImplScratch &x, size_t cap, int lcflx)
{
    const Tables &t = ctx->tab;
    const size_t nb = (size_t)t.KL * t.ML, nw = (nb + 31) / 32;
    if (x.ft && (size_t)x.NX != cap) free_impl_scratch(x);
    const size_t tiled = (cap + 63) / 64 * 64 * nb;      // whole 64-point tiles
    auto failed = [&](hipError_t e) {
        (void)hipGetLastError();
        ctx->err = std::string("IMPLSCH work arrays: ") + hipGetErrorString(e);
        return WAM_E_HIP;
    };
    if (!x.ft) {
        // all or nothing: a failed allocation leaves no partial set behind
        ImplScratch n{};
        n.NB4 = (int)(nb / 4);
        n.P0 = 0; n.NP = (int)cap; n.NX = (int)cap;
        const struct { void **p; size_t bytes; } req[] = {
            {(void **)&n.ft, tiled * sizeof(float)}, {(void **)&n.slt, tiled * sizeof(float)},
            {(void **)&n.flt, tiled * sizeof(float)}, {(void **)&n.axy, 3 * (size_t)t.ML * cap * sizeof(float)},
            {(void **)&n.pp, 8 * cap * sizeof(float)}, {(void **)&n.sprd, (size_t)t.KL * cap * sizeof(float)},
            {(void **)&n.llws, nw * cap * sizeof(unsigned)}, {(void **)&n.tick, 8 * 32 * sizeof(unsigned)},
            {(void **)&n.st6, t.cfg.iphys == 2 ? 6 * (size_t)t.ML * cap * sizeof(float) : 0}};
        for (const auto &r : req) {
            if (!r.bytes) continue;
            const hipError_t e = hipMalloc(r.p, r.bytes);
            if (e != hipSuccess) {
                *r.p = nullptr;
                free_impl_scratch(n);
                return failed(e);
            }
        }
        x = n;
    }
    if (lcflx && !x.smt) {
        const hipError_t e = hipMalloc(&x.smt, tiled * sizeof(float));
        if (e != hipSuccess) {
            x.smt = nullptr;
            return failed(e);
        }
    }
    return WAM_OK;
}

static void free_adj_extra(AdjExtra &e)
{
    for (void *p : {(void *)e.fn, (void *)e.llws1, (void *)e.cu, (void *)e.cub, (void *)e.sc})
        if (p) (void)hipFree(p);
    e = AdjExtra{};
}

// The split IMPLSCH adjoint's work arrays for cap member-points: those of
// the forward split (ctx->asx) and AdjExtra's.
static int ensure_adj_scratch(wam_ctx *ctx, size_t cap)
{
    const Tables &t = ctx->tab;
    const size_t nb = (size_t)t.KL * t.ML;
    if (ctx->asx.ft && (size_t)ctx->asx.NX != cap) free_adj_extra(ctx->aex);
    int rc = ensure_impl_scratch(ctx, ctx->asx, cap, 0);
    if (rc) return rc;
    if (!ctx->aex.fn) {
        const size_t tiled = (cap + 63) / 64 * 64 * nb;
        AdjExtra n{};
        const struct { void **p; size_t bytes; } req[] = {
            {(void **)&n.fn, tiled * sizeof(float)}, {(void **)&n.llws1, nb * cap},
            {(void **)&n.cu, (size_t)t.KL * cap * sizeof(float)}, {(void **)&n.cub, (size_t)t.KL * cap * sizeof(float)},
            {(void **)&n.sc, (size_t)AS_NSC * cap * sizeof(float)}};
        for (const auto &r : req) {
            const hipError_t e = hipMalloc(r.p, r.bytes);
            if (e != hipSuccess) {
                *r.p = nullptr;
                free_adj_extra(n);
                (void)hipGetLastError();
                ctx->err = std::string("adjoint work arrays: ") + hipGetErrorString(e);
                return WAM_E_HIP;
            }
        }
        ctx->aex = n;
    }
    return WAM_OK;
}

// Bytes of those work arrays per point.
static size_t impl_scratch_per_point(const Tables &t, int lcflx)
{
    const size_t nb = (size_t)t.KL * t.ML, nw = (nb + 31) / 32;
    return ((3 + (lcflx ? 1 : 0)) * nb + 3 * (size_t)t.ML + 8 + t.KL + nw + (t.cfg.iphys == 2 ? 6 * (size_t)t.ML : 0)) * 4;
}

// The member-points per IMPLSCH launch sequence of an ensemble: the tuned
// value, else all of them up to 2^20.
static size_t ensemble_chunk(const wam_ctx *ctx, size_t nmp)
{
    size_t c = ctx->tune.chunk > 0 ? (size_t)ctx->tune.chunk : ((size_t)1 << 20);
    return std::min(c, nmp);
}

static ImplArgs impl_args(float *d_fl3, const float *d_u10, const float *d_udir, float *d_tauw, float *d_ustar,
                          float *d_z0, const float *d_roairn, const float *d_wstar, int32_t lcflx, float *d_phiaw,
                          float *d_fluxes, int32_t *d_mij, int nmp)
{
    ImplArgs a{};
    a.fl3 = d_fl3; a.u10 = d_u10; a.udir = d_udir; a.tauw = d_tauw; a.ustar = d_ustar; a.z0 = d_z0;
    a.roairn = d_roairn; a.wstar = d_wstar; a.lcflx = lcflx ? 1 : 0; a.phiaw = d_phiaw; a.fluxes = d_fluxes;
    a.mij = d_mij; a.NMP = nmp;
    return a;
}

// IMPLSCH for nmem members (a.NMP = nmem * NSEA member-points): the three
// kernels over chunks of member-points whose work arrays x holds (allocated
// beforehand), or, for the configurations of the single kernel, one launch
// per member.
static int implsch_members(wam_ctx *ctx, int nmem, const ImplArgs &a, ImplScratch &x, bool timed)
{
    const Tables &t = ctx->tab;
    if (implsch_split_supported(ctx->dt)) {
        const int nmp = a.NMP, cap = x.NX;
        for (int p0 = 0; p0 < nmp; p0 += cap) {
            x.P0 = p0;
            x.NP = std::min(cap, nmp - p0);
            hipEvent_t ev[4] = {};
            if (timed) {
                const int rc = timing_events(ctx, ev, 4);
                if (rc) return rc;
            }
            HIPCHK(launch_implsch_split(ctx->dt, a, x, ctx->ncu, ctx->stream, ev, ctx->tune));
            timing_keep(ctx, "implsch_pre", ev[0], ev[1]);
            timing_keep(ctx, "implsch_snl", ev[1], ev[2]);
            timing_keep(ctx, "implsch_post", ev[2], ev[3]);
        }
        x.P0 = 0;
        x.NP = x.NX;
        return WAM_OK;
    }
    const size_t nb = (size_t)t.KL * t.ML * t.NSEA, ns = t.NSEA;
    for (int m = 0; m < nmem; m++) {
        ImplArgs am = a;                 // member m: its own arrays; the flux rows stay NMP long
        am.fl3 = a.fl3 + m * nb;
        am.u10 = a.u10 + m * ns; am.udir = a.udir + m * ns; am.roairn = a.roairn + m * ns;
        am.wstar = a.wstar + m * ns; am.tauw = a.tauw + m * ns; am.ustar = a.ustar + m * ns; am.z0 = a.z0 + m * ns;
        if (a.phiaw) am.phiaw = a.phiaw + m * ns;
        if (a.fluxes) am.fluxes = a.fluxes + m * ns;
        if (a.mij) am.mij = a.mij + m * ns;
        hipEvent_t ev[2] = {};
        if (timed) {
            const int rc = timing_events(ctx, ev, 2);
            if (rc) return rc;
        }
        timing_record(ctx, ev[0]);
        HIPCHK(launch_implsch(ctx->dt, am, ctx->ncu, ctx->stream));
        timing_record(ctx, ev[1]);
        timing_keep(ctx, "implsch", ev[0], ev[1]);
    }
    return WAM_OK;
}

int wam_gpu_implsch_dev(wam_ctx *ctx, float *d_fl3, const float *d_u10, const float *d_udir, float *d_tauw,
                        float *d_ustar, float *d_z0, const float *d_roairn, const float *d_wstar, int32_t lcflx,
                        float *d_phiaw, float *d_fluxes, int32_t *d_mij)
{
    if (!ctx || !d_fl3 || !d_u10 || !d_udir || !d_tauw || !d_ustar || !d_z0 || !d_roairn || !d_wstar) return WAM_E_ARG;
    if (!ctx->prepared) { ctx->err = "PREPARE_PROPAGATION not done"; return WAM_E_STATE; }
    const ImplArgs a = impl_args(d_fl3, d_u10, d_udir, d_tauw, d_ustar, d_z0, d_roairn, d_wstar, lcflx, d_phiaw,
                                 d_fluxes, d_mij, ctx->tab.NSEA);
    if (implsch_split_supported(ctx->dt)) {
        const int rc = ensure_impl_scratch(ctx, ctx->isx, ctx->tab.NSEA, a.lcflx);
        if (rc) return rc;
    }
    return implsch_members(ctx, 1, a, ctx->isx, true);
}

// ---- ensembles ------------------------------------------------------------

static int ensemble_args_ok(wam_ctx *ctx, int32_t nmem)
{
    if (!ctx) return WAM_E_ARG;
    if (!ctx->prepared) { ctx->err = "PREPARE_PROPAGATION not done"; return WAM_E_STATE; }
    if (nmem < 1 || (int64_t)nmem * ctx->tab.NSEA > 0x7fffffff || nmem > 65535) {
        ctx->err = "nmem must be >= 1 with nmem * NSEA < 2^31";
        return WAM_E_ARG;
    }
    return WAM_OK;
}

// the COUNTER > 1 ping-pong spectra of nmem members
static int ensure_ensemble_fscratch(wam_ctx *ctx, int nmem)
{
    if (ctx->counter <= 1) return WAM_OK;
    const size_t n = (size_t)nmem * ctx->tab.KL * ctx->tab.ML * ctx->tab.NSEA;
    if (ctx->fscratch_e_n >= n) return WAM_OK;
    if (ctx->fscratch_e) (void)hipFree(ctx->fscratch_e);
    ctx->fscratch_e = nullptr; ctx->fscratch_e_n = 0;
    HIPCHK(hipMalloc(&ctx->fscratch_e, n * sizeof(float)));
    ctx->fscratch_e_n = n;
    return WAM_OK;
}

int wam_gpu_propags_ensemble_dev(wam_ctx *ctx, int32_t nmem, const float *d_in, float *d_out)
{
    int rc = ensemble_args_ok(ctx, nmem);
    if (rc) return rc;
    if (!d_in || !d_out || d_in == d_out) return WAM_E_ARG;
    rc = ensure_ensemble_fscratch(ctx, nmem);
    if (rc) return rc;
    return propags_members(ctx, nmem, d_in, d_out, ctx->fscratch_e, true);
}

int wam_gpu_implsch_ensemble_dev(wam_ctx *ctx, int32_t nmem, float *d_fl3, const float *d_u10, const float *d_udir,
                                 float *d_tauw, float *d_ustar, float *d_z0, const float *d_roairn,
                                 const float *d_wstar, int32_t lcflx, float *d_phiaw, float *d_fluxes,
                                 int32_t *d_mij)
{
    int rc = ensemble_args_ok(ctx, nmem);
    if (rc) return rc;
    if (!d_fl3 || !d_u10 || !d_udir || !d_tauw || !d_ustar || !d_z0 || !d_roairn || !d_wstar) return WAM_E_ARG;
    const int nmp = nmem * ctx->tab.NSEA;
    const ImplArgs a = impl_args(d_fl3, d_u10, d_udir, d_tauw, d_ustar, d_z0, d_roairn, d_wstar, lcflx, d_phiaw,
                                 d_fluxes, d_mij, nmp);
    if (implsch_split_supported(ctx->dt)) {
        rc = ensure_impl_scratch(ctx, ctx->esx, ensemble_chunk(ctx, nmp), a.lcflx);
        if (rc) return rc;
    }
    return implsch_members(ctx, nmem, a, ctx->esx, true);
}

// nsteps model steps (PROPAGS d_fa -> d_fb, IMPLSCH on d_fb, then the roles
// swapped) on the context stream.
static int ensemble_steps(wam_ctx *ctx, int nmem, int nsteps, float *d_fa, float *d_fb, const ImplArgs &a0,
                          bool timed)
{
    float *cur = d_fa, *nxt = d_fb;
    for (int i = 0; i < nsteps; i++) {
        int rc = propags_members(ctx, nmem, cur, nxt, ctx->fscratch_e, timed);
        if (rc) return rc;
        ImplArgs a = a0;
        a.fl3 = nxt;
        rc = implsch_members(ctx, nmem, a, ctx->esx, timed);
        if (rc) return rc;
        std::swap(cur, nxt);
    }
    return WAM_OK;
}

int wam_gpu_ensemble_run_dev(wam_ctx *ctx, int32_t nmem, int32_t nsteps, float *d_fa, float *d_fb,
                             const float *d_u10, const float *d_udir, float *d_tauw, float *d_ustar, float *d_z0,
                             const float *d_roairn, const float *d_wstar, int32_t graph)
{
    int rc = ensemble_args_ok(ctx, nmem);
    if (rc) return rc;
    if (nsteps < 0 || !d_fa || !d_fb || d_fa == d_fb || !d_u10 || !d_udir || !d_tauw || !d_ustar || !d_z0 ||
        !d_roairn || !d_wstar)
        return WAM_E_ARG;
    if (nsteps == 0) return WAM_OK;
    const int nmp = nmem * ctx->tab.NSEA;
    const ImplArgs a = impl_args(d_fb, d_u10, d_udir, d_tauw, d_ustar, d_z0, d_roairn, d_wstar, 0, nullptr, nullptr,
                                 nullptr, nmp);
    // every allocation before a capture
    rc = ensure_ensemble_fscratch(ctx, nmem);
    if (rc) return rc;
    if (implsch_split_supported(ctx->dt)) {
        rc = ensure_impl_scratch(ctx, ctx->esx, ensemble_chunk(ctx, nmp), 0);
        if (rc) return rc;
    }
    if (!graph) return ensemble_steps(ctx, nmem, nsteps, d_fa, d_fb, a, true);

    // The captured loop: built for these arguments, launch shapes and
    // sub-step count, replayed while they stay the same.
    const std::vector<uintptr_t> key = {
        (uintptr_t)nmem, (uintptr_t)nsteps, (uintptr_t)d_fa, (uintptr_t)d_fb, (uintptr_t)d_u10, (uintptr_t)d_udir,
        (uintptr_t)d_tauw, (uintptr_t)d_ustar, (uintptr_t)d_z0, (uintptr_t)d_roairn, (uintptr_t)d_wstar,
        (uintptr_t)ctx->tune.prop_wg, (uintptr_t)ctx->tune.src_wg, (uintptr_t)ctx->tune.snl_per_cu,
        (uintptr_t)ctx->esx.NX, (uintptr_t)ctx->counter, (uintptr_t)ctx->dt.DEPTH, (uintptr_t)ctx->coef,
        (uintptr_t)(ctx->dt.brk | ctx->dt.phil << 1 | ctx->dt.split << 2)};
    if (!ctx->gstream) {
        HIPCHK(hipStreamCreateWithFlags(&ctx->gstream, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&ctx->gev[0], hipEventDisableTiming));
        HIPCHK(hipEventCreateWithFlags(&ctx->gev[1], hipEventDisableTiming));
    }
    if (!ctx->gexec || ctx->gkey != key) {
        if (ctx->gexec) { (void)hipGraphExecDestroy(ctx->gexec); ctx->gexec = nullptr; }
        if (ctx->graph) { (void)hipGraphDestroy(ctx->graph); ctx->graph = nullptr; }
        hipStream_t keep = ctx->stream;
        ctx->stream = ctx->gstream;
        HIPCHK(hipStreamBeginCapture(ctx->gstream, hipStreamCaptureModeRelaxed));
        rc = ensemble_steps(ctx, nmem, nsteps, d_fa, d_fb, a, false);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(ctx->gstream, &g);
        ctx->stream = keep;
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        HIPCHK(e);
        ctx->graph = g;
        HIPCHK(hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0));
        ctx->gkey = key;
    }
    // ordered after the work already on the context stream, and before what follows
    HIPCHK(hipEventRecord(ctx->gev[0], ctx->stream));
    HIPCHK(hipStreamWaitEvent(ctx->gstream, ctx->gev[0], 0));
    HIPCHK(hipGraphLaunch(ctx->gexec, ctx->gstream));
    HIPCHK(hipEventRecord(ctx->gev[1], ctx->gstream));
    HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->gev[1], 0));
    return WAM_OK;
}

// ---- autotuner ------------------------------------------------------------

// Milliseconds of fn() on the context stream, the best of reps runs.
static int time_best(wam_ctx *ctx, int reps, const std::function<int()> &fn, float &best)
{
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    best = 3.4e38f;
    int rc = WAM_OK;
    for (int r = 0; r < reps && !rc; r++) {
        (void)hipEventRecord(e0, ctx->stream);
        rc = fn();
        (void)hipEventRecord(e1, ctx->stream);
        if (rc) break;
        if (hipEventSynchronize(e1) != hipSuccess) { rc = WAM_E_HIP; break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        best = std::min(best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int wam_gpu_autotune(wam_ctx *ctx, int32_t nmem, const float *d_fa, float *d_fb, const float *d_u10,
                     const float *d_udir, const float *d_tauw, const float *d_z0, const float *d_roairn,
                     const float *d_wstar, int32_t reps)
{
    int rc = ensemble_args_ok(ctx, nmem);
    if (rc) return rc;
    if (!d_fa || !d_fb || d_fa == d_fb || !d_u10 || !d_udir || !d_tauw || !d_z0 || !d_roairn || !d_wstar)
        return WAM_E_ARG;
    reps = std::max(1, reps);
    const Tables &t = ctx->tab;
    const size_t ns = t.NSEA, nmp = (size_t)nmem * ns;
    rc = ensure_ensemble_fscratch(ctx, nmem);
    if (rc) return rc;
    ctx->tune_log.clear();
    // PROPAGS: points per workgroup
    {
        int best_wg = ctx->tune.prop_wg;
        float best = 3.4e38f;
        for (int wg : {64, 128, 256}) {
            ctx->tune.prop_wg = wg;
            float ms = 0.f;
            rc = time_best(ctx, reps, [&] { return propags_members(ctx, nmem, d_fa, d_fb, ctx->fscratch_e, false); },
                           ms);
            if (rc) return rc;
            ctx->tune_log.push_back({"prop_wg", wg, ms});
            if (ms < best) { best = ms; best_wg = wg; }
        }
        ctx->tune.prop_wg = best_wg;
    }
    if (!implsch_split_supported(ctx->dt)) return WAM_OK;     // the single kernel has no launch shapes to tune
    // IMPLSCH on the propagated state in d_fb, from the given TAUW and Z0
    // every time (copies; USTAR, TAUW, Z0 of the runs are discarded)
    float *tmp = nullptr;
    HIPCHK(hipMalloc(&tmp, 3 * nmp * sizeof(float)));
    float *tw = tmp, *z0 = tmp + nmp, *us = tmp + 2 * nmp;
    auto run = [&]() -> int {
        int r = propags_members(ctx, nmem, d_fa, d_fb, ctx->fscratch_e, false);
        if (r) return r;
        HIPCHK(hipMemcpyAsync(tw, d_tauw, nmp * 4, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(z0, d_z0, nmp * 4, hipMemcpyDeviceToDevice, ctx->stream));
        return WAM_OK;
    };
    const ImplArgs a = impl_args(d_fb, d_u10, d_udir, tw, us, z0, d_roairn, d_wstar, 0, nullptr, nullptr, nullptr,
                                 (int)nmp);
    auto time_implsch = [&](float &best) -> int {
        best = 3.4e38f;
        for (int r = 0; r < reps; r++) {
            int e = run();
            if (e) return e;
            float ms = 0.f;
            e = time_best(ctx, 1, [&] { return implsch_members(ctx, nmem, a, ctx->esx, false); }, ms);
            if (e) return e;
            best = std::min(best, ms);
        }
        return WAM_OK;
    };
    // member-points per launch sequence: powers of two from 2^18 and all of
    // them, whose work arrays fit half of the free memory
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) freeb = (size_t)64 << 30;
    const size_t ppp = impl_scratch_per_point(t, 0);
    std::vector<size_t> chunks;
    for (size_t c = (size_t)1 << 18; c < nmp; c <<= 1) chunks.push_back(c);
    chunks.push_back(nmp);
    std::vector<size_t> fit;
    for (size_t c : chunks)
        if (c * ppp <= freeb / 2 + (ctx->esx.ft ? (size_t)ctx->esx.NX * ppp : 0)) fit.push_back(c);
    if (fit.empty()) fit.push_back(std::min(nmp, (size_t)1 << 18));
    rc = WAM_OK;
    // coordinate search: chunk, then the per-lane kernels' width, then the SNONLIN kernel's occupancy
    float best = 3.4e38f;
    size_t best_c = fit.back() < ((size_t)1 << 20) ? fit.back() : ((size_t)1 << 20);
    for (size_t c : fit) {
        ctx->tune.chunk = (long)c;
        rc = ensure_impl_scratch(ctx, ctx->esx, c, 0);
        float ms = 0.f;
        if (!rc) rc = time_implsch(ms);
        if (rc) break;
        ctx->tune_log.push_back({"chunk", (long)c, ms});
        if (ms < best) { best = ms; best_c = c; }
    }
    if (!rc) {
        ctx->tune.chunk = (long)best_c;
        rc = ensure_impl_scratch(ctx, ctx->esx, best_c, 0);
    }
    if (!rc) {
        int best_w = ctx->tune.src_wg;
        best = 3.4e38f;
        for (int w : {64, 128, 256}) {
            ctx->tune.src_wg = w;
            float ms = 0.f;
            rc = time_implsch(ms);
            if (rc) break;
            ctx->tune_log.push_back({"src_wg", w, ms});
            if (ms < best) { best = ms; best_w = w; }
        }
        ctx->tune.src_wg = best_w;
    }
    if (!rc) {
        int best_o = 0;
        best = 3.4e38f;
        for (int o : {0, 1, 2}) {
            ctx->tune.snl_per_cu = o;
            float ms = 0.f;
            rc = time_implsch(ms);
            if (rc) break;
            ctx->tune_log.push_back({"snl_per_cu", o, ms});
            if (ms < best) { best = ms; best_o = o; }
        }
        ctx->tune.snl_per_cu = best_o;
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    return rc;
}

// ---- several grids, nested -----------------------------------------------

}  // extern "C"

struct wam_multigrid {
    struct Grid { wam_grid_state st; ImplArgs a; };
    struct Link {
        int coarse, fine, nb, nf;
        int *ijarc = nullptr;           // device: [nb] coarse output points, then [3][nf] IJARF, IBFL, IBFR
        float *fbc = nullptr;           // device: [nb][ML][KL] gathered spectra, [3][nb] their mean values, [nf] BFW
        float *mean = nullptr, *bfw = nullptr;
        int *ijarf = nullptr, *ibfl = nullptr, *ibfr = nullptr;
    };
    std::vector<Grid> grids;
    std::vector<Link> links;
    hipStream_t origin = nullptr;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;         // [0] fork, [1..n] joins
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int gsteps = -1, gparity = -1, gopts = -1;
    int parity = 0;                     // swaps done: grid states are in d_fb when odd
    std::string err;
};

static int mg_fail(wam_multigrid *mg, const std::string &m, int rc)
{
    mg->err = m;
    return rc;
}

#define MGCHK(x)                                                                   \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            (void)hipGetLastError();                                               \
            return mg_fail(mg, std::string(#x) + ": " + hipGetErrorString(e_), WAM_E_HIP); \
        }                                                                          \
    } while (0)

// One model step of grid g on stream st (the context's stream redirected).
static int mg_grid_step(wam_multigrid *mg, wam_multigrid::Grid &g, float *cur, float *nxt, hipStream_t st)
{
    wam_ctx *ctx = g.st.ctx;
    hipStream_t keep = ctx->stream;
    ctx->stream = st;
    int rc = propags_members(ctx, 1, cur, nxt, ctx->fscratch, false);
    if (!rc) {
        ImplArgs a = g.a;
        a.fl3 = nxt;
        rc = implsch_members(ctx, 1, a, ctx->isx, false);
    }
    ctx->stream = keep;
    if (rc) mg->err = ctx->err;
    return rc;
}

// The nest exchanges after every grid's step: BOUNDARY_OUTPUT's gather on the
// coarse grid, BOUNDARY_INPUT's space interpolation into the fine grid.
static int mg_links(wam_multigrid *mg, int parity, hipStream_t st)
{
    for (auto &l : mg->links) {
        auto &c = mg->grids[l.coarse], &f = mg->grids[l.fine];
        float *cf = parity ? c.st.d_fb : c.st.d_fa, *ff = parity ? f.st.d_fb : f.st.d_fa;
        MGCHK(launch_boundary_gather(c.st.ctx->dt, cf, l.ijarc, l.nb, l.fbc, l.mean, st));
        MGCHK(launch_boundary_space(f.st.ctx->dt, l.nb, l.fbc, l.mean, l.nf, l.ijarf, l.ibfl, l.ibfr, l.bfw, ff, st));
    }
    return WAM_OK;
}

// nsteps steps from the given parity, grids one after another on one stream
// (fused = 0) or each on its own stream between a fork and a join per step.
static int mg_steps(wam_multigrid *mg, int nsteps, int parity, bool concurrent)
{
    const int n = (int)mg->grids.size();
    for (int s = 0; s < nsteps; s++) {
        if (concurrent) MGCHK(hipEventRecord(mg->ev[0], mg->origin));
        for (int i = 0; i < n; i++) {
            auto &g = mg->grids[i];
            float *cur = parity ? g.st.d_fb : g.st.d_fa, *nxt = parity ? g.st.d_fa : g.st.d_fb;
            hipStream_t st = mg->origin;
            if (concurrent) {
                st = mg->side[i];
                MGCHK(hipStreamWaitEvent(st, mg->ev[0], 0));
            }
            const int rc = mg_grid_step(mg, g, cur, nxt, st);
            if (rc) return rc;
            if (concurrent) {
                MGCHK(hipEventRecord(mg->ev[1 + i], st));
                MGCHK(hipStreamWaitEvent(mg->origin, mg->ev[1 + i], 0));
            }
        }
        parity ^= 1;
        const int rc = mg_links(mg, parity, mg->origin);
        if (rc) return rc;
    }
    return WAM_OK;
}

extern "C" {

int wam_gpu_multigrid_create(wam_multigrid **out, int32_t ngrid, const wam_grid_state *grids)
{
    if (!out || ngrid < 1 || ngrid > 64 || !grids) return WAM_E_ARG;
    *out = nullptr;
    auto *mg = new wam_multigrid();
    for (int i = 0; i < ngrid; i++) {
        const wam_grid_state &g = grids[i];
        wam_ctx *ctx = g.ctx;
        if (!ctx || !g.d_fa || !g.d_fb || g.d_fa == g.d_fb || !g.d_u10 || !g.d_udir || !g.d_tauw || !g.d_ustar ||
            !g.d_z0 || !g.d_roairn || !g.d_wstar) {
            delete mg;
            return WAM_E_ARG;
        }
        if (!ctx->prepared) { ctx->err = "PREPARE_PROPAGATION not done"; delete mg; return WAM_E_STATE; }
        for (int j = 0; j < i; j++)      // a context's work arrays serve one grid
            if (grids[j].ctx == ctx) { delete mg; return WAM_E_ARG; }
        // the work arrays and sub-step buffers, allocated before any capture
        const size_t nb = (size_t)ctx->tab.KL * ctx->tab.ML * ctx->tab.NSEA;
        int rc = WAM_OK;
        if (ctx->counter > 1 && !ctx->fscratch && hipMalloc(&ctx->fscratch, nb * sizeof(float)) != hipSuccess)
            rc = WAM_E_HIP;
        if (!rc && implsch_split_supported(ctx->dt)) rc = ensure_impl_scratch(ctx, ctx->isx, ctx->tab.NSEA, 0);
        if (rc) { delete mg; return rc; }
        wam_multigrid::Grid gr;
        gr.st = g;
        gr.a = impl_args(g.d_fb, g.d_u10, g.d_udir, g.d_tauw, g.d_ustar, g.d_z0, g.d_roairn, g.d_wstar, 0, nullptr,
                         nullptr, nullptr, ctx->tab.NSEA);
        mg->grids.push_back(gr);
    }
    bool ok = hipStreamCreateWithFlags(&mg->origin, hipStreamNonBlocking) == hipSuccess;
    mg->side.resize(ngrid);
    mg->ev.resize(ngrid + 1);
    for (int i = 0; ok && i < ngrid; i++) ok = hipStreamCreateWithFlags(&mg->side[i], hipStreamNonBlocking) == hipSuccess;
    for (auto &e : mg->ev) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    *out = mg;
    return ok ? WAM_OK : mg_fail(mg, "stream or event creation failed", WAM_E_HIP);
}

int wam_gpu_multigrid_add_nest(wam_multigrid *mg, int32_t coarse, int32_t fine, int32_t nb, const int32_t *ijarc,
                               int32_t nbounf, const int32_t *ijarf, const int32_t *ibfl, const int32_t *ibfr,
                               const float *bfw)
{
    if (!mg) return WAM_E_ARG;
    const int n = (int)mg->grids.size();
    if (coarse < 0 || coarse >= n || fine < 0 || fine >= n || coarse == fine || nb < 1 || nbounf < 0 || !ijarc ||
        (nbounf > 0 && (!ijarf || !ibfl || !ibfr || !bfw)))
        return mg_fail(mg, "nest: bad grid index or point list", WAM_E_ARG);
    const Tables &tc = mg->grids[coarse].st.ctx->tab, &tf = mg->grids[fine].st.ctx->tab;
    // READ_BOUNDARY_INPUT aborts unless the spectra match (read_boundary_input.f90:136-156)
    if (tc.KL != tf.KL || tc.ML != tf.ML || std::fabs(tc.FR[0] - tf.FR[0]) > 0.00001f ||
        std::fabs(tc.TH[0] - tf.TH[0]) > 0.0001f)
        return mg_fail(mg, "nest: the coarse and fine grids' directions and frequencies differ", WAM_E_ARG);
    for (int32_t i = 0; i < nb; i++)
        if (ijarc[i] < 1 || ijarc[i] > tc.NSEA) return mg_fail(mg, "nest: IJARC outside 1..NSEA", WAM_E_ARG);
    // each fine point once, with its last entry (as the reference's loop leaves it)
    std::vector<std::pair<int32_t, int32_t>> byp((size_t)nbounf);
    for (int32_t i = 0; i < nbounf; i++) {
        if (ijarf[i] < 1 || ijarf[i] > tf.NSEA || ibfl[i] < 1 || ibfl[i] > nb || ibfr[i] < 1 || ibfr[i] > nb)
            return mg_fail(mg, "nest: IJARF outside 1..NSEA or IBFL, IBFR outside 1..NB", WAM_E_ARG);
        byp[i] = {ijarf[i], i};
    }
    std::sort(byp.begin(), byp.end());
    std::vector<int32_t> keep;
    for (size_t j = 0; j < byp.size(); j++)
        if (j + 1 == byp.size() || byp[j + 1].first != byp[j].first) keep.push_back(byp[j].second);
    std::sort(keep.begin(), keep.end());
    const int nf = (int)keep.size();
    std::vector<int32_t> idx((size_t)nb + 3 * nf);
    std::vector<float> w(nf);
    std::copy(ijarc, ijarc + nb, idx.begin());
    for (int j = 0; j < nf; j++) {
        idx[nb + j] = ijarf[keep[j]];
        idx[nb + nf + j] = ibfl[keep[j]];
        idx[nb + 2 * nf + j] = ibfr[keep[j]];
        w[j] = bfw[keep[j]];
    }
    wam_multigrid::Link l;
    l.coarse = coarse; l.fine = fine; l.nb = nb; l.nf = nf;
    const size_t NB = (size_t)tc.KL * tc.ML;
    MGCHK(hipMalloc(&l.ijarc, idx.size() * sizeof(int32_t)));
    MGCHK(hipMalloc(&l.fbc, ((size_t)nb * NB + 3 * (size_t)nb + nf) * sizeof(float)));
    l.mean = l.fbc + (size_t)nb * NB;
    l.bfw = l.mean + 3 * (size_t)nb;
    l.ijarf = l.ijarc + nb; l.ibfl = l.ijarf + nf; l.ibfr = l.ibfl + nf;
    MGCHK(hipMemcpy(l.ijarc, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (nf) MGCHK(hipMemcpy(l.bfw, w.data(), nf * sizeof(float), hipMemcpyHostToDevice));
    mg->links.push_back(l);
    if (mg->gexec) { (void)hipGraphExecDestroy(mg->gexec); mg->gexec = nullptr; }
    if (mg->graph) { (void)hipGraphDestroy(mg->graph); mg->graph = nullptr; }
    return WAM_OK;
}

int wam_gpu_multigrid_run(wam_multigrid *mg, int32_t nsteps, int32_t fused)
{
    if (!mg || nsteps < 0) return WAM_E_ARG;
    if (nsteps == 0) return WAM_OK;
    const int n = (int)mg->grids.size();
    // after the work already queued on every grid's context stream
    for (int i = 0; i < n; i++) {
        MGCHK(hipEventRecord(mg->ev[1 + i], mg->grids[i].st.ctx->stream));
        MGCHK(hipStreamWaitEvent(mg->origin, mg->ev[1 + i], 0));
    }
    if (!fused) {
        const int rc = mg_steps(mg, nsteps, mg->parity, false);
        if (rc) return rc;
    } else {
        // the grids' source options are baked into the captured launches
        int opts = 0;
        for (int i = 0; i < n; i++) {
            const DevTables &d = mg->grids[i].st.ctx->dt;
            opts = opts * 8 + (d.brk | d.phil << 1 | d.split << 2);
        }
        if (!mg->gexec || mg->gsteps != nsteps || mg->gparity != mg->parity || mg->gopts != opts) {
            if (mg->gexec) { (void)hipGraphExecDestroy(mg->gexec); mg->gexec = nullptr; }
            if (mg->graph) { (void)hipGraphDestroy(mg->graph); mg->graph = nullptr; }
            hipStream_t cap = nullptr;
            MGCHK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
            // capture on a stream of its own (origin is waiting on the grids' streams)
            hipStream_t keep = mg->origin;
            mg->origin = cap;
            hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeRelaxed);
            int rc = e == hipSuccess ? mg_steps(mg, nsteps, mg->parity, true) : WAM_E_HIP;
            hipGraph_t g = nullptr;
            const hipError_t e2 = hipStreamEndCapture(cap, &g);
            mg->origin = keep;
            (void)hipStreamDestroy(cap);
            if (e != hipSuccess) return mg_fail(mg, "hipStreamBeginCapture failed", WAM_E_HIP);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            MGCHK(e2);
            mg->graph = g;
            MGCHK(hipGraphInstantiate(&mg->gexec, g, nullptr, nullptr, 0));
            mg->gsteps = nsteps;
            mg->gparity = mg->parity;
            mg->gopts = opts;
        }
        MGCHK(hipGraphLaunch(mg->gexec, mg->origin));
    }
    mg->parity ^= nsteps & 1;
    // and before what the grids' streams do next
    MGCHK(hipEventRecord(mg->ev[0], mg->origin));
    for (int i = 0; i < n; i++) MGCHK(hipStreamWaitEvent(mg->grids[i].st.ctx->stream, mg->ev[0], 0));
    return WAM_OK;
}

int wam_gpu_multigrid_state(wam_multigrid *mg, int32_t grid, float **d_cur)
{
    if (!mg || !d_cur || grid < 0 || grid >= (int)mg->grids.size()) return WAM_E_ARG;
    *d_cur = mg->parity ? mg->grids[grid].st.d_fb : mg->grids[grid].st.d_fa;
    return WAM_OK;
}

const char *wam_gpu_multigrid_last_error(wam_multigrid *mg) { return mg ? mg->err.c_str() : "null handle"; }

void wam_gpu_multigrid_free(wam_multigrid *mg)
{
    if (!mg) return;
    if (mg->origin) (void)hipStreamSynchronize(mg->origin);
    for (auto &l : mg->links) {
        if (l.ijarc) (void)hipFree(l.ijarc);
        if (l.fbc) (void)hipFree(l.fbc);
    }
    if (mg->gexec) (void)hipGraphExecDestroy(mg->gexec);
    if (mg->graph) (void)hipGraphDestroy(mg->graph);
    for (auto e : mg->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto st : mg->side)
        if (st) (void)hipStreamDestroy(st);
    if (mg->origin) (void)hipStreamDestroy(mg->origin);
    delete mg;
}

int wam_gpu_get_tuning(wam_ctx *ctx, int64_t *out)
{
    if (!ctx || !out) return WAM_E_ARG;
    out[0] = ctx->tune.prop_wg; out[1] = ctx->tune.src_wg; out[2] = ctx->tune.snl_per_cu; out[3] = ctx->tune.chunk;
    return WAM_OK;
}

int wam_gpu_set_tuning(wam_ctx *ctx, const int64_t *in)
{
    if (!ctx || !in) return WAM_E_ARG;
    const int pw = (int)in[0], sw = (int)in[1];
    if ((pw != 64 && pw != 128 && pw != 256) || (sw != 64 && sw != 128 && sw != 256) || in[2] < 0 || in[2] > 64 ||
        in[3] < 0 || in[3] > 0x7fffffff) {
        ctx->err = "launch shapes: PROPAGS and per-lane widths 64, 128 or 256; workgroups per CU 0-64; chunk >= 0";
        return WAM_E_ARG;
    }
    ctx->tune.prop_wg = pw; ctx->tune.src_wg = sw; ctx->tune.snl_per_cu = (int)in[2]; ctx->tune.chunk = (long)in[3];
    return WAM_OK;
}

int64_t wam_gpu_get_tuning_log(wam_ctx *ctx, char *buf, int64_t n)
{
    if (!ctx) return -1;
    std::string s;
    for (auto &e : ctx->tune_log) s += e.what + " " + std::to_string(e.value) + " " + std::to_string(e.ms) + "\n";
    if (buf && n > 0) {
        const size_t k = std::min((size_t)n - 1, s.size());
        std::memcpy(buf, s.data(), k);
        buf[k] = 0;
    }
    return (int64_t)s.size();
}

// The no-allocation form of wam_gpu_implsch_dev: the three-kernel IMPLSCH on
// chunks of sea points whose work arrays fit the static scratch array, carved
// as F, SL, FL, SMIN tiles, STRESSO's sums, the per-point scalars, SPRD, the
// LLWS bits and the ticket counters (each 256-byte aligned).  Returns the
// points per chunk (a multiple of 64), 0 if not even 64 fit.
static int noalloc_layout(const Tables &t, void *base, ImplScratch &x)
{
    const size_t nb = (size_t)t.KL * t.ML, nw = (nb + 31) / 32;
    const size_t per_pt = (4 * nb + 3 * (size_t)t.ML + 8 + t.KL + nw + (t.cfg.iphys == 2 ? 6 * (size_t)t.ML : 0)) * 4;
    const size_t fixed = 1024 + 11 * 256;                       // tickets, alignment
    if (NOALLOC_SCRATCH_BYTES < fixed + 64 * per_pt) return 0;
    const size_t c = (NOALLOC_SCRATCH_BYTES - fixed) / per_pt / 64 * 64;
    unsigned char *p = (unsigned char *)base;
    auto take = [&](size_t bytes) { void *q = p; p += (bytes + 255) / 256 * 256; return q; };
    x.NB4 = (int)(nb / 4);
    x.NX = (int)c;
    x.ft = (float *)take(c * nb * 4);
    x.slt = (float *)take(c * nb * 4);
    x.flt = (float *)take(c * nb * 4);
    x.smt = (float *)take(c * nb * 4);
    x.axy = (float *)take(3 * (size_t)t.ML * c * 4);
    x.pp = (float *)take(8 * c * 4);
    x.sprd = (float *)take((size_t)t.KL * c * 4);
    x.llws = (unsigned *)take(nw * c * 4);
    x.tick = (unsigned *)take(8 * 32 * 4);
    if (t.cfg.iphys == 2) x.st6 = (float *)take(6 * (size_t)t.ML * c * 4);
    if ((size_t)(p - (unsigned char *)base) > NOALLOC_SCRATCH_BYTES) return 0;
    return (int)c;
}

int wam_gpu_implsch_dev_noalloc(wam_ctx *ctx, float *d_fl3, const float *d_u10, const float *d_udir, float *d_tauw,
                                float *d_ustar, float *d_z0, const float *d_roairn, const float *d_wstar, int32_t lcflx,
                                float *d_phiaw, float *d_fluxes, int32_t *d_mij)
{
    if (!ctx || !d_fl3 || !d_u10 || !d_udir || !d_tauw || !d_ustar || !d_z0 || !d_roairn || !d_wstar) return WAM_E_ARG;
    if (!ctx->prepared) { ctx->err = "PREPARE_PROPAGATION not done"; return WAM_E_STATE; }
    const ImplArgs a = impl_args(d_fl3, d_u10, d_udir, d_tauw, d_ustar, d_z0, d_roairn, d_wstar, lcflx, d_phiaw,
                                 d_fluxes, d_mij, ctx->tab.NSEA);
    if (!implsch_split_supported(ctx->dt)) {       // the single kernel uses no work arrays
        HIPCHK(launch_implsch(ctx->dt, a, ctx->ncu, ctx->stream));
        return WAM_OK;
    }
    if (!ctx->noalloc) { ctx->err = "the device address of the static scratch array is unknown"; return WAM_E_STATE; }
    ImplScratch x{};
    const int c = noalloc_layout(ctx->tab, ctx->noalloc, x);
    if (c < 64) {
        ctx->err = "KL*ML too large for the no-allocation scratch (256 MiB holds the work arrays of fewer than 64 "
                   "sea points)";
        return WAM_E_ARG;
    }
    const int ns = ctx->tab.NSEA;
    for (int p0 = 0; p0 < ns; p0 += c) {
        x.P0 = p0;
        x.NP = std::min(c, ns - p0);
        HIPCHK(launch_implsch_split(ctx->dt, a, x, ctx->ncu, ctx->stream, nullptr, ctx->tune));
    }
    return WAM_OK;
}

static int ensure_stage(wam_ctx *ctx, size_t n)
{
    if (ctx->stage_n >= n) return WAM_OK;
    if (ctx->stage) (void)hipFree(ctx->stage);
    ctx->stage = nullptr; ctx->stage_n = 0;
    HIPCHK(hipMalloc(&ctx->stage, n * sizeof(float)));
    ctx->stage_n = n;
    return WAM_OK;
}

int wam_gpu_propags(wam_ctx *ctx, int32_t nijs, int32_t nijl, int32_t kl, int32_t ml, float *f3)
{
    if (!ctx || !f3) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    if (nijl - nijs + 1 != t.NSEA || kl != t.KL || ml != t.ML) { ctx->err = "PROPAGS size mismatch"; return WAM_E_ARG; }
    const size_t nb = (size_t)t.KL * t.ML * t.NSEA;
    int rc = ensure_stage(ctx, 2 * nb);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->stage, f3, nb * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
    rc = wam_gpu_propags_dev(ctx, ctx->stage, ctx->stage + nb);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(f3, ctx->stage + nb, nb * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAM_OK;
}

int wam_gpu_implsch(wam_ctx *ctx, int32_t nsea, int32_t kl, int32_t ml, float *fl3, const float *u10,
                    const float *udir, float *tauw, float *ustar, float *z0, const float *roairn,
                    const float *wstar, const float *depth, const int32_t *indep, int32_t lcflx, float *phiaw,
                    float *fluxes)
{
    if (!ctx || !fl3 || !u10 || !udir || !tauw || !ustar || !z0 || !roairn || !wstar) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    if (nsea != t.NSEA || kl != t.KL || ml != t.ML) { ctx->err = "IMPLSCH size mismatch"; return WAM_E_ARG; }
    if (!ctx->prepared) { ctx->err = "PREPARE_PROPAGATION not done"; return WAM_E_STATE; }
    // The reference passes DEPTH(NIJS:NIJL) and INDEP(NIJS:NIJL) of the same
    // module arrays PREPARE_PROPAGATION used (wamodel.f90:354); the kernels
    // read the context's prepared fields, so arguments that differ from them
    // mean the host changed the depth without preparing again.  INDEP is only
    // defined in shallow water (wam_propagation_module.f90:334-336).
    if (depth && std::memcmp(depth, ctx->depth_h.data(), (size_t)nsea * sizeof(float)) != 0) {
        ctx->err = "DEPTH differs from the field of PREPARE_PROPAGATION (call wam_gpu_prepare_propagation again)";
        return WAM_E_STATE;
    }
    if (indep && t.cfg.shallow && std::memcmp(indep, t.INDEP.data(), (size_t)nsea * sizeof(int32_t)) != 0) {
        ctx->err = "INDEP differs from the table of PREPARE_PROPAGATION (call wam_gpu_prepare_propagation again)";
        return WAM_E_STATE;
    }
    const size_t nb = (size_t)t.KL * t.ML * nsea, ns = nsea;
    // FL3, then 8 per-point arrays (U10 UDIR TAUW USTAR Z0 ROAIRN WSTAR PHIAW) and 6 flux rows
    int rc = ensure_stage(ctx, nb + 14 * ns);
    if (rc) return rc;
    float *d = ctx->stage;
    float *dfl = d, *du = d + nb, *dd = du + ns, *dt = dd + ns, *dus = dt + ns, *dz = dus + ns, *dr = dz + ns,
          *dw = dr + ns, *dph = dw + ns, *dfx = dph + ns;
    hipStream_t st = ctx->stream;
    HIPCHK(hipMemcpyAsync(dfl, fl3, nb * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(du, u10, ns * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dd, udir, ns * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dt, tauw, ns * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dz, z0, ns * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dr, roairn, ns * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dw, wstar, ns * 4, hipMemcpyHostToDevice, st));
    rc = wam_gpu_implsch_dev(ctx, dfl, du, dd, dt, dus, dz, dr, dw, lcflx, dph, lcflx ? dfx : nullptr, nullptr);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(fl3, dfl, nb * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(tauw, dt, ns * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(ustar, dus, ns * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(z0, dz, ns * 4, hipMemcpyDeviceToHost, st));
    if (phiaw) HIPCHK(hipMemcpyAsync(phiaw, dph, ns * 4, hipMemcpyDeviceToHost, st));
    if (fluxes && lcflx) HIPCHK(hipMemcpyAsync(fluxes, dfx, 6 * ns * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return WAM_OK;
}

int wam_gpu_output_parameters_dev(wam_ctx *ctx, const float *d_fl3, const float *d_u10, const float *d_udir,
                                  const float *d_ustar, const float *d_tauw, const float *d_z0, int32_t orient,
                                  float *d_out)
{
    if (!ctx || !d_fl3 || !d_u10 || !d_udir || !d_ustar || !d_tauw || !d_z0 || !d_out) return WAM_E_ARG;
    if (!ctx->prepared) { ctx->err = "PREPARE_PROPAGATION not done"; return WAM_E_STATE; }
    if (ctx->tab.KL > 64 || ctx->tab.ML > 64) { ctx->err = "output parameters need KL, ML <= 64"; return WAM_E_ARG; }
    HIPCHK(launch_output_params(ctx->dt, d_fl3, d_u10, d_udir, d_ustar, d_tauw, d_z0, orient ? 1 : 0, d_out,
                                ctx->stream));
    return WAM_OK;
}

int wam_gpu_output_parameters(wam_ctx *ctx, int32_t nsea, int32_t kl, int32_t ml, const float *fl3, const float *u10,
                              const float *udir, const float *ustar, const float *tauw, const float *z0,
                              int32_t orient, float *out)
{
    if (!ctx || !fl3 || !u10 || !udir || !ustar || !tauw || !z0 || !out) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    if (nsea != t.NSEA || kl != t.KL || ml != t.ML) { ctx->err = "output parameters: size mismatch"; return WAM_E_ARG; }
    const size_t nb = (size_t)t.KL * t.ML * nsea, ns = nsea;
    // FL3, five per-point fields, WAM_NOUT output rows
    int rc = ensure_stage(ctx, nb + (5 + WAM_NOUT) * ns);
    if (rc) return rc;
    float *d = ctx->stage;
    float *dfl = d, *du = d + nb, *dd = du + ns, *dus = dd + ns, *dt = dus + ns, *dz = dt + ns, *dout = dz + ns;
    hipStream_t st = ctx->stream;
    HIPCHK(hipMemcpyAsync(dfl, fl3, nb * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(du, u10, ns * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dd, udir, ns * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dus, ustar, ns * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dt, tauw, ns * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dz, z0, ns * 4, hipMemcpyHostToDevice, st));
    rc = wam_gpu_output_parameters_dev(ctx, dfl, du, dd, dus, dt, dz, orient, dout);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(out, dout, WAM_NOUT * ns * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return WAM_OK;
}

static int ensure_nest(wam_ctx *ctx, size_t n)
{
    if (ctx->nest_n >= n) return WAM_OK;
    if (ctx->nest) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->nest);
    }
    ctx->nest = nullptr;
    ctx->nest_n = 0;
    HIPCHK(hipMalloc(&ctx->nest, n * sizeof(float)));
    ctx->nest_n = n;
    return WAM_OK;
}

int wam_gpu_boundary_output_dev(wam_ctx *ctx, int32_t nb, const int32_t *ijarc, const float *d_fl3, float *fbc,
                                float *mean)
{
    if (!ctx || nb < 0 || (nb > 0 && (!ijarc || !d_fl3 || !fbc || !mean))) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    if (t.KL > 64 || t.ML > 64) { ctx->err = "nest boundary exchange needs KL, ML <= 64"; return WAM_E_ARG; }
    for (int32_t i = 0; i < nb; i++)
        if (ijarc[i] < 1 || ijarc[i] > t.NSEA) { ctx->err = "IJARC outside 1..NSEA"; return WAM_E_ARG; }
    if (nb == 0) return WAM_OK;
    const size_t NB = (size_t)t.KL * t.ML, n = (size_t)nb;
    int rc = ensure_nest(ctx, n * NB + 4 * n);
    if (rc) return rc;
    float *dfbc = ctx->nest, *dmean = dfbc + n * NB;
    int *dij = reinterpret_cast<int *>(dmean + 3 * n);
    hipStream_t st = ctx->stream;
    HIPCHK(hipMemcpyAsync(dij, ijarc, n * 4, hipMemcpyHostToDevice, st));
    hipEvent_t ev[2];
    rc = timing_events(ctx, ev, 2);
    if (rc) return rc;
    timing_record(ctx, ev[0]);
    HIPCHK(launch_boundary_gather(ctx->dt, d_fl3, dij, nb, dfbc, dmean, st));
    timing_record(ctx, ev[1]);
    timing_keep(ctx, "boundary_gather", ev[0], ev[1]);
    HIPCHK(hipMemcpyAsync(fbc, dfbc, n * NB * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(mean, dmean, 3 * n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return WAM_OK;
}

int wam_gpu_boundary_input_dev(wam_ctx *ctx, int32_t nbinp, const float *f1, const float *m1, const float *f2,
                               const float *m2, int32_t which, float del12, float del1l, int32_t nbounf,
                               const int32_t *ijarf, const int32_t *ibfl, const int32_t *ibfr, const float *bfw,
                               float *d_fl3, float *fi, float *mi)
{
    if (!ctx || nbinp <= 0 || nbounf < 0 || which < 0 || which > 2) return WAM_E_ARG;
    if (which != 2 && (!f1 || !m1)) return WAM_E_ARG;
    if (which != 1 && (!f2 || !m2)) return WAM_E_ARG;
    if (nbounf > 0 && (!ijarf || !ibfl || !ibfr || !bfw || !d_fl3)) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    for (int32_t i = 0; i < nbounf; i++) {
        if (ijarf[i] < 1 || ijarf[i] > t.NSEA) { ctx->err = "IJARF outside 1..NSEA"; return WAM_E_ARG; }
        if (ibfl[i] < 1 || ibfl[i] > nbinp || ibfr[i] < 1 || ibfr[i] > nbinp) {
            ctx->err = "IBFL or IBFR outside 1..NBINP";
            return WAM_E_ARG;
        }
    }
    // the reference stores the points one after another: of a sea point
    // listed twice the last entry stays; the kernel writes each point once
    std::vector<std::pair<int32_t, int32_t>> byp((size_t)nbounf);    // (IJARF, entry), sorted
    for (int32_t i = 0; i < nbounf; i++) byp[i] = {ijarf[i], i};
    std::sort(byp.begin(), byp.end());
    std::vector<int32_t> keep;                                       // the last entry of each point
    for (size_t j = 0; j < byp.size(); j++)
        if (j + 1 == byp.size() || byp[j + 1].first != byp[j].first) keep.push_back(byp[j].second);
    std::sort(keep.begin(), keep.end());
    std::vector<int32_t> ij, il, ir;
    std::vector<float> w;
    for (const int32_t i : keep) {
        ij.push_back(ijarf[i]);
        il.push_back(ibfl[i]);
        ir.push_back(ibfr[i]);
        w.push_back(bfw[i]);
    }
    const size_t NB = (size_t)t.KL * t.ML, np = (size_t)nbinp, nf = ij.size();
    int rc = ensure_nest(ctx, 3 * np * NB + 9 * np + 4 * nf);
    if (rc) return rc;
    float *df1 = ctx->nest, *df2 = df1 + np * NB, *dfi = df2 + np * NB, *dm1 = dfi + np * NB, *dm2 = dm1 + 3 * np,
          *dmi = dm2 + 3 * np, *dw = dmi + 3 * np;
    int *dij = reinterpret_cast<int *>(dw + nf), *dl = dij + nf, *dr = dl + nf;
    hipStream_t st = ctx->stream;
    if (which != 2) {
        HIPCHK(hipMemcpyAsync(df1, f1, np * NB * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dm1, m1, 3 * np * 4, hipMemcpyHostToDevice, st));
    }
    if (which != 1) {
        HIPCHK(hipMemcpyAsync(df2, f2, np * NB * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dm2, m2, 3 * np * 4, hipMemcpyHostToDevice, st));
    }
    if (nf > 0) {
        HIPCHK(hipMemcpyAsync(dw, w.data(), nf * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dij, ij.data(), nf * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dl, il.data(), nf * 4, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(dr, ir.data(), nf * 4, hipMemcpyHostToDevice, st));
    }
    hipEvent_t ev[3];
    rc = timing_events(ctx, ev, 3);
    if (rc) return rc;
    timing_record(ctx, ev[0]);
    HIPCHK(launch_boundary_time(ctx->dt, nbinp, df1, dm1, df2, dm2, which, del12, del1l, dfi, dmi, st));
    timing_record(ctx, ev[1]);
    HIPCHK(launch_boundary_space(ctx->dt, nbinp, dfi, dmi, (int)nf, dij, dl, dr, dw, d_fl3, st));
    timing_record(ctx, ev[2]);
    timing_keep(ctx, "boundary_time", ev[0], ev[1]);
    timing_keep(ctx, "boundary_space", ev[1], ev[2]);
    if (fi) HIPCHK(hipMemcpyAsync(fi, dfi, np * NB * 4, hipMemcpyDeviceToHost, st));
    if (mi) HIPCHK(hipMemcpyAsync(mi, dmi, 3 * np * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));     // the host vectors above are read by the copies
    return WAM_OK;
}

int wam_gpu_boundary_output(wam_ctx *ctx, int32_t nsea, int32_t kl, int32_t ml, const float *fl3, int32_t nb,
                            const int32_t *ijarc, float *fbc, float *mean)
{
    if (!ctx || !fl3) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    if (nsea != t.NSEA || kl != t.KL || ml != t.ML) { ctx->err = "boundary output: size mismatch"; return WAM_E_ARG; }
    const size_t nbyte = (size_t)t.KL * t.ML * nsea * 4;
    int rc = ensure_stage(ctx, nbyte / 4);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->stage, fl3, nbyte, hipMemcpyHostToDevice, ctx->stream));
    return wam_gpu_boundary_output_dev(ctx, nb, ijarc, ctx->stage, fbc, mean);
}

int wam_gpu_boundary_input(wam_ctx *ctx, int32_t nsea, int32_t kl, int32_t ml, float *fl3, int32_t nbinp,
                           const float *f1, const float *m1, const float *f2, const float *m2, int32_t which,
                           float del12, float del1l, int32_t nbounf, const int32_t *ijarf, const int32_t *ibfl,
                           const int32_t *ibfr, const float *bfw, float *fi, float *mi)
{
    if (!ctx || !fl3) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    if (nsea != t.NSEA || kl != t.KL || ml != t.ML) { ctx->err = "boundary input: size mismatch"; return WAM_E_ARG; }
    const size_t nbyte = (size_t)t.KL * t.ML * nsea * 4;
    int rc = ensure_stage(ctx, nbyte / 4);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(ctx->stage, fl3, nbyte, hipMemcpyHostToDevice, ctx->stream));
    rc = wam_gpu_boundary_input_dev(ctx, nbinp, f1, m1, f2, m2, which, del12, del1l, nbounf, ijarf, ibfl, ibfr, bfw,
                                    ctx->stage, fi, mi);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(fl3, ctx->stage, nbyte, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAM_OK;
}

// BOUT row of parameter IP (1-based): parameters 1-16 and 32 are computed
static int map_row(int ip) { return ip >= 1 && ip <= 16 ? ip - 1 : (ip == 32 ? 16 : -1); }

int wam_gpu_output_grids_dev(wam_ctx *ctx, const int32_t *fflag, const float *d_out, float *d_grids)
{
    if (!ctx || !fflag || !d_out || !d_grids) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    int rows[WAM_NOUT], n = 0;
    for (int ip = 1; ip <= WAM_MAP_NOUT; ip++) {
        if (!fflag[ip - 1]) continue;
        const int r = map_row(ip);
        if (r < 0) { ctx->err = "gridded output: only parameters 1-16 and 32 are computed"; return WAM_E_ARG; }
        rows[n++] = r;
    }
    const int nx = t.cfg.nx, ny = t.cfg.ny, ncell = nx * ny;
    if (!ctx->cell_pt) {
        std::vector<int> cp((size_t)ncell, -1);
        for (int ij = 0; ij < t.NSEA; ij++) cp[(size_t)(t.KXLT[ij] - 1) * nx + (t.IXLG[ij] - 1)] = ij;
        HIPCHK(hipMalloc(&ctx->cell_pt, (size_t)ncell * sizeof(int)));
        HIPCHK(hipMemcpy(ctx->cell_pt, cp.data(), (size_t)ncell * sizeof(int), hipMemcpyHostToDevice));
    }
    HIPCHK(launch_map_grids(d_out, t.NSEA, ctx->cell_pt, ncell, rows, n, d_grids, ctx->stream));
    return WAM_OK;
}

int wam_gpu_write_map(wam_ctx *ctx, const char *path, int32_t append, const char *cdtpro, const char *cdatea,
                      const int32_t *fflag, const float *d_out)
{
    if (!ctx || !path || !cdtpro || !cdatea || !fflag || !d_out) return WAM_E_ARG;
    const Tables &t = ctx->tab;
    int n = 0;
    for (int i = 0; i < WAM_MAP_NOUT; i++) n += fflag[i] != 0;
    const int nx = t.cfg.nx, ny = t.cfg.ny;
    const size_t ncell = (size_t)nx * ny;
    if (ctx->grids_n < n * ncell) {
        if (ctx->grids) {
            HIPCHK(hipStreamSynchronize(ctx->stream));
            (void)hipFree(ctx->grids);
        }
        ctx->grids = nullptr;
        ctx->grids_n = 0;
        HIPCHK(hipMalloc(&ctx->grids, n * ncell * sizeof(float)));
        ctx->grids_n = n * ncell;
    }
    int rc = n ? wam_gpu_output_grids_dev(ctx, fflag, d_out, ctx->grids) : WAM_OK;
    if (rc) return rc;
    std::vector<float> host(n * ncell);
    if (n) HIPCHK(hipMemcpyAsync(host.data(), ctx->grids, n * ncell * sizeof(float), hipMemcpyDeviceToHost,
                                 ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const int32_t corners[4] = {(int32_t)t.cfg.amowep, (int32_t)t.cfg.amosop, (int32_t)t.cfg.amoeap,
                                (int32_t)t.cfg.amonop};
    std::vector<int32_t> nlon((size_t)ny, nx), zdello((size_t)ny, t.ZDELLO);
    rc = wam_map_write(path, append, cdtpro, cdatea, nx, ny, corners, nlon.data(), zdello.data(), fflag,
                       host.data());
    if (rc) ctx->err = "gridded output: cannot write the file";
    return rc;
}

int wam_gpu_set_propagation_weights(wam_ctx *ctx, int32_t stored)
{
    if (!ctx) return WAM_E_ARG;
    if (stored < -1 || stored > 1) { ctx->err = "stored must be -1, 0 or 1"; return WAM_E_ARG; }
    ctx->weights_mode = stored;
    ctx->prepared = false;      // takes effect at the next wam_gpu_prepare_propagation
    return WAM_OK;
}

int wam_gpu_set_implsch_form(wam_ctx *ctx, int32_t form)
{
    if (!ctx) return WAM_E_ARG;
    if (form < -1 || form > 1) { ctx->err = "form must be -1, 0 or 1"; return WAM_E_ARG; }
    ctx->implsch_form = form;
    ctx->dt.split = implsch_form_split(form, ctx->tab.cfg.iphys, ctx->tab.NSEA);
    return WAM_OK;
}

int wam_gpu_set_source_options(wam_ctx *ctx, int32_t wave_breaking, int32_t phillips)
{
    if (!ctx) return WAM_E_ARG;
    // SET_MODEL_OPTION switches breaking off in deep water (wam_timopt_module.f90:213-227)
    ctx->dt.brk = wave_breaking && ctx->tab.cfg.shallow ? 1 : 0;
    ctx->dt.phil = phillips ? 1 : 0;
    return WAM_OK;
}

int wam_gpu_get_source_options(wam_ctx *ctx, int32_t *wave_breaking, int32_t *phillips)
{
    if (!ctx) return WAM_E_ARG;
    if (wave_breaking) *wave_breaking = ctx->dt.brk;
    if (phillips) *phillips = ctx->dt.phil;
    return WAM_OK;
}

int wam_gpu_set_timing(wam_ctx *ctx, int32_t on)
{
    if (!ctx) return WAM_E_ARG;
    timing_clear(ctx);
    ctx->timing = on != 0;
    return WAM_OK;
}

int wam_gpu_get_timing(wam_ctx *ctx, const char *name, double *total_ms, int64_t *launches)
{
    if (!ctx || !name || !total_ms || !launches) return WAM_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double sum = 0.0;
    int64_t n = 0;
    for (auto &r : ctx->timed) {
        if (r.name != name) continue;
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
        sum += ms;
        n++;
    }
    *total_ms = sum;
    *launches = n;
    return WAM_OK;
}

// The grid's neighbour relation is symmetric (KLAT_S(i) = j exactly when
// KLAT_N(j) = i, KLON_W(i) = j exactly when KLON_E(j) = i): the transposed
// PROPAGS gathers over it.
static bool neighbours_symmetric(const Tables &t)
{
    for (int j = 0; j < t.NSEA; j++) {
        const int s = t.KLAT_S[j] - 1, n = t.KLAT_N[j] - 1, w = t.KLON_W[j] - 1, e = t.KLON_E[j] - 1;
        if (s >= 0 && t.KLAT_N[s] - 1 != j) return false;
        if (n >= 0 && t.KLAT_S[n] - 1 != j) return false;
        if (w >= 0 && t.KLON_E[w] - 1 != j) return false;
        if (e >= 0 && t.KLON_W[e] - 1 != j) return false;
    }
    return true;
}

static int adjoint_planes(wam_ctx *ctx);

// Checks and work arrays of the step adjoint for nmem members.
static int adjoint_ready(wam_ctx *ctx, int32_t nmem)
{
    int rc = ensemble_args_ok(ctx, nmem);
    if (rc) return rc;
    const Tables &t = ctx->tab;
    if (t.cfg.iphys != 0 || t.cfg.shallow) {
        ctx->err = "the step adjoint is built for IPHYS = 0 in deep water";
        return WAM_E_ARG;
    }
    if (ctx->dt.phil) { ctx->err = "the step adjoint does not differentiate SOURCE_PHILLIPS"; return WAM_E_ARG; }
    if (t.KL > 64 || t.ML > 64) { ctx->err = "the step adjoint needs KL, ML <= 64"; return WAM_E_ARG; }
    if (!t.adj_err.empty()) { ctx->err = t.adj_err; return WAM_E_ARG; }
    if (ctx->adj_sym < 0) ctx->adj_sym = neighbours_symmetric(t) ? 1 : 0;
    if (!ctx->adj_sym) { ctx->err = "the transposed PROPAGS needs a symmetric neighbour relation"; return WAM_E_ARG; }
    const size_t nb = (size_t)t.KL * t.ML * t.NSEA, n = (size_t)nmem * nb;
    if (ctx->adj_n < n) {
        if (ctx->adj_fp) (void)hipFree(ctx->adj_fp);
        if (ctx->adj_fb) (void)hipFree(ctx->adj_fb);
        ctx->adj_fp = ctx->adj_fb = nullptr;
        ctx->adj_n = 0;
        HIPCHK(hipMalloc(&ctx->adj_fp, n * sizeof(float)));
        HIPCHK(hipMalloc(&ctx->adj_fb, n * sizeof(float)));
        ctx->adj_n = n;
    }
    rc = adjoint_planes(ctx);
    if (rc) return rc;
    return ensure_ensemble_fscratch(ctx, nmem);
}

// The weight planes the transposed stencil reads: PROPAGS's own, or where
// it computes its weights in the stencil, planes made for the adjoint (as
// CHECK_CFL scaled them).
static int adjoint_planes(wam_ctx *ctx)
{
    const Tables &t = ctx->tab;
    const size_t nb = (size_t)t.KL * t.ML * t.NSEA;
    if (!ctx->coef && !ctx->adj_coef) {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, nb * ctx->ncoef * sizeof(float)));
        ctx->point_allocs.push_back(p);
        ctx->adj_coef = (float *)p;
        hipLaunchKernelGGL(prep_coef_kernel, dim3((t.NSEA + 255) / 256, t.KL * t.ML), dim3(256), 0, ctx->stream,
                           ctx->dt, ctx->adj_coef);
        HIPCHK(hipGetLastError());
        if (ctx->counter > 1) {
            hipLaunchKernelGGL(rescale_kernel, dim3(4096), dim3(256), 0, ctx->stream, ctx->adj_coef, nb, ctx->ncoef,
                               (float)ctx->counter);
            HIPCHK(hipGetLastError());
        }
    }
    if (!t.cfg.refr_c && !ctx->adj_tw) {
        void *p = nullptr;
        HIPCHK(hipMalloc(&p, nb * 5 * sizeof(float)));
        ctx->point_allocs.push_back(p);
        ctx->adj_tw = (float *)p;
        HIPCHK(launch_transpose_weights(ctx->dt, ctx->coef ? ctx->coef : ctx->adj_coef, ctx->adj_tw, ctx->stream));
    }
    return WAM_OK;
}

// PROPAGS transposed: COUNTER sub-steps from src, the last into dst, the
// others alternating between the two scratch arrays (src may be one of them).
static int propags_transposed(wam_ctx *ctx, int nmem, const float *src, float *dst, float *sa, float *sb)
{
    const float *coef = ctx->coef ? ctx->coef : ctx->adj_coef;
    for (int i = 0; i < ctx->counter; i++) {
        const int left = ctx->counter - 1 - i;
        float *to = left % 2 == 0 ? dst : (src == sa ? sb : sa);
        hipEvent_t ev[2] = {nullptr, nullptr};
        const int rc = timing_events(ctx, ev, 2);
        if (rc) return rc;
        timing_record(ctx, ev[0]);
        HIPCHK(launch_propags_adjoint(ctx->dt, coef, ctx->adj_tw, nmem, src, to, ctx->stream));
        timing_record(ctx, ev[1]);
        timing_keep(ctx, "propags_adjoint", ev[0], ev[1]);
        src = to;
    }
    return WAM_OK;
}

int wam_gpu_step_adjoint_dev(wam_ctx *ctx, int32_t nmem, const float *d_fl3, const float *d_u10,
                             const float *d_udir, const float *d_tauw, const float *d_roairn, const float *d_wstar,
                             const float *d_fl3_bar, const float *d_tauw_bar, const float *d_ustar_bar,
                             const float *d_z0_bar, float *d_fl3_bar_out, float *d_u10_bar, float *d_udir_bar,
                             float *d_tauw_bar_out)
{
    if (!d_fl3 || !d_u10 || !d_udir || !d_tauw || !d_roairn || !d_wstar || !d_fl3_bar || !d_fl3_bar_out ||
        !d_u10_bar || !d_udir_bar || !d_tauw_bar_out || d_fl3_bar_out == d_fl3)
        return WAM_E_ARG;
    int rc = adjoint_ready(ctx, nmem);
    if (rc) return rc;
    const Tables &t = ctx->tab;
    // 1. IMPLSCH's input: PROPAGS of the step's input state
    rc = propags_members(ctx, nmem, d_fl3, ctx->adj_fp, ctx->fscratch_e, true);
    if (rc) return rc;
    // 2. IMPLSCH's adjoint, the input spectrum's bar into adj_fb
    AdjArgs a{};
    a.fl3 = ctx->adj_fp; a.u10 = d_u10; a.udir = d_udir; a.tauw = d_tauw; a.roairn = d_roairn; a.wstar = d_wstar;
    a.fl3_bar = d_fl3_bar; a.tauw_bar = d_tauw_bar; a.ustar_bar = d_ustar_bar; a.z0_bar = d_z0_bar;
    a.fl3_bar_out = ctx->adj_fb; a.u10_bar = d_u10_bar; a.udir_bar = d_udir_bar; a.tauw_bar_out = d_tauw_bar_out;
    a.NMP = nmem * t.NSEA;
    hipEvent_t ev[2] = {nullptr, nullptr};
    rc = timing_events(ctx, ev, 2);
    if (rc) return rc;
    timing_record(ctx, ev[0]);
    if (implsch_split_supported(ctx->dt)) {
        // in chunks of member-points, as the ensemble's IMPLSCH
        const size_t nmp = (size_t)a.NMP, cap = ensemble_chunk(ctx, nmp);
        rc = ensure_adj_scratch(ctx, cap);
        if (rc) return rc;
        ImplScratch &x = ctx->asx;
        for (size_t p0 = 0; p0 < nmp; p0 += cap) {
            x.P0 = (int)p0;
            x.NP = (int)std::min(cap, nmp - p0);
            HIPCHK(launch_implsch_adjoint_split(ctx->dt, a, x, ctx->aex, ctx->ncu, ctx->stream, ctx->tune));
        }
        x.P0 = 0;
        x.NP = x.NX;
    } else {
        HIPCHK(launch_implsch_adjoint(ctx->dt, a, ctx->ncu, ctx->stream));
    }
    timing_record(ctx, ev[1]);
    timing_keep(ctx, "implsch_adjoint", ev[0], ev[1]);
    // 3. PROPAGS transposed into d_fl3_bar_out
    return propags_transposed(ctx, nmem, ctx->adj_fb, d_fl3_bar_out, ctx->adj_fp, ctx->adj_fb);
}

int wam_gpu_propags_adjoint_dev(wam_ctx *ctx, int32_t nmem, const float *d_in_bar, float *d_out_bar)
{
    if (!d_in_bar || !d_out_bar || d_in_bar == d_out_bar) return WAM_E_ARG;
    int rc = ensemble_args_ok(ctx, nmem);
    if (rc) return rc;
    if (ctx->adj_sym < 0) ctx->adj_sym = neighbours_symmetric(ctx->tab) ? 1 : 0;
    if (!ctx->adj_sym) { ctx->err = "the transposed PROPAGS needs a symmetric neighbour relation"; return WAM_E_ARG; }
    rc = adjoint_planes(ctx);
    if (rc) return rc;
    const size_t n = (size_t)nmem * ctx->tab.KL * ctx->tab.ML * ctx->tab.NSEA;
    if (ctx->counter > 1 && ctx->adj_n < n) {
        if (ctx->adj_fp) (void)hipFree(ctx->adj_fp);
        if (ctx->adj_fb) (void)hipFree(ctx->adj_fb);
        ctx->adj_fp = ctx->adj_fb = nullptr;
        ctx->adj_n = 0;
        HIPCHK(hipMalloc(&ctx->adj_fp, n * sizeof(float)));
        HIPCHK(hipMalloc(&ctx->adj_fb, n * sizeof(float)));
        ctx->adj_n = n;
    }
    return propags_transposed(ctx, nmem, d_in_bar, d_out_bar, ctx->adj_fp, ctx->adj_fb);
}

int wam_gpu_wind_uv_dev(wam_ctx *ctx, const float *d_u, const float *d_v, float *d_u10, float *d_udir, int64_t n)
{
    if (!ctx || n < 0 || (n > 0 && (!d_u || !d_v || !d_u10 || !d_udir))) return WAM_E_ARG;
    HIPCHK(launch_wind_uv(ctx->tab.k.ZPI, d_u, d_v, d_u10, d_udir, n, ctx->stream));
    return WAM_OK;
}

int wam_gpu_wind_uv_adjoint_dev(wam_ctx *ctx, const float *d_u, const float *d_v, const float *d_u10_bar,
                                const float *d_udir_bar, float *d_u_bar, float *d_v_bar, int64_t n)
{
    if (!ctx || n < 0 || (n > 0 && (!d_u || !d_v || !d_u10_bar || !d_udir_bar || !d_u_bar || !d_v_bar)))
        return WAM_E_ARG;
    HIPCHK(launch_wind_uv_adjoint(d_u, d_v, d_u10_bar, d_udir_bar, d_u_bar, d_v_bar, n, ctx->stream));
    return WAM_OK;
}

int wam_gpu_set_stream(wam_ctx *ctx, void *stream)
{
    if (!ctx) return WAM_E_ARG;
    ctx->stream = stream ? (hipStream_t)stream : ctx->own_stream;
    return WAM_OK;
}

int wam_gpu_get_stream(wam_ctx *ctx, void **stream)
{
    if (!ctx || !stream) return WAM_E_ARG;
    *stream = (void *)ctx->stream;
    return WAM_OK;
}

int wam_gpu_sync(wam_ctx *ctx)
{
    if (!ctx) return WAM_E_ARG;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return WAM_OK;
}

int64_t wam_gpu_get_table(wam_ctx *ctx, const char *name, void *out, int64_t n)
{
    if (!ctx || !name) return -1;
    const Tables &t = ctx->tab;
    auto put = [&](const auto &v) -> int64_t {
        const int64_t len = (int64_t)v.size();
        if (out && n >= len) std::memcpy(out, v.data(), len * sizeof(v[0]));
        return len;
    };
    const std::string s(name);
    if (s == "FR") return put(t.FR);
    if (s == "DF") return put(t.DF);
    if (s == "DFIM") return put(t.DFIM);
    if (s == "DFIMOFR") return put(t.DFIMOFR);
    if (s == "DFIM_FR") return put(t.DFIM_FR);
    if (s == "DFIM_FR2") return put(t.DFIM_FR2);
    if (s == "FR5") return put(t.FR5);
    if (s == "FRM5") return put(t.FRM5);
    if (s == "RHOWG_DFIM") return put(t.RHOWG_DFIM);
    if (s == "GOM") return put(t.GOM);
    if (s == "TH") return put(t.TH);
    if (s == "COSTH") return put(t.COSTH);
    if (s == "SINTH") return put(t.SINTH);
    if (s == "FLMINFR") return put(t.FLMINFR);
    if (s == "TFAK") return put(t.TFAK);
    if (s == "TCGOND") return put(t.TCGOND);
    if (s == "TSIHKD") return put(t.TSIHKD);
    if (s == "IKP") return put(t.IKP);
    if (s == "IKP1") return put(t.IKP1);
    if (s == "IKM") return put(t.IKM);
    if (s == "IKM1") return put(t.IKM1);
    if (s == "K1W") return put(t.K1W);
    if (s == "K2W") return put(t.K2W);
    if (s == "K11W") return put(t.K11W);
    if (s == "K21W") return put(t.K21W);
    if (s == "AF11") return put(t.AF11);
    if (s == "FKLAP") return put(t.FKLAP);
    if (s == "FKLAP1") return put(t.FKLAP1);
    if (s == "FKLAM") return put(t.FKLAM);
    if (s == "FKLAM1") return put(t.FKLAM1);
    if (s == "FRH") return put(t.FRH);
    if (s == "FTRF") return put(t.FTRF);
    if (s == "INLCOEF") return put(t.INLCOEF);
    if (s == "RNLCOEF") return put(t.RNLCOEF);
    if (s == "SINPH") return put(t.SINPH);
    if (s == "COSPH") return put(t.COSPH);
    if (s == "DELLAM") return put(t.DELLAM);
    if (s == "IXLG") return put(t.IXLG);
    if (s == "KXLT") return put(t.KXLT);
    if (s == "DCO") return put(t.DCO);
    if (s == "INDEP") return put(t.INDEP);
    if (s == "INDEP_G") return put(t.INDEP_G);
    if (s == "DDPHI") return put(t.DDPHI);
    if (s == "DDLAM") return put(t.DDLAM);
    if (s == "DUPHI") return put(t.DUPHI);
    if (s == "DULAM") return put(t.DULAM);
    if (s == "DVPHI") return put(t.DVPHI);
    if (s == "DVLAM") return put(t.DVLAM);
    if (s == "DELLA0") return put(t.DELLA0);
    if (s == "DRGM") return put(t.DRGM);
    if (s == "NCONTRIB") return put(t.NCONTRIB);
    if (s == "CONTRIB") return put(t.CONTRIB);
    if (s == "SN_CH") return put(t.SN_CH);
    if (s == "SN_WW") return put(t.SN_WW);
    if (s == "SN_E") { std::vector<int> v(t.SN_E.begin(), t.SN_E.end()); return put(v); }
    if (s == "SP_RW") return put(t.SP_RW);
    if (s == "SP_E") { std::vector<int> v(t.SP_E.begin(), t.SP_E.end()); return put(v); }
    if (s == "KLAT" || s == "KLON" || s == "DPSN") {
        // reference layout: KLAT(IJ,IC,ICL) = [ICL][IC][IJ]; KLON(IJ,IC); DPSN(IJ,IC)
        std::vector<int> vi; std::vector<float> vf;
        if (s == "KLAT") {
            for (int icl = 0; icl < 2; icl++) { vi.insert(vi.end(), t.KLAT_S.begin(), t.KLAT_S.end()); vi.insert(vi.end(), t.KLAT_N.begin(), t.KLAT_N.end()); }
            return put(vi);
        }
        if (s == "KLON") { vi = t.KLON_W; vi.insert(vi.end(), t.KLON_E.begin(), t.KLON_E.end()); return put(vi); }
        vf = t.DPSN_S; vf.insert(vf.end(), t.DPSN_N.begin(), t.DPSN_N.end());
        return put(vf);
    }
    if ((s == "WN" || s == "CGG") && ctx->wn) {
        std::vector<float> h((size_t)t.ML * t.NSEA);
        if (hipMemcpy(h.data(), s == "WN" ? ctx->wn : ctx->cgg, h.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        return put(h);
    }
    if (s == "COEF" && ctx->coef) {
        // expanded to the nine reference planes (absent planes are zero)
        const size_t nb = (size_t)t.KL * t.ML * t.NSEA;
        std::vector<float> h((size_t)P_NPLANES * nb, 0.f), dev((size_t)ctx->ncoef * nb);
        if (hipMemcpy(dev.data(), ctx->coef, dev.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
        const int map5[5] = {P_SUM, P_LAT1, P_LON1, P_TH1, P_TH2};
        for (int pl = 0; pl < ctx->ncoef; pl++) {
            const int dst = ctx->ncoef == 9 ? pl : map5[pl];
            std::memcpy(h.data() + dst * nb, dev.data() + pl * nb, nb * 4);
        }
        return put(h);
    }
    return -1;
}

double wam_gpu_get_scalar(wam_ctx *ctx, const char *name)
{
    if (!ctx || !name) return NAN;
    const Tables &t = ctx->tab;
    const std::string s(name);
    if (s == "NSEA") return t.NSEA;
    if (s == "COUNTER") return ctx->counter;
    if (s == "CFL_MAX") return ctx->cfl_max;
    if (s == "CFL_MIN") return ctx->cfl_min;
    if (s == "ACL1") return t.ACL1;
    if (s == "ACL2") return t.ACL2;
    if (s == "CL11") return t.CL11;
    if (s == "CL21") return t.CL21;
    if (s == "DAL1") return t.DAL1;
    if (s == "DAL2") return t.DAL2;
    if (s == "ISP") return t.ISP;
    if (s == "ISM") return t.ISM;
    if (s == "MFRSTLW") return t.MFRSTLW;
    if (s == "MLSTHG") return t.MLSTHG;
    if (s == "KFRH") return t.KFRH;
    if (s == "X0TAUHF") return t.X0TAUHF;
    if (s == "DELPHI") return t.DELPHI;
    if (s == "DELTH") return t.DELTH;
    if (s == "DELU") return t.DELU;
    if (s == "INV_LOG_CO") return t.INV_LOG_CO;
    if (s == "MO_TAIL") return t.MO_TAIL;
    if (s == "MM1_TAIL") return t.MM1_TAIL;
    if (s == "MP1_TAIL") return t.MP1_TAIL;
    if (s == "NQ") return t.NQ;
    if (s == "NCMAX") return t.NCMAX;
    if (s == "CHUNK") return t.CHUNK;
    if (s == "NCHUNK") return t.NCHUNK;
    if (s == "SP_CHUNK") return t.SP_CHUNK;
    if (s == "SP_NCHUNK") return t.SP_NCHUNK;
    if (s == "SP_NRUN") return t.SP_NRUN;
    if (s == "SP_PAD") return t.SP_PAD;
    if (s == "NK10") return t.NK10;
    if (s == "ALPHA") return t.k.ALPHA;
    if (s == "NCOEF") return ctx->ncoef;
    return NAN;
}

int wam_gpu_eval_math(int32_t fn, const float *x, const float *y, float *out, int64_t n)
{
    if (fn < 0 || fn > 12 || !x || !out || n < 0 || ((fn == 8 || fn == 10 || fn == 11) && !y)) return WAM_E_ARG;
    if (n == 0) return WAM_OK;
    float *d = nullptr;
    const size_t bytes = (size_t)n * sizeof(float);
    if (hipMalloc(&d, 3 * bytes) != hipSuccess) return WAM_E_HIP;
    hipError_t e = hipMemcpy(d, x, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && y) e = hipMemcpy(d + n, y, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(eval_math_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, fn, d,
                           y ? d + n : nullptr, d + 2 * n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, d + 2 * n, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return e == hipSuccess ? WAM_OK : WAM_E_HIP;
}

const char *wam_gpu_last_error(wam_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void wam_gpu_finalize(wam_ctx *ctx)
{
    if (!ctx) return;
    free_list(ctx->point_allocs);
    free_list(ctx->dev_allocs);
    if (ctx->fscratch) (void)hipFree(ctx->fscratch);
    if (ctx->stage) (void)hipFree(ctx->stage);
    if (ctx->nest) (void)hipFree(ctx->nest);
    if (ctx->cell_pt) (void)hipFree(ctx->cell_pt);
    if (ctx->grids) (void)hipFree(ctx->grids);
    timing_clear(ctx);
    free_impl_scratch(ctx->isx);
    free_impl_scratch(ctx->esx);
    if (ctx->fscratch_e) (void)hipFree(ctx->fscratch_e);
    if (ctx->adj_fp) (void)hipFree(ctx->adj_fp);
    if (ctx->adj_fb) (void)hipFree(ctx->adj_fb);
    free_impl_scratch(ctx->asx);
    free_adj_extra(ctx->aex);
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    for (hipEvent_t e : ctx->gev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->gstream) (void)hipStreamDestroy(ctx->gstream);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

}  // extern "C"
