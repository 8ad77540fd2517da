// nsw6_api.cpp -- the C ABI of include/nsw6.h: context set-up, the mp_driver entry
// points and kernel timing.  Replaces MP_DRIVER::mp_init / mp_driver
// (src/mod_mp_driver.cpp:11-24, 335-613) for one GPU.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <mutex>
#include <new>
#include <vector>

#include "../../include/nsw6.h"
#include "nsw6_internal.h"
#include "nsw6_setup.h"

using namespace nsw6;

namespace {

constexpr int N_EVENTS = NSW6_STEP_EVENTS;
// step, rates, extremes+bounds, fluxes, finish; adjoint call, its forward part, its
// reverse passes; the wall fluxes of the general regions
constexpr int N_TIMERS = 9;

struct PendingTiming {
    hipEvent_t ev[N_EVENTS];
};
struct PendingAdjTiming {
    hipEvent_t ev[NSW6_ADJ_EVENTS];
};

// Scratch of the reverse mode: the records of the reverse passes, the per-block
// coefficient partials and the diagnostics the recomputed forward part writes (kept
// away from the caller's arrays).  Allocated on the first nsw6_mp_driver_adj call.
struct AdjScratch {
    double *rec = nullptr;    // [rec_slots][max_columns * KALL]
    int rec_slots = 0;        // adj_record_slots<NT>() of the scheme it was sized for
    double *diag = nullptr;   // re_liquid, re_cld, GPREC [K2]; 13 rows of [columns]
    double *part = nullptr;   // [blocks][NSW6_NCOEF]
    int part_blocks = 0;
    double *coef = nullptr;   // [NSW6_NCOEF] sink when the caller passes none
};

}  // namespace

struct nsw6_ctx {
    int device = -1;
    int64_t max_columns = 0;
    int region_capacity = 0;
    int precision = NSW6_FP64;
    DevParams P{};
    DevParams *P_dev = nullptr;
    GridDev *grid_dev = nullptr;
    Workspace ws{};
    hipStream_t last_stream = nullptr;
    int last_lall = 0, last_ijdim = 0;
    int last_nt = 3;                 // sedimenting tracers of the last step (kcell bounds layout)
    // where the last call left its report, region flags and kcell bounds: the
    // context's workspace, or the allocation-free call's scratch
    const int *last_report = nullptr, *last_region_flag = nullptr, *last_bounds = nullptr;
    bool timing = false;
    std::vector<PendingTiming> pending;
    double t_sum[N_TIMERS] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    int64_t t_count = 0;
    AdjScratch adj;
    bool last_adjoint = false;
    std::vector<PendingAdjTiming> pending_adj;
    int64_t t_adj_count = 0;
    int div_num = 1;                 // MP_DIV_NUM sub-steps per call
    double *div_sum = nullptr;       // their output sums (allocated at the first use)
    // reverse mode of a sub-cycled call: sub-step inputs, cotangent carries and scaled
    // output cotangents (grown on demand), and the general-region flags of all sub-steps
    double *div_adj = nullptr;
    size_t div_adj_doubles = 0;
    int *div_adj_flags = nullptr;
    int div_adj_regions = 0;
};

static int hip_ok(hipError_t e) { return e == hipSuccess ? NSW6_OK : NSW6_HIP_ERROR; }

// ---------------------------------------------------------------------------------
// Workspace of the allocation-free entry point nsw6_mp_driver_noalloc: one file-scope
// device array per GPU, sized for NSW6_NOALLOC_MAX_COLUMNS columns in at most
// NSW6_NOALLOC_MAX_REGIONS regions and carved into the step's workspace by
// noalloc_layout.  Its device address is looked up once per device at nsw6_init
// (hipGetSymbolAddress, with that device current) and kept in g_scratch_ptr; host code
// never passes the array's own name to a kernel or a copy.
// ---------------------------------------------------------------------------------
namespace {
constexpr size_t round_up(size_t n, size_t a) { return (n + a - 1) / a * a; }
constexpr size_t ALIGN_D = 32;   // 256 bytes, in doubles
struct NoallocLayout {           // offsets in doubles from the array's start
    size_t rhog, rhoge, rhogq, vt, flux, part, div_sum, ints, total;
};
constexpr NoallocLayout noalloc_layout()
{
    const size_t cols = NSW6_NOALLOC_MAX_COLUMNS, regions = NSW6_NOALLOC_MAX_REGIONS;
    const size_t K2 = (size_t)KALL * cols;
    NoallocLayout L{};
    size_t o = 0;
    L.rhog = o;  o = round_up(o + K2, ALIGN_D);
    L.rhoge = o; o = round_up(o + K2, ALIGN_D);
    L.rhogq = o; o = round_up(o + K2 * NQ, ALIGN_D);
    L.vt = o;    o = round_up(o + K2 * 4, ALIGN_D);
    L.flux = o;  o = round_up(o + K2 * FLUX_SLOTS_MAX, ALIGN_D);
    // per-block partial extremes: sum over regions of ceil(ijdim / block) blocks
    L.part = o;  o = round_up(o + (size_t)KALL * 16 * (cols / EXT_COLS_PER_BLOCK + regions), ALIGN_D);
    L.div_sum = o; o = round_up(o + cols * (KALL + DIV_COLUMN_SUMS), ALIGN_D);
    // ints: kcell bounds, region flags, report, cloud top (two per double)
    L.ints = o;
    const size_t n_int = (size_t)EXT_PER_REGION_MAX * regions + regions + REPORT_INTS + cols;
    o = round_up(o + (n_int + 1) / 2, ALIGN_D);
    L.total = o;
    return L;
}
constexpr NoallocLayout NOALLOC = noalloc_layout();
}  // namespace

static __device__ __align__(256) double scratch[NOALLOC.total];

static std::mutex g_scratch_mu;
constexpr int MAX_DEVICES = 64;
static double *g_scratch_ptr[MAX_DEVICES];   // device address of `scratch`, per device

// Look up `scratch` on the current device (once per device).
static int resolve_scratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return NSW6_BAD_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return NSW6_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return NSW6_HIP_ERROR;
    g_scratch_ptr[device] = static_cast<double *>(p);
    return NSW6_OK;
}

// The step workspace inside the scratch of `device` (its address as resolved above).
static Workspace noalloc_workspace(double *base, double **div_sum)
{
    Workspace w{};
    w.rhog = base + NOALLOC.rhog;
    w.rhoge = base + NOALLOC.rhoge;
    w.rhogq = base + NOALLOC.rhogq;
    w.vt = base + NOALLOC.vt;
    w.flux = base + NOALLOC.flux;
    w.part = base + NOALLOC.part;
    *div_sum = base + NOALLOC.div_sum;
    int *ip = reinterpret_cast<int *>(base + NOALLOC.ints);
    w.bounds = ip;
    ip += (size_t)EXT_PER_REGION_MAX * NSW6_NOALLOC_MAX_REGIONS;
    w.region_flag = ip;
    ip += NSW6_NOALLOC_MAX_REGIONS;
    w.report = ip;
    ip += REPORT_INTS;
    w.cloud_top = ip;
    return w;
}

static void free_ws(nsw6_ctx *c)
{
    Workspace &w = c->ws;
    void *ptrs[] = {w.rhog, w.rhoge, w.rhogq, w.vt, w.flux, w.part, w.bounds, w.region_flag, w.report, w.cloud_top,
                    c->adj.rec, c->adj.diag, c->adj.part, c->adj.coef};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    memset(&w, 0, sizeof w);
    c->adj = AdjScratch();
}

// Per-region buffers: kcell bounds, region flags and the per-block partial extremes
// (sized for any lall * ijdim <= max_columns split into `cap` regions).
static int ensure_regions(nsw6_ctx *c, int lall)
{
    if (lall <= c->region_capacity) return NSW6_OK;
    int cap = c->region_capacity ? c->region_capacity : 64;
    while (cap < lall) cap *= 2;
    Workspace &w = c->ws;
    if (w.part) (void)hipFree(w.part);
    if (w.bounds) (void)hipFree(w.bounds);
    if (w.region_flag) (void)hipFree(w.region_flag);
    w.part = nullptr; w.bounds = nullptr; w.region_flag = nullptr;
    const size_t blocks = (size_t)(c->max_columns / EXT_COLS_PER_BLOCK) + (size_t)cap;
    if (hipMalloc(&w.part, sizeof(double) * (size_t)KALL * 16 * blocks) != hipSuccess ||
        hipMalloc(&w.bounds, sizeof(int) * EXT_PER_REGION_MAX * (size_t)cap) != hipSuccess ||
        hipMalloc(&w.region_flag, sizeof(int) * (size_t)cap) != hipSuccess) {
        c->region_capacity = 0;
        return NSW6_HIP_ERROR;
    }
    c->region_capacity = cap;
    return NSW6_OK;
}

static Fields to_fields(const nsw6_fields &s)
{
    Fields f;
    f.rhog = s.rhog; f.rhogvx = s.rhogvx; f.rhogvy = s.rhogvy; f.rhogvz = s.rhogvz;
    f.rhogw = s.rhogw; f.rhoge = s.rhoge; f.rhogq = s.rhogq;
    f.vx = s.vx; f.vy = s.vy; f.vz = s.vz; f.w = s.w; f.unccn = s.unccn;
    f.rho = s.rho; f.tem = s.tem; f.pre = s.pre; f.q = s.q; f.qd = s.qd;
    f.precip = s.precip; f.ISO1_precip = s.ISO1_precip; f.ISO2_precip = s.ISO2_precip;
    f.precip_rhoe = s.precip_rhoe; f.precip_lh_heat = s.precip_lh_heat;
    f.precip_rhophi = s.precip_rhophi; f.precip_rhokin = s.precip_rhokin;
    f.re_liquid = s.re_liquid; f.re_solid = s.re_solid; f.re_cld = s.re_cld;
    f.rctop = s.rctop; f.rwtop = s.rwtop; f.tctop = s.tctop;
    f.frhoge_af = s.frhoge_af; f.frhogqv_af = s.frhogqv_af; f.frhoge_rad = s.frhoge_rad;
    f.rhogqke = s.rhogqke;
    f.gsgam2 = s.gsgam2; f.gsgam2h = s.gsgam2h; f.gam2 = s.gam2; f.gam2h = s.gam2h;
    f.ix = s.ix; f.iy = s.iy; f.iz = s.iz; f.jx = s.jx; f.jy = s.jy; f.jz = s.jz;
    f.z = s.z; f.zh = s.zh;
    f.GDCLW = s.GDCLW; f.GDCFRC = s.GDCFRC; f.GPREC = s.GPREC; f.CBMFX = s.CBMFX;
    return f;
}

// Everything of a step's arguments but the fields and the workspace.
static void fill_step_args(const nsw6_ctx *ctx, int lall, int ijdim, double dt, StepArgs &a)
{
    a.P = ctx->P_dev;
    a.grid = ctx->grid_dev;
    a.lall = lall;
    a.ijdim = ijdim;
    a.cloudtop_intended = ctx->P.cloudtop_intended != 0.0;
    a.opt_indir = ctx->P.opt_indir != 0.0;
    a.kin2kin = ctx->P.kin2kin != 0.0;
    a.four_water = ctx->P.four_water != 0.0;
    a.default_scheme = ctx->P.sw_kk2000 == 1.0 && ctx->P.sw_expice == 1.0 && ctx->P.sw_roh2014 == 1.0;
    a.dt = dt;
    a.dt2 = pow(dt, 2);
    a.dt3 = pow(dt, 3);
    a.dt3_6 = a.dt3 / 6.0;
}

static int use_device(const nsw6_ctx *ctx)
{
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return NSW6_HIP_ERROR;
    if (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess) return NSW6_HIP_ERROR;
    return NSW6_OK;
}

extern "C" {

int nsw6_init(const nsw6_grid *grid, const nsw6_params *params, int64_t max_columns, nsw6_ctx **out)
{
    if (!grid || !out || max_columns <= 0) return NSW6_BAD_ARG;
    // element offsets of a [l][6][k][ij] array must fit the 64-bit index math; the
    // bound below keeps every array well inside 2^31 elements per field
    if (max_columns * (int64_t)KALL * NQ >= ((int64_t)1 << 31)) return NSW6_BAD_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return NSW6_NO_DEVICE;
    nsw6_ctx *c = new (std::nothrow) nsw6_ctx();
    if (!c) return NSW6_BAD_ARG;
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return NSW6_HIP_ERROR; }
    c->max_columns = max_columns;
    nsw6_params p;
    if (params) p = *params; else nsw6_params_default(&p);
    to_dev_params(p, c->P);

    GridDev gh;
    make_grid_dev(grid->gz, grid->gzh, gh);

    const size_t K2 = (size_t)KALL * (size_t)max_columns;
    Workspace &w = c->ws;
    int st = NSW6_OK;
    if (hipMalloc(&c->grid_dev, sizeof(GridDev)) != hipSuccess ||
        hipMemcpy(c->grid_dev, &gh, sizeof(GridDev), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&c->P_dev, sizeof(DevParams)) != hipSuccess ||
        hipMemcpy(c->P_dev, &c->P, sizeof(DevParams), hipMemcpyHostToDevice) != hipSuccess ||
        hipMalloc(&w.rhog, sizeof(double) * K2) != hipSuccess ||
        hipMalloc(&w.rhoge, sizeof(double) * K2) != hipSuccess ||
        hipMalloc(&w.rhogq, sizeof(double) * K2 * NQ) != hipSuccess ||
        hipMalloc(&w.vt, sizeof(double) * K2 * 4) != hipSuccess ||
        hipMalloc(&w.flux, sizeof(double) * K2 * FLUX_SLOTS_MAX) != hipSuccess ||
        hipMalloc(&w.report, sizeof(int) * REPORT_INTS) != hipSuccess ||
        hipMemset(w.report, 0, sizeof(int) * REPORT_INTS) != hipSuccess ||
        hipMalloc(&w.cloud_top, sizeof(int) * (size_t)max_columns) != hipSuccess ||
        hipMemset(w.cloud_top, 0xff, sizeof(int) * (size_t)max_columns) != hipSuccess)
        st = NSW6_HIP_ERROR;
    if (st == NSW6_OK) st = ensure_regions(c, 64);
    if (st == NSW6_OK) st = resolve_scratch(c->device);
    if (st != NSW6_OK) {
        free_ws(c);
        if (c->grid_dev) (void)hipFree(c->grid_dev);
        if (c->P_dev) (void)hipFree(c->P_dev);
        delete c;
        return st;
    }
    *out = c;
    return NSW6_OK;
}

int nsw6_set_params(nsw6_ctx *ctx, const nsw6_params *params)
{
    if (!ctx || !params) return NSW6_BAD_ARG;
    DevParams P;
    to_dev_params(*params, P);
    // steps already queued read the old coefficients; let them finish first
    if (ctx->last_stream && hipStreamSynchronize(ctx->last_stream) != hipSuccess) return NSW6_HIP_ERROR;
    if (hipDeviceSynchronize() != hipSuccess) return NSW6_HIP_ERROR;
    if (hipMemcpy(ctx->P_dev, &P, sizeof(DevParams), hipMemcpyHostToDevice) != hipSuccess) return NSW6_HIP_ERROR;
    // the host copy drives the choice of kernel variants: change it only once the
    // device holds the same coefficients
    ctx->P = P;
    return NSW6_OK;
}

int nsw6_set_precision(nsw6_ctx *ctx, int precision)
{
    if (!ctx || (precision != NSW6_FP64 && precision != NSW6_MIXED)) return NSW6_BAD_ARG;
    ctx->precision = precision;
    return NSW6_OK;
}

// Arguments every mp_driver step needs (the context's column limit is checked by the
// caller).
static int check_step_args(const nsw6_ctx *ctx, int lall, int ijdim, const nsw6_fields *fl, double dt)
{
    if (!ctx || !fl || lall <= 0 || ijdim <= 0) return NSW6_BAD_ARG;
    if (!(dt > 0.0)) return NSW6_BAD_ARG;
    const double *req[] = {fl->rhog, fl->rhogvx, fl->rhogvy, fl->rhogvz, fl->rhogw, fl->rhoge, fl->rhogq,
                           fl->vx, fl->vy, fl->vz, fl->rho, fl->tem, fl->pre, fl->q, fl->qd, fl->precip,
                           fl->ISO1_precip, fl->ISO2_precip, fl->precip_rhoe, fl->precip_lh_heat,
                           fl->precip_rhophi, fl->precip_rhokin, fl->re_liquid, fl->re_solid, fl->re_cld,
                           fl->rctop, fl->rwtop, fl->tctop, fl->gsgam2, fl->gsgam2h, fl->gam2, fl->gam2h,
                           fl->ix, fl->iy, fl->iz, fl->jx, fl->jy, fl->jz, fl->z, fl->GDCLW, fl->GDCFRC,
                           fl->GPREC};
    for (const double *p : req)
        if (!p) return NSW6_BAD_ARG;
    if (ctx->P.opt_indir != 0.0 && !fl->unccn) return NSW6_BAD_ARG;   // read under OPT_INDIR
    return NSW6_OK;
}

// mp_driver's sub-step loop (src/mod_mp_driver.cpp:420-612) on the workspace of `a`:
// MP_DIV_NUM steps of dt_mp, their output sums in `div_sum` and the scaling by
// fraction_mp.  With `timed`, HIP events of each step go to the context's timers.
static int run_steps(nsw6_ctx *ctx, StepArgs &a, int lall, double dt, double *div_sum, hipStream_t s, bool timed)
{
    const int ndiv = ctx->div_num;
    // fraction_mp and dt_mp as src/mod_mp_driver.cpp:420-421
    const double fraction_mp = 1.0 / double(ndiv);
    fill_step_args(ctx, lall, a.ijdim, dt * fraction_mp, a);
    ctx->last_stream = s;
    ctx->last_lall = lall;
    ctx->last_ijdim = a.ijdim;
    ctx->last_nt = a.four_water ? 4 : 3;
    ctx->last_adjoint = false;
    ctx->last_report = a.ws.report;
    ctx->last_region_flag = a.ws.region_flag;
    ctx->last_bounds = a.ws.bounds;
    for (int m = 0; m < ndiv; m++) {
        a.keep_report = (m > 0);
        hipEvent_t *ev = nullptr;
        PendingTiming pt;
        if (timed) {
            for (int i = 0; i < N_EVENTS; i++)
                if (hipEventCreate(&pt.ev[i]) != hipSuccess) return NSW6_HIP_ERROR;
            ev = pt.ev;
        }
        hipError_t e = nsw6_launch_step(&a, ctx->precision, s, ev);
        if (timed) ctx->pending.push_back(pt);
        if (e == hipSuccess && ndiv > 1) e = nsw6_launch_div_sum(&a, div_sum, m == 0, s);
        if (e != hipSuccess) return NSW6_HIP_ERROR;
    }
    if (ndiv > 1) return hip_ok(nsw6_launch_div_scale(&a, div_sum, fraction_mp, s));
    return NSW6_OK;
}

int nsw6_mp_driver_fields(nsw6_ctx *ctx, int l_region, int lall, int ijdim, const nsw6_fields *fl,
                          double dt, double ct, void *stream)
{
    (void)l_region;   // only stored by the reference
    (void)ct;         // unused by the reference's NSW6 path
    int st = check_step_args(ctx, lall, ijdim, fl, dt);
    if (st != NSW6_OK) return st;
    if ((int64_t)lall * ijdim > ctx->max_columns) return NSW6_BAD_ARG;
    st = use_device(ctx);
    if (st == NSW6_OK) st = ensure_regions(ctx, lall);
    if (st != NSW6_OK) return st;

    if (ctx->div_num > 1 && !ctx->div_sum) {
        const size_t n = sizeof(double) * (size_t)ctx->max_columns * (KALL + DIV_COLUMN_SUMS);
        if (hipMalloc(&ctx->div_sum, n) != hipSuccess) {
            ctx->div_sum = nullptr;
            return NSW6_HIP_ERROR;
        }
    }
    StepArgs a;
    memset(&a, 0, sizeof a);
    a.f = to_fields(*fl);
    a.ws = ctx->ws;
    a.ijdim = ijdim;
    return run_steps(ctx, a, lall, dt, ctx->div_sum, (hipStream_t)stream, ctx->timing);
}

int nsw6_mp_driver_noalloc(nsw6_ctx *ctx, int l_region, int lall, int ijdim, const nsw6_fields *fl,
                           double dt, double ct, void *stream)
{
    (void)l_region;
    (void)ct;
    int st = check_step_args(ctx, lall, ijdim, fl, dt);
    if (st != NSW6_OK) return st;
    if (lall > NSW6_NOALLOC_MAX_REGIONS || (int64_t)lall * ijdim > NSW6_NOALLOC_MAX_COLUMNS) return NSW6_BAD_ARG;
    st = use_device(ctx);
    if (st != NSW6_OK) return st;
    double *base = g_scratch_ptr[ctx->device];   // resolved by nsw6_init on this device
    if (!base) return NSW6_HIP_ERROR;
    StepArgs a;
    memset(&a, 0, sizeof a);
    a.f = to_fields(*fl);
    double *div_sum = nullptr;
    a.ws = noalloc_workspace(base, &div_sum);
    a.ijdim = ijdim;
    // cloud_top needs no clearing: the array starts at 0, k_cloudtop puts back -1 after
    // each use, and both read as "no cloudy level" (levels 1..94 only count)
    return run_steps(ctx, a, lall, dt, div_sum, (hipStream_t)stream, false);
}

int nsw6_set_div_num(nsw6_ctx *ctx, int mp_div_num)
{
    if (!ctx || mp_div_num < 1 || mp_div_num > NSW6_MAX_DIV_NUM) return NSW6_BAD_ARG;
    ctx->div_num = mp_div_num;
    return NSW6_OK;
}

// The reverse mode covers the Roh & Satoh or the exponential snow moments, KK2000 or
// Berry's autoconversion, with or without explicit ice generation (without it: the
// Bergeron growth and the mixed-phase saturation adjustment), each with or without
// OPT_INDIR, under "3WATER" or "4WATER" (see include/nsw6.h).
static bool adjoint_supported(const DevParams &P)
{
    const bool kk = (P.sw_kk2000 == 1.0), berry = (P.sw_kk2000 == 0.0);
    const bool roh = (P.sw_roh2014 == 1.0), expo = (P.sw_roh2014 == 0.0);
    const bool expice = (P.sw_expice == 1.0 || P.sw_expice == 0.0);
    return (roh || expo) && (kk || berry) && expice;
}

static int ensure_adj_scratch(nsw6_ctx *ctx, int lall, int ijdim)
{
    AdjScratch &s = ctx->adj;
    const size_t K2 = (size_t)KALL * (size_t)ctx->max_columns;
    if (!s.diag) {
        if (hipMalloc(&s.diag, sizeof(double) * (3 * K2 + 13 * (size_t)ctx->max_columns)) != hipSuccess ||
            hipMalloc(&s.coef, sizeof(double) * NSW6_NCOEF) != hipSuccess) {
            (void)hipFree(s.diag);
            (void)hipFree(s.coef);
            s.diag = s.coef = nullptr;
            return NSW6_HIP_ERROR;
        }
    }
    // the records of the current scheme ("4WATER" needs more slots): grown when a
    // change of scheme by nsw6_set_params asks for more
    const int slots = ctx->P.four_water != 0.0 ? adj_record_slots<4>() : adj_record_slots<3>();
    if (slots > s.rec_slots) {
        // a queued adjoint may still read the old records
        if (s.rec && hipDeviceSynchronize() != hipSuccess) return NSW6_HIP_ERROR;
        (void)hipFree(s.rec);
        s.rec = nullptr;
        s.rec_slots = 0;
        if (hipMalloc(&s.rec, sizeof(double) * K2 * (size_t)slots) != hipSuccess) return NSW6_HIP_ERROR;
        s.rec_slots = slots;
    }
    const int nb = nsw6_adjoint_blocks(lall, ijdim);
    if (nb > s.part_blocks) {
        // a queued adjoint may still read the old partials
        if (s.part && hipDeviceSynchronize() != hipSuccess) return NSW6_HIP_ERROR;
        (void)hipFree(s.part);
        s.part = nullptr;
        s.part_blocks = 0;
        if (hipMalloc(&s.part, sizeof(double) * NSW6_NCOEF * (size_t)nb) != hipSuccess) return NSW6_HIP_ERROR;
        s.part_blocks = nb;
    }
    return NSW6_OK;
}

// One reverse step of dt (the whole of nsw6_mp_driver_adj at MP_DIV_NUM = 1).  The
// diagnostics of the recomputed forward part go to context scratch.
// One FP64 reverse-mode step on workspace `ws`; forward = 0 when `ws` already holds the
// forward part of `state` (nsw6_launch_adjoint).
static int adj_step(nsw6_ctx *ctx, int lall, int ijdim, const nsw6_fields *state, const nsw6_fields *cot,
                    const double *cot_sflux, const nsw6_fields *grad, double *grad_coef, double dt,
                    int keep_report, hipStream_t s, hipEvent_t *ev, const Workspace &ws, int forward)
{
    AdjArgs A;
    memset(&A, 0, sizeof A);
    StepArgs &a = A.a;
    a.f = to_fields(*state);
    const size_t K2 = (size_t)KALL * (size_t)lall * (size_t)ijdim, C = (size_t)lall * (size_t)ijdim;
    double *d = ctx->adj.diag;
    a.f.re_liquid = d;
    a.f.re_cld = d + K2;
    a.f.GPREC = d + 2 * K2;
    double *dc = d + 3 * K2;
    a.f.rctop = dc;
    a.f.tctop = dc + C;
    a.f.rwtop = dc + 2 * C;
    a.f.precip = dc + 3 * C;
    a.f.ISO1_precip = dc + 5 * C;
    a.f.ISO2_precip = dc + 7 * C;
    a.f.precip_rhoe = dc + 9 * C;
    a.f.precip_lh_heat = dc + 10 * C;
    a.f.precip_rhophi = dc + 11 * C;
    a.f.precip_rhokin = dc + 12 * C;
    fill_step_args(ctx, lall, ijdim, dt, a);
    a.ws = ws;
    a.keep_report = keep_report;
    A.cot = to_fields(*cot);
    A.cot_sflux = cot_sflux;
    A.grad = to_fields(*grad);
    // records [slot][l][k][ij] of the current batch shape and tracer count
    const bool four = a.four_water != 0;
    A.cr = ctx->adj.rec;
    A.wr = A.cr + (size_t)(four ? CellRec<4>::N : CellRec<3>::N) * K2;
    A.dws = A.wr + (size_t)(four ? WallRec<4>::N : WallRec<3>::N) * K2;
    A.din = A.dws + (size_t)NDW * K2;
    return hip_ok(nsw6_launch_adjoint(&A, ctx->adj.part, grad_coef ? grad_coef : ctx->adj.coef, s, ev, forward));
}

// The arrays one sub-step hands to the next: rhog, the momenta, rhoge, rhogq, tem and q
// (QV is read; the halo levels consume the rest), plus rho, pre and qd, which a step
// writes and the next overwrites.
struct Carry {
    double *rhog, *rhogvx, *rhogvy, *rhogvz, *rhogw, *rhoge, *rhogq, *tem, *q, *rho, *pre, *qd;
};
static constexpr int CARRY_K2 = 21;   // K2-sized rows of a Carry (rhogq and q hold 6 each)
static Carry carry_at(double *base, size_t K2)
{
    Carry c;
    double **one[] = {&c.rhog, &c.rhogvx, &c.rhogvy, &c.rhogvz, &c.rhogw, &c.rhoge, &c.tem, &c.rho, &c.pre};
    size_t o = 0;
    for (double **p : one) { *p = base + o; o += K2; }
    c.rhogq = base + o; o += 6 * K2;
    c.q = base + o; o += 6 * K2;
    c.qd = nullptr;   // written by the forward sub-steps into the diagnostic scratch
    return c;
}
static void put_carry(nsw6_fields &f, const Carry &c)
{
    f.rhog = c.rhog; f.rhogvx = c.rhogvx; f.rhogvy = c.rhogvy; f.rhogvz = c.rhogvz; f.rhogw = c.rhogw;
    f.rhoge = c.rhoge; f.rhogq = c.rhogq; f.tem = c.tem; f.q = c.q;
}

// Scratch of the sub-cycled reverse mode, sized for the call (grown, never shrunk).
static int ensure_div_adj(nsw6_ctx *ctx, size_t doubles, int lall)
{
    if (doubles <= ctx->div_adj_doubles && lall <= ctx->div_adj_regions) return NSW6_OK;
    // a queued call may still use the old buffers
    if (ctx->div_adj && hipDeviceSynchronize() != hipSuccess) return NSW6_HIP_ERROR;
    (void)hipFree(ctx->div_adj);
    (void)hipFree(ctx->div_adj_flags);
    ctx->div_adj = nullptr;
    ctx->div_adj_flags = nullptr;
    ctx->div_adj_doubles = 0;
    ctx->div_adj_regions = 0;
    if (hipMalloc(&ctx->div_adj, sizeof(double) * doubles) != hipSuccess ||
        hipMalloc(&ctx->div_adj_flags, sizeof(int) * (size_t)lall) != hipSuccess) {
        (void)hipFree(ctx->div_adj);
        ctx->div_adj = nullptr;
        ctx->div_adj_flags = nullptr;
        return NSW6_HIP_ERROR;
    }
    ctx->div_adj_doubles = doubles;
    ctx->div_adj_regions = lall;
    return NSW6_OK;
}

// Reverse mode of an mp_driver call of n = MP_DIV_NUM sub-steps of dt/n
// (src/mod_mp_driver.cpp:420-612): the sub-step inputs are rebuilt forward (n-1 state
// copies), then the sub-steps are reversed last to first.  The input cotangents of
// sub-step m are the output cotangents of sub-step m-1 (rho, pre and qd, overwritten
// by sub-step m, get none); the accumulated outputs (precip and its energy fluxes,
// GPREC, the surface mass fluxes) are fraction_mp times the sum over the sub-steps, so
// every sub-step receives fraction_mp times their cotangents; the cotangents of vx, vy,
// vz, unccn and of the coefficients are summed over the sub-steps.  Events (NULL or 3): start,
// after the forward rebuild, end.
static int adj_subcycled(nsw6_ctx *ctx, int lall, int ijdim, const nsw6_fields *state, const nsw6_fields *cot,
                         const double *cot_sflux, const nsw6_fields *grad, double *grad_coef, double dt,
                         hipStream_t s, hipEvent_t *ev)
{
    const int n = ctx->div_num;
    const double fraction_mp = 1.0 / double(n);
    const double dt_mp = dt * fraction_mp;
    const size_t C = (size_t)lall * (size_t)ijdim, K2 = (size_t)KALL * C;
    // layout: n-1 carries | 2 cotangent carries | vx, vy, vz, unccn | write-only rows (qd
    // and the constant diagnostics) | scaled GPREC | scaled column cotangents | coefficients
    // | a second step workspace (sized for "4WATER") for the forward part of the last
    // sub-step, so that the step workspace keeps the forward part of sub-step n-2 from the
    // rebuild
    const size_t n_carry = (size_t)(n - 1) * CARRY_K2 * K2, n_cot = 2 * (size_t)CARRY_K2 * K2;
    const size_t alt_part = (size_t)KALL * 16 * (C / EXT_COLS_PER_BLOCK + (size_t)lall);
    const size_t alt_ints = (size_t)EXT_PER_REGION_MAX * (size_t)lall + (size_t)lall + REPORT_INTS + C;
    const size_t n_alt = (size_t)(1 + 1 + NQ + 4 + FLUX_SLOTS_MAX) * K2 + alt_part + (alt_ints + 1) / 2;
    const size_t doubles = n_carry + n_cot + 4 * K2 + 2 * K2 + K2 + 10 * C + NSW6_NCOEF + n_alt;
    int st = ensure_div_adj(ctx, doubles, lall);
    if (st != NSW6_OK) return st;
    double *base = ctx->div_adj;
    double *cotbuf = base + n_carry;
    double *vtmp = cotbuf + n_cot;
    double *junk = vtmp + 4 * K2;   // qd (K2) and re_solid / GDCLW / GDCFRC (K2, same thread, same index)
    double *gprec_s = junk + 2 * K2;
    double *col_s = gprec_s + K2;
    double *coef_tmp = col_s + 10 * C;   // 6 C of column cotangents, NT C of surface fluxes
    Workspace alt{};
    {
        double *a = coef_tmp + NSW6_NCOEF;
        alt.rhog = a; a += K2;
        alt.rhoge = a; a += K2;
        alt.rhogq = a; a += NQ * K2;
        alt.vt = a; a += 4 * K2;
        alt.flux = a; a += (size_t)FLUX_SLOTS_MAX * K2;
        alt.part = a; a += alt_part;
        int *ip = reinterpret_cast<int *>(a);
        alt.bounds = ip; ip += (size_t)EXT_PER_REGION_MAX * (size_t)lall;
        alt.region_flag = ip; ip += lall;
        alt.report = ip; ip += REPORT_INTS;
        alt.cloud_top = ip;
    }
    auto carry = [&](int m) { return carry_at(base + (size_t)(m - 1) * CARRY_K2 * K2, K2); };   // m = 1..n-1
    auto cotc = [&](int i) { return carry_at(cotbuf + (size_t)i * CARRY_K2 * K2, K2); };
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t r) { if (e == hipSuccess) e = r; return e == hipSuccess; };
    auto copy = [&](double *dst, const double *src, size_t count) {
        return ok(hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToDevice, s));
    };

    if (ev && !ok(hipEventRecord(ev[0], s))) return NSW6_HIP_ERROR;
    // ---- forward: the inputs of sub-steps 1..n-1
    {
        StepArgs a;
        memset(&a, 0, sizeof a);
        a.f = to_fields(*state);
        double *d = ctx->adj.diag;
        a.f.re_liquid = d;
        a.f.re_cld = d + K2;
        a.f.GPREC = d + 2 * K2;
        double *dc = d + 3 * K2;
        a.f.rctop = dc;
        a.f.tctop = dc + C;
        a.f.rwtop = dc + 2 * C;
        a.f.precip = dc + 3 * C;
        a.f.ISO1_precip = dc + 5 * C;
        a.f.ISO2_precip = dc + 7 * C;
        a.f.precip_rhoe = dc + 9 * C;
        a.f.precip_lh_heat = dc + 10 * C;
        a.f.precip_rhophi = dc + 11 * C;
        a.f.precip_rhokin = dc + 12 * C;
        a.f.qd = junk;
        a.f.re_solid = a.f.GDCLW = a.f.GDCFRC = junk + K2;
        fill_step_args(ctx, lall, ijdim, dt_mp, a);
        a.ws = ctx->ws;
        for (int m = 1; m < n; m++) {
            const Carry dst = carry(m);
            nsw6_fields src_f = *state;
            if (m > 1) put_carry(src_f, carry(m - 1));
            const double *rows[] = {src_f.rhog, src_f.rhogvx, src_f.rhogvy, src_f.rhogvz, src_f.rhogw,
                                    src_f.rhoge, src_f.tem};
            double *rows_d[] = {dst.rhog, dst.rhogvx, dst.rhogvy, dst.rhogvz, dst.rhogw, dst.rhoge, dst.tem};
            for (int i = 0; i < 7; i++) copy(rows_d[i], rows[i], K2);
            copy(dst.rhogq, src_f.rhogq, 6 * K2);
            copy(dst.q, src_f.q, 6 * K2);
            a.f.rhog = dst.rhog; a.f.rhogvx = dst.rhogvx; a.f.rhogvy = dst.rhogvy; a.f.rhogvz = dst.rhogvz;
            a.f.rhogw = dst.rhogw; a.f.rhoge = dst.rhoge; a.f.rhogq = dst.rhogq; a.f.tem = dst.tem;
            a.f.q = dst.q; a.f.rho = dst.rho; a.f.pre = dst.pre;
            a.keep_report = (m > 1);
            if (!ok(nsw6_launch_step(&a, NSW6_FP64, s, nullptr))) return NSW6_HIP_ERROR;
        }
    }
    if (ev && !ok(hipEventRecord(ev[1], s))) return NSW6_HIP_ERROR;

    // ---- the accumulated outputs' cotangents, times fraction_mp (NULL stays NULL)
    nsw6_fields cdiag;
    memset(&cdiag, 0, sizeof cdiag);
    const double *csrc[] = {cot->precip, cot->precip_rhoe, cot->precip_lh_heat, cot->precip_rhophi,
                            cot->precip_rhokin};
    double **cdst[] = {&cdiag.precip, &cdiag.precip_rhoe, &cdiag.precip_lh_heat, &cdiag.precip_rhophi,
                       &cdiag.precip_rhokin};
    const size_t crows[] = {2 * C, C, C, C, C};
    size_t off = 0;
    for (int i = 0; i < 5; i++) {
        if (csrc[i]) {
            *cdst[i] = col_s + off;
            if (!ok(nsw6_launch_scale(*cdst[i], csrc[i], fraction_mp, (int64_t)crows[i], s))) return NSW6_HIP_ERROR;
        }
        off += crows[i];
    }
    double *sflux_s = nullptr;
    if (cot_sflux) {
        const size_t nt = ctx->P.four_water != 0.0 ? 4 : 3;
        sflux_s = col_s + off;   // nt C
        if (!ok(nsw6_launch_scale(sflux_s, cot_sflux, fraction_mp, (int64_t)(nt * C), s))) return NSW6_HIP_ERROR;
    }
    if (cot->GPREC) {
        cdiag.GPREC = gprec_s;
        if (!ok(nsw6_launch_scale(gprec_s, cot->GPREC, fraction_mp, (int64_t)K2, s))) return NSW6_HIP_ERROR;
    }

    // ---- reverse, last sub-step first
    for (int m = n - 1; m >= 0; m--) {
        nsw6_fields st_m = *state;
        if (m > 0) put_carry(st_m, carry(m));
        // output cotangents of sub-step m
        nsw6_fields c_m = cdiag;
        if (m == n - 1) {
            c_m.rhog = cot->rhog; c_m.rhogvx = cot->rhogvx; c_m.rhogvy = cot->rhogvy; c_m.rhogvz = cot->rhogvz;
            c_m.rhogw = cot->rhogw; c_m.rhoge = cot->rhoge; c_m.rhogq = cot->rhogq; c_m.tem = cot->tem;
            c_m.q = cot->q; c_m.rho = cot->rho; c_m.pre = cot->pre; c_m.qd = cot->qd;
        } else {
            put_carry(c_m, cotc((m + 1) & 1));   // written by sub-step m+1; rho, pre, qd stay NULL
        }
        // input cotangents of sub-step m
        nsw6_fields g_m;
        memset(&g_m, 0, sizeof g_m);
        if (m == 0) {
            g_m = *grad;
        } else {
            put_carry(g_m, cotc(m & 1));
        }
        const bool first = (m == n - 1);
        g_m.vx = grad->vx ? (first ? grad->vx : vtmp) : nullptr;
        g_m.vy = grad->vy ? (first ? grad->vy : vtmp + K2) : nullptr;
        g_m.vz = grad->vz ? (first ? grad->vz : vtmp + 2 * K2) : nullptr;
        g_m.unccn = grad->unccn ? (first ? grad->unccn : vtmp + 3 * K2) : nullptr;
        double *coef_m = grad_coef ? (first ? grad_coef : coef_tmp) : nullptr;
        // the rebuild left the forward part of sub-step n-2 in the step workspace: the last
        // sub-step recomputes its own in the second workspace, and sub-step n-2 reuses it
        const Workspace &use = (m == n - 1) ? alt : ctx->ws;
        st = adj_step(ctx, lall, ijdim, &st_m, &c_m, sflux_s, &g_m, coef_m, dt_mp, 1, s, nullptr, use,
                      m == n - 2 ? 0 : 1);
        if (st != NSW6_OK) return st;
        if (!first) {
            if (grad->vx && !ok(nsw6_launch_accumulate(grad->vx, vtmp, (int64_t)K2, s))) return NSW6_HIP_ERROR;
            if (grad->vy && !ok(nsw6_launch_accumulate(grad->vy, vtmp + K2, (int64_t)K2, s))) return NSW6_HIP_ERROR;
            if (grad->vz && !ok(nsw6_launch_accumulate(grad->vz, vtmp + 2 * K2, (int64_t)K2, s)))
                return NSW6_HIP_ERROR;
            if (grad->unccn && !ok(nsw6_launch_accumulate(grad->unccn, vtmp + 3 * K2, (int64_t)K2, s)))
                return NSW6_HIP_ERROR;
            if (grad_coef && !ok(nsw6_launch_accumulate(grad_coef, coef_tmp, NSW6_NCOEF, s))) return NSW6_HIP_ERROR;
        }
        // a region that took the general path in any sub-step is counted
        if (!ok(nsw6_launch_merge_flags(ctx->div_adj_flags, use.region_flag, lall, first, s)))
            return NSW6_HIP_ERROR;
    }
    if (!ok(hipMemcpyAsync(ctx->ws.region_flag, ctx->div_adj_flags, sizeof(int) * (size_t)lall,
                           hipMemcpyDeviceToDevice, s)))
        return NSW6_HIP_ERROR;
    if (ev && !ok(hipEventRecord(ev[2], s))) return NSW6_HIP_ERROR;
    return NSW6_OK;
}

int nsw6_mp_driver_adj(nsw6_ctx *ctx, int l_region, int lall, int ijdim, const nsw6_fields *state,
                       const nsw6_fields *cot, const double *cot_sflux, const nsw6_fields *grad,
                       double *grad_coef, double dt, double ct, void *stream)
{
    (void)l_region;
    (void)ct;
    if (!ctx || !state || !cot || !grad || lall <= 0 || ijdim <= 0) return NSW6_BAD_ARG;
    if ((int64_t)lall * ijdim > ctx->max_columns) return NSW6_BAD_ARG;
    if (!(dt > 0.0)) return NSW6_BAD_ARG;
    const nsw6_fields *fl = state;
    const double *req[] = {fl->rhog, fl->rhogvx, fl->rhogvy, fl->rhogvz, fl->rhogw, fl->rhoge, fl->rhogq,
                           fl->vx, fl->vy, fl->vz, fl->tem, fl->q, fl->gsgam2, fl->gsgam2h, fl->gam2,
                           fl->gam2h, fl->ix, fl->iy, fl->iz, fl->jx, fl->jy, fl->jz, fl->z};
    for (const double *p : req)
        if (!p) return NSW6_BAD_ARG;
    if (ctx->P.opt_indir != 0.0 && !fl->unccn) return NSW6_BAD_ARG;   // read under OPT_INDIR
    if (!adjoint_supported(ctx->P)) return NSW6_UNSUPPORTED;
    int st = use_device(ctx);
    if (st == NSW6_OK) st = ensure_regions(ctx, lall);
    if (st == NSW6_OK) st = ensure_adj_scratch(ctx, lall, ijdim);
    if (st != NSW6_OK) return st;

    hipStream_t s = (hipStream_t)stream;
    hipEvent_t *ev = nullptr;
    PendingAdjTiming pt;
    if (ctx->timing) {
        for (int i = 0; i < NSW6_ADJ_EVENTS; i++)
            if (hipEventCreate(&pt.ev[i]) != hipSuccess) return NSW6_HIP_ERROR;
        ev = pt.ev;
    }
    if (ctx->div_num == 1)
        st = adj_step(ctx, lall, ijdim, state, cot, cot_sflux, grad, grad_coef, dt, 0, s, ev, ctx->ws, 1);
    else
        st = adj_subcycled(ctx, lall, ijdim, state, cot, cot_sflux, grad, grad_coef, dt, s, ev);
    if (ctx->timing) ctx->pending_adj.push_back(pt);
    ctx->last_stream = s;
    ctx->last_lall = lall;
    ctx->last_ijdim = ijdim;
    ctx->last_nt = 3;
    ctx->last_adjoint = true;
    ctx->last_report = ctx->ws.report;
    ctx->last_region_flag = ctx->ws.region_flag;
    ctx->last_bounds = ctx->ws.bounds;
    return st;
}

int nsw6_mp_driver(nsw6_ctx *ctx, int lall, int ijdim, int l_region, double *rhog, double *rhogvx,
                   double *rhogvy, double *rhogvz, double *rhogw, double *rhoge, double *rhogq, double *vx,
                   double *vy, double *vz, double *w, double *unccn, double *rho, double *tem, double *pre,
                   double *q, double *qd, double *precip, double *ISO1_precip, double *ISO2_precip,
                   double *precip_rhoe, double *precip_lh_heat, double *precip_rhophi, double *precip_rhokin,
                   double *re_liquid, double *re_solid, double *re_cld, double *rctop, double *rwtop,
                   double *tctop, double *frhoge_af, double *frhogqv_af, double *frhoge_rad, double *rhogqke,
                   double *gsgam2, double *gsgam2h, double *gam2, double *gam2h, double *ix, double *iy,
                   double *iz, double *jx, double *jy, double *jz, double *z, double *zh, double dt, double ct,
                   double *GDCLW, double *GDCFRC, double *GPREC, double *CBMFX, void *stream)
{
    nsw6_fields f = {rhog,        rhogvx,        rhogvy,        rhogvz,      rhogw,          rhoge,
                     rhogq,       vx,            vy,            vz,          w,              unccn,
                     rho,         tem,           pre,           q,           qd,             precip,
                     ISO1_precip, ISO2_precip,   precip_rhoe,   precip_lh_heat, precip_rhophi, precip_rhokin,
                     re_liquid,   re_solid,      re_cld,        rctop,       rwtop,          tctop,
                     frhoge_af,   frhogqv_af,    frhoge_rad,    rhogqke,     gsgam2,         gsgam2h,
                     gam2,        gam2h,         ix,            iy,          iz,             jx,
                     jy,          jz,            z,             zh,          GDCLW,          GDCFRC,
                     GPREC,       CBMFX};
    return nsw6_mp_driver_fields(ctx, l_region, lall, ijdim, &f, dt, ct, stream);
}

int nsw6_last_status(nsw6_ctx *ctx, nsw6_report *r)
{
    if (!ctx || !r) return NSW6_BAD_ARG;
    memset(r, 0, sizeof *r);
    r->first_l = r->first_k = r->first_ij = -1;
    if (hipStreamSynchronize(ctx->last_stream) != hipSuccess) {
        r->status = NSW6_HIP_ERROR;
        return NSW6_HIP_ERROR;
    }
    if (!ctx->last_report) {   // no step yet
        r->status = NSW6_OK;
        return NSW6_OK;
    }
    int rep[REPORT_INTS];
    if (hipMemcpy(rep, ctx->last_report, sizeof rep, hipMemcpyDeviceToHost) != hipSuccess) {
        r->status = NSW6_HIP_ERROR;
        return NSW6_HIP_ERROR;
    }
    r->n_not_converged = rep[1];
    r->n_nan = rep[2];
    unsigned long long key;
    memcpy(&key, rep + REPORT_KEY, sizeof key);
    if (key != ~0ULL && ctx->last_ijdim > 0) {
        const unsigned long long per_l = (unsigned long long)KALL * (unsigned long long)ctx->last_ijdim;
        r->first_l = (int32_t)(key / per_l);
        r->first_k = (int32_t)((key % per_l) / (unsigned long long)ctx->last_ijdim);
        r->first_ij = (int32_t)(key % (unsigned long long)ctx->last_ijdim);
    }
    if (ctx->last_lall > 0) {
        std::vector<int> flags((size_t)ctx->last_lall);
        if (hipMemcpy(flags.data(), ctx->last_region_flag, sizeof(int) * flags.size(), hipMemcpyDeviceToHost) !=
            hipSuccess) {
            r->status = NSW6_HIP_ERROR;
            return NSW6_HIP_ERROR;
        }
        for (int v : flags) r->n_general_regions += (v != 0);
    }
    if (ctx->last_adjoint)
        r->status = NSW6_OK;
    else
        r->status = r->n_nan ? NSW6_NAN_DETECTED : (r->n_not_converged ? NSW6_NOT_CONVERGED : NSW6_OK);
    return r->status;
}

static int drain_timing(nsw6_ctx *ctx)
{
    for (PendingTiming &pt : ctx->pending) {
        if (hipEventSynchronize(pt.ev[N_EVENTS - 1]) != hipSuccess) return NSW6_HIP_ERROR;
        // timer <- (from event, to event)
        const int which[6] = {0, 1, 2, 3, 4, 8}, from[6] = {0, 0, 1, 2, 4, 3}, to[6] = {5, 1, 2, 3, 5, 4};
        for (int i = 0; i < 6; i++) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, pt.ev[from[i]], pt.ev[to[i]]);
            ctx->t_sum[which[i]] += ms;
        }
        ctx->t_count++;
        for (int i = 0; i < N_EVENTS; i++) (void)hipEventDestroy(pt.ev[i]);
    }
    ctx->pending.clear();
    for (PendingAdjTiming &pt : ctx->pending_adj) {
        if (hipEventSynchronize(pt.ev[NSW6_ADJ_EVENTS - 1]) != hipSuccess) return NSW6_HIP_ERROR;
        const int from[3] = {0, 0, 1}, to[3] = {2, 1, 2};
        for (int i = 0; i < 3; i++) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, pt.ev[from[i]], pt.ev[to[i]]);
            ctx->t_sum[5 + i] += ms;
        }
        ctx->t_adj_count++;
        for (int i = 0; i < NSW6_ADJ_EVENTS; i++) (void)hipEventDestroy(pt.ev[i]);
    }
    ctx->pending_adj.clear();
    return NSW6_OK;
}

int nsw6_kcell_bounds(nsw6_ctx *ctx, int lall, int32_t *out, int64_t n_out)
{
    if (!ctx || !out || lall <= 0 || lall > ctx->last_lall || !ctx->last_bounds) return NSW6_BAD_ARG;
    const int per = ctx->last_nt * 2 * KALL * 2;
    if (n_out < (int64_t)lall * per) return NSW6_BAD_ARG;
    if (hipStreamSynchronize(ctx->last_stream) != hipSuccess) return NSW6_HIP_ERROR;
    return hip_ok(hipMemcpy(out, ctx->last_bounds, sizeof(int32_t) * (size_t)lall * per, hipMemcpyDeviceToHost));
}

int nsw6_timing_enable(nsw6_ctx *ctx, int on)
{
    if (!ctx) return NSW6_BAD_ARG;
    ctx->timing = on != 0;
    return NSW6_OK;
}

int nsw6_timing_read(nsw6_ctx *ctx, int which, double *avg_ms, int64_t *count)
{
    if (!ctx || which < 0 || which >= N_TIMERS || !avg_ms) return NSW6_BAD_ARG;
    int st = drain_timing(ctx);
    if (st != NSW6_OK) return st;
    const int64_t n = (which >= 5 && which <= 7) ? ctx->t_adj_count : ctx->t_count;
    *avg_ms = n ? ctx->t_sum[which] / (double)n : 0.0;
    if (count) *count = n;
    return NSW6_OK;
}

int nsw6_timing_reset(nsw6_ctx *ctx)
{
    if (!ctx) return NSW6_BAD_ARG;
    int st = drain_timing(ctx);
    for (int i = 0; i < N_TIMERS; i++) ctx->t_sum[i] = 0.0;
    ctx->t_count = 0;
    ctx->t_adj_count = 0;
    return st;
}

void nsw6_finalize(nsw6_ctx *ctx)
{
    if (!ctx) return;
    (void)drain_timing(ctx);
    (void)hipDeviceSynchronize();
    free_ws(ctx);
    if (ctx->div_sum) (void)hipFree(ctx->div_sum);
    if (ctx->div_adj) (void)hipFree(ctx->div_adj);
    if (ctx->div_adj_flags) (void)hipFree(ctx->div_adj_flags);
    if (ctx->grid_dev) (void)hipFree(ctx->grid_dev);
    if (ctx->P_dev) (void)hipFree(ctx->P_dev);
    delete ctx;
}

}  // extern "C"
