// nsw6_api.cpp -- the C ABI of include/nsw6.h: context set-up, the mp_driver entry
// points and kernel timing.  Replaces MP_DRIVER::mp_init / mp_driver
// (src/mod_mp_driver.cpp:11-24, 335-613) for one GPU.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <new>
#include <vector>

#include "../../include/nsw6.h"
#include "nsw6_internal.h"
#include "nsw6_setup.h"

using namespace nsw6;

namespace {

constexpr int N_EVENTS = NSW6_STEP_EVENTS;
// step, rates, extremes+bounds, fluxes, finish; adjoint call, its forward part, its
// reverse passes
constexpr int N_TIMERS = 8;

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
    double *rec = nullptr;    // [ADJ_RECORD_SLOTS][max_columns * KALL]
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
    int last_lall = 0;
    bool timing = false;
    std::vector<PendingTiming> pending;
    double t_sum[N_TIMERS] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t t_count = 0;
    AdjScratch adj;
    bool last_adjoint = false;
    std::vector<PendingAdjTiming> pending_adj;
    int64_t t_adj_count = 0;
    int div_num = 1;                 // MP_DIV_NUM sub-steps per call
    double *div_sum = nullptr;       // their output sums (allocated at the first use)
};

static int hip_ok(hipError_t e) { return e == hipSuccess ? NSW6_OK : NSW6_HIP_ERROR; }

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

static void fill_step_args(const nsw6_ctx *ctx, int lall, int ijdim, double dt, StepArgs &a)
{
    a.ws = ctx->ws;
    a.P = ctx->P_dev;
    a.grid = ctx->grid_dev;
    a.lall = lall;
    a.ijdim = ijdim;
    a.cloudtop_intended = ctx->P.cloudtop_intended != 0.0;
    a.opt_indir = ctx->P.opt_indir != 0.0;
    a.kin2kin = ctx->P.kin2kin != 0.0;
    a.four_water = ctx->P.four_water != 0.0;
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
    to_dev_params(*params, ctx->P);
    // steps already queued read the old coefficients; let them finish first
    if (ctx->last_stream && hipStreamSynchronize(ctx->last_stream) != hipSuccess) return NSW6_HIP_ERROR;
    if (hipDeviceSynchronize() != hipSuccess) return NSW6_HIP_ERROR;
    return hip_ok(hipMemcpy(ctx->P_dev, &ctx->P, sizeof(DevParams), hipMemcpyHostToDevice));
}

int nsw6_set_precision(nsw6_ctx *ctx, int precision)
{
    if (!ctx || (precision != NSW6_FP64 && precision != NSW6_MIXED)) return NSW6_BAD_ARG;
    ctx->precision = precision;
    return NSW6_OK;
}

int nsw6_mp_driver_fields(nsw6_ctx *ctx, int l_region, int lall, int ijdim, const nsw6_fields *fl,
                          double dt, double ct, void *stream)
{
    (void)l_region;   // only stored by the reference
    (void)ct;         // unused by the reference's NSW6 path
    if (!ctx || !fl || lall <= 0 || ijdim <= 0) return NSW6_BAD_ARG;
    if ((int64_t)lall * ijdim > ctx->max_columns) return NSW6_BAD_ARG;
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
    int st = use_device(ctx);
    if (st == NSW6_OK) st = ensure_regions(ctx, lall);
    if (st != NSW6_OK) return st;

    const int ndiv = ctx->div_num;
    if (ndiv > 1 && !ctx->div_sum) {
        const size_t n = sizeof(double) * (size_t)ctx->max_columns * (KALL + DIV_COLUMN_SUMS);
        if (hipMalloc(&ctx->div_sum, n) != hipSuccess) {
            ctx->div_sum = nullptr;
            return NSW6_HIP_ERROR;
        }
    }

    // fraction_mp and dt_mp as src/mod_mp_driver.cpp:420-421
    const double fraction_mp = 1.0 / double(ndiv);
    StepArgs a;
    memset(&a, 0, sizeof a);
    a.f = to_fields(*fl);
    fill_step_args(ctx, lall, ijdim, dt * fraction_mp, a);

    hipStream_t s = (hipStream_t)stream;
    ctx->last_stream = s;
    ctx->last_lall = lall;
    ctx->last_adjoint = false;
    for (int m = 0; m < ndiv; m++) {
        a.keep_report = (m > 0);
        hipEvent_t *ev = nullptr;
        PendingTiming pt;
        if (ctx->timing) {
            for (int i = 0; i < N_EVENTS; i++)
                if (hipEventCreate(&pt.ev[i]) != hipSuccess) return NSW6_HIP_ERROR;
            ev = pt.ev;
        }
        hipError_t e = nsw6_launch_step(&a, ctx->precision, s, ev);
        if (ctx->timing) ctx->pending.push_back(pt);
        if (e == hipSuccess && ndiv > 1) e = nsw6_launch_div_sum(&a, ctx->div_sum, m == 0, s);
        if (e != hipSuccess) return NSW6_HIP_ERROR;
    }
    if (ndiv > 1) return hip_ok(nsw6_launch_div_scale(&a, ctx->div_sum, fraction_mp, s));
    return NSW6_OK;
}

int nsw6_set_div_num(nsw6_ctx *ctx, int mp_div_num)
{
    if (!ctx || mp_div_num < 1 || mp_div_num > NSW6_MAX_DIV_NUM) return NSW6_BAD_ARG;
    ctx->div_num = mp_div_num;
    return NSW6_OK;
}

// The reverse mode covers the default scheme only (see include/nsw6.h).
static bool adjoint_supported(const DevParams &P)
{
    return P.sw_roh2014 == 1.0 && P.sw_kk2000 == 1.0 && P.sw_expice == 1.0 && P.opt_indir == 0.0 &&
           P.kin2kin == 0.0 && P.four_water == 0.0;
}

static int ensure_adj_scratch(nsw6_ctx *ctx, int lall, int ijdim)
{
    AdjScratch &s = ctx->adj;
    const size_t K2 = (size_t)KALL * (size_t)ctx->max_columns;
    if (!s.rec) {
        if (hipMalloc(&s.rec, sizeof(double) * K2 * ADJ_RECORD_SLOTS) != hipSuccess ||
            hipMalloc(&s.diag, sizeof(double) * (3 * K2 + 13 * (size_t)ctx->max_columns)) != hipSuccess ||
            hipMalloc(&s.coef, sizeof(double) * NSW6_NCOEF) != hipSuccess) {
            (void)hipFree(s.rec);
            (void)hipFree(s.diag);
            (void)hipFree(s.coef);
            s = AdjScratch();
            return NSW6_HIP_ERROR;
        }
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
    if (!adjoint_supported(ctx->P) || ctx->div_num != 1) return NSW6_UNSUPPORTED;
    int st = use_device(ctx);
    if (st == NSW6_OK) st = ensure_regions(ctx, lall);
    if (st == NSW6_OK) st = ensure_adj_scratch(ctx, lall, ijdim);
    if (st != NSW6_OK) return st;

    AdjArgs A;
    memset(&A, 0, sizeof A);
    StepArgs &a = A.a;
    a.f = to_fields(*state);
    // diagnostics of the recomputed forward part go to scratch
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
    A.cot = to_fields(*cot);
    A.cot_sflux = cot_sflux;
    A.grad = to_fields(*grad);
    // records [slot][l][k][ij] of the current batch shape
    A.cr = ctx->adj.rec;
    A.wr = A.cr + (size_t)NCR * K2;
    A.dws = A.wr + (size_t)NWR * K2;
    A.din = A.dws + (size_t)NDW * K2;

    hipStream_t s = (hipStream_t)stream;
    hipEvent_t *ev = nullptr;
    PendingAdjTiming pt;
    if (ctx->timing) {
        for (int i = 0; i < NSW6_ADJ_EVENTS; i++)
            if (hipEventCreate(&pt.ev[i]) != hipSuccess) return NSW6_HIP_ERROR;
        ev = pt.ev;
    }
    hipError_t e = nsw6_launch_adjoint(&A, ctx->adj.part, grad_coef ? grad_coef : ctx->adj.coef, s, ev);
    if (ctx->timing) ctx->pending_adj.push_back(pt);
    ctx->last_stream = s;
    ctx->last_lall = lall;
    ctx->last_adjoint = true;
    return hip_ok(e);
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
    int rep[REPORT_INTS];
    if (hipMemcpy(rep, ctx->ws.report, sizeof rep, hipMemcpyDeviceToHost) != hipSuccess) {
        r->status = NSW6_HIP_ERROR;
        return NSW6_HIP_ERROR;
    }
    r->n_not_converged = rep[1];
    r->n_nan = rep[2];
    r->first_l = rep[3];
    r->first_k = rep[4];
    r->first_ij = rep[5];
    if (ctx->last_lall > 0) {
        std::vector<int> flags((size_t)ctx->last_lall);
        if (hipMemcpy(flags.data(), ctx->ws.region_flag, sizeof(int) * flags.size(), hipMemcpyDeviceToHost) !=
            hipSuccess) {
            r->status = NSW6_HIP_ERROR;
            return NSW6_HIP_ERROR;
        }
        for (int v : flags) r->n_general_regions += (v != 0);
    }
    if (ctx->last_adjoint)
        r->status = r->n_general_regions ? NSW6_UNSUPPORTED : NSW6_OK;
    else
        r->status = r->n_nan ? NSW6_NAN_DETECTED : (r->n_not_converged ? NSW6_NOT_CONVERGED : NSW6_OK);
    return r->status;
}

static int drain_timing(nsw6_ctx *ctx)
{
    for (PendingTiming &pt : ctx->pending) {
        if (hipEventSynchronize(pt.ev[N_EVENTS - 1]) != hipSuccess) return NSW6_HIP_ERROR;
        const int from[N_TIMERS] = {0, 0, 1, 2, 3}, to[N_TIMERS] = {4, 1, 2, 3, 4};
        for (int i = 0; i < N_TIMERS; i++) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, pt.ev[from[i]], pt.ev[to[i]]);
            ctx->t_sum[i] += ms;
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
    const int64_t n = (which >= 5) ? ctx->t_adj_count : ctx->t_count;
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
    if (ctx->grid_dev) (void)hipFree(ctx->grid_dev);
    if (ctx->P_dev) (void)hipFree(ctx->P_dev);
    delete ctx;
}

}  // extern "C"
