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

using namespace nsw6;

namespace {

constexpr int N_EVENTS = NSW6_STEP_EVENTS;
constexpr int N_TIMERS = 5;   // step, rates, extremes+bounds, sweep, adjust

struct PendingTiming {
    hipEvent_t ev[N_EVENTS];
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
    double t_sum[N_TIMERS] = {0, 0, 0, 0, 0};
    int64_t t_count = 0;
};

static int hip_ok(hipError_t e) { return e == hipSuccess ? NSW6_OK : NSW6_HIP_ERROR; }

static void to_dev_params(const nsw6_params &p, DevParams &d)
{
    memset(&d, 0, sizeof d);
    d.N0r = p.N0r; d.N0s = p.N0s; d.N0g = p.N0g;
    d.Ar = p.Ar; d.As = p.As; d.Ag = p.Ag; d.Cr = p.Cr; d.Cs = p.Cs; d.Cg = p.Cg;
    d.GAM = p.GAM; d.GAM_2 = p.GAM_2; d.GAM_3 = p.GAM_3;
    d.GAM_1br = p.GAM_1br; d.GAM_2br = p.GAM_2br; d.GAM_3br = p.GAM_3br; d.GAM_3dr = p.GAM_3dr;
    d.GAM_6dr = p.GAM_6dr; d.GAM_1brdr = p.GAM_1brdr; d.GAM_5dr_h = p.GAM_5dr_h;
    d.GAM_1bs = p.GAM_1bs; d.GAM_2bs = p.GAM_2bs; d.GAM_3bs = p.GAM_3bs; d.GAM_3ds = p.GAM_3ds;
    d.GAM_1bsds = p.GAM_1bsds; d.GAM_5ds_h = p.GAM_5ds_h;
    d.GAM_1bg = p.GAM_1bg; d.GAM_3dg = p.GAM_3dg; d.GAM_1bgdg = p.GAM_1bgdg; d.GAM_5dg_h = p.GAM_5dg_h;
    d.sw_kk2000 = p.sw_kk2000; d.sw_roh2014 = p.sw_roh2014; d.ln10 = p.ln10;
    for (int i = 0; i < 10; i++) { d.coef_a[i] = p.coef_a[i]; d.coef_b[i] = p.coef_b[i]; }
    d.Eiw = p.Eiw; d.Erw = p.Erw; d.Esw = p.Esw; d.Egw = p.Egw; d.Eri = p.Eri;
    d.Esi = p.Esi; d.Egi = p.Egi; d.Esr = p.Esr; d.Egr = p.Egr; d.Egs = p.Egs;
    d.gamma_sacr = p.gamma_sacr; d.gamma_gacs = p.gamma_gacs; d.mi = p.mi;
    d.Nc_def = p.Nc_def; d.beta_saut = p.beta_saut; d.gamma_saut = p.gamma_saut;
    d.beta_gaut = p.beta_gaut; d.gamma_gaut = p.gamma_gaut; d.qicrt_saut = p.qicrt_saut;
    d.qscrt_gaut = p.qscrt_gaut;
    d.f1r = p.f1r; d.f2r = p.f2r; d.f1s = p.f1s; d.f2s = p.f2s; d.f1g = p.f1g; d.f2g = p.f2g;
    d.A_frz = p.A_frz; d.B_frz = p.B_frz;
    d.mi40 = p.mi40; d.mi50 = p.mi50; d.vti50 = p.vti50; d.Ri50 = p.Ri50;
    d.sw_expice = p.sw_expice; d.Nc_ihtr = p.Nc_ihtr; d.Di_max = p.Di_max; d.Di_a = p.Di_a;
    d.sw_constVti = p.sw_constVti; d.CONST_Vti = p.CONST_Vti;
    d.kk_aut_rate = p.kk_aut_rate; d.icenuc_coef = p.icenuc_coef;
    d.coef_dgam = p.coef_dgam; d.coef_xf = p.coef_xf;
    // constants the reference evaluates per point, computed once here
    d.ln_mi40 = log(p.mi40);
    d.ln_mi50 = log(p.mi50);
    d.Nc_pow_m179 = pow(p.Nc_def, -1.79);
    d.Dc = 0.146 - 5.964E-2 * log(p.Nc_def / 2000.0);
    d.LovR_liq = C_LHV0 / C_Rvap;
    d.LovR_ice = C_LHS0 / C_Rvap;
    d.RTEM00 = 1.0 / C_TEM00;
    d.dtemp_criteria = pow(10.0, (double)(-(15 + 1) / 2));
}

static void free_ws(nsw6_ctx *c)
{
    Workspace &w = c->ws;
    void *ptrs[] = {w.rhog, w.rhoge, w.rhogq, w.vt, w.flux, w.part, w.bounds, w.region_flag, w.report};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    memset(&w, 0, sizeof w);
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
    if (hipMalloc(&w.part, sizeof(double) * (size_t)KALL * 12 * blocks) != hipSuccess ||
        hipMalloc(&w.bounds, sizeof(int) * BOUNDS_PER_REGION * (size_t)cap) != hipSuccess ||
        hipMalloc(&w.region_flag, sizeof(int) * (size_t)cap) != hipSuccess) {
        c->region_capacity = 0;
        return NSW6_HIP_ERROR;
    }
    c->region_capacity = cap;
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
    nsw6_grid g;
    nsw6_grid_setup(grid->gz, grid->gzh, &g);
    for (int k = 0; k < KALL; k++) {
        gh.gz[k] = g.gz[k]; gh.gzh[k] = g.gzh[k];
        gh.dgz[k] = g.dgz[k]; gh.dgzh[k] = g.dgzh[k];
        gh.rdgz[k] = g.rdgz[k]; gh.rdgzh[k] = g.rdgzh[k];
        gh.afact[k] = g.afact[k]; gh.bfact[k] = g.bfact[k];
        gh.cfact[k] = g.cfact[k]; gh.dfact[k] = g.dfact[k];
        gh.gz_shift[k] = 0.0;
    }
    for (int k = KMIN; k <= KMAX + 1; k++) gh.gz_shift[k] = g.gz[k - 1];   // precip_transport :705-708
    for (int k = 0; k < KALL; k++) {
        gh.rsum_gz[k] = (k >= 1) ? 1.0 / (g.dgz[k - 1] + g.dgz[k]) : 0.0;
        gh.rsum_gzh[k] = (k >= 1) ? 1.0 / (g.dgzh[k - 1] + g.dgzh[k]) : 0.0;
    }

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
        hipMalloc(&w.vt, sizeof(double) * K2 * 3) != hipSuccess ||
        hipMalloc(&w.flux, sizeof(double) * K2 * FLUX_SLOTS) != hipSuccess ||
        hipMalloc(&w.report, sizeof(int) * REPORT_INTS) != hipSuccess ||
        hipMemset(w.report, 0, sizeof(int) * REPORT_INTS) != hipSuccess)
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
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return NSW6_HIP_ERROR;
    if (cur != ctx->device && hipSetDevice(ctx->device) != hipSuccess) return NSW6_HIP_ERROR;
    int st = ensure_regions(ctx, lall);
    if (st != NSW6_OK) return st;

    StepArgs a;
    memset(&a, 0, sizeof a);
    Fields &f = a.f;
    f.rhog = fl->rhog; f.rhogvx = fl->rhogvx; f.rhogvy = fl->rhogvy; f.rhogvz = fl->rhogvz;
    f.rhogw = fl->rhogw; f.rhoge = fl->rhoge; f.rhogq = fl->rhogq;
    f.vx = fl->vx; f.vy = fl->vy; f.vz = fl->vz; f.w = fl->w; f.unccn = fl->unccn;
    f.rho = fl->rho; f.tem = fl->tem; f.pre = fl->pre; f.q = fl->q; f.qd = fl->qd;
    f.precip = fl->precip; f.ISO1_precip = fl->ISO1_precip; f.ISO2_precip = fl->ISO2_precip;
    f.precip_rhoe = fl->precip_rhoe; f.precip_lh_heat = fl->precip_lh_heat;
    f.precip_rhophi = fl->precip_rhophi; f.precip_rhokin = fl->precip_rhokin;
    f.re_liquid = fl->re_liquid; f.re_solid = fl->re_solid; f.re_cld = fl->re_cld;
    f.rctop = fl->rctop; f.rwtop = fl->rwtop; f.tctop = fl->tctop;
    f.frhoge_af = fl->frhoge_af; f.frhogqv_af = fl->frhogqv_af; f.frhoge_rad = fl->frhoge_rad;
    f.rhogqke = fl->rhogqke;
    f.gsgam2 = fl->gsgam2; f.gsgam2h = fl->gsgam2h; f.gam2 = fl->gam2; f.gam2h = fl->gam2h;
    f.ix = fl->ix; f.iy = fl->iy; f.iz = fl->iz; f.jx = fl->jx; f.jy = fl->jy; f.jz = fl->jz;
    f.z = fl->z; f.zh = fl->zh;
    f.GDCLW = fl->GDCLW; f.GDCFRC = fl->GDCFRC; f.GPREC = fl->GPREC; f.CBMFX = fl->CBMFX;
    a.ws = ctx->ws;
    a.P = ctx->P_dev;
    a.grid = ctx->grid_dev;
    a.lall = lall;
    a.ijdim = ijdim;
    a.dt = dt;
    a.dt2 = pow(dt, 2);
    a.dt3 = pow(dt, 3);
    a.dt3_6 = a.dt3 / 6.0;

    hipStream_t s = (hipStream_t)stream;
    hipEvent_t *ev = nullptr;
    PendingTiming pt;
    if (ctx->timing) {
        for (int i = 0; i < N_EVENTS; i++)
            if (hipEventCreate(&pt.ev[i]) != hipSuccess) return NSW6_HIP_ERROR;
        ev = pt.ev;
    }
    hipError_t e = nsw6_launch_step(&a, ctx->precision, s, ev);
    if (ctx->timing) ctx->pending.push_back(pt);
    ctx->last_stream = s;
    ctx->last_lall = lall;
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
    *avg_ms = ctx->t_count ? ctx->t_sum[which] / (double)ctx->t_count : 0.0;
    if (count) *count = ctx->t_count;
    return NSW6_OK;
}

int nsw6_timing_reset(nsw6_ctx *ctx)
{
    if (!ctx) return NSW6_BAD_ARG;
    int st = drain_timing(ctx);
    for (int i = 0; i < N_TIMERS; i++) ctx->t_sum[i] = 0.0;
    ctx->t_count = 0;
    return st;
}

void nsw6_finalize(nsw6_ctx *ctx)
{
    if (!ctx) return;
    (void)drain_timing(ctx);
    (void)hipDeviceSynchronize();
    free_ws(ctx);
    if (ctx->grid_dev) (void)hipFree(ctx->grid_dev);
    if (ctx->P_dev) (void)hipFree(ctx->P_dev);
    delete ctx;
}

}  // extern "C"
