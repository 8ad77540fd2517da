// nsw6_columns.h -- column access of the transport, shared by the HIP kernels
// (forward k_flux / k_flux_general / k_finish and the reverse passes) and the host
// check of the reverse passes: gathers of one cell, one wall window or one cell's
// divergence inputs from the step workspace, and the per-wall flux computations.
//
// In a single-cell region (every departure cell in {k, k+1}) each flux of wall w
// depends only on cells w-1 (rhog, q), w, w+1 and the fall speeds Vt(w-3..w+1)
// (src/mod_vadv1d.cpp:393-450 with the kcell bounds of :317-352), so one thread per
// wall computes it without a column sweep.  In the other regions a wall's flux also
// reads the cells of the region's kcell range at that wall, still one thread per wall.
#pragma once

#include <stdint.h>

#include "nsw6_internal.h"
#include "nsw6_step.h"

namespace nsw6 {

// Vt(k-3 .. k+1) of one tracer of one column (levels below 0 are never used: they read
// level 0 and take 0, so the loads need no branch and issue together).  The fall speeds
// are stored as WS: double, or float in the mixed-precision form, whose FP32 rate
// kernel computes them in float (the stored values are its results exactly).
template <typename WS = double>
NSW6_HD void load_vt_window(const WS *vt_col, int64_t ijdim, int k, double V[5])
{
#pragma unroll
    for (int d = -3; d <= 1; d++) {
        const int kk = (k + d >= 0) ? k + d : 0;
        const double v = (double)vt_col[(int64_t)kk * ijdim];
        V[d + 3] = (k + d >= 0) ? v : 0.0;
    }
}
// The fall-speed and flux workspace as WS (the arrays are allocated as doubles; a float
// form uses the first half of each).
template <typename WS> NSW6_HD WS *ws_vt(const StepArgs &a) { return reinterpret_cast<WS *>(a.ws.vt); }
template <typename WS> NSW6_HD WS *ws_flux(const StepArgs &a) { return reinterpret_cast<WS *>(a.ws.flux); }

// Pre-sedimentation cell j of the column at p2 = l*K2 + ij, p3 = l*NQ*K2 + ij.
NSW6_HD void load_cell(const StepArgs &a, int64_t p2, int64_t p3, int64_t K2, int ijdim, int j, CellIn<double> &r)
{
    const Fields &f = a.f;
    const int64_t o = p2 + (int64_t)j * ijdim;
    const int64_t oq = p3 + (int64_t)j * ijdim;
    r.rhog = a.ws.rhog[o];
#pragma unroll
    for (int n = 1; n <= 5; n++) r.rq[n - 1] = a.ws.rhogq[oq + n * K2];
    r.rhoge = a.ws.rhoge[o];
    r.gs = f.gsgam2[o];
    r.gam2 = f.gam2[o];
    r.gsh = f.gsgam2h[o];
    r.gam2h = f.gam2h[o];
    r.rvx = f.rhogvx[o];
    r.rvy = f.rhogvy[o];
    r.rvz = f.rhogvz[o];
    r.rw = f.rhogw[o];
    r.vx = f.vx[o];
    r.vy = f.vy[o];
    r.vz = f.vz[o];
    r.z = f.z[o];
}

NSW6_HD Unit load_unit(const Fields &f, int64_t pc)
{
    Unit u;
    u.ix = f.ix[pc];
    u.iy = f.iy[pc];
    u.iz = f.iz[pc];
    u.jx = f.jx[pc];
    u.jy = f.jy[pc];
    u.jz = f.jz[pc];
    return u;
}

// Cell w-1 enters the walls of cell w only through rhog, q of the sedimenting tracers
// and gsgam2.
template <int NT> struct BelowCellN {
    double rhog, q[NT], gs;
};
using BelowCell = BelowCellN<3>;
template <int NT = 3>
NSW6_HD BelowCellN<NT> load_below(const StepArgs &a, int64_t p2, int64_t p3, int64_t K2, int ijdim, int w)
{
    const int64_t om = p2 + (int64_t)(w - 1) * ijdim, oqm = p3 + (int64_t)(w - 1) * ijdim;
    BelowCellN<NT> b;
    b.rhog = a.ws.rhog[om];
    b.gs = a.f.gsgam2[om];
#pragma unroll
    for (int t = 0; t < NT; t++) b.q[t] = a.ws.rhogq[oqm + sed_nq<NT>(t) * K2] / b.rhog;
    return b;
}

// Everything the flux of wall w (1..94) of column (l, ij) is computed from.
template <int NT> struct WallStateN {
    Unit u;
    BelowCellN<NT> m;
    CellIn<double> r0, r1;
    CellN<double, NT> cur, nxt;
    double V[NT][5];
    double z0[NT], zh0[NT];
    WallGeomN<NT> geo;
};
using WallState = WallStateN<3>;
// Every load of the wall first (the cells, the fall speeds, the kcell bounds), then the
// derived cells and the departure geometry: one memory round trip per wall.
template <int NT, typename WS = double>
NSW6_HD void wall_state(const StepArgs &a, int l, int w, int ij, WallStateN<NT> &s)
{
    const int ijdim = a.ijdim;
    const GridDev &g = *a.grid;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const int64_t pv = (int64_t)l * NT * K2 + ij;
    const int *bnd = a.ws.bounds + (size_t)l * ext_per_region<NT>();
    const int64_t p2 = (int64_t)l * K2 + ij, p3 = (int64_t)l * NQ * K2 + ij;
    s.u = load_unit(a.f, (int64_t)l * ijdim + ij);
    s.m = load_below<NT>(a, p2, p3, K2, ijdim, w);
    load_cell(a, p2, p3, K2, ijdim, w, s.r0);
    load_cell(a, p2, p3, K2, ijdim, w + 1, s.r1);
    int kcmax[NT], khmax[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        load_vt_window(ws_vt<WS>(a) + pv + t * K2, ijdim, w, s.V[t]);
        kcmax[t] = bnd[bounds_index(t, 0, w, 1)];
        khmax[t] = bnd[bounds_index(t, 1, w, 1)];
    }
    derive_cell(g, w, s.r0, s.m.rhog, s.m.q, s.m.gs, s.u, s.cur);
    derive_cell(g, w + 1, s.r1, s.cur.rhog, s.cur.q, s.cur.gs, s.u, s.nxt);
    wall_geometry(g, w, a.dt, a.dt2, a.dt3_6, s.V, kcmax, khmax, s.z0, s.zh0, s.geo);
}

// Flux slots of wall w (FxSlots<NT>) through put(slot, value), frain = GPREC
// (src/mod_precip_transport.cpp:916-927) and, at wall kmin, the surface fluxes.  Walls 0
// and 95 carry no flux.
template <int NT, typename Put, typename WS = double>
NSW6_HD void flux_wall_to(const StepArgs &a, int l, int w, int ij, Put &&put)
{
    using X = FxSlots<NT>;
    const int ijdim = a.ijdim;
    const Fields &f = a.f;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const int64_t p2 = (int64_t)l * K2 + ij;
    const int64_t ow = p2 + (int64_t)w * ijdim;

    if (w == 0 || w == KALL - 1) {
#pragma unroll
        for (int i = 0; i < X::RW; i++) put(i, 0.0);
        put(X::RW, f.rhogw[ow]);
        f.GPREC[ow] = (0.0 + 0.0) * 1.0;
        return;
    }
    // Every store comes after every load: a store before them would make the compiler
    // fetch the grid and kcell tables (uniform) through the vector memory path, one
    // round trip per step of the departure-cell search, instead of the scalar cache.
    WallStateN<NT> s;
    wall_state<NT, WS>(a, l, w, ij, s);
    WallFluxN<double, NT> F;
    double ft[NT][4];
    wall_flux(*a.grid, w, s.cur, s.nxt, s.z0, s.zh0, s.geo, F, ft);

#pragma unroll
    for (int t = 0; t < NT; t++) put(X::Q + t, F.fq[t]);
#pragma unroll
    for (int i = 0; i < 5; i++) put(X::E + i, F.fs[i]);
    put(X::KINV, F.fh[0]);
    put(X::W, F.fh[1]);
    put(X::RW, s.r0.rw);   // the pre-transport rhogw of cell w
    f.GPREC[ow] = (0.0 + (-F.fq[0] / a.dt)) * 1.0;
    if (w == KMIN) {
        Surface<double> sf;
        surface_fluxes<NT>(ft, a.dt, sf);
        const int64_t pc = (int64_t)l * ijdim + ij;
        const int64_t pp = (int64_t)l * 2 * ijdim + ij;
        f.precip[pp] = sf.precip0;
        f.precip[pp + ijdim] = sf.precip1;
        f.ISO1_precip[pp] = 0.0;
        f.ISO1_precip[pp + ijdim] = 0.0;
        f.ISO2_precip[pp] = 0.0;
        f.ISO2_precip[pp + ijdim] = 0.0;
        f.precip_rhoe[pc] = sf.prhoe;
        f.precip_lh_heat[pc] = sf.lh;
        f.precip_rhophi[pc] = sf.prhophi;
        f.precip_rhokin[pc] = sf.prhokin;
    }
}

// Wall w of a general region: the same slots as flux_wall_to, from the fluxes over the
// region's kcell ranges (wall_flux_general).
template <int NT, typename Put, typename WS = double>
NSW6_HD void flux_wall_general_to(const StepArgs &a, int l, int w, int ij, Put &&put)
{
    using X = FxSlots<NT>;
    const int ijdim = a.ijdim;
    const Fields &f = a.f;
    const GridDev &g = *a.grid;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const int64_t p2 = (int64_t)l * K2 + ij, p3 = (int64_t)l * NQ * K2 + ij;
    const int64_t ow = p2 + (int64_t)w * ijdim;

    if (w == 0 || w == KALL - 1) {
#pragma unroll
        for (int i = 0; i < X::RW; i++) put(i, 0.0);
        put(X::RW, f.rhogw[ow]);
        f.GPREC[ow] = (0.0 + 0.0) * 1.0;
        return;
    }
    const int *bnd = a.ws.bounds + (size_t)l * ext_per_region<NT>();
    const int64_t pv = (int64_t)l * NT * K2 + ij;
    double V[NT][5];
    int kcmin[NT], kcmax[NT], khmin[NT], khmax[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        load_vt_window(ws_vt<WS>(a) + pv + t * K2, ijdim, w, V[t]);
        kcmin[t] = bnd[bounds_index(t, 0, w, 0)];
        kcmax[t] = bnd[bounds_index(t, 0, w, 1)];
        khmin[t] = bnd[bounds_index(t, 1, w, 0)];
        khmax[t] = bnd[bounds_index(t, 1, w, 1)];
    }
    const double rw = f.rhogw[ow];
    const Unit u = load_unit(f, (int64_t)l * ijdim + ij);
    GeneralGeomN<double, NT> G;
    general_geometry(g, w, a.dt, a.dt2, a.dt3_6, V, kcmin, kcmax, khmin, khmax, G);
    auto raw_at = [&](int j, CellIn<double> &r) { load_cell(a, p2, p3, K2, ijdim, j, r); };
    auto below_at = [&](int j) {
        const BelowCellN<NT> m = load_below<NT>(a, p2, p3, K2, ijdim, j);
        BelowOf<double, NT> b;
        b.rhog = m.rhog;
#pragma unroll
        for (int t = 0; t < NT; t++) b.q[t] = m.q[t];
        b.gs = m.gs;
        return b;
    };
    WallFluxN<double, NT> F;
    double ft[NT][4];
    wall_flux_general(g, w, G, u, raw_at, below_at, F, ft);

#pragma unroll
    for (int t = 0; t < NT; t++) put(X::Q + t, F.fq[t]);
#pragma unroll
    for (int i = 0; i < 5; i++) put(X::E + i, F.fs[i]);
    put(X::KINV, F.fh[0]);
    put(X::W, F.fh[1]);
    put(X::RW, rw);
    f.GPREC[ow] = (0.0 + (-F.fq[0] / a.dt)) * 1.0;
    if (w == KMIN) {
        Surface<double> sf;
        surface_fluxes<NT>(ft, a.dt, sf);
        const int64_t pc = (int64_t)l * ijdim + ij;
        const int64_t pp = (int64_t)l * 2 * ijdim + ij;
        f.precip[pp] = sf.precip0;
        f.precip[pp + ijdim] = sf.precip1;
        f.ISO1_precip[pp] = 0.0;
        f.ISO1_precip[pp + ijdim] = 0.0;
        f.ISO2_precip[pp] = 0.0;
        f.ISO2_precip[pp + ijdim] = 0.0;
        f.precip_rhoe[pc] = sf.prhoe;
        f.precip_lh_heat[pc] = sf.lh;
        f.precip_rhophi[pc] = sf.prhophi;
        f.precip_rhokin[pc] = sf.prhokin;
    }
}

// The same into the flux workspace [l][FxSlots<NT>::N][k][ij] of WS (fall speeds and
// fluxes as float in the mixed-precision form).
template <int NT = 3, typename WS = double>
NSW6_HD void flux_wall(const StepArgs &a, int l, int w, int ij)
{
    const int64_t K2 = (int64_t)KALL * a.ijdim;
    WS *fx = ws_flux<WS>(a) + (int64_t)l * FxSlots<NT>::N * K2 + (int64_t)w * a.ijdim + ij;
    auto put = [&](int slot, double v) { fx[slot * K2] = (WS)v; };
    flux_wall_to<NT, decltype(put) &, WS>(a, l, w, ij, put);
}

template <int NT = 3, typename WS = double>
NSW6_HD void flux_wall_general(const StepArgs &a, int l, int w, int ij)
{
    const int64_t K2 = (int64_t)KALL * a.ijdim;
    WS *fx = ws_flux<WS>(a) + (int64_t)l * FxSlots<NT>::N * K2 + (int64_t)w * a.ijdim + ij;
    auto put = [&](int slot, double v) { fx[slot * K2] = (WS)v; };
    flux_wall_general_to<NT, decltype(put) &, WS>(a, l, w, ij, put);
}

// Reads of the flux workspace: slot of wall w of column (l, ij).
template <int NT, typename WS = double> struct FluxFromMemory {
    const WS *fx;
    int64_t K2;
    int ijdim;
    NSW6_HD FluxFromMemory(const StepArgs &a, int l, int ij)
        : fx(ws_flux<WS>(a) + (int64_t)l * FxSlots<NT>::N * (int64_t)KALL * a.ijdim + ij),
          K2((int64_t)KALL * a.ijdim), ijdim(a.ijdim) {}
    NSW6_HD double operator()(int slot, int w) const { return (double)fx[(int64_t)slot * K2 + (int64_t)w * ijdim]; }
};

// KIN2KIN at wall kmax+1 of a single-cell region, for the thread of halo cell
// kmax+1: rhogkin_v is 0 at that wall, and the density of cell kmax after
// sedimentation (its lower neighbour) is formed from the fluxes of walls kmax and
// kmax+1 as sediment() forms it.  rhog_top is the halo cell's density.
template <int NT, typename Fetch>
NSW6_HD double kin2kin_top_rhogh(const StepArgs &a, int l, int ij, double rhog_top, const Fetch &FXV)
{
    using X = FxSlots<NT>;
    const int ijdim = a.ijdim;
    const GridDev &g = *a.grid;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const int64_t o = (int64_t)l * K2 + (int64_t)(KALL - 1) * ijdim + ij;
    const int64_t om = o - ijdim;
    const double rc = g.rdgz[KMAX];
    double rg = a.ws.rhog[om];
#pragma unroll
    for (int t = 0; t < NT; t++) rg = rg + -(FXV(X::Q + t, KALL - 1) - FXV(X::Q + t, KMAX)) * rc;
    const double gs = a.f.gsgam2[o], gsh = a.f.gsgam2h[o], gs_m = a.f.gsgam2[om];
    const double c2w0 = g.afact[KALL - 1] / gs * gsh, c2w1 = g.bfact[KALL - 1] / gs_m * gsh;
    return c2w0 * rhog_top + c2w1 * rg;
}
template <int NT, typename Fetch>
NSW6_HD double kin2kin_top_wall(const StepArgs &a, int l, int ij, double rhog_top, const Fetch &FXV)
{
    const int64_t o = (int64_t)l * KALL * a.ijdim + (int64_t)(KALL - 1) * a.ijdim + ij;
    return kin2kin_v(kin2kin_top_rhogh<NT>(a, l, ij, rhog_top, FXV), 0.0 + 0.0, a.f.rhogw[o] + 0.0);
}

// Inputs of sediment() for interior cell c of column (l, ij); FXV(slot, wall) reads the
// wall fluxes (from the flux workspace or from a block's LDS).
template <int NT, typename Fetch>
NSW6_HD void load_sed_from(const StepArgs &a, int l, int ij, int c, SedInN<double, NT> &x, const Fetch &FXV)
{
    using X = FxSlots<NT>;
    const int ijdim = a.ijdim;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const Fields &f = a.f;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        x.fq[0][t] = FXV(X::Q + t, c - 1);
        x.fq[1][t] = FXV(X::Q + t, c);
        x.fq[2][t] = FXV(X::Q + t, c + 1);
        x.fq[3][t] = (c + 1 <= KMAX) ? FXV(X::Q + t, c + 2) : 0.0;   // wall 96 does not exist
    }
#pragma unroll
    for (int i = 0; i < 5; i++) {
        x.fs[0][i] = FXV(X::E + i, c);
        x.fs[1][i] = FXV(X::E + i, c + 1);
    }
#pragma unroll
    for (int m = 0; m < 3; m++) {
        x.fh[m][0] = FXV(X::KINV, c - 1 + m);
        x.fh[m][1] = FXV(X::W, c - 1 + m);
    }
    x.rw_c = FXV(X::RW, c);
    x.rw_p = FXV(X::RW, c + 1);
    const int64_t o = (int64_t)l * K2 + (int64_t)c * ijdim + ij;
    const int64_t oq = (int64_t)l * NQ * K2 + (int64_t)c * ijdim + ij;
    x.rg_m = a.ws.rhog[o - ijdim];
    x.rg_c = a.ws.rhog[o];
    x.rg_p = a.ws.rhog[o + ijdim];
#pragma unroll
    for (int t = 0; t < NT; t++) x.rq[t] = a.ws.rhogq[oq + sed_nq<NT>(t) * K2];
    x.re0 = a.ws.rhoge[o];
    x.rvx = f.rhogvx[o];
    x.rvy = f.rhogvy[o];
    x.rvz = f.rhogvz[o];
    x.gs_m = f.gsgam2[o - ijdim];
    x.gs_c = f.gsgam2[o];
    x.gs_p = f.gsgam2[o + ijdim];
    x.gsh_c = f.gsgam2h[o];
    x.gsh_p = f.gsgam2h[o + ijdim];
    x.z = f.z[o];
}
template <int NT>
NSW6_HD void load_sed(const StepArgs &a, int l, int ij, int c, SedInN<double, NT> &x)
{
    load_sed_from(a, l, ij, c, x, FluxFromMemory<NT>(a, l, ij));
}

}  // namespace nsw6
