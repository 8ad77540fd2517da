// nsw6_grad.cpp -- TEST INFRASTRUCTURE: forward-mode dual-number gradient of one NSW6
// step on the host, the independent check of the hand-written adjoint.
//
// The step is the templated arithmetic of nicam_kokkos_microphysics_amd/csrc/
// nsw6_step.h (the same statements the HIP kernels run with S = double, pinned to the
// oracle by the parity tests), driven here by plain host loops with S = Dual<N>: each
// call propagates N tangent directions through the whole step (rates, departure
// distances and fluxes, divergence and KIN2EIN, saturation adjustment with the
// implicit-function tangent of the Newton solution, final filter).  Discrete decisions
// (kcell bounds, departure cells, |f| < EPS, limiter selections, switches) are taken on
// the values, as the derivative of the piecewise-smooth step does.
//
// Loss of the adjoint check (one step, reference semantics: tem is authoritative):
//     L = sum_{l,ij} sum_{t in QR,QS,QG} sflux_t^2 + sum_{l,k,ij} (rhoge_out - rhoge_in) / dt
// with sflux_t = -f_t(kmin) / dt the tracer's surface precipitation flux.
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "../include/nsw6.h"
#include "../nicam_kokkos_microphysics_amd/csrc/nsw6_setup.h"
#include "nsw6_dual.h"

using namespace nsw6;
using nsw6dual::Dual;

namespace {

const double A1_TAB[32] = {
    0.0001E-7, 0.7939E-7, 0.7841E-6, 0.3369E-5, 0.4336E-5, 0.5285E-5, 0.3728E-5, 0.1852E-5,
    0.2991E-6, 0.4248E-6, 0.7434E-6, 0.1812E-5, 0.4394E-5, 0.9145E-5, 0.1725E-4, 0.3348E-4,
    0.1725E-4, 0.9175E-5, 0.4412E-5, 0.2252E-5, 0.9115E-6, 0.4876E-6, 0.3473E-6, 0.4758E-6,
    0.6306E-6, 0.8573E-6, 0.7868E-6, 0.7192E-6, 0.6513E-6, 0.5956E-6, 0.5333E-6, 0.4834E-6};
const double A2_TAB[32] = {
    0.0100, 0.4006, 0.4831, 0.5320, 0.5307, 0.5319, 0.5249, 0.4888, 0.3849, 0.4047, 0.4318,
    0.4771, 0.5183, 0.5463, 0.5651, 0.5813, 0.5655, 0.5478, 0.5203, 0.4906, 0.4447, 0.4126,
    0.3960, 0.4149, 0.4320, 0.4506, 0.4483, 0.4460, 0.4433, 0.4413, 0.4382, 0.4361};

// Differentiable inputs and outputs of the step, [l][k][ij] / [l][6][k][ij] / [l][ij].
template <typename S> struct HostState {
    int lall, ijdim;
    std::vector<S> rhog, rhogvx, rhogvy, rhogvz, rhogw, rhoge, rhogq, vx, vy, vz, tem, q;
    const double *gsgam2, *gsgam2h, *gam2, *gam2h, *ix, *iy, *iz, *jx, *jy, *jz, *z;
    // outputs besides the in-place state
    std::vector<S> rho, pre, qd, precip, prhoe, lh, prhophi, prhokin, GPREC, sflux;
};

struct Bounds {
    std::vector<int> kcmin, kcmax, khmin, khmax;   // [l][NT][KALL]
    std::vector<int> general;                      // [l]
};

// kcell bounds of one region from the fall-speed values (k_extremes + k_bounds,
// src/mod_vadv1d.cpp:317-352)
template <int NT = 3, typename S>
void region_bounds(const GridDev &g, double dt, double dt2, double dt3, int l, int ijdim, const std::vector<S> &vt,
                   Bounds &b)
{
    const int64_t K2 = (int64_t)KALL * ijdim;
    for (int t = 0; t < NT; t++) {
        for (int type = 0; type < 2; type++) {
            const double *zh = (type == 0) ? g.gzh : g.gz_shift;
            const int mkmin = (type == 0) ? KMIN : KMIN + 1, mkmax = KMAX;
            for (int k = 0; k < KALL; k++) {
                int kcmin = k, kcmax = k;
                if (k >= mkmin && k <= mkmax) {
                    double amax = -INFINITY, amin = INFINITY;
                    for (int ij = 0; ij < ijdim; ij++) {
                        double V[5];
                        for (int d = -3; d <= 1; d++)
                            V[d + 3] = (k + d >= 0) ? val(vt[((int64_t)l * NT + t) * K2 + (int64_t)(k + d) * ijdim + ij]) : 0.0;
                        const double z = (type == 0) ? zdis_full_at(g, dt, dt2, dt3, V, k)
                                                     : zdis_half_at(g, dt, dt2, dt3, V, k);
                        if (z == z) {
                            amax = dmax(amax, z);
                            amin = dmin(amin, z);
                        }
                    }
                    const double zzmax = (amax > DBL_MIN) ? amax : DBL_MIN;
                    const double zzmin = (amin < DBL_MIN) ? amin : DBL_MIN;
                    if (zzmax > 0.0)
                        for (int k2 = k; k2 >= mkmin; k2--)
                            if ((zh[k2] <= zh[k] - zzmax) && (zh[k2 + 1] > zh[k] - zzmax)) {
                                kcmin = k2;
                                break;
                            }
                    if (zzmin < 0.0)
                        for (int k2 = k; k2 <= mkmax; k2++)
                            if ((zh[k2] <= zh[k] - zzmin) && (zh[k2 + 1] > zh[k] - zzmin)) {
                                kcmax = k2;
                                break;
                            }
                    const int ktop = (k + 2 <= mkmax + 1) ? k + 2 : mkmax + 1;
                    if (kcmin != k || kcmax > k + 1) b.general[l] = 1;
                    if (amax > 0.0 && zh[k] - zzmax < zh[k]) b.general[l] = 1;
                    if (zh[k] - zzmin >= zh[ktop]) b.general[l] = 1;
                }
                (type == 0 ? b.kcmin : b.khmin)[((size_t)l * NT + t) * KALL + k] = kcmin;
                (type == 0 ? b.kcmax : b.khmax)[((size_t)l * NT + t) * KALL + k] = kcmax;
            }
        }
    }
}

// One step over all regions (0; the general regions through wall_flux_general).
// The fall speed of sedimenting tracer t of NT from the rate kernel's outputs (QR, QS,
// QG in vt; QI in vti under "4WATER").
template <int NT, typename S> S fall_speed(const PointOut<S> &y, int t)
{
    if (NT == 3) return y.vt[t];
    return t == 0 ? y.vt[0] : (t == 1 ? y.vti : y.vt[t - 1]);
}

template <int NT, typename S>
int host_step_n(const DevParams &P, const GridDev &g, double dt, const S tun[NCOEF], HostState<S> &h, int *n_newton_fail)
{
    const int lall = h.lall, ijdim = h.ijdim;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const int64_t N2 = (int64_t)lall * K2;
    const double dt2 = pow(dt, 2), dt3_6 = pow(dt, 3) / 6.0;
    std::vector<S> wrhog(N2), wrhoge(N2), wrhogq(N2 * NQ), vt(N2 * NT);

    // ---- rates
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l < lall; l++)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++) {
                const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
                const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;
                const int64_t pv = (int64_t)l * NT * K2 + (int64_t)k * ijdim + ij;
                if (k == 0 || k == KALL - 1) {
                    wrhog[p] = h.rhog[p];
                    wrhoge[p] = h.rhoge[p];
                    for (int n = 0; n < NQ; n++) wrhogq[pq + n * K2] = h.rhogq[pq + n * K2];
                    for (int t = 0; t < NT; t++) vt[pv + t * K2] = S(0.0);
                    continue;
                }
                PointIn<S> x;
                x.rhog = h.rhog[p];
                for (int n = 0; n < NQ; n++) x.rhogq[n] = h.rhogq[pq + n * K2];
                x.tem = h.tem[p];
                x.qv = h.q[pq];
                x.gsgam2 = h.gsgam2[p];
                PointOut<S> y;
                point_rates<S, S>(P, dt, x, tun, A1_TAB, A2_TAB, y);
                wrhog[p] = y.rhog;
                wrhoge[p] = y.rhoge;
                for (int n = 0; n < NQ; n++) wrhogq[pq + n * K2] = y.rhogq[n];
                for (int t = 0; t < NT; t++) vt[pv + t * K2] = fall_speed<NT>(y, t);
            }

    // ---- kcell bounds
    Bounds b;
    b.kcmin.assign((size_t)lall * NT * KALL, 0);
    b.kcmax.assign((size_t)lall * NT * KALL, 0);
    b.khmin.assign((size_t)lall * NT * KALL, 0);
    b.khmax.assign((size_t)lall * NT * KALL, 0);
    b.general.assign(lall, 0);
    for (int l = 0; l < lall; l++) region_bounds<NT>(g, dt, dt2, dt3_6, l, ijdim, vt, b);

    // ---- fluxes of walls 1..94 (slots as the flux workspace: fq NT, fs 5, fh 2)
    constexpr int NS = NT + 7;
    std::vector<S> flux(N2 * NS, S(0.0));
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l < lall; l++)
        for (int w = KMIN; w <= KMAX; w++)
            for (int ij = 0; ij < ijdim; ij++) {
                const int64_t pc = (int64_t)l * ijdim + ij;
                const Unit u = {h.ix[pc], h.iy[pc], h.iz[pc], h.jx[pc], h.jy[pc], h.jz[pc]};
                auto raw = [&](int j, CellIn<S> &r) {
                    const int64_t o = (int64_t)l * K2 + (int64_t)j * ijdim + ij;
                    const int64_t oq = (int64_t)l * NQ * K2 + (int64_t)j * ijdim + ij;
                    r.rhog = wrhog[o];
                    for (int n = 1; n <= 5; n++) r.rq[n - 1] = wrhogq[oq + n * K2];
                    r.rhoge = wrhoge[o];
                    r.gs = h.gsgam2[o];
                    r.gam2 = h.gam2[o];
                    r.gsh = h.gsgam2h[o];
                    r.gam2h = h.gam2h[o];
                    r.rvx = h.rhogvx[o];
                    r.rvy = h.rhogvy[o];
                    r.rvz = h.rhogvz[o];
                    r.rw = h.rhogw[o];
                    r.vx = h.vx[o];
                    r.vy = h.vy[o];
                    r.vz = h.vz[o];
                    r.z = h.z[o];
                };
                const int64_t om = (int64_t)l * K2 + (int64_t)(w - 1) * ijdim + ij;
                const int64_t oqm = (int64_t)l * NQ * K2 + (int64_t)(w - 1) * ijdim + ij;
                const S rg_m = wrhog[om];
                S q_m[NT];
                for (int t = 0; t < NT; t++) q_m[t] = wrhogq[oqm + sed_nq<NT>(t) * K2] / rg_m;
                CellN<S, NT> cur, nxt;
                CellIn<S> r;
                raw(w, r);
                derive_cell(g, w, r, rg_m, q_m, h.gsgam2[om], u, cur);
                raw(w + 1, r);
                derive_cell(g, w + 1, r, cur.rhog, cur.q, cur.gs, u, nxt);
                S V[NT][5];
                int kcmin[NT], kcmax[NT], khmin[NT], khmax[NT];
                for (int t = 0; t < NT; t++) {
                    for (int d = -3; d <= 1; d++)
                        V[t][d + 3] = (w + d >= 0) ? vt[((int64_t)l * NT + t) * K2 + (int64_t)(w + d) * ijdim + ij] : S(0.0);
                    kcmin[t] = b.kcmin[((size_t)l * NT + t) * KALL + w];
                    kcmax[t] = b.kcmax[((size_t)l * NT + t) * KALL + w];
                    khmin[t] = b.khmin[((size_t)l * NT + t) * KALL + w];
                    khmax[t] = b.khmax[((size_t)l * NT + t) * KALL + w];
                }
                WallFluxN<S, NT> F;
                S ft[NT][4];
                if (b.general[l]) {
                    GeneralGeomN<S, NT> G;
                    general_geometry(g, w, dt, dt2, dt3_6, V, kcmin, kcmax, khmin, khmax, G);
                    auto below_at = [&](int j) {
                        const int64_t ob = (int64_t)l * K2 + (int64_t)(j - 1) * ijdim + ij;
                        const int64_t oqb = (int64_t)l * NQ * K2 + (int64_t)(j - 1) * ijdim + ij;
                        BelowOf<S, NT> bl;
                        bl.rhog = wrhog[ob];
                        for (int t = 0; t < NT; t++) bl.q[t] = wrhogq[oqb + sed_nq<NT>(t) * K2] / bl.rhog;
                        bl.gs = h.gsgam2[ob];
                        return bl;
                    };
                    wall_flux_general(g, w, G, u, raw, below_at, F, ft);
                } else {
                    S z0[NT], zh0[NT];
                    WallGeomN<NT> geo;
                    wall_geometry(g, w, dt, dt2, dt3_6, V, kcmax, khmax, z0, zh0, geo);
                    wall_flux(g, w, cur, nxt, z0, zh0, geo, F, ft);
                }
                S *fx = &flux[(int64_t)l * NS * K2 + (int64_t)w * ijdim + ij];
                for (int t = 0; t < NT; t++) fx[t * K2] = F.fq[t];
                for (int i = 0; i < 5; i++) fx[(NT + i) * K2] = F.fs[i];
                fx[(NT + 5) * K2] = F.fh[0];
                fx[(NT + 6) * K2] = F.fh[1];
                h.GPREC[om + ijdim] = (0.0 + (-F.fq[0] / dt)) * 1.0;
                if (w == KMIN) {
                    Surface<S> sf;
                    surface_fluxes<NT>(ft, dt, sf);
                    h.precip[(int64_t)l * 2 * ijdim + ij] = sf.precip0;
                    h.precip[(int64_t)l * 2 * ijdim + ijdim + ij] = sf.precip1;
                    h.prhoe[pc] = sf.prhoe;
                    h.lh[pc] = sf.lh;
                    h.prhophi[pc] = sf.prhophi;
                    h.prhokin[pc] = sf.prhokin;
                    for (int t = 0; t < NT; t++) h.sflux[((int64_t)l * NT + t) * ijdim + ij] = -F.fq[t] / dt;
                }
            }

    // ---- divergence, KIN2EIN and the point-wise tail
    std::vector<S> o_rvx(h.rhogvx), o_rvy(h.rhogvy), o_rvz(h.rhogvz), o_rw(h.rhogw);
    int fails = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+ : fails)
    for (int l = 0; l < lall; l++)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++) {
                const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
                const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;
                const int64_t pc = (int64_t)l * ijdim + ij;
                const bool interior = (k >= KMIN && k <= KMAX);
                TailIn<S> x;
                if (interior) {
                    SedInN<S, NT> si;
                    const S *fx = &flux[(int64_t)l * NS * K2 + ij];
                    auto FXV = [&](int slot, int wall) { return fx[(int64_t)slot * K2 + (int64_t)wall * ijdim]; };
                    for (int t = 0; t < NT; t++) {
                        si.fq[0][t] = FXV(t, k - 1);
                        si.fq[1][t] = FXV(t, k);
                        si.fq[2][t] = FXV(t, k + 1);
                        si.fq[3][t] = (k + 1 <= KMAX) ? FXV(t, k + 2) : S(0.0);
                    }
                    for (int i = 0; i < 5; i++) {
                        si.fs[0][i] = FXV(NT + i, k);
                        si.fs[1][i] = FXV(NT + i, k + 1);
                    }
                    for (int m = 0; m < 3; m++) {
                        si.fh[m][0] = FXV(NT + 5, k - 1 + m);
                        si.fh[m][1] = FXV(NT + 6, k - 1 + m);
                    }
                    si.rw_c = h.rhogw[p];
                    si.rw_p = h.rhogw[p + ijdim];
                    si.rg_m = wrhog[p - ijdim];
                    si.rg_c = wrhog[p];
                    si.rg_p = wrhog[p + ijdim];
                    for (int t = 0; t < NT; t++) si.rq[t] = wrhogq[pq + sed_nq<NT>(t) * K2];
                    si.re0 = wrhoge[p];
                    si.rvx = h.rhogvx[p];
                    si.rvy = h.rhogvy[p];
                    si.rvz = h.rhogvz[p];
                    si.gs_m = h.gsgam2[p - ijdim];
                    si.gs_c = h.gsgam2[p];
                    si.gs_p = h.gsgam2[p + ijdim];
                    si.gsh_c = h.gsgam2h[p];
                    si.gsh_p = h.gsgam2h[p + ijdim];
                    si.z = h.z[p];
                    SedOutN<S, NT> so;
                    const Unit u = {h.ix[pc], h.iy[pc], h.iz[pc], h.jx[pc], h.jy[pc], h.jz[pc]};
                    sediment(g, k, u, si, so, P.kin2kin != 0.0);
                    o_rvx[p] = so.rvx;
                    o_rvy[p] = so.rvy;
                    o_rvz[p] = so.rvz;
                    o_rw[p] = so.rw;
                    x.rhog = so.rg;
                    x.rhoge = so.e;
                    for (int t = 0; t < NT; t++) x.rhogq[sed_nq<NT>(t)] = so.rq[t];
                } else {
                    x.rhog = wrhog[p] + 0.0;
                    x.rhoge = wrhoge[p] + 0.0;
                    for (int t = 0; t < NT; t++) x.rhogq[sed_nq<NT>(t)] = wrhogq[pq + sed_nq<NT>(t) * K2] + 0.0;
                    o_rvx[p] = h.rhogvx[p] + 0.0 * h.ix[pc] + 0.0 * h.jx[pc];
                    o_rvy[p] = h.rhogvy[p] + 0.0 * h.iy[pc] + 0.0 * h.jy[pc];
                    o_rvz[p] = h.rhogvz[p] + 0.0 * h.iz[pc] + 0.0 * h.jz[pc];
                    o_rw[p] = h.rhogw[p] + 0.0;
                    if (P.kin2kin != 0.0 && k == KALL - 1) {
                        // KIN2KIN at wall kmax+1: rhogkin_v = 0, the lower cell after sedimentation
                        const S *fx = &flux[(int64_t)l * NS * K2 + ij];
                        S rg = wrhog[p - ijdim];
                        for (int t = 0; t < NT; t++)
                            rg = rg + -(fx[t * K2 + (int64_t)(KALL - 1) * ijdim] - fx[t * K2 + (int64_t)KMAX * ijdim]) *
                                          g.rdgz[KMAX];
                        const double c2w0 = g.afact[KALL - 1] / h.gsgam2[p] * h.gsgam2h[p];
                        const double c2w1 = g.bfact[KALL - 1] / h.gsgam2[p - ijdim] * h.gsgam2h[p];
                        o_rw[p] = kin2kin_v(S(c2w0 * (wrhog[p] + 0.0) + c2w1 * rg), S(0.0 + 0.0), S(h.rhogw[p] + 0.0));
                    }
                    h.GPREC[p] = S(0.0);
                }
                x.rhogq[QV] = wrhogq[pq + QV * K2] + 0.0;
                x.rhogq[QC] = wrhogq[pq + QC * K2] + 0.0;
                if (NT == 3) x.rhogq[QI] = wrhogq[pq + QI * K2] + 0.0;
                x.gsgam2 = h.gsgam2[p];
                TailOut<S> y;
                const TailStatus st = tail(P, interior, x, y);
                if (st.newton < 0) fails++;
                h.rhog[p] = y.rhog;
                h.rhoge[p] = y.rhoge;
                h.rho[p] = y.rho;
                h.tem[p] = y.tem;
                h.pre[p] = y.pre;
                h.qd[p] = y.qd;
                for (int n = 0; n < NQ; n++) {
                    h.rhogq[pq + n * K2] = y.rhogq[n];
                    h.q[pq + n * K2] = y.q[n];
                }
            }
    h.rhogvx.swap(o_rvx);
    h.rhogvy.swap(o_rvy);
    h.rhogvz.swap(o_rvz);
    h.rhogw.swap(o_rw);
    if (n_newton_fail) *n_newton_fail = fails;
    return 0;
}

template <typename S>
int host_step(const DevParams &P, const GridDev &g, double dt, const S tun[NCOEF], HostState<S> &h, int *n_newton_fail)
{
    return P.four_water != 0.0 ? host_step_n<4>(P, g, dt, tun, h, n_newton_fail)
                               : host_step_n<3>(P, g, dt, tun, h, n_newton_fail);
}

template <typename S> void alloc_outputs(HostState<S> &h, int nt = 3)
{
    const int64_t N2 = (int64_t)h.lall * KALL * h.ijdim, NC = (int64_t)h.lall * h.ijdim;
    h.rho.assign(N2, S(0.0));
    h.pre.assign(N2, S(0.0));
    h.qd.assign(N2, S(0.0));
    h.GPREC.assign(N2, S(0.0));
    h.precip.assign(2 * NC, S(0.0));
    h.prhoe.assign(NC, S(0.0));
    h.lh.assign(NC, S(0.0));
    h.prhophi.assign(NC, S(0.0));
    h.prhokin.assign(NC, S(0.0));
    h.sflux.assign((size_t)nt * NC, S(0.0));
}

template <typename S> void bind_metrics(HostState<S> &h, const nsw6_fields *f)
{
    h.gsgam2 = f->gsgam2;
    h.gsgam2h = f->gsgam2h;
    h.gam2 = f->gam2;
    h.gam2h = f->gam2h;
    h.ix = f->ix;
    h.iy = f->iy;
    h.iz = f->iz;
    h.jx = f->jx;
    h.jy = f->jy;
    h.jz = f->jz;
    h.z = f->z;
}

}  // namespace

extern "C" {

// The templated step in FP64 on the host, in place on the arrays of `f` (the same
// arrays and layout as nsw6_mp_driver_fields; rho, pre, qd, precip, precip_*, GPREC
// written).  Returns 0 (1 on a Newton failure).
int grad_step_fp64(const double *gz, const double *gzh, const nsw6_params *params, int lall, int ijdim,
                   nsw6_fields *f, double dt)
{
    DevParams P;
    to_dev_params(*params, P);
    GridDev g;
    make_grid_dev(gz, gzh, g);
    const int64_t N2 = (int64_t)lall * KALL * ijdim, NC = (int64_t)lall * ijdim;
    HostState<double> h;
    h.lall = lall;
    h.ijdim = ijdim;
    auto cp = [](std::vector<double> &v, const double *src, int64_t n) { v.assign(src, src + n); };
    cp(h.rhog, f->rhog, N2);
    cp(h.rhogvx, f->rhogvx, N2);
    cp(h.rhogvy, f->rhogvy, N2);
    cp(h.rhogvz, f->rhogvz, N2);
    cp(h.rhogw, f->rhogw, N2);
    cp(h.rhoge, f->rhoge, N2);
    cp(h.rhogq, f->rhogq, N2 * NQ);
    cp(h.vx, f->vx, N2);
    cp(h.vy, f->vy, N2);
    cp(h.vz, f->vz, N2);
    cp(h.tem, f->tem, N2);
    cp(h.q, f->q, N2 * NQ);
    bind_metrics(h, f);
    alloc_outputs(h, P.four_water != 0.0 ? 4 : 3);
    double tun[NCOEF];
    tunables(P, tun);
    int fails = 0;
    const int st = host_step(P, g, dt, tun, h, &fails);
    if (st != 0) return st;
    auto out = [](double *dst, const std::vector<double> &v) { memcpy(dst, v.data(), v.size() * sizeof(double)); };
    out(f->rhog, h.rhog);
    out(f->rhogvx, h.rhogvx);
    out(f->rhogvy, h.rhogvy);
    out(f->rhogvz, h.rhogvz);
    out(f->rhogw, h.rhogw);
    out(f->rhoge, h.rhoge);
    out(f->rhogq, h.rhogq);
    out(f->tem, h.tem);
    out(f->q, h.q);
    out(f->rho, h.rho);
    out(f->pre, h.pre);
    out(f->qd, h.qd);
    out(f->GPREC, h.GPREC);
    out(f->precip, h.precip);
    out(f->precip_rhoe, h.prhoe);
    out(f->precip_lh_heat, h.lh);
    out(f->precip_rhophi, h.prhophi);
    out(f->precip_rhokin, h.prhokin);
    (void)NC;
    return fails ? 1 : 0;
}

// Surface mass fluxes sflux_t = -f_t(kmin)/dt of the falling tracers ([l][NT][ij]: QR,
// QS, QG, or QR, QI, QS, QG under four_water) of one FP64 host step from the state `f`
// (not modified).
int grad_sflux_fp64(const double *gz, const double *gzh, const nsw6_params *params, int lall, int ijdim,
                    const nsw6_fields *f, double dt, double *sflux)
{
    DevParams P;
    to_dev_params(*params, P);
    GridDev g;
    make_grid_dev(gz, gzh, g);
    const int64_t N2 = (int64_t)lall * KALL * ijdim;
    HostState<double> h;
    h.lall = lall;
    h.ijdim = ijdim;
    auto cp = [](std::vector<double> &v, const double *src, int64_t n) { v.assign(src, src + n); };
    cp(h.rhog, f->rhog, N2);
    cp(h.rhogvx, f->rhogvx, N2);
    cp(h.rhogvy, f->rhogvy, N2);
    cp(h.rhogvz, f->rhogvz, N2);
    cp(h.rhogw, f->rhogw, N2);
    cp(h.rhoge, f->rhoge, N2);
    cp(h.rhogq, f->rhogq, N2 * NQ);
    cp(h.vx, f->vx, N2);
    cp(h.vy, f->vy, N2);
    cp(h.vz, f->vz, N2);
    cp(h.tem, f->tem, N2);
    cp(h.q, f->q, N2 * NQ);
    bind_metrics(h, f);
    alloc_outputs(h, P.four_water != 0.0 ? 4 : 3);
    double tun[NCOEF];
    tunables(P, tun);
    int fails = 0;
    const int st = host_step(P, g, dt, tun, h, &fails);
    if (st != 0) return st;
    memcpy(sflux, h.sflux.data(), h.sflux.size() * sizeof(double));
    return fails ? 1 : 0;
}

// The check loss L (header comment) of the FP64 host step, for finite differences.
int grad_loss_fp64(const double *gz, const double *gzh, const nsw6_params *params, const double *coef, int lall,
                   int ijdim, const nsw6_fields *f, double dt, double *loss)
{
    DevParams P;
    to_dev_params(*params, P);
    GridDev g;
    make_grid_dev(gz, gzh, g);
    const int64_t N2 = (int64_t)lall * KALL * ijdim;
    HostState<double> h;
    h.lall = lall;
    h.ijdim = ijdim;
    auto cp = [](std::vector<double> &v, const double *src, int64_t n) { v.assign(src, src + n); };
    cp(h.rhog, f->rhog, N2);
    cp(h.rhogvx, f->rhogvx, N2);
    cp(h.rhogvy, f->rhogvy, N2);
    cp(h.rhogvz, f->rhogvz, N2);
    cp(h.rhogw, f->rhogw, N2);
    cp(h.rhoge, f->rhoge, N2);
    cp(h.rhogq, f->rhogq, N2 * NQ);
    cp(h.vx, f->vx, N2);
    cp(h.vy, f->vy, N2);
    cp(h.vz, f->vz, N2);
    cp(h.tem, f->tem, N2);
    cp(h.q, f->q, N2 * NQ);
    const std::vector<double> rhoge_in(h.rhoge);
    bind_metrics(h, f);
    alloc_outputs(h, P.four_water != 0.0 ? 4 : 3);
    double tun[NCOEF];
    tunables(P, tun);
    if (coef)
        for (int c = 0; c < NCOEF; c++) tun[c] = coef[c];
    int fails = 0;
    const int st = host_step(P, g, dt, tun, h, &fails);
    if (st != 0) return st;
    double L = 0.0;
    for (size_t e = 0; e < h.sflux.size(); e++) L += h.sflux[e] * h.sflux[e];
    double E = 0.0;
    for (int64_t e = 0; e < N2; e++) E += (h.rhoge[e] - rhoge_in[e]) / dt;
    *loss = L + E;
    return fails ? 1 : 0;
}

#define GRAD_NDIR 8

// Directional derivatives of the check loss L (header comment) along `ndir` <= 8
// directions.  dirs[i] holds direction arrays for rhog, rhogvx, rhogvy, rhogvz, rhogw,
// rhoge, rhogq, tem, q (only QV enters), vx, vy, vz (NULL pointers = 0); dcoef is
// [ndir][8] in the coefficient order of include/nsw6.h.  Writes L and dL[i].
int grad_dual_loss(const double *gz, const double *gzh, const nsw6_params *params, int lall, int ijdim,
                   const nsw6_fields *f, double dt, int ndir, const nsw6_fields *dirs, const double *dcoef,
                   double *loss, double *dloss)
{
    if (ndir < 1 || ndir > GRAD_NDIR) return -2;
    typedef Dual<GRAD_NDIR> D;
    DevParams P;
    to_dev_params(*params, P);
    GridDev g;
    make_grid_dev(gz, gzh, g);
    const int64_t N2 = (int64_t)lall * KALL * ijdim;
    HostState<D> h;
    h.lall = lall;
    h.ijdim = ijdim;
    auto seed = [&](std::vector<D> &v, const double *src, size_t off, int64_t n) {
        v.resize(n);
        for (int64_t e = 0; e < n; e++) {
            v[e] = D(src[e]);
            for (int i = 0; i < ndir; i++) {
                const double *dp = *(const double *const *)((const char *)&dirs[i] + off);
                v[e].d[i] = dp ? dp[e] : 0.0;
            }
        }
    };
#define SEED(field, n) seed(h.field, f->field, offsetof(nsw6_fields, field), n)
    SEED(rhog, N2);
    SEED(rhogvx, N2);
    SEED(rhogvy, N2);
    SEED(rhogvz, N2);
    SEED(rhogw, N2);
    SEED(rhoge, N2);
    SEED(rhogq, N2 * NQ);
    SEED(vx, N2);
    SEED(vy, N2);
    SEED(vz, N2);
    SEED(tem, N2);
    SEED(q, N2 * NQ);
#undef SEED
    const std::vector<D> rhoge_in(h.rhoge);
    bind_metrics(h, f);
    alloc_outputs(h, P.four_water != 0.0 ? 4 : 3);
    double t0[NCOEF];
    tunables(P, t0);
    D tun[NCOEF];
    for (int c = 0; c < NCOEF; c++) {
        tun[c] = D(t0[c]);
        for (int i = 0; i < ndir; i++) tun[c].d[i] = dcoef ? dcoef[i * NCOEF + c] : 0.0;
    }
    int fails = 0;
    const int st = host_step(P, g, dt, tun, h, &fails);
    if (st != 0) return st;
    D L(0.0);
    for (size_t e = 0; e < h.sflux.size(); e++) L += h.sflux[e] * h.sflux[e];
    D E(0.0);
    for (int64_t e = 0; e < N2; e++) E += (h.rhoge[e] - rhoge_in[e]) / dt;
    L += E;
    *loss = L.v;
    for (int i = 0; i < ndir; i++) dloss[i] = L.d[i];
    return fails ? 1 : 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------------
// Piece-by-piece check of the hand-written reverse functions (nsw6_adjoint.h) against
// forward-mode dual numbers: for each point, wall and cell of the given state, random
// input direction v and output cotangent yb, compare <reverse(yb), v> with
// <yb, J v>.  err[i] = max relative mismatch of piece i: 0 point_rates, 1 wall
// (derive_cell x2 + departure distances + fluxes + surface), 2 sediment, 3 tail.
// ---------------------------------------------------------------------------------
#include "../nicam_kokkos_microphysics_amd/csrc/nsw6_adjoint.h"

namespace {
typedef Dual<1> D1;
struct Rng {
    uint64_t s;
    double u()   // uniform in (-1, 1)
    {
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return ((double)(s >> 11) / 9007199254740992.0) * 2.0 - 1.0;
    }
};
inline double rel(double a, double b) { return fabs(a - b) / (fabs(a) + fabs(b) + 1e-300); }
inline D1 seed1(double x, double v)
{
    D1 d(x);
    d.d[0] = v;
    return d;
}
// output cotangent of relative weight ~1/|y|, floored at a physical scale `fl` of that
// output so that vanishing outputs do not turn the dot products into cancellations
inline double wt(Rng &r, double y, double fl) { return r.u() / fmax(fabs(y), fl); }
}  // namespace

extern "C" int grad_check_pieces(const double *gz, const double *gzh, const nsw6_params *params, int lall,
                                 int ijdim, const nsw6_fields *f, double dt, double err[4], int64_t count[4])
{
    DevParams P;
    to_dev_params(*params, P);
    GridDev g;
    make_grid_dev(gz, gzh, g);
    const int64_t K2 = (int64_t)KALL * ijdim, N2 = (int64_t)lall * K2;
    const double dt2 = pow(dt, 2), dt3_6 = pow(dt, 3) / 6.0;
    double tun[NCOEF];
    tunables(P, tun);
    for (int i = 0; i < 4; i++) {
        err[i] = 0.0;
        count[i] = 0;
    }
    Rng rng{12345};
    std::vector<double> wrhog(N2), wrhoge(N2), wrhogq(N2 * NQ), vt(N2 * 3);
    // ---- point_rates
    for (int l = 0; l < lall; l++)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++) {
                const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
                const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;
                const int64_t pv = (int64_t)l * 3 * K2 + (int64_t)k * ijdim + ij;
                if (k == 0 || k == KALL - 1) {
                    wrhog[p] = f->rhog[p];
                    wrhoge[p] = f->rhoge[p];
                    for (int n = 0; n < NQ; n++) wrhogq[pq + n * K2] = f->rhogq[pq + n * K2];
                    for (int t = 0; t < 3; t++) vt[pv + t * K2] = 0.0;
                    continue;
                }
                PointIn<double> x;
                x.rhog = f->rhog[p];
                for (int n = 0; n < NQ; n++) x.rhogq[n] = f->rhogq[pq + n * K2];
                x.tem = f->tem[p];
                x.qv = f->q[pq];
                x.gsgam2 = f->gsgam2[p];
                PointOut<double> y;
                point_rates<double, double>(P, dt, x, tun, A1_TAB, A2_TAB, y);
                wrhog[p] = y.rhog;
                wrhoge[p] = y.rhoge;
                for (int n = 0; n < NQ; n++) wrhogq[pq + n * K2] = y.rhogq[n];
                for (int t = 0; t < 3; t++) vt[pv + t * K2] = y.vt[t];
                // dual: direction v on (rhog, rhogq[6], tem, qv, tun[8])
                double v[17];
                v[0] = rng.u() * x.rhog;
                for (int n = 0; n < 6; n++) v[1 + n] = rng.u() * x.rhogq[n];
                v[7] = rng.u() * x.tem;
                v[8] = rng.u() * x.qv;
                for (int c = 0; c < NCOEF; c++) v[9 + c] = rng.u() * tun[c];
                PointIn<D1> xd;
                xd.rhog = seed1(x.rhog, v[0]);
                for (int n = 0; n < 6; n++) xd.rhogq[n] = seed1(x.rhogq[n], v[1 + n]);
                xd.tem = seed1(x.tem, v[7]);
                xd.qv = seed1(x.qv, v[8]);
                xd.gsgam2 = x.gsgam2;
                D1 td[NCOEF];
                for (int c = 0; c < NCOEF; c++) td[c] = seed1(tun[c], v[9 + c]);
                PointOut<D1> yd;
                point_rates<D1, D1>(P, dt, xd, td, A1_TAB, A2_TAB, yd);
                PointOutB yb;
                yb.rhog = wt(rng, y.rhog, 1e-3);
                yb.rhoge = wt(rng, y.rhoge, 1.0);
                for (int n = 0; n < 6; n++) yb.rhogq[n] = wt(rng, y.rhogq[n], 1e-9);
                for (int t = 0; t < 3; t++) yb.vt[t] = wt(rng, y.vt[t], 1e-3);
                double jv = yb.rhog * yd.rhog.d[0] + yb.rhoge * yd.rhoge.d[0];
                for (int n = 0; n < 6; n++) jv += yb.rhogq[n] * yd.rhogq[n].d[0];
                for (int t = 0; t < 3; t++) jv += yb.vt[t] * yd.vt[t].d[0];
                PointInB xb;
                memset(&xb, 0, sizeof xb);
                double tb[NCOEF] = {0, 0, 0, 0, 0, 0, 0, 0};
                point_rates_b(P, dt, x, tun, yb, xb, tb, A1_TAB, A2_TAB);
                double vx = xb.rhog * v[0] + xb.tem * v[7] + xb.qv * v[8];
                for (int n = 0; n < 6; n++) vx += xb.rhogq[n] * v[1 + n];
                for (int c = 0; c < NCOEF; c++) vx += tb[c] * v[9 + c];
                err[0] = fmax(err[0], rel(vx, jv));
                count[0]++;
            }
    // ---- kcell bounds (values)
    Bounds bd;
    bd.kcmin.assign((size_t)lall * 3 * KALL, 0);
    bd.kcmax.assign((size_t)lall * 3 * KALL, 0);
    bd.khmin.assign((size_t)lall * 3 * KALL, 0);
    bd.khmax.assign((size_t)lall * 3 * KALL, 0);
    bd.general.assign(lall, 0);
    for (int l = 0; l < lall; l++) region_bounds(g, dt, dt2, dt3_6, l, ijdim, vt, bd);
    for (int l = 0; l < lall; l++)
        if (bd.general[l]) return -1;
    // ---- walls
    std::vector<double> flux(N2 * 10, 0.0);
    for (int l = 0; l < lall; l++)
        for (int w = KMIN; w <= KMAX; w++)
            for (int ij = 0; ij < ijdim; ij++) {
                const int64_t pc = (int64_t)l * ijdim + ij;
                const Unit u = {f->ix[pc], f->iy[pc], f->iz[pc], f->jx[pc], f->jy[pc], f->jz[pc]};
                auto raw = [&](int j, CellIn<double> &r) {
                    const int64_t o = (int64_t)l * K2 + (int64_t)j * ijdim + ij;
                    const int64_t oq = (int64_t)l * NQ * K2 + (int64_t)j * ijdim + ij;
                    r.rhog = wrhog[o];
                    for (int n = 1; n <= 5; n++) r.rq[n - 1] = wrhogq[oq + n * K2];
                    r.rhoge = wrhoge[o];
                    r.gs = f->gsgam2[o];
                    r.gam2 = f->gam2[o];
                    r.gsh = f->gsgam2h[o];
                    r.gam2h = f->gam2h[o];
                    r.rvx = f->rhogvx[o];
                    r.rvy = f->rhogvy[o];
                    r.rvz = f->rhogvz[o];
                    r.rw = f->rhogw[o];
                    r.vx = f->vx[o];
                    r.vy = f->vy[o];
                    r.vz = f->vz[o];
                    r.z = f->z[o];
                };
                const int64_t om = (int64_t)l * K2 + (int64_t)(w - 1) * ijdim + ij;
                const int64_t oqm = (int64_t)l * NQ * K2 + (int64_t)(w - 1) * ijdim + ij;
                CellIn<double> rw0, rw1;
                raw(w, rw0);
                raw(w + 1, rw1);
                const double rgm = wrhog[om], gsm = f->gsgam2[om];
                const double rqm[3] = {wrhogq[oqm + QR * K2], wrhogq[oqm + QS * K2], wrhogq[oqm + QG * K2]};
                double V[3][5];
                int kcmax[3], khmax[3];
                for (int t = 0; t < 3; t++) {
                    for (int d = -3; d <= 1; d++)
                        V[t][d + 3] = (w + d >= 0) ? vt[((int64_t)l * 3 + t) * K2 + (int64_t)(w + d) * ijdim + ij] : 0.0;
                    kcmax[t] = bd.kcmax[((size_t)l * 3 + t) * KALL + w];
                    khmax[t] = bd.khmax[((size_t)l * 3 + t) * KALL + w];
                }
                // forward (double)
                const double qm[3] = {rqm[0] / rgm, rqm[1] / rgm, rqm[2] / rgm};
                Cell<double> cur, nxt;
                derive_cell(g, w, rw0, rgm, qm, gsm, u, cur);
                derive_cell(g, w + 1, rw1, cur.rhog, cur.q, cur.gs, u, nxt);
                double z0[3], zh0[3];
                WallGeom geo;
                wall_geometry(g, w, dt, dt2, dt3_6, V, kcmax, khmax, z0, zh0, geo);
                WallFlux<double> F;
                double ft[3][4];
                wall_flux(g, w, cur, nxt, z0, zh0, geo, F, ft);
                double *fx = &flux[(int64_t)l * 10 * K2 + (int64_t)w * ijdim + ij];
                for (int t = 0; t < 3; t++) fx[t * K2] = F.fq[t];
                for (int i = 0; i < 5; i++) fx[(3 + i) * K2] = F.fs[i];
                fx[8 * K2] = F.fh[0];
                fx[9 * K2] = F.fh[1];
                // dual with a random direction on: cell w-1 (rhog, rq3), cells w, w+1 (CellIn), V
                auto dcell = [&](const CellIn<double> &r, CellIn<D1> &rd, double *vv) {
                    rd.rhog = seed1(r.rhog, vv[0] = rng.u() * r.rhog);
                    for (int n = 0; n < 5; n++) rd.rq[n] = seed1(r.rq[n], vv[1 + n] = rng.u() * r.rq[n]);
                    rd.rhoge = seed1(r.rhoge, vv[6] = rng.u() * r.rhoge);
                    rd.rvx = seed1(r.rvx, vv[7] = rng.u() * r.rvx);
                    rd.rvy = seed1(r.rvy, vv[8] = rng.u() * r.rvy);
                    rd.rvz = seed1(r.rvz, vv[9] = rng.u() * r.rvz);
                    rd.rw = seed1(r.rw, vv[10] = rng.u() * r.rw);
                    rd.vx = seed1(r.vx, vv[11] = rng.u() * r.vx);
                    rd.vy = seed1(r.vy, vv[12] = rng.u() * r.vy);
                    rd.vz = seed1(r.vz, vv[13] = rng.u() * r.vz);
                    rd.gs = r.gs;
                    rd.gam2 = r.gam2;
                    rd.gsh = r.gsh;
                    rd.gam2h = r.gam2h;
                    rd.z = r.z;
                };
                double v0[14], v1[14], vm[4], vV[3][5];
                CellIn<D1> rd0, rd1;
                dcell(rw0, rd0, v0);
                dcell(rw1, rd1, v1);
                const D1 rgmd = seed1(rgm, vm[0] = rng.u() * rgm);
                D1 qmd[3];
                for (int t = 0; t < 3; t++) {
                    const D1 rq_ = seed1(rqm[t], vm[1 + t] = rng.u() * rqm[t]);
                    qmd[t] = rq_ / rgmd;
                }
                D1 Vd[3][5];
                for (int t = 0; t < 3; t++)
                    for (int d = 0; d < 5; d++) Vd[t][d] = seed1(V[t][d], vV[t][d] = rng.u() * V[t][d]);
                Cell<D1> curd, nxtd;
                derive_cell(g, w, rd0, rgmd, qmd, gsm, u, curd);
                derive_cell(g, w + 1, rd1, curd.rhog, curd.q, curd.gs, u, nxtd);
                D1 z0d[3], zh0d[3];
                WallGeom geod;
                wall_geometry(g, w, dt, dt2, dt3_6, Vd, kcmax, khmax, z0d, zh0d, geod);
                WallFlux<D1> Fd;
                D1 ftd[3][4];
                wall_flux(g, w, curd, nxtd, z0d, zh0d, geod, Fd, ftd);
                WallFlux<double> Fb;
                double ftb[3][4];
                double jv = 0.0;
                for (int t = 0; t < 3; t++) {
                    Fb.fq[t] = wt(rng, F.fq[t], 1e-12);
                    jv += Fb.fq[t] * Fd.fq[t].d[0];
                    for (int i = 0; i < 4; i++) {
                        ftb[t][i] = wt(rng, ft[t][i], 1e-9);
                        jv += ftb[t][i] * ftd[t][i].d[0];
                    }
                }
                for (int i = 0; i < 5; i++) {
                    Fb.fs[i] = wt(rng, F.fs[i], 1e-9);
                    jv += Fb.fs[i] * Fd.fs[i].d[0];
                }
                for (int i = 0; i < 2; i++) {
                    Fb.fh[i] = wt(rng, F.fh[i], 1e-12);
                    jv += Fb.fh[i] * Fd.fh[i].d[0];
                }
                // reverse
                CellB cb0, cb1;
                zero(cb0);
                zero(cb1);
                double z0b[3] = {0, 0, 0}, zh0b[3] = {0, 0, 0};
                wall_flux_b(g, w, cur, nxt, z0, zh0, geo, Fb, ftb, cb0, cb1, z0b, zh0b);
                double Vb[3][5];
                memset(Vb, 0, sizeof Vb);
                for (int t = 0; t < 3; t++) {
                    zdis_full_b(g, dt, dt2, dt3_6, V[t], w, z0b[t], Vb[t]);
                    if (w >= KMIN + 1) zdis_half_b(g, dt, dt2, dt3_6, V[t], w, zh0b[t], Vb[t]);
                }
                CellInB rb0, rb1;
                zero(rb0);
                zero(rb1);
                double brg_b = 0.0, bq_b[3] = {0, 0, 0};
                derive_cell_b(g, w + 1, rw1, cur.rhog, cur.q, cur.gs, u, nxt, cb1, rb1, cb0.rhog, cb0.q);
                derive_cell_b(g, w, rw0, rgm, qm, gsm, u, cur, cb0, rb0, brg_b, bq_b);
                double rgm_b = brg_b, rqm_b[3];
                for (int t = 0; t < 3; t++) {
                    rqm_b[t] = bq_b[t] / rgm;
                    rgm_b -= bq_b[t] * qm[t] / rgm;
                }
                auto dotc = [](const CellInB &b, const double *vv) {
                    double s = b.rhog * vv[0];
                    for (int n = 0; n < 5; n++) s += b.rq[n] * vv[1 + n];
                    return s + b.rhoge * vv[6] + b.rvx * vv[7] + b.rvy * vv[8] + b.rvz * vv[9] + b.rw * vv[10] +
                           b.vx * vv[11] + b.vy * vv[12] + b.vz * vv[13];
                };
                double vx = dotc(rb0, v0) + dotc(rb1, v1) + rgm_b * vm[0];
                for (int t = 0; t < 3; t++) vx += rqm_b[t] * vm[1 + t];
                for (int t = 0; t < 3; t++)
                    for (int d = 0; d < 5; d++) vx += Vb[t][d] * vV[t][d];
                err[1] = fmax(err[1], rel(vx, jv));
                count[1]++;
            }
    // ---- cells: sediment and tail
    for (int l = 0; l < lall; l++)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++) {
                const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
                const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;
                const int64_t pc = (int64_t)l * ijdim + ij;
                const bool interior = (k >= KMIN && k <= KMAX);
                TailIn<double> x;
                if (interior) {
                    SedIn<double> si;
                    const double *fx = &flux[(int64_t)l * 10 * K2 + ij];
                    auto FXV = [&](int slot, int wall) { return fx[(int64_t)slot * K2 + (int64_t)wall * ijdim]; };
                    for (int t = 0; t < 3; t++) {
                        si.fq[0][t] = FXV(t, k - 1);
                        si.fq[1][t] = FXV(t, k);
                        si.fq[2][t] = FXV(t, k + 1);
                        si.fq[3][t] = (k + 1 <= KMAX) ? FXV(t, k + 2) : 0.0;
                    }
                    for (int i = 0; i < 5; i++) {
                        si.fs[0][i] = FXV(3 + i, k);
                        si.fs[1][i] = FXV(3 + i, k + 1);
                    }
                    for (int m = 0; m < 3; m++) {
                        si.fh[m][0] = FXV(8, k - 1 + m);
                        si.fh[m][1] = FXV(9, k - 1 + m);
                    }
                    si.rw_c = f->rhogw[p];
                    si.rw_p = f->rhogw[p + ijdim];
                    si.rg_m = wrhog[p - ijdim];
                    si.rg_c = wrhog[p];
                    si.rg_p = wrhog[p + ijdim];
                    si.rq[0] = wrhogq[pq + QR * K2];
                    si.rq[1] = wrhogq[pq + QS * K2];
                    si.rq[2] = wrhogq[pq + QG * K2];
                    si.re0 = wrhoge[p];
                    si.rvx = f->rhogvx[p];
                    si.rvy = f->rhogvy[p];
                    si.rvz = f->rhogvz[p];
                    si.gs_m = f->gsgam2[p - ijdim];
                    si.gs_c = f->gsgam2[p];
                    si.gs_p = f->gsgam2[p + ijdim];
                    si.gsh_c = f->gsgam2h[p];
                    si.gsh_p = f->gsgam2h[p + ijdim];
                    si.z = f->z[p];
                    const Unit u = {f->ix[pc], f->iy[pc], f->iz[pc], f->jx[pc], f->jy[pc], f->jz[pc]};
                    SedOut<double> so;
                    sediment(g, k, u, si, so);
                    // dual check of sediment
                    SedIn<D1> sd;
                    SedInB vv;
                    auto sv = [&](double xv, double &vs) { vs = rng.u() * (fabs(xv) > 0 ? xv : 1e-3); return seed1(xv, vs); };
                    for (int a = 0; a < 4; a++)
                        for (int t = 0; t < 3; t++) sd.fq[a][t] = sv(si.fq[a][t], vv.fq[a][t]);
                    for (int a = 0; a < 2; a++)
                        for (int i = 0; i < 5; i++) sd.fs[a][i] = sv(si.fs[a][i], vv.fs[a][i]);
                    for (int a = 0; a < 3; a++)
                        for (int i = 0; i < 2; i++) sd.fh[a][i] = sv(si.fh[a][i], vv.fh[a][i]);
                    sd.rw_c = sv(si.rw_c, vv.rw_c);
                    sd.rw_p = sv(si.rw_p, vv.rw_p);
                    sd.rg_m = sv(si.rg_m, vv.rg_m);
                    sd.rg_c = sv(si.rg_c, vv.rg_c);
                    sd.rg_p = sv(si.rg_p, vv.rg_p);
                    for (int t = 0; t < 3; t++) sd.rq[t] = sv(si.rq[t], vv.rq[t]);
                    sd.re0 = sv(si.re0, vv.re0);
                    sd.rvx = sv(si.rvx, vv.rvx);
                    sd.rvy = sv(si.rvy, vv.rvy);
                    sd.rvz = sv(si.rvz, vv.rvz);
                    sd.gs_m = si.gs_m;
                    sd.gs_c = si.gs_c;
                    sd.gs_p = si.gs_p;
                    sd.gsh_c = si.gsh_c;
                    sd.gsh_p = si.gsh_p;
                    sd.z = si.z;
                    SedOut<D1> sod;
                    sediment(g, k, u, sd, sod);
                    SedOut<double> sb;
                    sb.rg = wt(rng, so.rg, 1e-3);
                    sb.e = wt(rng, so.e, 1.0);
                    for (int t = 0; t < 3; t++) sb.rq[t] = wt(rng, so.rq[t], 1e-9);
                    sb.rvx = wt(rng, so.rvx, 1e-3);
                    sb.rvy = wt(rng, so.rvy, 1e-3);
                    sb.rvz = wt(rng, so.rvz, 1e-3);
                    sb.rw = wt(rng, so.rw, 1e-3);
                    double jv = sb.rg * sod.rg.d[0] + sb.e * sod.e.d[0] + sb.rvx * sod.rvx.d[0] + sb.rvy * sod.rvy.d[0] +
                                sb.rvz * sod.rvz.d[0] + sb.rw * sod.rw.d[0];
                    for (int t = 0; t < 3; t++) jv += sb.rq[t] * sod.rq[t].d[0];
                    SedInB xbs;
                    zero(xbs);
                    sediment_b(g, k, u, si, sb, xbs);
                    double vx = 0.0;
                    for (int a = 0; a < 4; a++)
                        for (int t = 0; t < 3; t++) vx += xbs.fq[a][t] * vv.fq[a][t];
                    for (int a = 0; a < 2; a++)
                        for (int i = 0; i < 5; i++) vx += xbs.fs[a][i] * vv.fs[a][i];
                    for (int a = 0; a < 3; a++)
                        for (int i = 0; i < 2; i++) vx += xbs.fh[a][i] * vv.fh[a][i];
                    vx += xbs.rw_c * vv.rw_c + xbs.rw_p * vv.rw_p + xbs.rg_m * vv.rg_m + xbs.rg_c * vv.rg_c +
                          xbs.rg_p * vv.rg_p + xbs.re0 * vv.re0 + xbs.rvx * vv.rvx + xbs.rvy * vv.rvy + xbs.rvz * vv.rvz;
                    for (int t = 0; t < 3; t++) vx += xbs.rq[t] * vv.rq[t];
                    err[2] = fmax(err[2], rel(vx, jv));
                    count[2]++;
                    x.rhog = so.rg;
                    x.rhoge = so.e;
                    x.rhogq[QR] = so.rq[0];
                    x.rhogq[QS] = so.rq[1];
                    x.rhogq[QG] = so.rq[2];
                } else {
                    x.rhog = wrhog[p] + 0.0;
                    x.rhoge = wrhoge[p] + 0.0;
                    x.rhogq[QR] = wrhogq[pq + QR * K2] + 0.0;
                    x.rhogq[QS] = wrhogq[pq + QS * K2] + 0.0;
                    x.rhogq[QG] = wrhogq[pq + QG * K2] + 0.0;
                }
                x.rhogq[QV] = wrhogq[pq + QV * K2] + 0.0;
                x.rhogq[QC] = wrhogq[pq + QC * K2] + 0.0;
                x.rhogq[QI] = wrhogq[pq + QI * K2] + 0.0;
                x.gsgam2 = f->gsgam2[p];
                // dual check of tail
                TailIn<D1> td;
                double vt_[8];
                td.rhog = seed1(x.rhog, vt_[0] = rng.u() * x.rhog);
                td.rhoge = seed1(x.rhoge, vt_[1] = rng.u() * x.rhoge);
                for (int n = 0; n < 6; n++)
                    td.rhogq[n] = seed1(x.rhogq[n], vt_[2 + n] = rng.u() * (fabs(x.rhogq[n]) > 0 ? x.rhogq[n] : 1e-6));
                td.gsgam2 = x.gsgam2;
                TailOut<D1> yd;
                const TailStatus st = tail(P, interior, td, yd);
                TailOut<double> y;
                tail(P, interior, x, y);
                if (st.newton < 0) continue;
                TailOut<double> yb;
                yb.rhog = wt(rng, y.rhog, 1e-3);
                yb.rhoge = wt(rng, y.rhoge, 1.0);
                yb.rho = wt(rng, y.rho, 1e-3);
                yb.tem = wt(rng, y.tem, 1.0);
                yb.pre = wt(rng, y.pre, 1.0);
                yb.qd = wt(rng, y.qd, 1e-3);
                for (int n = 0; n < 6; n++) {
                    yb.rhogq[n] = wt(rng, y.rhogq[n], 1e-9);
                    yb.q[n] = wt(rng, y.q[n], 1e-9);
                }
                double jv = yb.rhog * yd.rhog.d[0] + yb.rhoge * yd.rhoge.d[0] + yb.rho * yd.rho.d[0] +
                            yb.tem * yd.tem.d[0] + yb.pre * yd.pre.d[0] + yb.qd * yd.qd.d[0];
                for (int n = 0; n < 6; n++) jv += yb.rhogq[n] * yd.rhogq[n].d[0] + yb.q[n] * yd.q[n].d[0];
                TailInB xb;
                memset(&xb, 0, sizeof xb);
                tail_b(P, interior, x, yb, xb);
                double vx = xb.rhog * vt_[0] + xb.rhoge * vt_[1];
                for (int n = 0; n < 6; n++) vx += xb.rhogq[n] * vt_[2 + n];
                err[3] = fmax(err[3], rel(vx, jv));
                count[3]++;
            }
    return 0;
}

// ---------------------------------------------------------------------------------
// Whole-step reverse mode on the host: the passes of nsw6_adjoint_sweep.h (the same
// bodies the HIP kernels run) over every point, after the forward part rebuilt in the
// device workspace layout.  Checked against grad_dual_vjp; the GPU reverse mode is
// checked against this.
// ---------------------------------------------------------------------------------
#include "../nicam_kokkos_microphysics_amd/csrc/nsw6_adjoint_sweep.h"

namespace {

Fields host_fields(const nsw6_fields &s)
{
    Fields f;
    memset(&f, 0, sizeof f);
    static_assert(sizeof(Fields) == sizeof(nsw6_fields), "field lists differ");
    memcpy(&f, &s, sizeof f);   // same 50 pointers in the same order
    return f;
}

// Dual-number state seeded with the value of `f` and `ndir` tangent directions.
template <int N>
void seed_state(HostState<Dual<N>> &h, const nsw6_fields *f, int ndir, const nsw6_fields *dirs)
{
    const int64_t N2 = (int64_t)h.lall * KALL * h.ijdim;
    auto seed = [&](std::vector<Dual<N>> &v, const double *src, size_t off, int64_t n) {
        v.resize(n);
        for (int64_t e = 0; e < n; e++) {
            v[e] = Dual<N>(src[e]);
            for (int i = 0; i < ndir; i++) {
                const double *dp = *(const double *const *)((const char *)&dirs[i] + off);
                v[e].d[i] = dp ? dp[e] : 0.0;
            }
        }
    };
#define SEED(field, n) seed(h.field, f->field, offsetof(nsw6_fields, field), n)
    SEED(rhog, N2);
    SEED(rhogvx, N2);
    SEED(rhogvy, N2);
    SEED(rhogvz, N2);
    SEED(rhogw, N2);
    SEED(rhoge, N2);
    SEED(rhogq, N2 * NQ);
    SEED(vx, N2);
    SEED(vy, N2);
    SEED(vz, N2);
    SEED(tem, N2);
    SEED(q, N2 * NQ);
#undef SEED
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {
// The host run of the reverse passes for NT sedimenting tracers (grad_adjoint_host).
template <int NT>
int adjoint_host_n(const DevParams &P, const GridDev &g, int lall, int ijdim, const nsw6_fields *state,
                   const nsw6_fields *cot, const double *cot_sflux, const nsw6_fields *grad, double *grad_coef,
                   double dt)
{
    const int64_t K2 = (int64_t)KALL * ijdim, N2 = (int64_t)lall * K2, C = (int64_t)lall * ijdim;
    std::vector<double> w_rhog(N2), w_rhoge(N2), w_rhogq(N2 * NQ), w_vt(N2 * NT),
        w_flux(N2 * FxSlots<NT>::N, 0.0);
    std::vector<int> w_bounds((size_t)lall * ext_per_region<NT>(), 0), w_flag(lall, 0);
    std::vector<double> diag(3 * N2 + 13 * C, 0.0), rec((size_t)adj_record_slots<NT>() * N2, 0.0);

    AdjArgs A;
    memset(&A, 0, sizeof A);
    StepArgs &a = A.a;
    a.f = host_fields(*state);
    a.f.re_liquid = diag.data();
    a.f.re_cld = diag.data() + N2;
    a.f.GPREC = diag.data() + 2 * N2;
    double *dc = diag.data() + 3 * N2;
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
    a.ws.rhog = w_rhog.data();
    a.ws.rhoge = w_rhoge.data();
    a.ws.rhogq = w_rhogq.data();
    a.ws.vt = w_vt.data();
    a.ws.flux = w_flux.data();
    a.ws.bounds = w_bounds.data();
    a.ws.region_flag = w_flag.data();
    a.P = &P;
    a.grid = &g;
    a.lall = lall;
    a.ijdim = ijdim;
    a.dt = dt;
    a.dt2 = pow(dt, 2);
    a.dt3 = pow(dt, 3);
    a.dt3_6 = a.dt3 / 6.0;
    a.kin2kin = (P.kin2kin != 0.0);
    a.four_water = (NT == 4);
    A.a1_tab = A1_TAB;
    A.a2_tab = A2_TAB;
    A.cot = host_fields(*cot);
    A.cot_sflux = cot_sflux;
    A.grad = host_fields(*grad);
    A.cr = rec.data();
    A.wr = A.cr + (size_t)CellRec<NT>::N * N2;
    A.dws = A.wr + (size_t)WallRec<NT>::N * N2;
    A.din = A.dws + (size_t)NDW * N2;

    double tun[NCOEF];
    tunables(P, tun);
    const Fields &f = a.f;
    // ---- forward part: rates (k_rates without the diagnostics)
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l < lall; l++)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++) {
                const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
                const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;
                const int64_t pv = (int64_t)l * NT * K2 + (int64_t)k * ijdim + ij;
                if (k == 0 || k == KALL - 1) {
                    w_rhog[p] = f.rhog[p];
                    w_rhoge[p] = f.rhoge[p];
                    for (int n = 0; n < NQ; n++) w_rhogq[pq + n * K2] = f.rhogq[pq + n * K2];
                    for (int t = 0; t < NT; t++) w_vt[pv + t * K2] = 0.0;
                    continue;
                }
                PointIn<double> x;
                x.rhog = f.rhog[p];
                for (int n = 0; n < NQ; n++) x.rhogq[n] = f.rhogq[pq + n * K2];
                x.tem = f.tem[p];
                x.qv = f.q[pq];
                x.gsgam2 = f.gsgam2[p];
                PointOut<double> y;
                point_rates<double, double>(P, dt, x, tun, A1_TAB, A2_TAB, y);
                w_rhog[p] = y.rhog;
                w_rhoge[p] = y.rhoge;
                for (int n = 0; n < NQ; n++) w_rhogq[pq + n * K2] = y.rhogq[n];
                for (int t = 0; t < NT; t++) w_vt[pv + t * K2] = fall_speed<NT>(y, t);
            }
    // ---- kcell bounds in the device layout (kcell_max is what the fluxes read)
    Bounds b;
    b.kcmin.assign((size_t)lall * NT * KALL, 0);
    b.kcmax.assign((size_t)lall * NT * KALL, 0);
    b.khmin.assign((size_t)lall * NT * KALL, 0);
    b.khmax.assign((size_t)lall * NT * KALL, 0);
    b.general.assign(lall, 0);
    for (int l = 0; l < lall; l++) region_bounds<NT>(g, dt, a.dt2, a.dt3_6, l, ijdim, w_vt, b);
    for (int l = 0; l < lall; l++) {
        w_flag[l] = b.general[l];
        int *bl = &w_bounds[(size_t)l * ext_per_region<NT>()];
        for (int t = 0; t < NT; t++)
            for (int k = 0; k < KALL; k++) {
                bl[bounds_index(t, 0, k, 0)] = b.kcmin[((size_t)l * NT + t) * KALL + k];
                bl[bounds_index(t, 1, k, 0)] = b.khmin[((size_t)l * NT + t) * KALL + k];
                bl[bounds_index(t, 0, k, 1)] = b.kcmax[((size_t)l * NT + t) * KALL + k];
                bl[bounds_index(t, 1, k, 1)] = b.khmax[((size_t)l * NT + t) * KALL + k];
            }
    }
    // ---- wall fluxes
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l < lall; l++)
        for (int w = 0; w < KALL; w++)
            for (int ij = 0; ij < ijdim; ij++) {
                if (w_flag[l])
                    flux_wall_general<NT>(a, l, w, ij);
                else
                    flux_wall<NT>(a, l, w, ij);
            }
    // ---- reverse passes
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l < lall; l++)
        for (int k = 0; k < KALL; k++)
            for (int ij = 0; ij < ijdim; ij++) adj_cell<NT>(A, l, k, ij);
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l < lall; l++)
        for (int w = 0; w < KALL; w++)
            for (int ij = 0; ij < ijdim; ij++) adj_wall<NT>(A, l, w, ij);
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l < lall; l++)
        for (int ij = 0; ij < ijdim; ij++) adj_walls_general<NT>(A, l, ij);
    std::vector<double> tb((size_t)lall * KALL * NCOEF, 0.0);
#pragma omp parallel for collapse(2) schedule(static)
    for (int l = 0; l < lall; l++)
        for (int k = 0; k < KALL; k++) {
            double *t = &tb[((size_t)l * KALL + k) * NCOEF];
            for (int ij = 0; ij < ijdim; ij++) adj_point<NT>(A, l, k, ij, t);
        }
    if (grad_coef)
        for (int c = 0; c < NCOEF; c++) {
            double s = 0.0;
            for (size_t e = 0; e < (size_t)lall * KALL; e++) s += tb[e * NCOEF + c];
            grad_coef[c] = s;
        }
    return NSW6_OK;
}
}  // namespace

extern "C" {

// Host reverse mode with the arguments of nsw6_mp_driver_adj (host arrays).  Returns
// NSW6_OK, or NSW6_UNSUPPORTED for a scheme these passes do not cover: anything but
// the Roh & Satoh or the exponential snow moments, KK2000 or Berry's autoconversion,
// with or without explicit ice generation, KIN2EIN or KIN2KIN, 3WATER or 4WATER, no
// OPT_INDIR (the device reverse mode also covers OPT_INDIR).
int grad_adjoint_host(const double *gz, const double *gzh, const nsw6_params *params, int lall, int ijdim,
                      const nsw6_fields *state, const nsw6_fields *cot, const double *cot_sflux,
                      const nsw6_fields *grad, double *grad_coef, double dt)
{
    DevParams P;
    to_dev_params(*params, P);
    if (!((P.sw_roh2014 == 1.0 || P.sw_roh2014 == 0.0) && (P.sw_kk2000 == 1.0 || P.sw_kk2000 == 0.0) &&
          (P.sw_expice == 1.0 || P.sw_expice == 0.0) && P.opt_indir == 0.0))
        return NSW6_UNSUPPORTED;
    GridDev g;
    make_grid_dev(gz, gzh, g);
    return P.four_water != 0.0
               ? adjoint_host_n<4>(P, g, lall, ijdim, state, cot, cot_sflux, grad, grad_coef, dt)
               : adjoint_host_n<3>(P, g, lall, ijdim, state, cot, cot_sflux, grad, grad_coef, dt);
}

// J = sum <cot, y(x)> + sum cot_sflux sflux over the step's outputs (the output list
// of nsw6_mp_driver_adj; NULL cotangents = 0) and its directional derivatives along
// `ndir` <= 8 input directions `dirs` (NULL members = 0) and coefficient directions
// dcoef [ndir][8], by forward-mode dual numbers.  Returns 0 or 1 (Newton failure).
int grad_dual_vjp(const double *gz, const double *gzh, const nsw6_params *params, int lall, int ijdim,
                  const nsw6_fields *f, double dt, const nsw6_fields *cot, const double *cot_sflux, int ndir,
                  const nsw6_fields *dirs, const double *dcoef, double *J, double *dJ)
{
    if (ndir < 1 || ndir > GRAD_NDIR) return -2;
    typedef Dual<GRAD_NDIR> D;
    DevParams P;
    to_dev_params(*params, P);
    GridDev g;
    make_grid_dev(gz, gzh, g);
    const int64_t N2 = (int64_t)lall * KALL * ijdim, NC = (int64_t)lall * ijdim;
    HostState<D> h;
    h.lall = lall;
    h.ijdim = ijdim;
    seed_state(h, f, ndir, dirs);
    bind_metrics(h, f);
    alloc_outputs(h, P.four_water != 0.0 ? 4 : 3);
    double t0[NCOEF];
    tunables(P, t0);
    D tun[NCOEF];
    for (int c = 0; c < NCOEF; c++) {
        tun[c] = D(t0[c]);
        for (int i = 0; i < ndir; i++) tun[c].d[i] = dcoef ? dcoef[i * NCOEF + c] : 0.0;
    }
    int fails = 0;
    const int st = host_step(P, g, dt, tun, h, &fails);
    if (st != 0) return st;
    D S(0.0);
    auto dot = [&](const double *c, const std::vector<D> &y) {
        if (!c) return;
        D s(0.0);
        for (size_t e = 0; e < y.size(); e++) s += c[e] * y[e];
        S += s;
    };
    dot(cot->rhog, h.rhog);
    dot(cot->rhogvx, h.rhogvx);
    dot(cot->rhogvy, h.rhogvy);
    dot(cot->rhogvz, h.rhogvz);
    dot(cot->rhogw, h.rhogw);
    dot(cot->rhoge, h.rhoge);
    dot(cot->rhogq, h.rhogq);
    dot(cot->rho, h.rho);
    dot(cot->tem, h.tem);
    dot(cot->pre, h.pre);
    dot(cot->q, h.q);
    dot(cot->qd, h.qd);
    dot(cot->precip, h.precip);
    dot(cot->precip_rhoe, h.prhoe);
    dot(cot->precip_lh_heat, h.lh);
    dot(cot->precip_rhophi, h.prhophi);
    dot(cot->precip_rhokin, h.prhokin);
    dot(cot->GPREC, h.GPREC);
    dot(cot_sflux, h.sflux);
    (void)N2;
    (void)NC;
    *J = S.v;
    for (int i = 0; i < ndir; i++) dJ[i] = S.d[i];
    return fails ? 1 : 0;
}

// The device's table-driven exp (which = 0) and log (which = 1) of nsw6_math.h, run on
// the host with host copies of its tables, for the accuracy test.
int grad_math_tang(int which, const double *x, double *y, int64_t n)
{
    static const double exp2_tab[64] = {NSW6_EXP2_TAB};
    static const double log_tab[130] = {NSW6_LOG_TAB};
    for (int64_t i = 0; i < n; i++) y[i] = which == 0 ? exp_tang(x[i], exp2_tab) : log_tang(x[i], log_tab);
    return (which == 0 || which == 1) ? 0 : 1;
}
// The tables themselves (64 and 130 doubles), for the test that recomputes them.
void grad_math_tables(double *exp2_tab, double *log_tab)
{
    static const double e[64] = {NSW6_EXP2_TAB};
    static const double l[130] = {NSW6_LOG_TAB};
    memcpy(exp2_tab, e, sizeof(e));
    memcpy(log_tab, l, sizeof(l));
}

}  // extern "C"
