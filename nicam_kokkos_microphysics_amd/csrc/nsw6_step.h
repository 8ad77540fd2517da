// nsw6_step.h -- the point, wall and cell pieces of one NSW6 step, templated on the
// scalar type S.  The HIP kernels instantiate them with S = double (the FP64 form;
// the mixed form adds an FP32 rate type R); the host dual-number gradient check
// instantiates them with a forward-mode dual number, so the adjoint is checked
// against derivatives of exactly this arithmetic.  Operation order follows the
// reference (paths relative to the reference repository) so that S = double
// reproduces it to round-off.
//
// A scalar type S provides the arithmetic operators, sqrt/exp/log/expm1/fabs/
// copysign, pow(double, S), dmax/dmin (select by value) and val(S) -> double.
#pragma once

#include "nsw6_device.h"
#include "nsw6_internal.h"

namespace nsw6 {

// ---------------------------------------------------------------------------------
// Point-wise rate update (k_rates): first negative_filter (src/mod_mp_nsw6.cpp:
// 1770-1831), THRMDYN (src/mod_thrmdyn.cpp:98-180), psat/qsat (src/mod_satadjust.cpp:
// 80-106, 133-160), TagBigLoop (src/mod_mp_nsw6.cpp:760-1594) with rate type R, and
// TagUpdateRHOGQ/RHOGE (:692-758).  Interior points only.
// ---------------------------------------------------------------------------------
template <typename S> struct PointIn {
    S rhog, rhogq[6], tem, qv;   // qv: the input q(QV) (the reference uses it stale)
    double gsgam2;
};
template <typename S> struct PointOut {
    S rhog, rhoge, rhogq[6], vt[3];   // pre-sedimentation state and fall speeds of QR, QS, QG
    S vti;                            // fall speed of QI ("4WATER" only, else 0)
    S rho, q[6];                      // density and q after the filter (q(QV) stale)
    RateSwitches sw;
};

template <typename S> NSW6_HD S psat_of(double LovR, double RTEM00, S tem)
{
    const S rtem = S(1.0) / dmax(tem, S(TEM_MIN));
    return S(C_PSAT0) * 1.0 * exp(LovR * (RTEM00 - rtem));
}

// a1_tab/a2_tab: the Bergeron tables (read only without explicit ice generation,
// sw_expice != 1).
template <typename S, typename R>
NSW6_HD void point_rates(const DevParams &P, double dt, const PointIn<S> &x, const S tun[NCOEF],
                         const double *a1_tab, const double *a2_tab, PointOut<S> &y, const NcPoint &nc)
{
    S rq[NQ];
#pragma unroll
    for (int n = 0; n < NQ; n++) rq[n] = x.rhogq[n];
    S diffq = S(0.0);
#pragma unroll
    for (int n = 2; n <= 5; n++) {
        diffq += rq[n];
        rq[n] = dmax(rq[n], S(0.0));
    }
#pragma unroll
    for (int n = 2; n <= 5; n++) diffq -= rq[n];
    rq[QV] += diffq;
    diffq = rq[QV];
    rq[QV] = dmax(rq[QV], S(0.0));
    diffq -= rq[QV];
    const S rhog = x.rhog * (1.0 - diffq);
    const S rho = rhog / x.gsgam2;
    S q[NQ];
    q[QV] = x.qv;
    const S rrhog = S(1.0) / rhog;
#pragma unroll
    for (int n = 1; n <= 5; n++) q[n] = rq[n] * rrhog;
    S qd = S(1.0);
#pragma unroll
    for (int n = 1; n <= 5; n++) qd = qd - q[n];
    S cva = qd * C_CVdry;
#pragma unroll
    for (int n = 1; n <= 5; n++) cva = cva + q[n] * C_CVW;
    const S tem = x.tem;
    S rhoge = tem * rhog * cva;
    const S pre = rho * (qd * C_Rdry + q[QV] * C_Rvap) * tem;
    const S rqden = S(1.0) / (rho * C_Rvap * tem);
    const S qsatl = psat_of(P.LovR_liq, P.RTEM00, tem) * rqden;
    const S qsati = psat_of(P.LovR_ice, P.RTEM00, tem) * rqden;

    RateIn<R> in;
    in.dens = R(rho);
    in.temp = R(tem);
#pragma unroll
    for (int n = 0; n < NQ; n++) in.q[n] = R(q[n]);
    in.qsatl = R(qsatl);
    in.qsati = R(qsati);
    in.pre = R(pre);
    in.rhog = R(rhog);
    S a1 = S(1.0), a2 = S(0.0), ma2 = S(1.0);
    if (P.sw_expice != 1.0) bergeron(a1_tab, a2_tab, tem, a1, a2, ma2);
    in.a1 = R(a1);
    in.a2 = R(a2);
    in.ma2 = R(ma2);
    double qv_[NQ];
#pragma unroll
    for (int n = 0; n < NQ; n++) qv_[n] = val(q[n]);
    in.sw = rate_switches(P, val(rho), val(tem), qv_, val(qsati));
    rate_differences<R, S>(in, tem, q, qsatl, qsati);
#pragma unroll
    for (int i = 0; i < NCOEF; i++) in.tun[i] = R(tun[i]);
    RateOut<R> o;
    rates<R>(P, R(dt), in, o, nc);

    // drhogq = rhog * q_t * dt with qv_t = -(sum of the others) (src/mod_mp_nsw6.cpp:
    // 1548-1559) in the precise type, so the point's total water is conserved to its
    // round-off; then TagUpdateRHOGQ / TagUpdateRHOGE
    S qt[NQ];
#pragma unroll
    for (int n = 1; n < NQ; n++) qt[n] = S(o.qt[n]);
    qt[QV] = -(qt[QC] + qt[QR] + qt[QI] + qt[QS] + qt[QG]);
    S drq[NQ];
#pragma unroll
    for (int n = 0; n < NQ; n++) drq[n] = rhog * qt[n] * dt;
#pragma unroll
    for (int n = 0; n < NQ; n++) rq[n] += drq[n];
    rhoge = rhoge - C_LHV0 * drq[QV] + C_LHF0 * drq[QI] + C_LHF0 * drq[QS] + C_LHF0 * drq[QG];

    y.rhog = rhog;
    y.rhoge = rhoge;
#pragma unroll
    for (int n = 0; n < NQ; n++) y.rhogq[n] = rq[n];
    y.vt[0] = S(o.Vtr);
    y.vt[1] = S(o.Vts);
    y.vt[2] = S(o.Vtg);
    y.vti = S(o.Vti);
    y.rho = rho;
#pragma unroll
    for (int n = 0; n < NQ; n++) y.q[n] = q[n];
    y.sw = in.sw;
}
template <typename S, typename R>
NSW6_HD void point_rates(const DevParams &P, double dt, const PointIn<S> &x, const S tun[NCOEF],
                         const double *a1_tab, const double *a2_tab, PointOut<S> &y)
{
    point_rates<S, R>(P, dt, x, tun, a1_tab, a2_tab, y, nc_default(P));
}

// ---------------------------------------------------------------------------------
// Departure distance of vadv1d_prep (src/mod_vadv1d.cpp:270-302):
// zdis = dt*wh - dt^2*wh*dw/dz/2 + dt^3*wh*((dw/dz)^2 + wh*d2w/dz2)/6, with the grid
// divisions taken as products with reciprocals computed once on the host.
// ---------------------------------------------------------------------------------
template <typename S>
NSW6_HD S zdis_interior(double dt, double dt2, double dt3_6, S whm, S whk, S whp, double rdzm, double rdzk,
                        double rs)
{
    const S d = whp - whm;
    const S t1 = dt * whk;
    const S t2 = dt2 * whk * d * rs * 0.5;
    const S g = d * rs;
    const S inner = ((whp - whk) * rdzk - (whk - whm) * rdzm) * rs * 2.0;
    const S t3 = dt3_6 * whk * (g * g + whk * inner);
    return t1 - t2 + t3;
}
template <typename S> NSW6_HD S zdis_bottom(double dt, double dt2, S whk, S whp, double rdzk)
{
    return dt * whk - dt2 * whk * (whp - whk) * rdzk * 0.5;
}
// With V[d + 3] = Vt(k + d), d = -3..+1: wh_full(V, k, d) is the full-level wh at wall
// k + d (wh(0) = 0, wh(1) = Vt(1), wh(95) = Vt(94)) and wh_half the half-level one
// (Vtermh of levels 2..93 averaged again).
template <typename S> NSW6_HD S wh_full(const S V[5], int k, int d)
{
    const int j = k + d;
    if (j == 0) return S(0.0);
    if (j == 1) return V[d + 3];
    if (j <= KMAX) return 0.5 * (V[d + 2] + V[d + 3]);
    return V[d + 2];
}
template <typename S> NSW6_HD S vh_half(const S V[5], int k, int d)
{
    const int j = k + d;
    return (j >= 2 && j <= KMAX - 1) ? 0.5 * (V[d + 3] + V[d + 2]) : S(0.0);
}
template <typename S> NSW6_HD S wh_half(const S V[5], int k, int d)
{
    const int j = k + d;
    if (j >= 3 && j <= KMAX) return 0.5 * (vh_half(V, k, d - 1) + vh_half(V, k, d));
    return (j == 2) ? vh_half(V, k, d) : S(0.0);
}
template <typename S>
NSW6_HD S zdis_full_at(const GridDev &g, double dt, double dt2, double dt3_6, const S V[5], int k)
{
    const S whp = wh_full(V, k, 1);
    if (k == KMIN) return zdis_bottom(dt, dt2, wh_full(V, k, 0), whp, g.rdgz[KMIN]);
    return zdis_interior(dt, dt2, dt3_6, wh_full(V, k, -1), wh_full(V, k, 0), whp, g.rdgz[k - 1], g.rdgz[k],
                         g.rsum_gz[k]);
}
template <typename S>
NSW6_HD S zdis_half_at(const GridDev &g, double dt, double dt2, double dt3_6, const S V[5], int k)
{
    const S whp = wh_half(V, k, 1);
    if (k == KMIN + 1) return zdis_bottom(dt, dt2, wh_half(V, k, 0), whp, g.rdgzh[KMIN + 1]);
    return zdis_interior(dt, dt2, dt3_6, wh_half(V, k, -1), wh_half(V, k, 0), whp, g.rdgzh[k - 1], g.rdgzh[k],
                         g.rsum_gzh[k]);
}
// per-column departure cell (src/mod_vadv1d.cpp:355-384)
NSW6_HD int departure_cell(const double *zh, int k, int kcmin, int kcmax, double zdis, int mkmin, int mkmax)
{
    int kc;
    if (kcmin == k && kcmax == k) {
        kc = k;
    } else {
        kc = 0;
        const double x = zh[k] - zdis;
        for (int k2 = kcmin; k2 <= kcmax; k2++) {
            const int tmp = (int)((double)k2 * copysign(1.0, x - zh[k2]) * copysign(1.0, zh[k2 + 1] - x));
            kc = (kc > tmp) ? kc : tmp;
        }
    }
    if (kc == 0) kc = mkmin;
    if (kc == mkmax + 1) kc = mkmax;
    return kc;
}
// The swept-cell factor of cell k2 for wall k and departure cell kc
// (src/mod_vadv1d.cpp:426-436): -dz for k <= k2 < kc, +dz for kc < k2 < k, else 0.
NSW6_HD double fact_of(int k2, int kc, int k, double dz)
{
    const double f1 = (double)(k2 - (kc + 1)), f2 = (double)((k - 1) - k2);
    const double f3 = (double)(k2 - k), f4 = (double)((kc - 1) - k2);
    return dz * 0.25 * ((copysign(1.0, f1) + 1.0) * (copysign(1.0, f2) + 1.0) -
                        (copysign(1.0, f3) + 1.0) * (copysign(1.0, f4) + 1.0));
}
// |f| < EPS -> 0 (src/mod_vadv1d.cpp:445-449)
template <typename S> NSW6_HD S eps_filter(S f) { return f * (0.5 + copysign(0.5, fabs(val(f)) - C_EPS)); }
NSW6_HD double eps_mask(double f) { return 0.5 + copysign(0.5, fabs(f) - C_EPS); }

// ---------------------------------------------------------------------------------
// Transported quantities of one cell (src/mod_precip_transport.cpp:740-780, 846-888)
// ---------------------------------------------------------------------------------
template <typename S> struct CellIn {
    S rhog, rq[5], rhoge;           // pre-sedimentation state; rq = QC, QR, QI, QS, QG
    S rvx, rvy, rvz, rw, vx, vy, vz;
    double gs, gam2, gsh, gam2h, z;
};
// Transported data of one cell for NT sedimenting tracers t (nq = sed_nq<NT>(t)).
template <typename S, int NT> struct CellN {
    S rhoq[NT], qrgs[NT], q[NT];
    S tem, kinh, up, vp;
    S hv[NT][2];                // half-level values at the cell's lower wall
    S rhog, kinv;
    double z, gs;
};
template <typename S> using Cell = CellN<S, 3>;   // "3WATER": QR, QS, QG
struct Unit {   // horizontal unit vectors of a column
    double ix, iy, iz, jx, jy, jz;
};

// Cell j from its raw data and the part of cell j-1 its lower wall needs
// (below_q: q of the NT sedimenting tracers of cell j-1).
template <typename S, int NT>
NSW6_HD void derive_cell(const GridDev &g, int j, const CellIn<S> &r, S below_rhog, const S *below_q,
                         double below_gs, const Unit &u, CellN<S, NT> &c)
{
    const bool interior = (j >= KMIN && j <= KMAX);
    const S rg = r.rhog;
    const double rgs_inv = 1.0 / r.gs;
    const double rgs = interior ? r.gam2 * rgs_inv : 0.0;
    const S rrg = S(1.0) / rg;
    S qq[5];
#pragma unroll
    for (int n = 0; n < 5; n++) qq[n] = r.rq[n] * rrg;
    S qd = S(1.0);
#pragma unroll
    for (int n = 0; n < 5; n++) qd = qd - qq[n];
    S cva = qd * C_CVdry;
#pragma unroll
    for (int n = 0; n < 5; n++) cva = cva + qq[n] * C_CVW;
    c.tem = interior ? r.rhoge / (rg * cva) : S(0.0);
    c.kinh = interior ? 0.5 * (r.rvx * r.rvx + r.rvy * r.rvy + r.rvz * r.rvz) * rrg : S(0.0);
    c.up = r.vx * u.ix + r.vy * u.iy + r.vz * u.iz;
    c.vp = r.vx * u.jx + r.vy * u.jy + r.vz * u.jz;
    c.z = r.z;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        c.q[t] = qq[sed_nq<NT>(t) - 1];                 // rq holds QC..QG
        c.rhoq[t] = r.rq[sed_nq<NT>(t) - 1] * rgs;
    }
#pragma unroll
    for (int t = 0; t < NT; t++) c.qrgs[t] = c.q[t] * rgs;
    c.rhog = rg;
    c.gs = r.gs;
    double c2w0 = 0.0, c2w1 = 0.0;
    if (j >= KMIN) {
        c2w0 = g.afact[j] * rgs_inv * r.gsh;
        c2w1 = g.bfact[j] / below_gs * r.gsh;
    }
    if (j >= KMIN + 1 && j <= KMAX) {
        c.kinv = 0.5 * (r.rw * r.rw) / (c2w0 * rg + c2w1 * below_rhog);
        const double rgsh = r.gam2h / r.gsh;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const S qh = 0.5 * (c.q[t] + below_q[t]) * rgsh;
            c.hv[t][0] = qh * c.kinv;
            c.hv[t][1] = qh * r.rw;
        }
    } else {
        c.kinv = S(0.0);
#pragma unroll
        for (int t = 0; t < NT; t++) c.hv[t][0] = c.hv[t][1] = S(0.0);
    }
}

template <typename S, int NT> NSW6_HD S rhof_cell(const CellN<S, NT> &c, int t, int i)
{
    switch (i) {
    case 0: return c.rhoq[t];
    case 1: return c.rhoq[t] * C_CVW * c.tem;
    case 2: return c.rhoq[t] * C_GRAV * c.z;
    case 3: return c.qrgs[t] * c.kinh;
    case 4: return c.rhoq[t] * c.up;
    default: return c.rhoq[t] * c.vp;
    }
}

// ---------------------------------------------------------------------------------
// Fluxes of wall w (1..94) of a single-cell region: the closed forms of
// vadv1d_getflux_new (src/mod_vadv1d.cpp:393-450) for the range [w, w+1]
//     kc = w   : f = rhof(w) * zdis0
//     kc = w+1 : f = -rhof(w) * dz(w) + rhof(w+1) * zdis0   (zdis0 kept, not refreshed)
// with the half level on the full level's kcell_max (src/mod_precip_transport.cpp:
// 890-892).  Non-mass fluxes are summed over the tracers in nq order.  sflux[t] is
// the tracer's surface flux -f(kmin)/dt and feeds precip (:900-958).
// ---------------------------------------------------------------------------------
template <typename S, int NT> struct WallFluxN {
    S fq[NT], fs[5], fh[2];
};
template <typename S> using WallFlux = WallFluxN<S, 3>;
template <int NT> struct WallGeomN {   // departure decisions of one wall and column
    bool up1[NT];           // full level: the departure cell is w+1
    bool half_next[NT];     // half level: the cell taken is w+1
    bool half_full[NT];     // half level: the two-cell form -h(w) dz + h(w+1) zdis0
};
using WallGeom = WallGeomN<3>;

template <typename S, int NT>
NSW6_HD void wall_geometry(const GridDev &g, int w, double dt, double dt2, double dt3_6, const S (*V)[5],
                           const int *kcmax, const int *khmax, S *z0, S *zh0, WallGeomN<NT> &geo)
{
#pragma unroll
    for (int t = 0; t < NT; t++) {
        z0[t] = zdis_full_at(g, dt, dt2, dt3_6, V[t], w);
        const int kc = departure_cell(g.gzh, w, w, kcmax[t], val(z0[t]), KMIN, KMAX);
        geo.up1[t] = (kcmax[t] != w) && (kc == w + 1);
        if (w >= KMIN + 1) {
            zh0[t] = zdis_half_at(g, dt, dt2, dt3_6, V[t], w);
            const int kch = departure_cell(g.gz_shift, w, w, khmax[t], val(zh0[t]), KMIN + 1, KMAX);
            geo.half_next[t] = (kch == w + 1);
            geo.half_full[t] = (kcmax[t] != w) && (kch == w + 1);
        } else {
            zh0[t] = S(0.0);
            geo.half_next[t] = geo.half_full[t] = false;
        }
    }
}

template <typename S, int NT>
NSW6_HD void wall_flux(const GridDev &g, int w, const CellN<S, NT> &cur, const CellN<S, NT> &nxt, const S *z0,
                       const S *zh0, const WallGeomN<NT> &geo, WallFluxN<S, NT> &F, S (*ft_surf)[4])
{
#pragma unroll
    for (int i = 0; i < 5; i++) F.fs[i] = S(0.0);
    F.fh[0] = F.fh[1] = S(0.0);
    const double dzw = g.dgz[w];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        S ft[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const S r0 = rhof_cell(cur, t, i);
            const S v = geo.up1[t] ? S(-(r0 * dzw) + rhof_cell(nxt, t, i) * z0[t]) : S(r0 * z0[t]);
            ft[i] = eps_filter(v);
        }
        F.fq[t] = ft[0];
#pragma unroll
        for (int i = 1; i < 6; i++) F.fs[i - 1] += ft[i];
#pragma unroll
        for (int i = 0; i < 4; i++) ft_surf[t][i] = ft[i];
        if (w >= KMIN + 1) {
            const double dzh = g.dgzh[w];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const S hc = cur.hv[t][i], hn = nxt.hv[t][i];
                S v;
                if (geo.half_full[t])
                    v = -(hc * dzh) + hn * zh0[t];
                else
                    v = (geo.half_next[t] ? hn : hc) * zh0[t];
                F.fh[i] += eps_filter(v);
            }
        }
    }
}

// ---------------------------------------------------------------------------------
// Fluxes of wall w of a general region: departure cells anywhere in the region's kcell
// range (src/mod_vadv1d.cpp:393-450), including its assignment of zdis inside the
// k2 loop (the last k2 = kcell_max wins) and, for the half level, the range
// kcellh_min .. kcell_max of the full level (src/mod_precip_transport.cpp:890-892).
// The same outputs as wall_flux; the swept cells come from cell_at(j, c), which
// derives cell j (with cell j-1 below it).
// ---------------------------------------------------------------------------------
template <typename S, int NT> struct GeneralGeomN {
    int kcmin[NT], kcmax[NT], kc[NT];   // full level: the region's range, the column's cell
    int khmin[NT], kch[NT];             // half level (walls 2..94)
    S z0[NT], zh0[NT];
    int lo, hi;                         // the cells every range covers together
};

template <typename S, int NT>
NSW6_HD void general_geometry(const GridDev &g, int w, double dt, double dt2, double dt3_6, const S (*V)[5],
                              const int *kcmin, const int *kcmax, const int *khmin, const int *khmax,
                              GeneralGeomN<S, NT> &G)
{
    G.lo = w;
    G.hi = w;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        G.kcmin[t] = kcmin[t];
        G.kcmax[t] = kcmax[t];
        G.z0[t] = zdis_full_at(g, dt, dt2, dt3_6, V[t], w);
        G.kc[t] = departure_cell(g.gzh, w, kcmin[t], kcmax[t], val(G.z0[t]), KMIN, KMAX);
        G.lo = (kcmin[t] < G.lo) ? kcmin[t] : G.lo;
        G.hi = (kcmax[t] > G.hi) ? kcmax[t] : G.hi;
        if (w >= KMIN + 1) {
            G.khmin[t] = khmin[t];
            G.zh0[t] = zdis_half_at(g, dt, dt2, dt3_6, V[t], w);
            G.kch[t] = departure_cell(g.gz_shift, w, khmin[t], khmax[t], val(G.zh0[t]), KMIN + 1, KMAX);
            G.lo = (khmin[t] < G.lo) ? khmin[t] : G.lo;
        } else {
            G.khmin[t] = w;
            G.kch[t] = w;
            G.zh0[t] = S(0.0);
        }
    }
}

// zdis after the k2 loop: zdis0 for the one-cell range, else zdis0 - fact(range end)
NSW6_HD double general_zd_shift(const double *dz, int w, int kmin, int kmax, int kc)
{
    return (kmin == w && kmax == w) ? 0.0 : fact_of(kmax, kc, w, dz[kmax]);
}

// The sums of the k2 loops (acc: sum of rhof(k2) fact(k2) over the range, acch the
// half level's) and the departure cells' values (rfc = rhof(kcell), hvc), in the
// reference's summation order.
template <typename S, int NT, typename CellAt>
NSW6_HD void general_sums(const GridDev &g, int w, const GeneralGeomN<S, NT> &G, CellAt &&cell_at,
                          S (*acc)[6], S (*rfc)[6], S (*acch)[2], S (*hvc)[2])
{
#pragma unroll
    for (int t = 0; t < NT; t++) {
#pragma unroll
        for (int i = 0; i < 6; i++) acc[t][i] = rfc[t][i] = S(0.0);
        acch[t][0] = acch[t][1] = hvc[t][0] = hvc[t][1] = S(0.0);
    }
    const bool half = (w >= KMIN + 1);
    auto take = [&](int j, const CellN<S, NT> &c) {
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const bool one = (G.kcmin[t] == w && G.kcmax[t] == w);
            const bool in = !one && j >= G.kcmin[t] && j <= G.kcmax[t];
            const double fa = fact_of(j, G.kc[t], w, g.dgz[j]);
#pragma unroll
            for (int i = 0; i < 6; i++) {
                const S rf = rhof_cell(c, t, i);
                if (in) acc[t][i] = acc[t][i] + rf * fa;
                if (j == G.kc[t]) rfc[t][i] = rf;
            }
            if (half) {
                const bool hone = (G.khmin[t] == w && G.kcmax[t] == w);
                const bool hin = !hone && j >= G.khmin[t] && j <= G.kcmax[t];
                const double fh = fact_of(j, G.kch[t], w, g.dgzh[j]);
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    if (hin) acch[t][i] = acch[t][i] + c.hv[t][i] * fh;
                    if (j == G.kch[t]) hvc[t][i] = c.hv[t][i];
                }
            }
        }
    };
    for (int j = G.lo; j <= G.hi; j++) {
        CellN<S, NT> c;
        cell_at(j, c);
        take(j, c);
    }
    // a departure cell the search left at the domain's end outside the ranges
#pragma unroll
    for (int t = 0; t < NT; t++) {
        if (G.kc[t] < G.lo || G.kc[t] > G.hi) {
            CellN<S, NT> c;
            cell_at(G.kc[t], c);
#pragma unroll
            for (int i = 0; i < 6; i++) rfc[t][i] = rhof_cell(c, t, i);
        }
        if (half && (G.kch[t] < G.lo || G.kch[t] > G.hi)) {
            CellN<S, NT> c;
            cell_at(G.kch[t], c);
            hvc[t][0] = c.hv[t][0];
            hvc[t][1] = c.hv[t][1];
        }
    }
}

// Cell j-1 as derive_cell takes it for cell j's lower wall.
template <typename S, int NT> struct BelowOf {
    S rhog, q[NT];
    double gs;
};

// The forward form.  zdis after the k2 loop is known before it, so the departure
// cell's term rhof(kcell)·zdis joins the sum when the sweep reaches kcell instead of
// being kept aside (general_sums keeps it aside for the reverse mode): with NT tracers
// that frees 8·NT live values across the sweep, and the sum's order differs from the
// reference's by the position of that one term (round-off only).  The sweep derives
// each cell from its raw data (raw_at(j, r)) and from the cell below it, which is the
// cell derived just before (below_at(j) only for the first and for cells outside the
// ranges), and requests cell j+1's raw data before deriving cell j.
template <typename S, int NT, typename RawAt, typename BelowAt>
NSW6_HD void wall_flux_general(const GridDev &g, int w, const GeneralGeomN<S, NT> &G, const Unit &u,
                               RawAt &&raw_at, BelowAt &&below_at, WallFluxN<S, NT> &F, S (*ft_surf)[4])
{
    const bool half = (w >= KMIN + 1);
    S acc[NT][6], acch[NT][2], zd[NT], zh[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        zd[t] = G.z0[t] - general_zd_shift(g.dgz, w, G.kcmin[t], G.kcmax[t], G.kc[t]);
        zh[t] = half ? S(G.zh0[t] - general_zd_shift(g.dgzh, w, G.khmin[t], G.kcmax[t], G.kch[t])) : S(0.0);
#pragma unroll
        for (int i = 0; i < 6; i++) acc[t][i] = S(0.0);
        acch[t][0] = acch[t][1] = S(0.0);
    }
    auto take = [&](int j, const CellN<S, NT> &c) {
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const bool one = (G.kcmin[t] == w && G.kcmax[t] == w);
            const bool in = !one && j >= G.kcmin[t] && j <= G.kcmax[t];
            const double fa = fact_of(j, G.kc[t], w, g.dgz[j]);
#pragma unroll
            for (int i = 0; i < 6; i++) {
                const S rf = rhof_cell(c, t, i);
                if (in) acc[t][i] = acc[t][i] + rf * fa;
                if (j == G.kc[t]) acc[t][i] = acc[t][i] + rf * zd[t];
            }
            if (half) {
                const bool hone = (G.khmin[t] == w && G.kcmax[t] == w);
                const bool hin = !hone && j >= G.khmin[t] && j <= G.kcmax[t];
                const double fh = fact_of(j, G.kch[t], w, g.dgzh[j]);
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    if (hin) acch[t][i] = acch[t][i] + c.hv[t][i] * fh;
                    if (j == G.kch[t]) acch[t][i] = acch[t][i] + c.hv[t][i] * zh[t];
                }
            }
        }
    };
    auto cell_at = [&](int j, CellN<S, NT> &c) {
        CellIn<S> r;
        raw_at(j, r);
        const BelowOf<S, NT> b = below_at(j);
        derive_cell(g, j, r, b.rhog, b.q, b.gs, u, c);
    };
    {
        BelowOf<S, NT> b = below_at(G.lo);
        CellIn<S> rn;
        raw_at(G.lo, rn);
        for (int j = G.lo; j <= G.hi; j++) {
            const CellIn<S> r = rn;
            if (j < G.hi) raw_at(j + 1, rn);
            CellN<S, NT> c;
            derive_cell(g, j, r, b.rhog, b.q, b.gs, u, c);
            take(j, c);
            b.rhog = c.rhog;
#pragma unroll
            for (int t = 0; t < NT; t++) b.q[t] = c.q[t];
            b.gs = c.gs;
        }
    }
    // a departure cell the search left at the domain's end outside the ranges
#pragma unroll
    for (int t = 0; t < NT; t++) {
        if (G.kc[t] < G.lo || G.kc[t] > G.hi) {
            CellN<S, NT> c;
            cell_at(G.kc[t], c);
#pragma unroll
            for (int i = 0; i < 6; i++) acc[t][i] = acc[t][i] + rhof_cell(c, t, i) * zd[t];
        }
        if (half && (G.kch[t] < G.lo || G.kch[t] > G.hi)) {
            CellN<S, NT> c;
            cell_at(G.kch[t], c);
            acch[t][0] = acch[t][0] + c.hv[t][0] * zh[t];
            acch[t][1] = acch[t][1] + c.hv[t][1] * zh[t];
        }
    }
#pragma unroll
    for (int i = 0; i < 5; i++) F.fs[i] = S(0.0);
    F.fh[0] = F.fh[1] = S(0.0);
#pragma unroll
    for (int t = 0; t < NT; t++) {
        S ft[6];
#pragma unroll
        for (int i = 0; i < 6; i++) ft[i] = eps_filter(acc[t][i]);
        F.fq[t] = ft[0];
#pragma unroll
        for (int i = 1; i < 6; i++) F.fs[i - 1] += ft[i];
#pragma unroll
        for (int i = 0; i < 4; i++) ft_surf[t][i] = ft[i];
        if (half) {
#pragma unroll
            for (int i = 0; i < 2; i++) F.fh[i] += eps_filter(acch[t][i]);
        }
    }
}

// Surface fluxes of wall kmin (src/mod_precip_transport.cpp:900-958) and mp_driver's
// scaling by fraction_mp = 1 (src/mod_mp_driver.cpp:560-612).  Tracer 0 (QR) is
// liquid; QI, QS and QG are solid (precip row 1 and the fusion heat).
template <typename S> struct Surface {
    S precip0, precip1, lh, prhoe, prhophi, prhokin;
};
template <int NT = 3, typename S> NSW6_HD void surface_fluxes(const S (*ft)[4], double dt, Surface<S> &s)
{
    S precip0 = S(0.0), precip1 = S(0.0), lh = S(0.0), prhoe = S(0.0), prhophi = S(0.0), prhokin = S(0.0);
#pragma unroll
    for (int t = 0; t < NT; t++) {
        if (t == 0) {
            precip0 = precip0 - ft[t][0] / dt;
        } else {
            precip1 = precip1 - ft[t][0] / dt;
            lh = lh + ft[t][0] * C_LHF0 / dt;
        }
        prhoe = prhoe - ft[t][1] / dt;
        prhophi = prhophi - ft[t][2] / dt;
        prhokin = prhokin - ft[t][3] / dt - 0.0 / dt;   // the half-level flux at kmin is 0
    }
    s.precip0 = (0.0 + precip0) * 1.0;
    s.precip1 = (0.0 + precip1) * 1.0;
    s.lh = (0.0 + lh) * 1.0;
    s.prhoe = (0.0 + prhoe) * 1.0;
    s.prhophi = (0.0 + prhophi) * 1.0;
    s.prhokin = (0.0 + prhokin) * 1.0;
}

// ---------------------------------------------------------------------------------
// Sedimentation apply and KIN2EIN of interior cell c (src/mod_precip_transport.cpp:
// 960-1011) from the fluxes of walls c-1 .. c+2.
// ---------------------------------------------------------------------------------
template <typename S, int NT> struct SedInN {
    // fluxes: fq of walls c-1..c+2, fs of walls c..c+1, fh of walls c-1..c+1
    S fq[4][NT], fs[2][5], fh[3][2];
    S rw_c, rw_p;                     // pre-transport rhogw of cells c, c+1
    S rg_m, rg_c, rg_p;               // pre-transport rhog of cells c-1, c, c+1
    S rq[NT], re0;                    // the sedimenting tracers and rhoge of cell c
    S rvx, rvy, rvz;                  // momentum of cell c
    double gs_m, gs_c, gs_p, gsh_c, gsh_p, z;
};
template <typename S, int NT> struct SedOutN {
    S rg, e, rq[NT], rvx, rvy, rvz, rw;
};
template <typename S> using SedIn = SedInN<S, 3>;
template <typename S> using SedOut = SedOutN<S, 3>;

// KIN2KIN (src/mod_precip_transport.cpp:1012-1068), the alternative to KIN2EIN: the
// velocities after sedimentation are rescaled so that each point keeps the kinetic
// energy the transport carried into it.  Horizontal, at cell c in [kmin, kmax]:
// v = rhogv / rhog scaled by sqrt(|kin_h0 / kin_h|) with kin_h0 = rhogkin_h / rhog.
template <typename S> NSW6_HD void kin2kin_h(S rg, S kinh, S &rvx, S &rvy, S &rvz)
{
    const S kin_h0 = kinh / rg;
    S vx = rvx / rg, vy = rvy / rg, vz = rvz / rg;
    const S kin_h = 0.5 * (vx * vx + vy * vy + vz * vz);
    if (kin_h > 1.0E-20) {
        const S f = sqrt(fabs(kin_h0 / kin_h));
        vx = vx * f;
        vy = vy * f;
        vz = vz * f;
    }
    rvx = vx * rg;
    rvy = vy * rg;
    rvz = vz * rg;
}
// Vertical, at wall k in [kmin, kmax+1]: rhogh = C2W0 rhog(k) + C2W1 rhog(k-1) (after
// sedimentation), w = rhogw / rhogh scaled by sqrt(|kin_v0 / kin_v|), kin_v0 =
// rhogkin_v / rhogh, kin_v = w^2 / 2; returns the new rhogw.
template <typename S> NSW6_HD S kin2kin_v(S rhogh, S kinv, S rw)
{
    const S kin_v0 = kinv / rhogh;
    S w = rw / rhogh;
    const S kin_v = 0.5 * (w * w);
    if (kin_v > 1.0E-20) w = w * sqrt(fabs(kin_v0 / kin_v));
    return w * rhogh;
}

template <typename S, int NT>
NSW6_HD void sediment(const GridDev &g, int c, const Unit &u, const SedInN<S, NT> &x, SedOutN<S, NT> &y,
                      bool kin2kin = false)
{
    // tendencies of cell c (walls c and c+1; wall 95 carries zero fluxes)
    const double rc = g.rdgz[c];
    S dq[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) dq[t] = -(x.fq[2][t] - x.fq[1][t]) * rc;
    const S de = -(x.fs[1][0] - x.fs[0][0]) * rc;
    const S dphi = -(x.fs[1][1] - x.fs[0][1]) * rc;
    const S dkinh = -(x.fs[1][2] - x.fs[0][2]) * rc;
    const S du = -(x.fs[1][3] - x.fs[0][3]) * rc;
    const S dv = -(x.fs[1][4] - x.fs[0][4]) * rc;
    // mass tendencies of cells c+1 and c-1
    S dqp[NT], dqm[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) dqp[t] = dqm[t] = S(0.0);
    if (c + 1 <= KMAX) {
        const double r = g.rdgz[c + 1];
#pragma unroll
        for (int t = 0; t < NT; t++) dqp[t] = -(x.fq[3][t] - x.fq[2][t]) * r;
    }
    if (c - 1 >= KMIN) {
        const double r = g.rdgz[c - 1];
#pragma unroll
        for (int t = 0; t < NT; t++) dqm[t] = -(x.fq[1][t] - x.fq[0][t]) * r;
    }
    // half-level tendencies at walls c and c+1 (walls 2..94; the flux of wall 1 is 0)
    S hk_c = S(0.0), hw_c = S(0.0), hk_p = S(0.0), hw_p = S(0.0);
    if (c >= KMIN + 1) {
        const double r = g.rdgzh[c];
        hk_c = -(x.fh[1][0] - x.fh[0][0]) * r;
        hw_c = -(x.fh[1][1] - x.fh[0][1]) * r;
    }
    if (c + 1 <= KMAX) {
        const double r = g.rdgzh[c + 1];
        hk_p = -(x.fh[2][0] - x.fh[1][0]) * r;
        hw_p = -(x.fh[2][1] - x.fh[1][1]) * r;
    }
    // pre-transport wall quantities of cells c and c+1 (C2W factors, rhogkin_v)
    const double rgs_c = 1.0 / x.gs_c;
    const double c2w0_c = g.afact[c] * rgs_c * x.gsh_c, c2w1_c = g.bfact[c] / x.gs_m * x.gsh_c;
    const double c2w0_p = g.afact[c + 1] / x.gs_p * x.gsh_p, c2w1_p = g.bfact[c + 1] * rgs_c * x.gsh_p;
    const S kinv0_c = (c >= KMIN + 1) ? S(0.5 * (x.rw_c * x.rw_c) / (c2w0_c * x.rg_c + c2w1_c * x.rg_m)) : S(0.0);
    const S kinv0_p = (c + 1 <= KMAX) ? S(0.5 * (x.rw_p * x.rw_p) / (c2w0_p * x.rg_p + c2w1_p * x.rg_c)) : S(0.0);

    const double rgsh_p = x.gs_c / x.gsh_p;
    const double w2c0 = g.cfact[c] * rgsh_p, w2c1 = g.dfact[c] * rgsh_p;
    const S kinh = 0.5 * (x.rvx * x.rvx + x.rvy * x.rvy + x.rvz * x.rvz) / x.rg_c;
    S e = x.re0 + de + dphi;
    const S kinh_u = kinh + dkinh;
    const S kinv_c = kinv0_c + hk_c;
    const S kinv_cp = kinv0_p + hk_p;
    y.rvx = x.rvx + du * u.ix + dv * u.jx;
    y.rvy = x.rvy + du * u.iy + dv * u.jy;
    y.rvz = x.rvz + du * u.iz + dv * u.jz;
    y.rw = x.rw_c + hw_c;
    S rg = x.rg_c;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        y.rq[t] = x.rq[t] + dq[t];
        rg = rg + dq[t];
        e = e - dq[t] * C_GRAV * x.z;
    }
    // KIN2EIN: kinetic energy before minus after, with the neighbours' new densities
    const S tmp2 = kinh_u + (w2c0 * kinv_cp + w2c1 * kinv_c);
    S rg_cp = x.rg_p, rg_cm = x.rg_m;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        rg_cp = rg_cp + dqp[t];
        rg_cm = rg_cm + dqm[t];
    }
    const S rw_cp = x.rw_p + hw_p;
    const S kinh_new = 0.5 * (y.rvx * y.rvx + y.rvy * y.rvy + y.rvz * y.rvz) / rg;
    const S kinv_new_c = (c >= KMIN + 1) ? S(0.5 * (y.rw * y.rw) / (c2w0_c * rg + c2w1_c * rg_cm)) : S(0.0);
    const S kinv_new_cp = (c + 1 <= KMAX) ? S(0.5 * (rw_cp * rw_cp) / (c2w0_p * rg_cp + c2w1_p * rg)) : S(0.0);
    const S tmpk = kinh_new + (w2c0 * kinv_new_cp + w2c1 * kinv_new_c);
    if (kin2kin) {
        kin2kin_h(rg, kinh_u, y.rvx, y.rvy, y.rvz);
        y.rw = kin2kin_v(S(c2w0_c * rg + c2w1_c * rg_cm), kinv_c, y.rw);
    } else {
        e = e + (tmp2 - tmpk);
    }
    y.rg = rg;
    y.e = e;
}

// ---------------------------------------------------------------------------------
// Point-wise tail: rediagnosis (src/mod_precip_transport.cpp:1074-1086),
// SATURATION_adjustment with satadjust_liq (src/mod_satadjust.cpp:274-362, 771-896)
// and the final negative_filter (src/mod_mp_nsw6.cpp:671, 1770-1831).
// ---------------------------------------------------------------------------------
template <typename S> struct TailIn {
    S rhog, rhoge, rhogq[6];   // state after sedimentation
    double gsgam2;
};
template <typename S> struct TailOut {
    S rhog, rhoge, rho, tem, pre, qd, rhogq[6], q[6];
};

// Newton of satadjust_liq on values (src/mod_satadjust.cpp:828-894).  q[QV] and q[QC]
// are updated in place; returns 0 when no iteration ran, n > 0 after convergence in n
// iterations, -1 on a NaN temperature, -2 when itelim = 100 is reached.
NSW6_HD int satadjust_liq_newton(const DevParams &P, double &tem, double q[6], double qsum, double qdl,
                                 double rho_sa, double ein_moist)
{
    double rtem = 1.0 / dmax(tem, TEM_MIN);
    double psat = C_PSAT0 * 1.0 * exp(P.LovR_liq * (P.RTEM00 - rtem));
    double qsat = psat / (rho_sa * C_Rvap * tem);
    if (!(qsum - qsat > C_EPS)) return 0;
    for (int ite = 0; ite < 100; ite++) {
        rtem = 1.0 / dmax(tem, TEM_MIN);
        psat = C_PSAT0 * 1.0 * exp(P.LovR_liq * (P.RTEM00 - rtem));
        qsat = psat / (rho_sa * C_Rvap * tem);
        q[QV] = qsat;
        q[QC] = qsum - qsat;
        double CVt = qdl * C_CVdry;
#pragma unroll
        for (int n = 1; n <= 5; n++) CVt = CVt + q[n] * C_CVW;
        const double Emoist_new = tem * CVt + q[QV] * C_LHV0;
        const double dqsat_dT = (P.LovR_liq / (tem * tem)) * qsat;   // CVovR_liq = 0
        const double dCVtot_dT = dqsat_dT * (C_CVW - C_CVW);
        const double dEmoist_dT = tem * dCVtot_dT + CVt + dqsat_dT * C_LHV0;
        const double dtemp = (Emoist_new - ein_moist) / dEmoist_dT;
        tem = tem - dtemp;
        if (fabs(dtemp) < P.dtemp_criteria) return ite + 1;
        if (tem * 0.0 != 0.0) return -1;
    }
    return -2;
}

// satadjust_all (src/mod_satadjust.cpp:620-769): the mixed-phase Newton.  The liquid
// fraction alpha = clamp((T - LL) / (UL - LL), 0, 1) splits the condensate between
// QC and QI and blends the liquid and ice saturation pressures; E(T) carries the
// fusion heat of the ice part.  Same return codes as satadjust_liq_newton; q holds
// the last iterate's QV, QC, QI.
NSW6_HD int satadjust_all_newton(const DevParams &P, double &tem, double q[6], double qsum, double qdl,
                                 double rho_sa, double ein_moist)
{
    const double UL = P.sat_ulimit, LL = P.sat_llimit;
    double alpha = dmin(dmax((tem - LL) / (UL - LL), 0.0), 1.0);
    double rtem = 1.0 / dmax(tem, TEM_MIN);
    double psatl = C_PSAT0 * 1.0 * exp(P.LovR_liq * (P.RTEM00 - rtem));
    double psati = C_PSAT0 * 1.0 * exp(P.LovR_ice * (P.RTEM00 - rtem));
    double psat = psatl * alpha + psati * (1.0 - alpha);
    double qsat = psat / (rho_sa * C_Rvap * tem);
    if (!(qsum - qsat > C_EPS)) return 0;
    for (int ite = 0; ite < 100; ite++) {
        alpha = dmin(dmax((tem - LL) / (UL - LL), 0.0), 1.0);
        rtem = 1.0 / dmax(tem, TEM_MIN);
        psatl = C_PSAT0 * 1.0 * exp(P.LovR_liq * (P.RTEM00 - rtem));
        psati = C_PSAT0 * 1.0 * exp(P.LovR_ice * (P.RTEM00 - rtem));
        psat = psatl * alpha + psati * (1.0 - alpha);
        const double qsatl = psatl / (rho_sa * C_Rvap * tem);
        const double qsati = psati / (rho_sa * C_Rvap * tem);
        qsat = psat / (rho_sa * C_Rvap * tem);
        q[QV] = qsat;
        q[QC] = (qsum - qsat) * alpha;
        q[QI] = (qsum - qsat) * (1.0 - alpha);
        double CVt = qdl * C_CVdry;
#pragma unroll
        for (int n = 1; n <= 5; n++) CVt = CVt + q[n] * C_CVW;
        const double Emoist_new = tem * CVt + q[QV] * C_LHV0 - q[QI] * C_LHF0;
        const double lim1 = 0.5 + copysign(0.5, UL - tem);
        const double lim2 = 0.5 + copysign(0.5, tem - LL);
        const double dalpha_dT = lim1 * lim2 / (UL - LL);
        const double rtt = 1.0 / (tem * tem);   // CVovR_liq = CVovR_ice = 0
        const double dqsatl_dT = (P.LovR_liq * rtt) * qsatl;
        const double dqsati_dT = (P.LovR_ice * rtt) * qsati;
        const double dqsat_dT = qsatl * dalpha_dT + dqsatl_dT * alpha - qsati * dalpha_dT + dqsati_dT * (1.0 - alpha);
        const double dqc_dT = (qsum - qsat) * dalpha_dT - dqsat_dT * alpha;
        const double dqi_dT = -(qsum - qsat) * dalpha_dT - dqsat_dT * (1.0 - alpha);
        const double dCVtot_dT = dqsat_dT * C_CVW + dqc_dT * C_CVW + dqi_dT * C_CVW;
        const double dEmoist_dT = tem * dCVtot_dT + CVt + dqsat_dT * C_LHV0 - dqi_dT * C_LHF0;
        const double dtemp = (Emoist_new - ein_moist) / dEmoist_dT;
        tem = tem - dtemp;
        if (fabs(dtemp) < P.dtemp_criteria) return ite + 1;
        if (tem * 0.0 != 0.0) return -1;
    }
    return -2;
}

// Residual of the satadjust_liq solve, E(T) - Emoist, for the implicit-function
// derivative of the converged temperature: with q(QC) = qsum - qsat(T) and the
// reference's heat capacity sum, E(T) = T (qdl CVdry + (qsum - qsat + qrest) CVW) +
// qsat LHV0.
template <typename S>
NSW6_HD S satadjust_residual(const DevParams &P, S tem, S qsum, S qdl, S qrest, S rho_sa, S ein_moist)
{
    const S qsat = psat_of(P.LovR_liq, P.RTEM00, tem) / (rho_sa * C_Rvap * tem);
    const S CVt = qdl * C_CVdry + ((qsum - qsat) + qrest) * C_CVW;
    return tem * CVt + qsat * C_LHV0 - ein_moist;
}
// dE/dT of that residual at fixed inputs (exact, unlike the Newton's own slope)
NSW6_HD double satadjust_dRdT(const DevParams &P, double tem, double qsum, double qdl, double qrest, double rho_sa)
{
    const double psat = C_PSAT0 * 1.0 * exp(P.LovR_liq * (P.RTEM00 - 1.0 / dmax(tem, TEM_MIN)));
    const double qsat = psat / (rho_sa * C_Rvap * tem);
    const double dpsat = (tem > TEM_MIN) ? psat * P.LovR_liq / (tem * tem) : 0.0;
    const double dqsat = dpsat / (rho_sa * C_Rvap * tem) - qsat / tem;
    const double CVt = qdl * C_CVdry + ((qsum - qsat) + qrest) * C_CVW;
    return CVt - tem * dqsat * C_CVW + dqsat * C_LHV0;
}

// The mixed-phase saturation of the satadjust_all solve at temperature T: liquid
// fraction alpha(T) = clamp((T - LL) / (UL - LL), 0, 1) over SATURATION_Setrange's
// limits, and qsat = (alpha psat_liq + (1 - alpha) psat_ice) / (rho Rvap T).
template <typename S> NSW6_HD S satadjust_all_alpha(const DevParams &P, S tem)
{
    return dmin(dmax((tem - P.sat_llimit) / (P.sat_ulimit - P.sat_llimit), S(0.0)), S(1.0));
}
template <typename S> NSW6_HD S satadjust_all_qsat(const DevParams &P, S tem, S rho_sa)
{
    const S alpha = satadjust_all_alpha(P, tem);
    const S psat = psat_of(P.LovR_liq, P.RTEM00, tem) * alpha + psat_of(P.LovR_ice, P.RTEM00, tem) * (1.0 - alpha);
    return psat / (rho_sa * C_Rvap * tem);
}
// Residual E_all(T) - Emoist of the satadjust_all solve: with q(QV) = qsat(T),
// q(QC) = (qsum - qsat) alpha, q(QI) = (qsum - qsat)(1 - alpha) and the reference's
// heat capacity sum, E_all(T) = T (qdl CVdry + (qsum - qsat + qrest) CVW) + qsat LHV0
// - q(QI) LHF0 (qrest = qr + qs + qg).
template <typename S>
NSW6_HD S satadjust_all_residual(const DevParams &P, S tem, S qsum, S qdl, S qrest, S rho_sa, S ein_moist)
{
    const S alpha = satadjust_all_alpha(P, tem);
    const S qsat = satadjust_all_qsat(P, tem, rho_sa);
    const S d = qsum - qsat;
    const S CVt = qdl * C_CVdry + (d + qrest) * C_CVW;
    return tem * CVt + qsat * C_LHV0 - d * (1.0 - alpha) * C_LHF0 - ein_moist;
}
// alpha, qsat and their temperature derivatives at fixed rho (the clamp's derivative
// is taken on the side its selections take)
struct SatAllSlopes {
    double alpha, dalpha, qsat, dqsat;
};
NSW6_HD SatAllSlopes satadjust_all_slopes(const DevParams &P, double tem, double rho_sa)
{
    SatAllSlopes s;
    const double a = (tem - P.sat_llimit) / (P.sat_ulimit - P.sat_llimit);
    s.alpha = dmin(dmax(a, 0.0), 1.0);
    s.dalpha = (a < 0.0 || 1.0 < a) ? 0.0 : 1.0 / (P.sat_ulimit - P.sat_llimit);
    const double psl = psat_of(P.LovR_liq, P.RTEM00, tem), psi = psat_of(P.LovR_ice, P.RTEM00, tem);
    const double dpsl = (tem > TEM_MIN) ? psl * P.LovR_liq / (tem * tem) : 0.0;
    const double dpsi = (tem > TEM_MIN) ? psi * P.LovR_ice / (tem * tem) : 0.0;
    const double psat = psl * s.alpha + psi * (1.0 - s.alpha);
    const double dpsat = dpsl * s.alpha + dpsi * (1.0 - s.alpha) + (psl - psi) * s.dalpha;
    const double den = rho_sa * C_Rvap * tem;
    s.qsat = psat / den;
    s.dqsat = dpsat / den - s.qsat / tem;
    return s;
}
// dE_all/dT of that residual at fixed inputs
NSW6_HD double satadjust_all_dRdT(const DevParams &P, double tem, double qsum, double qdl, double qrest,
                                  double rho_sa)
{
    const SatAllSlopes s = satadjust_all_slopes(P, tem, rho_sa);
    const double d = qsum - s.qsat;
    const double CVt = qdl * C_CVdry + (d + qrest) * C_CVW;
    return CVt - tem * s.dqsat * C_CVW + s.dqsat * C_LHV0 + (s.dqsat * (1.0 - s.alpha) + d * s.dalpha) * C_LHF0;
}

// Status of one point's tail (for the device report).
struct TailStatus {
    int newton;   // satadjust_liq_newton's return value
};

// The converged temperature and q(QV) = qsat of the satadjust_liq solve as values of
// the scalar type.  For double they are the Newton's own last iterates; a dual-number
// scalar overloads both (found by argument-dependent lookup) to carry the
// implicit-function tangent of the solution.
NSW6_HD double newton_tem(const DevParams &, double tem_final, double, double, double, double, double)
{
    return tem_final;
}
NSW6_HD double newton_qsat(const DevParams &, double qv_final, double, double) { return qv_final; }
// The same for the satadjust_all solve: its temperature, and q(QV), q(QC), q(QI).
NSW6_HD double newton_tem_all(const DevParams &, double tem_final, double, double, double, double, double)
{
    return tem_final;
}
NSW6_HD void newton_q_all(const DevParams &, const double q_final[6], double, double, double, double &qv,
                          double &qc, double &qi)
{
    qv = q_final[QV];
    qc = q_final[QC];
    qi = q_final[QI];
}

template <typename S>
NSW6_HD TailStatus tail(const DevParams &P, bool interior, const TailIn<S> &x, TailOut<S> &y)
{
    TailStatus st{0};
    S rhog = x.rhog, rhoge = x.rhoge, rhogq[6];
#pragma unroll
    for (int n = 0; n < 6; n++) rhogq[n] = x.rhogq[n];
    const double gsgam2 = x.gsgam2;
    S q[6];
    const S rrhog = S(1.0) / rhog;
    const double rgsgam2 = 1.0 / gsgam2;
#pragma unroll
    for (int n = 0; n < 6; n++) q[n] = rhogq[n] * rrhog;
    S rho = rhog * rgsgam2;
    const S ein = rhoge * rrhog;
    S qd = S(1.0);
#pragma unroll
    for (int n = 1; n <= 5; n++) qd = qd - q[n];
    S cv = S(0.0), qd2 = S(1.0);
#pragma unroll
    for (int n = 1; n <= 5; n++) {
        cv = cv + q[n] * C_CVW;
        qd2 = qd2 - q[n];
    }
    cv = cv + qd2 * C_CVdry;
    S tem = ein / cv;
    S pre = rho * tem * (qd2 * C_Rdry + q[QV] * C_Rvap);

    // SATURATION_adjustment (src/mod_satadjust.cpp:274-362): satadjust_liq, or
    // satadjust_all when explicit ice generation is off
    const bool ice = P.ice_adjust != 0.0;
    const S ein_moist = ice ? S(ein + q[QV] * C_LHV0 - q[QI] * C_LHF0) : S(ein + q[QV] * C_LHV0);
    const S qsum = ice ? S(q[QV] + q[QC] + q[QI]) : S(q[QV] + q[QC]);
    q[QV] = qsum;
    q[QC] = S(0.0);
    if (ice) q[QI] = S(0.0);
    S CVtot = qd * C_CVdry;
#pragma unroll
    for (int n = 1; n <= 5; n++) CVtot = CVtot + q[n] * C_CVW;
    const S rho_sa = rhog * rgsgam2;
    tem = (ein_moist - q[QV] * C_LHV0) / CVtot;
    if (interior) {
        S qdl = S(1.0);
#pragma unroll
        for (int n = 1; n <= 5; n++) qdl = qdl - q[n];
        double tv = val(tem), qv6[6];
#pragma unroll
        for (int n = 0; n < 6; n++) qv6[n] = val(q[n]);
        if (ice) {
            st.newton = satadjust_all_newton(P, tv, qv6, val(qsum), val(qdl), val(rho_sa), val(ein_moist));
            if (st.newton != 0) {
                const S qrest = (q[QR] + q[QS]) + q[QG];
                tem = newton_tem_all(P, tv, qsum, qdl, qrest, rho_sa, ein_moist);
                newton_q_all(P, qv6, tem, rho_sa, qsum, q[QV], q[QC], q[QI]);
            }
        } else {
            st.newton = satadjust_liq_newton(P, tv, qv6, val(qsum), val(qdl), val(rho_sa), val(ein_moist));
            if (st.newton != 0) {
                const S qrest = ((q[QR] + q[QI]) + q[QS]) + q[QG];
                tem = newton_tem(P, tv, qsum, qdl, qrest, rho_sa, ein_moist);
                q[QV] = newton_qsat(P, qv6[QV], tem, rho_sa);
                q[QC] = qsum - q[QV];
            }
        }
    }
    rhogq[QV] = rhog * q[QV];
    rhogq[QC] = rhog * q[QC];
    if (ice) rhogq[QI] = rhog * q[QI];
    S CVtot2 = qd * C_CVdry;
#pragma unroll
    for (int n = 1; n <= 5; n++) CVtot2 = CVtot2 + q[n] * C_CVW;
    rhoge = rhog * tem * CVtot2;

    if (interior) {
        // final negative_filter
        S diffq = S(0.0);
#pragma unroll
        for (int n = 2; n <= 5; n++) {
            diffq += rhogq[n];
            rhogq[n] = dmax(rhogq[n], S(0.0));
        }
#pragma unroll
        for (int n = 2; n <= 5; n++) diffq -= rhogq[n];
        rhogq[QV] += diffq;
        diffq = rhogq[QV];
        rhogq[QV] = dmax(rhogq[QV], S(0.0));
        diffq -= rhogq[QV];
        rhog = rhog * (1.0 - diffq);
        rho = rhog * rgsgam2;
        const S rrhog2 = S(1.0) / rhog;
#pragma unroll
        for (int n = 1; n <= 5; n++) q[n] = rhogq[n] * rrhog2;
        S qdn = S(1.0);
#pragma unroll
        for (int n = 1; n <= 5; n++) qdn = qdn - q[n];
        S cva = qdn * C_CVdry;
#pragma unroll
        for (int n = 1; n <= 5; n++) cva = cva + q[n] * C_CVW;
        rhoge = tem * rhog * cva;
        pre = rho * (qdn * C_Rdry + q[QV] * C_Rvap) * tem;
    }
    y.rhog = rhog;
    y.rhoge = rhoge;
    y.rho = rho;
    y.tem = tem;
    y.pre = pre;
    y.qd = qd;
#pragma unroll
    for (int n = 0; n < 6; n++) {
        y.rhogq[n] = rhogq[n];
        y.q[n] = q[n];
    }
    return st;
}

}  // namespace nsw6
