// nsw6_adjoint.h -- hand-written reverse mode of the FP64 NSW6 step, piece by piece.
//
// Every function here takes the inputs of one forward piece of nsw6_step.h, recomputes
// that piece's intermediate values and propagates output cotangents ("_b") back to
// input cotangents, accumulating (+=) into the caller's arrays.  Discrete decisions
// (limiter selections, switches, departure cells, |f| < EPS, the Newton's iteration
// count) are taken from the recomputed forward values, so the result is the gradient
// of the piecewise-smooth step.  The saturation adjustment is differentiated through
// the implicit-function theorem at the Newton solution.
//
// Coverage: Roh & Satoh 2014 or exponential snow moments, KK2000 or Berry's
// autoconversion, with or without explicit ice generation (Bergeron growth and the
// mixed-phase saturation adjustment), OPT_INDIR, KIN2EIN or KIN2KIN, "3WATER"
// transport with single-cell and general regions; the host rejects other
// configurations before launching the adjoint.
#pragma once

#include "nsw6_step.h"

namespace nsw6 {

// ---------------------------------------------------------------------------------
// Departure distances (reverse of zdis_interior / zdis_bottom / wh_full / wh_half)
// ---------------------------------------------------------------------------------
NSW6_HD void zdis_interior_b(double dt, double dt2, double dt3_6, double whm, double whk, double whp, double rdzm,
                             double rdzk, double rs, double z_b, double &whm_b, double &whk_b, double &whp_b)
{
    const double d = whp - whm;
    const double g = d * rs;
    const double A = whp - whk, B = whk - whm;
    const double inner = (A * rdzk - B * rdzm) * rs * 2.0;
    const double X = g * g + whk * inner;
    const double t1_b = z_b, t2_b = -z_b, t3_b = z_b;
    whk_b += t3_b * dt3_6 * X;
    const double X_b = t3_b * dt3_6 * whk;
    const double g_b = X_b * 2.0 * g;
    whk_b += X_b * inner;
    const double inner_b = X_b * whk;
    const double A_b = inner_b * rs * 2.0 * rdzk, B_b = -inner_b * rs * 2.0 * rdzm;
    whp_b += A_b;
    whk_b += B_b - A_b;
    whm_b -= B_b;
    double d_b = g_b * rs;
    whk_b += t2_b * dt2 * d * rs * 0.5;
    d_b += t2_b * dt2 * whk * rs * 0.5;
    whk_b += t1_b * dt;
    whp_b += d_b;
    whm_b -= d_b;
}
NSW6_HD void zdis_bottom_b(double dt, double dt2, double whk, double whp, double rdzk, double z_b, double &whk_b,
                           double &whp_b)
{
    whk_b += z_b * (dt - dt2 * (whp - whk) * rdzk * 0.5 + dt2 * whk * rdzk * 0.5);
    whp_b += -z_b * dt2 * whk * rdzk * 0.5;
}
NSW6_HD void wh_full_b(int k, int d, double b, double V_b[5])
{
    const int j = k + d;
    if (j == 0) return;
    if (j == 1) {
        V_b[d + 3] += b;
    } else if (j <= KMAX) {
        V_b[d + 2] += 0.5 * b;
        V_b[d + 3] += 0.5 * b;
    } else {
        V_b[d + 2] += b;
    }
}
NSW6_HD void vh_half_b(int k, int d, double b, double V_b[5])
{
    const int j = k + d;
    if (j >= 2 && j <= KMAX - 1) {
        V_b[d + 3] += 0.5 * b;
        V_b[d + 2] += 0.5 * b;
    }
}
NSW6_HD void wh_half_b(int k, int d, double b, double V_b[5])
{
    const int j = k + d;
    if (j >= 3 && j <= KMAX) {
        vh_half_b(k, d - 1, 0.5 * b, V_b);
        vh_half_b(k, d, 0.5 * b, V_b);
    } else if (j == 2) {
        vh_half_b(k, d, b, V_b);
    }
}
NSW6_HD void zdis_full_b(const GridDev &g, double dt, double dt2, double dt3_6, const double V[5], int k,
                         double z_b, double V_b[5])
{
    const double whp = wh_full(V, k, 1);
    double whm_b = 0.0, whk_b = 0.0, whp_b = 0.0;
    if (k == KMIN) {
        zdis_bottom_b(dt, dt2, wh_full(V, k, 0), whp, g.rdgz[KMIN], z_b, whk_b, whp_b);
    } else {
        zdis_interior_b(dt, dt2, dt3_6, wh_full(V, k, -1), wh_full(V, k, 0), whp, g.rdgz[k - 1], g.rdgz[k],
                        g.rsum_gz[k], z_b, whm_b, whk_b, whp_b);
        wh_full_b(k, -1, whm_b, V_b);
    }
    wh_full_b(k, 0, whk_b, V_b);
    wh_full_b(k, 1, whp_b, V_b);
}
NSW6_HD void zdis_half_b(const GridDev &g, double dt, double dt2, double dt3_6, const double V[5], int k,
                         double z_b, double V_b[5])
{
    const double whp = wh_half(V, k, 1);
    double whm_b = 0.0, whk_b = 0.0, whp_b = 0.0;
    if (k == KMIN + 1) {
        zdis_bottom_b(dt, dt2, wh_half(V, k, 0), whp, g.rdgzh[KMIN + 1], z_b, whk_b, whp_b);
    } else {
        zdis_interior_b(dt, dt2, dt3_6, wh_half(V, k, -1), wh_half(V, k, 0), whp, g.rdgzh[k - 1], g.rdgzh[k],
                        g.rsum_gzh[k], z_b, whm_b, whk_b, whp_b);
        wh_half_b(k, -1, whm_b, V_b);
    }
    wh_half_b(k, 0, whk_b, V_b);
    wh_half_b(k, 1, whp_b, V_b);
}

// ---------------------------------------------------------------------------------
// Transport cells and wall fluxes
// ---------------------------------------------------------------------------------
template <int NT> struct CellBN {   // cotangent of a CellN
    double rhoq[NT], qrgs[NT], q[NT], tem, kinh, up, vp, hv[NT][2], rhog, kinv;
};
using CellB = CellBN<3>;
struct CellInB {   // cotangent of a CellIn (differentiable members)
    double rhog, rq[5], rhoge, rvx, rvy, rvz, rw, vx, vy, vz;
};
// Zeroing by elements (a memset of the whole struct would keep it in scratch memory).
template <typename T> NSW6_HD void zero_words(T &v)
{
    static_assert(sizeof(T) % sizeof(double) == 0, "struct of doubles");
    double *w = reinterpret_cast<double *>(&v);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / sizeof(double)); i++) w[i] = 0.0;
}
template <int NT> NSW6_HD void zero(CellBN<NT> &c) { zero_words(c); }
NSW6_HD void zero(CellInB &c) { zero_words(c); }

template <int NT> NSW6_HD void rhof_b(const CellN<double, NT> &c, int t, int i, double b, CellBN<NT> &cb)
{
    switch (i) {
    case 0: cb.rhoq[t] += b; break;
    case 1:
        cb.rhoq[t] += b * C_CVW * c.tem;
        cb.tem += b * c.rhoq[t] * C_CVW;
        break;
    case 2: cb.rhoq[t] += b * C_GRAV * c.z; break;
    case 3:
        cb.qrgs[t] += b * c.kinh;
        cb.kinh += b * c.qrgs[t];
        break;
    case 4:
        cb.rhoq[t] += b * c.up;
        cb.up += b * c.rhoq[t];
        break;
    default:
        cb.rhoq[t] += b * c.vp;
        cb.vp += b * c.rhoq[t];
        break;
    }
}

// Reverse of wall_flux for tracer T (a compile-time index, so that the Cell
// cotangents stay in registers).
template <int T, int NT>
NSW6_HD void wall_flux_b_tracer(const GridDev &g, int w, const CellN<double, NT> &cur, const CellN<double, NT> &nxt,
                                const double z0[NT], const double zh0[NT], const WallGeomN<NT> &geo,
                                const WallFluxN<double, NT> &Fb, const double ft_surf_b[NT][4], CellBN<NT> &cur_b,
                                CellBN<NT> &nxt_b, double z0_b[NT], double zh0_b[NT])
{
    const double dzw = g.dgz[w];
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const double r0 = rhof_cell(cur, T, i);
        const double rn = rhof_cell(nxt, T, i);
        const double v = geo.up1[T] ? (-(r0 * dzw) + rn * z0[T]) : r0 * z0[T];
        double ft_b = (i == 0) ? Fb.fq[T] : Fb.fs[i - 1];
        if (i < 4) ft_b += ft_surf_b[T][i];
        const double v_b = ft_b * eps_mask(v);
        // both cells take a (possibly zero) share: one call per cell keeps each
        // cotangent struct in its own registers
        const bool up = geo.up1[T];
        rhof_b(cur, T, i, up ? -v_b * dzw : v_b * z0[T], cur_b);
        rhof_b(nxt, T, i, up ? v_b * z0[T] : 0.0, nxt_b);
        z0_b[T] += v_b * (up ? rn : r0);
    }
    if (w >= KMIN + 1) {
        const double dzh = g.dgzh[w];
#pragma unroll
        for (int i = 0; i < 2; i++) {
            const double hc = cur.hv[T][i], hn = nxt.hv[T][i];
            double v;
            if (geo.half_full[T])
                v = -(hc * dzh) + hn * zh0[T];
            else
                v = (geo.half_next[T] ? hn : hc) * zh0[T];
            const double v_b = Fb.fh[i] * eps_mask(v);
            const bool full = geo.half_full[T], next = geo.half_next[T];
            cur_b.hv[T][i] += full ? -v_b * dzh : (next ? 0.0 : v_b * zh0[T]);
            nxt_b.hv[T][i] += (full || next) ? v_b * zh0[T] : 0.0;
            zh0_b[T] += v_b * ((full || next) ? hn : hc);
        }
    }
}

// Reverse of wall_flux (the departure geometry is given).
template <int NT>
NSW6_HD void wall_flux_b(const GridDev &g, int w, const CellN<double, NT> &cur, const CellN<double, NT> &nxt,
                         const double z0[NT], const double zh0[NT], const WallGeomN<NT> &geo,
                         const WallFluxN<double, NT> &Fb, const double ft_surf_b[NT][4], CellBN<NT> &cur_b,
                         CellBN<NT> &nxt_b, double z0_b[NT], double zh0_b[NT])
{
    wall_flux_b_tracer<0, NT>(g, w, cur, nxt, z0, zh0, geo, Fb, ft_surf_b, cur_b, nxt_b, z0_b, zh0_b);
    wall_flux_b_tracer<1, NT>(g, w, cur, nxt, z0, zh0, geo, Fb, ft_surf_b, cur_b, nxt_b, z0_b, zh0_b);
    wall_flux_b_tracer<2, NT>(g, w, cur, nxt, z0, zh0, geo, Fb, ft_surf_b, cur_b, nxt_b, z0_b, zh0_b);
    if constexpr (NT == 4)
        wall_flux_b_tracer<3, NT>(g, w, cur, nxt, z0, zh0, geo, Fb, ft_surf_b, cur_b, nxt_b, z0_b, zh0_b);
}

// Reverse of derive_cell: Cell cotangent -> raw cell cotangent and the below-cell
// cotangents (rhog and q of cell j-1).
template <int NT>
NSW6_HD void derive_cell_b(const GridDev &g, int j, const CellIn<double> &r, double below_rhog,
                           const double below_q[NT], double below_gs, const Unit &u, const CellN<double, NT> &c,
                           const CellBN<NT> &cb, CellInB &rb, double &below_rhog_b, double below_q_b[NT])
{
    const bool interior = (j >= KMIN && j <= KMAX);
    const double rg = r.rhog;
    const double rgs = interior ? r.gam2 / r.gs : 0.0;
    double q_b[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) q_b[t] = cb.q[t];
    double rg_b = cb.rhog;
    if (j >= KMIN + 1 && j <= KMAX) {
        const double c2w0 = g.afact[j] / r.gs * r.gsh;
        const double c2w1 = g.bfact[j] / below_gs * r.gsh;
        const double rgsh = r.gam2h / r.gsh;
        double kinv_b = cb.kinv;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const double qh = 0.5 * (c.q[t] + below_q[t]) * rgsh;
            const double qh_b = cb.hv[t][0] * c.kinv + cb.hv[t][1] * r.rw;
            kinv_b += cb.hv[t][0] * qh;
            rb.rw += cb.hv[t][1] * qh;
            q_b[t] += qh_b * 0.5 * rgsh;
            below_q_b[t] += qh_b * 0.5 * rgsh;
        }
        const double den = c2w0 * rg + c2w1 * below_rhog;
        rb.rw += kinv_b * r.rw / den;
        const double den_b = -kinv_b * c.kinv / den;
        rg_b += den_b * c2w0;
        below_rhog_b += den_b * c2w1;
    }
    double qq_b[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < NT; t++) q_b[t] += cb.qrgs[t] * rgs;
    // tracer t is rq[sed_nq(t) - 1] (rq holds QC..QG)
#pragma unroll
    for (int t = 0; t < NT; t++) {
        rb.rq[sed_nq<NT>(t) - 1] += cb.rhoq[t] * rgs;
        qq_b[sed_nq<NT>(t) - 1] += q_b[t];
    }
    rb.vx += cb.up * u.ix + cb.vp * u.jx;
    rb.vy += cb.up * u.iy + cb.vp * u.jy;
    rb.vz += cb.up * u.iz + cb.vp * u.jz;
    if (interior) {
        rb.rvx += cb.kinh * r.rvx / rg;
        rb.rvy += cb.kinh * r.rvy / rg;
        rb.rvz += cb.kinh * r.rvz / rg;
        rg_b -= cb.kinh * c.kinh / rg;
        // tem = rhoge / (rg cva); cva does not depend on q (CVW = CVdry)
        double qd = 1.0;
#pragma unroll
        for (int n = 0; n < 5; n++) qd = qd - r.rq[n] / rg;
        double cva = qd * C_CVdry;
#pragma unroll
        for (int n = 0; n < 5; n++) cva = cva + r.rq[n] / rg * C_CVW;
        rb.rhoge += cb.tem / (rg * cva);
        rg_b -= cb.tem * c.tem / rg;
    }
#pragma unroll
    for (int n = 0; n < 5; n++) {
        rb.rq[n] += qq_b[n] / rg;
        rg_b -= qq_b[n] * (r.rq[n] / rg) / rg;
    }
    rb.rhog += rg_b;
}

// ---------------------------------------------------------------------------------
// Divergence and KIN2EIN of one cell (reverse of sediment)
// ---------------------------------------------------------------------------------
template <int NT> struct SedInBN {
    double fq[4][NT], fs[2][5], fh[3][2];
    double rw_c, rw_p, rg_m, rg_c, rg_p, rq[NT], re0, rvx, rvy, rvz;
};
using SedInB = SedInBN<3>;
template <int NT> NSW6_HD void zero(SedInBN<NT> &s) { zero_words(s); }

// Reverse of kin2kin_v (nsw6_step.h): out = w f rhogh with w = rw / rhogh,
// f = sqrt(|kin_v0 / kin_v|), kin_v0 = kinv / rhogh, kin_v = w^2 / 2 (f = 1 when kin_v
// is at most 1e-20).  Adds the cotangents of rhogh, kinv and rw.
NSW6_HD void kin2kin_v_b(double rhogh, double kinv, double rw, double out_b, double &rhogh_b, double &kinv_b,
                         double &rw_b)
{
    const double k0 = kinv / rhogh;
    const double w = rw / rhogh;
    const double kv = 0.5 * (w * w);
    double f = 1.0, w_b;
    // out = w' rhogh, w' = w f
    if (kv > 1.0E-20) {
        const double r = k0 / kv;
        f = sqrt(fabs(r));
        rhogh_b += out_b * (w * f);
        const double wp_b = out_b * rhogh;
        w_b = wp_b * f;
        const double f_b = wp_b * w;
        const double r_b = (f > 0.0) ? f_b * copysign(1.0, r) / (2.0 * f) : 0.0;
        const double k0_b = r_b / kv;
        const double kv_b = -r_b * r / kv;
        w_b += kv_b * w;
        kinv_b += k0_b / rhogh;
        rhogh_b -= k0_b * k0 / rhogh;
    } else {
        rhogh_b += out_b * w;
        w_b = out_b * rhogh;
    }
    rw_b += w_b / rhogh;
    rhogh_b -= w_b * w / rhogh;
}

// Reverse of kin2kin_h (nsw6_step.h): out_i = v_i f rg with v_i = rv_i / rg,
// f = sqrt(|kin_h0 / kin_h|), kin_h0 = kinh / rg, kin_h = |v|^2 / 2 (f = 1 when kin_h is
// at most 1e-20).  rv holds the momenta before the scaling; adds the cotangents of
// rg, kinh and rv.
NSW6_HD void kin2kin_h_b(double rg, double kinh, const double rv[3], const double out_b[3], double &rg_b,
                         double &kinh_b, double rv_b[3])
{
    const double k0 = kinh / rg;
    double v[3], v_b[3];
#pragma unroll
    for (int i = 0; i < 3; i++) v[i] = rv[i] / rg;
    const double kh = 0.5 * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    double f = 1.0;
    const bool scaled = kh > 1.0E-20;
    if (scaled) f = sqrt(fabs(k0 / kh));
    double f_b = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        rg_b += out_b[i] * (v[i] * f);
        const double vp_b = out_b[i] * rg;
        v_b[i] = vp_b * f;
        f_b += vp_b * v[i];
    }
    if (scaled) {
        const double r = k0 / kh;
        const double r_b = (f > 0.0) ? f_b * copysign(1.0, r) / (2.0 * f) : 0.0;
        const double k0_b = r_b / kh;
        const double kh_b = -r_b * r / kh;
#pragma unroll
        for (int i = 0; i < 3; i++) v_b[i] += kh_b * v[i];
        kinh_b += k0_b / rg;
        rg_b -= k0_b * k0 / rg;
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        rv_b[i] += v_b[i] / rg;
        rg_b -= v_b[i] * v[i] / rg;
    }
}

// Reverse of sediment for NT sedimenting tracers, with the KIN2KIN energy correction
// (kin2kin) or the default KIN2EIN one.
template <int NT>
NSW6_HD void sediment_b(const GridDev &g, int c, const Unit &u, const SedInN<double, NT> &x,
                        const SedOutN<double, NT> &yb, SedInBN<NT> &xb, bool kin2kin = false)
{
    // forward recomputation
    const double rc = g.rdgz[c];
    double dq[NT], dqp[NT], dqm[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        dq[t] = -(x.fq[2][t] - x.fq[1][t]) * rc;
        dqp[t] = dqm[t] = 0.0;
    }
    const double du = -(x.fs[1][3] - x.fs[0][3]) * rc;
    const double dv = -(x.fs[1][4] - x.fs[0][4]) * rc;
    const bool has_p = (c + 1 <= KMAX), has_m = (c - 1 >= KMIN), has_hc = (c >= KMIN + 1);
    if (has_p)
#pragma unroll
        for (int t = 0; t < NT; t++) dqp[t] = -(x.fq[3][t] - x.fq[2][t]) * g.rdgz[c + 1];
    if (has_m)
#pragma unroll
        for (int t = 0; t < NT; t++) dqm[t] = -(x.fq[1][t] - x.fq[0][t]) * g.rdgz[c - 1];
    const double hw_c = has_hc ? -(x.fh[1][1] - x.fh[0][1]) * g.rdgzh[c] : 0.0;
    const double hw_p = has_p ? -(x.fh[2][1] - x.fh[1][1]) * g.rdgzh[c + 1] : 0.0;
    const double c2w0_c = g.afact[c] / x.gs_c * x.gsh_c, c2w1_c = g.bfact[c] / x.gs_m * x.gsh_c;
    const double c2w0_p = g.afact[c + 1] / x.gs_p * x.gsh_p, c2w1_p = g.bfact[c + 1] / x.gs_c * x.gsh_p;
    const double den_c = c2w0_c * x.rg_c + c2w1_c * x.rg_m;
    const double den_p = c2w0_p * x.rg_p + c2w1_p * x.rg_c;
    const double kinv0_c = has_hc ? 0.5 * (x.rw_c * x.rw_c) / den_c : 0.0;
    const double kinv0_p = has_p ? 0.5 * (x.rw_p * x.rw_p) / den_p : 0.0;
    const double w2c0 = g.cfact[c] * x.gs_c / x.gsh_p, w2c1 = g.dfact[c] * x.gs_c / x.gsh_p;
    const double kinh = 0.5 * (x.rvx * x.rvx + x.rvy * x.rvy + x.rvz * x.rvz) / x.rg_c;
    const double rvx_n = x.rvx + du * u.ix + dv * u.jx;
    const double rvy_n = x.rvy + du * u.iy + dv * u.jy;
    const double rvz_n = x.rvz + du * u.iz + dv * u.jz;
    const double rw_n = x.rw_c + hw_c;
    double rg = x.rg_c, rg_cp = x.rg_p, rg_cm = x.rg_m;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        rg = rg + dq[t];
        rg_cp = rg_cp + dqp[t];
        rg_cm = rg_cm + dqm[t];
    }
    const double rw_cp = x.rw_p + hw_p;
    const double kinh_new = 0.5 * (rvx_n * rvx_n + rvy_n * rvy_n + rvz_n * rvz_n) / rg;
    const double den_nc = c2w0_c * rg + c2w1_c * rg_cm;
    const double den_np = c2w0_p * rg_cp + c2w1_p * rg;
    const double kinv_new_c = has_hc ? 0.5 * (rw_n * rw_n) / den_nc : 0.0;
    const double kinv_new_cp = has_p ? 0.5 * (rw_cp * rw_cp) / den_np : 0.0;

    // reverse
    const double e_b = yb.e;
    // KIN2EIN: e += tmp2 - tmpk; KIN2KIN: the momenta are rescaled instead
    const double tmp2_b = kin2kin ? 0.0 : e_b, tmpk_b = kin2kin ? 0.0 : -e_b, e1_b = e_b;
    const double kinh_new_b = tmpk_b;
    const double kinv_new_cp_b = tmpk_b * w2c0, kinv_new_c_b = tmpk_b * w2c1;
    double rg_b = yb.rg, rwn_b = yb.rw, rg_cm_b = 0.0, rg_cp_b = 0.0, rw_cp_b = 0.0;
    double rvxn_b = yb.rvx, rvyn_b = yb.rvy, rvzn_b = yb.rvz;
    double k2k_kinh_u_b = 0.0, k2k_kinv_c_b = 0.0;
    if (kin2kin) {
        // y.rw = kin2kin_v(c2w0_c rg + c2w1_c rg_cm, kinv_c, rw_n)
        const double kinh = 0.5 * (x.rvx * x.rvx + x.rvy * x.rvy + x.rvz * x.rvz) / x.rg_c;
        const double dkinh = -(x.fs[1][2] - x.fs[0][2]) * rc;
        const double hk_c = has_hc ? -(x.fh[1][0] - x.fh[0][0]) * g.rdgzh[c] : 0.0;
        const double kinv_c = kinv0_c + hk_c;
        double rhogh_b = 0.0, rw_b = 0.0;
        kin2kin_v_b(c2w0_c * rg + c2w1_c * rg_cm, kinv_c, rw_n, yb.rw, rhogh_b, k2k_kinv_c_b, rw_b);
        rwn_b = rw_b;
        rg_b += rhogh_b * c2w0_c;
        rg_cm_b += rhogh_b * c2w1_c;
        // (y.rvx, y.rvy, y.rvz) = kin2kin_h(rg, kinh_u, rv_n)
        const double rvn[3] = {rvx_n, rvy_n, rvz_n}, ob[3] = {yb.rvx, yb.rvy, yb.rvz};
        double rvn_b[3] = {0.0, 0.0, 0.0};
        kin2kin_h_b(rg, kinh + dkinh, rvn, ob, rg_b, k2k_kinh_u_b, rvn_b);
        rvxn_b = rvn_b[0];
        rvyn_b = rvn_b[1];
        rvzn_b = rvn_b[2];
    }
    if (has_hc) {
        rwn_b += kinv_new_c_b * rw_n / den_nc;
        const double d_b = -kinv_new_c_b * kinv_new_c / den_nc;
        rg_b += d_b * c2w0_c;
        rg_cm_b += d_b * c2w1_c;
    }
    if (has_p) {
        rw_cp_b += kinv_new_cp_b * rw_cp / den_np;
        const double d_b = -kinv_new_cp_b * kinv_new_cp / den_np;
        rg_cp_b += d_b * c2w0_p;
        rg_b += d_b * c2w1_p;
    }
    rvxn_b += kinh_new_b * rvx_n / rg;
    rvyn_b += kinh_new_b * rvy_n / rg;
    rvzn_b += kinh_new_b * rvz_n / rg;
    rg_b -= kinh_new_b * kinh_new / rg;
    double hw_p_b = 0.0, hw_c_b = 0.0, hk_p_b = 0.0, hk_c_b = 0.0;
    xb.rw_p += rw_cp_b;
    hw_p_b += rw_cp_b;
    double dqm_b[NT], dqp_b[NT], dq_b[NT];
    xb.rg_m += rg_cm_b;
    xb.rg_p += rg_cp_b;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        dqm_b[t] = rg_cm_b;
        dqp_b[t] = rg_cp_b;
    }
    const double kinh_u_b = tmp2_b + k2k_kinh_u_b, kinv_cp_b = tmp2_b * w2c0, kinv_c_b = tmp2_b * w2c1 + k2k_kinv_c_b;
    xb.rg_c += rg_b;
#pragma unroll
    for (int t = 0; t < NT; t++) {
        dq_b[t] = rg_b + yb.rq[t] - e1_b * C_GRAV * x.z;
        xb.rq[t] += yb.rq[t];
    }
    xb.rw_c += rwn_b;
    hw_c_b += rwn_b;
    xb.rvx += rvxn_b;
    xb.rvy += rvyn_b;
    xb.rvz += rvzn_b;
    const double du_b = rvxn_b * u.ix + rvyn_b * u.iy + rvzn_b * u.iz;
    const double dv_b = rvxn_b * u.jx + rvyn_b * u.jy + rvzn_b * u.jz;
    const double kinv0_p_b = kinv_cp_b;
    hk_p_b += kinv_cp_b;
    const double kinv0_c_b = kinv_c_b;
    hk_c_b += kinv_c_b;
    const double kinh_b = kinh_u_b, dkinh_b = kinh_u_b;
    xb.re0 += e1_b;
    const double de_b = e1_b, dphi_b = e1_b;
    xb.rvx += kinh_b * x.rvx / x.rg_c;
    xb.rvy += kinh_b * x.rvy / x.rg_c;
    xb.rvz += kinh_b * x.rvz / x.rg_c;
    xb.rg_c -= kinh_b * kinh / x.rg_c;
    if (has_p) {
        xb.rw_p += kinv0_p_b * x.rw_p / den_p;
        const double d_b = -kinv0_p_b * kinv0_p / den_p;
        xb.rg_p += d_b * c2w0_p;
        xb.rg_c += d_b * c2w1_p;
    }
    if (has_hc) {
        xb.rw_c += kinv0_c_b * x.rw_c / den_c;
        const double d_b = -kinv0_c_b * kinv0_c / den_c;
        xb.rg_c += d_b * c2w0_c;
        xb.rg_m += d_b * c2w1_c;
    }
    // tendencies -> fluxes
    if (has_hc) {
        const double r = g.rdgzh[c];
        xb.fh[1][0] -= hk_c_b * r;
        xb.fh[0][0] += hk_c_b * r;
        xb.fh[1][1] -= hw_c_b * r;
        xb.fh[0][1] += hw_c_b * r;
    }
    if (has_p) {
        const double r = g.rdgzh[c + 1];
        xb.fh[2][0] -= hk_p_b * r;
        xb.fh[1][0] += hk_p_b * r;
        xb.fh[2][1] -= hw_p_b * r;
        xb.fh[1][1] += hw_p_b * r;
        const double rq = g.rdgz[c + 1];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            xb.fq[3][t] -= dqp_b[t] * rq;
            xb.fq[2][t] += dqp_b[t] * rq;
        }
    }
    if (has_m) {
        const double rq = g.rdgz[c - 1];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            xb.fq[1][t] -= dqm_b[t] * rq;
            xb.fq[0][t] += dqm_b[t] * rq;
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        xb.fq[2][t] -= dq_b[t] * rc;
        xb.fq[1][t] += dq_b[t] * rc;
    }
    const double sb[5] = {de_b, dphi_b, dkinh_b, du_b, dv_b};
#pragma unroll
    for (int i = 0; i < 5; i++) {
        xb.fs[1][i] -= sb[i] * rc;
        xb.fs[0][i] += sb[i] * rc;
    }
}

// ---------------------------------------------------------------------------------
// The point-wise tail (reverse of tail): rediagnosis, saturation adjustment through
// the implicit-function theorem at the Newton solution, final negative filter.
// ---------------------------------------------------------------------------------
struct TailInB {
    double rhog, rhoge, rhogq[6];
};

NSW6_HD void tail_b(const DevParams &P, bool interior, const TailIn<double> &x, const TailOut<double> &yb,
                    TailInB &xb)
{
    // ---- forward recomputation (values as tail<double>)
    const double rhog0 = x.rhog, rhoge0 = x.rhoge;
    const double gs = x.gsgam2;
    double q0[6];
#pragma unroll
    for (int n = 0; n < 6; n++) q0[n] = x.rhogq[n] / rhog0;
    double qd = 1.0;
#pragma unroll
    for (int n = 1; n <= 5; n++) qd = qd - q0[n];
    double cv = 0.0, qd2 = 1.0;
#pragma unroll
    for (int n = 1; n <= 5; n++) {
        cv = cv + q0[n] * C_CVW;
        qd2 = qd2 - q0[n];
    }
    cv = cv + qd2 * C_CVdry;
    const double rho0 = rhog0 / gs;
    const double tem0 = (rhoge0 / rhog0) / cv;
    // satadjust_all (ice = true: mixed phase, QI adjusted too) or satadjust_liq
    const bool ice = P.ice_adjust != 0.0;
    const double ein_moist = ice ? rhoge0 / rhog0 + q0[QV] * C_LHV0 - q0[QI] * C_LHF0 : rhoge0 / rhog0 + q0[QV] * C_LHV0;
    const double qsum = ice ? q0[QV] + q0[QC] + q0[QI] : q0[QV] + q0[QC];
    double q[6];
#pragma unroll
    for (int n = 0; n < 6; n++) q[n] = q0[n];
    q[QV] = qsum;
    q[QC] = 0.0;
    if (ice) q[QI] = 0.0;
    double CVtot = qd * C_CVdry;
#pragma unroll
    for (int n = 1; n <= 5; n++) CVtot = CVtot + q[n] * C_CVW;
    const double rho_sa = rhog0 / gs;
    const double tem1 = (ein_moist - q[QV] * C_LHV0) / CVtot;
    double tem = tem1;
    double qdl = 1.0;
    int newton = 0;
    if (interior) {
#pragma unroll
        for (int n = 1; n <= 5; n++) qdl = qdl - q[n];
        if (ice) {
            newton = satadjust_all_newton(P, tem, q, qsum, qdl, rho_sa, ein_moist);
        } else {
            newton = satadjust_liq_newton(P, tem, q, qsum, qdl, rho_sa, ein_moist);
            if (newton != 0) q[QC] = qsum - q[QV];
        }
    }
    const bool solved = interior && newton != 0;
    double rq[6];
#pragma unroll
    for (int n = 0; n < 6; n++) rq[n] = x.rhogq[n];
    rq[QV] = rhog0 * q[QV];
    rq[QC] = rhog0 * q[QC];
    if (ice) rq[QI] = rhog0 * q[QI];
    double CVtot2 = qd * C_CVdry;
#pragma unroll
    for (int n = 1; n <= 5; n++) CVtot2 = CVtot2 + q[n] * C_CVW;
    // final filter decisions
    double rqf[6], rhogf = rhog0;
    bool pos[6] = {true, true, true, true, true, true};
    double diffq2 = 0.0;
    if (interior) {
#pragma unroll
        for (int n = 0; n < 6; n++) rqf[n] = rq[n];
        double diffq = 0.0;
#pragma unroll
        for (int n = 2; n <= 5; n++) {
            diffq += rqf[n];
            pos[n] = !(rqf[n] < 0.0);
            rqf[n] = pos[n] ? rqf[n] : 0.0;
        }
#pragma unroll
        for (int n = 2; n <= 5; n++) diffq -= rqf[n];
        rqf[QV] += diffq;
        diffq2 = rqf[QV];
        pos[QV] = !(rqf[QV] < 0.0);
        rqf[QV] = pos[QV] ? rqf[QV] : 0.0;
        diffq2 -= rqf[QV];
        rhogf = rhog0 * (1.0 - diffq2);
    }

    // ---- reverse
    double rhog0_b = 0.0, rhoge0_b = 0.0, q0_b[6] = {0, 0, 0, 0, 0, 0};
    double tem_b = yb.tem, q_b[6], rq_b[6];
#pragma unroll
    for (int n = 0; n < 6; n++) {
        q_b[n] = yb.q[n];
        rq_b[n] = yb.rhogq[n];
    }
    double qd_b = yb.qd;
    double rho0_b = 0.0, tem0_b = 0.0, q0v_pre_b = 0.0, qd2_b = 0.0;
    double CVtot2_b = 0.0;
    if (interior) {
        // outputs: rhog = rhogf, rho = rhogf/gs, q[1..5] = rqf/rhogf (q[QV] not refreshed),
        // rhoge = tem rhogf cva, pre = rho (qdn Rdry + q[QV] Rvap) tem
        double qf[6];
#pragma unroll
        for (int n = 1; n <= 5; n++) qf[n] = rqf[n] / rhogf;
        double qdn = 1.0;
#pragma unroll
        for (int n = 1; n <= 5; n++) qdn = qdn - qf[n];
        double cva = qdn * C_CVdry;
#pragma unroll
        for (int n = 1; n <= 5; n++) cva = cva + qf[n] * C_CVW;
        const double rhof_ = rhogf / gs;
        const double Rg = qdn * C_Rdry + q[QV] * C_Rvap;
        double rhogf_b = yb.rhog;
        double rhof_b_ = yb.rho;
        // pre = rho * Rg * tem
        rhof_b_ += yb.pre * Rg * tem;
        const double Rg_b = yb.pre * rhof_ * tem;
        tem_b += yb.pre * rhof_ * Rg;
        double qdn_b = Rg_b * C_Rdry;
        q_b[QV] += Rg_b * C_Rvap;
        // rhoge = tem * rhogf * cva (cva independent of qf: CVW = CVdry)
        tem_b += yb.rhoge * rhogf * cva;
        rhogf_b += yb.rhoge * tem * cva;
        double qf_b[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int n = 1; n <= 5; n++) {
            qf_b[n] = q_b[n] - qdn_b;
            q_b[n] = 0.0;
        }
        rhogf_b += rhof_b_ / gs;
        double rqf_b[6];
#pragma unroll
        for (int n = 0; n < 6; n++) rqf_b[n] = rq_b[n];
#pragma unroll
        for (int n = 1; n <= 5; n++) {
            rqf_b[n] += qf_b[n] / rhogf;
            rhogf_b -= qf_b[n] * qf[n] / rhogf;
        }
        // rhogf = rhog0 * (1 - diffq2)
        rhog0_b += rhogf_b * (1.0 - diffq2);
        const double diffq2_b = -rhogf_b * rhog0;
        // a = rq[QV] + sum_n (rq[n] - clip(rq[n])); rqf[QV] = clip(a); diffq2 = a - clip(a)
        const double ma = pos[QV] ? 1.0 : 0.0;
        const double a_b = rqf_b[QV] * ma + diffq2_b * (1.0 - ma);
        double rqin_b[6];
#pragma unroll
        for (int n = 0; n < 6; n++) rqin_b[n] = 0.0;
        rqin_b[QV] += a_b;
        rqin_b[QC] += rqf_b[QC];
#pragma unroll
        for (int n = 2; n <= 5; n++) {
            // rqf[n] = clip(rq[n]); diffq += rq[n] - clip(rq[n])
            const double m = pos[n] ? 1.0 : 0.0;
            rqin_b[n] += rqf_b[n] * m + a_b * (1.0 - m);
        }
#pragma unroll
        for (int n = 0; n < 6; n++) rq_b[n] = rqin_b[n];
        // the state before the filter: rhoge (pre-filter) is overwritten, q[] as before
    } else {
        // halo outputs: rhog = rhog0, rho = rho0, pre = pre0, rhoge = rhog0 tem CVtot2
        rhog0_b += yb.rhog;
        rho0_b += yb.rho;
        // pre0 = rho0 * tem0 * (qd2 Rdry + q0[QV] Rvap)
        const double Rg0 = qd2 * C_Rdry + q0[QV] * C_Rvap;
        rho0_b += yb.pre * tem0 * Rg0;
        tem0_b += yb.pre * rho0 * Rg0;
        qd2_b += yb.pre * rho0 * tem0 * C_Rdry;
        q0v_pre_b += yb.pre * rho0 * tem0 * C_Rvap;
        rhog0_b += yb.rhoge * tem * CVtot2;
        tem_b += yb.rhoge * rhog0 * CVtot2;
        CVtot2_b += yb.rhoge * rhog0 * tem;
    }
    // rhogq[QV] = rhog0 q[QV], rhogq[QC] = rhog0 q[QC], under satadjust_all also
    // rhogq[QI] = rhog0 q[QI] (others pass through)
    rhog0_b += rq_b[QV] * q[QV] + rq_b[QC] * q[QC];
    q_b[QV] += rq_b[QV] * rhog0;
    q_b[QC] += rq_b[QC] * rhog0;
    if (ice) {
        rhog0_b += rq_b[QI] * q[QI];
        q_b[QI] += rq_b[QI] * rhog0;
    }
    double rhogq_in_b[6];
#pragma unroll
    for (int n = 0; n < 6; n++) rhogq_in_b[n] = (n == QV || n == QC || (ice && n == QI)) ? 0.0 : rq_b[n];
    // CVtot2 = qd CVdry + sum_{1..5} q[n] CVW  (q[QC] is the adjusted one)
    qd_b += CVtot2_b * C_CVdry;
    q_b[QC] += CVtot2_b * C_CVW;
#pragma unroll
    for (int n = 2; n <= 5; n++) q_b[n] += CVtot2_b * C_CVW;
    // saturation adjustment
    double qsum_b = 0.0, ein_moist_b = 0.0, rho_sa_b = 0.0, qdl_b = 0.0, CVtot_b = 0.0;
    double qrest_b = 0.0;
    if (solved && ice) {
        // q[QV] = qsat(T*), q[QC] = (qsum - qsat) alpha(T*), q[QI] = (qsum - qsat)(1 - alpha);
        // T* root of R_all(T; qsum, qdl, qrest, rho_sa, ein), qrest = qr + qs + qg
        const SatAllSlopes sl = satadjust_all_slopes(P, tem, rho_sa);
        const double d = qsum - sl.qsat;
        const double d_b = q_b[QC] * sl.alpha + q_b[QI] * (1.0 - sl.alpha);
        const double alpha_b = (q_b[QC] - q_b[QI]) * d;
        qsum_b += d_b;
        const double qsat_b = q_b[QV] - d_b;
        q_b[QV] = q_b[QC] = q_b[QI] = 0.0;
        tem_b += alpha_b * sl.dalpha + qsat_b * sl.dqsat;
        rho_sa_b -= qsat_b * sl.qsat / rho_sa;
        const double qrest = (q[QR] + q[QS]) + q[QG];
        const double dRdT = satadjust_all_dRdT(P, tem, qsum, qdl, qrest, rho_sa);
        const double lam = -tem_b / dRdT;
        tem_b = 0.0;
        // R = T (qdl CVdry + (qsum - qsat + qrest) CVW) + qsat LHV0 - (qsum - qsat)(1 - alpha) LHF0 - ein
        qdl_b += lam * tem * C_CVdry;
        qsum_b += lam * (tem * C_CVW - (1.0 - sl.alpha) * C_LHF0);
        qrest_b += lam * tem * C_CVW;
        rho_sa_b -= lam * (sl.qsat / rho_sa) * (C_LHV0 - tem * C_CVW + (1.0 - sl.alpha) * C_LHF0);
        ein_moist_b -= lam;
    } else if (solved) {
        // q[QC] = qsum - q[QV]; q[QV] = qsat(T*, rho_sa); T* root of R(T; qsum, qdl, qrest, rho_sa, ein)
        qsum_b += q_b[QC];
        q_b[QV] -= q_b[QC];
        q_b[QC] = 0.0;
        const double psat = C_PSAT0 * 1.0 * exp(P.LovR_liq * (P.RTEM00 - 1.0 / dmax(tem, TEM_MIN)));
        const double qsat = psat / (rho_sa * C_Rvap * tem);
        const double dpsat = (tem > TEM_MIN) ? psat * P.LovR_liq / (tem * tem) : 0.0;
        const double dqsat_dT = dpsat / (rho_sa * C_Rvap * tem) - qsat / tem;
        const double dqsat_drho = -qsat / rho_sa;
        tem_b += q_b[QV] * dqsat_dT;
        rho_sa_b += q_b[QV] * dqsat_drho;
        q_b[QV] = 0.0;
        const double qrest = ((q[QR] + q[QI]) + q[QS]) + q[QG];
        const double dRdT = satadjust_dRdT(P, tem, qsum, qdl, qrest, rho_sa);
        const double lam = -tem_b / dRdT;   // cotangent of the residual
        tem_b = 0.0;
        // R = T (qdl CVdry + ((qsum - qsat) + qrest) CVW) + qsat LHV0 - ein_moist
        qdl_b += lam * tem * C_CVdry;
        qsum_b += lam * tem * C_CVW;
        qrest_b += lam * tem * C_CVW;
        rho_sa_b += lam * dqsat_drho * (C_LHV0 - tem * C_CVW);
        ein_moist_b -= lam;
    } else {
        // no iteration: q[QV] = qsum, q[QC] = 0 (and q[QI] = 0 under satadjust_all), tem = tem1
        qsum_b += q_b[QV];
        q_b[QV] = 0.0;
        q_b[QC] = 0.0;
    }
    // qdl = 1 - sum_{1..5} q (with q[QC] = 0) ; qrest = qr + qi + qs + qg
#pragma unroll
    for (int n = 2; n <= 5; n++) q_b[n] += qrest_b - qdl_b;
    // tem1 = (ein_moist - qsum LHV0) / CVtot, CVtot = qd CVdry + sum_{2..5} q CVW
    ein_moist_b += tem_b / CVtot;
    qsum_b -= tem_b * C_LHV0 / CVtot;
    CVtot_b -= tem_b * tem1 / CVtot;
    qd_b += CVtot_b * C_CVdry;
#pragma unroll
    for (int n = 2; n <= 5; n++) q_b[n] += CVtot_b * C_CVW;
    // under satadjust_all, q[QI] was set to 0 before the solve
    if (ice) q_b[QI] = 0.0;
    // rho_sa = rhog0 / gs
    rhog0_b += rho_sa_b / gs;
    // qsum = q0[QV] + q0[QC] (+ q0[QI])
    q0_b[QV] += qsum_b;
    q0_b[QC] += qsum_b;
    if (ice) q0_b[QI] += qsum_b;
    // ein_moist = rhoge0/rhog0 + q0[QV] LHV0 (- q0[QI] LHF0)
    rhoge0_b += ein_moist_b / rhog0;
    rhog0_b -= ein_moist_b * (rhoge0 / rhog0) / rhog0;
    q0_b[QV] += ein_moist_b * C_LHV0;
    if (ice) q0_b[QI] -= ein_moist_b * C_LHF0;
    // halo pre / rho terms of the first rediagnosis
    q0_b[QV] += q0v_pre_b;
    // tem0 = (rhoge0 / rhog0) / cv (cv independent of q)
    rhoge0_b += tem0_b / (rhog0 * cv);
    rhog0_b -= tem0_b * tem0 / rhog0;
    rhog0_b += rho0_b / gs;
    // qd = 1 - sum q0[1..5], qd2 = the same sum
#pragma unroll
    for (int n = 1; n <= 5; n++) q0_b[n] -= qd_b + qd2_b;
    // q[2..5] are q0[2..5]
#pragma unroll
    for (int n = 2; n <= 5; n++) q0_b[n] += q_b[n];
    // q0[n] = rhogq[n] / rhog0
#pragma unroll
    for (int n = 0; n < 6; n++) {
        rhogq_in_b[n] += q0_b[n] / rhog0;
        rhog0_b -= q0_b[n] * q0[n] / rhog0;
    }
    xb.rhog += rhog0_b;
    xb.rhoge += rhoge0_b;
#pragma unroll
    for (int n = 0; n < 6; n++) xb.rhogq[n] += rhogq_in_b[n];
}

}  // namespace nsw6

namespace nsw6 {

// ---------------------------------------------------------------------------------
// Process rates (reverse of rates<double>: Roh & Satoh 2014 or exponential snow
// moments, KK2000 or Berry's autoconversion, with or without explicit ice generation)
// ---------------------------------------------------------------------------------
struct RateInB {
    double dens, temp, q[6], qsatl, qsati, pre, temc, sice_m1, sliq_1m, qv_qsati, tun[NCOEF];
    double nc_pm179;   // Nc^-1.79 of the point (per point under OPT_INDIR)
    double nc_n, nc_dc;   // Nc and Dc of the point (Berry's autoconversion)
};

// rates that enter the limiter blocks and the tendencies
enum {
    rPigen, rPidep, rPsdep, rPgdep, rPraut, rPracw, rPihom, rPihtr, rPsacw, rPsfw, rPgacw, rPrevp, rPiacr_s,
    rPiacr_g, rPsacr_s, rPsacr_g, rPgacr, rPgfrz, rPisub, rPimlt, rPsaut, rPraci_s, rPraci_g, rPsaci, rPsfi,
    rPgaci, rPssub, rPsmlt, rPgaut, rPracs, rPgacs, rPgsub, rPgmlt, NRATE
};
// the six loss limiters in the reference's order (src/mod_mp_nsw6.cpp:1290-1488); each
// block's signed member list is also its species' tendency (:1491-1540)
constexpr int NBLK = 6, BLKMAX = 14;
constexpr int blk_q[NBLK] = {QC, QI, QR, QV, QS, QG};
constexpr int blk_n[NBLK] = {8, 12, 11, 8, 14, 13};
constexpr int blk_r[NBLK][BLKMAX] = {
    {rPimlt, rPraut, rPracw, rPihom, rPihtr, rPsacw, rPsfw, rPgacw},
    {rPigen, rPidep, rPihom, rPihtr, rPisub, rPimlt, rPsaut, rPraci_s, rPsaci, rPsfi, rPraci_g, rPgaci},
    {rPraut, rPracw, rPsmlt, rPgmlt, rPrevp, rPiacr_s, rPsacr_s, rPiacr_g, rPsacr_g, rPgacr, rPgfrz},
    {rPrevp, rPisub, rPssub, rPgsub, rPigen, rPidep, rPsdep, rPgdep},
    {rPsdep, rPsacw, rPsfw, rPiacr_s, rPsacr_s, rPsaut, rPraci_s, rPsaci, rPsfi, rPssub, rPsmlt, rPgaut, rPracs,
     rPgacs},
    {rPgdep, rPgacw, rPiacr_g, rPsacr_g, rPgacr, rPgfrz, rPraci_g, rPgaci, rPgaut, rPracs, rPgacs, rPgsub,
     rPgmlt}};
constexpr double blk_s[NBLK][BLKMAX] = {
    {+1, -1, -1, -1, -1, -1, -1, -1},
    {+1, +1, +1, +1, -1, -1, -1, -1, -1, -1, -1, -1},
    {+1, +1, +1, +1, -1, -1, -1, -1, -1, -1, -1},
    {+1, +1, +1, +1, -1, -1, -1, -1},
    {+1, +1, +1, +1, +1, +1, +1, +1, +1, -1, -1, -1, -1, -1},
    {+1, +1, +1, +1, +1, +1, +1, +1, +1, +1, +1, -1, -1}};
// species of the tendency outputs qt[1..5] -> block index
constexpr int tend_blk[6] = {-1, 0, 2, 1, 4, 5};

NSW6_HD double sel(bool c, double a, double b) { return c ? a : b; }

// Slope of the exponential snow size distribution and the powers of it that the
// moments take (rates<R> without Roh & Satoh): x = dens qs / (As N0s GAM_1bs) + zs,
// s = x^(1/4) (1 - zs), sds = s^(1/4).
struct SnowSlope {
    double x, s, sds, s2, s3, s1bs, s2bs, s3bs, s3ds, s5dsh;
};
NSW6_HD SnowSlope snow_slope(const DevParams &P, double dens, double qs, double zs)
{
    SnowSlope L;
    L.x = dens * qs * P.rK_s + zs;
    L.s = sqrt(sqrt(L.x)) * (1.0 - zs);
    L.sds = sqrt(sqrt(L.s));
    L.s2 = L.s * L.s;
    L.s3 = L.s2 * L.s;
    L.s1bs = L.s2 * L.s2;
    L.s2bs = L.s1bs * L.s;
    L.s3bs = L.s3 * L.s3;
    L.s3ds = L.s3 * L.sds;
    L.s5dsh = sqrt(L.s2bs * L.sds);
    return L;
}

// Reverse of bergeron(): the temperature cotangent from those of a1, a2 and ma2 = 1 -
// a2 (a1 = v1 1e-3^ma2; v1, v2 linear in the clamped temperature within a table cell;
// zero where the clamp holds the temperature).
NSW6_HD double bergeron_b(const double *a1_tab, const double *a2_tab, double tem, double a1, double ma2, double a1_b,
                          double a2_b, double ma2_b)
{
    const double x = tem - C_TEM00;
    if (x < -30.99 || 0.0 < x || !(x == x)) return 0.0;
    const int itemc = (int)(-x);
    const double v1_b = a1_b * pow(1.0E-3, ma2);
    const double v2_b = a2_b - (ma2_b + a1_b * a1 * log(1.0E-3));
    // v = (1 - fact) t[i] + fact t[i+1], fact = -(temc + i - 1)
    const double fact_b =
        v1_b * (a1_tab[itemc + 1] - a1_tab[itemc]) + v2_b * (a2_tab[itemc + 1] - a2_tab[itemc]);
    return -fact_b;
}

// A copy of the rate inputs that the device compiler treats as unknown values, so
// that expressions of them evaluated before and after are computed twice rather
// than kept in registers in between.
NSW6_HD double opaque(double x)
{
#ifdef __HIP_DEVICE_COMPILE__
    asm volatile("" : "+v"(x));
#endif
    return x;
}
NSW6_HD RateIn<double> opaque_copy(const RateIn<double> &a)
{
    RateIn<double> r = a;
    double *w = reinterpret_cast<double *>(&r);
    static_assert(sizeof(RateIn<double>) % sizeof(double) == 0, "RateIn<double> holds doubles only");
#pragma unroll
    for (int i = 0; i < (int)(sizeof(RateIn<double>) / sizeof(double)); i++) w[i] = opaque(w[i]);
    return r;
}
NSW6_HD PointIn<double> opaque_point(const PointIn<double> &a)
{
    PointIn<double> r;
    r.rhog = opaque(a.rhog);
#pragma unroll
    for (int n = 0; n < NQ; n++) r.rhogq[n] = opaque(a.rhogq[n]);
    r.tem = opaque(a.tem);
    r.qv = opaque(a.qv);
    r.gsgam2 = opaque(a.gsgam2);
    return r;
}

// Per-thread storage that the rate reverse parks its phase-1 results in across the
// phase-2 recompute: element j at base[j * stride] (LDS on the device, a local array
// on the host).  Kept in registers, the NRATE + 6 values would be live through the
// whole recompute.
// Slots: [0, NRATE) raw-rate cotangents, then dq/dt cotangents (6), the forward
// tendencies qt[1..5] (5, from phase 1 to the return), and point_rates_b's own
// state-update cotangents held across rates_b (POINT_HELD).
constexpr int RATE_STASH = NRATE + 6 + 5;
constexpr int POINT_HELD = 14;
constexpr int POINT_STASH = RATE_STASH + POINT_HELD;
struct RateStash {
    double *base;
    int stride;
    NSW6_HD double &operator[](int j) const { return base[(int64_t)j * stride]; }
};
// Compiler barrier: values written to the stash before it are read back after it
// rather than kept in registers.
NSW6_HD void stash_fence()
{
#ifdef __HIP_DEVICE_COMPILE__
    asm volatile("" ::: "memory");
#endif
}

// Also returns the forward tendencies qt[1..5] (the values rates<double> gives).
// ncp: the point's droplet number Nc, Nc^-1.79 and Dc under OPT_INDIR; nullptr reads
// them from P where they are used (so they are not held in registers through the
// pass); their cotangents are added to b.nc_n, b.nc_pm179 and b.nc_dc.  DEFAULT_SCHEME:
// sw_kk2000 = 1, sw_expice = 1 and sw_roh2014 = 1 known at compile time (otherwise read
// from P at run time).  a1_tab/a2_tab: the Bergeron tables (read only without explicit ice generation).
template <bool DEFAULT_SCHEME>
NSW6_HD void rates_b(const DevParams &P, double dt, const RateIn<double> &in0, const NcPoint *ncp, const double qt_b[6],
                     const double vt_b[3], double vti_b, RateInB &b, double qt[6], const RateStash &st, const double *a1_tab,
                     const double *a2_tab)
{
    double raw_b[NRATE], dqdt_b[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < NRATE; j++) raw_b[j] = 0.0;
    // ---- phase 1: forward, limiter decisions, reverse of the limiter algebra -> raw_b
    {
        const RateIn<double> &in = in0;
        const double nc_pm179 = ncp ? ncp->pow_m179 : P.Nc_pow_m179;
#include "nsw6_rates_fwd.inc"
        // limiters: R = min(P0, dq_dt) * c; lim = true when dq_dt is taken
        double R[NRATE];
        bool lim[NRATE];
        double cst[NRATE];
        double raw[NRATE];   // the limited quantity before min
        int limq[NRATE];     // species whose dq_dt limits it (-1: none)
#pragma unroll
        for (int j = 0; j < NRATE; j++) {
            lim[j] = false;
            cst[j] = 1.0;
            limq[j] = -1;
            raw[j] = 0.0;
        }
        auto setlim = [&](int j, double p0, int sp, double c) {
            raw[j] = p0;
            limq[j] = sp;
            cst[j] = c;
            lim[j] = (dqdt[sp] < p0);
            R[j] = (lim[j] ? dqdt[sp] : p0) * c;
        };
        setlim(rPigen, Pigen0, QV, iex);
        setlim(rPidep, Pidep0, QV, iex);
        setlim(rPsdep, Psdep0, QV, ie);
        setlim(rPgdep, Pgdep0, QV, ie);
        const double Pracw1 = Pracw0 + Psacw0 * we + Pgacw0 * we;
        setlim(rPraut, Praut0, QC, 1.0);
        setlim(rPracw, Pracw1, QC, 1.0);
        setlim(rPihom, Pihom0, QC, iex);
        setlim(rPihtr, Pihtr0, QC, iex);
        setlim(rPsacw, Psacw0, QC, ie);
        setlim(rPsfw, Psfw0, QC, ieb);
        setlim(rPgacw, Pgacw0, QC, ie);
        setlim(rPrevp, Prevp0, QR, 1.0);
        double Piacr1, Psacr1;
        {
            const bool l1 = (dqdt[QR] < Piacr0);
            Piacr1 = (l1 ? dqdt[QR] : Piacr0) * ie;
            lim[rPiacr_s] = lim[rPiacr_g] = l1;
            raw[rPiacr_s] = raw[rPiacr_g] = Piacr0;
            limq[rPiacr_s] = limq[rPiacr_g] = QR;
            const bool l2 = (dqdt[QR] < Psacr0);
            Psacr1 = (l2 ? dqdt[QR] : Psacr0) * ie;
            lim[rPsacr_s] = lim[rPsacr_g] = l2;
            raw[rPsacr_s] = raw[rPsacr_g] = Psacr0;
            limq[rPsacr_s] = limq[rPsacr_g] = QR;
        }
        setlim(rPgacr, Pgacr0, QR, ie);
        setlim(rPgfrz, Pgfrz0, QR, ie);
        setlim(rPisub, Pisub0, QI, iex);
        setlim(rPimlt, Pimlt0, QI, wex);
        setlim(rPsaut, Psaut0, QI, ie);
        double Praci1;
        {
            const bool l1 = (dqdt[QI] < Praci0);
            Praci1 = (l1 ? dqdt[QI] : Praci0) * ie;
            lim[rPraci_s] = lim[rPraci_g] = l1;
            raw[rPraci_s] = raw[rPraci_g] = Praci0;
            limq[rPraci_s] = limq[rPraci_g] = QI;
        }
        setlim(rPsaci, Psaci0, QI, ie);
        setlim(rPsfi, Psfi0, QI, ieb);
        setlim(rPgaci, Pgaci0, QI, ie);
        setlim(rPssub, Pssub0, QS, ie);
        setlim(rPsmlt, Psmlt0, QS, we);
        setlim(rPgaut, Pgaut0, QS, ie);
        double Pracs1;
        {
            const bool l1 = (dqdt[QS] < Pracs0);
            Pracs1 = (l1 ? dqdt[QS] : Pracs0) * ie;
            lim[rPracs] = l1;
            raw[rPracs] = Pracs0;
            limq[rPracs] = QS;
        }
        setlim(rPgacs, Pgacs0, QS, 1.0);
        setlim(rPgsub, Pgsub0, QG, ie);
        setlim(rPgmlt, Pgmlt0, QG, we);
        // splits
        R[rPiacr_s] = (1.0 - delta1) * Piacr1;
        R[rPiacr_g] = delta1 * Piacr1;
        R[rPraci_s] = (1.0 - delta1) * Praci1;
        R[rPraci_g] = delta1 * Praci1;
        R[rPsacr_s] = delta2 * Psacr1;
        R[rPsacr_g] = (1.0 - delta2) * Psacr1;
        R[rPracs] = (1.0 - delta2) * Pracs1;

        // limiter blocks: R0[b] = block entry values of the members, fac, ratio branch
        double Rin[NBLK][BLKMAX], net[NBLK], fac[NBLK];
        bool ratio[NBLK];
#pragma unroll
        for (int k = 0; k < NBLK; k++) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < BLKMAX; m++) {
                if (m >= blk_n[k]) break;
                Rin[k][m] = R[blk_r[k][m]];
                s = (m == 0) ? blk_s[k][m] * Rin[k][m] : s + blk_s[k][m] * Rin[k][m];
            }
            net[k] = s;
            const double fs = 0.5 + copysign(0.5, s + EPS);
            const double rt = -dqdt[blk_q[k]] / (s - fs);
            ratio[k] = (fs == 0.0) && !(1.0 < rt);
            fac[k] = fs + (1.0 - fs) * dmin(rt, 1.0);
#pragma unroll
            for (int m = 0; m < BLKMAX; m++) {
                if (m >= blk_n[k]) break;
                R[blk_r[k][m]] *= fac[k];
            }
        }

        // =============================== reverse ===============================
        double R_b[NRATE];
#pragma unroll
        for (int j = 0; j < NRATE; j++) R_b[j] = 0.0;
        // tendencies q_t = max(sum, -dq_dt)
#pragma unroll
        for (int n = 1; n < 6; n++) {
            const int k = tend_blk[n];
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < BLKMAX; m++) {
                if (m >= blk_n[k]) break;
                s = (m == 0) ? blk_s[k][m] * R[blk_r[k][m]] : s + blk_s[k][m] * R[blk_r[k][m]];
            }
            qt[n] = (s < -dqdt[n]) ? -dqdt[n] : s;
            if (s < -dqdt[n]) {
                dqdt_b[n] -= qt_b[n];
            } else {
#pragma unroll
                for (int m = 0; m < BLKMAX; m++) {
                    if (m >= blk_n[k]) break;
                    R_b[blk_r[k][m]] += blk_s[k][m] * qt_b[n];
                }
            }
        }
        // limiter blocks in reverse
#pragma unroll
        for (int k = NBLK - 1; k >= 0; k--) {
            double fac_b = 0.0;
#pragma unroll
            for (int m = 0; m < BLKMAX; m++) {
                if (m >= blk_n[k]) break;
                const int j = blk_r[k][m];
                fac_b += R_b[j] * Rin[k][m];
                R_b[j] *= fac[k];
            }
            if (ratio[k]) {
                // fac = -dq_dt / net
                const double n_ = net[k];
                dqdt_b[blk_q[k]] -= fac_b / n_;
                const double net_b = fac_b * dqdt[blk_q[k]] / (n_ * n_);
#pragma unroll
                for (int m = 0; m < BLKMAX; m++) {
                    if (m >= blk_n[k]) break;
                    R_b[blk_r[k][m]] += blk_s[k][m] * net_b;
                }
            }
        }
        // splits
        const double Piacr1_b = (1.0 - delta1) * R_b[rPiacr_s] + delta1 * R_b[rPiacr_g];
        const double Praci1_b = (1.0 - delta1) * R_b[rPraci_s] + delta1 * R_b[rPraci_g];
        const double Psacr1_b = delta2 * R_b[rPsacr_s] + (1.0 - delta2) * R_b[rPsacr_g];
        const double Pracs1_b = (1.0 - delta2) * R_b[rPracs];
        // limiters: cotangent of the quantity before min (raw) or of dq_dt
        auto unlim = [&](int j, double rb) {
            const double v = rb * cst[j];
            if (lim[j])
                dqdt_b[limq[j]] += v;
            else
                raw_b[j] += v;
        };
        const int plain[] = {rPigen, rPidep, rPsdep, rPgdep, rPraut, rPracw, rPihom, rPihtr, rPsacw, rPsfw, rPgacw,
                             rPrevp, rPgacr, rPgfrz, rPisub, rPimlt, rPsaut, rPsaci, rPsfi, rPgaci, rPssub, rPsmlt,
                             rPgaut, rPgacs, rPgsub, rPgmlt};
#pragma unroll
        for (int i = 0; i < (int)(sizeof(plain) / sizeof(plain[0])); i++) unlim(plain[i], R_b[plain[i]]);
        // split rates share one limited quantity: constant ie, limiter of the pair
        {
            const double v1 = Piacr1_b * ie;
            if (lim[rPiacr_s]) dqdt_b[QR] += v1; else raw_b[rPiacr_s] += v1;
            const double v2 = Psacr1_b * ie;
            if (lim[rPsacr_s]) dqdt_b[QR] += v2; else raw_b[rPsacr_s] += v2;
            const double v3 = Praci1_b * ie;
            if (lim[rPraci_s]) dqdt_b[QI] += v3; else raw_b[rPraci_s] += v3;
            const double v4 = Pracs1_b * ie;
            if (lim[rPracs]) dqdt_b[QS] += v4; else raw_b[rPracs] += v4;
        }
    }
    // the phase-1 cotangents wait in the stash while the forward values are recomputed
#pragma unroll
    for (int j = 0; j < NRATE; j++) st[j] = raw_b[j];
#pragma unroll
    for (int n = 0; n < 6; n++) st[NRATE + n] = dqdt_b[n];
#pragma unroll
    for (int n = 1; n <= 5; n++) st[NRATE + 6 + n - 1] = qt[n];
    stash_fence();
    // ---- phase 2: the forward values again, from inputs the compiler cannot see
    // through (so it recomputes them rather than keeping phase 1 values live), then
    // the reverse of the rate formulas
    const RateIn<double> in = opaque_copy(in0);
    const double nc_pm179 = opaque(ncp ? ncp->pow_m179 : P.Nc_pow_m179);
#include "nsw6_rates_fwd.inc"
    stash_fence();
#pragma unroll
    for (int j = 0; j < NRATE; j++) raw_b[j] = st[j];
#pragma unroll
    for (int n = 0; n < 6; n++) dqdt_b[n] = st[NRATE + n];
    // raw rates -> their pre-limiter forms
    double Pigen0_b = raw_b[rPigen], Pidep0_b = raw_b[rPidep], Psdep0_b = raw_b[rPsdep], Pgdep0_b = raw_b[rPgdep];
    double Praut0_b = raw_b[rPraut], Pihom0_b = raw_b[rPihom], Pihtr0_b = raw_b[rPihtr];
    double Psacw0_b = raw_b[rPsacw], Pgacw0_b = raw_b[rPgacw], Prevp0_b = raw_b[rPrevp];
    double Piacr0_b = raw_b[rPiacr_s], Psacr0_b = raw_b[rPsacr_s], Pgacr0_b = raw_b[rPgacr];
    double Pgfrz0_b = raw_b[rPgfrz], Pisub0_b = raw_b[rPisub], Pimlt0_b = raw_b[rPimlt];
    double Psaut0_b = raw_b[rPsaut], Praci0_b = raw_b[rPraci_s], Psaci0_b = raw_b[rPsaci];
    double Pgaci0_b = raw_b[rPgaci], Pssub0_b = raw_b[rPssub], Psmlt0_b = raw_b[rPsmlt];
    double Pgaut0_b = raw_b[rPgaut], Pracs0_b = raw_b[rPracs], Pgacs0_b = raw_b[rPgacs];
    double Pgsub0_b = raw_b[rPgsub], Pgmlt0_b = raw_b[rPgmlt];
    // Pracw1 = Pracw0 + (Psacw0 + Pgacw0) we
    const double Pracw1_b = raw_b[rPracw];
    const double Pracw0_b = Pracw1_b;
    Psacw0_b += Pracw1_b * we;
    Pgacw0_b += Pracw1_b * we;

    // intermediate cotangents
    double dens_b = 0.0, temc_b = 0.0, Rdens_b = 0.0, rf_b = 0.0, temp_b = 0.0;
    double q_b[6] = {0, 0, 0, 0, 0, 0};
    double L1_b = 0.0, Ldr_b = 0.0, L2_b = 0.0, L3_b = 0.0, L4_b = 0.0, L5_b = 0.0, L6_b = 0.0, L7_b = 0.0;
    double L3dr_b = 0.0, L5dr_b = 0.0, L6dr_b = 0.0;
    double G1_b = 0.0, Gdg_b = 0.0, G2_b = 0.0, G3_b = 0.0, G3dg_b = 0.0, G5dg_b = 0.0;
    double M0_b = 0.0, M1_b = 0.0, M2_b = 0.0, M0bs_b = 0.0, M1bs_b = 0.0, M2bs_b = 0.0, M2ds_b = 0.0, M5dsh_b = 0.0;
    double Lsds_b = 0.0;
    double Vtr_b = vt_b[0], Vts_b = vt_b[1], Vtg_b = vt_b[2];
    double Esi_mod_b = 0.0, Egs_mod_b = 0.0;
    double Giv_b = 0.0, Glv_b = 0.0, Gil_b = 0.0, Nu_b = 0.0, Ka_b = 0.0, Kd_b = 0.0;
    double frz_b = 0.0, sice_b = 0.0;
    double *tun_b = b.tun;

    // Pgfrz0 = kfrz * Rdens * frz * L7
    Rdens_b += Pgfrz0_b * kfrz * frz * L7;
    frz_b += Pgfrz0_b * kfrz * Rdens * L7;
    L7_b += Pgfrz0_b * kfrz * Rdens * frz;
    // Bergeron growth: Psfw0 = Ni50 W, W = a1 mi50^a2 + kW dens qc, Psfi0 = qi / dt1,
    // Ni50 = qi dt / (mi50 dt1), dt1 = (mi50^ma2 - mi40^ma2) / (a1 ma2)
    if (berg) {
        const double Psfw0_b = raw_b[rPsfw], Psfi0_b = raw_b[rPsfi];
        const double Ni50_b = Psfw0_b * bW, W_b = Psfw0_b * bNi50;
        double a1_b = W_b * bE2;
        const double a2_b = W_b * in.a1 * bE2 * P.ln_mi50;
        const double kW = PI * P.Eiw * P.Ri50 * P.Ri50 * P.vti50;
        dens_b += W_b * kW * qc;
        q_b[QC] += W_b * kW * dens;
        q_b[QI] += Ni50_b * dt / (P.mi50 * bdt1) + Psfi0_b / bdt1;
        const double dt1_b = -(Ni50_b * bNi50 + Psfi0_b * Psfi0) / bdt1;
        a1_b -= dt1_b * bdt1 / in.a1;
        const double ma2_b = -dt1_b * bdt1 / in.ma2 + dt1_b / (in.a1 * in.ma2) * (P.ln_mi50 * bE50 - P.ln_mi40 * bE40);
        temp_b += bergeron_b(a1_tab, a2_tab, temp, in.a1, in.ma2, a1_b, a2_b, ma2_b);
    }
    // Pgmlt00 = kev Rdens N0g Gil ventg + cmlt (Pgacw0 + Pgacr0)
    double ventg_b = 0.0, cmlt_b = 0.0;
    if (gmlt_pos) {
        Rdens_b += Pgmlt0_b * kev * N0g * Gil * ventg;
        Gil_b += Pgmlt0_b * kev * Rdens * N0g * ventg;
        ventg_b += Pgmlt0_b * kev * Rdens * N0g * Gil;
        cmlt_b += Pgmlt0_b * (Pgacw0 + Pgacr0);
        Pgacw0_b += Pgmlt0_b * cmlt;
        Pgacr0_b += Pgmlt0_b * cmlt;
    }
    // Pgdep0 = spsati tmpg, Pgsub0 = -(1 - spsati) tmpg
    const double tmpg_b = spsati * Pgdep0_b - (1.0 - spsati) * Pgsub0_b;
    Rdens_b += tmpg_b * kev * N0g * in.sice_m1 * Giv * ventg;
    sice_b += tmpg_b * kev * Rdens * N0g * Giv * ventg;
    Giv_b += tmpg_b * kev * Rdens * N0g * in.sice_m1 * ventg;
    ventg_b += tmpg_b * kev * Rdens * N0g * in.sice_m1 * Giv;
    // ventg = f1g GAM_2 G2 + f2g srg GAM_5dg_h, srg = sqrt(Cg rho_fact / Nu G5dg)
    G2_b += ventg_b * P.f1g * GAM_2;
    {
        const double srg_b = ventg_b * P.f2g * P.GAM_5dg_h;
        const double arg_b = (srg > 0.0) ? srg_b * 0.5 / srg : 0.0;
        // arg = ((Cg rho_fact) / Nu) G5dg
        tun_b[6] += arg_b * rho_fact / Nu * G5dg;
        rf_b += arg_b * Cg / Nu * G5dg;
        Nu_b -= arg_b * (Cg * rho_fact / Nu) * G5dg / Nu;
        G5dg_b += arg_b * (Cg * rho_fact / Nu);
    }
    // Psmlt00 = kev Rdens Gil vents + cmlt (Psacw0 + Psacr0)
    double vents_b = 0.0;
    if (smlt_pos) {
        Rdens_b += Psmlt0_b * kev * Gil * vents;
        Gil_b += Psmlt0_b * kev * Rdens * vents;
        vents_b += Psmlt0_b * kev * Rdens * Gil;
        cmlt_b += Psmlt0_b * (Psacw0 + Psacr0);
        Psacw0_b += Psmlt0_b * cmlt;
        Psacr0_b += Psmlt0_b * cmlt;
    }
    temc_b += cmlt_b * C_CL / C_LHF0;
    // Psdep0 / Pssub0 from tmps = kev Rdens sice Giv vents
    const double tmps_b = spsati * Psdep0_b - (1.0 - spsati) * Pssub0_b;
    Rdens_b += tmps_b * kev * in.sice_m1 * Giv * vents;
    sice_b += tmps_b * kev * Rdens * Giv * vents;
    Giv_b += tmps_b * kev * Rdens * in.sice_m1 * vents;
    vents_b += tmps_b * kev * Rdens * in.sice_m1 * Giv;
    // vents = f1s M1 + f2s srs M5dsh, srs = sqrt(Cs rho_fact / Nu)
    M1_b += vents_b * P.f1s;
    M5dsh_b += vents_b * P.f2s * srs;
    {
        const double srs_b = vents_b * P.f2s * M5dsh;
        const double arg_b = (srs > 0.0) ? srs_b * 0.5 / srs : 0.0;
        tun_b[5] += arg_b * rho_fact / Nu;
        rf_b += arg_b * Cs / Nu;
        Nu_b -= arg_b * (Cs * rho_fact / Nu) / Nu;
    }
    // Pimlt0 = sw_mlt qi / dt ; Pihtr0 = sw_htr (dens / rho_w (qc qc) / (Nc 1e6)) B_frz frz
    q_b[QI] += Pimlt0_b * sw.sw_mlt / dt;
    {
        const double kh = sw.sw_htr * P.B_frz / (P.Nc_ihtr * 1.0E+6) / C_DWATR;
        // Pihtr0 = kh' dens qc^2 frz (product form of the same value)
        dens_b += Pihtr0_b * kh * qc * qc * frz;
        q_b[QC] += Pihtr0_b * kh * dens * 2.0 * qc * frz;
        frz_b += Pihtr0_b * kh * dens * qc * qc;
    }
    q_b[QC] += Pihom0_b * sw.sw_hom / dt;
    // frz = exp(-A_frz temc) - 1
    temc_b += frz_b * efr * (-P.A_frz);
    // Pidep0 / Pisub0 from tmpi = 4 Di XNi Rdens sice Giv
    {
        const double tb = spsati * Pidep0_b - (1.0 - spsati) * Pisub0_b;
        double Di_b = tb * 4.0 * XNi * Rdens * in.sice_m1 * Giv;
        double XNi_b = tb * 4.0 * Di * Rdens * in.sice_m1 * Giv;
        Rdens_b += tb * 4.0 * Di * XNi * in.sice_m1 * Giv;
        sice_b += tb * 4.0 * Di * XNi * Rdens * Giv;
        Giv_b += tb * 4.0 * Di * XNi * Rdens * in.sice_m1;
        // Di = min(Di_a sqrt(XMi), Di_max)
        double XMi_b = 0.0;
        if (!Dihi) XMi_b += (sXM > 0.0) ? Di_b * P.Di_a * 0.5 / sXM : 0.0;
        // XMi = rhoqi / XNi
        double rhoqi_b = XMi_b / XNi;
        XNi_b -= XMi_b * XMi / XNi;
        // XNi = clamp(5.38e7 rhoqi^0.75, 1e3, 1e6)
        if (!XNhi && !XNlo) rhoqi_b += XNi_b * XN0 * 0.75 / rhoqi;
        if (rqi_pos) {
            dens_b += rhoqi_b * qi;
            q_b[QI] += rhoqi_b * dens;
        }
    }
    // Prevp0 = kev Rdens N0r sliq Glv ventr
    double ventr_b = 0.0;
    Rdens_b += Prevp0_b * kev * N0r * in.sliq_1m * Glv * ventr;
    b.sliq_1m += Prevp0_b * kev * Rdens * N0r * Glv * ventr;
    Glv_b += Prevp0_b * kev * Rdens * N0r * in.sliq_1m * ventr;
    ventr_b += Prevp0_b * kev * Rdens * N0r * in.sliq_1m * Glv;
    // ventr = f1r GAM_2 L2 + f2r srr GAM_5dr_h, srr = sqrt(((Cr rho_fact) / Nu) L5dr)
    L2_b += ventr_b * P.f1r * GAM_2;
    {
        const double srr_b = ventr_b * P.f2r * P.GAM_5dr_h;
        const double arg_b = (srr > 0.0) ? srr_b * 0.5 / srr : 0.0;
        tun_b[4] += arg_b * rho_fact / Nu * L5dr;
        rf_b += arg_b * Cr / Nu * L5dr;
        Nu_b -= arg_b * (Cr * rho_fact / Nu) * L5dr / Nu;
        L5dr_b += arg_b * (Cr * rho_fact / Nu);
    }
    // Gil = 1 / (LHF0 / (Ka temc)) : d/d(Ka temc) = 1 / LHF0
    Ka_b += Gil_b * temc / C_LHF0;
    temc_b += Gil_b * Ka / C_LHF0;
    // Giv = 1 / (ai1 ai2 + 1 / Di_), ai1 = LHS0 / (Ka temp), ai2 = LHS0 / (Rvap temp) - 1, Di_ = Kd dens qsati
    {
        const double den_b = -Giv_b * Giv * Giv;
        const double ai1_b = den_b * ai2, ai2_b = den_b * ai1;
        const double D_b = -den_b / (Di_ * Di_);
        Ka_b -= ai1_b * ai1 / Ka;
        temp_b -= ai1_b * ai1 / temp;
        temp_b -= ai2_b * C_LHS0 / (C_Rvap * temp * temp);
        Kd_b += D_b * dens * in.qsati;
        dens_b += D_b * Kd * in.qsati;
        b.qsati += D_b * Kd * dens;
    }
    {
        const double den_b = -Glv_b * Glv * Glv;
        const double av1_b = den_b * av2, av2_b = den_b * av1;
        const double D_b = -den_b / (Dv * Dv);
        Ka_b -= av1_b * av1 / Ka;
        temp_b -= av1_b * av1 / temp;
        temp_b -= av2_b * C_LHV0 / (C_Rvap * temp * temp);
        Kd_b += D_b * dens * in.qsatl;
        dens_b += D_b * Kd * in.qsatl;
        b.qsatl += D_b * Kd * dens;
    }
    // Nu = (nu0 + dnu_dT temc) Rdens ; Kd = (Kd0 + dKd_dT temc) Pstd / pre ; Ka = Ka0 + dKa_dT temc
    temc_b += Nu_b * dnu_dT * Rdens;
    Rdens_b += Nu_b * Nut;
    temc_b += Kd_b * dKd_dT * C_Pstd / in.pre;
    b.pre -= Kd_b * Kd / in.pre;
    temc_b += Ka_b * dKa_dT;
    // Pgaut0 = max(betas (qs - qscrt), 0)
    if (gaut_pos) {
        q_b[QS] += Pgaut0_b * betas;
        if (bsE) temc_b += Pgaut0_b * (qs - P.qscrt_gaut) * P.beta_gaut * ebs * P.gamma_gaut;
    }
    if (saut_pos) {
        q_b[QI] += Psaut0_b * betai;
        tun_b[1] -= Psaut0_b * betai;
        if (biE) temc_b += Psaut0_b * (qi - qicrt) * P.beta_saut * ebi * P.gamma_saut;
    }
    // Praut0 = swk Praut_kk + (1 - swk) Praut_be (Praut_kk alone under KK2000);
    // Praut_kk = kk qc^2.47 Nc^-1.79
    {
        const double K_b = kk_only ? Praut0_b : Praut0_b * swk;
        tun_b[0] += K_b * Eqc * nc_pm179;
        b.nc_pm179 += K_b * kk * Eqc;
        if (qc > 0.0) q_b[QC] += K_b * kk * nc_pm179 * 2.47 * Eqc / qc;
    }
    // Praut_be = Rdens 1.67e-5 rhoqc^2 / Bden, Bden = 5 + 3.66e-2 Nc / Dcq,
    // Dcq = Dc rhoqc + EPS, rhoqc = 1000 dens qc
    if (!kk_only) {
        const double nc_Nc = ncp ? ncp->Nc : P.Nc_def, nc_Dc = ncp ? ncp->Dc : P.Dc;
        const double B_b = Praut0_b * (1.0 - swk);
        const double num = Rdens * 1.67E-5 * rhoqc * rhoqc;
        Rdens_b += B_b * 1.67E-5 * rhoqc * rhoqc / Bden;
        double rhoqc_b = B_b * Rdens * 1.67E-5 * 2.0 * rhoqc / Bden;
        const double Bden_b = -B_b * num / (Bden * Bden);
        const double Dcq_b = -(Bden_b * 3.66E-2 * nc_Nc / (Dcq * Dcq));
        rhoqc_b += Dcq_b * nc_Dc;
        b.nc_n += Bden_b * 3.66E-2 / Dcq;
        b.nc_dc += Dcq_b * rhoqc;
        dens_b += rhoqc_b * 1000.0 * qc;
        q_b[QC] += rhoqc_b * 1000.0 * dens;
    }
    // Pgacs0 = As kac Rdens Egs_mod N0g |Vtg - Vts| SGS
    {
        const double k0 = As * kac * N0g;
        Rdens_b += Pgacs0_b * k0 * Egs_mod * aVgs * SGS;
        Egs_mod_b += Pgacs0_b * k0 * Rdens * aVgs * SGS;
        const double a_b = Pgacs0_b * k0 * Rdens * Egs_mod * SGS;
        Vtg_b += a_b * sVgs;
        Vts_b -= a_b * sVgs;
        const double S_b = Pgacs0_b * k0 * Rdens * Egs_mod * aVgs;
        M0bs_b += S_b * GAM_3 * G3;
        G3_b += S_b * M0bs * GAM_3;
        M1bs_b += S_b * 2.0 * GAM_2 * G2;
        G2_b += S_b * 2.0 * M1bs * GAM_2;
        M2bs_b += S_b * GAM * G1;
        G1_b += S_b * M2bs * GAM;
    }
    // Pracs0 = As kac Rdens Esr N0r |Vtr - Vts| SS
    {
        const double k0 = As * kac * N0r;
        Rdens_b += Pracs0_b * k0 * Esr * aVrs * SS;
        tun_b[3] += Pracs0_b * k0 * Rdens * aVrs * SS;
        const double a_b = Pracs0_b * k0 * Rdens * Esr * SS;
        Vtr_b += a_b * sVrs;
        Vts_b -= a_b * sVrs;
        const double S_b = Pracs0_b * k0 * Rdens * Esr * aVrs;
        M0bs_b += S_b * GAM_3 * L3;
        L3_b += S_b * M0bs * GAM_3;
        M1bs_b += S_b * 2.0 * GAM_2 * L2;
        L2_b += S_b * 2.0 * M1bs * GAM_2;
        M2bs_b += S_b * GAM * L1;
        L1_b += S_b * M2bs * GAM;
    }
    // Pgacr0 = Ar kac Rdens Egr N0g N0r |Vtg - Vtr| SG
    {
        const double k0 = Ar * kac * P.Egr * N0g * N0r;
        Rdens_b += Pgacr0_b * k0 * aVgr * SG;
        const double a_b = Pgacr0_b * k0 * Rdens * SG;
        Vtg_b += a_b * sVgr;
        Vtr_b -= a_b * sVgr;
        const double S_b = Pgacr0_b * k0 * Rdens * aVgr;
        L4_b += S_b * P.GAM_1br * GAM_3 * G3;
        G3_b += S_b * P.GAM_1br * L4 * GAM_3;
        L5_b += S_b * 2.0 * P.GAM_2br * GAM_2 * G2;
        G2_b += S_b * 2.0 * P.GAM_2br * L5 * GAM_2;
        L6_b += S_b * P.GAM_3br * GAM * G1;
        G1_b += S_b * P.GAM_3br * L6 * GAM;
    }
    // Psacr0 = Ar kac Rdens Esr N0r |Vtr - Vts| SR
    {
        const double k0 = Ar * kac * N0r;
        Rdens_b += Psacr0_b * k0 * Esr * aVrs * SR;
        tun_b[3] += Psacr0_b * k0 * Rdens * aVrs * SR;
        const double a_b = Psacr0_b * k0 * Rdens * Esr * SR;
        Vtr_b += a_b * sVrs;
        Vts_b -= a_b * sVrs;
        const double S_b = Psacr0_b * k0 * Rdens * Esr * aVrs;
        L4_b += S_b * P.GAM_1br * M2;
        M2_b += S_b * P.GAM_1br * L4;
        L5_b += S_b * 2.0 * P.GAM_2br * M1;
        M1_b += S_b * 2.0 * P.GAM_2br * L5;
        L6_b += S_b * P.GAM_3br * M0;
        M0_b += S_b * P.GAM_3br * L6;
    }
    // collection products
    {
        // Piacr0 = qi (Ar / mi) kac Eri N0r Cr GAM_6dr L6dr rho_fact
        const double k0 = Ar / P.mi * kac * P.Eri * N0r * P.GAM_6dr;
        q_b[QI] += Piacr0_b * k0 * Cr * L6dr * rho_fact;
        tun_b[4] += Piacr0_b * qi * k0 * L6dr * rho_fact;
        L6dr_b += Piacr0_b * qi * k0 * Cr * rho_fact;
        rf_b += Piacr0_b * qi * k0 * Cr * L6dr;
    }
    {
        // Pgaci0 = qi kac Egi N0g Cg GAM_3dg G3dg rho_fact
        const double k0 = kac * P.Egi * N0g * P.GAM_3dg;
        q_b[QI] += Pgaci0_b * k0 * Cg * G3dg * rho_fact;
        tun_b[6] += Pgaci0_b * qi * k0 * G3dg * rho_fact;
        G3dg_b += Pgaci0_b * qi * k0 * Cg * rho_fact;
        rf_b += Pgaci0_b * qi * k0 * Cg * G3dg;
    }
    {
        // Psaci0 = qi kac Esi_mod Cs M2ds rho_fact
        q_b[QI] += Psaci0_b * kac * Esi_mod * Cs * M2ds * rho_fact;
        Esi_mod_b += Psaci0_b * qi * kac * Cs * M2ds * rho_fact;
        tun_b[5] += Psaci0_b * qi * kac * Esi_mod * M2ds * rho_fact;
        M2ds_b += Psaci0_b * qi * kac * Esi_mod * Cs * rho_fact;
        rf_b += Psaci0_b * qi * kac * Esi_mod * Cs * M2ds;
    }
    {
        // Praci0 = qi kac Eri N0r Cr GAM_3dr L3dr rho_fact
        const double k0 = kac * P.Eri * N0r * P.GAM_3dr;
        q_b[QI] += Praci0_b * k0 * Cr * L3dr * rho_fact;
        tun_b[4] += Praci0_b * qi * k0 * L3dr * rho_fact;
        L3dr_b += Praci0_b * qi * k0 * Cr * rho_fact;
        rf_b += Praci0_b * qi * k0 * Cr * L3dr;
    }
    {
        // Pgacw0 = qc kac Egw N0g Cg GAM_3dg G3dg rho_fact
        const double k0 = kac * P.Egw * N0g * P.GAM_3dg;
        q_b[QC] += Pgacw0_b * k0 * Cg * G3dg * rho_fact;
        tun_b[6] += Pgacw0_b * qc * k0 * G3dg * rho_fact;
        G3dg_b += Pgacw0_b * qc * k0 * Cg * rho_fact;
        rf_b += Pgacw0_b * qc * k0 * Cg * G3dg;
    }
    {
        // Psacw0 = qc kac Esw Cs M2ds rho_fact
        const double k0 = kac * P.Esw;
        q_b[QC] += Psacw0_b * k0 * Cs * M2ds * rho_fact;
        tun_b[5] += Psacw0_b * qc * k0 * M2ds * rho_fact;
        M2ds_b += Psacw0_b * qc * k0 * Cs * rho_fact;
        rf_b += Psacw0_b * qc * k0 * Cs * M2ds;
    }
    // Pracw0 = swk Pracw_kk + (1 - swk) Pracw_or (Pracw_kk alone under KK2000);
    // Pracw_kk = 67 exp(1.15 log(qc qr + zcr)) (1 - zcr)
    {
        const double K_b = kk_only ? Pracw0_b : Pracw0_b * swk;
        if (zcr == 0.0 && Xcr > 0.0) {
            const double X_b = K_b * 67.0 * Ecr * 1.15 / Xcr;
            q_b[QC] += X_b * qr;
            q_b[QR] += X_b * qc;
        }
    }
    // Pracw_or = qc kac Erw N0r Cr GAM_3dr L3dr rho_fact
    if (!kk_only) {
        const double O_b = Pracw0_b * (1.0 - swk);
        const double k0 = kac * P.Erw * N0r * P.GAM_3dr;
        q_b[QC] += O_b * k0 * Cr * L3dr * rho_fact;
        tun_b[4] += O_b * qc * k0 * L3dr * rho_fact;
        L3dr_b += O_b * qc * k0 * Cr * rho_fact;
        rf_b += O_b * qc * k0 * Cr * L3dr;
    }
    // Esi_mod = min(Esi, Esi exp(gamma_sacr temc)); Egs_mod likewise (Egs is not tunable)
    if (EsiE) {
        tun_b[2] += Esi_mod_b * eEsi;
        temc_b += Esi_mod_b * Esi * eEsi * P.gamma_sacr;
    } else {
        tun_b[2] += Esi_mod_b;
    }
    if (EgsE) temc_b += Egs_mod_b * P.Egs * eEgs * P.gamma_gacs;
    // Pigen0 = max(min(Qi0 - qi, qv_qsati), 0) / dt ; Qi0 = icen E133 Rdens
    if (gpos) {
        const double m_b = Pigen0_b / dt;
        if (gselB) {
            b.qv_qsati += m_b;
        } else {
            const double Qi0_b = m_b;
            q_b[QI] -= m_b;
            tun_b[7] += Qi0_b * E133 * Rdens;
            Rdens_b += Qi0_b * icen * E133;
            if (!Nlow) {
                // E133 = Ni0^1.33, Ni0 = 1000 exp(-0.1 temc)
                const double Ni0_b = Qi0_b * icen * Rdens * 1.33 * E133 / Ni0;
                temc_b += Ni0_b * 1000.0 * eN * (-0.1);
            }
        }
    }
    // fall speeds
    {
        // Vti = (1 - sw_constVti) (-3.29) (dens qi + zi)^0.16 (1 - zi) + sw_constVti (-CONST_Vti)
        // ("4WATER" only), zi = [qi < 1e-8]
        if (vti_b != 0.0 && !(qi - 1.0E-8 < 0.0)) {
            const double dV = (1.0 - P.sw_constVti) * (-3.29) * 0.16 * exp(log(dens * qi) * (0.16 - 1.0));
            dens_b += vti_b * dV * qi;
            q_b[QI] += vti_b * dV * dens;
        }
        // Vtg = -Cg rho_fact kVg Gdg (evaluated as ((-Cg rho_fact) GAM_1bgdg / GAM_1bg) Gdg)
        tun_b[6] -= Vtg_b * rho_fact * kVg * Gdg;
        rf_b -= Vtg_b * Cg * kVg * Gdg;
        Gdg_b -= Vtg_b * Cg * rho_fact * kVg;
        // Vts = -Cs rho_fact RMV
        tun_b[5] -= Vts_b * rho_fact * RMV;
        rf_b -= Vts_b * Cs * RMV;
        const double RMV_b = -Vts_b * Cs * rho_fact;
        if (roh) {
            M2ds_b += RMV_b / Vden;
            M0bs_b -= RMV_b * RMV / Vden;
        } else {
            Lsds_b += RMV_b * P.GAM_1bsds / P.GAM_1bs;
        }
        // Vtr = -Cr rho_fact kVr Ldr
        tun_b[4] -= Vtr_b * rho_fact * kVr * Ldr;
        rf_b -= Vtr_b * Cr * kVr * Ldr;
        Ldr_b -= Vtr_b * Cr * rho_fact * kVr;
    }
    // graupel powers: G5dg = (G3 G2) Gdg, G3dg = G3 Gdg, G3 = G2 G1, G2 = G1 G1, Gdg = sqrt(G1)
    G3_b += G5dg_b * G2 * Gdg;
    G2_b += G5dg_b * G3 * Gdg;
    Gdg_b += G5dg_b * G3 * G2;
    G3_b += G3dg_b * Gdg;
    Gdg_b += G3dg_b * G3;
    G2_b += G3_b * G1;
    G1_b += G3_b * G2;
    G1_b += G2_b * 2.0 * G1;
    if (Gdg > 0.0) G1_b += Gdg_b * 0.5 / Gdg;
    if (zg == 0.0) {
        // G1 = sqrt(sqrt(Xg)), Xg = dens qg / Kg
        const double Xg_b = G1_b * 0.25 * G1 / Xg;
        dens_b += Xg_b * qg / Kg;
        q_b[QG] += Xg_b * dens / Kg;
    }
    // snow moments without Roh & Satoh: M = N0s GAM_. Ls^n (n = 1..6), M2ds = c Ls^3 Lsds,
    // M5dsh = c sqrt(Ls^5 Lsds), Lsds = Ls^(1/4), Ls = Xs^(1/4) (1 - zs)
    if (!roh) {
        const double dn = opaque(dens), qn = opaque(qs), N0s = P.N0s;
        const SnowSlope L = snow_slope(P, dn, qn, sw.zs);
        double Ls_b = M0_b * N0s * GAM + M1_b * N0s * GAM_2 * 2.0 * L.s + M2_b * N0s * GAM_3 * 3.0 * L.s2 +
                      M0bs_b * N0s * P.GAM_1bs * 4.0 * L.s3 + M1bs_b * N0s * P.GAM_2bs * 5.0 * L.s1bs +
                      M2bs_b * N0s * P.GAM_3bs * 6.0 * L.s2bs;
        Ls_b += M2ds_b * N0s * P.GAM_3ds * 3.0 * L.s2 * L.sds;
        Lsds_b += M2ds_b * N0s * P.GAM_3ds * L.s3;
        if (L.s5dsh > 0.0) {
            const double u_b = M5dsh_b * N0s * P.GAM_5ds_h * 0.5 / L.s5dsh;
            Ls_b += u_b * 5.0 * L.s1bs * L.sds;
            Lsds_b += u_b * L.s2bs;
        }
        if (L.s > 0.0) Ls_b += Lsds_b * 0.25 * L.sds / L.s;
        if (sw.zs == 0.0) {
            const double Xs_b = Ls_b * 0.25 * L.s / L.x;
            dens_b += Xs_b * qn * P.rK_s;
            q_b[QS] += Xs_b * dn * P.rK_s;
        }
        M0_b = M1_b = M2_b = M0bs_b = M1bs_b = M2bs_b = M2ds_b = M5dsh_b = 0.0;
    }
    // snow moments: M = m omz (m = exp(ln10 loga + lx b)), M2 = Xs2
    double Xs2_b = M2_b;
    double mom_b[7] = {M0_b * omz, M1_b * omz, M0bs_b * omz, M1bs_b * omz, M2bs_b * omz, M2ds_b * omz, M5dsh_b * omz};
    double lx_b = 0.0, at_b[4] = {0, 0, 0, 0}, bt_b[4] = {0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < 7; m++) {
        const double arg_b = mom_b[m] * mom[m];
        const double lga_b = arg_b * ln10, bb_b = arg_b * lx;
        lx_b += arg_b * bb[m];
        const double nm = nmv[m];
        if (m == 0) {
            at_b[0] += lga_b;
            bt_b[0] += bb_b;
        } else {
            at_b[0] += lga_b;
            at_b[1] += lga_b * nm;
            at_b[2] += lga_b * nm * nm;
            at_b[3] += lga_b * nm * nm * nm;
            bt_b[0] += bb_b;
            bt_b[1] += bb_b * nm;
            bt_b[2] += bb_b * nm * nm;
            bt_b[3] += bb_b * nm * nm * nm;
        }
    }
    // lx = log(Xs2 + zx)
    Xs2_b += lx_b / (Xs2 + zx);
    // at/bt polynomials in tems
    if (tcold) {
        double t_b = 0.0;
        t_b += at_b[0] * (ca[1] + tems * (2.0 * ca[4] + 3.0 * tems * ca[8]));
        t_b += at_b[1] * (ca[3] + 2.0 * tems * ca[6]);
        t_b += at_b[2] * ca[7];
        t_b += bt_b[0] * (cb[1] + tems * (2.0 * cb[4] + 3.0 * tems * cb[8]));
        t_b += bt_b[1] * (cb[3] + 2.0 * tems * cb[6]);
        t_b += bt_b[2] * cb[7];
        temc_b += t_b;
    }
    // Xs2 = dens qs / As
    dens_b += Xs2_b * qs / As;
    q_b[QS] += Xs2_b * dens / As;
    // rain powers
    L5_b += L5dr_b * Ldr;
    Ldr_b += L5dr_b * L5;
    L6_b += L6dr_b * Ldr;
    Ldr_b += L6dr_b * L6;
    L3_b += L3dr_b * Ldr;
    Ldr_b += L3dr_b * L3;
    L6_b += L7_b * L1;
    L1_b += L7_b * L6;
    L3_b += L6_b * 2.0 * L3;
    L4_b += L5_b * L1;
    L1_b += L5_b * L4;
    L2_b += L4_b * 2.0 * L2;
    L2_b += L3_b * L1;
    L1_b += L3_b * L2;
    L1_b += L2_b * 2.0 * L1;
    if (Ldr > 0.0) L1_b += Ldr_b * 0.5 / Ldr;
    if (zr == 0.0) {
        const double Xr_b = L1_b * 0.25 * L1 / Xr;
        dens_b += Xr_b * qr / Kr;
        q_b[QR] += Xr_b * dens / Kr;
    }
    // rho_fact = sqrt(DENS00 Rdens), Rdens = 1 / dens
    Rdens_b += rf_b * 0.5 * rho_fact / Rdens;
    dens_b -= Rdens_b * Rdens * Rdens;
    // dq_dt = q / dt
#pragma unroll
    for (int n = 0; n < 6; n++) q_b[n] += dqdt_b[n] / dt;
    // q = max(q_in, 0)
#pragma unroll
    for (int n = 0; n < 6; n++)
        if (qpos[n]) b.q[n] += q_b[n];
    b.dens += dens_b;
    b.temp += temp_b;
    b.temc += temc_b;
    b.sice_m1 += sice_b;
    stash_fence();
#pragma unroll
    for (int n = 1; n <= 5; n++) qt[n] = st[NRATE + 6 + n - 1];
}

// ---------------------------------------------------------------------------------
// Point-wise rate update (reverse of point_rates<double, double>)
// ---------------------------------------------------------------------------------
struct PointInB {
    double rhog, rhogq[6], tem, qv;
    double nc_pm179;      // cotangent of the point's Nc^-1.79
    double nc_n, nc_dc;   // cotangents of the point's Nc and Dc
};
struct PointOutB {
    double rhog, rhoge, rhogq[6], vt[3], vti;   // vt: QR, QS, QG; vti: QI ("4WATER")
};

// a1_tab/a2_tab: the Bergeron tables (read only without explicit ice generation).
template <bool DEFAULT_SCHEME>
NSW6_HD void point_rates_b(const DevParams &P, double dt, const PointIn<double> &x, const double tun[NCOEF],
                           const PointOutB &yb, PointInB &xb, double tun_b[NCOEF], const RateStash &st,
                           const NcPoint *ncp, const double *a1_tab, const double *a2_tab)
{
    // ---- forward preprocessing, reverse of the state update, the rate kernel
    double drq_b[NQ], rq_b[NQ], rhoge0_b, rhog_b;
    RateInB ib;
    zero_words(ib);
    {
#include "nsw6_point_fwd.inc"
        // ---- reverse (the forward tendencies qt come out of rates_b's recomputation)
        // rhoge = rhoge0 - LHV0 drq[QV] + LHF0 (drq[QI] + drq[QS] + drq[QG]); rq_out = rq + drq
#pragma unroll
        for (int n = 0; n < NQ; n++) drq_b[n] = yb.rhogq[n];
        drq_b[QV] -= yb.rhoge * C_LHV0;
        drq_b[QI] += yb.rhoge * C_LHF0;
        drq_b[QS] += yb.rhoge * C_LHF0;
        drq_b[QG] += yb.rhoge * C_LHF0;
#pragma unroll
        for (int n = 0; n < NQ; n++) rq_b[n] = yb.rhogq[n];
        rhoge0_b = yb.rhoge;
        rhog_b = yb.rhog;
        // drq = rhog qt dt, qt[QV] = -(qt[QC] + qt[QR] + qt[QI] + qt[QS] + qt[QG])
        double qt_b[NQ];
#pragma unroll
        for (int n = 0; n < NQ; n++) qt_b[n] = drq_b[n] * rhog * dt;
#pragma unroll
        for (int n = 1; n < NQ; n++) qt_b[n] -= qt_b[QV];
        // the rate kernel; the state-update cotangents wait in the stash meanwhile
#pragma unroll
        for (int n = 0; n < NQ; n++) {
            st[RATE_STASH + n] = drq_b[n];
            st[RATE_STASH + NQ + n] = rq_b[n];
        }
        st[RATE_STASH + 2 * NQ] = rhoge0_b;
        st[RATE_STASH + 2 * NQ + 1] = rhog_b;
        stash_fence();
        double qt[NQ];
        rates_b<DEFAULT_SCHEME>(P, dt, in, ncp, qt_b, yb.vt, yb.vti, ib, qt, st, a1_tab, a2_tab);
        stash_fence();
#pragma unroll
        for (int n = 0; n < NQ; n++) {
            drq_b[n] = st[RATE_STASH + n];
            rq_b[n] = st[RATE_STASH + NQ + n];
        }
        rhoge0_b = st[RATE_STASH + 2 * NQ];
        rhog_b = st[RATE_STASH + 2 * NQ + 1];
        qt[QV] = -(qt[QC] + qt[QR] + qt[QI] + qt[QS] + qt[QG]);
#pragma unroll
        for (int n = 0; n < NQ; n++) rhog_b += drq_b[n] * qt[n] * dt;
#pragma unroll
        for (int i = 0; i < NCOEF; i++) tun_b[i] += ib.tun[i];
        xb.nc_pm179 += ib.nc_pm179;
        xb.nc_n += ib.nc_n;
        xb.nc_dc += ib.nc_dc;
    }
    // ---- the preprocessing again, from inputs the compiler cannot see through
    const PointIn<double> xo = opaque_point(x);
    {
        const PointIn<double> &x = xo;
#include "nsw6_point_fwd.inc"
        // differences formed from the inputs
        double q_b[NQ];
#pragma unroll
        for (int n = 0; n < NQ; n++) q_b[n] = ib.q[n];
        double tem_b = ib.temp + ib.temc;
        double rho_b = ib.dens, pre_b = ib.pre, qsatl_b = ib.qsatl, qsati_b = ib.qsati;
        {
            const double qv = (q[QV] < 0.0) ? 0.0 : q[QV];
            const bool qvpos = !(q[QV] < 0.0);
            double qv_b = 0.0;
            // qv_qsati = qv - qsati
            qv_b += ib.qv_qsati;
            qsati_b -= ib.qv_qsati;
            // sice_m1 = qv / max(qsati, EPS) - 1
            const double dqi = dmax(qsati, C_EPS);
            qv_b += ib.sice_m1 / dqi;
            if (!(qsati < C_EPS)) qsati_b -= ib.sice_m1 * (qv / dqi) / dqi;
            // sliq_1m = 1 - min(qv / max(qsatl, EPS), 1)
            const double dql = dmax(qsatl, C_EPS);
            const double Sliq = qv / dql;
            if (!(1.0 < Sliq)) {
                qv_b -= ib.sliq_1m / dql;
                if (!(qsatl < C_EPS)) qsatl_b += ib.sliq_1m * Sliq / dql;
            }
            if (qvpos) q_b[QV] += qv_b;
        }
        // qsat = psat(tem) / (rho Rvap tem)
        {
            const double dpsl = (tem > TEM_MIN) ? psl * P.LovR_liq / (tem * tem) : 0.0;
            const double dpsi = (tem > TEM_MIN) ? psi * P.LovR_ice / (tem * tem) : 0.0;
            tem_b += qsatl_b * (dpsl / qden - qsatl / tem) + qsati_b * (dpsi / qden - qsati / tem);
            rho_b -= (qsatl_b * qsatl + qsati_b * qsati) / rho;
        }
        // pre = rho Rg tem, Rg = qd Rdry + q[QV] Rvap
        rho_b += pre_b * Rg * tem;
        tem_b += pre_b * rho * Rg;
        const double Rg_b = pre_b * rho * tem;
        double qd_b = Rg_b * C_Rdry;
        q_b[QV] += Rg_b * C_Rvap;
        // rhoge0 = tem rhog cva (cva does not depend on q: CVW = CVdry)
        tem_b += rhoge0_b * rhog * cva;
        rhog_b += rhoge0_b * tem * cva;
#pragma unroll
        for (int n = 1; n <= 5; n++) q_b[n] -= qd_b;
        // q[n] = rq[n] / rhog
#pragma unroll
        for (int n = 1; n <= 5; n++) {
            rq_b[n] += q_b[n] / rhog;
            rhog_b -= q_b[n] * q[n] / rhog;
        }
        xb.qv += q_b[QV];
        xb.tem += tem_b;
        rhog_b += rho_b / x.gsgam2;
        // rhog = rhog_in (1 - diffq2)
        xb.rhog += rhog_b * (1.0 - diffq2);
        const double diffq2_b = -rhog_b * x.rhog;
        // negative filter: a = rq0[QV] + sum (rq0[n] - clip(rq0[n])); rq[QV] = clip(a); diffq2 = a - clip(a)
        const double ma = pos[QV] ? 1.0 : 0.0;
        const double a_b = rq_b[QV] * ma + diffq2_b * (1.0 - ma);
        xb.rhogq[QV] += a_b;
        xb.rhogq[QC] += rq_b[QC];
#pragma unroll
        for (int n = 2; n <= 5; n++) {
            const double m = pos[n] ? 1.0 : 0.0;
            xb.rhogq[n] += rq_b[n] * m + a_b * (1.0 - m);
        }
    }
}


// The same with a stash of local (register or stack) storage.
NSW6_HD void point_rates_b(const DevParams &P, double dt, const PointIn<double> &x, const double tun[NCOEF],
                           const PointOutB &yb, PointInB &xb, double tun_b[NCOEF], const double *a1_tab,
                           const double *a2_tab)
{
    double buf[POINT_STASH];
    point_rates_b<false>(P, dt, x, tun, yb, xb, tun_b, RateStash{buf, 1}, nullptr, a1_tab, a2_tab);
}

}  // namespace nsw6
