// nsw6_adjoint_sweep.h -- reverse mode of one FP64 NSW6 step, as point-wise passes
// shared by the HIP kernels and the host check.
//
// The forward step is rebuilt up to the wall fluxes first (rates, kcell bounds,
// flux_wall, and flux_wall_general in the general regions).  Then
//   adj_cell   per cell c: sediment() and the tail recomputed, tail_b, sediment_b.
//              Writes the cell record: the cotangents cell c owes to the fluxes of
//              walls c-1..c+2, to rhog of cells c-1..c+1 and to rhogw of cells c, c+1,
//              and its own pre-sedimentation / momentum cotangents.
//   adj_wall   per wall w: the flux cotangent is gathered from the four cell records
//              that read the wall, plus the frain and surface-flux terms; then
//              wall_flux_b, the departure-distance reverse and derive_cell_b.  Writes
//              the wall record: cotangents of cell w-1 (rhog, QR, QS, QG), cells w
//              and w+1 (raw cell inputs) and of Vt(w-3..w+1).  In a general region
//              the records are left 0 and adj_walls_general (one thread per column,
//              walking its walls) adds the same cotangents into them.
//   adj_point  per point: gathers every record naming the point, then point_rates_b;
//              writes the input cotangents and returns the coefficient cotangents.
// Every record slot has one writer (one thread, for the general regions' wall
// records) and every gather only reads: no atomics, and the result does not depend on
// the order the points are visited in.
//
// Cotangents of the outputs are those of the mp_driver arrays after the step
// (state, rho, tem, pre, q, qd, surface precipitation and its energy fluxes, GPREC)
// plus the tracers' surface mass fluxes sflux_t = -f_t(kmin)/dt.  The diagnostics
// re_*, rctop, rwtop, tctop carry no cotangent.  Inputs that receive a cotangent:
// rhog, rhogvx, rhogvy, rhogvz, rhogw, rhoge (halo levels only: the interior rhoge
// is recomputed from tem), rhogq, vx, vy, vz, tem, q(QV), and the 8 tunable rate
// coefficients.
#pragma once

#include <stdint.h>

#include "nsw6_adjoint.h"
#include "nsw6_columns.h"

namespace nsw6 {

NSW6_HD double cot_at(const double *p, int64_t i) { return p ? p[i] : 0.0; }

// Record slot of level c of column p2 = l*K2 + ij, 0 outside the interior levels.  The
// records exist for every level (0 at the halos), so the load is made at c clamped to
// 0..KALL-1 without a branch and the result selected: the gathers of a pass then issue
// together instead of one round trip each.
NSW6_HD double record_at(const double *rec, int slot, int64_t N2, int64_t p2, int c, int ijdim)
{
    const int cc = (c < 0) ? 0 : ((c > KALL - 1) ? KALL - 1 : c);
    const double v = rec[slot * N2 + p2 + (int64_t)cc * ijdim];
    return (c >= KMIN && c <= KMAX) ? v : 0.0;
}

template <int NT>
NSW6_HD void adj_cell(const AdjArgs &A, int l, int k, int ij)
{
    using CR = CellRec<NT>;
    const StepArgs &a = A.a;
    const int ijdim = a.ijdim;
    const int64_t K2 = (int64_t)KALL * ijdim, N2 = (int64_t)a.lall * K2;
    const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
    const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;
    const bool interior = (k >= KMIN && k <= KMAX);
    double cr[CR::N], dws[NDW], din[NDI];
#pragma unroll
    for (int i = 0; i < CR::N; i++) cr[i] = 0.0;
#pragma unroll
    for (int i = 0; i < NDW; i++) dws[i] = 0.0;
#pragma unroll
    for (int i = 0; i < NDI; i++) din[i] = 0.0;

    {
        const Fields &C = A.cot;
        TailOut<double> yb;
        yb.rhog = cot_at(C.rhog, p);
        yb.rhoge = cot_at(C.rhoge, p);
        yb.rho = cot_at(C.rho, p);
        yb.tem = cot_at(C.tem, p);
        yb.pre = cot_at(C.pre, p);
        yb.qd = cot_at(C.qd, p);
#pragma unroll
        for (int n = 0; n < NQ; n++) {
            yb.rhogq[n] = cot_at(C.rhogq, pq + n * K2);
            yb.q[n] = cot_at(C.q, pq + n * K2);
        }
        TailIn<double> x;
        x.gsgam2 = a.f.gsgam2[p];
        x.rhogq[QV] = a.ws.rhogq[pq + QV * K2] + 0.0;
        x.rhogq[QC] = a.ws.rhogq[pq + QC * K2] + 0.0;
        if (NT == 3) x.rhogq[QI] = a.ws.rhogq[pq + QI * K2] + 0.0;
        TailInB xb;
        xb.rhog = xb.rhoge = 0.0;
#pragma unroll
        for (int n = 0; n < NQ; n++) xb.rhogq[n] = 0.0;
        if (interior) {
            SedInN<double, NT> si;
            load_sed<NT>(a, l, ij, k, si);
            const Unit u = load_unit(a.f, (int64_t)l * ijdim + ij);
            SedOutN<double, NT> so;
            sediment(*a.grid, k, u, si, so, a.kin2kin != 0);
            x.rhog = so.rg;
            x.rhoge = so.e;
#pragma unroll
            for (int t = 0; t < NT; t++) x.rhogq[sed_nq<NT>(t)] = so.rq[t];
            tail_b(*a.P, true, x, yb, xb);
            SedOutN<double, NT> sb;
            sb.rg = xb.rhog;
            sb.e = xb.rhoge;
#pragma unroll
            for (int t = 0; t < NT; t++) sb.rq[t] = xb.rhogq[sed_nq<NT>(t)];
            sb.rvx = cot_at(C.rhogvx, p);
            sb.rvy = cot_at(C.rhogvy, p);
            sb.rvz = cot_at(C.rhogvz, p);
            sb.rw = cot_at(C.rhogw, p);
            SedInBN<NT> s;
            zero(s);
            sediment_b<NT>(*a.grid, k, u, si, sb, s, a.kin2kin != 0);
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int t = 0; t < NT; t++) cr[CR::FQ + m * NT + t] = s.fq[m][t];
#pragma unroll
            for (int m = 0; m < 2; m++)
#pragma unroll
                for (int i = 0; i < 5; i++) cr[CR::FS + m * 5 + i] = s.fs[m][i];
#pragma unroll
            for (int m = 0; m < 3; m++)
#pragma unroll
                for (int i = 0; i < 2; i++) cr[CR::FH + m * 2 + i] = s.fh[m][i];
            cr[CR::RG + 0] = s.rg_m;
            cr[CR::RG + 1] = s.rg_c;
            cr[CR::RG + 2] = s.rg_p;
            cr[CR::RW + 0] = s.rw_c;
            cr[CR::RW + 1] = s.rw_p;
            dws[1] = s.re0;
            dws[2 + QV] = xb.rhogq[QV];
            dws[2 + QC] = xb.rhogq[QC];
            if (NT == 3) dws[2 + QI] = xb.rhogq[QI];
#pragma unroll
            for (int t = 0; t < NT; t++) dws[2 + sed_nq<NT>(t)] = s.rq[t];
            din[0] = s.rvx;
            din[1] = s.rvy;
            din[2] = s.rvz;
        } else {
            x.rhog = a.ws.rhog[p] + 0.0;
            x.rhoge = a.ws.rhoge[p] + 0.0;
#pragma unroll
            for (int t = 0; t < NT; t++) x.rhogq[sed_nq<NT>(t)] = a.ws.rhogq[pq + sed_nq<NT>(t) * K2] + 0.0;
            tail_b(*a.P, false, x, yb, xb);
            dws[0] = xb.rhog;
            dws[1] = xb.rhoge;
#pragma unroll
            for (int n = 0; n < NQ; n++) dws[2 + n] = xb.rhogq[n];
            // halo momentum passes through unchanged, except the top rhogw under KIN2KIN:
            // kin2kin_v with rhogkin_v = 0 there (0 when w^2/2 > 1e-20, else rhogw)
            din[0] = cot_at(C.rhogvx, p);
            din[1] = cot_at(C.rhogvy, p);
            din[2] = cot_at(C.rhogvz, p);
            din[3] = cot_at(C.rhogw, p);
            if (a.kin2kin && k == KALL - 1) {
                const double rhogh = kin2kin_top_rhogh<NT>(a, l, ij, x.rhog, FluxFromMemory<NT>(a, l, ij));
                double rhogh_b = 0.0, kinv_b = 0.0, rw_b = 0.0;
                kin2kin_v_b(rhogh, 0.0, a.f.rhogw[p] + 0.0, din[3], rhogh_b, kinv_b, rw_b);
                din[3] = rw_b;   // rhogh's cotangent is 0 on either side of the threshold
            }
        }
    }
#pragma unroll
    for (int i = 0; i < CR::N; i++) A.cr[i * N2 + p] = cr[i];
#pragma unroll
    for (int i = 0; i < NDW; i++) A.dws[i * N2 + p] = dws[i];
#pragma unroll
    for (int i = 0; i < NDI; i++) A.din[i * N2 + p] = din[i];
}

// CellInB in wall-record order
NSW6_HD void put_cell_b(const CellInB &b, double *o)
{
    o[0] = b.rhog;
#pragma unroll
    for (int n = 0; n < 5; n++) o[1 + n] = b.rq[n];
    o[6] = b.rhoge;
    o[7] = b.rvx;
    o[8] = b.rvy;
    o[9] = b.rvz;
    o[10] = b.rw;
    o[11] = b.vx;
    o[12] = b.vy;
    o[13] = b.vz;
}

// The cotangent of the fluxes of wall w (1..94) of column (l, ij), gathered from the
// interior cells that read them: cell w+1-m reads fq of wall w as its wall m
// (m = 0..3), cell w-m reads fs as its wall m (m = 0..1), cell w+1-m reads fh as its
// wall m (m = 0..2); plus frain = GPREC and, at wall kmin, the surface fluxes (ftb,
// the per-tracer fluxes before their sums).
template <int NT>
NSW6_HD void wall_flux_cot(const AdjArgs &A, int l, int w, int ij, WallFluxN<double, NT> &Fb, double ftb[NT][4])
{
    using CR = CellRec<NT>;
    const StepArgs &a = A.a;
    const int ijdim = a.ijdim;
    const int64_t K2 = (int64_t)KALL * ijdim, N2 = (int64_t)a.lall * K2;
    const int64_t p2 = (int64_t)l * K2 + ij;
    const int64_t pw = p2 + (int64_t)w * ijdim;
    const double dt = a.dt;
    auto CRV = [&](int slot, int c) { return record_at(A.cr, slot, N2, p2, c, ijdim); };
#pragma unroll
    for (int t = 0; t < NT; t++) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 4; m++) s += CRV(CR::FQ + m * NT + t, w + 1 - m);
        Fb.fq[t] = s;
    }
#pragma unroll
    for (int i = 0; i < 5; i++) Fb.fs[i] = CRV(CR::FS + i, w) + CRV(CR::FS + 5 + i, w - 1);
#pragma unroll
    for (int i = 0; i < 2; i++) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 3; m++) s += CRV(CR::FH + m * 2 + i, w + 1 - m);
        Fb.fh[i] = s;
    }
    // GPREC(w) = frain = -fq_QR(w) / dt
    Fb.fq[0] += -cot_at(A.cot.GPREC, pw) / dt;
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) ftb[t][i] = 0.0;
    if (w == KMIN) {
        const int64_t pc = (int64_t)l * ijdim + ij, pp = (int64_t)l * 2 * ijdim + ij;
        const double c_p0 = cot_at(A.cot.precip, pp), c_p1 = cot_at(A.cot.precip, pp + ijdim);
        const double c_lh = cot_at(A.cot.precip_lh_heat, pc), c_e = cot_at(A.cot.precip_rhoe, pc);
        const double c_phi = cot_at(A.cot.precip_rhophi, pc), c_kin = cot_at(A.cot.precip_rhokin, pc);
        // tracer 0 (QR) is liquid, the others (QI under "4WATER", QS, QG) solid
#pragma unroll
        for (int t = 0; t < NT; t++) {
            ftb[t][0] = (t == 0) ? -c_p0 / dt : (-c_p1 / dt + c_lh * C_LHF0 / dt);
            ftb[t][1] = -c_e / dt;
            ftb[t][2] = -c_phi / dt;
            ftb[t][3] = -c_kin / dt;
            if (A.cot_sflux) Fb.fq[t] += -A.cot_sflux[((int64_t)l * NT + t) * ijdim + ij] / dt;
        }
    }
}

template <int NT>
NSW6_HD void adj_wall(const AdjArgs &A, int l, int w, int ij)
{
    using WRr = WallRec<NT>;
    const StepArgs &a = A.a;
    const int ijdim = a.ijdim;
    const int64_t K2 = (int64_t)KALL * ijdim, N2 = (int64_t)a.lall * K2;
    const int64_t p2 = (int64_t)l * K2 + ij;
    const int64_t pw = p2 + (int64_t)w * ijdim;
    double wr[WRr::N];
#pragma unroll
    for (int i = 0; i < WRr::N; i++) wr[i] = 0.0;

    if (a.ws.region_flag[l] == 0 && w >= KMIN && w <= KMAX) {
        const GridDev &g = *a.grid;
        WallFluxN<double, NT> Fb;
        double ftb[NT][4];
        wall_flux_cot<NT>(A, l, w, ij, Fb, ftb);
        WallStateN<NT> s;
        wall_state<NT>(a, l, w, ij, s);
        CellBN<NT> cb0, cb1;
        zero(cb0);
        zero(cb1);
        double z0b[NT], zh0b[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) z0b[t] = zh0b[t] = 0.0;
        wall_flux_b<NT>(g, w, s.cur, s.nxt, s.z0, s.zh0, s.geo, Fb, ftb, cb0, cb1, z0b, zh0b);
#pragma unroll
        for (int t = 0; t < NT; t++) {
            double Vb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            zdis_full_b(g, a.dt, a.dt2, a.dt3_6, s.V[t], w, z0b[t], Vb);
            if (w >= KMIN + 1) zdis_half_b(g, a.dt, a.dt2, a.dt3_6, s.V[t], w, zh0b[t], Vb);
#pragma unroll
            for (int d = 0; d < 5; d++) wr[WRr::V + t * 5 + d] = Vb[d];
        }
        CellInB rb0, rb1;
        zero(rb0);
        zero(rb1);
        double brg_b = 0.0, bq_b[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) bq_b[t] = 0.0;
        derive_cell_b<NT>(g, w + 1, s.r1, s.cur.rhog, s.cur.q, s.cur.gs, s.u, s.nxt, cb1, rb1, cb0.rhog, cb0.q);
        derive_cell_b<NT>(g, w, s.r0, s.m.rhog, s.m.q, s.m.gs, s.u, s.cur, cb0, rb0, brg_b, bq_b);
        // q_m[t] = rhogq_t(w-1) / rhog(w-1)
        double rgm_b = brg_b;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            wr[WRr::M + 1 + t] = bq_b[t] / s.m.rhog;
            rgm_b -= bq_b[t] * s.m.q[t] / s.m.rhog;
        }
        wr[WRr::M + 0] = rgm_b;
        put_cell_b(rb0, wr + WRr::C0);
        put_cell_b(rb1, wr + WRr::C1);
    }
#pragma unroll
    for (int i = 0; i < WRr::N; i++) A.wr[i * N2 + pw] = wr[i];
}

// Reverse of the wall fluxes of a general region (wall_flux_general), for one column:
// one thread walks the column's walls 1..94.  A wall's fluxes read the swept cells of
// the region's kcell ranges, any number of levels away, so their cotangents cannot go
// to fixed record slots of the wall as adj_wall's do; instead the thread adds them into
// its column's wall records (which adj_wall left 0 for this region and nothing else
// writes): the raw-input cotangent of cell j to the cell-w slots of wall j, the part
// of cell j-1 that cell j's lower wall reads (rhog, q) to those of wall j-1 (to the
// cell-(w-1) slots of wall 1 for the halo cell 0), and the cotangent of Vt(j) to the
// d = 3 slot of wall j.  adj_point gathers them like any others.
template <int NT>
NSW6_HD void adj_walls_general(const AdjArgs &A, int l, int ij)
{
    using WRr = WallRec<NT>;
    const StepArgs &a = A.a;
    if (a.ws.region_flag[l] == 0) return;
    const int ijdim = a.ijdim;
    const GridDev &g = *a.grid;
    const int64_t K2 = (int64_t)KALL * ijdim, N2 = (int64_t)a.lall * K2;
    const int64_t p2 = (int64_t)l * K2 + ij, p3 = (int64_t)l * NQ * K2 + ij;
    const int64_t pv = (int64_t)l * NT * K2 + ij;
    const int *bnd = a.ws.bounds + (size_t)l * ext_per_region<NT>();
    const Unit u = load_unit(a.f, (int64_t)l * ijdim + ij);
    auto WR = [&](int slot, int w) -> double & { return A.wr[slot * N2 + p2 + (int64_t)w * ijdim]; };
    auto derive = [&](int j, CellIn<double> &r, BelowCellN<NT> &bl, CellN<double, NT> &c) {
        load_cell(a, p2, p3, K2, ijdim, j, r);
        bl = load_below<NT>(a, p2, p3, K2, ijdim, j);
        derive_cell(g, j, r, bl.rhog, bl.q, bl.gs, u, c);
    };
    // cell j's cotangent cb through derive_cell_b into the column's wall records
    auto put_cell = [&](int j, const CellIn<double> &r, const BelowCellN<NT> &bl, const CellN<double, NT> &c,
                        const CellBN<NT> &cb) {
        CellInB rb;
        zero(rb);
        double brg_b = 0.0, bq_b[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) bq_b[t] = 0.0;
        derive_cell_b<NT>(g, j, r, bl.rhog, bl.q, bl.gs, u, c, cb, rb, brg_b, bq_b);
        double o[14];
        put_cell_b(rb, o);
#pragma unroll
        for (int i = 0; i < 14; i++) WR(WRr::C0 + i, j) += o[i];
        // q_m[t] = rhogq_t(j-1) / rhog(j-1)
        double rgm_b = brg_b, rqm_b[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            rqm_b[t] = bq_b[t] / bl.rhog;
            rgm_b -= bq_b[t] * bl.q[t] / bl.rhog;
        }
        if (j - 1 >= KMIN) {
            WR(WRr::C0 + 0, j - 1) += rgm_b;
#pragma unroll
            for (int t = 0; t < NT; t++) WR(WRr::C0 + sed_nq<NT>(t), j - 1) += rqm_b[t];
        } else {
            WR(WRr::M + 0, KMIN) += rgm_b;
#pragma unroll
            for (int t = 0; t < NT; t++) WR(WRr::M + 1 + t, KMIN) += rqm_b[t];
        }
    };
    for (int w = KMIN; w <= KMAX; w++) {
        WallFluxN<double, NT> Fb;
        double ftb[NT][4];
        wall_flux_cot<NT>(A, l, w, ij, Fb, ftb);
        double V[NT][5];
        int kcmin[NT], kcmax[NT], khmin[NT], khmax[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            load_vt_window(a.ws.vt + pv + t * K2, ijdim, w, V[t]);
            kcmin[t] = bnd[bounds_index(t, 0, w, 0)];
            kcmax[t] = bnd[bounds_index(t, 0, w, 1)];
            khmin[t] = bnd[bounds_index(t, 1, w, 0)];
            khmax[t] = bnd[bounds_index(t, 1, w, 1)];
        }
        GeneralGeomN<double, NT> G;
        general_geometry(g, w, a.dt, a.dt2, a.dt3_6, V, kcmin, kcmax, khmin, khmax, G);
        double acc[NT][6], rfc[NT][6], acch[NT][2], hvc[NT][2];
        general_sums(g, w, G, [&](int j, CellN<double, NT> &c) {
            CellIn<double> r;
            BelowCellN<NT> bl;
            derive(j, r, bl, c);
        }, acc, rfc, acch, hvc);
        const bool half = (w >= KMIN + 1);
        // cotangents of the sums before the |f| < EPS filter, and of zdis0
        double vb[NT][6], vhb[NT][2], zd[NT], zh[NT], z0b[NT], zh0b[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            zd[t] = G.z0[t] - general_zd_shift(g.dgz, w, G.kcmin[t], G.kcmax[t], G.kc[t]);
            z0b[t] = 0.0;
#pragma unroll
            for (int i = 0; i < 6; i++) {
                double ft_b = (i == 0) ? Fb.fq[t] : Fb.fs[i - 1];
                if (i < 4) ft_b += ftb[t][i];
                vb[t][i] = ft_b * eps_mask(acc[t][i] + rfc[t][i] * zd[t]);
                z0b[t] += vb[t][i] * rfc[t][i];
            }
            zh[t] = half ? G.zh0[t] - general_zd_shift(g.dgzh, w, G.khmin[t], G.kcmax[t], G.kch[t]) : 0.0;
            zh0b[t] = 0.0;
#pragma unroll
            for (int i = 0; i < 2; i++) {
                vhb[t][i] = half ? Fb.fh[i] * eps_mask(acch[t][i] + hvc[t][i] * zh[t]) : 0.0;
                zh0b[t] += vhb[t][i] * hvc[t][i];
            }
        }
        // the swept cells of the ranges
        for (int j = G.lo; j <= G.hi; j++) {
            CellIn<double> r;
            BelowCellN<NT> bl;
            CellN<double, NT> c;
            derive(j, r, bl, c);
            CellBN<NT> cb;
            zero(cb);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const bool one = (G.kcmin[t] == w && G.kcmax[t] == w);
                const bool in = !one && j >= G.kcmin[t] && j <= G.kcmax[t];
                const double fa = in ? fact_of(j, G.kc[t], w, g.dgz[j]) : 0.0;
                const double wk = (j == G.kc[t]) ? zd[t] : 0.0;
#pragma unroll
                for (int i = 0; i < 6; i++) rhof_b(c, t, i, vb[t][i] * fa + vb[t][i] * wk, cb);
                if (half) {
                    const bool hone = (G.khmin[t] == w && G.kcmax[t] == w);
                    const bool hin = !hone && j >= G.khmin[t] && j <= G.kcmax[t];
                    const double fh = hin ? fact_of(j, G.kch[t], w, g.dgzh[j]) : 0.0;
                    const double wh = (j == G.kch[t]) ? zh[t] : 0.0;
#pragma unroll
                    for (int i = 0; i < 2; i++) cb.hv[t][i] += vhb[t][i] * fh + vhb[t][i] * wh;
                }
            }
            put_cell(j, r, bl, c, cb);
        }
        // departure cells the search left at the domain's end outside the ranges
#pragma unroll
        for (int t = 0; t < NT; t++) {
            if (G.kc[t] < G.lo || G.kc[t] > G.hi) {
                CellIn<double> r;
                BelowCellN<NT> bl;
                CellN<double, NT> c;
                derive(G.kc[t], r, bl, c);
                CellBN<NT> cb;
                zero(cb);
#pragma unroll
                for (int i = 0; i < 6; i++) rhof_b(c, t, i, vb[t][i] * zd[t], cb);
                put_cell(G.kc[t], r, bl, c, cb);
            }
            if (half && (G.kch[t] < G.lo || G.kch[t] > G.hi)) {
                CellIn<double> r;
                BelowCellN<NT> bl;
                CellN<double, NT> c;
                derive(G.kch[t], r, bl, c);
                CellBN<NT> cb;
                zero(cb);
                cb.hv[t][0] = vhb[t][0] * zh[t];
                cb.hv[t][1] = vhb[t][1] * zh[t];
                put_cell(G.kch[t], r, bl, c, cb);
            }
        }
        // the fall speeds Vt(w-3..w+1) through the departure distances
#pragma unroll
        for (int t = 0; t < NT; t++) {
            double Vb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            zdis_full_b(g, a.dt, a.dt2, a.dt3_6, V[t], w, z0b[t], Vb);
            if (half) zdis_half_b(g, a.dt, a.dt2, a.dt3_6, V[t], w, zh0b[t], Vb);
#pragma unroll
            for (int d = 0; d < 5; d++) {
                const int j = w + d - 3;
                if (j >= KMIN && j <= KMAX) WR(WRr::V + t * 5 + 3, j) += Vb[d];
            }
        }
    }
}

// Input cotangents of point (l, k, ij); the coefficient cotangents of the point are
// added to tun_b.
NSW6_HD void store_point_grad(const Fields &G, int64_t p, int64_t pq, int64_t K2, double g_rhog, double g_rhoge,
                              const double g_rhogq[NQ], double g_tem, double g_qv, double gvx, double gvy, double gvz,
                              double gw, double gux, double guy, double guz, double g_unccn)
{
    if (G.rhog) G.rhog[p] = g_rhog;
    if (G.unccn) G.unccn[p] = g_unccn;
    if (G.rhoge) G.rhoge[p] = g_rhoge;
    if (G.tem) G.tem[p] = g_tem;
    if (G.rhogvx) G.rhogvx[p] = gvx;
    if (G.rhogvy) G.rhogvy[p] = gvy;
    if (G.rhogvz) G.rhogvz[p] = gvz;
    if (G.rhogw) G.rhogw[p] = gw;
    if (G.vx) G.vx[p] = gux;
    if (G.vy) G.vy[p] = guy;
    if (G.vz) G.vz[p] = guz;
#pragma unroll
    for (int n = 0; n < NQ; n++) {
        if (G.rhogq) G.rhogq[pq + n * K2] = g_rhogq[n];
        if (G.q) G.q[pq + n * K2] = (n == QV) ? g_qv : 0.0;
    }
}

// adj_point of a point; st holds the rate reverse's stash.
// INDIR: OPT_INDIR (the point's droplet number from unccn; A.a.opt_indir).  DEFAULT_SCHEME:
// KK2000 alone, explicit ice generation and Roh & Satoh snow (A.a.default_scheme);
// otherwise sw_kk2000, sw_expice and sw_roh2014 are read at run time.
template <int NT, bool INDIR, bool DEFAULT_SCHEME>
NSW6_HD void adj_point_single(const AdjArgs &A, int l, int k, int ij, double tun_b[NCOEF], const RateStash &st)
{
    using CR = CellRec<NT>;
    using WRr = WallRec<NT>;
    const StepArgs &a = A.a;
    const int ijdim = a.ijdim;
    const int64_t K2 = (int64_t)KALL * ijdim, N2 = (int64_t)a.lall * K2;
    const int64_t p2 = (int64_t)l * K2 + ij;
    const int64_t p = p2 + (int64_t)k * ijdim;
    const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;
    const Fields &f = a.f;
    auto CRV = [&](int slot, int c) { return record_at(A.cr, slot, N2, p2, c, ijdim); };
    auto WRV = [&](int slot, int w) { return record_at(A.wr, slot, N2, p2, w, ijdim); };
    // cell k is cell c (own record), c-1 of cell k+1, c+1 of cell k-1; cell w-1 of
    // wall k+1, cell w of wall k, cell w+1 of wall k-1
    const double rg_b = A.dws[0 * N2 + p] + CRV(CR::RG + 1, k) + CRV(CR::RG + 0, k + 1) + CRV(CR::RG + 2, k - 1) +
                        WRV(WRr::M + 0, k + 1) + WRV(WRr::C0 + 0, k) + WRV(WRr::C1 + 0, k - 1);
    const double re_b = A.dws[1 * N2 + p] + WRV(WRr::C0 + 6, k) + WRV(WRr::C1 + 6, k - 1);
    double rq_b[NQ];
#pragma unroll
    for (int n = 0; n < NQ; n++) {
        rq_b[n] = A.dws[(2 + n) * N2 + p];
        if (n >= 1) rq_b[n] += WRV(WRr::C0 + n, k) + WRV(WRr::C1 + n, k - 1);
    }
#pragma unroll
    for (int t = 0; t < NT; t++) rq_b[sed_nq<NT>(t)] += WRV(WRr::M + 1 + t, k + 1);
    // Vt(k) is Vt(w + d - 3) of wall w = k + 3 - d, d = 0..4
    double vt_b[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        double s = 0.0;
#pragma unroll
        for (int d = 0; d < 5; d++) s += WRV(WRr::V + t * 5 + d, k + 3 - d);
        vt_b[t] = s;
    }
    const double gvx = A.din[0 * N2 + p] + WRV(WRr::C0 + 7, k) + WRV(WRr::C1 + 7, k - 1);
    const double gvy = A.din[1 * N2 + p] + WRV(WRr::C0 + 8, k) + WRV(WRr::C1 + 8, k - 1);
    const double gvz = A.din[2 * N2 + p] + WRV(WRr::C0 + 9, k) + WRV(WRr::C1 + 9, k - 1);
    const double gw = A.din[3 * N2 + p] + CRV(CR::RW + 0, k) + CRV(CR::RW + 1, k - 1) + WRV(WRr::C0 + 10, k) +
                      WRV(WRr::C1 + 10, k - 1);
    const double gux = WRV(WRr::C0 + 11, k) + WRV(WRr::C1 + 11, k - 1);
    const double guy = WRV(WRr::C0 + 12, k) + WRV(WRr::C1 + 12, k - 1);
    const double guz = WRV(WRr::C0 + 13, k) + WRV(WRr::C1 + 13, k - 1);

    double g_rhog, g_rhoge, g_rhogq[NQ], g_tem = 0.0, g_qv = 0.0, g_unccn = 0.0;
    if (k >= KMIN && k <= KMAX) {
        const DevParams &P = *a.P;
        PointIn<double> x;
        x.rhog = f.rhog[p];
#pragma unroll
        for (int n = 0; n < NQ; n++) x.rhogq[n] = f.rhogq[pq + n * K2];
        x.tem = f.tem[p];
        x.qv = f.q[pq];
        x.gsgam2 = f.gsgam2[p];
        double tun[NCOEF];
        tunables(P, tun);
        PointOutB yb;
        yb.rhog = rg_b;
        yb.rhoge = re_b;
#pragma unroll
        for (int n = 0; n < NQ; n++) yb.rhogq[n] = rq_b[n];
        // the rate kernel's fall speeds: QR, QS, QG and, under "4WATER", QI (tracer 1)
        yb.vt[0] = vt_b[0];
        yb.vt[1] = vt_b[NT - 2];
        yb.vt[2] = vt_b[NT - 1];
        yb.vti = (NT == 4) ? vt_b[1] : 0.0;
        PointInB xb;
        xb.rhog = xb.tem = xb.qv = xb.nc_pm179 = xb.nc_n = xb.nc_dc = 0.0;
#pragma unroll
        for (int n = 0; n < NQ; n++) xb.rhogq[n] = 0.0;
        // OPT_INDIR: Nc = max(unccn 1e-6, Nc_def) of the point (src/mod_mp_nsw6.cpp:475-484),
        // Nc^-1.79 and Dc = 0.146 - 5.964e-2 log(Nc / 2000) from it
        NcPoint nc = nc_default(P);
        double dnc_dunccn = 0.0;
        if (INDIR) {
            const double un = f.unccn[p];
            nc = nc_indirect(P, un);
            if (!(un * 1.0E-6 < P.Nc_def)) dnc_dunccn = 1.0E-6;
        }
        point_rates_b<DEFAULT_SCHEME>(P, a.dt, x, tun, yb, xb, tun_b, st, INDIR ? &nc : nullptr, A.a1_tab, A.a2_tab);
        if (INDIR) {
            double nc_b = xb.nc_pm179 * (-1.79 * nc.pow_m179 / nc.Nc);
            if (!DEFAULT_SCHEME) nc_b += xb.nc_n + xb.nc_dc * (-5.964E-2 / nc.Nc);
            g_unccn = nc_b * dnc_dunccn;
        }
        g_rhog = xb.rhog;
        g_rhoge = 0.0;   // the interior rhoge is recomputed from tem
#pragma unroll
        for (int n = 0; n < NQ; n++) g_rhogq[n] = xb.rhogq[n];
        g_tem = xb.tem;
        g_qv = xb.qv;
    } else {
        // halo levels: the pre-sedimentation state is the input itself
        g_rhog = rg_b;
        g_rhoge = re_b;
#pragma unroll
        for (int n = 0; n < NQ; n++) g_rhogq[n] = rq_b[n];
    }
    store_point_grad(A.grad, p, pq, K2, g_rhog, g_rhoge, g_rhogq, g_tem, g_qv, gvx, gvy, gvz, gw, gux, guy, guz,
                     g_unccn);
}

template <int NT, bool INDIR, bool DEFAULT_SCHEME>
NSW6_HD void adj_point(const AdjArgs &A, int l, int k, int ij, double tun_b[NCOEF], const RateStash &st)
{
    adj_point_single<NT, INDIR, DEFAULT_SCHEME>(A, l, k, ij, tun_b, st);
}
template <int NT>
NSW6_HD void adj_point(const AdjArgs &A, int l, int k, int ij, double tun_b[NCOEF])
{
    double buf[POINT_STASH];
    if (A.a.opt_indir)
        adj_point<NT, true, false>(A, l, k, ij, tun_b, RateStash{buf, 1});
    else
        adj_point<NT, false, false>(A, l, k, ij, tun_b, RateStash{buf, 1});
}

}  // namespace nsw6
