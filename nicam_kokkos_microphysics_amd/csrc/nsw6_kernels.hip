// nsw6_kernels.hip -- HIP kernels of the NSW6 microphysics step for gfx950 (MI355X).
//
// One step = six launches on one stream:
//   k_rates    point-wise over (l,k,ij): negative filter, psat/qsat, Bergeron, the
//              49 process rates and the state update (src/mod_mp_nsw6.cpp:439-589);
//              writes the pre-sedimentation state and the fall speeds to workspace.
//   k_extremes point-wise departure distances zdis of the full and half levels for
//              QR, QS, QG (src/mod_vadv1d.cpp:270-302) reduced per block to partial
//              per-level max/min (:317-327).
//   k_bounds   one thread per (region, tracer, level type, level): kcell_min/max
//              search (:329-351) and the per-region choice of transport path.
//   k_flux     point-wise over walls of the single-cell regions (every departure
//              cell in {k, k+1}): closed-form semi-Lagrangian fluxes (:393-450),
//              frain and the surface fluxes (src/mod_precip_transport.cpp:573-958).
//   k_sweep    one lane per column, bottom to top, for the general regions only
//              (departures beyond one cell, upward, or out of the domain).
//   k_finish   point-wise: flux divergences and KIN2EIN (:960-1011) of the
//              single-cell regions, rediagnosis (:1074-1086), saturation adjustment
//              (src/mod_satadjust.cpp:274-362,771-896) and the final negative filter.
// Every access of a wavefront is one contiguous 512-byte segment of a level (ij
// fastest).  No MFMA: nothing here is a dense contraction.
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "nsw6_device.h"
#include "nsw6_columns.h"
#include "nsw6_internal.h"
#include "nsw6_rates_kernel.h"
#include "nsw6_step.h"

using namespace nsw6;

// Cloud top with the semantics of src/mod_mp_nsw6.cpp:1575-1582's comments: the
// highest level whose cloud switch fired in k_rates gives rctop = rceff and tctop = T
// there (UNDEF without one); rwtop as TagUpdateCLD (:717-720).  Runs after k_rates
// and before k_finish rewrites tem, and clears the level for the next step.
__global__ void __launch_bounds__(256) k_cloudtop(StepArgs a)
{
    const int l = blockIdx.y;
    const int ij = blockIdx.x * 256 + threadIdx.x;
    if (ij >= a.ijdim) return;
    const int64_t pc = (int64_t)l * a.ijdim + ij;
    const int top = a.ws.cloud_top[pc];
    double rct = C_UNDEF, tct = C_UNDEF;
    if (top >= KMIN && top <= KMAX) {
        const int64_t p = ((int64_t)l * KALL + top) * a.ijdim + ij;
        rct = a.f.re_liquid[p];
        tct = a.f.tem[p];
    }
    a.f.rctop[pc] = rct;
    a.f.tctop[pc] = tct;
    const double sw = step_pos(tct - C_TEM00);
    a.f.rwtop[pc] = (sw)*rct + (1.0 - sw) * C_UNDEF;
    a.ws.cloud_top[pc] = -1;
}

// zdis_interior / zdis_bottom and the point-wise wall speeds are in nsw6_step.h.

// Sliding window over the fall speed Vt(k) of one tracer that produces, level by
// level, wh/zdis of the full level (mkmin=1, mkmax=94) and of the half level
// (Vtermh, mkmin=2, mkmax=94) exactly as vadv1d_prep builds them.
struct WindWindow {
    double vt_k;          // Vt(k)
    double wh_km1, wh_k;  // wh(k-1), wh(k)          (full level)
    double vh_k;          // Vtermh(k)
    double whh_km1, whh_k;// wh_h(k-1), wh_h(k)      (half level)
};
// Initialise with cell 0 and cell 1 known (Vt(0) = 0): state for k = 1.
__device__ __forceinline__ void wind_init(WindWindow &w, double vt1)
{
    w.vt_k = vt1;
    w.wh_km1 = 0.0;   // wh(0) = Vt(0) = 0
    w.wh_k = vt1;     // wh(1) = Vt(1)
    w.vh_k = 0.0;     // Vtermh(1) = 0
    w.whh_km1 = 0.0;  // wh_h(0) = 0
    w.whh_k = 0.0;    // wh_h(1) = Vtermh(1) = 0
}
// Given Vt(k+1) (0 for k+1 = 95), produce wh(k+1) and wh_h(k+1) for level k.
__device__ __forceinline__ void wind_next(const WindWindow &w, int k, double vt_kp1, double &wh_kp1,
                                          double &vh_kp1, double &whh_kp1)
{
    const int kp = k + 1;
    wh_kp1 = (kp <= KMAX) ? 0.5 * (w.vt_k + vt_kp1) : w.vt_k;            // wh(95) = Vt(94)
    vh_kp1 = (kp >= 2 && kp <= KMAX - 1) ? 0.5 * (vt_kp1 + w.vt_k) : 0.0; // Vtermh
    whh_kp1 = (kp >= 3 && kp <= KMAX) ? 0.5 * (w.vh_k + vh_kp1) : ((kp == 2) ? vh_kp1 : 0.0);
}
__device__ __forceinline__ void wind_shift(WindWindow &w, double vt_kp1, double wh_kp1, double vh_kp1,
                                           double whh_kp1)
{
    w.vt_k = vt_kp1;
    w.wh_km1 = w.wh_k;
    w.wh_k = wh_kp1;
    w.vh_k = vh_kp1;
    w.whh_km1 = w.whh_k;
    w.whh_k = whh_kp1;
}
__device__ __forceinline__ double zdis_full(const GridDev &g, double dt, double dt2, double dt3_6,
                                            const WindWindow &w, int k, double wh_kp1)
{
    return (k == KMIN) ? zdis_bottom(dt, dt2, w.wh_k, wh_kp1, g.rdgz[KMIN])
                       : zdis_interior(dt, dt2, dt3_6, w.wh_km1, w.wh_k, wh_kp1, g.rdgz[k - 1], g.rdgz[k],
                                       g.rsum_gz[k]);
}
__device__ __forceinline__ double zdis_half(const GridDev &g, double dt, double dt2, double dt3_6,
                                            const WindWindow &w, int k, double whh_kp1)
{
    return (k == KMIN + 1) ? zdis_bottom(dt, dt2, w.whh_k, whh_kp1, g.rdgzh[KMIN + 1])
                           : zdis_interior(dt, dt2, dt3_6, w.whh_km1, w.whh_k, whh_kp1, g.rdgzh[k - 1],
                                           g.rdgzh[k], g.rsum_gzh[k]);
}
template <int N>
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = N / 2; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o, N));
    return v;
}
template <int N>
__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = N / 2; o > 0; o >>= 1) v = dmin(v, __shfl_xor(v, o, N));
    return v;
}

// ---------------------------------------------------------------------------------
// k_extremes: per-level max/min over the columns of a region of the departure
// distances (src/mod_vadv1d.cpp:317-327).  One block per (column chunk, level k in
// 1..94, region); each thread takes EXT_COLS_PER_THREAD columns strided by the block
// size.  The block's 12 extremes (3 tracers x {full, half} x {max, min}) go to the
// partials array [l][k][4 NT][block] that k_bounds reduces (NT = 3 or 4 tracers).
// ---------------------------------------------------------------------------------
template <int NT>
__global__ void __launch_bounds__(EXT_THREADS) k_extremes(StepArgs a)
{
    const int ijdim = a.ijdim;
    const int nb = (ijdim + EXT_COLS_PER_BLOCK - 1) / EXT_COLS_PER_BLOCK;
    const LevelBlock lb = level_block(blockIdx.x, nb, KMAX - KMIN + 1, KMIN, a.lall);
    if (!lb.valid) return;
    const int l = lb.l, k = lb.k;
    const GridDev &g = *a.grid;
    const double dt = a.dt, dt2 = a.dt2, dt3 = a.dt3_6;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const double *vt = a.ws.vt + (int64_t)l * NT * K2;

    double e[4 * NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        e[t * 4 + 0] = -__builtin_inf();
        e[t * 4 + 1] = __builtin_inf();
        e[t * 4 + 2] = -__builtin_inf();
        e[t * 4 + 3] = __builtin_inf();
    }
    for (int i = 0; i < EXT_COLS_PER_THREAD; i++) {
        const int ij = lb.bx * EXT_COLS_PER_BLOCK + i * EXT_THREADS + threadIdx.x;
        if (ij >= ijdim) break;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            double V[5];
            load_vt_window(vt + t * K2 + ij, ijdim, k, V);
            const double zf = zdis_full_at(g, dt, dt2, dt3, V, k);
            if (zf == zf) {
                e[t * 4 + 0] = dmax(e[t * 4 + 0], zf);
                e[t * 4 + 1] = dmin(e[t * 4 + 1], zf);
            }
            if (k >= KMIN + 1) {
                const double zh = zdis_half_at(g, dt, dt2, dt3, V, k);
                if (zh == zh) {
                    e[t * 4 + 2] = dmax(e[t * 4 + 2], zh);
                    e[t * 4 + 3] = dmin(e[t * 4 + 3], zh);
                }
            }
        }
    }
    __shared__ double red[EXT_THREADS / 64][4 * NT];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 4 * NT; j++) {
        const double v = (j & 1) ? wave_min<64>(e[j]) : wave_max<64>(e[j]);
        if (lane == 0) red[wv][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4 * NT) {
        const int j = threadIdx.x;
        double v = red[0][j];
#pragma unroll
        for (int w = 1; w < EXT_THREADS / 64; w++) v = (j & 1) ? dmin(v, red[w][j]) : dmax(v, red[w][j]);
        a.ws.part[(((int64_t)l * KALL + k) * (4 * NT) + j) * nb + lb.bx] = v;
    }
}

// ---------------------------------------------------------------------------------
// k_bounds: kcell_min/kcell_max of vadv1d_prep (src/mod_vadv1d.cpp:317-352)
// ---------------------------------------------------------------------------------
template <int NT>
__global__ void __launch_bounds__(192) k_bounds(StepArgs a)
{
    const int l = blockIdx.y;
    const int t = blockIdx.x / 2, type = blockIdx.x % 2;   // grid.x = 2 NT
    const int k = threadIdx.x;                              // 0..95 used
    if (k >= KALL) return;
    const GridDev &g = *a.grid;
    const double *zh = (type == 0) ? g.gzh : g.gz_shift;
    const int mkmin = (type == 0) ? KMIN : KMIN + 1, mkmax = KMAX;
    int kcmin = k, kcmax = k;
    int general = 0;
    if (k >= mkmin && k <= mkmax) {
        const int nw = (a.ijdim + EXT_COLS_PER_BLOCK - 1) / EXT_COLS_PER_BLOCK;
        const double *pmax = a.ws.part + (((int64_t)l * KALL + k) * (4 * NT) + (t * 2 + type) * 2 + 0) * nw;
        const double *pmin = pmax + nw;
        double amax = -__builtin_inf(), amin = __builtin_inf();
        for (int i = 0; i < nw; i++) {
            amax = dmax(amax, pmax[i]);
            amin = dmin(amin, pmin[i]);
        }
        const double zzmax = (amax > DBL_MIN) ? amax : DBL_MIN;   // starts at numeric_limits::min()
        const double zzmin = (amin < DBL_MIN) ? amin : DBL_MIN;
        if (zzmax > 0.0) {
            for (int k2 = k; k2 >= mkmin; k2--) {
                if ((zh[k2] <= zh[k] - zzmax) && (zh[k2 + 1] > zh[k] - zzmax)) {
                    kcmin = k2;
                    break;
                }
            }
        }
        if (zzmin < 0.0) {
            for (int k2 = k; k2 <= mkmax; k2++) {
                if ((zh[k2] <= zh[k] - zzmin) && (zh[k2 + 1] > zh[k] - zzmin)) {
                    kcmax = k2;
                    break;
                }
            }
        }
        // The register-window sweep needs every departure cell in [k, k+1].
        const int ktop = (k + 2 <= mkmax + 1) ? k + 2 : mkmax + 1;
        if (kcmin != k || kcmax > k + 1) general = 1;
        if (amax > 0.0 && zh[k] - zzmax < zh[k]) general = 1;
        if (zh[k] - zzmin >= zh[ktop]) general = 1;
    }
    int *b = a.ws.bounds + (size_t)l * ext_per_region<NT>();
    b[bounds_index(t, type, k, 0)] = kcmin;
    b[bounds_index(t, type, k, 1)] = kcmax;
    if (general) atomicOr(&a.ws.region_flag[l], 1);
}

// ---------------------------------------------------------------------------------
// k_sweep
// ---------------------------------------------------------------------------------
// Compact transport data of one cell j (what the fluxes need), NT sedimenting tracers.
template <int NT> struct TCell {
    double rhoq[NT];     // rhogq'(t) * rgs
    double qrgs[NT];     // q'(t) * rgs
    double tem, z, kinh, up, vp;
    double hv[NT][2];    // half-level values at wall j: qh*rhogkin_v, qh*rhogw
};
// Values of cell j kept until the cell is finalised.
template <int NT> struct FCell {
    double rhog, rhogw, kinv, c2w0, c2w1;   // pre-transport; kinv/c2w* belong to wall j
    double drhoq[NT], drhoe, drhophi, drhokinh, drhogu, drhogv;  // full-level tendencies
    double drhokinv, drhogw;                                    // half-level tendencies (wall j)
};
// Sedimenting tracer index of tracer nq (-1 if it does not sediment).
template <int NT> __device__ __forceinline__ int sed_index(int nq)
{
    if (NT == 4) return (nq >= QR) ? nq - QR : -1;
    return (nq == QR) ? 0 : (nq == QS) ? 1 : (nq == QG) ? 2 : -1;
}
// The tracer update of src/mod_precip_transport.cpp:970-989 for one cell: nq = 1..5
// (QC..QG) in order, the sedimenting tracers adding their tendency to rhogq, rhog and
// (as -d g z) to rhoge, the others adding 0; then QV.
template <int NT>
__device__ __forceinline__ void apply_sedimented(const FCell<NT> &fc, double z, double rq[6], double &rg, double &e)
{
#pragma unroll
    for (int n = 1; n <= 5; n++) {
        const int t = sed_index<NT>(n);
        if (t >= 0) {
            rq[n] = rq[n] + fc.drhoq[t];
            rg = rg + fc.drhoq[t];
            e = e - fc.drhoq[t] * C_GRAV * z;
        } else {
            rq[n] = rq[n] + 0.0;
        }
    }
    rq[QV] = rq[QV] + 0.0;
}

template <int NT> __device__ __forceinline__ double rhof_of(const TCell<NT> &c, int t, int i)
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
__device__ __forceinline__ double cs1(double x) { return copysign(1.0, x); }
__device__ __forceinline__ double fact_of(int k2, int kc, int k, double dz)
{
    const double f1 = (double)(k2 - (kc + 1)), f2 = (double)((k - 1) - k2);
    const double f3 = (double)(k2 - k), f4 = (double)((kc - 1) - k2);
    return dz * 0.25 * ((cs1(f1) + 1.0) * (cs1(f2) + 1.0) - (cs1(f3) + 1.0) * (cs1(f4) + 1.0));
}
struct SweepCtx {
    const StepArgs *a;
    int l, ij, ijdim;
    int64_t p2, p3, pv, K2;     // base offsets of column (l, ij)
    double ix, iy, iz, jx, jy, jz;
};

template <int NT>
__device__ __forceinline__ void load_tcell(const SweepCtx &s, int j, TCell<NT> &c, double &rhog_j, double *q_j,
                                           double &gsgam2_j)
{
    const StepArgs &a = *s.a;
    const int64_t o = s.p2 + (int64_t)j * s.ijdim;
    const double rg = a.ws.rhog[o];
    double rq[6];
#pragma unroll
    for (int n = 1; n <= 5; n++) rq[n] = a.ws.rhogq[s.p3 + (int64_t)(n * KALL + j) * s.ijdim];
    const double re = a.ws.rhoge[o];
    const double gs = a.f.gsgam2[o];
    const bool interior = (j >= KMIN && j <= KMAX);
    const double rgs = interior ? a.f.gam2[o] / gs : 0.0;
    double qq[6];
#pragma unroll
    for (int n = 1; n <= 5; n++) qq[n] = rq[n] / rg;
    double qd = 1.0;
#pragma unroll
    for (int n = 1; n <= 5; n++) qd = qd - qq[n];
    double cva = qd * C_CVdry;
#pragma unroll
    for (int n = 1; n <= 5; n++) cva = cva + qq[n] * C_CVW;
    c.tem = interior ? re / (rg * cva) : 0.0;
    const double rvx = a.f.rhogvx[o], rvy = a.f.rhogvy[o], rvz = a.f.rhogvz[o];
    c.kinh = interior ? 0.5 * (rvx * rvx + rvy * rvy + rvz * rvz) / rg : 0.0;
    const double vx = a.f.vx[o], vy = a.f.vy[o], vz = a.f.vz[o];
    c.up = vx * s.ix + vy * s.iy + vz * s.iz;
    c.vp = vx * s.jx + vy * s.jy + vz * s.jz;
    c.z = a.f.z[o];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        c.rhoq[t] = rq[sed_nq<NT>(t)] * rgs;
        c.qrgs[t] = qq[sed_nq<NT>(t)] * rgs;
        q_j[t] = qq[sed_nq<NT>(t)];
    }
    rhog_j = rg;
    gsgam2_j = gs;
}

// Half-level values at wall j and the pre-transport wall quantities of FCell j.
template <int NT>
__device__ __forceinline__ void load_wall(const SweepCtx &s, const GridDev &g, int j, double rhog_j,
                                          double rhog_jm1, const double *q_j, const double *q_jm1,
                                          double gsgam2_j, double gsgam2_jm1, TCell<NT> &c, FCell<NT> &fc)
{
    const StepArgs &a = *s.a;
    const int64_t o = s.p2 + (int64_t)j * s.ijdim;
    const double rw = a.f.rhogw[o];
    const double gsh = a.f.gsgam2h[o];
    fc.rhog = rhog_j;
    fc.rhogw = rw;
    if (j >= KMIN) {
        fc.c2w0 = g.afact[j] / gsgam2_j * gsh;
        fc.c2w1 = g.bfact[j] / gsgam2_jm1 * gsh;
    } else {
        fc.c2w0 = 0.0;
        fc.c2w1 = 0.0;
    }
    if (j >= KMIN + 1 && j <= KMAX) {
        fc.kinv = 0.5 * (rw * rw) / (fc.c2w0 * rhog_j + fc.c2w1 * rhog_jm1);
        const double rgsh = a.f.gam2h[o] / gsh;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const double qh = 0.5 * (q_j[t] + q_jm1[t]) * rgsh;
            c.hv[t][0] = qh * fc.kinv;
            c.hv[t][1] = qh * rw;
        }
    } else {
        fc.kinv = 0.0;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            c.hv[t][0] = 0.0;
            c.hv[t][1] = 0.0;
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) fc.drhoq[t] = 0.0;
    fc.drhoe = fc.drhophi = fc.drhokinh = fc.drhogu = fc.drhogv = 0.0;
    fc.drhokinv = fc.drhogw = 0.0;
}

// Pre-transport data of an arbitrary cell/wall, from memory (general sweep only).
template <int NT>
__device__ double general_rhof(const SweepCtx &s, const GridDev &g, int j, int t, int i)
{
    TCell<NT> c;
    double rg, qj[NT], gs;
    load_tcell(s, j, c, rg, qj, gs);
    return rhof_of(c, t, i);
}
template <int NT>
__device__ double general_hv(const SweepCtx &s, const GridDev &g, int j, int t, int i)
{
    if (j < KMIN + 1 || j > KMAX) return 0.0;
    TCell<NT> c, cm;
    FCell<NT> fc;
    double rg, qj[NT], gs, rgm, qjm[NT], gsm;
    load_tcell(s, j, c, rg, qj, gs);
    load_tcell(s, j - 1, cm, rgm, qjm, gsm);
    load_wall(s, g, j, rg, rgm, qj, qjm, gs, gsm, c, fc);
    return c.hv[t][i];
}

// Post-sedimentation writes of one interior cell of a general region: rhog, rhoge and
// the sedimenting tracers (the rest of the state is finished by k_finish).
template <int NT>
__device__ __forceinline__ void store_sedimented(const SweepCtx &s, int c, double rg, const double rq[6], double e)
{
    const Fields &f = s.a->f;
    const int64_t o = s.p2 + (int64_t)c * s.ijdim;
    f.rhog[o] = rg;
    f.rhoge[o] = e;
#pragma unroll
    for (int t = 0; t < NT; t++)
        f.rhogq[s.p3 + (int64_t)(sed_nq<NT>(t) * KALL + c) * s.ijdim] = rq[sed_nq<NT>(t)];
}

// ---------------------------------------------------------------------------------
// Single-cell regions: every departure cell lies in {k, k+1} (checked by k_bounds),
// so each flux depends only on cells w, w+1 and the fall speeds Vt(w-3..w+1), and the
// transport splits into two point-wise kernels: k_flux (one thread per wall) and
// k_finish (one thread per cell).  The arithmetic is in nsw6_step.h, the column
// gathers and the per-wall flux (flux_wall) in nsw6_columns.h.
// ---------------------------------------------------------------------------------
// Three waves per SIMD: the cells, the fall speeds and the kcell bounds of the wall are
// all requested up front (one memory round trip per wave) at 150 VGPRs without spills.
template <int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_flux(StepArgs a)
{
    const int ijdim = a.ijdim;
    const LevelBlock lb = level_block(blockIdx.x, (ijdim + 255) / 256, KALL, 0, a.lall);
    if (!lb.valid) return;
    if (a.ws.region_flag[lb.l] != 0) return;
    const int ij = lb.bx * 256 + threadIdx.x;
    if (ij >= ijdim) return;
    flux_wall<NT>(a, lb.l, lb.k, ij);
}

// ---------------------------------------------------------------------------------
// k_finish: one thread per point.  The sedimented state comes from sediment() in the
// single-cell regions and from the general sweep elsewhere; then the point-wise tail
// (rediagnosis, saturation adjustment, final negative filter).
// ---------------------------------------------------------------------------------
// One point (l, k, ij) of the finish: the sedimented state (sediment() on the wall
// fluxes FXV gives, or the general sweep's result), then the point-wise tail.
template <int NT, typename Fetch>
__device__ __forceinline__ void finish_point(const StepArgs &a, int l, int k, int ij, bool general, const Fetch &FXV)
{
    const int ijdim = a.ijdim;
    const Fields &f = a.f;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
    const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;
    const bool interior = (k >= KMIN && k <= KMAX);

    // State after sedimentation.  QV, QC, QI ("3WATER") and the halo levels only
    // receive zero tendencies.  The momenta of the single-cell regions are stored with
    // the other outputs at the end: every store comes after every load, so that the
    // uniform reads of the grid and the parameters stay on the scalar cache.
    TailIn<double> x;
    // read by every path: loaded before the branches, with the path's own loads
    x.rhogq[QV] = a.ws.rhogq[pq + QV * K2] + 0.0;
    x.rhogq[QC] = a.ws.rhogq[pq + QC * K2] + 0.0;
    if (NT == 3) x.rhogq[QI] = a.ws.rhogq[pq + QI * K2] + 0.0;
    x.gsgam2 = f.gsgam2[p];
    double mom[4];   // rhogvx, rhogvy, rhogvz, rhogw after sedimentation (single-cell regions)
    if (interior) {
        if (general) {
            x.rhog = f.rhog[p];
            x.rhoge = f.rhoge[p];
#pragma unroll
            for (int t = 0; t < NT; t++) x.rhogq[sed_nq<NT>(t)] = f.rhogq[pq + sed_nq<NT>(t) * K2];
        } else {
            SedInN<double, NT> si;
            load_sed_from(a, l, ij, k, si, FXV);
            SedOutN<double, NT> so;
            sediment(*a.grid, k, load_unit(f, (int64_t)l * ijdim + ij), si, so, a.kin2kin != 0);
            mom[0] = so.rvx;
            mom[1] = so.rvy;
            mom[2] = so.rvz;
            mom[3] = so.rw;
            x.rhog = so.rg;
            x.rhoge = so.e;
#pragma unroll
            for (int t = 0; t < NT; t++) x.rhogq[sed_nq<NT>(t)] = so.rq[t];
        }
    } else {
        x.rhog = a.ws.rhog[p] + 0.0;
        x.rhoge = a.ws.rhoge[p] + 0.0;
#pragma unroll
        for (int t = 0; t < NT; t++) x.rhogq[sed_nq<NT>(t)] = a.ws.rhogq[pq + sed_nq<NT>(t) * K2] + 0.0;
        if (!general) {
            const int64_t pc = (int64_t)l * ijdim + ij;
            mom[0] = f.rhogvx[p] + 0.0 * f.ix[pc] + 0.0 * f.jx[pc];
            mom[1] = f.rhogvy[p] + 0.0 * f.iy[pc] + 0.0 * f.jy[pc];
            mom[2] = f.rhogvz[p] + 0.0 * f.iz[pc] + 0.0 * f.jz[pc];
            mom[3] = (a.kin2kin && k == KALL - 1) ? kin2kin_top_wall<NT>(a, l, ij, x.rhog, FXV) : f.rhogw[p] + 0.0;
        }
    }

    TailOut<double> y;
    const TailStatus st = tail(*a.P, interior, x, y);
    if (st.newton < 0) {
        const int code = (st.newton == -1) ? 2 : 1;
        atomicAdd(&a.ws.report[code], 1);
        const unsigned long long key = ((unsigned long long)l * KALL + (unsigned long long)k) * ijdim + ij;
        atomicMin(reinterpret_cast<unsigned long long *>(a.ws.report + REPORT_KEY), key);
        atomicOr(&a.ws.report[0], 1);
    }
    f.rhog[p] = y.rhog;
    f.rhoge[p] = y.rhoge;
    f.rho[p] = y.rho;
    f.tem[p] = y.tem;
    f.pre[p] = y.pre;
    f.qd[p] = y.qd;
#pragma unroll
    for (int n = 0; n < 6; n++) {
        f.rhogq[pq + n * K2] = y.rhogq[n];
        f.q[pq + n * K2] = y.q[n];
    }
    f.re_solid[p] = 20.0E-6;
    f.GDCLW[p] = 0.0 * 1.0;
    f.GDCFRC[p] = 0.0 * 1.0;
    if (!general) {
        f.rhogvx[p] = mom[0];
        f.rhogvy[p] = mom[1];
        f.rhogvz[p] = mom[2];
        f.rhogw[p] = mom[3];
    }
}

// Three waves per SIMD: with the registers four waves leave (126 of 128 used) the
// kernel is about 1 % slower.
template <int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_finish(StepArgs a)
{
    const int ijdim = a.ijdim;
    const LevelBlock lb = level_block(blockIdx.x, (ijdim + 255) / 256, KALL, 0, a.lall);
    if (!lb.valid) return;
    const int ij = lb.bx * 256 + threadIdx.x;
    if (ij >= ijdim) return;
    finish_point<NT>(a, lb.l, lb.k, ij, a.ws.region_flag[lb.l] != 0, FluxFromMemory<NT>(a, lb.l, ij));
}

template <bool GENERAL, int NT>
__global__ void __launch_bounds__(64) k_sweep(StepArgs a)
{
    const int ijdim = a.ijdim;
    const int l = blockIdx.y;
    if ((a.ws.region_flag[l] != 0) != GENERAL) return;
    const int ij = blockIdx.x * 64 + threadIdx.x;
    if (ij >= ijdim) return;
    const GridDev &g = *a.grid;
    const Fields &f = a.f;
    const int *bnd = a.ws.bounds + (size_t)l * ext_per_region<NT>();
    const double dt = a.dt, dt2 = a.dt2, dt3 = a.dt3_6;

    SweepCtx s;
    s.a = &a;
    s.l = l;
    s.ij = ij;
    s.ijdim = ijdim;
    s.K2 = (int64_t)KALL * ijdim;
    s.p2 = (int64_t)l * s.K2 + ij;
    s.p3 = (int64_t)l * NQ * s.K2 + ij;
    s.pv = (int64_t)l * NT * s.K2 + ij;
    const int64_t pc = (int64_t)l * ijdim + ij;
    s.ix = f.ix[pc]; s.iy = f.iy[pc]; s.iz = f.iz[pc];
    s.jx = f.jx[pc]; s.jy = f.jy[pc]; s.jz = f.jz[pc];
    const double *vt = a.ws.vt + s.pv;
    // deferred momentum: slots 0..3 of the region's flux block (k_flux skips general regions)
    double *mom = GENERAL ? (a.ws.flux + (int64_t)l * FxSlots<NT>::N * s.K2 + ij) : nullptr;

    // window: cells w (cur) and w+1 (nxt); finalisation records for w-2 .. w+1
    TCell<NT> cur, nxt;
    FCell<NT> fm2, fm1, f0, fp1;
    double rhog_cur, q_cur[NT], gs_cur;
    double rhog_nxt, q_nxt[NT], gs_nxt;
    WindWindow wind[NT];
    double fprev[NT][6], fprevh[NT][2];
    double rhog_new_prev = 0.0;       // rhog after sedimentation of the last finalised cell
    double rg_kmax = 0.0;             // rhog of cell kmax after sedimentation
    double precip0 = 0.0, precip1 = 0.0, lh = 0.0, prhoe = 0.0, prhophi = 0.0, prhokin = 0.0;

    // cell 0 (halo) and cell 1
    {
        TCell<NT> c0;
        double rg0, q0[NT], gs0;
        load_tcell(s, 0, c0, rg0, q0, gs0);
        FCell<NT> fdummy;
        load_wall(s, g, 0, rg0, rg0, q0, q0, gs0, gs0, c0, fdummy);
        fm1 = fdummy;     // cell 0 record (w-1 for w = 1)
        load_tcell(s, 1, cur, rhog_cur, q_cur, gs_cur);
        load_wall(s, g, 1, rhog_cur, rg0, q_cur, q0, gs_cur, gs0, cur, f0);
        rhog_new_prev = rg0;          // halo cell 0 keeps its density
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        wind_init(wind[t], vt[t * s.K2 + 1 * (int64_t)ijdim]);
#pragma unroll
        for (int i = 0; i < 6; i++) fprev[t][i] = 0.0;    // f(0) = 0
        fprevh[t][0] = fprevh[t][1] = 0.0;                // f_h(1) = 0
    }

    for (int w = KMIN; w <= KMAX + 1; w++) {
        double fcur[NT][6], fcurh[NT][2];
        if (w <= KMAX) {
            // ---- load cell w+1
            load_tcell(s, w + 1, nxt, rhog_nxt, q_nxt, gs_nxt);
            load_wall(s, g, w + 1, rhog_nxt, rhog_cur, q_nxt, q_cur, gs_nxt, gs_cur, nxt, fp1);
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const double vt_kp1 = (w + 1 <= KMAX) ? vt[t * s.K2 + (int64_t)(w + 1) * ijdim] : 0.0;
                double wh_kp1, vh_kp1, whh_kp1;
                wind_next(wind[t], w, vt_kp1, wh_kp1, vh_kp1, whh_kp1);
                // ---- full-level flux at wall w
                {
                    const double z0 = zdis_full(g, dt, dt2, dt3, wind[t], w, wh_kp1);
                    const int kcmin = bnd[bounds_index(t, 0, w, 0)], kcmax = bnd[bounds_index(t, 0, w, 1)];
                    const int kc = departure_cell(g.gzh, w, kcmin, kcmax, z0, KMIN, KMAX);
#pragma unroll
                    for (int i = 0; i < 6; i++) {
                        double acc = 0.0, zd = 0.0;
                        if (kcmin == w && kcmax == w) {
                            zd = z0;
                        } else {
                            for (int k2 = kcmin; k2 <= kcmax; k2++) {
                                const double fa = fact_of(k2, kc, w, g.dgz[k2]);
                                double rf;
                                if (k2 == w) rf = rhof_of(cur, t, i);
                                else if (k2 == w + 1) rf = rhof_of(nxt, t, i);
                                else rf = GENERAL ? general_rhof<NT>(s, g, k2, t, i) : 0.0;
                                acc = acc + rf * fa;
                                zd = z0 - fa;
                            }
                        }
                        double rfc;
                        if (kc == w) rfc = rhof_of(cur, t, i);
                        else if (kc == w + 1) rfc = rhof_of(nxt, t, i);
                        else rfc = GENERAL ? general_rhof<NT>(s, g, kc, t, i) : 0.0;
                        acc = acc + rfc * zd;
                        fcur[t][i] = eps_filter(acc);
                    }
                }
                // ---- half-level flux at wall w (walls 2..94)
                if (w >= KMIN + 1) {
                    const double z0 = zdis_half(g, dt, dt2, dt3, wind[t], w, whh_kp1);
                    const int khmin = bnd[bounds_index(t, 1, w, 0)], khmax = bnd[bounds_index(t, 1, w, 1)];
                    const int kfmax = bnd[bounds_index(t, 0, w, 1)];   // full-level kcell_max
                    const int kc = departure_cell(g.gz_shift, w, khmin, khmax, z0, KMIN + 1, KMAX);
#pragma unroll
                    for (int i = 0; i < 2; i++) {
                        double acc = 0.0, zd = 0.0;
                        if (khmin == w && kfmax == w) {
                            zd = z0;
                        } else {
                            for (int k2 = khmin; k2 <= kfmax; k2++) {
                                const double fa = fact_of(k2, kc, w, g.dgzh[k2]);
                                double hvv;
                                if (k2 == w) hvv = cur.hv[t][i];
                                else if (k2 == w + 1) hvv = nxt.hv[t][i];
                                else hvv = GENERAL ? general_hv<NT>(s, g, k2, t, i) : 0.0;
                                acc = acc + hvv * fa;
                                zd = z0 - fa;
                            }
                        }
                        double hvc;
                        if (kc == w) hvc = cur.hv[t][i];
                        else if (kc == w + 1) hvc = nxt.hv[t][i];
                        else hvc = GENERAL ? general_hv<NT>(s, g, kc, t, i) : 0.0;
                        acc = acc + hvc * zd;
                        fcurh[t][i] = eps_filter(acc);
                    }
                } else {
                    fcurh[t][0] = fcurh[t][1] = 0.0;
                }
                wind_shift(wind[t], vt_kp1, wh_kp1, vh_kp1, whh_kp1);
            }
            // GPREC = frain = -fprec_q(QR)/dt on k = 1..94, then * fraction_mp
            f.GPREC[s.p2 + (int64_t)w * ijdim] = (0.0 + (-fcur[0][0] / dt)) * 1.0;
            if (w == KMIN) {
                precip0 = precip0 - fcur[0][0] / dt;
#pragma unroll
                for (int t = 1; t < NT; t++) {   // QI ("4WATER"), QS, QG are solid
                    precip1 = precip1 - fcur[t][0] / dt;
                    lh = lh + fcur[t][0] * C_LHF0 / dt;
                }
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    prhoe = prhoe - fcur[t][1] / dt;
                    prhophi = prhophi - fcur[t][2] / dt;
                    prhokin = prhokin - fcur[t][3] / dt - 0.0 / dt;   // f_h(kmin) = 0
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < NT; t++) {
#pragma unroll
                for (int i = 0; i < 6; i++) fcur[t][i] = 0.0;
                fcurh[t][0] = fcurh[t][1] = 0.0;
            }
        }

        // ---- full-level divergence of cell w-1 (k = 1..94); its record is fm1
        if (w - 1 >= KMIN) {
            const double r = g.rdgz[w - 1];
#pragma unroll
            for (int t = 0; t < NT; t++) fm1.drhoq[t] = -(fcur[t][0] - fprev[t][0]) * r;
#pragma unroll
            for (int t = 0; t < NT; t++) {
                fm1.drhoe = fm1.drhoe - (fcur[t][1] - fprev[t][1]) * r;
                fm1.drhophi = fm1.drhophi - (fcur[t][2] - fprev[t][2]) * r;
                fm1.drhokinh = fm1.drhokinh - (fcur[t][3] - fprev[t][3]) * r;
                fm1.drhogu = fm1.drhogu - (fcur[t][4] - fprev[t][4]) * r;
                fm1.drhogv = fm1.drhogv - (fcur[t][5] - fprev[t][5]) * r;
            }
        }
        // ---- half-level divergence at wall w (2..94)
        if (w >= KMIN + 1 && w <= KMAX) {
            const double r = g.rdgzh[w];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                f0.drhokinv = f0.drhokinv - (fcurh[t][0] - fprevh[t][0]) * r;
                f0.drhogw = f0.drhogw - (fcurh[t][1] - fprevh[t][1]) * r;
            }
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
#pragma unroll
            for (int i = 0; i < 6; i++) fprev[t][i] = fcur[t][i];
            fprevh[t][0] = fcurh[t][0];
            fprevh[t][1] = fcurh[t][1];
        }

        // ---- finalise cell c = w-2 (needs cell w-1's tendencies, computed above)
        const int c = w - 2;
        if (c >= KMIN) {
            // fm2 = cell c, fm1 = cell c+1
            const int64_t o = s.p2 + (int64_t)c * ijdim;
            double rq[6];
#pragma unroll
            for (int n = 0; n < 6; n++) rq[n] = a.ws.rhogq[s.p3 + (int64_t)(n * KALL + c) * ijdim];
            const double re0 = a.ws.rhoge[o];
            double rvx = f.rhogvx[o], rvy = f.rhogvy[o], rvz = f.rhogvz[o];
            const double z = f.z[o], gs = f.gsgam2[o];
            const double gshp = f.gsgam2h[o + ijdim];
            const double w2c0 = g.cfact[c] * gs / gshp, w2c1 = g.dfact[c] * gs / gshp;
            const double kinh = 0.5 * (rvx * rvx + rvy * rvy + rvz * rvz) / fm2.rhog;
            // sedimentation apply (src/mod_precip_transport.cpp:960-989)
            double e = re0 + fm2.drhoe + fm2.drhophi;
            const double kinh_u = kinh + fm2.drhokinh;
            const double kinv_c = fm2.kinv + fm2.drhokinv;
            const double kinv_cp = fm1.kinv + fm1.drhokinv;
            double rvx_n = rvx + fm2.drhogu * s.ix + fm2.drhogv * s.jx;
            double rvy_n = rvy + fm2.drhogu * s.iy + fm2.drhogv * s.jy;
            double rvz_n = rvz + fm2.drhogu * s.iz + fm2.drhogv * s.jz;
            double rw_n = fm2.rhogw + fm2.drhogw;
            double rg = fm2.rhog;
            apply_sedimented(fm2, z, rq, rg, e);
            // KIN2EIN (:991-1010)
            const double tmp2 = kinh_u + (w2c0 * kinv_cp + w2c1 * kinv_c);
            double rg_cp = fm1.rhog;
#pragma unroll
            for (int t = 0; t < NT; t++) rg_cp = rg_cp + fm1.drhoq[t];
            const double rw_cp = fm1.rhogw + fm1.drhogw;
            const double kinh_new = 0.5 * (rvx_n * rvx_n + rvy_n * rvy_n + rvz_n * rvz_n) / rg;
            const double kinv_new_c =
                (c >= KMIN + 1) ? 0.5 * (rw_n * rw_n) / (fm2.c2w0 * rg + fm2.c2w1 * rhog_new_prev) : 0.0;
            const double kinv_new_cp =
                (c + 1 <= KMAX) ? 0.5 * (rw_cp * rw_cp) / (fm1.c2w0 * rg_cp + fm1.c2w1 * rg) : 0.0;
            const double tmpk = kinh_new + (w2c0 * kinv_new_cp + w2c1 * kinv_new_c);
            if (a.kin2kin) {
                kin2kin_h(rg, kinh_u, rvx_n, rvy_n, rvz_n);
                rw_n = kin2kin_v(fm2.c2w0 * rg + fm2.c2w1 * rhog_new_prev, kinv_c, rw_n);
            } else {
                e = e + (tmp2 - tmpk);
            }
            rhog_new_prev = rg;
            if (GENERAL) {
                mom[0 * s.K2 + (int64_t)c * ijdim] = rvx_n;
                mom[1 * s.K2 + (int64_t)c * ijdim] = rvy_n;
                mom[2 * s.K2 + (int64_t)c * ijdim] = rvz_n;
                mom[3 * s.K2 + (int64_t)c * ijdim] = rw_n;
            } else {
                f.rhogvx[o] = rvx_n;
                f.rhogvy[o] = rvy_n;
                f.rhogvz[o] = rvz_n;
                f.rhogw[o] = rw_n;
            }
            store_sedimented<NT>(s, c, rg, rq, e);
        }

        // ---- shift the window
        fm2 = fm1;
        fm1 = f0;
        f0 = fp1;
        cur = nxt;
        rhog_cur = rhog_nxt;
        gs_cur = gs_nxt;
#pragma unroll
        for (int t = 0; t < NT; t++) q_cur[t] = q_nxt[t];
    }
    // After the loop (w = 95 handled): cells up to 93 finalised. Cell 94 remains,
    // with cell 95 (halo) as its upper neighbour: fm2 = cell 94, fm1 = cell 95.
    {
        const int c = KMAX;
        const int64_t o = s.p2 + (int64_t)c * ijdim;
        double rq[6];
#pragma unroll
        for (int n = 0; n < 6; n++) rq[n] = a.ws.rhogq[s.p3 + (int64_t)(n * KALL + c) * ijdim];
        const double re0 = a.ws.rhoge[o];
        double rvx = f.rhogvx[o], rvy = f.rhogvy[o], rvz = f.rhogvz[o];
        const double z = f.z[o], gs = f.gsgam2[o];
        const double gshp = f.gsgam2h[o + ijdim];
        const double w2c0 = g.cfact[c] * gs / gshp, w2c1 = g.dfact[c] * gs / gshp;
        const double kinh = 0.5 * (rvx * rvx + rvy * rvy + rvz * rvz) / fm2.rhog;
        double e = re0 + fm2.drhoe + fm2.drhophi;
        const double kinh_u = kinh + fm2.drhokinh;
        const double kinv_c = fm2.kinv + fm2.drhokinv;
        const double kinv_cp = 0.0 + 0.0;    // wall 95
        double rvx_n = rvx + fm2.drhogu * s.ix + fm2.drhogv * s.jx;
        double rvy_n = rvy + fm2.drhogu * s.iy + fm2.drhogv * s.jy;
        double rvz_n = rvz + fm2.drhogu * s.iz + fm2.drhogv * s.jz;
        double rw_n = fm2.rhogw + fm2.drhogw;
        double rg = fm2.rhog;
        apply_sedimented(fm2, z, rq, rg, e);
        const double tmp2 = kinh_u + (w2c0 * kinv_cp + w2c1 * kinv_c);
        const double kinh_new = 0.5 * (rvx_n * rvx_n + rvy_n * rvy_n + rvz_n * rvz_n) / rg;
        const double kinv_new_c = 0.5 * (rw_n * rw_n) / (fm2.c2w0 * rg + fm2.c2w1 * rhog_new_prev);
        const double tmpk = kinh_new + (w2c0 * 0.0 + w2c1 * kinv_new_c);
        if (a.kin2kin) {
            kin2kin_h(rg, kinh_u, rvx_n, rvy_n, rvz_n);
            rw_n = kin2kin_v(fm2.c2w0 * rg + fm2.c2w1 * rhog_new_prev, kinv_c, rw_n);
        } else {
            e = e + (tmp2 - tmpk);
        }
        rg_kmax = rg;
        if (GENERAL) {
            mom[0 * s.K2 + (int64_t)c * ijdim] = rvx_n;
            mom[1 * s.K2 + (int64_t)c * ijdim] = rvy_n;
            mom[2 * s.K2 + (int64_t)c * ijdim] = rvz_n;
            mom[3 * s.K2 + (int64_t)c * ijdim] = rw_n;
        } else {
            f.rhogvx[o] = rvx_n;
            f.rhogvy[o] = rvy_n;
            f.rhogvz[o] = rvz_n;
            f.rhogw[o] = rw_n;
        }
        store_sedimented<NT>(s, c, rg, rq, e);
    }
    // halo cells 0 and 95: zero tendencies (their scalars are finished by k_finish)
#pragma unroll
    for (int hc = 0; hc < 2; hc++) {
        const int c = hc ? (KALL - 1) : 0;
        const int64_t o = s.p2 + (int64_t)c * ijdim;
        const double rvx_n = f.rhogvx[o] + 0.0 * s.ix + 0.0 * s.jx;
        const double rvy_n = f.rhogvy[o] + 0.0 * s.iy + 0.0 * s.jy;
        const double rvz_n = f.rhogvz[o] + 0.0 * s.iz + 0.0 * s.jz;
        double rw_n = f.rhogw[o] + 0.0;
        if (a.kin2kin && hc) {
            // KIN2KIN at wall kmax+1: rhogkin_v there is 0, the lower cell is kmax
            const double gs = f.gsgam2[o], gsh = f.gsgam2h[o], gs_m = f.gsgam2[o - ijdim];
            const double c2w0 = g.afact[c] / gs * gsh, c2w1 = g.bfact[c] / gs_m * gsh;
            rw_n = kin2kin_v(c2w0 * (a.ws.rhog[o] + 0.0) + c2w1 * rg_kmax, 0.0 + 0.0, rw_n);
        }
        if (GENERAL) {
            mom[0 * s.K2 + (int64_t)c * ijdim] = rvx_n;
            mom[1 * s.K2 + (int64_t)c * ijdim] = rvy_n;
            mom[2 * s.K2 + (int64_t)c * ijdim] = rvz_n;
            mom[3 * s.K2 + (int64_t)c * ijdim] = rw_n;
        } else {
            f.rhogvx[o] = rvx_n;
            f.rhogvy[o] = rvy_n;
            f.rhogvz[o] = rvz_n;
            f.rhogw[o] = rw_n;
        }
        f.GPREC[o] = (0.0 + 0.0) * 1.0;
    }
    if (GENERAL) {
        for (int c = 0; c < KALL; c++) {
            const int64_t o = s.p2 + (int64_t)c * ijdim;
            f.rhogvx[o] = mom[0 * s.K2 + (int64_t)c * ijdim];
            f.rhogvy[o] = mom[1 * s.K2 + (int64_t)c * ijdim];
            f.rhogvz[o] = mom[2 * s.K2 + (int64_t)c * ijdim];
            f.rhogw[o] = mom[3 * s.K2 + (int64_t)c * ijdim];
        }
    }
    // surface fluxes (precip_transport) then mp_driver's sum * fraction_mp
    const int64_t pp = (int64_t)l * 2 * ijdim + ij;
    f.precip[pp] = (0.0 + precip0) * 1.0;
    f.precip[pp + ijdim] = (0.0 + precip1) * 1.0;
    f.ISO1_precip[pp] = 0.0;
    f.ISO1_precip[pp + ijdim] = 0.0;
    f.ISO2_precip[pp] = 0.0;
    f.ISO2_precip[pp + ijdim] = 0.0;
    f.precip_rhoe[pc] = (0.0 + prhoe) * 1.0;
    f.precip_lh_heat[pc] = (0.0 + lh) * 1.0;
    f.precip_rhophi[pc] = (0.0 + prhophi) * 1.0;
    f.precip_rhokin[pc] = (0.0 + prhokin) * 1.0;
}

// ---------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------
// The transport of NT sedimenting tracers: extremes and kcell bounds, the wall fluxes
// and the general sweep, then the point-wise finish.
template <int NT> static void launch_transport(const StepArgs &a, hipStream_t stream, hipEvent_t *ev)
{
    const dim3 grid_c((a.ijdim + 63) / 64, a.lall);
    const dim3 grid_e(level_grid((a.ijdim + EXT_COLS_PER_BLOCK - 1) / EXT_COLS_PER_BLOCK, KMAX - KMIN + 1, a.lall));
    const dim3 grid_l(level_grid((a.ijdim + 255) / 256, KALL, a.lall));
    hipLaunchKernelGGL(k_extremes<NT>, grid_e, dim3(EXT_THREADS), 0, stream, a);
    hipLaunchKernelGGL(k_bounds<NT>, dim3(2 * NT, a.lall), dim3(192), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[2], stream);
    hipLaunchKernelGGL(k_flux<NT>, grid_l, dim3(256), 0, stream, a);
    hipLaunchKernelGGL((k_sweep<true, NT>), grid_c, dim3(64), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[3], stream);
    hipLaunchKernelGGL(k_finish<NT>, grid_l, dim3(256), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[4], stream);
}

extern "C" hipError_t nsw6_launch_step(const StepArgs *args, int precision, hipStream_t stream,
                                       hipEvent_t *ev /* NSW6_STEP_EVENTS events or NULL */)
{
    const StepArgs &a = *args;
    const dim3 grid_p((a.ijdim + 255) / 256, KALL, a.lall);
    if (ev) (void)hipEventRecord(ev[0], stream);
    if (precision == 1)
        (void)nsw6_launch_rates_mixed(args, stream);   // errors surface at hipGetLastError below
    else if (a.opt_indir)
        hipLaunchKernelGGL((k_rates<double, true, true>), grid_p, dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL((k_rates<double, true>), grid_p, dim3(256), 0, stream, a);
    if (a.cloudtop_intended)
        hipLaunchKernelGGL(k_cloudtop, dim3((a.ijdim + 255) / 256, a.lall), dim3(256), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[1], stream);
    if (a.four_water)
        launch_transport<4>(a, stream, ev);
    else
        launch_transport<3>(a, stream, ev);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------
// MP_DIV_NUM sub-cycling (src/mod_mp_driver.cpp:420-612).  After each sub-step the
// step outputs that mp_driver accumulates are added to the sums (sum = 0 + x at the
// first sub-step, sum + x after); at the end the outputs are sum * fraction_mp.
// GDCLW, GDCFRC and the ISO precipitations are zero in every sub-step, and so are
// their scaled sums.  Grid (ij/256, KALL + DIV_COLUMN_SUMS, l): rows below KALL are
// the GPREC levels, the rest the per-column outputs.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ double *div_output(const StepArgs &a, int l, int r)
{
    const int64_t n = a.ijdim;
    const Fields &f = a.f;
    if (r < KALL) return f.GPREC + ((int64_t)l * KALL + r) * n;
    switch (r - KALL) {
    case 0: return f.precip + (int64_t)l * 2 * n;
    case 1: return f.precip + (int64_t)l * 2 * n + n;
    case 2: return f.precip_rhoe + (int64_t)l * n;
    case 3: return f.precip_lh_heat + (int64_t)l * n;
    case 4: return f.precip_rhophi + (int64_t)l * n;
    default: return f.precip_rhokin + (int64_t)l * n;
    }
}
__device__ __forceinline__ int64_t div_sum_row(const StepArgs &a, int l, int r)
{
    const int64_t n = a.ijdim;
    if (r < KALL) return ((int64_t)l * KALL + r) * n;
    return (int64_t)a.lall * KALL * n + ((int64_t)l * DIV_COLUMN_SUMS + (r - KALL)) * n;
}

__global__ void __launch_bounds__(256) k_div_sum(StepArgs a, double *sum, int first)
{
    const int l = blockIdx.z, r = blockIdx.y;
    const int ij = blockIdx.x * 256 + threadIdx.x;
    if (ij >= a.ijdim) return;
    const double x = div_output(a, l, r)[ij];
    double *s = sum + div_sum_row(a, l, r) + ij;
    *s = first ? 0.0 + x : *s + x;
}

__global__ void __launch_bounds__(256) k_div_scale(StepArgs a, const double *sum, double fraction)
{
    const int l = blockIdx.z, r = blockIdx.y;
    const int ij = blockIdx.x * 256 + threadIdx.x;
    if (ij >= a.ijdim) return;
    div_output(a, l, r)[ij] = sum[div_sum_row(a, l, r) + ij] * fraction;
}

extern "C" hipError_t nsw6_launch_div_sum(const StepArgs *args, double *sum, int first, hipStream_t stream)
{
    const dim3 grid((args->ijdim + 255) / 256, KALL + DIV_COLUMN_SUMS, args->lall);
    hipLaunchKernelGGL(k_div_sum, grid, dim3(256), 0, stream, *args, sum, first);
    return hipGetLastError();
}

extern "C" hipError_t nsw6_launch_div_scale(const StepArgs *args, const double *sum, double fraction,
                                            hipStream_t stream)
{
    const dim3 grid((args->ijdim + 255) / 256, KALL + DIV_COLUMN_SUMS, args->lall);
    hipLaunchKernelGGL(k_div_scale, grid, dim3(256), 0, stream, *args, sum, fraction);
    return hipGetLastError();
}

// The FP64 step up to the wall fluxes of the single-cell regions (the forward part
// that the reverse sweep recomputes); general regions keep region_flag = 1.
extern "C" hipError_t nsw6_launch_forward_to_flux(const StepArgs *args, hipStream_t stream)
{
    const StepArgs &a = *args;
    const dim3 grid_p((a.ijdim + 255) / 256, KALL, a.lall);
    const dim3 grid_e(level_grid((a.ijdim + EXT_COLS_PER_BLOCK - 1) / EXT_COLS_PER_BLOCK, KMAX - KMIN + 1, a.lall));
    const dim3 grid_l(level_grid((a.ijdim + 255) / 256, KALL, a.lall));
    if (a.opt_indir)
        hipLaunchKernelGGL((k_rates<double, false, true>), grid_p, dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL((k_rates<double, false>), grid_p, dim3(256), 0, stream, a);
    hipLaunchKernelGGL(k_extremes<3>, grid_e, dim3(EXT_THREADS), 0, stream, a);
    hipLaunchKernelGGL(k_bounds<3>, dim3(6, a.lall), dim3(192), 0, stream, a);
    hipLaunchKernelGGL(k_flux<3>, grid_l, dim3(256), 0, stream, a);
    return hipGetLastError();
}
