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
#include "nsw6_internal.h"

using namespace nsw6;

__constant__ double c_a1_tab[32] = {
    0.0001E-7, 0.7939E-7, 0.7841E-6, 0.3369E-5, 0.4336E-5, 0.5285E-5, 0.3728E-5, 0.1852E-5,
    0.2991E-6, 0.4248E-6, 0.7434E-6, 0.1812E-5, 0.4394E-5, 0.9145E-5, 0.1725E-4, 0.3348E-4,
    0.1725E-4, 0.9175E-5, 0.4412E-5, 0.2252E-5, 0.9115E-6, 0.4876E-6, 0.3473E-6, 0.4758E-6,
    0.6306E-6, 0.8573E-6, 0.7868E-6, 0.7192E-6, 0.6513E-6, 0.5956E-6, 0.5333E-6, 0.4834E-6};
__constant__ double c_a2_tab[32] = {
    0.0100, 0.4006, 0.4831, 0.5320, 0.5307, 0.5319, 0.5249, 0.4888, 0.3849, 0.4047, 0.4318,
    0.4771, 0.5183, 0.5463, 0.5651, 0.5813, 0.5655, 0.5478, 0.5203, 0.4906, 0.4447, 0.4126,
    0.3960, 0.4149, 0.4320, 0.4506, 0.4483, 0.4460, 0.4433, 0.4413, 0.4382, 0.4361};

// ---------------------------------------------------------------------------------
// k_rates
// ---------------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(256) k_rates(StepArgs a)
{
    const int ijdim = a.ijdim;
    const int l = blockIdx.z, k = blockIdx.y;
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;

    // Region-wide reductions of this step start from their neutral values.
    if (blockIdx.x == 0 && k == 0) {
        if (threadIdx.x == 0) a.ws.region_flag[l] = 0;
        if (l == 0 && threadIdx.x < REPORT_INTS) a.ws.report[threadIdx.x] = (threadIdx.x >= 3) ? -1 : 0;
    }
    if (ij >= ijdim) return;

    const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;   // + n*K2
    const int64_t pv = (int64_t)l * 3 * K2 + (int64_t)k * ijdim + ij;    // + t*K2
    const Fields &f = a.f;

    if (k == 0 || k == KALL - 1) {
        a.ws.rhog[p] = f.rhog[p];
        a.ws.rhoge[p] = f.rhoge[p];
#pragma unroll
        for (int n = 0; n < NQ; n++) a.ws.rhogq[pq + n * K2] = f.rhogq[pq + n * K2];
#pragma unroll
        for (int t = 0; t < 3; t++) a.ws.vt[pv + t * K2] = 0.0;
        f.re_liquid[p] = 0.0;
        f.re_cld[p] = 0.0;
        return;
    }

    const DevParams &P = *a.P;
    // ---- negative_filter (src/mod_mp_nsw6.cpp:1770-1831), interior point
    double rq[NQ];
#pragma unroll
    for (int n = 0; n < NQ; n++) rq[n] = f.rhogq[pq + n * K2];
    double diffq = 0.0;
#pragma unroll
    for (int n = 2; n <= 5; n++) {
        diffq += rq[n];
        rq[n] = dmax(rq[n], 0.0);
    }
#pragma unroll
    for (int n = 2; n <= 5; n++) diffq -= rq[n];
    rq[QV] += diffq;
    diffq = rq[QV];
    rq[QV] = dmax(rq[QV], 0.0);
    diffq -= rq[QV];
    const double rhog = f.rhog[p] * (1.0 - diffq);
    const double gsgam2 = f.gsgam2[p];
    const double rho = rhog / gsgam2;
    double q[NQ];
    q[QV] = f.q[pq];                                  // stale input QV (reference keeps it)
#pragma unroll
    for (int n = 1; n <= 5; n++) q[n] = rq[n] / rhog;
    double qd = 1.0;
#pragma unroll
    for (int n = 1; n <= 5; n++) qd = qd - q[n];
    double cva = qd * C_CVdry;
#pragma unroll
    for (int n = 1; n <= 5; n++) cva = cva + q[n] * C_CVW;
    const double tem = f.tem[p];
    double rhoge = tem * rhog * cva;
    const double pre = rho * (qd * C_Rdry + q[QV] * C_Rvap) * tem;

    // ---- psat / qsat (TagSATWater), Bergeron
    const double qsatl = psat_liq(P, tem) / (rho * C_Rvap * tem);
    const double qsati = psat_ice(P, tem) / (rho * C_Rvap * tem);
    double a1 = 1.0, a2 = 0.0, ma2 = 1.0;   // only the Bergeron growth (off with explicit ice) uses them
    if (P.sw_expice != 1.0) bergeron(c_a1_tab, c_a2_tab, tem, a1, a2, ma2);

    // ---- TagBigLoop
    RateIn<R> in;
    in.dens = (R)rho;
    in.temp = (R)tem;
#pragma unroll
    for (int n = 0; n < NQ; n++) in.q[n] = (R)q[n];
    in.qsatl = (R)qsatl;
    in.qsati = (R)qsati;
    in.pre = (R)pre;
    in.rhog = (R)rhog;
    in.a1 = (R)a1;
    in.a2 = (R)a2;
    in.ma2 = (R)ma2;
    in.sw = rate_switches(P, rho, tem, q, qsati);
    rate_differences(in, tem, q, qsatl, qsati);
    RateOut<R> o;
    rates<R>(P, (R)a.dt, in, o);
    double rceff, rceff_cld, rct, tct;
    effective_radius(P, in.sw, rho, tem, q, rceff, rceff_cld, rct, tct);

    // ---- drhogq = rhog * q_t * dt with qv_t = -(sum of the others) (src/mod_mp_nsw6.cpp:
    // 1548-1559), in FP64 for both forms so the point's total water is conserved to FP64
    // round-off; then TagUpdateRHOGQ / TagUpdateRHOGE
    double qt[NQ];
#pragma unroll
    for (int n = 1; n < NQ; n++) qt[n] = (double)o.qt[n];
    qt[QV] = -(qt[QC] + qt[QR] + qt[QI] + qt[QS] + qt[QG]);
    double drq[NQ];
#pragma unroll
    for (int n = 0; n < NQ; n++) drq[n] = rhog * qt[n] * a.dt;
#pragma unroll
    for (int n = 0; n < NQ; n++) rq[n] += drq[n];
    rhoge = rhoge - C_LHV0 * drq[QV] + C_LHF0 * drq[QI] + C_LHF0 * drq[QS] + C_LHF0 * drq[QG];

    a.ws.rhog[p] = rhog;
    a.ws.rhoge[p] = rhoge;
#pragma unroll
    for (int n = 0; n < NQ; n++) a.ws.rhogq[pq + n * K2] = rq[n];
    a.ws.vt[pv + 0 * K2] = (double)o.Vtr;
    a.ws.vt[pv + 1 * K2] = (double)o.Vts;
    a.ws.vt[pv + 2 * K2] = (double)o.Vtg;
    f.re_liquid[p] = rceff;
    f.re_cld[p] = rceff_cld;
    if (k == KMAX) {
        // every level writes the cloud-top row in the reference; the serial order leaves kmax
        const int64_t pc = (int64_t)l * ijdim + ij;
        f.rctop[pc] = rct;
        f.tctop[pc] = tct;
        const double sw = step_pos(tct - C_TEM00);
        f.rwtop[pc] = (sw)*rct + (1.0 - sw) * C_UNDEF;
    }
}

// ---------------------------------------------------------------------------------
// departure distance of vadv1d_prep (src/mod_vadv1d.cpp:284-302)
// ---------------------------------------------------------------------------------
// zdis = dt*wh - dt^2*wh*dw/dz/2 + dt^3*wh*((dw/dz)^2 + wh*d2w/dz2)/6 with the grid
// divisions taken as products with reciprocals computed once on the host
__device__ __forceinline__ double zdis_interior(double dt, double dt2, double dt3_6, double whm, double whk,
                                                double whp, double rdzm, double rdzk, double rs)
{
    const double d = whp - whm;
    const double t1 = dt * whk;
    const double t2 = dt2 * whk * d * rs * 0.5;
    const double g = d * rs;
    const double inner = ((whp - whk) * rdzk - (whk - whm) * rdzm) * rs * 2.0;
    const double t3 = dt3_6 * whk * (g * g + whk * inner);
    return t1 - t2 + t3;
}
__device__ __forceinline__ double zdis_bottom(double dt, double dt2, double whk, double whp, double rdzk)
{
    return dt * whk - dt2 * whk * (whp - whk) * rdzk * 0.5;
}

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

// Point-wise form of the same wall speeds: with V[d + 3] = Vt(k + d), d = -3..+1,
// wh_full(V, k, d) is the full-level wh at wall k + d and wh_half(V, k, d) the
// half-level one (Vtermh averaged again), identical to the sliding window above.
__device__ __forceinline__ double wh_full(const double V[5], int k, int d)
{
    const int j = k + d;
    if (j == 0) return 0.0;
    if (j == 1) return V[d + 3];                        // wh(1) = Vt(1)
    if (j <= KMAX) return 0.5 * (V[d + 2] + V[d + 3]);  // 0.5 (Vt(j-1) + Vt(j))
    return V[d + 2];                                    // wh(95) = Vt(94)
}
__device__ __forceinline__ double vh_half(const double V[5], int k, int d)
{
    const int j = k + d;
    return (j >= 2 && j <= KMAX - 1) ? 0.5 * (V[d + 3] + V[d + 2]) : 0.0;
}
__device__ __forceinline__ double wh_half(const double V[5], int k, int d)
{
    const int j = k + d;
    if (j >= 3 && j <= KMAX) return 0.5 * (vh_half(V, k, d - 1) + vh_half(V, k, d));
    return (j == 2) ? vh_half(V, k, d) : 0.0;
}
// zdis of the full level at wall k (1..94) and of the half level at wall k (2..94)
__device__ __forceinline__ double zdis_full_at(const GridDev &g, double dt, double dt2, double dt3_6,
                                               const double V[5], int k)
{
    const double whp = wh_full(V, k, 1);
    if (k == KMIN) return zdis_bottom(dt, dt2, wh_full(V, k, 0), whp, g.rdgz[KMIN]);
    return zdis_interior(dt, dt2, dt3_6, wh_full(V, k, -1), wh_full(V, k, 0), whp, g.rdgz[k - 1], g.rdgz[k],
                         g.rsum_gz[k]);
}
__device__ __forceinline__ double zdis_half_at(const GridDev &g, double dt, double dt2, double dt3_6,
                                               const double V[5], int k)
{
    const double whp = wh_half(V, k, 1);
    if (k == KMIN + 1) return zdis_bottom(dt, dt2, wh_half(V, k, 0), whp, g.rdgzh[KMIN + 1]);
    return zdis_interior(dt, dt2, dt3_6, wh_half(V, k, -1), wh_half(V, k, 0), whp, g.rdgzh[k - 1], g.rdgzh[k],
                         g.rsum_gzh[k]);
}
// Vt(k-3 .. k+1) of one tracer of one column (levels below 0 are never used).
__device__ __forceinline__ void load_vt_window(const double *vt_col, int64_t ijdim, int k, double V[5])
{
#pragma unroll
    for (int d = -3; d <= 1; d++) V[d + 3] = (k + d >= 0) ? vt_col[(int64_t)(k + d) * ijdim] : 0.0;
}

// ---------------------------------------------------------------------------------
// k_extremes: per-level max/min over the columns of a region of the departure
// distances (src/mod_vadv1d.cpp:317-327).  One block per (column chunk, level k in
// 1..94, region); each thread takes EXT_COLS_PER_THREAD columns strided by the block
// size.  The block's 12 extremes (3 tracers x {full, half} x {max, min}) go to the
// partials array [l][k][12][block] that k_bounds reduces.
// ---------------------------------------------------------------------------------
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
    const double *vt = a.ws.vt + (int64_t)l * 3 * K2;

    double e[12];
#pragma unroll
    for (int t = 0; t < 3; t++) {
        e[t * 4 + 0] = -__builtin_inf();
        e[t * 4 + 1] = __builtin_inf();
        e[t * 4 + 2] = -__builtin_inf();
        e[t * 4 + 3] = __builtin_inf();
    }
    for (int i = 0; i < EXT_COLS_PER_THREAD; i++) {
        const int ij = lb.bx * EXT_COLS_PER_BLOCK + i * EXT_THREADS + threadIdx.x;
        if (ij >= ijdim) break;
#pragma unroll
        for (int t = 0; t < 3; t++) {
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
    __shared__ double red[EXT_THREADS / 64][12];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 12; j++) {
        const double v = (j & 1) ? wave_min<64>(e[j]) : wave_max<64>(e[j]);
        if (lane == 0) red[wv][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        const int j = threadIdx.x;
        double v = red[0][j];
#pragma unroll
        for (int w = 1; w < EXT_THREADS / 64; w++) v = (j & 1) ? dmin(v, red[w][j]) : dmax(v, red[w][j]);
        a.ws.part[(((int64_t)l * KALL + k) * 12 + j) * nb + lb.bx] = v;
    }
}

// ---------------------------------------------------------------------------------
// k_bounds: kcell_min/kcell_max of vadv1d_prep (src/mod_vadv1d.cpp:317-352)
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(192) k_bounds(StepArgs a)
{
    const int l = blockIdx.y;
    const int t = blockIdx.x / 2, type = blockIdx.x % 2;   // grid.x = 6
    const int k = threadIdx.x;                              // 0..95 used
    if (k >= KALL) return;
    const GridDev &g = *a.grid;
    const double *zh = (type == 0) ? g.gzh : g.gz_shift;
    const int mkmin = (type == 0) ? KMIN : KMIN + 1, mkmax = KMAX;
    int kcmin = k, kcmax = k;
    int general = 0;
    if (k >= mkmin && k <= mkmax) {
        const int nw = (a.ijdim + EXT_COLS_PER_BLOCK - 1) / EXT_COLS_PER_BLOCK;
        const double *pmax = a.ws.part + (((int64_t)l * KALL + k) * 12 + (t * 2 + type) * 2 + 0) * nw;
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
    int *b = a.ws.bounds + (size_t)l * BOUNDS_PER_REGION;
    b[bounds_index(t, type, k, 0)] = kcmin;
    b[bounds_index(t, type, k, 1)] = kcmax;
    if (general) atomicOr(&a.ws.region_flag[l], 1);
}

// ---------------------------------------------------------------------------------
// k_sweep
// ---------------------------------------------------------------------------------
// Compact transport data of one cell j (what the fluxes need).
struct TCell {
    double rhoq[3];     // rhogq'(t) * rgs
    double qrgs[3];     // q'(t) * rgs
    double tem, z, kinh, up, vp;
    double hv[3][2];    // half-level values at wall j: qh*rhogkin_v, qh*rhogw
};
// Values of cell j kept until the cell is finalised.
struct FCell {
    double rhog, rhogw, kinv, c2w0, c2w1;   // pre-transport; kinv/c2w* belong to wall j
    double drhoq[3], drhoe, drhophi, drhokinh, drhogu, drhogv;  // full-level tendencies
    double drhokinv, drhogw;                                   // half-level tendencies (wall j)
};

__device__ __forceinline__ double rhof_of(const TCell &c, int t, int i)
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
__device__ __forceinline__ double eps_filter(double f)
{
    return f * (0.5 + copysign(0.5, fabs(f) - C_EPS));
}
// per-column departure cell (src/mod_vadv1d.cpp:355-384)
__device__ __forceinline__ int departure_cell(const double *zh, int k, int kcmin, int kcmax, double zdis, int mkmin,
                                              int mkmax)
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

struct SweepCtx {
    const StepArgs *a;
    int l, ij, ijdim;
    int64_t p2, p3, pv, K2;     // base offsets of column (l, ij)
    double ix, iy, iz, jx, jy, jz;
};

__device__ __forceinline__ void load_tcell(const SweepCtx &s, int j, TCell &c, double &rhog_j, double q_j[3],
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
    c.rhoq[0] = rq[QR] * rgs;
    c.rhoq[1] = rq[QS] * rgs;
    c.rhoq[2] = rq[QG] * rgs;
    c.qrgs[0] = qq[QR] * rgs;
    c.qrgs[1] = qq[QS] * rgs;
    c.qrgs[2] = qq[QG] * rgs;
    rhog_j = rg;
    q_j[0] = qq[QR];
    q_j[1] = qq[QS];
    q_j[2] = qq[QG];
    gsgam2_j = gs;
}

// Half-level values at wall j and the pre-transport wall quantities of FCell j.
__device__ __forceinline__ void load_wall(const SweepCtx &s, const GridDev &g, int j, double rhog_j,
                                          double rhog_jm1, const double q_j[3], const double q_jm1[3],
                                          double gsgam2_j, double gsgam2_jm1, TCell &c, FCell &fc)
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
        for (int t = 0; t < 3; t++) {
            const double qh = 0.5 * (q_j[t] + q_jm1[t]) * rgsh;
            c.hv[t][0] = qh * fc.kinv;
            c.hv[t][1] = qh * rw;
        }
    } else {
        fc.kinv = 0.0;
#pragma unroll
        for (int t = 0; t < 3; t++) {
            c.hv[t][0] = 0.0;
            c.hv[t][1] = 0.0;
        }
    }
#pragma unroll
    for (int t = 0; t < 3; t++) fc.drhoq[t] = 0.0;
    fc.drhoe = fc.drhophi = fc.drhokinh = fc.drhogu = fc.drhogv = 0.0;
    fc.drhokinv = fc.drhogw = 0.0;
}

// Pre-transport data of an arbitrary cell/wall, from memory (general sweep only).
__device__ double general_rhof(const SweepCtx &s, const GridDev &g, int j, int t, int i)
{
    TCell c;
    double rg, qj[3], gs;
    load_tcell(s, j, c, rg, qj, gs);
    return rhof_of(c, t, i);
}
__device__ double general_hv(const SweepCtx &s, const GridDev &g, int j, int t, int i)
{
    if (j < KMIN + 1 || j > KMAX) return 0.0;
    TCell c, cm;
    FCell fc;
    double rg, qj[3], gs, rgm, qjm[3], gsm;
    load_tcell(s, j, c, rg, qj, gs);
    load_tcell(s, j - 1, cm, rgm, qjm, gsm);
    load_wall(s, g, j, rg, rgm, qj, qjm, gs, gsm, c, fc);
    return c.hv[t][i];
}

// Post-sedimentation writes of one interior cell of a general region: rhog, rhoge and
// the sedimenting tracers (the rest of the state is finished by k_finish).
__device__ __forceinline__ void store_sedimented(const SweepCtx &s, int c, double rg, const double rq[6], double e)
{
    const Fields &f = s.a->f;
    const int64_t o = s.p2 + (int64_t)c * s.ijdim;
    f.rhog[o] = rg;
    f.rhoge[o] = e;
    f.rhogq[s.p3 + (int64_t)(QR * KALL + c) * s.ijdim] = rq[QR];
    f.rhogq[s.p3 + (int64_t)(QS * KALL + c) * s.ijdim] = rq[QS];
    f.rhogq[s.p3 + (int64_t)(QG * KALL + c) * s.ijdim] = rq[QG];
}

// ---------------------------------------------------------------------------------
// Single-cell regions: every departure cell lies in {k, k+1} (checked by k_bounds),
// so each semi-Lagrangian flux of vadv1d_getflux_new (src/mod_vadv1d.cpp:393-450)
// has the closed form
//     kc = k   : f = rhof(k) * zdis0
//     kc = k+1 : f = -rhof(k) * dz(k) + rhof(k+1) * zdis0   (zdis0 kept, not refreshed)
// and depends only on cells k, k+1 and the fall speeds Vt(k-3..k+1).  The half level
// uses the full level's kcell_max as the reference does
// (src/mod_precip_transport.cpp:890-892).  The transport then splits into two
// point-wise kernels: k_flux (one thread per wall) and k_finish (one thread per cell).
// Flux divergences of the non-mass quantities are taken on tracer-summed fluxes (a
// reordering of the reference's per-tracer sum).
// ---------------------------------------------------------------------------------
struct Raw {
    double rhog, rq[5], rhoge;
    double gs, gam2, gsh, gam2h;
    double rvx, rvy, rvz, rw, vx, vy, vz, z;
};
struct Cell {
    double rhoq[3], qrgs[3], q[3];   // sedimenting tracers QR, QS, QG
    double tem, z, kinh, up, vp;
    double hv[3][2];                 // half-level values at the cell's lower wall
    double rhog, rhogw, kinv, c2w0, c2w1, gs;
};

__device__ __forceinline__ void load_raw(const SweepCtx &s, int j, Raw &r)
{
    const StepArgs &a = *s.a;
    const Fields &f = a.f;
    const int64_t o = s.p2 + (int64_t)j * s.ijdim;
    const int64_t oq = s.p3 + (int64_t)j * s.ijdim;
    r.rhog = a.ws.rhog[o];
#pragma unroll
    for (int n = 1; n <= 5; n++) r.rq[n - 1] = a.ws.rhogq[oq + n * s.K2];
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

// Transported quantities of cell j (src/mod_precip_transport.cpp:740-780, 846-888)
// given the wall-relevant part of cell j-1 (rhog, q, gsgam2).
__device__ __forceinline__ void derive_cell(const SweepCtx &s, const GridDev &g, int j, const Raw &r,
                                            const Cell &below, Cell &c)
{
    const bool interior = (j >= KMIN && j <= KMAX);
    const double rg = r.rhog;
    const double rgs = interior ? r.gam2 / r.gs : 0.0;
    double qq[5];
#pragma unroll
    for (int n = 0; n < 5; n++) qq[n] = r.rq[n] / rg;
    double qd = 1.0;
#pragma unroll
    for (int n = 0; n < 5; n++) qd = qd - qq[n];
    double cva = qd * C_CVdry;
#pragma unroll
    for (int n = 0; n < 5; n++) cva = cva + qq[n] * C_CVW;
    c.tem = interior ? r.rhoge / (rg * cva) : 0.0;
    c.kinh = interior ? 0.5 * (r.rvx * r.rvx + r.rvy * r.rvy + r.rvz * r.rvz) / rg : 0.0;
    c.up = r.vx * s.ix + r.vy * s.iy + r.vz * s.iz;
    c.vp = r.vx * s.jx + r.vy * s.jy + r.vz * s.jz;
    c.z = r.z;
    // tracer order inside rq/qq: QC, QR, QI, QS, QG
    c.q[0] = qq[1];
    c.q[1] = qq[3];
    c.q[2] = qq[4];
    c.rhoq[0] = r.rq[1] * rgs;
    c.rhoq[1] = r.rq[3] * rgs;
    c.rhoq[2] = r.rq[4] * rgs;
#pragma unroll
    for (int t = 0; t < 3; t++) c.qrgs[t] = c.q[t] * rgs;
    c.rhog = rg;
    c.rhogw = r.rw;
    c.gs = r.gs;
    if (j >= KMIN) {
        c.c2w0 = g.afact[j] / r.gs * r.gsh;
        c.c2w1 = g.bfact[j] / below.gs * r.gsh;
    } else {
        c.c2w0 = 0.0;
        c.c2w1 = 0.0;
    }
    if (j >= KMIN + 1 && j <= KMAX) {
        c.kinv = 0.5 * (r.rw * r.rw) / (c.c2w0 * rg + c.c2w1 * below.rhog);
        const double rgsh = r.gam2h / r.gsh;
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const double qh = 0.5 * (c.q[t] + below.q[t]) * rgsh;
            c.hv[t][0] = qh * c.kinv;
            c.hv[t][1] = qh * r.rw;
        }
    } else {
        c.kinv = 0.0;
#pragma unroll
        for (int t = 0; t < 3; t++) c.hv[t][0] = c.hv[t][1] = 0.0;
    }
}

__device__ __forceinline__ double rhof_cell(const Cell &c, int t, int i)
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
// k_flux: one thread per (wall w, column) of the single-cell regions.  Writes the flux
// workspace of wall w (FLUX_SLOTS), frain = GPREC (src/mod_precip_transport.cpp:
// 916-927) and, at wall kmin, the surface fluxes precip, precip_lh_heat,
// precip_rhoe/rhophi/rhokin (:900-958) scaled by fraction_mp = 1 (src/mod_mp_driver.cpp:
// 560-612).  Walls 0 and 95 carry no flux.
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_flux(StepArgs a)
{
    const int ijdim = a.ijdim;
    const LevelBlock lb = level_block(blockIdx.x, (ijdim + 255) / 256, KALL, 0, a.lall);
    if (!lb.valid) return;
    const int l = lb.l, w = lb.k;
    if (a.ws.region_flag[l] != 0) return;
    const int ij = lb.bx * 256 + threadIdx.x;
    if (ij >= ijdim) return;
    const GridDev &g = *a.grid;
    const Fields &f = a.f;

    SweepCtx s;
    s.a = &a;
    s.l = l;
    s.ij = ij;
    s.ijdim = ijdim;
    s.K2 = (int64_t)KALL * ijdim;
    s.p2 = (int64_t)l * s.K2 + ij;
    s.p3 = (int64_t)l * NQ * s.K2 + ij;
    s.pv = (int64_t)l * 3 * s.K2 + ij;
    double *fx = a.ws.flux + (int64_t)l * FLUX_SLOTS * s.K2 + (int64_t)w * ijdim + ij;   // + slot*K2
    const int64_t ow = s.p2 + (int64_t)w * ijdim;

    fx[FX_RW * s.K2] = f.rhogw[ow];
    if (w == 0 || w == KALL - 1) {
#pragma unroll
        for (int i = 0; i < FX_RW; i++) fx[i * s.K2] = 0.0;
        f.GPREC[ow] = (0.0 + 0.0) * 1.0;
        return;
    }
    const int64_t pc = (int64_t)l * ijdim + ij;
    s.ix = f.ix[pc]; s.iy = f.iy[pc]; s.iz = f.iz[pc];
    s.jx = f.jx[pc]; s.jy = f.jy[pc]; s.jz = f.jz[pc];
    const int *bnd = a.ws.bounds + (size_t)l * BOUNDS_PER_REGION;
    const double dt = a.dt, dt2 = a.dt2, dt3 = a.dt3_6;

    // cell w-1 enters the walls of cell w only through rhog, q and gsgam2
    Cell below, cur, nxt;
    {
        const int64_t o = s.p2 + (int64_t)(w - 1) * ijdim;
        const int64_t oq = s.p3 + (int64_t)(w - 1) * ijdim;
        below.rhog = a.ws.rhog[o];
        below.gs = f.gsgam2[o];
        below.q[0] = a.ws.rhogq[oq + QR * s.K2] / below.rhog;
        below.q[1] = a.ws.rhogq[oq + QS * s.K2] / below.rhog;
        below.q[2] = a.ws.rhogq[oq + QG * s.K2] / below.rhog;
    }
    {
        Raw r;
        load_raw(s, w, r);
        derive_cell(s, g, w, r, below, cur);
        load_raw(s, w + 1, r);
        derive_cell(s, g, w + 1, r, cur, nxt);
    }

    double fq[3], fs[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, fh[2] = {0.0, 0.0};
    double precip0 = 0.0, precip1 = 0.0, lh = 0.0, prhoe = 0.0, prhophi = 0.0, prhokin = 0.0;
#pragma unroll
    for (int t = 0; t < 3; t++) {
        double V[5];
        load_vt_window(a.ws.vt + s.pv + t * s.K2, ijdim, w, V);
        // full level
        const double z0 = zdis_full_at(g, dt, dt2, dt3, V, w);
        const int kcmax = bnd[bounds_index(t, 0, w, 1)];
        const int kc = departure_cell(g.gzh, w, w, kcmax, z0, KMIN, KMAX);
        const bool up1 = (kcmax != w) && (kc == w + 1);
        const double dzw = g.dgz[w];
        double ft[6];
#pragma unroll
        for (int i = 0; i < 6; i++) {
            const double r0 = rhof_cell(cur, t, i);
            const double v = up1 ? (-(r0 * dzw) + rhof_cell(nxt, t, i) * z0) : r0 * z0;
            ft[i] = eps_filter(v);
        }
        fq[t] = ft[0];
#pragma unroll
        for (int i = 1; i < 6; i++) fs[i - 1] += ft[i];
        if (w == KMIN) {
            if (t == 0) {
                precip0 = precip0 - ft[0] / dt;
            } else {
                precip1 = precip1 - ft[0] / dt;
                lh = lh + ft[0] * C_LHF0 / dt;
            }
            prhoe = prhoe - ft[1] / dt;
            prhophi = prhophi - ft[2] / dt;
            prhokin = prhokin - ft[3] / dt - 0.0 / dt;   // the half-level flux at kmin is 0
        }
        // half level (walls 2..94)
        if (w >= KMIN + 1) {
            const double zh0 = zdis_half_at(g, dt, dt2, dt3, V, w);
            const int khmax = bnd[bounds_index(t, 1, w, 1)];
            const int kch = departure_cell(g.gz_shift, w, w, khmax, zh0, KMIN + 1, KMAX);
            const double dzh = g.dgzh[w];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const double hc = cur.hv[t][i], hn = nxt.hv[t][i];
                double v;
                if (kcmax == w)  // region-wide fast path of getflux: rhof(kcell) * zdis0
                    v = ((kch == w + 1) ? hn : hc) * zh0;
                else
                    v = (kch == w + 1) ? (-(hc * dzh) + hn * zh0) : hc * zh0;
                fh[i] += eps_filter(v);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 3; t++) fx[(FX_Q + t) * s.K2] = fq[t];
#pragma unroll
    for (int i = 0; i < 5; i++) fx[(FX_E + i) * s.K2] = fs[i];
    fx[FX_KINV * s.K2] = fh[0];
    fx[FX_W * s.K2] = fh[1];
    f.GPREC[ow] = (0.0 + (-fq[0] / dt)) * 1.0;
    if (w == KMIN) {
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
}

// Sedimentation apply and KIN2EIN of interior cell c of a single-cell region
// (src/mod_precip_transport.cpp:960-1011) from the k_flux workspace: returns rhog,
// rhoge and QR, QS, QG after transport and stores the transported momentum.
__device__ __forceinline__ void sediment_cell(const StepArgs &a, int l, int ij, int c, double &rg_out,
                                              double &e_out, double rq[6])
{
    const int ijdim = a.ijdim;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const GridDev &g = *a.grid;
    const Fields &f = a.f;
    const double *fx = a.ws.flux + (int64_t)l * FLUX_SLOTS * K2 + ij;
    const int64_t p2 = (int64_t)l * K2 + ij;
    const int64_t p3 = (int64_t)l * NQ * K2 + ij;
#define FXV(slot, wall) fx[(int64_t)(slot) * K2 + (int64_t)(wall) * ijdim]

    // tendencies of cell c (walls c and c+1; wall 95 carries zero fluxes)
    const double rc = g.rdgz[c];
    double dq[3];
#pragma unroll
    for (int t = 0; t < 3; t++) dq[t] = -(FXV(FX_Q + t, c + 1) - FXV(FX_Q + t, c)) * rc;
    const double de = -(FXV(FX_E, c + 1) - FXV(FX_E, c)) * rc;
    const double dphi = -(FXV(FX_PHI, c + 1) - FXV(FX_PHI, c)) * rc;
    const double dkinh = -(FXV(FX_KINH, c + 1) - FXV(FX_KINH, c)) * rc;
    const double du = -(FXV(FX_U, c + 1) - FXV(FX_U, c)) * rc;
    const double dv = -(FXV(FX_V, c + 1) - FXV(FX_V, c)) * rc;
    // mass tendencies of cells c+1 and c-1
    double dqp[3] = {0.0, 0.0, 0.0}, dqm[3] = {0.0, 0.0, 0.0};
    if (c + 1 <= KMAX) {
        const double r = g.rdgz[c + 1];
#pragma unroll
        for (int t = 0; t < 3; t++) dqp[t] = -(FXV(FX_Q + t, c + 2) - FXV(FX_Q + t, c + 1)) * r;
    }
    if (c - 1 >= KMIN) {
        const double r = g.rdgz[c - 1];
#pragma unroll
        for (int t = 0; t < 3; t++) dqm[t] = -(FXV(FX_Q + t, c) - FXV(FX_Q + t, c - 1)) * r;
    }
    // half-level tendencies at walls c and c+1 (walls 2..94; the flux of wall 1 is 0)
    double hk_c = 0.0, hw_c = 0.0, hk_p = 0.0, hw_p = 0.0;
    if (c >= KMIN + 1) {
        const double r = g.rdgzh[c];
        hk_c = -(FXV(FX_KINV, c) - FXV(FX_KINV, c - 1)) * r;
        hw_c = -(FXV(FX_W, c) - FXV(FX_W, c - 1)) * r;
    }
    if (c + 1 <= KMAX) {
        const double r = g.rdgzh[c + 1];
        hk_p = -(FXV(FX_KINV, c + 1) - FXV(FX_KINV, c)) * r;
        hw_p = -(FXV(FX_W, c + 1) - FXV(FX_W, c)) * r;
    }
    const double rw_c = FXV(FX_RW, c), rw_p = FXV(FX_RW, c + 1);
#undef FXV

    // pre-transport wall quantities of cells c and c+1 (C2W factors, rhogkin_v)
    const int64_t o = p2 + (int64_t)c * ijdim;
    const double rg_m = a.ws.rhog[o - ijdim], rg_c = a.ws.rhog[o], rg_p = a.ws.rhog[o + ijdim];
    const double gs_m = f.gsgam2[o - ijdim], gs_c = f.gsgam2[o], gs_p = f.gsgam2[o + ijdim];
    const double gsh_c = f.gsgam2h[o], gsh_p = f.gsgam2h[o + ijdim];
    const double c2w0_c = g.afact[c] / gs_c * gsh_c, c2w1_c = g.bfact[c] / gs_m * gsh_c;
    const double c2w0_p = g.afact[c + 1] / gs_p * gsh_p, c2w1_p = g.bfact[c + 1] / gs_c * gsh_p;
    const double kinv0_c = (c >= KMIN + 1) ? 0.5 * (rw_c * rw_c) / (c2w0_c * rg_c + c2w1_c * rg_m) : 0.0;
    const double kinv0_p = (c + 1 <= KMAX) ? 0.5 * (rw_p * rw_p) / (c2w0_p * rg_p + c2w1_p * rg_c) : 0.0;

    const int64_t pc = (int64_t)l * ijdim + ij;
    const double ix = f.ix[pc], iy = f.iy[pc], iz = f.iz[pc];
    const double jx = f.jx[pc], jy = f.jy[pc], jz = f.jz[pc];
    const int64_t oq = p3 + (int64_t)c * ijdim;
    rq[QR] = a.ws.rhogq[oq + QR * K2];
    rq[QS] = a.ws.rhogq[oq + QS * K2];
    rq[QG] = a.ws.rhogq[oq + QG * K2];
    const double re0 = a.ws.rhoge[o];
    const double rvx = f.rhogvx[o], rvy = f.rhogvy[o], rvz = f.rhogvz[o];
    const double z = f.z[o];
    const double w2c0 = g.cfact[c] * gs_c / gsh_p, w2c1 = g.dfact[c] * gs_c / gsh_p;
    const double kinh = 0.5 * (rvx * rvx + rvy * rvy + rvz * rvz) / rg_c;
    double e = re0 + de + dphi;
    const double kinh_u = kinh + dkinh;
    const double kinv_c = kinv0_c + hk_c;
    const double kinv_cp = kinv0_p + hk_p;
    const double rvx_n = rvx + du * ix + dv * jx;
    const double rvy_n = rvy + du * iy + dv * jy;
    const double rvz_n = rvz + du * iz + dv * jz;
    const double rw_n = rw_c + hw_c;
    double rg = rg_c;
    rq[QR] = rq[QR] + dq[0];
    rg = rg + dq[0];
    e = e - dq[0] * C_GRAV * z;
    rq[QS] = rq[QS] + dq[1];
    rg = rg + dq[1];
    e = e - dq[1] * C_GRAV * z;
    rq[QG] = rq[QG] + dq[2];
    rg = rg + dq[2];
    e = e - dq[2] * C_GRAV * z;
    // KIN2EIN: kinetic energy before minus after, with the neighbours' new densities
    const double tmp2 = kinh_u + (w2c0 * kinv_cp + w2c1 * kinv_c);
    const double rg_cp = ((rg_p + dqp[0]) + dqp[1]) + dqp[2];
    const double rg_cm = ((rg_m + dqm[0]) + dqm[1]) + dqm[2];
    const double rw_cp = rw_p + hw_p;
    const double kinh_new = 0.5 * (rvx_n * rvx_n + rvy_n * rvy_n + rvz_n * rvz_n) / rg;
    const double kinv_new_c = (c >= KMIN + 1) ? 0.5 * (rw_n * rw_n) / (c2w0_c * rg + c2w1_c * rg_cm) : 0.0;
    const double kinv_new_cp = (c + 1 <= KMAX) ? 0.5 * (rw_cp * rw_cp) / (c2w0_p * rg_cp + c2w1_p * rg) : 0.0;
    const double tmpk = kinh_new + (w2c0 * kinv_new_cp + w2c1 * kinv_new_c);
    e = e + (tmp2 - tmpk);
    f.rhogvx[o] = rvx_n;
    f.rhogvy[o] = rvy_n;
    f.rhogvz[o] = rvz_n;
    f.rhogw[o] = rw_n;
    rg_out = rg;
    e_out = e;
}

// ---------------------------------------------------------------------------------
// k_finish: one thread per point.  The sedimented state comes from sediment_cell in
// the single-cell regions and from the general sweep elsewhere; then the point-wise
// tail of the step: rediagnosis (src/mod_precip_transport.cpp:1074-1086),
// SATURATION_adjustment with satadjust_liq (src/mod_satadjust.cpp:274-362, 771-896)
// and the final negative_filter (src/mod_mp_nsw6.cpp:671, 1770-1831).
// ---------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_finish(StepArgs a)
{
    const int ijdim = a.ijdim;
    const LevelBlock lb = level_block(blockIdx.x, (ijdim + 255) / 256, KALL, 0, a.lall);
    if (!lb.valid) return;
    const int l = lb.l, k = lb.k;
    const int ij = lb.bx * 256 + threadIdx.x;
    if (ij >= ijdim) return;
    const Fields &f = a.f;
    const DevParams &P = *a.P;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
    const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;
    const bool interior = (k >= KMIN && k <= KMAX);
    const bool general = a.ws.region_flag[l] != 0;

    // State after sedimentation.  QV, QC, QI and the halo levels only receive zero
    // tendencies.
    double rhog, rhoge, rhogq[6];
    if (interior) {
        if (general) {
            rhog = f.rhog[p];
            rhoge = f.rhoge[p];
            rhogq[QR] = f.rhogq[pq + QR * K2];
            rhogq[QS] = f.rhogq[pq + QS * K2];
            rhogq[QG] = f.rhogq[pq + QG * K2];
        } else {
            sediment_cell(a, l, ij, k, rhog, rhoge, rhogq);
        }
    } else {
        rhog = a.ws.rhog[p] + 0.0;
        rhoge = a.ws.rhoge[p] + 0.0;
        rhogq[QR] = a.ws.rhogq[pq + QR * K2] + 0.0;
        rhogq[QS] = a.ws.rhogq[pq + QS * K2] + 0.0;
        rhogq[QG] = a.ws.rhogq[pq + QG * K2] + 0.0;
        if (!general) {
            const int64_t pc = (int64_t)l * ijdim + ij;
            f.rhogvx[p] = f.rhogvx[p] + 0.0 * f.ix[pc] + 0.0 * f.jx[pc];
            f.rhogvy[p] = f.rhogvy[p] + 0.0 * f.iy[pc] + 0.0 * f.jy[pc];
            f.rhogvz[p] = f.rhogvz[p] + 0.0 * f.iz[pc] + 0.0 * f.jz[pc];
            f.rhogw[p] = f.rhogw[p] + 0.0;
        }
    }
    rhogq[QV] = a.ws.rhogq[pq + QV * K2] + 0.0;
    rhogq[QC] = a.ws.rhogq[pq + QC * K2] + 0.0;
    rhogq[QI] = a.ws.rhogq[pq + QI * K2] + 0.0;
    const double gsgam2 = f.gsgam2[p];

    // rediagnosis at the end of precip_transport_new
    double q[6];
#pragma unroll
    for (int n = 0; n < 6; n++) q[n] = rhogq[n] / rhog;
    double rho = rhog / gsgam2;
    const double ein = rhoge / rhog;
    double qd = 1.0;
#pragma unroll
    for (int n = 1; n <= 5; n++) qd = qd - q[n];
    double cv = 0.0, qd2 = 1.0;
#pragma unroll
    for (int n = 1; n <= 5; n++) {
        cv = cv + q[n] * C_CVW;
        qd2 = qd2 - q[n];
    }
    cv = cv + qd2 * C_CVdry;
    double tem = ein / cv;
    double pre = rho * tem * (qd2 * C_Rdry + q[QV] * C_Rvap);

    // SATURATION_adjustment, ice_adjust = false
    const double ein_moist = rhoge / rhog + q[QV] * C_LHV0;
    const double qsum = q[QV] + q[QC];
    q[QV] = qsum;
    q[QC] = 0.0;
    double CVtot = qd * C_CVdry;
#pragma unroll
    for (int n = 1; n <= 5; n++) CVtot = CVtot + q[n] * C_CVW;
    const double rho_sa = rhog / gsgam2;
    tem = (ein_moist - q[QV] * C_LHV0) / CVtot;
    if (interior) {
        double qdl = 1.0;
#pragma unroll
        for (int n = 1; n <= 5; n++) qdl = qdl - q[n];
        double rtem = 1.0 / dmax(tem, TEM_MIN);
        double psat = C_PSAT0 * 1.0 * exp(P.LovR_liq * (P.RTEM00 - rtem));
        double qsat = psat / (rho_sa * C_Rvap * tem);
        if (qsum - qsat > C_EPS) {
            bool converged = false, nan_break = false;
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
                const double dqsat_dT = (P.LovR_liq / (tem * tem) + 0.0 / tem) * qsat;
                const double dCVtot_dT = dqsat_dT * (C_CVW - C_CVW);
                const double dEmoist_dT = tem * dCVtot_dT + CVt + dqsat_dT * C_LHV0;
                const double dtemp = (Emoist_new - ein_moist) / dEmoist_dT;
                tem = tem - dtemp;
                if (fabs(dtemp) < P.dtemp_criteria) {
                    converged = true;
                    break;
                }
                if (tem * 0.0 != 0.0) {
                    nan_break = true;
                    break;
                }
            }
            if (!converged) {
                const int code = nan_break ? 2 : 1;
                if (atomicAdd(&a.ws.report[code], 1) == 0) {
                    a.ws.report[3] = l;
                    a.ws.report[4] = k;
                    a.ws.report[5] = ij;
                }
                atomicOr(&a.ws.report[0], 1);
            }
        }
    }
    rhogq[QV] = rhog * q[QV];
    rhogq[QC] = rhog * q[QC];
    double CVtot2 = qd * C_CVdry;
#pragma unroll
    for (int n = 1; n <= 5; n++) CVtot2 = CVtot2 + q[n] * C_CVW;
    rhoge = rhog * tem * CVtot2;

    if (interior) {
        // final negative_filter
        double diffq = 0.0;
#pragma unroll
        for (int n = 2; n <= 5; n++) {
            diffq += rhogq[n];
            rhogq[n] = dmax(rhogq[n], 0.0);
        }
#pragma unroll
        for (int n = 2; n <= 5; n++) diffq -= rhogq[n];
        rhogq[QV] += diffq;
        diffq = rhogq[QV];
        rhogq[QV] = dmax(rhogq[QV], 0.0);
        diffq -= rhogq[QV];
        rhog = rhog * (1.0 - diffq);
        rho = rhog / gsgam2;
#pragma unroll
        for (int n = 1; n <= 5; n++) q[n] = rhogq[n] / rhog;
        double qdn = 1.0;
#pragma unroll
        for (int n = 1; n <= 5; n++) qdn = qdn - q[n];
        double cva = qdn * C_CVdry;
#pragma unroll
        for (int n = 1; n <= 5; n++) cva = cva + q[n] * C_CVW;
        rhoge = tem * rhog * cva;
        pre = rho * (qdn * C_Rdry + q[QV] * C_Rvap) * tem;
    }

    f.rhog[p] = rhog;
    f.rhoge[p] = rhoge;
    f.rho[p] = rho;
    f.tem[p] = tem;
    f.pre[p] = pre;
    f.qd[p] = qd;
#pragma unroll
    for (int n = 0; n < 6; n++) {
        f.rhogq[pq + n * K2] = rhogq[n];
        f.q[pq + n * K2] = q[n];
    }
    f.re_solid[p] = 20.0E-6;
    f.GDCLW[p] = 0.0 * 1.0;
    f.GDCFRC[p] = 0.0 * 1.0;
}

template <bool GENERAL>
__global__ void __launch_bounds__(64) k_sweep(StepArgs a)
{
    const int ijdim = a.ijdim;
    const int l = blockIdx.y;
    if ((a.ws.region_flag[l] != 0) != GENERAL) return;
    const int ij = blockIdx.x * 64 + threadIdx.x;
    if (ij >= ijdim) return;
    const GridDev &g = *a.grid;
    const Fields &f = a.f;
    const int *bnd = a.ws.bounds + (size_t)l * BOUNDS_PER_REGION;
    const double dt = a.dt, dt2 = a.dt2, dt3 = a.dt3_6;

    SweepCtx s;
    s.a = &a;
    s.l = l;
    s.ij = ij;
    s.ijdim = ijdim;
    s.K2 = (int64_t)KALL * ijdim;
    s.p2 = (int64_t)l * s.K2 + ij;
    s.p3 = (int64_t)l * NQ * s.K2 + ij;
    s.pv = (int64_t)l * 3 * s.K2 + ij;
    const int64_t pc = (int64_t)l * ijdim + ij;
    s.ix = f.ix[pc]; s.iy = f.iy[pc]; s.iz = f.iz[pc];
    s.jx = f.jx[pc]; s.jy = f.jy[pc]; s.jz = f.jz[pc];
    const double *vt = a.ws.vt + s.pv;
    // deferred momentum: slots 0..3 of the region's flux block (k_flux skips general regions)
    double *mom = GENERAL ? (a.ws.flux + (int64_t)l * FLUX_SLOTS * s.K2 + ij) : nullptr;

    // window: cells w (cur) and w+1 (nxt); finalisation records for w-2 .. w+1
    TCell cur, nxt;
    FCell fm2, fm1, f0, fp1;
    double rhog_cur, q_cur[3], gs_cur;
    double rhog_nxt, q_nxt[3], gs_nxt;
    WindWindow wind[3];
    double fprev[3][6], fprevh[3][2];
    double rhog_new_prev = 0.0;       // rhog after sedimentation of the last finalised cell
    double precip0 = 0.0, precip1 = 0.0, lh = 0.0, prhoe = 0.0, prhophi = 0.0, prhokin = 0.0;

    // cell 0 (halo) and cell 1
    {
        TCell c0;
        double rg0, q0[3], gs0;
        load_tcell(s, 0, c0, rg0, q0, gs0);
        FCell fdummy;
        load_wall(s, g, 0, rg0, rg0, q0, q0, gs0, gs0, c0, fdummy);
        fm1 = fdummy;     // cell 0 record (w-1 for w = 1)
        load_tcell(s, 1, cur, rhog_cur, q_cur, gs_cur);
        load_wall(s, g, 1, rhog_cur, rg0, q_cur, q0, gs_cur, gs0, cur, f0);
        rhog_new_prev = 0.0;
    }
#pragma unroll
    for (int t = 0; t < 3; t++) {
        wind_init(wind[t], vt[t * s.K2 + 1 * (int64_t)ijdim]);
#pragma unroll
        for (int i = 0; i < 6; i++) fprev[t][i] = 0.0;    // f(0) = 0
        fprevh[t][0] = fprevh[t][1] = 0.0;                // f_h(1) = 0
    }

    for (int w = KMIN; w <= KMAX + 1; w++) {
        double fcur[3][6], fcurh[3][2];
        if (w <= KMAX) {
            // ---- load cell w+1
            load_tcell(s, w + 1, nxt, rhog_nxt, q_nxt, gs_nxt);
            load_wall(s, g, w + 1, rhog_nxt, rhog_cur, q_nxt, q_cur, gs_nxt, gs_cur, nxt, fp1);
#pragma unroll
            for (int t = 0; t < 3; t++) {
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
                                else rf = GENERAL ? general_rhof(s, g, k2, t, i) : 0.0;
                                acc = acc + rf * fa;
                                zd = z0 - fa;
                            }
                        }
                        double rfc;
                        if (kc == w) rfc = rhof_of(cur, t, i);
                        else if (kc == w + 1) rfc = rhof_of(nxt, t, i);
                        else rfc = GENERAL ? general_rhof(s, g, kc, t, i) : 0.0;
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
                                else hvv = GENERAL ? general_hv(s, g, k2, t, i) : 0.0;
                                acc = acc + hvv * fa;
                                zd = z0 - fa;
                            }
                        }
                        double hvc;
                        if (kc == w) hvc = cur.hv[t][i];
                        else if (kc == w + 1) hvc = nxt.hv[t][i];
                        else hvc = GENERAL ? general_hv(s, g, kc, t, i) : 0.0;
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
                precip1 = precip1 - fcur[1][0] / dt;
                lh = lh + fcur[1][0] * C_LHF0 / dt;
                precip1 = precip1 - fcur[2][0] / dt;
                lh = lh + fcur[2][0] * C_LHF0 / dt;
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    prhoe = prhoe - fcur[t][1] / dt;
                    prhophi = prhophi - fcur[t][2] / dt;
                    prhokin = prhokin - fcur[t][3] / dt - 0.0 / dt;   // f_h(kmin) = 0
                }
            }
        } else {
#pragma unroll
            for (int t = 0; t < 3; t++) {
#pragma unroll
                for (int i = 0; i < 6; i++) fcur[t][i] = 0.0;
                fcurh[t][0] = fcurh[t][1] = 0.0;
            }
        }

        // ---- full-level divergence of cell w-1 (k = 1..94); its record is fm1
        if (w - 1 >= KMIN) {
            const double r = g.rdgz[w - 1];
#pragma unroll
            for (int t = 0; t < 3; t++) fm1.drhoq[t] = -(fcur[t][0] - fprev[t][0]) * r;
#pragma unroll
            for (int t = 0; t < 3; t++) {
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
            for (int t = 0; t < 3; t++) {
                f0.drhokinv = f0.drhokinv - (fcurh[t][0] - fprevh[t][0]) * r;
                f0.drhogw = f0.drhogw - (fcurh[t][1] - fprevh[t][1]) * r;
            }
        }
#pragma unroll
        for (int t = 0; t < 3; t++) {
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
            const double rvx_n = rvx + fm2.drhogu * s.ix + fm2.drhogv * s.jx;
            const double rvy_n = rvy + fm2.drhogu * s.iy + fm2.drhogv * s.jy;
            const double rvz_n = rvz + fm2.drhogu * s.iz + fm2.drhogv * s.jz;
            const double rw_n = fm2.rhogw + fm2.drhogw;
            double rg = fm2.rhog;
            // nq = 1..5: QC(0), QR, QI(0), QS, QG
            rq[QC] = rq[QC] + 0.0;
            rq[QR] = rq[QR] + fm2.drhoq[0];
            rg = rg + fm2.drhoq[0];
            e = e - fm2.drhoq[0] * C_GRAV * z;
            rq[QI] = rq[QI] + 0.0;
            rq[QS] = rq[QS] + fm2.drhoq[1];
            rg = rg + fm2.drhoq[1];
            e = e - fm2.drhoq[1] * C_GRAV * z;
            rq[QG] = rq[QG] + fm2.drhoq[2];
            rg = rg + fm2.drhoq[2];
            e = e - fm2.drhoq[2] * C_GRAV * z;
            rq[QV] = rq[QV] + 0.0;
            // KIN2EIN (:991-1010)
            const double tmp2 = kinh_u + (w2c0 * kinv_cp + w2c1 * kinv_c);
            const double rg_cp = ((fm1.rhog + fm1.drhoq[0]) + fm1.drhoq[1]) + fm1.drhoq[2];
            const double rw_cp = fm1.rhogw + fm1.drhogw;
            const double kinh_new = 0.5 * (rvx_n * rvx_n + rvy_n * rvy_n + rvz_n * rvz_n) / rg;
            const double kinv_new_c =
                (c >= KMIN + 1) ? 0.5 * (rw_n * rw_n) / (fm2.c2w0 * rg + fm2.c2w1 * rhog_new_prev) : 0.0;
            const double kinv_new_cp =
                (c + 1 <= KMAX) ? 0.5 * (rw_cp * rw_cp) / (fm1.c2w0 * rg_cp + fm1.c2w1 * rg) : 0.0;
            const double tmpk = kinh_new + (w2c0 * kinv_new_cp + w2c1 * kinv_new_c);
            e = e + (tmp2 - tmpk);
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
            store_sedimented(s, c, rg, rq, e);
        }

        // ---- shift the window
        fm2 = fm1;
        fm1 = f0;
        f0 = fp1;
        cur = nxt;
        rhog_cur = rhog_nxt;
        gs_cur = gs_nxt;
#pragma unroll
        for (int t = 0; t < 3; t++) q_cur[t] = q_nxt[t];
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
        const double rvx_n = rvx + fm2.drhogu * s.ix + fm2.drhogv * s.jx;
        const double rvy_n = rvy + fm2.drhogu * s.iy + fm2.drhogv * s.jy;
        const double rvz_n = rvz + fm2.drhogu * s.iz + fm2.drhogv * s.jz;
        const double rw_n = fm2.rhogw + fm2.drhogw;
        double rg = fm2.rhog;
        rq[QC] = rq[QC] + 0.0;
        rq[QR] = rq[QR] + fm2.drhoq[0];
        rg = rg + fm2.drhoq[0];
        e = e - fm2.drhoq[0] * C_GRAV * z;
        rq[QI] = rq[QI] + 0.0;
        rq[QS] = rq[QS] + fm2.drhoq[1];
        rg = rg + fm2.drhoq[1];
        e = e - fm2.drhoq[1] * C_GRAV * z;
        rq[QG] = rq[QG] + fm2.drhoq[2];
        rg = rg + fm2.drhoq[2];
        e = e - fm2.drhoq[2] * C_GRAV * z;
        rq[QV] = rq[QV] + 0.0;
        const double tmp2 = kinh_u + (w2c0 * kinv_cp + w2c1 * kinv_c);
        const double kinh_new = 0.5 * (rvx_n * rvx_n + rvy_n * rvy_n + rvz_n * rvz_n) / rg;
        const double kinv_new_c = 0.5 * (rw_n * rw_n) / (fm2.c2w0 * rg + fm2.c2w1 * rhog_new_prev);
        const double tmpk = kinh_new + (w2c0 * 0.0 + w2c1 * kinv_new_c);
        e = e + (tmp2 - tmpk);
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
        store_sedimented(s, c, rg, rq, e);
    }
    // halo cells 0 and 95: zero tendencies (their scalars are finished by k_finish)
#pragma unroll
    for (int hc = 0; hc < 2; hc++) {
        const int c = hc ? (KALL - 1) : 0;
        const int64_t o = s.p2 + (int64_t)c * ijdim;
        const double rvx_n = f.rhogvx[o] + 0.0 * s.ix + 0.0 * s.jx;
        const double rvy_n = f.rhogvy[o] + 0.0 * s.iy + 0.0 * s.jy;
        const double rvz_n = f.rhogvz[o] + 0.0 * s.iz + 0.0 * s.jz;
        const double rw_n = f.rhogw[o] + 0.0;
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
extern "C" hipError_t nsw6_launch_step(const StepArgs *args, int precision, hipStream_t stream,
                                       hipEvent_t *ev /* NSW6_STEP_EVENTS events or NULL */)
{
    const StepArgs &a = *args;
    const dim3 grid_p((a.ijdim + 255) / 256, KALL, a.lall);
    const dim3 grid_c((a.ijdim + 63) / 64, a.lall);
    const dim3 grid_e(level_grid((a.ijdim + EXT_COLS_PER_BLOCK - 1) / EXT_COLS_PER_BLOCK, KMAX - KMIN + 1, a.lall));
    const dim3 grid_l(level_grid((a.ijdim + 255) / 256, KALL, a.lall));
    if (ev) (void)hipEventRecord(ev[0], stream);
    if (precision == 1)
        hipLaunchKernelGGL(k_rates<float>, grid_p, dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(k_rates<double>, grid_p, dim3(256), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[1], stream);
    hipLaunchKernelGGL(k_extremes, grid_e, dim3(EXT_THREADS), 0, stream, a);
    hipLaunchKernelGGL(k_bounds, dim3(6, a.lall), dim3(192), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[2], stream);
    hipLaunchKernelGGL(k_flux, grid_l, dim3(256), 0, stream, a);
    hipLaunchKernelGGL(k_sweep<true>, grid_c, dim3(64), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[3], stream);
    hipLaunchKernelGGL(k_finish, grid_l, dim3(256), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[4], stream);
    return hipGetLastError();
}
