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
//   k_flux_general  the same walls of the general regions (departures beyond the
//              next cell, upward, or out of the domain) over the regions' kcell ranges.
//   k_finish   point-wise: flux divergences and KIN2EIN (:960-1011), rediagnosis
//              (:1074-1086), saturation adjustment
//              (src/mod_satadjust.cpp:274-362,771-896) and the final negative filter;
//              in nsw6_finish.hip, a unit of its own.
// Every access of a wavefront is one contiguous 512-byte segment of a level (ij
// fastest).  No MFMA: nothing here is a dense contraction.
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

// double exp and log from the LDS tables of nsw6_math.h (k_rates)
#define NSW6_TABLE_MATH 1
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
template <int NT, typename WS = double>
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
    const WS *vt = ws_vt<WS>(a) + (int64_t)l * NT * K2;

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
        // k_flux's closed forms need every departure cell in [k, k+1].
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
// The transport as two point-wise kernels: k_flux (one thread per wall) and k_finish
// (one thread per cell).  In a single-cell region (every departure cell in {k, k+1},
// checked by k_bounds) each flux depends only on cells w-1, w, w+1 and the fall speeds
// Vt(w-3..w+1); k_flux_general takes the walls of the other regions over their kcell
// ranges.  Both write the same flux workspace, from which k_finish forms every
// cell's divergence.  The arithmetic is in nsw6_step.h, the column gathers and the
// per-wall fluxes (flux_wall, flux_wall_general) in nsw6_columns.h.
// ---------------------------------------------------------------------------------
// Three waves per SIMD: the cells, the fall speeds and the kcell bounds of the wall are
// all requested up front (one memory round trip per wave) at 150 VGPRs without spills.
template <int NT, typename WS = double>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_flux(StepArgs a)
{
    const int ijdim = a.ijdim;
    const LevelBlock lb = level_block<FLUX_BAND>(blockIdx.x, (ijdim + 255) / 256, KALL, 0, a.lall);
    if (!lb.valid) return;
    if (a.ws.region_flag[lb.l] != 0) return;
    const int ij = lb.bx * 256 + threadIdx.x;
    if (ij >= ijdim) return;
    flux_wall<NT, WS>(a, lb.l, lb.k, ij);
}

// The walls of the general regions (departures beyond the next cell, upward, or out
// of the domain): the swept cells of the region's kcell ranges are loaded and derived
// one after another (the ranges are uniform over a region, so the loop is too).  Each
// block of the (general_blocks(ijdim), lall) grid takes GENERAL_TILES of its region's
// (level, 256-column block) tiles, which keeps the launch cheap when no region needs
// it (a tenth of the blocks of k_flux's grid).  The blocks that one XCD receives
// (every NUM_XCD-th) take consecutive tiles of one chunk, i.e. neighbouring levels of
// the same columns, whose cells they share through that XCD's L2.  Two waves per SIMD
// (223 registers, no spills); forced to three it spills 56 registers and takes 1.8x
// as long.
constexpr int GENERAL_TILES = 4;
static inline int general_blocks(int ijdim)
{
    const int total = ((ijdim + 255) / 256) * KALL;
    const int per_xcd = (total + NUM_XCD * GENERAL_TILES - 1) / (NUM_XCD * GENERAL_TILES);
    return per_xcd * NUM_XCD;
}
template <int NT, typename WS = double>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_flux_general(StepArgs a)
{
    const int l = blockIdx.y;
    if (a.ws.region_flag[l] == 0) return;
    const int ijdim = a.ijdim;
    const int total = ((ijdim + 255) / 256) * KALL;
    const int per = (total + NUM_XCD - 1) / NUM_XCD;
    const int xcd = blockIdx.x % NUM_XCD, lanes = gridDim.x / NUM_XCD;
    for (int j = blockIdx.x / NUM_XCD; j < per; j += lanes) {
        const int L = xcd * per + j;
        if (L >= total) break;
        const int ij = (L / KALL) * 256 + threadIdx.x;
        if (ij < ijdim) flux_wall_general<NT, WS>(a, l, L % KALL, ij);
    }
}

// ---------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------
// The transport of NT sedimenting tracers: extremes and kcell bounds, the wall fluxes
// of the single-cell and of the general regions, then the point-wise finish.  WS is the
// storage of the fall speeds and the wall fluxes: double, or float for the mixed form.
template <int NT, typename WS> static void launch_transport(const StepArgs &a, hipStream_t stream, hipEvent_t *ev)
{
    const dim3 grid_e(level_grid((a.ijdim + EXT_COLS_PER_BLOCK - 1) / EXT_COLS_PER_BLOCK, KMAX - KMIN + 1, a.lall));
    const dim3 grid_l(level_grid((a.ijdim + 255) / 256, KALL, a.lall));
    hipLaunchKernelGGL((k_extremes<NT, WS>), grid_e, dim3(EXT_THREADS), 0, stream, a);
    hipLaunchKernelGGL(k_bounds<NT>, dim3(2 * NT, a.lall), dim3(192), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[2], stream);
    hipLaunchKernelGGL((k_flux<NT, WS>), grid_l, dim3(256), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[3], stream);
    hipLaunchKernelGGL((k_flux_general<NT, WS>), dim3(general_blocks(a.ijdim), a.lall), dim3(256), 0, stream, a);
    if (ev) (void)hipEventRecord(ev[4], stream);
    (void)nsw6_launch_finish(&a, NT, sizeof(WS) == sizeof(float), stream);   // errors surface at hipGetLastError
    if (ev) (void)hipEventRecord(ev[5], stream);
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
    if (precision == 1) {
        if (a.four_water)
            launch_transport<4, float>(a, stream, ev);
        else
            launch_transport<3, float>(a, stream, ev);
    } else {
        if (a.four_water)
            launch_transport<4, double>(a, stream, ev);
        else
            launch_transport<3, double>(a, stream, ev);
    }
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

// The FP64 step up to the wall fluxes (the forward part that the reverse passes
// recompute).
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
    if (a.four_water) {
        hipLaunchKernelGGL(k_extremes<4>, grid_e, dim3(EXT_THREADS), 0, stream, a);
        hipLaunchKernelGGL(k_bounds<4>, dim3(8, a.lall), dim3(192), 0, stream, a);
        hipLaunchKernelGGL(k_flux<4>, grid_l, dim3(256), 0, stream, a);
        hipLaunchKernelGGL(k_flux_general<4>, dim3(general_blocks(a.ijdim), a.lall), dim3(256), 0, stream, a);
    } else {
        hipLaunchKernelGGL(k_extremes<3>, grid_e, dim3(EXT_THREADS), 0, stream, a);
        hipLaunchKernelGGL(k_bounds<3>, dim3(6, a.lall), dim3(192), 0, stream, a);
        hipLaunchKernelGGL(k_flux<3>, grid_l, dim3(256), 0, stream, a);
        hipLaunchKernelGGL(k_flux_general<3>, dim3(general_blocks(a.ijdim), a.lall), dim3(256), 0, stream, a);
    }
    return hipGetLastError();
}
