// nsw6_finish.hip -- k_finish, the last kernel of the step, in a unit of its own: the
// flux divergences, KIN2EIN, the rediagnosis, the saturation adjustment and the final
// negative filter, one thread per point (the sources are cited at finish_point).
//
// The unit is built with the FP64 step's floating-point options but without the
// LDS-table exp and log of nsw6_math.h: the kernel calls exp once per point (the
// saturation check) and a few times more at the points the Newton iteration visits,
// so it gains nothing from the tables, while their copy into LDS puts a memory round
// trip and a barrier in front of every block's loads.  With the device library's exp
// the kernel takes 1.465 instead of 1.49 ms at 169,000 x 96 (memory-bound: skipping
// the Newton iteration altogether leaves its time unchanged).
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "nsw6_device.h"
#include "nsw6_columns.h"
#include "nsw6_internal.h"
#include "nsw6_step.h"

using namespace nsw6;

// ---------------------------------------------------------------------------------
// k_finish: one thread per point.  The sedimented state from sediment() on the wall
// fluxes, then the point-wise tail (rediagnosis, saturation adjustment, final
// negative filter).
// ---------------------------------------------------------------------------------
template <int NT, typename Fetch>
__device__ __forceinline__ void finish_point(const StepArgs &a, int l, int k, int ij, const Fetch &FXV)
{
    const int ijdim = a.ijdim;
    const Fields &f = a.f;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
    const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;
    const bool interior = (k >= KMIN && k <= KMAX);

    // State after sedimentation.  QV, QC, QI ("3WATER") and the halo levels only
    // receive zero tendencies.  The momenta are stored with the other outputs at the
    // end: every store comes after every load, so that the uniform reads of the grid
    // and the parameters stay on the scalar cache.
    TailIn<double> x;
    // read by every path: loaded before the branches, with the path's own loads
    x.rhogq[QV] = a.ws.rhogq[pq + QV * K2] + 0.0;
    x.rhogq[QC] = a.ws.rhogq[pq + QC * K2] + 0.0;
    if (NT == 3) x.rhogq[QI] = a.ws.rhogq[pq + QI * K2] + 0.0;
    x.gsgam2 = f.gsgam2[p];
    double mom[4];   // rhogvx, rhogvy, rhogvz, rhogw after sedimentation
    if (interior) {
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
    } else {
        x.rhog = a.ws.rhog[p] + 0.0;
        x.rhoge = a.ws.rhoge[p] + 0.0;
#pragma unroll
        for (int t = 0; t < NT; t++) x.rhogq[sed_nq<NT>(t)] = a.ws.rhogq[pq + sed_nq<NT>(t) * K2] + 0.0;
        const int64_t pc = (int64_t)l * ijdim + ij;
        mom[0] = f.rhogvx[p] + 0.0 * f.ix[pc] + 0.0 * f.jx[pc];
        mom[1] = f.rhogvy[p] + 0.0 * f.iy[pc] + 0.0 * f.jy[pc];
        mom[2] = f.rhogvz[p] + 0.0 * f.iz[pc] + 0.0 * f.jz[pc];
        mom[3] = (a.kin2kin && k == KALL - 1) ? kin2kin_top_wall<NT>(a, l, ij, x.rhog, FXV) : f.rhogw[p] + 0.0;
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
    f.rhogvx[p] = mom[0];
    f.rhogvy[p] = mom[1];
    f.rhogvz[p] = mom[2];
    f.rhogw[p] = mom[3];
}

// Three waves per SIMD: with the registers four waves leave (126 of 128 used) the
// kernel is about 1 % slower.
template <int NT, typename WS = double>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_finish(StepArgs a)
{
    const int ijdim = a.ijdim;
    const LevelBlock lb = level_block<FINISH_BAND>(blockIdx.x, (ijdim + 255) / 256, KALL, 0, a.lall);
    if (!lb.valid) return;
    const int ij = lb.bx * 256 + threadIdx.x;
    if (ij >= ijdim) return;
    finish_point<NT>(a, lb.l, lb.k, ij, FluxFromMemory<NT, WS>(a, lb.l, ij));
}


extern "C" hipError_t nsw6_launch_finish(const StepArgs *args, int nt, int ws_float, hipStream_t stream)
{
    const StepArgs &a = *args;
    const dim3 grid_l(level_grid((a.ijdim + 255) / 256, KALL, a.lall));
    if (nt == 4) {
        if (ws_float)
            hipLaunchKernelGGL((k_finish<4, float>), grid_l, dim3(256), 0, stream, a);
        else
            hipLaunchKernelGGL((k_finish<4, double>), grid_l, dim3(256), 0, stream, a);
    } else {
        if (ws_float)
            hipLaunchKernelGGL((k_finish<3, float>), grid_l, dim3(256), 0, stream, a);
        else
            hipLaunchKernelGGL((k_finish<3, double>), grid_l, dim3(256), 0, stream, a);
    }
    return hipGetLastError();
}
