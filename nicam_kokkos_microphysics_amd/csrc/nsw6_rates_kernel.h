// nsw6_rates_kernel.h -- the point-wise rate kernel k_rates<R, DIAG, INDIR> (negative
// filter, psat/qsat, Bergeron, the 49 process rates and the state update,
// src/mod_mp_nsw6.cpp:439-589), shared by the FP64 unit (nsw6_kernels.hip) and the
// mixed-precision unit (nsw6_rates_mixed.hip), which are built with different
// floating-point options.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsw6_device.h"
#include "nsw6_internal.h"
#include "nsw6_step.h"
#include "nsw6_tables.h"

namespace nsw6 {

// ---------------------------------------------------------------------------------
// k_rates
// ---------------------------------------------------------------------------------
// DIAG = false (the forward part of the reverse mode) skips the effective radii and
// cloud-top diagnostics, which carry no cotangent.
template <typename R, bool DIAG, bool INDIR = false>
__global__ void __launch_bounds__(256) k_rates(StepArgs a)
{
    load_math_tables();
    const int ijdim = a.ijdim;
    const int l = blockIdx.z, k = blockIdx.y;
    const int ij = blockIdx.x * blockDim.x + threadIdx.x;

    // Region-wide reductions of this step start from their neutral values.
    if (blockIdx.x == 0 && k == 0) {
        if (threadIdx.x == 0) a.ws.region_flag[l] = 0;
        if (l == 0 && !a.keep_report && threadIdx.x < REPORT_INTS)
            a.ws.report[threadIdx.x] = (threadIdx.x >= 3) ? -1 : 0;
    }
    if (ij >= ijdim) return;

    const int64_t p = ((int64_t)l * KALL + k) * ijdim + ij;
    const int64_t K2 = (int64_t)KALL * ijdim;
    const int64_t pq = (int64_t)l * NQ * K2 + (int64_t)k * ijdim + ij;   // + n*K2
    // fall speeds [l][NT][k][ij] of QR, QS, QG ("3WATER") or QR, QI, QS, QG ("4WATER")
    const int nt = a.four_water ? 4 : 3;
    const int64_t pv = (int64_t)l * nt * K2 + (int64_t)k * ijdim + ij;   // + t*K2
    const Fields &f = a.f;
    // the fall speeds in the rate type: FP64, or FP32 for the mixed form's transport
    R *vt = reinterpret_cast<R *>(a.ws.vt);

    if (k == 0 || k == KALL - 1) {
        a.ws.rhog[p] = f.rhog[p];
        a.ws.rhoge[p] = f.rhoge[p];
#pragma unroll
        for (int n = 0; n < NQ; n++) a.ws.rhogq[pq + n * K2] = f.rhogq[pq + n * K2];
        for (int t = 0; t < nt; t++) vt[pv + t * K2] = R(0);
        if (DIAG) {
            f.re_liquid[p] = 0.0;
            f.re_cld[p] = 0.0;
        }
        return;
    }

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
    const NcPoint nc = INDIR ? nc_indirect(P, f.unccn[p]) : nc_default(P);
    PointOut<double> y;
    point_rates<double, R>(P, a.dt, x, tun, c_a1_tab, c_a2_tab, y, nc);
    // the effective radii before any store: after one, the compiler would read the
    // parameters they use through the vector memory path instead of the scalar cache
    double rceff = 0.0, rceff_cld = 0.0, rct = 0.0, tct = 0.0, csw = 0.0;
    if (DIAG) effective_radius(P, y.sw, y.rho, x.tem, y.q, nc.Nc, rceff, rceff_cld, rct, tct, csw);

    a.ws.rhog[p] = y.rhog;
    a.ws.rhoge[p] = y.rhoge;
#pragma unroll
    for (int n = 0; n < NQ; n++) a.ws.rhogq[pq + n * K2] = y.rhogq[n];
    if (a.four_water) {
        vt[pv] = R(y.vt[0]);
        vt[pv + K2] = R(y.vti);
        vt[pv + 2 * K2] = R(y.vt[1]);
        vt[pv + 3 * K2] = R(y.vt[2]);
    } else {
#pragma unroll
        for (int t = 0; t < 3; t++) vt[pv + t * K2] = R(y.vt[t]);
    }
    if (!DIAG) return;
    f.re_liquid[p] = rceff;
    f.re_cld[p] = rceff_cld;
    if (a.cloudtop_intended) {
        // integer max: the same level whatever the order of the levels' threads
        if (csw != 0.0) atomicMax(&a.ws.cloud_top[(int64_t)l * ijdim + ij], k);
    } else if (k == KMAX) {
        // every level writes the cloud-top row in the reference; the serial order leaves kmax
        const int64_t pc = (int64_t)l * ijdim + ij;
        f.rctop[pc] = rct;
        f.tctop[pc] = tct;
        const double sw = step_pos(tct - C_TEM00);
        f.rwtop[pc] = (sw)*rct + (1.0 - sw) * C_UNDEF;
    }
}

}  // namespace nsw6
