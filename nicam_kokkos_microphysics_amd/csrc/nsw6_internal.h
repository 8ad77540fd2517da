// nsw6_internal.h -- structures shared by the HIP kernels and the host ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsw6_device.h"

namespace nsw6 {

// Device pointers of the 49 mp_driver arrays (layout of include/nsw6.h).
struct Fields {
    double *rhog, *rhogvx, *rhogvy, *rhogvz, *rhogw, *rhoge, *rhogq;
    double *vx, *vy, *vz, *w, *unccn, *rho, *tem, *pre, *q, *qd;
    double *precip, *ISO1_precip, *ISO2_precip;
    double *precip_rhoe, *precip_lh_heat, *precip_rhophi, *precip_rhokin;
    double *re_liquid, *re_solid, *re_cld, *rctop, *rwtop, *tctop;
    double *frhoge_af, *frhogqv_af, *frhoge_rad, *rhogqke;
    double *gsgam2, *gsgam2h, *gam2, *gam2h;
    double *ix, *iy, *iz, *jx, *jy, *jz;
    double *z, *zh;
    double *GDCLW, *GDCFRC, *GPREC, *CBMFX;
};

// Vertical grid on the device (GRD_* plus GRD_gz_shift of precip_transport_new).
struct GridDev {
    double gz[KALL], gzh[KALL], gz_shift[KALL];
    double dgz[KALL], dgzh[KALL], rdgz[KALL], rdgzh[KALL];
    double afact[KALL], bfact[KALL], cfact[KALL], dfact[KALL];
    double rsum_gz[KALL], rsum_gzh[KALL];   // 1/(dgz(k-1)+dgz(k)), 1/(dgzh(k-1)+dgzh(k))
};

// Per-region reductions: [tracer 3][level type 2: full, half][k 96][max, min]
constexpr int EXT_PER_REGION = 3 * 2 * KALL * 2;
// k_extremes: one column per lane and one tracer per wavefront; per-wave partial
// extremes are stored as [l][k][12][wave]
constexpr int EXT_COLS_PER_WAVE = 64;
constexpr int BOUNDS_PER_REGION = EXT_PER_REGION;
__host__ __device__ constexpr int ext_index(int t, int type, int k, int mm)
{
    return ((t * 2 + type) * KALL + k) * 2 + mm;
}
__host__ __device__ constexpr int bounds_index(int t, int type, int k, int mm) { return ext_index(t, type, k, mm); }
// report ints: [0] any failure, [1] not converged, [2] NaN, [3] l, [4] k, [5] ij
constexpr int REPORT_INTS = 6;

// Context-owned device workspace (allocated once for the maximum column count).
struct Workspace {
    double *rhog, *rhoge, *rhogq;   // state after the rate update (pre-sedimentation)
    double *vt;                     // fall speeds of QR, QS, QG  [l][3][k][ij]
    double *mom;                    // deferred momentum of the general sweep [l][4][k][ij]
    double *part;                   // [l][KALL][12][waves per region] partial extremes
    int *bounds;                    // [l][BOUNDS_PER_REGION]
    int *region_flag;               // [l] 1 = region needs the general sweep
    int *report;                    // [REPORT_INTS]
};

struct StepArgs {
    Fields f;
    Workspace ws;
    DevParams P;
    const GridDev *grid;
    int lall, ijdim;
    double dt, dt2, dt3, dt3_6;   // dt, dt^2, dt^3, dt^3/6
};

}  // namespace nsw6

// events recorded by nsw6_launch_step: before k_rates, after k_rates, after
// k_extremes+k_bounds, after the sweeps, after k_adjust
constexpr int NSW6_STEP_EVENTS = 5;

extern "C" hipError_t nsw6_launch_step(const nsw6::StepArgs *args, int precision, hipStream_t stream,
                                       hipEvent_t *ev);
