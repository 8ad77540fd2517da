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

// Per-region reductions: [tracer NT][level type 2: full, half][k 96][max, min]
// (NT = 3 sedimenting tracers under "3WATER", 4 under "4WATER")
template <int NT> constexpr int ext_per_region() { return NT * 2 * KALL * 2; }
constexpr int EXT_PER_REGION = ext_per_region<3>();
constexpr int EXT_PER_REGION_MAX = ext_per_region<4>();
// k_extremes: one block of EXT_THREADS threads reduces EXT_COLS_PER_BLOCK columns of
// one level; per-block partial extremes are stored as [l][k][4 NT][block], slot
// t*4 + {0 full max, 1 full min, 2 half max, 3 half min}
constexpr int EXT_THREADS = 256;
constexpr int EXT_COLS_PER_THREAD = 8;   // 0.204 ms per step at 169,000 x 96; 16: 0.213, 4: 0.214, 32: 0.244
constexpr int EXT_COLS_PER_BLOCK = EXT_THREADS * EXT_COLS_PER_THREAD;
constexpr int BOUNDS_PER_REGION = EXT_PER_REGION;
// Flux workspace [l][FLUX_SLOTS][k][ij].  For the single-cell regions, wall w holds:
//   0..NT-1      mass fluxes of the sedimenting tracers (QR, [QI,] QS, QG)
//   NT..NT+4     tracer-summed fluxes of internal energy, potential energy, horizontal
//                kinetic energy, and the two horizontal momentum components
//   NT+5..NT+6   tracer-summed half-level fluxes of vertical kinetic energy and momentum
//   NT+7         the pre-transport rhogw of cell w
// (the same slots for the walls of the general regions, k_flux_general).
template <int NT> struct FxSlots {
    static constexpr int Q = 0, E = NT, PHI = NT + 1, KINH = NT + 2, U = NT + 3, V = NT + 4, KINV = NT + 5,
                         W = NT + 6, RW = NT + 7, N = NT + 8;
};
constexpr int FLUX_SLOTS = FxSlots<3>::N;
constexpr int FLUX_SLOTS_MAX = FxSlots<4>::N;
constexpr int FX_Q = FxSlots<3>::Q, FX_E = FxSlots<3>::E, FX_PHI = FxSlots<3>::PHI, FX_KINH = FxSlots<3>::KINH,
              FX_U = FxSlots<3>::U, FX_V = FxSlots<3>::V, FX_KINV = FxSlots<3>::KINV, FX_W = FxSlots<3>::W,
              FX_RW = FxSlots<3>::RW;
// Sedimenting tracer t of NT in nq order (the order of the reference's tracer loop,
// src/mod_precip_transport.cpp:789): 3WATER QR, QS, QG; 4WATER QR, QI, QS, QG.
template <int NT> __host__ __device__ constexpr int sed_nq(int t)
{
    return NT == 4 ? QR + t : (t == 0 ? QR : (t == 1 ? QS : QG));
}
__host__ __device__ constexpr int ext_index(int t, int type, int k, int mm)
{
    return ((t * 2 + type) * KALL + k) * 2 + mm;
}
__host__ __device__ constexpr int bounds_index(int t, int type, int k, int mm) { return ext_index(t, type, k, mm); }
// report ints: [0] any failure, [1] not converged, [2] NaN, [3..5] unused, [6..7] the
// first failing point as the 64-bit key (l * KALL + k) * ijdim + ij (atomicMin, so the
// lexicographically first (l, k, ij) whatever the thread order; all ones = none)
constexpr int REPORT_INTS = 8;
constexpr int REPORT_KEY = 6;   // int offset of the key (8-byte aligned)

// XCD-aware block order for the stencil kernels.  MI355X deals consecutive
// workgroups round-robin to its 8 XCDs, each with its own L2.  A 1-D grid of
// NUM_XCD * per workgroups is mapped so that workgroup b takes logical block
// (b % NUM_XCD) * per + b / NUM_XCD: every XCD walks one contiguous range of logical
// blocks.  Within a region the logical blocks go band by band, a band being BAND
// neighbouring 256-column blocks; a band is walked level by level, its BAND blocks
// of one level one after another.  The neighbouring levels a block re-reads were
// then just loaded into the same L2, and the blocks an XCD holds at once touch each
// field in runs of BAND x 2 KB instead of 2 KB: at 169,000 x 96 the finish kernel
// takes 1.49 ms with bands of 8 blocks against 1.53 ms with bands of one (a column
// strip walked level by level), and the wall-flux kernel 1.04 against 1.05 ms with
// bands of 16; bands of 32 are slower again (the L2 no longer holds the levels that
// are re-read).
constexpr int NUM_XCD = 8;
struct LevelBlock {
    int l, bx, k;
    bool valid;
};
__host__ __device__ inline int level_grid(int nbx, int nk, int lall)
{
    const int total = nbx * nk * lall;
    return ((total + NUM_XCD - 1) / NUM_XCD) * NUM_XCD;
}
template <int BAND = 1>
__device__ __forceinline__ LevelBlock level_block(int b, int nbx, int nk, int k0, int lall)
{
    const int total = nbx * nk * lall;
    const int per = (total + NUM_XCD - 1) / NUM_XCD;
    const int L = (b % NUM_XCD) * per + b / NUM_XCD;
    LevelBlock r;
    r.valid = L < total;
    const int per_l = nbx * nk;
    r.l = L / per_l;
    const int rr = L - r.l * per_l;
    const int band = rr / (BAND * nk);
    const int rb = rr - band * BAND * nk;
    const int width = (nbx - band * BAND < BAND) ? nbx - band * BAND : BAND;   // the last band may be narrower
    r.k = k0 + rb / width;
    r.bx = band * BAND + rb % width;
    return r;
}
constexpr int FLUX_BAND = 16, FINISH_BAND = 8;
// the reverse passes (k_adj_cell, k_adj_wall, k_adj_point): bands of 4 take them from
// 1.40 to 1.35 ms at 16,900 x 96 (1.38 ms with bands of 16)
constexpr int ADJ_BAND = 4;

// Context-owned device workspace (allocated once for the maximum column count).
struct Workspace {
    double *rhog, *rhoge, *rhogq;   // state after the rate update (pre-sedimentation)
    double *vt;                     // fall speeds of the NT sedimenting tracers [l][NT][k][ij]
    double *flux;                   // [l][FLUX_SLOTS][k][ij] (see FLUX_SLOTS)
    double *part;                   // [l][KALL][12][blocks per region] partial extremes
    int *bounds;                    // [l][BOUNDS_PER_REGION]
    int *region_flag;               // [l] 1 = region needs the general wall fluxes
    int *report;                    // [REPORT_INTS]
    int *cloud_top;                 // [l][ij] highest cloudy level of the step, -1 = none
};

struct StepArgs {
    Fields f;
    Workspace ws;
    const DevParams *P;   // device copy owned by the context (kept out of the kernel
                          // arguments: ~130 doubles there overflow the scalar registers)
    const GridDev *grid;
    int lall, ijdim;
    int keep_report;              // 1: keep the report of earlier sub-steps of one call
    int cloudtop_intended;        // 1: cloud top = highest cloudy level (k_cloudtop)
    int opt_indir;                // 1: aerosol indirect effect (k_rates<R, DIAG, true>)
    int kin2kin;                  // 1: KIN2KIN energy correction of the sedimentation
    int four_water;               // 1: "4WATER", QI sediments (transport kernels with NT = 4)
    int default_scheme;           // 1: KK2000 alone (sw_kk2000 = 1), explicit ice generation
                                  // (sw_expice = 1) and Roh & Satoh snow (sw_roh2014 = 1): the
                                  // reverse point pass's kernel for them
    double dt, dt2, dt3, dt3_6;   // dt, dt^2, dt^3, dt^3/6
};

// Reverse sweep (nsw6_adjoint_sweep.h): record arrays [slot][l][k][ij] written by one
// pass and gathered by the next, for NT sedimenting tracers.
// cell record slots: the cotangents of fq of walls c-1..c+2 (4 NT), of fs of walls c, c+1
// (10), of fh of walls c-1..c+1 (6), of rhog of cells c-1..c+1 (3) and of the
// pre-transport rhogw of cells c, c+1 (2)
template <int NT> struct CellRec {
    static constexpr int FQ = 0, FS = 4 * NT, FH = FS + 10, RG = FH + 6, RW = RG + 3, N = RW + 2;
};
// wall record slots: cell w-1 (rhog and the NT tracers); cells w and w+1 in CellInB
// order (rhog, rq[5], rhoge, rvx, rvy, rvz, rw, vx, vy, vz); Vt(w-3..w+1) of the NT tracers
template <int NT> struct WallRec {
    static constexpr int M = 0, C0 = 1 + NT, C1 = C0 + 14, V = C1 + 14, N = V + 5 * NT;
};
constexpr int CR_FQ = CellRec<3>::FQ, CR_FS = CellRec<3>::FS, CR_FH = CellRec<3>::FH, CR_RG = CellRec<3>::RG,
              CR_RW = CellRec<3>::RW, NCR = CellRec<3>::N;
constexpr int WR_M = WallRec<3>::M, WR_C0 = WallRec<3>::C0, WR_C1 = WallRec<3>::C1, WR_V = WallRec<3>::V,
              NWR = WallRec<3>::N;
// own-cell cotangents: pre-sedimentation state (rhog, rhoge, rhogq[6]) and the input
// momentum (rhogvx, rhogvy, rhogvz, rhogw)
constexpr int NDW = 8, NDI = 4;
// 92 slots for NT = 3, 102 for NT = 4
template <int NT> constexpr int adj_record_slots() { return CellRec<NT>::N + WallRec<NT>::N + NDW + NDI; }
static_assert(adj_record_slots<3>() == 92 && adj_record_slots<4>() == 102, "record layout");

struct AdjArgs {
    StepArgs a;                 // forward step state (workspace filled up to the fluxes)
    Fields cot;                 // output cotangents (NULL pointer = zero)
    const double *cot_sflux;    // [l][3][ij] or NULL
    Fields grad;                // input cotangents written (NULL pointer = not wanted)
    double *cr, *wr, *dws, *din;   // records [slot][l][k][ij]
    const double *a1_tab, *a2_tab;  // Bergeron tables a1(T), a2(T) (32 entries each)
};

}  // namespace nsw6

// events recorded by nsw6_launch_step: before k_rates, after k_rates, after
// k_extremes+k_bounds, after k_flux, after k_flux_general, after k_finish
constexpr int NSW6_STEP_EVENTS = 6;

extern "C" hipError_t nsw6_launch_step(const nsw6::StepArgs *args, int precision, hipStream_t stream,
                                       hipEvent_t *ev);
// k_rates<float> of the mixed-precision form (its own unit, nsw6_rates_mixed.hip)
extern "C" hipError_t nsw6_launch_rates_mixed(const nsw6::StepArgs *args, hipStream_t stream);
// k_finish (nsw6_finish.hip) for nt = 3 or 4 sedimenting tracers, fall speeds and wall
// fluxes stored as float (ws_float = 1, the mixed form) or double
extern "C" hipError_t nsw6_launch_finish(const nsw6::StepArgs *args, int nt, int ws_float, hipStream_t stream);
// MP_DIV_NUM sub-cycling: sums [l][KALL][ij] of GPREC followed by [l][DIV_COLUMN_SUMS][ij]
// of precip (2 rows), precip_rhoe, precip_lh_heat, precip_rhophi, precip_rhokin.
constexpr int DIV_COLUMN_SUMS = 6;
extern "C" hipError_t nsw6_launch_div_sum(const nsw6::StepArgs *args, double *sum, int first,
                                          hipStream_t stream);
extern "C" hipError_t nsw6_launch_div_scale(const nsw6::StepArgs *args, const double *sum, double fraction,
                                            hipStream_t stream);

// events recorded by nsw6_launch_adjoint: before the forward part, after it, after the
// reverse passes.  forward = 0: adj->a.ws already holds the forward part (the rates'
// state and fall speeds, kcell bounds and region flags, wall fluxes) of this input
// state, so it is not recomputed.
constexpr int NSW6_ADJ_EVENTS = 3;
extern "C" int nsw6_adjoint_blocks(int lall, int ijdim);
extern "C" hipError_t nsw6_launch_adjoint(const nsw6::AdjArgs *adj, double *coef_part, double *grad_coef,
                                          hipStream_t stream, hipEvent_t *ev, int forward);
// Sub-cycled reverse mode: y = a x, y += x (n doubles), acc = flags (first) or acc | flags.
extern "C" hipError_t nsw6_launch_scale(double *y, const double *x, double a, int64_t n, hipStream_t stream);
extern "C" hipError_t nsw6_launch_accumulate(double *y, const double *x, int64_t n, hipStream_t stream);
extern "C" hipError_t nsw6_launch_merge_flags(int *acc, const int *flags, int n, int first, hipStream_t stream);
