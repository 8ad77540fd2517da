// nsw6_adjoint.hip -- reverse mode of one FP64 NSW6 step on gfx950.
//
// Launch sequence (one stream): the forward step up to the wall fluxes
// (nsw6_launch_forward_to_flux: k_rates<double>, k_extremes, k_bounds, k_flux,
// k_flux_general), then the reverse passes -- k_adj_cell, k_adj_wall and
// k_adj_wall_general (the general regions' walls, one thread per column), k_adj_point,
// whose bodies are in nsw6_adjoint_sweep.h -- and k_adj_coef, which sums the
// per-block coefficient cotangents in a fixed order.  The passes communicate only
// through gather-only record arrays, so the result is deterministic and needs no
// atomics.  Block order is the XCD-aware level order of the forward kernels: a block
// re-reads the records of the neighbouring levels just written by the same XCD.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nsw6_adjoint_sweep.h"
#include "nsw6_tables.h"

using namespace nsw6;

extern "C" hipError_t nsw6_launch_forward_to_flux(const StepArgs *args, hipStream_t stream);

namespace {

struct AdjLaunch {
    AdjArgs A;
    double *coef_part;   // [blocks][NCOEF]
    double *grad_coef;   // [NCOEF]
};

template <int NT>
__global__ void __launch_bounds__(256) k_adj_cell(AdjLaunch L)
{
    const int ijdim = L.A.a.ijdim;
    const LevelBlock lb = level_block<ADJ_BAND>(blockIdx.x, (ijdim + 255) / 256, KALL, 0, L.A.a.lall);
    const int ij = lb.bx * 256 + threadIdx.x;
    if (!lb.valid || ij >= ijdim) return;
    adj_cell<NT>(L.A, lb.l, lb.k, ij);
}

// One wave per SIMD: the 512-register budget keeps the pass out of scratch (the
// overflow goes to 66 accumulation registers); at two waves 236 B per lane spilled
// to scratch and the reverse passes took about 4 % longer.
template <int NT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) k_adj_wall(AdjLaunch L)
{
    const int ijdim = L.A.a.ijdim;
    const LevelBlock lb = level_block<ADJ_BAND>(blockIdx.x, (ijdim + 255) / 256, KALL, 0, L.A.a.lall);
    const int ij = lb.bx * 256 + threadIdx.x;
    if (!lb.valid || ij >= ijdim) return;
    adj_wall<NT>(L.A, lb.l, lb.k, ij);
}

// The general regions' walls: one thread per column walks them (adj_walls_general);
// the single-cell regions' blocks return at once.
template <int NT>
__global__ void __launch_bounds__(64) k_adj_wall_general(AdjLaunch L)
{
    const int ij = blockIdx.x * 64 + threadIdx.x;
    if (ij >= L.A.a.ijdim) return;
    adj_walls_general<NT>(L.A, blockIdx.y, ij);
}

// LDS holds two things here (the pass runs one block per CU, so it has the room):
// * the parameter block: the pass is too large for the compiler to prove that no store
//   of the launch overwrites it in global memory, so its uniform reads would otherwise
//   take the vector memory path (about 150 of them per point), each a round trip that
//   one wave per SIMD cannot hide;
// * the rate reverse's stash (POINT_STASH doubles per thread, column threadIdx.x), which
//   takes the values that wait through the rate reverse and its phase-2 recompute out
//   of the registers.
template <int NT, bool INDIR, bool DEFAULT_SCHEME>
__global__ void __launch_bounds__(256) k_adj_point(AdjLaunch L)
{
    __shared__ double red[4][NCOEF];
    __shared__ double stash[POINT_STASH][256];
    __shared__ DevParams sP;
    {
        const int *src = reinterpret_cast<const int *>(L.A.a.P);
        int *dst = reinterpret_cast<int *>(&sP);
        for (int i = threadIdx.x; i < (int)(sizeof(DevParams) / sizeof(int)); i += 256) dst[i] = src[i];
        __syncthreads();
    }
    AdjArgs A = L.A;
    A.a.P = &sP;
    A.a1_tab = c_a1_tab;
    A.a2_tab = c_a2_tab;
    const int ijdim = A.a.ijdim;
    const LevelBlock lb = level_block<ADJ_BAND>(blockIdx.x, (ijdim + 255) / 256, KALL, 0, A.a.lall);
    const int ij = lb.bx * 256 + threadIdx.x;
    double tun_b[NCOEF];
#pragma unroll
    for (int c = 0; c < NCOEF; c++) tun_b[c] = 0.0;
    if (lb.valid && ij < ijdim)
        adj_point<NT, INDIR, DEFAULT_SCHEME>(A, lb.l, lb.k, ij, tun_b, RateStash{&stash[0][threadIdx.x], 256});
    // per-block sums of the coefficient cotangents (every thread reaches this point)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NCOEF; c++) {
        double v = tun_b[c];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[wv][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < NCOEF) {
        const int c = threadIdx.x;
        L.coef_part[(int64_t)blockIdx.x * NCOEF + c] = ((red[0][c] + red[1][c]) + red[2][c]) + red[3][c];
    }
}

// One block per coefficient: sum of the per-block partials in a fixed order.
__global__ void __launch_bounds__(256) k_adj_coef(AdjLaunch L, int nblocks)
{
    __shared__ double red[256];
    const int c = blockIdx.x;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 256) s += L.coef_part[(int64_t)b * NCOEF + c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) L.grad_coef[c] = red[0];
}

// Elementwise helpers of the sub-cycled reverse mode (MP_DIV_NUM > 1, nsw6_api.cpp).
__global__ void __launch_bounds__(256) k_scale(double *y, const double *x, double a, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}
__global__ void __launch_bounds__(256) k_accumulate(double *y, const double *x, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = y[i] + x[i];
}
__global__ void __launch_bounds__(64) k_merge_flags(int *acc, const int *flags, int n, int first)
{
    const int i = blockIdx.x * 64 + threadIdx.x;
    if (i < n) acc[i] = first ? flags[i] : (acc[i] | flags[i]);
}

}  // namespace

extern "C" hipError_t nsw6_launch_scale(double *y, const double *x, double a, int64_t n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, y, x, a, n);
    return hipGetLastError();
}
extern "C" hipError_t nsw6_launch_accumulate(double *y, const double *x, int64_t n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_accumulate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, y, x, n);
    return hipGetLastError();
}
extern "C" hipError_t nsw6_launch_merge_flags(int *acc, const int *flags, int n, int first, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_merge_flags, dim3((n + 63) / 64), dim3(64), 0, stream, acc, flags, n, first);
    return hipGetLastError();
}

// The reverse passes for NT sedimenting tracers.  The point pass of the default scheme
// has a kernel of its own (without the Berry, Bergeron and exponential-moment terms);
// "4WATER" takes the general one.
template <int NT> static void launch_reverse_passes(const AdjLaunch &L, int nb, hipStream_t stream)
{
    hipLaunchKernelGGL(k_adj_cell<NT>, dim3(nb), dim3(256), 0, stream, L);
    hipLaunchKernelGGL(k_adj_wall<NT>, dim3(nb), dim3(256), 0, stream, L);
    hipLaunchKernelGGL(k_adj_wall_general<NT>, dim3((L.A.a.ijdim + 63) / 64, L.A.a.lall), dim3(64), 0, stream, L);
    const bool dflt = (NT == 3) && L.A.a.default_scheme;
    if (!dflt && L.A.a.opt_indir)
        hipLaunchKernelGGL((k_adj_point<NT, true, false>), dim3(nb), dim3(256), 0, stream, L);
    else if (!dflt)
        hipLaunchKernelGGL((k_adj_point<NT, false, false>), dim3(nb), dim3(256), 0, stream, L);
    else if (L.A.a.opt_indir)
        hipLaunchKernelGGL((k_adj_point<3, true, true>), dim3(nb), dim3(256), 0, stream, L);
    else
        hipLaunchKernelGGL((k_adj_point<3, false, true>), dim3(nb), dim3(256), 0, stream, L);
}

// Number of blocks of the reverse passes (the coefficient partials hold NCOEF each).
extern "C" int nsw6_adjoint_blocks(int lall, int ijdim) { return level_grid((ijdim + 255) / 256, KALL, lall); }

// Events (NULL or 3): before the forward part, after it, after the reverse passes.
// forward = 0 skips the forward part (the workspace already holds it).
extern "C" hipError_t nsw6_launch_adjoint(const AdjArgs *adj, double *coef_part, double *grad_coef,
                                          hipStream_t stream, hipEvent_t *ev, int forward)
{
    AdjLaunch L;
    L.A = *adj;
    L.coef_part = coef_part;
    L.grad_coef = grad_coef;
    if (ev) (void)hipEventRecord(ev[0], stream);
    if (forward) {
        hipError_t e = nsw6_launch_forward_to_flux(&L.A.a, stream);
        if (e != hipSuccess) return e;
    }
    if (ev) (void)hipEventRecord(ev[1], stream);
    const int nb = nsw6_adjoint_blocks(L.A.a.lall, L.A.a.ijdim);
    if (L.A.a.four_water)
        launch_reverse_passes<4>(L, nb, stream);
    else
        launch_reverse_passes<3>(L, nb, stream);
    hipLaunchKernelGGL(k_adj_coef, dim3(NCOEF), dim3(256), 0, stream, L, nb);
    if (ev) (void)hipEventRecord(ev[2], stream);
    return hipGetLastError();
}
