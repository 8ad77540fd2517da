// nsw6_rates_mixed.hip -- the rate kernel of the mixed-precision form (k_rates<float>),
// in a unit of its own.  It is built without the approximate-function lowering of the
// FP64 unit, which would move its FP32 exp, log, sqrt and divisions to the native
// approximations; the mixed form's limiters amplify those differences over a rollout.
#include <hip/hip_runtime.h>

// the FP64 parts' exp and log from the LDS tables of nsw6_math.h
#define NSW6_TABLE_MATH 1
#include "nsw6_internal.h"
#include "nsw6_rates_kernel.h"

using namespace nsw6;

extern "C" hipError_t nsw6_launch_rates_mixed(const StepArgs *args, hipStream_t stream)
{
    const StepArgs &a = *args;
    const dim3 grid_p((a.ijdim + 255) / 256, KALL, a.lall);
    if (a.opt_indir)
        hipLaunchKernelGGL((k_rates<float, true, true>), grid_p, dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL((k_rates<float, true>), grid_p, dim3(256), 0, stream, a);
    return hipGetLastError();
}
