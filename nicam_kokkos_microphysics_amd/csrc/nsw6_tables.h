// nsw6_tables.h -- the Bergeron tables a1(T), a2(T) in constant memory, for the
// device units that evaluate the Bergeron growth (the rate kernel and the reverse
// point pass).  Each unit that includes this header holds its own copy.
#pragma once

#include <hip/hip_runtime.h>

namespace nsw6 {

// Bergeron tables a1, a2 (src/mod_mp_nsw6.cpp:1846-1860).
static __constant__ double c_a1_tab[32] = {
    0.0001E-7, 0.7939E-7, 0.7841E-6, 0.3369E-5, 0.4336E-5, 0.5285E-5, 0.3728E-5, 0.1852E-5,
    0.2991E-6, 0.4248E-6, 0.7434E-6, 0.1812E-5, 0.4394E-5, 0.9145E-5, 0.1725E-4, 0.3348E-4,
    0.1725E-4, 0.9175E-5, 0.4412E-5, 0.2252E-5, 0.9115E-6, 0.4876E-6, 0.3473E-6, 0.4758E-6,
    0.6306E-6, 0.8573E-6, 0.7868E-6, 0.7192E-6, 0.6513E-6, 0.5956E-6, 0.5333E-6, 0.4834E-6};
static __constant__ double c_a2_tab[32] = {
    0.0100, 0.4006, 0.4831, 0.5320, 0.5307, 0.5319, 0.5249, 0.4888, 0.3849, 0.4047, 0.4318,
    0.4771, 0.5183, 0.5463, 0.5651, 0.5813, 0.5655, 0.5478, 0.5203, 0.4906, 0.4447, 0.4126,
    0.3960, 0.4149, 0.4320, 0.4506, 0.4483, 0.4460, 0.4433, 0.4413, 0.4382, 0.4361};

}  // namespace nsw6
