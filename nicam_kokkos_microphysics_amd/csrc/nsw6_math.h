// nsw6_math.h -- exp and log of the FP64 device code, table-driven.
//
// The device library's double exp and log spend most of their instructions on
// extended-precision intermediate steps (log is evaluated in double-double) and on
// wide polynomials: about 45 and 90 VALU instructions per call.  The rate kernel makes
// some 25 such calls per point and is bound by VALU issue, so these two functions
// decide much of its time.  Here both follow P. T. P. Tang's table-driven methods
// ("Table-driven implementation of the exponential function in IEEE floating-point
// arithmetic", ACM TOMS 15(2), 1989; "... of the logarithm function ...", ACM TOMS
// 16(4), 1990): a reduction to a small interval around a table point, then a short
// Taylor polynomial.
//
//   exp: x = (64 m + j) ln2/64 + r with |r| <= ln2/128, exp(x) = 2^m 2^(j/64) exp(r),
//        exp(r) - 1 = r + r^2 (1/2 + r (1/6 + r (1/24 + r/120))), truncation < r^6/720
//        (3.5e-17).  About 1 ulp.
//   log: x = f 2^e with f in [1/2, 1), F = 1/2 + j/128 the table point nearest to f,
//        log(x) = e ln2 + log(F) + log(1 + r), r = (f - F)/F in [-1/128, 1/128],
//        log(1 + r) = r + r^2 (-1/2 + r (1/3 - ... - r/8)), truncation < r^9/9.
//        e ln2 + log(F) is formed from head and tail parts, so the result is within
//        about 1 ulp and log(1 + d) for small d keeps its relative accuracy.
//
// Special values follow the C library: exp(-inf) = 0, exp(+inf) = +inf, overflow to
// +inf, underflow through the subnormals to 0, exp(NaN) = NaN; log(+-0) = -inf,
// log(+inf) = +inf, log(x < 0) = NaN, log(NaN) = NaN.  Every table index stays in range
// for every input, NaN included.
//
// The tables hold correctly rounded values: 2^(j/64) (j = 0..63) and the head and tail
// of log(1/2 + j/128) (j = 0..64); the test of this header recomputes them to 80
// decimal digits.  Each unit that includes this header holds its own copy in constant
// memory.  Host code keeps the C library (these functions serve the device overloads
// and the host accuracy test only).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace nsw6 {

#define NSW6_EXP2_TAB                                                                                     \
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,               \
        0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,           \
        0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,           \
        0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,           \
        0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,           \
        0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,           \
        0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,           \
        0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,           \
        0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,           \
        0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,           \
        0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,           \
        0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,           \
        0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,           \
        0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,           \
        0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,           \
        0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0

// log(1/2 + j/128): head, tail (j = 0..64)
#define NSW6_LOG_TAB                                                                                      \
    -0x1.62e42fefa39efp-1, -0x1.abc9e3b39803fp-56, -0x1.5af405c3649e0p-1, 0x1.6714fbcd8135bp-55,           \
        -0x1.5322e26867857p-1, -0x1.988ba4aea614dp-56, -0x1.4b6fd6f970c1fp-1, -0x1.c457b531506f6p-55,     \
        -0x1.43d9ff2f923c5p-1, 0x1.84f481051f71ap-56, -0x1.3c6080c36bfb5p-1, -0x1.1930603d87b6ep-56,      \
        -0x1.35028ad9d8c86p-1, 0x1.f01ab6065515cp-56, -0x1.2dbf557b0df43p-1, 0x1.b941ee770436bp-56,       \
        -0x1.269621134db92p-1, -0x1.e0efadd9db02bp-55, -0x1.1f8635fc61659p-1, 0x1.2164ff40e9817p-56,      \
        -0x1.188ee40f23ca6p-1, -0x1.89df1568ca0b0p-55, -0x1.11af823c75aa8p-1, 0x1.91eee7772c7c2p-55,      \
        -0x1.0ae76e2d054fap-1, -0x1.0d710fcfc4e0dp-55, -0x1.04360be7603adp-1, 0x1.17f9e54e78104p-57,      \
        -0x1.fb358af7a4884p-2, -0x1.7e8f05924d259p-57, -0x1.ee2a156b413e5p-2, 0x1.74b71fb5e57e3p-62,      \
        -0x1.e148a1a2726cep-2, 0x1.ac81cc8a4dfb8p-56, -0x1.d490246defa6bp-2, -0x1.d7f4d3b3d406bp-56,      \
        -0x1.c7ff9c74554c9p-2, -0x1.223eadb651b4ap-57, -0x1.bb9611b80e2fbp-2, -0x1.6fd02999b21e1p-59,     \
        -0x1.af5295248cdd0p-2, -0x1.9d56c45dd3e86p-56, -0x1.a33440224fa79p-2, 0x1.ba8062860ae23p-57,      \
        -0x1.973a3431356aep-2, 0x1.89d2816cf838fp-57, -0x1.8b639a88b2df5p-2, 0x1.70f2f38238303p-56,       \
        -0x1.7fafa3bd8151cp-2, 0x1.219024acd3b77p-58, -0x1.741d876c67bb1p-2, -0x1.84a4ee3059583p-56,      \
        -0x1.68ac83e9c6a14p-2, -0x1.a64eadd740178p-58, -0x1.5d5bddf595f30p-2, 0x1.6541148cbb8a2p-56,      \
        -0x1.522ae0738a3d8p-2, 0x1.8f7e9b38a6979p-57, -0x1.4718dc271c41bp-2, -0x1.8fb4c14c56eefp-60,      \
        -0x1.3c25277333184p-2, 0x1.2ad27e50a8ec6p-56, -0x1.314f1e1d35ce4p-2, 0x1.3d69909e5c3dcp-56,       \
        -0x1.269621134db92p-2, -0x1.e0efadd9db02bp-56, -0x1.1bf99635a6b95p-2, 0x1.12aeb84249223p-57,      \
        -0x1.1178e8227e47cp-2, 0x1.0e63a5f01c691p-57, -0x1.07138604d5862p-2, -0x1.cdb16ed4e9138p-56,      \
        -0x1.f991c6cb3b379p-3, -0x1.f665066f980a2p-57, -0x1.e530effe71012p-3, -0x1.2276041f43042p-59,     \
        -0x1.d1037f2655e7bp-3, -0x1.60629242471a2p-57, -0x1.bd087383bd8adp-3, -0x1.dd355f6a516d7p-60,     \
        -0x1.a93ed3c8ad9e3p-3, -0x1.bcafa9de97203p-57, -0x1.95a5adcf7017fp-3, -0x1.142c507fb7a3dp-58,     \
        -0x1.823c16551a3c2p-3, 0x1.1232ce70be781p-57, -0x1.6f0128b756abcp-3, 0x1.8de59c21e166cp-57,       \
        -0x1.5bf406b543db2p-3, 0x1.1f5b44c0df7e7p-61, -0x1.4913d8333b561p-3, 0x1.0d5604930f135p-58,       \
        -0x1.365fcb0159016p-3, -0x1.7d411a5b944adp-58, -0x1.23d712a49c202p-3, 0x1.6e38161051d69p-57,      \
        -0x1.1178e8227e47cp-3, 0x1.0e63a5f01c691p-58, -0x1.fe89139dbd566p-4, 0x1.ac9f4215f9393p-58,       \
        -0x1.da727638446a2p-4, -0x1.401fa71733019p-58, -0x1.b6ac88dad5b1cp-4, 0x1.0057eed1ca59fp-59,      \
        -0x1.9335e5d594989p-4, 0x1.478a85704ccb7p-58, -0x1.700d30aeac0e1p-4, 0x1.72566212cdd05p-61,       \
        -0x1.4d3115d207eacp-4, -0x1.769f42c7842ccp-58, -0x1.2aa04a44717a5p-4, 0x1.d15d38d2fa3f7p-58,      \
        -0x1.08598b59e3a07p-4, 0x1.dd7009902bf32p-58, -0x1.ccb73cdddb2ccp-5, 0x1.e48fb0500efd4p-59,       \
        -0x1.894aa149fb343p-5, -0x1.a8be97660a23dp-60, -0x1.466aed42de3eap-5, 0x1.cdd6f7f4a137ep-59,      \
        -0x1.0415d89e74444p-5, -0x1.c05cf1d753622p-59, -0x1.8492528c8cabfp-6, 0x1.d192d0619fa67p-60,      \
        -0x1.0205658935847p-6, -0x1.27c8e8416e71fp-60, -0x1.010157588de71p-7, -0x1.46662d417ced0p-62,     \
        0x0.0p+0, 0x0.0p+0

static __constant__ double c_exp2_tab[64] = {NSW6_EXP2_TAB};
static __constant__ double c_log_tab[130] = {NSW6_LOG_TAB};

// The kernels of the units that define NSW6_TABLE_MATH read the tables from LDS
// copies: a lookup per lane from constant memory goes through the vector memory path,
// whose latency the VALU-bound kernels cannot hide.  Each kernel of such a unit that
// evaluates exp or log calls load_math_tables() with every thread of its
// (one-dimensional) block before its first exp or log and before any thread leaves.
#ifdef NSW6_TABLE_MATH
static __shared__ double s_exp2_tab[64];
static __shared__ double s_log_tab[130];
__device__ __forceinline__ void load_math_tables()
{
    for (int i = threadIdx.x; i < 64 + 130; i += blockDim.x) {
        if (i < 64)
            s_exp2_tab[i] = c_exp2_tab[i];
        else
            s_log_tab[i - 64] = c_log_tab[i - 64];
    }
    __syncthreads();
}
#else
__device__ __forceinline__ void load_math_tables() {}
#endif

constexpr double C_LN2_HI = 0x1.62e42fefa39efp-1, C_LN2_LO = 0x1.abc9e3b39803fp-56;
constexpr double C_LN2_64_HI = 0x1.62e42fefa39efp-7, C_LN2_64_LO = 0x1.abc9e3b39803fp-62;
constexpr double C_64_LN2 = 0x1.71547652b82fep+6;

// 1/F to full precision: the hardware reciprocal refined by one Newton step on the
// device, a division on the host.
__host__ __device__ __forceinline__ double recip_table_point(double F)
{
#ifdef __HIP_DEVICE_COMPILE__
    const double r0 = __builtin_amdgcn_rcp(F);
    return fma(fma(-F, r0, 1.0), r0, r0);
#else
    return 1.0 / F;
#endif
}

// exp(x) with the table T = 2^(j/64), j = 0..63.
__host__ __device__ __forceinline__ double exp_tang(double x, const double *T)
{
    // clamped so that the reduction stays finite; +-1100 still over/underflow
    const double xc = fmin(fmax(x, -1100.0), 1100.0);
    const double N = rint(xc * C_64_LN2);
    double r = fma(-N, C_LN2_64_HI, xc);
    r = fma(-N, C_LN2_64_LO, r);
    const int n = (int)N;
    const double t = T[n & 63];
    const double q = fma(fma(fma(r, 1.0 / 120.0, 1.0 / 24.0), r, 1.0 / 6.0), r, 0.5);
    const double p = fma(r * r, q, r);
    const double y = ldexp(fma(t, p, t), n >> 6);
    return (x == x) ? y : x;
}

// log(x) with the table L = {head, tail of log(1/2 + j/128)}, j = 0..64.
__host__ __device__ __forceinline__ double log_tang(double x, const double *L)
{
    const bool finite_pos = (x > 0.0) && (x < HUGE_VAL);
    int e = 0;
    const double fr = frexp(x, &e);
    const double f = finite_pos ? fr : 0.75;          // in [1/2, 1) for every input
    const int j = (int)rint((f - 0.5) * 128.0);       // 0..64, both steps exact
    const double F = fma((double)j, 0.0078125, 0.5);
    const double r = (f - F) * recip_table_point(F);  // f - F exact
    double q = fma(r, -0.125, 1.0 / 7.0);
    q = fma(q, r, -1.0 / 6.0);
    q = fma(q, r, 0.2);
    q = fma(q, r, -0.25);
    q = fma(q, r, 1.0 / 3.0);
    q = fma(q, r, -0.5);
    const double lp = fma(r * r, q, r);
    const double ed = finite_pos ? (double)e : 0.0;
    const double y = fma(ed, C_LN2_HI, L[2 * j]) + (fma(ed, C_LN2_LO, L[2 * j + 1]) + lp);
    const double special = (x == 0.0) ? -HUGE_VAL : ((x > 0.0) ? x : NAN);
    return finite_pos ? y : special;
}

}  // namespace nsw6
