// nsw6_device.h -- device-side physics of the NSW6 step for gfx950.
//
// Point-wise pieces of the reference step (paths relative to the reference repo):
//   negative filter          src/mod_mp_nsw6.cpp:1750-1832
//   psat_liq / psat_ice      src/mod_satadjust.cpp:80-106, 133-160
//   Bergeron_param           src/mod_mp_nsw6.cpp:1834-1890
//   TagBigLoop (49 rates)    src/mod_mp_nsw6.cpp:760-1594
//   THRMDYN_qd / _cv         src/mod_thrmdyn.cpp:98-135
//   satadjust_liq Newton     src/mod_satadjust.cpp:828-894
//
// The rate code is templated on the arithmetic type R: double for the FP64 form
// and float for the mixed form.  Operation order follows the reference so that the
// FP64 form reproduces it to round-off; the deliberate differences are
//   * integer powers pow(x,n), n = 2..7, are products (<= 2 ulp),
//   * log(Xs2+zerosw) of the seven Roh & Satoh moments is evaluated once,
//   * logs/powers of constants (log(mi40), log(mi50), Nc^-1.79, log(Nc/2000)) are
//     computed once on the host,
//   * the dead first assignment of Praut_kk (overwritten at :1154) is dropped,
//   * exp(ln10*a)*exp(lx*b) of the Roh moments is exp(ln10*a + lx*b) and pow(x,y)
//     with real y is exp(y*log(x)) (a few ulp),
//   * terms multiplied by a switch that is exactly zero for the whole step are not
//     evaluated when their value is finite by construction: the Bergeron growth
//     (x sw_bergeron = 1 - sw_expice), Berry autoconversion and the original
//     Pracw (x (1 - sw_kk2000)), the non-Roh snow moments (x (1 - sw_roh2014)),
//     and the QI fall speed (QI does not sediment under "3WATER").
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

#include "nsw6_math.h"

// The point-wise physics is compiled for the device and for the host: the host
// instantiations serve the dual-number gradient checks of the adjoint.
#define NSW6_HD __host__ __device__ __forceinline__

namespace nsw6 {

// float arguments take the float overloads on the host as on the device
using std::copysign;
using std::expm1;
using std::fabs;
using std::pow;
using std::sqrt;
// double exp and log on the device: the table-driven forms of nsw6_math.h (about 20
// and 35 instead of 45 and 90 VALU instructions, within about 1 ulp) in the units that
// define NSW6_TABLE_MATH, whose kernels then load the tables into LDS
// (load_math_tables); the device library's elsewhere.  The host keeps the C library's,
// and float the library's on both sides.
#ifdef NSW6_TABLE_MATH
static __device__ __forceinline__ double exp(double x) { return exp_tang(x, s_exp2_tab); }
static __device__ __forceinline__ double log(double x) { return log_tang(x, s_log_tab); }
#else
static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
static __device__ __forceinline__ double log(double x) { return ::log(x); }
#endif
static __host__ inline double exp(double x) { return std::exp(x); }
static __host__ inline double log(double x) { return std::log(x); }
__host__ __device__ __forceinline__ float exp(float x) { return std::exp(x); }
__host__ __device__ __forceinline__ float log(float x) { return std::log(x); }

constexpr int KALL = 96, KMIN = 1, KMAX = 94, NQ = 6;
constexpr int NCOEF = 8;   // tunable coefficients (order of include/nsw6.h)
constexpr int QV = 0, QC = 1, QR = 2, QI = 3, QS = 4, QG = 5;

constexpr double C_PI = 3.141592653589793;
constexpr double C_EPS = 2.220446E-16;
constexpr double C_UNDEF = -9.9999E+30;
constexpr double C_GRAV = 9.79764;
constexpr double C_Rdry = 287.04;
constexpr double C_CVdry = 717.56;
constexpr double C_Rvap = 461.46;
constexpr double C_CL = 4218.0;
constexpr double C_PSAT0 = 610.7;
constexpr double C_LHV0 = 2501000.0;
constexpr double C_LHS0 = 2834000.0;
constexpr double C_LHF0 = 333000.0;
constexpr double C_Pstd = 101325.0;
constexpr double C_TEM00 = 273.15;
constexpr double C_DWATR = 1000.0;
constexpr double C_CVW = 717.56;      // CVW(:) = CVdry under EIN_TYPE = SIMPLE
constexpr double TEM_MIN = 10.0;
constexpr double DENS00 = 1.28;
constexpr double Ka0 = 2.428E-2, dKa_dT = 7.47E-5;
constexpr double Kd0 = 2.222E-5, dKd_dT = 1.37E-7;
constexpr double nu0 = 1.718E-5, dnu_dT = 5.28E-8;

// Scheme coefficients as the kernels see them (nsw6_params + host-side constants).
struct DevParams {
    double N0r, N0s, N0g, Ar, As, Ag, Cr, Cs, Cg;
    double GAM, GAM_2, GAM_3;
    double GAM_1br, GAM_2br, GAM_3br, GAM_3dr, GAM_6dr, GAM_1brdr, GAM_5dr_h;
    double GAM_1bs, GAM_2bs, GAM_3bs, GAM_3ds, GAM_1bsds, GAM_5ds_h;
    double GAM_1bg, GAM_3dg, GAM_1bgdg, GAM_5dg_h;
    double sw_kk2000, sw_roh2014, ln10;
    double coef_a[10], coef_b[10];
    double Eiw, Erw, Esw, Egw, Eri, Esi, Egi, Esr, Egr, Egs;
    double gamma_sacr, gamma_gacs, mi;
    double Nc_def, beta_saut, gamma_saut, beta_gaut, gamma_gaut, qicrt_saut, qscrt_gaut;
    double f1r, f2r, f1s, f2s, f1g, f2g, A_frz, B_frz;
    double mi40, mi50, vti50, Ri50;
    double sw_expice, Nc_ihtr, Di_max, Di_a, sw_constVti, CONST_Vti;
    double kk_aut_rate, icenuc_coef, coef_dgam, coef_xf;
    // host-computed constants
    double ln_mi40, ln_mi50;   // log(mi40), log(mi50)
    double Nc_pow_m179;        // pow(Nc_def, -1.79)
    double Dc;                 // 0.146 - 5.964e-2 * log(Nc_def / 2000)
    double LovR_liq, LovR_ice, RTEM00;
    double dtemp_criteria;     // the Newton's stopping test; the reference's 10^(-(15+1)/2) = 1e-8
    // reciprocals and ratios of parameters, formed once on the host so that the
    // point-wise code multiplies instead of dividing
    double rK_r, rK_s, rK_g;    // 1/(Ar N0r GAM_1br), 1/(As N0s GAM_1bs), 1/(Ag N0g GAM_1bg)
    double rAs, rmi, rNc_ihtr;  // 1/As, 1/mi, 1/(Nc_ihtr 1e6)
    double Vtr_fac, Vtg_fac;    // GAM_1brdr/GAM_1br, GAM_1bgdg/GAM_1bg
    // SATURATION_adjustment path (src/mod_mp_nsw6.cpp:647-660): satadjust_all with
    // SATURATION_Setrange(273.16, 233.16) when explicit ice generation is off
    // (sw_expice != 1), else satadjust_liq (range 100, 90)
    double ice_adjust, sat_ulimit, sat_llimit;
    double cloudtop_intended;   // 1: highest cloudy level (see nsw6_params)
    double opt_indir;           // 1: Nc = max(unccn * 1e-6, Nc_def) per point
    double kin2kin;             // 1: PRCIP_TRN_ECORRECT = "KIN2KIN"
    double four_water;          // 1: precip_transport_type = "4WATER" (QI sediments)
};


NSW6_HD double dmax(double a, double b) { return (a < b) ? b : a; }
NSW6_HD double dmin(double a, double b) { return (b < a) ? b : a; }
NSW6_HD float dmax(float a, float b) { return (a < b) ? b : a; }
NSW6_HD float dmin(float a, float b) { return (b < a) ? b : a; }
// the value of a scalar (the identity for the floating-point types; a dual-number
// scalar overloads it)
NSW6_HD double val(double x) { return x; }
NSW6_HD float val(float x) { return x; }

// Cloud droplet number of a point and the two functions of it the rates use:
// Nc_def everywhere (the constants formed on the host), or under the aerosol
// indirect effect max(unccn * 1e-6, Nc_def) (src/mod_mp_nsw6.cpp:676-679) with
// Nc^-1.79 (:1154) and Dc (:1146) evaluated at the point.
struct NcPoint {
    double Nc, pow_m179, Dc;
};
NSW6_HD NcPoint nc_default(const DevParams &P)
{
    NcPoint n;
    n.Nc = P.Nc_def;
    n.pow_m179 = P.Nc_pow_m179;
    n.Dc = P.Dc;
    return n;
}
// Where the maximum is Nc_def the host-formed constants are taken, so that such points
// step (and reverse) exactly as without OPT_INDIR.
NSW6_HD NcPoint nc_indirect(const DevParams &P, double unccn)
{
    const double nc = unccn * 1.0E-6;
    if (nc < P.Nc_def) return nc_default(P);
    NcPoint n;
    n.Nc = nc;
    n.pow_m179 = pow(n.Nc, -1.79);
    n.Dc = 0.146 - 5.964E-2 * log(n.Nc / 2000.0);
    return n;
}

// 0.5 + copysign(0.5, x): 1 for x >= +0, 0 for x <= -0 (IEEE sign, as the reference).
template <typename R> NSW6_HD R step_pos(R x) { return R(0.5) + copysign(R(0.5), x); }
template <typename R> NSW6_HD R step_neg(R x) { return R(0.5) - copysign(R(0.5), x); }

NSW6_HD double psat_liq(const DevParams &P, double tem)
{
    double rtem = 1.0 / dmax(tem, TEM_MIN);
    // pow(tem*RTEM00, CPovR_liq) with CPovR_liq = 0 under SIMPLE is exactly 1
    return C_PSAT0 * 1.0 * exp(P.LovR_liq * (P.RTEM00 - rtem));
}
NSW6_HD double psat_ice(const DevParams &P, double tem)
{
    double rtem = 1.0 / dmax(tem, TEM_MIN);
    return C_PSAT0 * 1.0 * exp(P.LovR_ice * (P.RTEM00 - rtem));
}

// Bergeron_param at one point (keeps the reference's off-by-one interpolation).  The
// scalar type S is double in the kernels; a dual-number scalar carries the
// temperature tangent through the interpolation (the table cell is piecewise
// constant) and the power (its pow(double, S) overload).
template <typename S>
NSW6_HD void bergeron(const double *a1_tab, const double *a2_tab, S tem, S &a1, S &a2, S &ma2)
{
    S temc = dmin(dmax(tem - C_TEM00, S(-30.99)), S(0.0));
    // the table index from a NaN-free copy: the same index for every finite temperature,
    // and a NaN temperature (which the clamps above pass through) cannot index outside
    // the tables; its NaN still reaches a1, a2 through `fact`
    const double tc = val(temc);
    int itemc = (int)(-((tc == tc) ? tc : 0.0));
    S fact = -(temc + (double)(itemc - 1));
    S v1 = (1.0 - fact) * a1_tab[itemc] + (fact)*a1_tab[itemc + 1];
    S v2 = (1.0 - fact) * a2_tab[itemc] + (fact)*a2_tab[itemc + 1];
    ma2 = 1.0 - v2;
    a2 = v2;
    a1 = v1 * pow(1.0E-3, ma2);
}

// Discrete switches of TagBigLoop that depend only on the point's input state
// (src/mod_mp_nsw6.cpp:935-951, 956, 971, 997, 1060, 1094, 1195-1207).  Both
// forms evaluate them in FP64: on FP32-rounded inputs a switch flips near its
// threshold (0 degC, q = 1e-4, ...) and changes the point's tendencies by O(1).
struct RateSwitches {
    double delta1, delta2, spsati, iceflg, sw_bergeron;
    double zr, zs, zx, zg, zcr;   // zerosw of rain, snow (non-Roh), Roh Xs2, graupel, qc*qr
    double sw_hom, sw_htr, sw_mlt;
};
NSW6_HD RateSwitches rate_switches(const DevParams &P, double dens, double temp,
                                                      const double q[6], double qsati)
{
    RateSwitches s;
    const double qv = dmax(q[QV], 0.0), qc = dmax(q[QC], 0.0), qr = dmax(q[QR], 0.0);
    const double qs = dmax(q[QS], 0.0), qg = dmax(q[QG], 0.0);
    const double Sice = qv / dmax(qsati, C_EPS);
    const double temc = temp - C_TEM00;
    s.delta1 = step_pos(qr - 1.0E-4);
    s.delta2 = step_pos(1.0E-4 - qr) * step_pos(1.0E-4 - qs);
    s.spsati = step_pos(Sice - 1.0);
    s.iceflg = step_neg(temc);
    s.sw_bergeron = step_pos(temc + 30.0) * step_pos(0.0 - temc) * (1.0 - P.sw_expice);
    s.zr = step_neg(qr - 1.0E-12);
    s.zs = step_neg(qs - 1.0E-12);
    s.zx = step_neg(dens * qs * P.rAs - 1.0E-12);
    s.zg = step_neg(qg - 1.0E-12);
    s.zcr = step_neg(qc * qr - 1.0E-12);
    s.sw_hom = step_neg(temc + 40.0);
    s.sw_htr = step_pos(temc + 40.0) * step_neg(temc);
    s.sw_mlt = step_pos(temc);
    return s;
}

// Effective radius of liquid water and the cloud-top diagnostics of one point
// (src/mod_mp_nsw6.cpp:1566-1592), in FP64 for both forms.
NSW6_HD void effective_radius(const DevParams &P, const RateSwitches &sw, double dens,
                                                 double temp, const double q[6], double Nc, double &rceff,
                                                 double &rceff_cld, double &rctop, double &tctop,
                                                 double &cloud_sw)
{
    const double qc = dmax(q[QC], 0.0), qr = dmax(q[QR], 0.0);
    const double zr = sw.zr;
    const double RLMDr = sqrt(sqrt(dens * qr * P.rK_r + zr)) * (1.0 - zr);
    const double RLMDr_3 = (RLMDr * RLMDr) * RLMDr;
    const double rhoqc = dens * qc * 1000.0;
    const double coef_xf = P.coef_xf;
    const double xf_qc = rhoqc / Nc * 1.0E-9;
    const double rf_qc = exp(0.33333333 * log(coef_xf * xf_qc + C_EPS));
    const double r2_qc = P.coef_dgam * (rf_qc * rf_qc) * (Nc * 1.0E+6);
    const double r2_qr = 0.25 * P.N0r * P.GAM_3 * RLMDr_3;
    const double r3_qc = coef_xf * dens * qc;
    const double r3_qr = coef_xf * dens * qr;
    double zerosw = step_neg((r2_qc + r2_qr) - 1.0E-10);
    rceff = (r3_qc + r3_qr) / (r2_qc + r2_qr + zerosw) * (1.0 - zerosw);
    const double s = step_pos((qc + qr) - 1.0E-5) * zerosw;
    rctop = (s)*rceff + (1.0 - s) * C_UNDEF;
    tctop = (s)*temp + (1.0 - s) * C_UNDEF;
    // the switch the comment of src/mod_mp_nsw6.cpp:1577 describes (cloud present)
    cloud_sw = step_pos((qc + qr) - 1.0E-5) * (1.0 - zerosw);
    zerosw = step_neg(r2_qc - 1.0E-10);
    rceff_cld = r3_qc / (r2_qc + zerosw) * (1.0 - zerosw);
}

// Input of one rate evaluation.
template <typename R> struct RateIn {
    R dens, temp, q[6], qsatl, qsati, pre, rhog, a1, a2, ma2;
    RateSwitches sw;   // FP64 switches (exactly 0 or 1)
    // differences of nearly equal FP64 inputs, formed in FP64 and rounded once (in
    // FP32 they would lose most digits near 0 degC and near saturation)
    R temc;            // temp - TEM00
    R sice_m1;         // Sice - 1
    R sliq_1m;         // 1 - min(Sliq, 1)
    R qv_qsati;        // max(qv, 0) - qsati
    // the 8 tunable coefficients: kk_aut_rate, qicrt_saut, Esi, Esr, Cr, Cs, Cg,
    // icenuc_coef (carried here so that a dual-number scalar can hold their tangents)
    R tun[NCOEF];
};
// The tunable coefficients of a parameter set, in the order above.
NSW6_HD void tunables(const DevParams &P, double tun[NCOEF])
{
    tun[0] = P.kk_aut_rate;
    tun[1] = P.qicrt_saut;
    tun[2] = P.Esi;
    tun[3] = P.Esr;
    tun[4] = P.Cr;
    tun[5] = P.Cs;
    tun[6] = P.Cg;
    tun[7] = P.icenuc_coef;
}
// Fill the differences of RateIn from the precise inputs (type D: double for the
// HIP forms, a dual number in the gradient check).
template <typename R, typename D>
NSW6_HD void rate_differences(RateIn<R> &in, D temp, const D q[6], D qsatl, D qsati)
{
    const D qv = dmax(q[QV], D(0.0));
    const D Sliq = qv / dmax(qsatl, D(C_EPS));
    const D Sice = qv / dmax(qsati, D(C_EPS));
    in.temc = R(temp - C_TEM00);
    in.sice_m1 = R(Sice - 1.0);
    in.sliq_1m = R(1.0 - dmin(Sliq, D(1.0)));
    in.qv_qsati = R(qv - qsati);
}
// Output of one rate evaluation: the limited tendencies q_t (1/s) of QC..QG (the
// caller closes QV and forms rhog * q_t * dt in FP64) and the fall speeds.
template <typename R> struct RateOut {
    R qt[6];
    R Vtr, Vti, Vts, Vtg;
};

// TagBigLoop at one point; `nc` is the point's droplet number (nc_default or
// nc_indirect).
template <typename R>
NSW6_HD void rates(const DevParams &P, R dt, const RateIn<R> &in, RateOut<R> &o, const NcPoint &nc)
{
    const R PI = R(C_PI), EPS = R(C_EPS), UNDEF = R(C_UNDEF);
    const R Rvap = R(C_Rvap), LHV0 = R(C_LHV0), LHS0 = R(C_LHS0), LHF0 = R(C_LHF0);
    const R PRE00 = R(C_Pstd), TEM00 = R(C_TEM00), CL = R(C_CL), rho_w = R(C_DWATR);
    const R Ar = R(P.Ar), As = R(P.As), Ag = R(P.Ag), Cr = in.tun[4], Cs = in.tun[5], Cg = in.tun[6];
    const R N0r = R(P.N0r), N0s = R(P.N0s), N0g = R(P.N0g);
    const R GAM = R(P.GAM), GAM_2 = R(P.GAM_2), GAM_3 = R(P.GAM_3);
    const R dens = in.dens, temp = in.temp;

    const R qv = dmax(in.q[QV], R(0)), qc = dmax(in.q[QC], R(0)), qr = dmax(in.q[QR], R(0));
    const R qi = dmax(in.q[QI], R(0)), qs = dmax(in.q[QS], R(0)), qg = dmax(in.q[QG], R(0));

    const R Rdens = R(1) / dens;
    const R rho_fact = sqrt(R(DENS00) * Rdens);
    const R temc = in.temc;

    const R delta1 = R(in.sw.delta1);
    const R delta2 = R(in.sw.delta2);
    const R spsati = R(in.sw.spsati);
    const R iceflg = R(in.sw.iceflg);
    const R rdt = R(1) / dt;
    const R dqv_dt = qv * rdt, dqc_dt = qc * rdt, dqr_dt = qr * rdt;
    const R dqi_dt = qi * rdt, dqs_dt = qs * rdt, dqg_dt = qg * rdt;
    const R sw_expice = R(P.sw_expice);
    const R sw_bergeron = R(in.sw.sw_bergeron);

    // slope parameter lambda (rain)
    R zerosw = R(in.sw.zr);
    const R RLMDr = sqrt(sqrt(dens * qr * R(P.rK_r) + zerosw)) * (R(1) - zerosw);
    const R RLMDr_dr = sqrt(RLMDr);
    const R RLMDr_2 = RLMDr * RLMDr;
    const R RLMDr_3 = RLMDr_2 * RLMDr;
    const R RLMDr_1br = RLMDr_2 * RLMDr_2;
    const R RLMDr_2br = RLMDr_1br * RLMDr;
    const R RLMDr_3br = RLMDr_3 * RLMDr_3;
    const R RLMDr_7 = RLMDr_3br * RLMDr;
    const R RLMDr_3dr = RLMDr_3 * RLMDr_dr;
    const R RLMDr_5dr = RLMDr_2br * RLMDr_dr;
    const R RLMDr_6dr = RLMDr_3br * RLMDr_dr;

    // slope parameter lambda (snow) and its moments (only the non-Roh form needs them)
    const bool roh = (P.sw_roh2014 == 1.0);
    R MOMs_0 = R(0), MOMs_1 = R(0), MOMs_2 = R(0), MOMs_0bs = R(0), MOMs_1bs = R(0), MOMs_2bs = R(0);
    R MOMs_2ds = R(0), MOMs_5ds_h = R(0), RMOMs_Vt = R(0);
    if (!roh) {
        zerosw = R(in.sw.zs);
        const R RLMDs = sqrt(sqrt(dens * qs * R(P.rK_s) + zerosw)) * (R(1) - zerosw);
        const R RLMDs_ds = sqrt(sqrt(RLMDs));
        const R RLMDs_2 = RLMDs * RLMDs;
        const R RLMDs_3 = RLMDs_2 * RLMDs;
        const R RLMDs_1bs = RLMDs_2 * RLMDs_2;
        const R RLMDs_2bs = RLMDs_1bs * RLMDs;
        const R RLMDs_3bs = RLMDs_3 * RLMDs_3;
        const R RLMDs_3ds = RLMDs_3 * RLMDs_ds;
        const R RLMDs_5ds = RLMDs_2bs * RLMDs_ds;
        MOMs_0 = N0s * GAM * RLMDs;
        MOMs_1 = N0s * GAM_2 * RLMDs_2;
        MOMs_2 = N0s * GAM_3 * RLMDs_3;
        MOMs_0bs = N0s * R(P.GAM_1bs) * RLMDs_1bs;
        MOMs_1bs = N0s * R(P.GAM_2bs) * RLMDs_2bs;
        MOMs_2bs = N0s * R(P.GAM_3bs) * RLMDs_3bs;
        MOMs_2ds = N0s * R(P.GAM_3ds) * RLMDs_3ds;
        MOMs_5ds_h = N0s * R(P.GAM_5ds_h) * sqrt(RLMDs_5ds);
        RMOMs_Vt = R(P.GAM_1bsds) / R(P.GAM_1bs) * RLMDs_ds;
    }

    // Roh and Satoh (2014) bimodal snow moments
    const R Xs2 = dens * qs * R(P.rAs);
    zerosw = R(in.sw.zx);
    const R tems = dmin(R(-0.1), temc);
    R at[4], bt[4];
    at[0] = R(P.coef_a[0]) + tems * (R(P.coef_a[1]) + tems * (R(P.coef_a[4]) + tems * R(P.coef_a[8])));
    at[1] = R(P.coef_a[2]) + tems * (R(P.coef_a[3]) + tems * R(P.coef_a[6]));
    at[2] = R(P.coef_a[5]) + tems * R(P.coef_a[7]);
    at[3] = R(P.coef_a[9]);
    bt[0] = R(P.coef_b[0]) + tems * (R(P.coef_b[1]) + tems * (R(P.coef_b[4]) + tems * R(P.coef_b[8])));
    bt[1] = R(P.coef_b[2]) + tems * (R(P.coef_b[3]) + tems * R(P.coef_b[6]));
    bt[2] = R(P.coef_b[5]) + tems * R(P.coef_b[7]);
    bt[3] = R(P.coef_b[9]);
    const R sw_roh = R(P.sw_roh2014), ln10 = R(P.ln10);
    const R lx = log(Xs2 + zerosw);
    const R omz = R(1) - zerosw;
    // moment of order nm: 10^loga(nm) * Xs2^b(nm)
#define NSW6_MOM(NM, OUT)                                                                        \
    {                                                                                            \
        const R nm = R(NM);                                                                      \
        const R loga_ = at[0] + nm * (at[1] + nm * (at[2] + nm * at[3]));                        \
        const R b_ = bt[0] + nm * (bt[1] + nm * (bt[2] + nm * bt[3]));                           \
        OUT = exp(ln10 * loga_ + lx * b_);                                                       \
    }
    R m;
    if (roh) {
        m = exp(ln10 * at[0] + lx * bt[0]);
        MOMs_0 = m * omz;
        NSW6_MOM(1.0, m)
        MOMs_1 = m * omz;
        MOMs_2 = Xs2;
        NSW6_MOM(2.0, m)
        MOMs_0bs = m * omz;
        NSW6_MOM(3.0, m)
        MOMs_1bs = m * omz;
        NSW6_MOM(4.0, m)
        MOMs_2bs = m * omz;
        NSW6_MOM(2.25, m)
        MOMs_2ds = m * omz;
        NSW6_MOM(1.625, m)
        MOMs_5ds_h = m * omz;
        // RMOMs_Vt uses the (nm = 2.25) moment of MOMs_2ds
        RMOMs_Vt = MOMs_2ds / (MOMs_0bs + zerosw);
    } else {
        m = exp(ln10 * at[0] + lx * bt[0]);
        MOMs_0 = sw_roh * m * omz + (R(1) - sw_roh) * MOMs_0;
        NSW6_MOM(1.0, m)
        MOMs_1 = sw_roh * m * omz + (R(1) - sw_roh) * MOMs_1;
        MOMs_2 = sw_roh * Xs2 + (R(1) - sw_roh) * MOMs_2;
        NSW6_MOM(2.0, m)
        MOMs_0bs = sw_roh * m * omz + (R(1) - sw_roh) * MOMs_0bs;
        NSW6_MOM(3.0, m)
        MOMs_1bs = sw_roh * m * omz + (R(1) - sw_roh) * MOMs_1bs;
        NSW6_MOM(4.0, m)
        MOMs_2bs = sw_roh * m * omz + (R(1) - sw_roh) * MOMs_2bs;
        NSW6_MOM(2.25, m)
        const R m225 = m;
        MOMs_2ds = sw_roh * m * omz + (R(1) - sw_roh) * MOMs_2ds;
        NSW6_MOM(1.625, m)
        MOMs_5ds_h = sw_roh * m * omz + (R(1) - sw_roh) * MOMs_5ds_h;
        RMOMs_Vt = sw_roh * m225 * omz / (MOMs_0bs + zerosw) + (R(1) - sw_roh) * RMOMs_Vt;
    }
#undef NSW6_MOM

    // slope parameter lambda (graupel)
    zerosw = R(in.sw.zg);
    const R RLMDg = sqrt(sqrt(dens * qg * R(P.rK_g) + zerosw)) * (R(1) - zerosw);
    const R RLMDg_dg = sqrt(RLMDg);
    const R RLMDg_2 = RLMDg * RLMDg;
    const R RLMDg_3 = RLMDg_2 * RLMDg;
    const R RLMDg_3dg = RLMDg_3 * RLMDg_dg;
    const R RLMDg_5dg = (RLMDg_3 * RLMDg_2) * RLMDg_dg;

    // terminal velocity; QI sediments only under "4WATER" (otherwise Vti is unused)
    R Vti = R(0);
    if (P.four_water != 0.0) {
        const R zi = step_neg(qi - R(1.0E-8));
        Vti = (R(1) - R(P.sw_constVti)) * R(-3.29) * exp(log(dens * qi + zi) * R(0.16)) * (R(1) - zi) +
              R(P.sw_constVti) * R(-P.CONST_Vti);
    }
    const R Vtr = -Cr * rho_fact * R(P.Vtr_fac) * RLMDr_dr;
    const R Vts = -Cs * rho_fact * RMOMs_Vt;
    const R Vtg = -Cg * rho_fact * R(P.Vtg_fac) * RLMDg_dg;

    // nucleation
    const R Ni0 = dmax(exp(R(-0.1) * temc), R(1)) * R(1000);
    const R Qi0 = in.tun[7] * exp(log(Ni0) * R(1.33)) * Rdens;
    R Pigen = dmax(dmin(Qi0 - qi, in.qv_qsati), R(0)) * rdt;

    // accretion
    const R Esi = in.tun[2], Egs = R(P.Egs);
    const R e_sacr = exp(R(P.gamma_sacr) * temc);
    const R Esi_mod = dmin(Esi, Esi * e_sacr);
    // min(Egs, Egs exp(.)) is exactly Egs = 0 under Roh & Satoh (Egs = 0): no exp needed
    const R Egs_mod = (P.Egs == 0.0) ? Egs : dmin(Egs, Egs * exp(R(P.gamma_gacs) * temc));
    const R GAM_3dr = R(P.GAM_3dr), GAM_3dg = R(P.GAM_3dg);
    const bool kk = (P.sw_kk2000 == 1.0);
    zerosw = R(in.sw.zcr);
    const R Pracw_kk = R(67) * exp(log(qc * qr + zerosw) * R(1.15)) * (R(1) - zerosw);
    const R sw_kk = R(P.sw_kk2000);
    R Pracw;
    if (kk) {
        Pracw = Pracw_kk;
    } else {
        const R Pracw_orig = qc * R(0.25) * PI * R(P.Erw) * N0r * Cr * GAM_3dr * RLMDr_3dr * rho_fact;
        Pracw = (R(1) - sw_kk) * Pracw_orig + (sw_kk)*Pracw_kk;
    }
    R Psacw = qc * R(0.25) * PI * R(P.Esw) * Cs * MOMs_2ds * rho_fact;
    R Pgacw = qc * R(0.25) * PI * R(P.Egw) * N0g * Cg * GAM_3dg * RLMDg_3dg * rho_fact;
    R Praci = qi * R(0.25) * PI * R(P.Eri) * N0r * Cr * GAM_3dr * RLMDr_3dr * rho_fact;
    R Psaci = qi * R(0.25) * PI * Esi_mod * Cs * MOMs_2ds * rho_fact;
    R Pgaci = qi * R(0.25) * PI * R(P.Egi) * N0g * Cg * GAM_3dg * RLMDg_3dg * rho_fact;
    R Piacr = qi * Ar * R(P.rmi) * R(0.25) * PI * R(P.Eri) * N0r * Cr * R(P.GAM_6dr) * RLMDr_6dr * rho_fact;
    const R GAM_1br = R(P.GAM_1br), GAM_2br = R(P.GAM_2br), GAM_3br = R(P.GAM_3br);
    const R Esr = in.tun[3];
    R Psacr = Ar * R(0.25) * PI * Rdens * Esr * N0r * fabs(Vtr - Vts) *
              (GAM_1br * RLMDr_1br * MOMs_2 + R(2) * GAM_2br * RLMDr_2br * MOMs_1 + GAM_3br * RLMDr_3br * MOMs_0);
    R Pgacr = Ar * R(0.25) * PI * Rdens * R(P.Egr) * N0g * N0r * fabs(Vtg - Vtr) *
              (GAM_1br * RLMDr_1br * GAM_3 * RLMDg_3 + R(2) * GAM_2br * RLMDr_2br * GAM_2 * RLMDg_2 +
               GAM_3br * RLMDr_3br * GAM * RLMDg);
    R Pracs = As * R(0.25) * PI * Rdens * Esr * N0r * fabs(Vtr - Vts) *
              (MOMs_0bs * GAM_3 * RLMDr_3 + R(2) * MOMs_1bs * GAM_2 * RLMDr_2 + MOMs_2bs * GAM * RLMDr);
    R Pgacs = As * R(0.25) * PI * Rdens * Egs_mod * N0g * fabs(Vtg - Vts) *
              (MOMs_0bs * GAM_3 * RLMDg_3 + R(2) * MOMs_1bs * GAM_2 * RLMDg_2 + MOMs_2bs * GAM * RLMDg);

    // autoconversion
    const R Nc = R(nc.Nc);
    const R rhoqc = dens * qc * R(1000);
    const R Praut_kk = in.tun[0] * exp(R(2.47) * log(qc)) * R(nc.pow_m179);
    R Praut;
    if (kk) {
        Praut = Praut_kk;
    } else {
        const R Dc = R(nc.Dc);
        const R Praut_berry = Rdens * R(1.67E-5) * rhoqc * rhoqc / (R(5) + R(3.66E-2) * Nc / (Dc * rhoqc + EPS));
        Praut = (R(1) - sw_kk) * Praut_berry + sw_kk * Praut_kk;
    }
    const R beta_saut = R(P.beta_saut), beta_gaut = R(P.beta_gaut);
    // exp(gamma_saut temc) is exp(gamma_sacr temc) when the two exponents are equal
    // (6e-2 both by default); min(beta_gaut, beta_gaut exp(.)) is exactly 0 when Pgaut is
    // off (beta_gaut = 0)
    const R e_saut = (P.gamma_saut == P.gamma_sacr) ? e_sacr : exp(R(P.gamma_saut) * temc);
    const R betai = dmin(beta_saut, beta_saut * e_saut);
    R Psaut = dmax(betai * (qi - in.tun[1]), R(0));
    const R betas = (P.beta_gaut == 0.0) ? beta_gaut : dmin(beta_gaut, beta_gaut * exp(R(P.gamma_gaut) * temc));
    R Pgaut = dmax(betas * (qs - R(P.qscrt_gaut)), R(0));

    // evaporation, sublimation, melting, freezing
    const R Ka = (R(Ka0) + R(dKa_dT) * temc);
    const R Kd = (R(Kd0) + R(dKd_dT) * temc) * PRE00 / in.pre;
    const R Nu = (R(nu0) + R(dnu_dT) * temc) * Rdens;
    const R rNu = R(1) / Nu;
    // G = 1 / (A + 1 / (Kd dens qsat)) = D / (A D + 1) with D = Kd dens qsat, and
    // 1 / (LHF0 / (Ka temc)) = Ka temc / LHF0 (0 at temc = 0, as 1 / inf)
    const R rT = R(1) / temp;
    const R rKaT = rT / Ka;
    const R Dl = Kd * dens * in.qsatl, Di_ = Kd * dens * in.qsati;
    const R Glv = Dl / ((LHV0 * rKaT) * (LHV0 * R(1.0 / C_Rvap) * rT - R(1)) * Dl + R(1));
    const R Giv = Di_ / ((LHS0 * rKaT) * (LHS0 * R(1.0 / C_Rvap) * rT - R(1)) * Di_ + R(1));
    const R Gil = Ka * temc * R(1.0 / C_LHF0);
    const R ventr = R(P.f1r) * GAM_2 * RLMDr_2 + R(P.f2r) * sqrt(Cr * rho_fact * rNu * RLMDr_5dr) * R(P.GAM_5dr_h);
    R Prevp = R(2) * PI * Rdens * N0r * in.sliq_1m * Glv * ventr;

    const R rhoqi = dmax(dens * qi, EPS);
    const R XNi = dmin(dmax(R(5.38E+7) * exp(log(rhoqi) * R(0.75)), R(1.0E+3)), R(1.0E+6));
    const R XMi = rhoqi / XNi;
    const R Di = dmin(R(P.Di_a) * sqrt(XMi), R(P.Di_max));
    R tmp = R(4) * Di * XNi * Rdens * in.sice_m1 * Giv;
    R Pidep = (spsati) * (tmp);
    R Pisub = (R(1) - spsati) * (-tmp);

    R sw = R(in.sw.sw_hom);
    R Pihom = sw * qc * rdt;
    sw = R(in.sw.sw_htr);
    const R A_frz = R(P.A_frz), B_frz = R(P.B_frz);
    // exp(-A_frz temc) - 1 cancels near 0 degC: FP32 takes expm1, FP64 keeps the
    // reference's form
    const R frz_em1 = (sizeof(R) == sizeof(float)) ? R(expm1(-A_frz * temc)) : (exp(-A_frz * temc) - R(1));
    R Pihtr = sw * (dens * R(1.0 / C_DWATR) * (qc * qc) * R(P.rNc_ihtr)) * B_frz * frz_em1;
    sw = R(in.sw.sw_mlt);
    R Pimlt = sw * qi * rdt;

    const R vents = R(P.f1s) * MOMs_1 + R(P.f2s) * sqrt(Cs * rho_fact * rNu) * MOMs_5ds_h;
    tmp = R(2) * PI * Rdens * in.sice_m1 * Giv * vents;
    R Psdep = (spsati) * (tmp);
    R Pssub = (R(1) - spsati) * (-tmp);
    R Psmlt = R(2) * PI * Rdens * Gil * vents + CL * temc * R(1.0 / C_LHF0) * (Psacw + Psacr);
    Psmlt = dmax(Psmlt, R(0));

    const R ventg = R(P.f1g) * GAM_2 * RLMDg_2 + R(P.f2g) * sqrt(Cg * rho_fact * rNu * RLMDg_5dg) * R(P.GAM_5dg_h);
    tmp = R(2) * PI * Rdens * N0g * in.sice_m1 * Giv * ventg;
    R Pgdep = (spsati) * (tmp);
    R Pgsub = (R(1) - spsati) * (-tmp);
    R Pgmlt = R(2) * PI * Rdens * N0g * Gil * ventg + CL * temc * R(1.0 / C_LHF0) * (Pgacw + Pgacr);
    Pgmlt = dmax(Pgmlt, R(0));
    R Pgfrz = R(2) * PI * Rdens * N0r * R(60) * B_frz * Ar * frz_em1 * RLMDr_7;

    // Bergeron process (zeroed by sw_bergeron = 0 when explicit ice generation is on)
    R Psfw = R(0), Psfi = R(0);
    if (P.sw_expice != 1.0) {
        const R mi50 = R(P.mi50);
        const R dt1 = (exp(R(P.ln_mi50) * in.ma2) - exp(R(P.ln_mi40) * in.ma2)) / (in.a1 * in.ma2);
        const R Ni50 = qi * dt / (mi50 * dt1);
        Psfw = Ni50 * (in.a1 * exp(in.a2 * R(P.ln_mi50)) + PI * R(P.Eiw) * dens * qc * R(P.Ri50) * R(P.Ri50) * R(P.vti50));
        Psfi = qi / dt1;
    }

    // limiters
    Pigen = dmin(Pigen, dqv_dt) * (iceflg)*sw_expice;
    Pidep = dmin(Pidep, dqv_dt) * (iceflg)*sw_expice;
    Psdep = dmin(Psdep, dqv_dt) * (iceflg);
    Pgdep = dmin(Pgdep, dqv_dt) * (iceflg);

    Pracw = Pracw + Psacw * (R(1) - iceflg) + Pgacw * (R(1) - iceflg);

    Praut = dmin(Praut, dqc_dt);
    Pracw = dmin(Pracw, dqc_dt);
    Pihom = dmin(Pihom, dqc_dt) * (iceflg)*sw_expice;
    Pihtr = dmin(Pihtr, dqc_dt) * (iceflg)*sw_expice;
    Psacw = dmin(Psacw, dqc_dt) * (iceflg);
    Psfw = dmin(Psfw, dqc_dt) * (iceflg)*sw_bergeron;
    Pgacw = dmin(Pgacw, dqc_dt) * (iceflg);

    Prevp = dmin(Prevp, dqr_dt);
    Piacr = dmin(Piacr, dqr_dt) * (iceflg);
    Psacr = dmin(Psacr, dqr_dt) * (iceflg);
    Pgacr = dmin(Pgacr, dqr_dt) * (iceflg);
    Pgfrz = dmin(Pgfrz, dqr_dt) * (iceflg);

    Pisub = dmin(Pisub, dqi_dt) * (iceflg)*sw_expice;
    Pimlt = dmin(Pimlt, dqi_dt) * (R(1) - iceflg) * sw_expice;
    Psaut = dmin(Psaut, dqi_dt) * (iceflg);
    Praci = dmin(Praci, dqi_dt) * (iceflg);
    Psaci = dmin(Psaci, dqi_dt) * (iceflg);
    Psfi = dmin(Psfi, dqi_dt) * (iceflg)*sw_bergeron;
    Pgaci = dmin(Pgaci, dqi_dt) * (iceflg);

    Pssub = dmin(Pssub, dqs_dt) * (iceflg);
    Psmlt = dmin(Psmlt, dqs_dt) * (R(1) - iceflg);
    Pgaut = dmin(Pgaut, dqs_dt) * (iceflg);
    Pracs = dmin(Pracs, dqs_dt) * (iceflg);
    Pgacs = dmin(Pgacs, dqs_dt);

    Pgsub = dmin(Pgsub, dqg_dt) * (iceflg);
    Pgmlt = dmin(Pgmlt, dqg_dt) * (R(1) - iceflg);

    R Piacr_s = (R(1) - delta1) * Piacr;
    R Piacr_g = (delta1)*Piacr;
    R Praci_s = (R(1) - delta1) * Praci;
    R Praci_g = (delta1)*Praci;
    R Psacr_s = (delta2)*Psacr;
    R Psacr_g = (R(1) - delta2) * Psacr;
    Pracs = (R(1) - delta2) * Pracs;

    R net, fac_sw, fac;
    // [QC]
    net = +Pimlt - Praut - Pracw - Pihom - Pihtr - Psacw - Psfw - Pgacw;
    fac_sw = step_pos(net + EPS);
    fac = fac_sw + (R(1) - fac_sw) * dmin(-dqc_dt / (net - fac_sw), R(1));
    Pimlt *= fac; Praut *= fac; Pracw *= fac; Pihom *= fac;
    Pihtr *= fac; Psacw *= fac; Psfw *= fac; Pgacw *= fac;
    // [QI]
    net = +Pigen + Pidep + Pihom + Pihtr - Pisub - Pimlt - Psaut - Praci_s - Psaci - Psfi - Praci_g - Pgaci;
    fac_sw = step_pos(net + EPS);
    fac = fac_sw + (R(1) - fac_sw) * dmin(-dqi_dt / (net - fac_sw), R(1));
    Pigen *= fac; Pidep *= fac; Pihom *= fac; Pihtr *= fac; Pisub *= fac; Pimlt *= fac;
    Psaut *= fac; Praci_s *= fac; Psaci *= fac; Psfi *= fac; Praci_g *= fac; Pgaci *= fac;
    // [QR]
    net = +Praut + Pracw + Psmlt + Pgmlt - Prevp - Piacr_s - Psacr_s - Piacr_g - Psacr_g - Pgacr - Pgfrz;
    fac_sw = step_pos(net + EPS);
    fac = fac_sw + (R(1) - fac_sw) * dmin(-dqr_dt / (net - fac_sw), R(1));
    Praut *= fac; Pracw *= fac; Psmlt *= fac; Pgmlt *= fac; Prevp *= fac; Piacr_s *= fac;
    Psacr_s *= fac; Piacr_g *= fac; Psacr_g *= fac; Pgacr *= fac; Pgfrz *= fac;
    // [QV]
    net = +Prevp + Pisub + Pssub + Pgsub - Pigen - Pidep - Psdep - Pgdep;
    fac_sw = step_pos(net + EPS);
    fac = fac_sw + (R(1) - fac_sw) * dmin(-dqv_dt / (net - fac_sw), R(1));
    Prevp *= fac; Pisub *= fac; Pssub *= fac; Pgsub *= fac;
    Pigen *= fac; Pidep *= fac; Psdep *= fac; Pgdep *= fac;
    // [QS]
    net = +Psdep + Psacw + Psfw + Piacr_s + Psacr_s + Psaut + Praci_s + Psaci + Psfi - Pssub - Psmlt - Pgaut -
          Pracs - Pgacs;
    fac_sw = step_pos(net + EPS);
    fac = fac_sw + (R(1) - fac_sw) * dmin(-dqs_dt / (net - fac_sw), R(1));
    Psdep *= fac; Psacw *= fac; Psfw *= fac; Piacr_s *= fac; Psacr_s *= fac; Psaut *= fac; Praci_s *= fac;
    Psaci *= fac; Psfi *= fac; Pssub *= fac; Psmlt *= fac; Pgaut *= fac; Pracs *= fac; Pgacs *= fac;
    // [QG]
    net = +Pgdep + Pgacw + Piacr_g + Psacr_g + Pgacr + Pgfrz + Praci_g + Pgaci + Pgaut + Pracs + Pgacs - Pgsub -
          Pgmlt;
    fac_sw = step_pos(net + EPS);
    fac = fac_sw + (R(1) - fac_sw) * dmin(-dqg_dt / (net - fac_sw), R(1));
    Pgdep *= fac; Pgacw *= fac; Piacr_g *= fac; Psacr_g *= fac; Pgacr *= fac; Pgfrz *= fac; Praci_g *= fac;
    Pgaci *= fac; Pgaut *= fac; Pracs *= fac; Pgacs *= fac; Pgsub *= fac; Pgmlt *= fac;

    R qc_t = +Pimlt - Praut - Pracw - Pihom - Pihtr - Psacw - Psfw - Pgacw;
    R qr_t = +Praut + Pracw + Psmlt + Pgmlt - Prevp - Piacr_s - Psacr_s - Piacr_g - Psacr_g - Pgacr - Pgfrz;
    R qi_t = +Pigen + Pidep + Pihom + Pihtr - Pisub - Pimlt - Psaut - Praci_s - Psaci - Psfi - Praci_g - Pgaci;
    R qs_t = +Psdep + Psacw + Psfw + Piacr_s + Psacr_s + Psaut + Praci_s + Psaci + Psfi - Pssub - Psmlt - Pgaut -
             Pracs - Pgacs;
    R qg_t = +Pgdep + Pgacw + Piacr_g + Psacr_g + Pgacr + Pgfrz + Praci_g + Pgaci + Pgaut + Pracs + Pgacs - Pgsub -
             Pgmlt;
    qc_t = dmax(qc_t, -dqc_dt);
    qr_t = dmax(qr_t, -dqr_dt);
    qi_t = dmax(qi_t, -dqi_dt);
    qs_t = dmax(qs_t, -dqs_dt);
    qg_t = dmax(qg_t, -dqg_dt);
    o.qt[QV] = R(0);
    o.qt[QC] = qc_t;
    o.qt[QR] = qr_t;
    o.qt[QI] = qi_t;
    o.qt[QS] = qs_t;
    o.qt[QG] = qg_t;
    o.Vtr = Vtr;
    o.Vti = Vti;
    o.Vts = Vts;
    o.Vtg = Vtg;
}
template <typename R>
NSW6_HD void rates(const DevParams &P, R dt, const RateIn<R> &in, RateOut<R> &o)
{
    rates<R>(P, dt, in, o, nc_default(P));
}

}  // namespace nsw6
