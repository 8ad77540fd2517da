// nsw6_dual.h -- TEST INFRASTRUCTURE.  A forward-mode dual number with N tangent
// directions (value + N partial derivatives, all FP64), the scalar type of the
// host gradient check of the NSW6 adjoint.  It provides what the templated step
// (nicam_kokkos_microphysics_amd/csrc/nsw6_step.h) asks of a scalar: arithmetic,
// sqrt/exp/log/expm1/fabs/copysign, pow of a constant base, dmax/dmin selecting by
// value, val(), and the implicit-function derivatives of the saturation-adjustment
// solutions (liquid and mixed-phase).
#pragma once

#include <math.h>

#include "../nicam_kokkos_microphysics_amd/csrc/nsw6_step.h"

namespace nsw6dual {

template <int N> struct Dual {
    double v;
    double d[N];
    NSW6_HD Dual() : v(0.0)
    {
        for (int i = 0; i < N; i++) d[i] = 0.0;
    }
    NSW6_HD Dual(double x) : v(x)
    {
        for (int i = 0; i < N; i++) d[i] = 0.0;
    }
    NSW6_HD Dual &operator+=(const Dual &b)
    {
        v += b.v;
        for (int i = 0; i < N; i++) d[i] += b.d[i];
        return *this;
    }
    NSW6_HD Dual &operator-=(const Dual &b)
    {
        v -= b.v;
        for (int i = 0; i < N; i++) d[i] -= b.d[i];
        return *this;
    }
    NSW6_HD Dual &operator*=(const Dual &b)
    {
        for (int i = 0; i < N; i++) d[i] = d[i] * b.v + v * b.d[i];
        v *= b.v;
        return *this;
    }
};

template <int N> NSW6_HD double val(const Dual<N> &a) { return a.v; }

template <int N> NSW6_HD Dual<N> operator+(const Dual<N> &a) { return a; }
template <int N> NSW6_HD Dual<N> operator-(const Dual<N> &a)
{
    Dual<N> r;
    r.v = -a.v;
    for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
    return r;
}
template <int N> NSW6_HD Dual<N> operator+(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r = a;
    r += b;
    return r;
}
template <int N> NSW6_HD Dual<N> operator-(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r = a;
    r -= b;
    return r;
}
template <int N> NSW6_HD Dual<N> operator*(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    r.v = a.v * b.v;
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N> NSW6_HD Dual<N> operator/(const Dual<N> &a, const Dual<N> &b)
{
    Dual<N> r;
    r.v = a.v / b.v;
    for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
template <int N> NSW6_HD Dual<N> operator+(const Dual<N> &a, double b) { return a + Dual<N>(b); }
template <int N> NSW6_HD Dual<N> operator+(double a, const Dual<N> &b) { return Dual<N>(a) + b; }
template <int N> NSW6_HD Dual<N> operator-(const Dual<N> &a, double b) { return a - Dual<N>(b); }
template <int N> NSW6_HD Dual<N> operator-(double a, const Dual<N> &b) { return Dual<N>(a) - b; }
template <int N> NSW6_HD Dual<N> operator*(const Dual<N> &a, double b)
{
    Dual<N> r;
    r.v = a.v * b;
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b;
    return r;
}
template <int N> NSW6_HD Dual<N> operator*(double a, const Dual<N> &b) { return b * a; }
template <int N> NSW6_HD Dual<N> operator/(const Dual<N> &a, double b)
{
    Dual<N> r;
    r.v = a.v / b;
    for (int i = 0; i < N; i++) r.d[i] = a.d[i] / b;
    return r;
}
template <int N> NSW6_HD Dual<N> operator/(double a, const Dual<N> &b) { return Dual<N>(a) / b; }

#define NSW6_DUAL_CMP(OP)                                                                          \
    template <int N> NSW6_HD bool operator OP(const Dual<N> &a, const Dual<N> &b) { return a.v OP b.v; } \
    template <int N> NSW6_HD bool operator OP(const Dual<N> &a, double b) { return a.v OP b; }           \
    template <int N> NSW6_HD bool operator OP(double a, const Dual<N> &b) { return a OP b.v; }
NSW6_DUAL_CMP(<)
NSW6_DUAL_CMP(>)
NSW6_DUAL_CMP(<=)
NSW6_DUAL_CMP(>=)
NSW6_DUAL_CMP(==)
NSW6_DUAL_CMP(!=)
#undef NSW6_DUAL_CMP

template <int N> NSW6_HD Dual<N> chain(const Dual<N> &a, double fv, double dfdx)
{
    Dual<N> r;
    r.v = fv;
    for (int i = 0; i < N; i++) r.d[i] = dfdx * a.d[i];
    return r;
}
template <int N> NSW6_HD Dual<N> sqrt(const Dual<N> &a)
{
    const double s = ::sqrt(a.v);
    return chain(a, s, (s > 0.0) ? 0.5 / s : 0.0);
}
// exp underflowing to 0 has a zero tangent (x^p = exp(p log x) at x = 0, p > 1)
template <int N> NSW6_HD Dual<N> exp(const Dual<N> &a)
{
    const double e = ::exp(a.v);
    Dual<N> r;
    r.v = e;
    for (int i = 0; i < N; i++) r.d[i] = (e == 0.0) ? 0.0 : e * a.d[i];
    return r;
}
template <int N> NSW6_HD Dual<N> expm1(const Dual<N> &a) { return chain(a, ::expm1(a.v), ::exp(a.v)); }
template <int N> NSW6_HD Dual<N> log(const Dual<N> &a)
{
    Dual<N> r;
    r.v = ::log(a.v);
    for (int i = 0; i < N; i++) r.d[i] = (a.d[i] == 0.0) ? 0.0 : a.d[i] / a.v;
    return r;
}
// a^x for a constant base a > 0
template <int N> NSW6_HD Dual<N> pow(double a, const Dual<N> &x)
{
    const double p = ::pow(a, x.v);
    return chain(x, p, p * ::log(a));
}
template <int N> NSW6_HD Dual<N> fabs(const Dual<N> &a) { return chain(a, ::fabs(a.v), (a.v < 0.0) ? -1.0 : 1.0); }
// copysign(a, b) = |a| sign(b): the tangent follows a only
template <int N> NSW6_HD Dual<N> copysign(const Dual<N> &a, const Dual<N> &b)
{
    const double r = ::copysign(a.v, b.v);
    return chain(a, r, (::signbit(a.v) != ::signbit(b.v)) ? -1.0 : 1.0);
}
template <int N> NSW6_HD Dual<N> dmax(const Dual<N> &a, const Dual<N> &b) { return (a.v < b.v) ? b : a; }
template <int N> NSW6_HD Dual<N> dmin(const Dual<N> &a, const Dual<N> &b) { return (b.v < a.v) ? b : a; }

// The converged satadjust_liq temperature: the Newton's value with the tangent of the
// root of E(T; x) - Emoist(x) = 0, T' = -(dR/dx x') / (dR/dT).
template <int N>
NSW6_HD Dual<N> newton_tem(const nsw6::DevParams &P, double tem_final, Dual<N> qsum, Dual<N> qdl, Dual<N> qrest,
                           Dual<N> rho_sa, Dual<N> ein_moist)
{
    const Dual<N> R = nsw6::satadjust_residual(P, Dual<N>(tem_final), qsum, qdl, qrest, rho_sa, ein_moist);
    const double dRdT = nsw6::satadjust_dRdT(P, tem_final, qsum.v, qdl.v, qrest.v, rho_sa.v);
    Dual<N> T(tem_final);
    for (int i = 0; i < N; i++) T.d[i] = -R.d[i] / dRdT;
    return T;
}
// q(QV) = qsat at the solution: the Newton's value, the tangent of qsat(T, rho)
template <int N>
NSW6_HD Dual<N> newton_qsat(const nsw6::DevParams &P, double qv_final, Dual<N> tem, Dual<N> rho_sa)
{
    Dual<N> qs = nsw6::psat_of(P.LovR_liq, P.RTEM00, tem) / (rho_sa * nsw6::C_Rvap * tem);
    qs.v = qv_final;
    return qs;
}

// The converged satadjust_all temperature: the root of E_all(T; x) - Emoist(x) = 0,
// tangent T' = -(dR/dx x') / (dR/dT).
template <int N>
NSW6_HD Dual<N> newton_tem_all(const nsw6::DevParams &P, double tem_final, Dual<N> qsum, Dual<N> qdl,
                               Dual<N> qrest, Dual<N> rho_sa, Dual<N> ein_moist)
{
    const Dual<N> R = nsw6::satadjust_all_residual(P, Dual<N>(tem_final), qsum, qdl, qrest, rho_sa, ein_moist);
    const double dRdT = nsw6::satadjust_all_dRdT(P, tem_final, qsum.v, qdl.v, qrest.v, rho_sa.v);
    Dual<N> T(tem_final);
    for (int i = 0; i < N; i++) T.d[i] = -R.d[i] / dRdT;
    return T;
}
// q(QV) = qsat, q(QC) = (qsum - qsat) alpha, q(QI) = (qsum - qsat)(1 - alpha) at the
// solution: the Newton's values with the tangents of these functions of (T, rho, qsum)
template <int N>
NSW6_HD void newton_q_all(const nsw6::DevParams &P, const double q_final[6], Dual<N> tem, Dual<N> rho_sa,
                          Dual<N> qsum, Dual<N> &qv, Dual<N> &qc, Dual<N> &qi)
{
    const Dual<N> alpha = nsw6::satadjust_all_alpha(P, tem);
    const Dual<N> qs = nsw6::satadjust_all_qsat(P, tem, rho_sa);
    const Dual<N> d = qsum - qs;
    qv = qs;
    qc = d * alpha;
    qi = d * (1.0 - alpha);
    qv.v = q_final[nsw6::QV];
    qc.v = q_final[nsw6::QC];
    qi.v = q_final[nsw6::QI];
}

}  // namespace nsw6dual
