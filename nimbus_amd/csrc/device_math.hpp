// Device-side numerics shared by the kernels: thermodynamics, Kessler rates,
// WENO5/MC face reconstruction and UP5 derivatives. Each function is
// templated on the arithmetic type A (float for the FP32 bulk of the mixed
// mode, double for the FP64 mode).
#pragma once

#include "nimbus_internal.hpp"

#include <cfloat>

namespace nimbus::dev
{

template <class A>
__device__ __forceinline__ A clampv(A x, A lo, A hi)
{
    return x < lo ? lo : (x > hi ? hi : x);
}

template <class A>
__device__ __forceinline__ bool finite(A x)
{
    return isfinite(x);
}

template <class A>
struct Eps;
template <>
struct Eps<float>
{
    static constexpr float value = FLT_EPSILON;
};
template <>
struct Eps<double>
{
    static constexpr double value = DBL_EPSILON;
};

// Power and exponential of the FP32 bulk through the hardware base-2
// logarithm and exponential (v_log_f32, v_exp_f32: about 1 ulp each; the
// composed result is within a few ulp for the arguments used here); FP64
// keeps the library functions. x must be >= 0 for fpow (fpow(0, y > 0) = 0).
template <class A>
__device__ __forceinline__ A fpow(A x, A y)
{
    if constexpr (sizeof(A) == sizeof(float)) return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x));
    else return pow(x, y);
}

template <class A>
__device__ __forceinline__ A fexp(A x)
{
    if constexpr (sizeof(A) == sizeof(float)) return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
    else return exp(x);
}

// ---- thermodynamics (src/microphysics/base/thermodynamics.hpp:29-470) ----

template <class A>
__device__ __forceinline__ A theta_to_temperature(A theta, A p)
{
    if (!finite(theta) || !finite(p)) return A(kT0);
    const A th = clampv(theta, A(150), A(700));
    const A pp = clampv(p, A(100), A(200000));
    const A T = th * fpow(pp / A(kP00), A(kKappa));
    return finite(T) ? T : A(kT0);
}

template <class A>
__device__ __forceinline__ A theta_tendency_from_temperature(A dT, A p)
{
    if (!finite(dT) || !finite(p)) return A(0);
    const A pp = clampv(p, A(100), A(200000));
    const A r = dT * fpow(A(kP00) / pp, A(kKappa));
    return finite(r) ? r : A(0);
}

template <class A>
__device__ __forceinline__ A qvs_water(A T, A p)
{
    if (!finite(T) || !finite(p)) return A(0);
    const A Tc = clampv(T, A(150), A(400)) - A(kT0);
    A es = A(611.21) * fexp((A(18.678) - Tc / A(234.5)) * Tc / (A(257.14) + Tc));
    es = finite(es) ? fmax(A(0), es) : A(0);
    const A pp = clampv(p, A(100), A(200000));
    const A q = A(0.62198) * es / fmax(pp - es, A(1));
    return clampv(q, A(0), A(0.2));
}

template <class A>
__device__ __forceinline__ A qvs_slope(A T, A qs, A p)
{
    if (!(T > A(150) && T < A(400)) || !(qs > A(0) && qs < A(0.2))) return A(0);
    const A pp = clampv(p, A(100), A(200000));
    const A e = qs * pp / (A(0.62198) + qs);
    const A Tc = T - A(kT0);
    const A den = A(257.14) + Tc;
    const A num = (A(18.678) - Tc / A(234.5)) * Tc;
    const A dnum = A(18.678) - A(2) * Tc / A(234.5);
    const A dlog = (dnum * den - num) / (den * den);
    return qs * pp / fmax(pp - e, A(1)) * dlog;
}

// Bracketed Newton saturation adjustment, at most 10 iterations.
template <class A>
__device__ A saturation_adjust(A T_in, A p, A& qv, A& qc)
{
    const A L = A(kLv / kCp);
    A T = finite(T_in) ? clampv(T_in, A(150), A(400)) : A(kT0);
    qv = finite(qv) ? fmax(A(0), qv) : A(0);
    qc = finite(qc) ? fmax(A(0), qc) : A(0);
    A qs = qvs_water(T, p);
    if (!finite(qs) || qs < A(0)) qs = A(0);
    const A tol = A(8) * Eps<A>::value;
    if (qv > qs)
    {
        const A cmax = fmin(qv, fmax(A(0), (A(400) - T) / L));
        A c = cmax;
        if (qv - cmax - qvs_water(T + L * cmax, p) <= A(0))
        {
            A lo = A(0), hi = cmax;
            c = clampv((qv - qs) / (A(1) + L * qvs_slope(T, qs, p)), lo, hi);
            for (int it = 0; it < 10; ++it)
            {
                const A Ta = T + L * c;
                const A qa = qvs_water(Ta, p);
                const A f = qv - c - qa;
                if (fabs(f) <= tol * fmax(A(1), fmax(fabs(qv), fabs(qa)))) break;
                if (f > A(0)) lo = c; else hi = c;
                const A cand = c - f / (-A(1) - L * qvs_slope(Ta, qa, p));
                c = (cand > lo && cand < hi) ? cand : A(0.5) * (lo + hi);
            }
        }
        qv -= c;
        qc += c;
        T += L * c;
    }
    else if (qv < qs && qc > A(0))
    {
        const A emax = fmin(qc, fmax(A(0), (T - A(150)) / L));
        A e = emax;
        if (qv + emax - qvs_water(T - L * emax, p) >= A(0))
        {
            A lo = A(0), hi = emax;
            e = clampv((qs - qv) / (A(1) + L * qvs_slope(T, qs, p)), lo, hi);
            for (int it = 0; it < 10; ++it)
            {
                const A Ta = T - L * e;
                const A qa = qvs_water(Ta, p);
                const A f = qv + e - qa;
                if (fabs(f) <= tol * fmax(A(1), fmax(fabs(qv), fabs(qa)))) break;
                if (f < A(0)) lo = e; else hi = e;
                const A cand = e - f / (A(1) + L * qvs_slope(Ta, qa, p));
                e = (cand > lo && cand < hi) ? cand : A(0.5) * (lo + hi);
            }
        }
        qc -= e;
        qv += e;
        T -= L * e;
    }
    qc = fmax(A(0), qc);
    qv = fmax(A(0), qv);
    return T;
}

template <class A>
__device__ __forceinline__ A kessler_rain_speed(A q, A rho)
{
    if (q < A(1.0e-9)) return A(0);
    return A(36.34) * fpow(A(1.0e-3) * rho * q, A(0.1346));
}

__device__ __forceinline__ double eos_pressure(double rho, double theta)
{
    return kP00 * pow(rho * kRd * theta / kP00, kGamma);
}

// ---- WENO5-JS and MC slope (weno5_face.hpp:43-120; advection_cartesian.cpp:135-162) ----

__device__ __forceinline__ double weno5_face(double a, double b, double c, double d, double e)
{
    const double q0 = (1.0 / 3.0) * a - (7.0 / 6.0) * b + (11.0 / 6.0) * c;
    const double q1 = -(1.0 / 6.0) * b + (5.0 / 6.0) * c + (1.0 / 3.0) * d;
    const double q2 = (1.0 / 3.0) * c + (5.0 / 6.0) * d - (1.0 / 6.0) * e;
    const double t0a = a - 2.0 * b + c, t0b = a - 4.0 * b + 3.0 * c;
    const double t1a = b - 2.0 * c + d, t1b = b - d;
    const double t2a = c - 2.0 * d + e, t2b = 3.0 * c - 4.0 * d + e;
    const double s0 = 13.0 / 12.0 * t0a * t0a + 0.25 * t0b * t0b;
    const double s1 = 13.0 / 12.0 * t1a * t1a + 0.25 * t1b * t1b;
    const double s2 = 13.0 / 12.0 * t2a * t2a + 0.25 * t2b * t2b;
    const double m2 = fmax(fmax(fmax(a * a, b * b), fmax(c * c, d * d)), e * e);
    const double eps = 1.0e-6 * m2 + 1.0e-40;
    const double w0 = 0.1 / ((eps + s0) * (eps + s0));
    const double w1 = 0.6 / ((eps + s1) * (eps + s1));
    const double w2 = 0.3 / ((eps + s2) * (eps + s2));
    const double ws = w0 + w1 + w2;
    if (!(ws > 0.0) || !isfinite(ws)) return c;
    const double f = (w0 * q0 + w1 * q1 + w2 * q2) / ws;
    return isfinite(f) ? f : c;
}

// Offset form: the face value is c + result, with all differences taken
// relative to c and scaled by the stencil magnitude so FP32 keeps the
// smoothness-indicator arithmetic bounded (weno5_face.hpp:88-120). The
// reciprocals are the hardware v_rcp_f32 (1 ulp): the stencil scale cancels
// exactly in form and the weights only need relative accuracy, while a
// correctly rounded division costs eleven instructions per reciprocal.
__device__ __forceinline__ float weno5_offset(float a, float b, float c, float d, float e)
{
    const float mag = fmaxf(fmaxf(fmaxf(fabsf(a), fabsf(b)), fmaxf(fabsf(c), fabsf(d))), fabsf(e));
    const float scale = fmaxf(mag, 1e-20f);
    if (!isfinite(scale)) return 0.0f;
    const float inv = __builtin_amdgcn_rcpf(scale);
    // differences of finite physical scalars cannot overflow; a non-finite
    // stencil yields a non-finite offset, which the final check turns into 0
    auto diff = [=](float x) { return (x - c) * inv; };
    const float ac = diff(a), bc = diff(b), dc = diff(d), ec = diff(e);
    const float p0 = (1.0f / 3.0f) * ac - (7.0f / 6.0f) * bc;
    const float p1 = -(1.0f / 6.0f) * bc + (1.0f / 3.0f) * dc;
    const float p2 = (5.0f / 6.0f) * dc - (1.0f / 6.0f) * ec;
    const float s0a = ac - 2.0f * bc, s0b = ac - 4.0f * bc;
    const float s1a = bc + dc, s1b = bc - dc;
    const float s2a = ec - 2.0f * dc, s2b = ec - 4.0f * dc;
    const float b0 = (13.0f / 12.0f) * s0a * s0a + 0.25f * s0b * s0b;
    const float b1 = (13.0f / 12.0f) * s1a * s1a + 0.25f * s1b * s1b;
    const float b2 = (13.0f / 12.0f) * s2a * s2a + 0.25f * s2b * s2b;
    const float rel = mag * inv;
    const float flo = 1e-20f * inv;
    const float epsv = 1e-6f * rel * rel + flo * flo;
    // With the stencil scaled to O(1), eps + beta >= 1e-6 (or >= 1 for an
    // all-tiny stencil), so 1 / (eps + beta)^2 stays below 1e12: the weights
    // need no further normalisation.
    const float t0 = __builtin_amdgcn_rcpf(epsv + b0), t1 = __builtin_amdgcn_rcpf(epsv + b1),
                t2 = __builtin_amdgcn_rcpf(epsv + b2);
    const float a0 = 0.1f * t0 * t0, a1 = 0.6f * t1 * t1, a2 = 0.3f * t2 * t2;
    const float off = scale * ((a0 * p0 + a1 * p1 + a2 * p2) * __builtin_amdgcn_rcpf(a0 + a1 + a2));
    return isfinite(off) ? off : 0.0f;
}

template <class A>
__device__ __forceinline__ A minmod3(A a, A b, A c)
{
    if (a > A(0) && b > A(0) && c > A(0)) return fmin(a, fmin(b, c));
    if (a < A(0) && b < A(0) && c < A(0)) return fmax(a, fmax(b, c));
    return A(0);
}

template <class A>
__device__ __forceinline__ A mc_slope(A qm, A q0, A qp)
{
    const A r = qp - q0, l = q0 - qm;
    return minmod3(A(0.5) * (r + l), A(2) * r, A(2) * l);
}

// Retained-flux scalar face flux through the face between lower cell `pos`
// and upper cell `pos + 1` along one axis of extent n, from the six values
// v(-2) .. v(3) at offsets -2 .. 3 from the lower cell (values outside the
// axis may be any finite number: they are never selected). The
// reconstruction is WENO5 where the 5-point upwind stencil fits and
// MC-limited linear otherwise (advection_cartesian.cpp:1941-2173). The
// upwind stencil is selected without branching: the downwind case is the
// upwind case on the mirrored stencil, and the MC slope of a mirrored stencil
// is exactly the negated slope, so both cases share one evaluation. FP32
// forms the flux as fma(m, base, m * offset) like the reference shader.
template <class A>
__device__ __forceinline__ A face_flux6(A vm2, A vm1, A v0, A v1, A v2, A v3, int pos, int n, A m)
{
    const bool up = m >= A(0);
    const A a = up ? vm2 : v3, b = up ? vm1 : v2, c = up ? v0 : v1, d = up ? v1 : v0, e = up ? v2 : vm1;
    const bool fits = up ? (pos >= 2 && pos + 2 <= n - 1) : (pos >= 1 && pos + 3 <= n - 1);
    const bool has_mc = up ? pos >= 1 : pos + 1 <= n - 2;
    if constexpr (sizeof(A) == sizeof(double))
    {
        double phi;
        if (fits) phi = weno5_face(a, b, c, d, e);
        else phi = c + 0.5 * (has_mc ? mc_slope<double>(b, c, d) : 0.0);
        return m * phi;
    }
    else
    {
        float off;
        if (fits) off = weno5_offset(a, b, c, d, e);
        else off = 0.5f * (has_mc ? mc_slope<float>(b, c, d) : 0.0f);
        return fmaf(m, c, m * off);
    }
}

// ---- UP5/UP3/UP1 (derivative_operators.hpp:324-401) ----
template <class A, class F>
__device__ __forceinline__ A upwind_axis(F at, int pos, int n, A a, A inv)
{
    const int half = min(pos, n - 1 - pos);
    if (half <= 0) return A(0);
    const A c = at(0), m1 = at(-1), p1 = at(1);
    if (half == 1) return a >= A(0) ? (c - m1) * inv : (p1 - c) * inv;
    const A m2 = at(-2), p2 = at(2);
    if (half == 2)
        return a >= A(0) ? (A(2) * p1 + A(3) * c - A(6) * m1 + m2) * (inv * A(1.0 / 6.0))
                         : (-p2 + A(6) * p1 - A(3) * c - A(2) * m1) * (inv * A(1.0 / 6.0));
    const A m3 = at(-3), p3 = at(3);
    return a >= A(0) ? (-A(2) * m3 + A(15) * m2 - A(60) * m1 + A(20) * c + A(30) * p1 - A(3) * p2) * (inv * A(1.0 / 60.0))
                     : (A(2) * p3 - A(15) * p2 + A(60) * p1 - A(20) * c - A(30) * m1 + A(3) * m2) * (inv * A(1.0 / 60.0));
}

}  // namespace nimbus::dev
