// Scalar policies of the IMRPhenomXAS setup: the same templated code runs
// in double (FP64 and split-precision paths), float (plain FP32) and a
// forward-mode dual of double (table derivatives for the adjoint).
//
// The library is compiled with floating-point contraction off: the
// reference's fma(a, b, c) is the unfused a * b + c (xpnr/_math.py:125-137),
// and explicit fused operations are written as fma()/fmaf() where the hot
// kernels want them.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "wfx_libm.h"

#define WFX_HD __host__ __device__
// inlined whatever its size: the set-ups' stages, so that a mode index known
// at the call site reaches every branch and array index inside as a constant
#define WFX_HDI __host__ __device__ __attribute__((always_inline))
#define WFX_CONST __device__ __constant__ static const

namespace wfx {

constexpr double kPi = 3.141592653589793;
constexpr double kTwoPi = 6.283185307179586;
constexpr double kEulerGamma = 0.5772156649015329;
constexpr double kLn2 = 0.6931471805599453;
// LAL's constants (lal 7.7; xpnr/_math.py:46-51)
constexpr double kPcSI = 3.085677581491367278913937957796471611e16;
constexpr double kMpcSI = 1.0e6 * kPcSI;
constexpr double kMrsunSI = 1.476625038050124729627979840144936351e3;
constexpr double kMtsunSI = 4.925490947641266978197229498498379006e-6;
// the spherical harmonic of the polarizations, 2 sqrt(5 / (64 pi))
// (xas/core.py:103-106)
constexpr double kYlm = 0.31539156525252005;

// ------------------------------------------------------------------ dual

// A value with N derivative directions.
template <int N>
struct Dual {
    double v;
    double d[N];
    // trivial, as a double's: an array of duals then needs no constructor
    // loop, which would keep it in scratch memory (Dual() is zero)
    Dual() = default;
    WFX_HD Dual(double x) : v(x) {
        for (int i = 0; i < N; ++i) d[i] = 0.0;
    }
    // copies member by member, not as a block: a copy of one of two duals
    // chosen at run time (c ? a : b) then reads through the chosen address,
    // which the compiler turns into a select of values (a block copy from
    // it would keep both duals in scratch memory)
    WFX_HD Dual(const Dual& o) : v(o.v) {
        for (int i = 0; i < N; ++i) d[i] = o.d[i];
    }
    WFX_HD Dual& operator=(const Dual& o) {
        v = o.v;
        for (int i = 0; i < N; ++i) d[i] = o.d[i];
        return *this;
    }
};

template <int N>
WFX_HD inline Dual<N> chain(double value, const Dual<N>& x, double slope) {
    Dual<N> r(value);
    for (int i = 0; i < N; ++i) r.d[i] = x.d[i] * slope;
    return r;
}
template <int N>
WFX_HD inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
template <int N>
WFX_HD inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
template <int N>
WFX_HD inline Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
    return r;
}
template <int N>
WFX_HD inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N>
WFX_HD inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    double q = a.v / b.v;
    Dual<N> r(q);
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - q * b.d[i]) / b.v;
    return r;
}
#define WFX_DUAL_MIXED(op)                                                                         \
    template <int N>                                                                               \
    WFX_HD inline Dual<N> operator op(const Dual<N>& a, double b) { return a op Dual<N>(b); }      \
    template <int N>                                                                               \
    WFX_HD inline Dual<N> operator op(double a, const Dual<N>& b) { return Dual<N>(a) op b; }
WFX_DUAL_MIXED(+)
WFX_DUAL_MIXED(-)
WFX_DUAL_MIXED(*)
WFX_DUAL_MIXED(/)
#undef WFX_DUAL_MIXED
#define WFX_DUAL_CMP(op)                                                                                \
    template <int N>                                                                                    \
    WFX_HD inline bool operator op(const Dual<N>& a, const Dual<N>& b) { return a.v op b.v; }           \
    template <int N>                                                                                    \
    WFX_HD inline bool operator op(const Dual<N>& a, double b) { return a.v op b; }                     \
    template <int N>                                                                                    \
    WFX_HD inline bool operator op(double a, const Dual<N>& b) { return a op b.v; }
WFX_DUAL_CMP(<)
WFX_DUAL_CMP(>)
WFX_DUAL_CMP(<=)
WFX_DUAL_CMP(>=)
WFX_DUAL_CMP(==)
WFX_DUAL_CMP(!=)
#undef WFX_DUAL_CMP

WFX_HD inline double value_of(double x) { return x; }
WFX_HD inline double value_of(float x) { return x; }
template <int N>
WFX_HD inline double value_of(const Dual<N>& x) {
    return x.v;
}

// ------------------------------------------------- elementary functions

WFX_HD inline double m_cbrt(double x) { return cbrt(x); }
WFX_HD inline double m_sqrt(double x) { return sqrt(x); }
WFX_HD inline double m_exp(double x) { return exp(x); }
WFX_HD inline double m_log(double x) { return log(x); }
WFX_HD inline double m_atan(double x) { return atan(x); }
WFX_HD inline double m_cos(double x) { return cos(x); }
WFX_HD inline double m_sin(double x) { return sin(x); }
WFX_HD inline double m_abs(double x) { return fabs(x); }
WFX_HD inline double m_pow(double x, double y) { return pow(x, y); }

WFX_HD inline float m_cbrt(float x) { return cbrtf(x); }
WFX_HD inline float m_sqrt(float x) { return sqrtf(x); }
WFX_HD inline float m_exp(float x) { return expf(x); }
WFX_HD inline float m_log(float x) { return logf(x); }
WFX_HD inline float m_atan(float x) { return atanf(x); }
WFX_HD inline float m_cos(float x) { return cosf(x); }
WFX_HD inline float m_sin(float x) { return sinf(x); }
WFX_HD inline float m_abs(float x) { return fabsf(x); }
WFX_HD inline float m_pow(float x, double y) { return powf(x, (float)y); }

template <int N>
WFX_HD inline Dual<N> m_cbrt(const Dual<N>& x) {
    double c = cbrt(x.v);
    return chain(c, x, c / (3.0 * x.v));
}
template <int N>
WFX_HD inline Dual<N> m_sqrt(const Dual<N>& x) {
    double s = sqrt(x.v);
    return chain(s, x, 0.5 / s);
}
template <int N>
WFX_HD inline Dual<N> m_exp(const Dual<N>& x) {
    double e = exp(x.v);
    return chain(e, x, e);
}
template <int N>
WFX_HD inline Dual<N> m_log(const Dual<N>& x) {
    return chain(log(x.v), x, 1.0 / x.v);
}
template <int N>
WFX_HD inline Dual<N> m_atan(const Dual<N>& x) {
    return chain(atan(x.v), x, 1.0 / (1.0 + x.v * x.v));
}
template <int N>
WFX_HD inline Dual<N> m_cos(const Dual<N>& x) {
    return chain(cos(x.v), x, -sin(x.v));
}
template <int N>
WFX_HD inline Dual<N> m_sin(const Dual<N>& x) {
    return chain(sin(x.v), x, cos(x.v));
}
// d|x| = sign(x) dx, zero at zero (JAX's convention)
template <int N>
WFX_HD inline Dual<N> m_abs(const Dual<N>& x) {
    return chain(fabs(x.v), x, x.v > 0.0 ? 1.0 : (x.v < 0.0 ? -1.0 : 0.0));
}
template <int N>
WFX_HD inline Dual<N> m_pow(const Dual<N>& x, double y) {
    return chain(pow(x.v, y), x, y * pow(x.v, y - 1.0));
}

// The phase's leading term T0 f^(-5/3) - A and the frequency powers read the
// shared-source exp, log and cbrt of wfx_libm.h in FP64 (the same bits on the
// host and the device); the FP32 setup keeps the FP32 library functions.
WFX_HD inline double lead_exp(double x) { return wfx_libm::exp(x); }
WFX_HD inline double lead_log(double x) { return wfx_libm::log(x); }
WFX_HD inline double lead_cbrt(double x) { return wfx_libm::cbrt(x); }
WFX_HD inline float lead_exp(float x) { return expf(x); }
WFX_HD inline float lead_log(float x) { return logf(x); }
WFX_HD inline float lead_cbrt(float x) { return cbrtf(x); }
template <int N>
WFX_HD inline Dual<N> lead_exp(const Dual<N>& x) {
    const double e = wfx_libm::exp(x.v);
    return chain(e, x, e);
}
template <int N>
WFX_HD inline Dual<N> lead_log(const Dual<N>& x) {
    return chain(wfx_libm::log(x.v), x, 1.0 / x.v);
}
template <int N>
WFX_HD inline Dual<N> lead_cbrt(const Dual<N>& x) {
    const double c = wfx_libm::cbrt(x.v);
    return chain(c, x, c / (3.0 * x.v));
}

template <class R>
WFX_HD inline R m_fma(const R& a, const R& b, const R& c) {
    return a * b + c;
}
// a * b + c of mixed scalars (a constant times a traced value, say): the
// same unfused operations as m_fma above, which is chosen when all three
// are of one type
template <class A, class B, class C>
WFX_HD inline auto m_fma(const A& a, const B& b, const C& c) -> decltype(a * b + c) {
    return a * b + c;
}
WFX_HD inline double m_tanh(double x) { return tanh(x); }
WFX_HD inline float m_tanh(float x) { return tanhf(x); }
template <int N>
WFX_HD inline Dual<N> m_tanh(const Dual<N>& x) {
    const double t = tanh(x.v);
    return chain(t, x, 1.0 - t * t);
}
template <int N>
WFX_HD inline Dual<N> m_atan2(const Dual<N>& y, const Dual<N>& x) {
    const double r2 = x.v * x.v + y.v * y.v;
    Dual<N> out(atan2(y.v, x.v));
    for (int i = 0; i < N; ++i) out.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / r2;
    return out;
}
// the first derivative direction of a scalar (zero for a plain number)
WFX_HD inline double tangent_of(double) { return 0.0; }
template <int N>
WFX_HD inline double tangent_of(const Dual<N>& x) {
    return x.d[0];
}
// a parameter type left out of template deduction (deduced from the others)
template <class T>
struct NoDeduce {
    using type = T;
};

template <class R>
WFX_HD inline R m_fms(const R& a, const R& b, const R& c) {
    return a * b - c;
}
template <class R>
WFX_HD inline R m_fnms(const R& a, const R& b, const R& c) {
    return c - a * b;
}
template <class R>
WFX_HD inline R m_where(bool c, const R& a, const R& b) {
    return c ? a : b;
}
// jnp.minimum: NaN if either is NaN; a tie splits the derivative
template <class R>
WFX_HD inline R m_min(const R& a, const R& b) {
    double va = value_of(a), vb = value_of(b);
    if (va != va || vb != vb) return R(va + vb);
    if (va == vb) return R(0.5) * (a + b);
    return va < vb ? a : b;
}
// sqrt for x > 0 and 0 otherwise (xpnr/_math.py:522-526)
template <class R>
WFX_HD inline R m_sqrt_nonneg(const R& x) {
    return x > R(0.0) ? m_sqrt(x) : R(0.0);
}

}  // namespace wfx
