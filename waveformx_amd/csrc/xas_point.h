// The per-(binary, frequency) formula of IMRPhenomXAS (xas/core.py:423-507)
// on one record of xas_layout.h: the phase (regions, leading term,
// reduction to [-pi, pi]), the amplitude and the polarizations
// h+ = kp A e^{i phi}, hx = i kc A e^{i phi}.
//
// Two arithmetics:
//  * point_f64: FP64, the reference's order of operations, with the
//    record's FP64 powers of f -- the parity path;
//  * point_f32<Lead>: FP32 with Horner forms of the phase sums and the
//    inspiral's leading term T0 f^(-5/3) - A and its reduction in FP64
//    (kLeadF64), in double-float (kLeadDF), or plain FP32 (kLeadF32).
#pragma once

#include "xas_setup.h"

namespace wfx {

// ---------------------------------------------------------- FP64 (parity)

// the amplitude and the phase (reduced to [-pi, pi]) of the (2,2) mode
WFX_HD inline void amp_phase_f64(const double* c, const FreqPowers<double>& fp, double& amp, double& phi) {
    bool past;
    const double rest = phase_rest(c, fp, past);
    phi = (past ? 0.0 : (c[kT0] + c[kT0Lo]) * fp.m53) + rest;
    phi = phi - (c[kAnchor] + c[kAnchorLo]);
    const double turns = rint(phi / kTwoPi);
    phi = phi - turns * kTwoPi;

    const double f = fp.f;
    if (f > c[kFAmpRd]) {
        const double u = c[kAmpRdS] * (f - c[kShift]);
        amp = c[kAmpRdN] * fp.m76 * exp(-c[kAmpRdGr] * u) / (u * u + c[kAmpRdG]);
    } else if (f >= c[kFAmpIn]) {
        double den = c[kAmpIntD + 4];
        for (int j = 3; j >= 0; --j) den = den * f + c[kAmpIntD + j];
        amp = c[kAmpIntN] / den;
    } else {
        double poly = c[kAmpIns + 9];
        for (int k = 8; k >= 0; --k) poly = poly * fp.t + c[kAmpIns + k];
        amp = fp.m76 * poly;
    }
    amp = f <= c[kFCut] ? amp : 0.0;
}

WFX_HD inline void point_f64(const double* c, const FreqPowers<double>& fp, double& hp_re, double& hp_im,
                             double& hc_re, double& hc_im) {
    double amp, phi;
    amp_phase_f64(c, fp, amp, phi);
    double s, co;
    sincos(phi, &s, &co);
    const double ap = c[kKp] * amp;
    const double ac = c[kKc] * amp;
    hp_re = ap * co;
    hp_im = ap * s;
    hc_re = -ac * s;
    hc_im = ac * co;
}

// ---------------------------------------------------------- FP32 paths

enum Lead : int { kLeadF64 = 0, kLeadDF = 1, kLeadF32 = 2 };

// the powers of one frequency for the FP32 formula; f^(-5/3) in FP64 (and
// as a double-float pair)
struct PowF32 {
    float f, t, m1, lf, m76;
    double m53;
    float m53_hi, m53_lo;
};

// arctangent in FP32: atan(x) = x P(x^2) on [0, 1] (P of degree 7 in x^2,
// the near-minimax polynomial that gen_atan.py fits by Lawson's iteratively
// reweighted least squares and rounds to float32; |error| < 7e-8 in exact
// arithmetic, < 2e-7 in FP32 over [0, inf)), and pi/2 - atan(1/x) beyond
__device__ inline float fast_atan(float x) {
    const float a = fabsf(x);
    const bool big = a > 1.0f;
    const float r = big ? __builtin_amdgcn_rcpf(a) : a;
    const float r2 = r * r;
    float p = -0.004054555203765631f;
    p = fmaf(p, r2, 0.02186291664838791f);
    p = fmaf(p, r2, -0.05591227486729622f);
    p = fmaf(p, r2, 0.09642194211483002f);
    p = fmaf(p, r2, -0.1390862911939621f);
    p = fmaf(p, r2, 0.19946566224098206f);
    p = fmaf(p, r2, -0.33329859375953674f);
    p = fmaf(p, r2, 0.9999993443489075f);
    p *= r;
    p = big ? 1.5707963267948966f - p : p;
    return copysignf(p, x);
}

// a / b through the hardware reciprocal (1 ulp)
__device__ inline float fast_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// FP32 powers of a frequency given in FP32; f^(-1/3) from 1/t is refined
// by two Newton steps in FP64 before f^(-5/3) = (f^(-1/3))^5
//
// r = f^(-1/3) in FP64 by two Newton steps r <- r (4 - f r^3) / 3 from the
// hardware exp2(-log2(f) / 3) (relative error ~1e-7, then ~1e-14, then
// below FP64 rounding); f^(1/3) = f r^2 and 1/f = r^3 are rounded from it to
// FP32 (no separate cube root or division), f^(-5/3) = r^5 stays FP64 for
// the leading phase term.
__device__ inline PowF32 pow_f32(float f) {
    PowF32 p;
    p.f = f;
    p.lf = logf(f);
    const double fd = (double)f;
    double r = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf(f) * (-1.0f / 3.0f));
    r = r * (4.0 - fd * (r * r * r)) * (1.0 / 3.0);
    r = r * (4.0 - fd * (r * r * r)) * (1.0 / 3.0);
    const double r2 = r * r;
    p.t = (float)(fd * r2);
    p.m1 = (float)(r2 * r);
    p.m76 = p.m1 * __builtin_amdgcn_rsqf(p.t);
    p.m53 = r2 * r2 * r;
    p.m53_hi = (float)p.m53;
    p.m53_lo = (float)(p.m53 - (double)p.m53_hi);
    return p;
}

// double-float helpers (exact transformations: Knuth's TwoSum, Dekker's
// TwoProduct through the fused multiply-add)
struct F2 {
    float hi, lo;
};
__device__ inline F2 two_sum(float a, float b) {
    const float s = a + b;
    const float bb = s - a;
    return F2{s, (a - (s - bb)) + (b - bb)};
}
__device__ inline F2 fast_two_sum(float a, float b) {
    const float s = a + b;
    return F2{s, b - (s - a)};
}
__device__ inline F2 f2_mul(F2 a, F2 b) {
    const float p = a.hi * b.hi;
    const float e = fmaf(a.hi, b.hi, -p);
    return fast_two_sum(p, fmaf(a.hi, b.lo, fmaf(a.lo, b.hi, e)));
}
__device__ inline F2 f2_add(F2 a, F2 b) {
    const F2 s = two_sum(a.hi, b.hi);
    return fast_two_sum(s.hi, s.lo + (a.lo + b.lo));
}

constexpr float kInvTwoPiF = 0.15915494309189535f;
constexpr float kTwoPiHiF = 6.28318548202514648f;     // float(2 pi)
constexpr float kTwoPiLoF = -1.7484555314695172e-07f;  // 2 pi - float(2 pi)
constexpr double kInvTwoPi = 0.15915494309189535;

// T0 and A of a FP32 record in FP64 (slots 60..63 hold them as doubles,
// low word first); assembled from the two words so that a record copied
// to registers stays in registers
__device__ inline double record_f64(const float* c, int slot) {
    return __hiloint2double(__float_as_int(c[slot + 1]), __float_as_int(c[slot]));
}
constexpr int kT0F64 = 60;
constexpr int kAnchorF64 = 62;

// The FP32 records' post-inspiral constants (kIntConst, kRdConst) hold
// their FP64 value minus the anchor A, reduced modulo 2 pi: past the
// inspiral the phase is then the region's sum alone, of modest size,
// reduced in FP32; the inspiral keeps T0 f^(-5/3) - A in the lead's
// arithmetic. The derivative of such a constant is that of the constant
// minus that of A (the reduction is a constant shift).
WFX_HD inline double fold_anchor(double constant, double anchor) {
    const double v = constant - anchor;
    return v - kTwoPi * rint(v * (1.0 / kTwoPi));
}

// x reduced to [-pi, pi] in FP32 (Cody-Waite with a two-part 2 pi)
__device__ inline float reduce_f32(float x) {
    const float turns = rintf(x * kInvTwoPiF);
    return fmaf(-turns, kTwoPiLoF, fmaf(-turns, kTwoPiHiF, x));
}

// One template body for both arithmetics: the FP32 kernels' fast forms
// (hardware reciprocal, exponential, sine and cosine on |x| <= pi, the
// fitted arctangent) and their FP64 library counterparts.
__device__ inline void m_sincos(float x, float* s, float* c) { __sincosf(x, s, c); }
__device__ inline void m_sincos(double x, double* s, double* c) { sincos(x, s, c); }
__device__ inline float d_atan(float x) { return fast_atan(x); }
__device__ inline double d_atan(double x) { return atan(x); }
__device__ inline float d_atan2(float y, float x) { return atan2f(y, x); }
__device__ inline double d_atan2(double y, double x) { return atan2(y, x); }
__device__ inline float d_exp(float x) { return __expf(x); }
__device__ inline double d_exp(double x) { return exp(x); }
__device__ inline float d_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline double d_rcp(double x) { return 1.0 / x; }
__device__ inline float d_div(float a, float b) { return fast_div(a, b); }
__device__ inline double d_div(double a, double b) { return a / b; }
__device__ inline float d_sqrt(float x) { return __builtin_sqrtf(x); }
__device__ inline double d_sqrt(double x) { return sqrt(x); }
__device__ inline float d_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double d_rsq(double x) { return 1.0 / sqrt(x); }
__device__ inline float d_fma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ inline double d_fma(double a, double b, double c) { return fma(a, b, c); }
// the smallest normal number: below it the FP32 forms' reciprocals flush
__device__ inline float tiny_normal(float) { return 1.17549435e-38f; }
__device__ inline double tiny_normal(double) { return 2.2250738585072014e-308; }
// x reduced to [-pi, pi]: FP32 as reduce_f32, FP64 with 2 pi in two parts
// (its low part 2 pi - fl(2 pi))
__device__ inline float d_reduce(float x) { return reduce_f32(x); }
__device__ inline double d_reduce(double x) {
    const double turns = rint(x * kInvTwoPi);
    return fma(-turns, 2.4492935982947064e-16, fma(-turns, kTwoPi, x));
}
// a record's FP64 slot: two floats (low word first) in a FP32 record, one
// double in a FP64 record
__device__ inline double wide_slot(const float* r, int slot) { return record_f64(r, slot); }
__device__ inline double wide_slot(const double* r, int slot) { return r[slot]; }

// the amplitude and the phase (reduced to [-pi, pi]) of the (2,2) mode in FP32
template <int L>
__device__ inline void amp_phase_f32(const float* __restrict__ c, const PowF32& p, float& amp, float& phi) {
    const float f = p.f;
    const float t = p.t;
    const float m1 = p.m1;
    const bool past = f >= c[kFPhaseIn];
    const bool ringdown = f > c[kFPhaseIm];
    float rest;
    if (!past) {
        // sum_k c_k f^((k-3)/3) + log f (c + c' f^1/3 + c'' f)
        float poly = c[kInsP2];
        poly = fmaf(poly, t, c[kInsP53]);
        poly = fmaf(poly, t, c[kInsP43]);
        poly = fmaf(poly, t, c[kInsP1]);
        poly = fmaf(poly, t, c[kInsP23]);
        poly = fmaf(poly, t, c[kInsP13]);
        poly = fmaf(poly, t, c[kInsConst]);
        poly = fmaf(poly, t, c[kInsM13]);
        poly = fmaf(poly, t, c[kInsM23]);
        poly = fmaf(poly, t, c[kInsM1]);
        const float lg = fmaf(fmaf(c[kInsP1L], t * t, c[kInsP13L]), t, c[kInsLog]);
        rest = fmaf(poly, m1, lg * p.lf);
    } else if (!ringdown) {
        const float inv = fmaf(fmaf(c[kIntM3], m1, c[kIntM2]), m1, c[kIntM1]);
        rest = fmaf(inv, m1, c[kIntConst]);
        rest = fmaf(c[kIntP1], f, rest);
        rest = fmaf(c[kIntLog], p.lf, rest);
        rest = fmaf(c[kAtanA1], fast_atan(c[kAtanW1] * (f - c[kShift])), rest);
    } else {
        const float inv = fmaf(c[kRdM3], m1 * m1, c[kRdM1]);
        rest = fmaf(inv, m1, c[kRdConst]);
        rest = fmaf(c[kRdP1], f, rest);
        rest = fmaf(c[kRdP23], t * t, rest);
        rest = fmaf(c[kAtanA2], fast_atan(c[kAtanW2] * (f - c[kShift])), rest);
    }

    if (past) {
        phi = reduce_f32(rest);  // the constants hold -A (fold_anchor)
    } else if (L == kLeadF64) {
        const double full = fma(record_f64(c, kT0F64), p.m53, (double)rest - record_f64(c, kAnchorF64));
        const double turns = rint(full * kInvTwoPi);
        phi = (float)fma(-turns, kTwoPi, full);
    } else if (L == kLeadDF) {
        const F2 lead = f2_mul(F2{c[kT0], c[kT0Lo]}, F2{p.m53_hi, p.m53_lo});
        F2 s = f2_add(lead, F2{-c[kAnchor], -c[kAnchorLo]});
        s = f2_add(s, F2{rest, 0.0f});
        phi = reduce_f32(s.hi) + s.lo;
    } else {
        phi = reduce_f32(fmaf(c[kT0], p.m53_hi, rest - c[kAnchor]));
    }

    if (f > c[kFAmpRd]) {
        const float u = c[kAmpRdS] * (f - c[kShift]);
        amp = fast_div(c[kAmpRdN] * p.m76 * __expf(-c[kAmpRdGr] * u), fmaf(u, u, c[kAmpRdG]));
    } else if (f >= c[kFAmpIn]) {
        float den = c[kAmpIntD + 4];
        den = fmaf(den, f, c[kAmpIntD + 3]);
        den = fmaf(den, f, c[kAmpIntD + 2]);
        den = fmaf(den, f, c[kAmpIntD + 1]);
        den = fmaf(den, f, c[kAmpIntD + 0]);
        amp = fast_div(c[kAmpIntN], den);
    } else {
        float poly = c[kAmpIns + 9];
        poly = fmaf(poly, t, c[kAmpIns + 8]);
        poly = fmaf(poly, t, c[kAmpIns + 7]);
        poly = fmaf(poly, t, c[kAmpIns + 6]);
        poly = fmaf(poly, t, c[kAmpIns + 5]);
        poly = fmaf(poly, t, c[kAmpIns + 4]);
        poly = fmaf(poly, t, c[kAmpIns + 3]);
        poly = fmaf(poly, t, c[kAmpIns + 2]);
        poly = fmaf(poly, t, c[kAmpIns + 1]);
        poly = fmaf(poly, t, c[kAmpIns + 0]);
        amp = p.m76 * poly;
    }
    amp = f <= c[kFCut] ? amp : 0.0f;
}

template <int L>
__device__ inline void point_f32(const float* __restrict__ c, const PowF32& p, float2& hp, float2& hc) {
    float amp, phi;
    amp_phase_f32<L>(c, p, amp, phi);
    float s, co;
    __sincosf(phi, &s, &co);
    const float ap = c[kKp] * amp;
    const float ac = c[kKc] * amp;
    hp = make_float2(ap * co, ap * s);
    hc = make_float2(-ac * s, ac * co);
}

}  // namespace wfx
