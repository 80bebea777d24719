// The per-(binary, frequency) formula of IMRPhenomD (phenomd/core.py:
// 443-485) on one record of phenomd_layout.h: the phase of three regions,
// the leading term T0 f^(-5/3) - A and the reduction to [-pi, pi], the
// amplitude of three regions, h+ = kp A e^{i phi}, hx = i kc A e^{i phi}.
// Every frequency is evaluated (no cutoff: LAL's frequency sequence has
// none).
//
//  * point_pd_f64: FP64, the reference's order of operations;
//  * point_pd_f32<Lead>: FP32 with Horner forms, the leading term and its
//    reduction in FP64 (kLeadF64), double-float (kLeadDF) or FP32, as the
//    IMRPhenomXAS evaluation; past the inspiral the FP32 record's constants
//    hold -A modulo 2 pi (fold_anchor) and the reduction is FP32.
#pragma once

#include "phenomd_setup.h"
#include "xas_point.h"

namespace wfx {

WFX_HD inline void point_pd_f64(const double* c, const PdPowers<double>& fp, double& hp_re, double& hp_im,
                                double& hc_re, double& hc_im) {
    bool inspiral;
    const double rest = pd_phase_rest(c, fp, inspiral);
    double phi = (inspiral ? (c[kT0] + c[kT0Lo]) * fp.m53 : 0.0) + rest;
    phi = phi - (c[kAnchor] + c[kAnchorLo]);
    const double turns = rint(phi / kTwoPi);
    phi = phi - turns * kTwoPi;

    const double f = fp.f;
    double amp;
    if (f < c[kPdFAmpIns]) {
        double poly = c[kPdAmpIns + 9];
        for (int k = 8; k >= 0; --k) poly = poly * fp.t + c[kPdAmpIns + k];
        amp = poly;
    } else if (f >= c[kPdFAmpMrd]) {
        const double u = c[kPdAmpRdS] * (f - c[kPdAmpRdShift]);
        amp = c[kPdAmpRdN] * exp(-c[kPdAmpRdGr] * u) / (u * u + 1.0);
    } else {
        const double y = c[kPdAmpIntS] * f;
        double inter = c[kPdAmpInt + 4];
        for (int j = 3; j >= 0; --j) inter = inter * y + c[kPdAmpInt + j];
        amp = inter;
    }
    amp = fp.m76 * amp;
    double s, co;
    sincos(phi, &s, &co);
    const double ap = c[kPdKp] * amp;
    const double ac = c[kPdKc] * amp;
    hp_re = ap * co;
    hp_im = ap * s;
    hc_re = -ac * s;
    hc_im = ac * co;
}

// the powers of one frequency for the FP32 formula: IMRPhenomXAS's and
// f^(3/4) (the merger-ringdown phase's basis term)
struct PdPowF32 {
    PowF32 p;
    float f34;
};

__device__ inline PdPowF32 pd_pow_f32(float f) {
    PdPowF32 q;
    q.p = pow_f32(f);
    const float root = sqrtf(f);
    q.f34 = root * sqrtf(root);
    return q;
}

template <int L>
__device__ inline void point_pd_f32(const float* __restrict__ c, const PdPowF32& q, float2& hp, float2& hc) {
    const PowF32& p = q.p;
    const float f = p.f;
    const float t = p.t;
    const float m1 = p.m1;
    const bool inspiral = f < c[kPdFPhaseIns];
    const bool mrd = f >= c[kPdFPhaseMrd];
    float phi;
    if (inspiral) {
        // sum_j c_j f^(j/3), j = -3..6, by Horner in t = f^(1/3) times
        // f^-1, plus log f (c + c' f^1/3)
        float poly = c[kPdInsP2];
        poly = fmaf(poly, t, c[kPdInsP53]);
        poly = fmaf(poly, t, c[kPdInsP43]);
        poly = fmaf(poly, t, c[kPdInsP1]);
        poly = fmaf(poly, t, c[kPdInsP23]);
        poly = fmaf(poly, t, c[kPdInsP13]);
        poly = fmaf(poly, t, c[kPdInsConst]);
        poly = fmaf(poly, t, c[kPdInsM13]);
        poly = fmaf(poly, t, c[kPdInsM23]);
        poly = fmaf(poly, t, c[kPdInsM1]);
        const float lg = fmaf(c[kPdInsP13L], t, c[kPdInsLog]);
        const float rest = fmaf(poly, m1, lg * p.lf);
        if (L == kLeadF64) {
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
    } else if (mrd) {
        float rest = fmaf(c[kPdMrdM1], m1, c[kPdMrdConst]);
        rest = fmaf(c[kPdMrdP1], f, rest);
        rest = fmaf(c[kPdMrdP34], q.f34, rest);
        rest = fmaf(c[kPdAtanA], fast_atan(c[kPdAtanW] * (f - c[kPdAtanShift])), rest);
        phi = reduce_f32(rest);  // the constant holds -A (fold_anchor)
    } else {
        float rest = fmaf(c[kPdIntM3], m1 * m1 * m1, c[kPdIntConst]);
        rest = fmaf(c[kPdIntP1], f, rest);
        rest = fmaf(c[kPdIntLog], p.lf, rest);
        phi = reduce_f32(rest);
    }

    float amp;
    if (f < c[kPdFAmpIns]) {
        float poly = c[kPdAmpIns + 9];
        poly = fmaf(poly, t, c[kPdAmpIns + 8]);
        poly = fmaf(poly, t, c[kPdAmpIns + 7]);
        poly = fmaf(poly, t, c[kPdAmpIns + 6]);
        poly = fmaf(poly, t, c[kPdAmpIns + 5]);
        poly = fmaf(poly, t, c[kPdAmpIns + 4]);
        poly = fmaf(poly, t, c[kPdAmpIns + 3]);
        poly = fmaf(poly, t, c[kPdAmpIns + 2]);
        poly = fmaf(poly, t, c[kPdAmpIns + 1]);
        poly = fmaf(poly, t, c[kPdAmpIns + 0]);
        amp = p.m76 * poly;
    } else if (f >= c[kPdFAmpMrd]) {
        const float u = c[kPdAmpRdS] * (f - c[kPdAmpRdShift]);
        amp = fast_div(c[kPdAmpRdN] * p.m76 * __expf(-c[kPdAmpRdGr] * u), fmaf(u, u, 1.0f));
    } else {
        const float y = c[kPdAmpIntS] * f;
        float inter = c[kPdAmpInt + 4];
        inter = fmaf(inter, y, c[kPdAmpInt + 3]);
        inter = fmaf(inter, y, c[kPdAmpInt + 2]);
        inter = fmaf(inter, y, c[kPdAmpInt + 1]);
        inter = fmaf(inter, y, c[kPdAmpInt + 0]);
        amp = p.m76 * inter;
    }
    float s, co;
    __sincosf(phi, &s, &co);
    const float ap = c[kPdKp] * amp;
    const float ac = c[kPdKc] * amp;
    hp = make_float2(ap * co, ap * s);
    hc = make_float2(-ac * s, ac * co);
}

}  // namespace wfx
