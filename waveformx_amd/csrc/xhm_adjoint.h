// The gradient of IMRPhenomXHM per (binary, frequency), split arithmetic:
// the five modes and their derivatives in the four traced parameters (m1,
// m2, chi1z, chi2z), forward mode through the split evaluation of
// xhm_point.h (the same FP32 operations in the same order for the values,
// so that a fused loss sees the forward pass's bits), from the records and
// their derivatives of the dual-number set-up (wfx_xhm_setup_jvp,
// wfx_xas_setup_jvp). The four parameters that enter in closed form (d_L,
// the inclination, phi_ref, f_ref) are projections the kernel also takes.
//
// The reference's GPU port differentiates the same way in part (the (2,2)
// and the factors in forward mode, src/waveformx/models/xhm/mlx/_kernels.py:
// 300-441); here every mode is in forward mode with four directions, which
// needs no per-slot accumulators.
#pragma once

#include "xhm_point.h"

namespace wfx {

// ------------------------------------------------------ tangent scalar

// an FP32 value with its derivatives in four directions
struct DF {
    float v;
    float d[4];
};
__device__ inline DF df_const(float x) { return DF{x, {0.0f, 0.0f, 0.0f, 0.0f}}; }
__device__ inline DF operator+(const DF& a, const DF& b) {
    return DF{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2], a.d[3] + b.d[3]}};
}
__device__ inline DF operator-(const DF& a, const DF& b) {
    return DF{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2], a.d[3] - b.d[3]}};
}
__device__ inline DF operator-(const DF& a) { return DF{-a.v, {-a.d[0], -a.d[1], -a.d[2], -a.d[3]}}; }
__device__ inline DF operator*(const DF& a, const DF& b) {
    DF r;
    r.v = a.v * b.v;
    for (int i = 0; i < 4; ++i) r.d[i] = fmaf(a.d[i], b.v, a.v * b.d[i]);
    return r;
}
__device__ inline DF operator*(const DF& a, float b) { return DF{a.v * b, {a.d[0] * b, a.d[1] * b, a.d[2] * b, a.d[3] * b}}; }
__device__ inline DF operator*(float a, const DF& b) { return b * a; }
__device__ inline DF operator+(const DF& a, float b) { return DF{a.v + b, {a.d[0], a.d[1], a.d[2], a.d[3]}}; }
__device__ inline DF operator-(const DF& a, float b) { return DF{a.v - b, {a.d[0], a.d[1], a.d[2], a.d[3]}}; }
__device__ inline DF operator-(float a, const DF& b) { return DF{a - b.v, {-b.d[0], -b.d[1], -b.d[2], -b.d[3]}}; }
// fmaf(a, b, c) on the values, the product rule on the derivatives
__device__ inline DF dfma(const DF& a, const DF& b, const DF& c) {
    DF r;
    r.v = fmaf(a.v, b.v, c.v);
    for (int i = 0; i < 4; ++i) r.d[i] = fmaf(a.d[i], b.v, fmaf(a.v, b.d[i], c.d[i]));
    return r;
}
__device__ inline DF dfma(const DF& a, float b, const DF& c) {
    DF r;
    r.v = fmaf(a.v, b, c.v);
    for (int i = 0; i < 4; ++i) r.d[i] = fmaf(a.d[i], b, c.d[i]);
    return r;
}
// the value's function f and its derivative f1 on the derivatives
__device__ inline DF df_chain(float f, const DF& x, float f1) {
    return DF{f, {x.d[0] * f1, x.d[1] * f1, x.d[2] * f1, x.d[3] * f1}};
}
// a / b as fast_div (the hardware reciprocal)
__device__ inline DF ddiv(const DF& a, const DF& b) {
    const float r = __builtin_amdgcn_rcpf(b.v);
    const float q = a.v * r;
    DF out;
    out.v = q;
    for (int i = 0; i < 4; ++i) out.d[i] = (a.d[i] - q * b.d[i]) * r;
    return out;
}
__device__ inline DF dexp(const DF& x) {
    const float e = __expf(x.v);
    return df_chain(e, x, e);
}
__device__ inline DF dsqrt(const DF& x) {
    const float s = __builtin_sqrtf(x.v);
    return df_chain(s, x, s > 0.0f ? 0.5f * __builtin_amdgcn_rcpf(s) : 0.0f);
}
__device__ inline DF datan(const DF& x) {
    return df_chain(fast_atan(x.v), x, __builtin_amdgcn_rcpf(fmaf(x.v, x.v, 1.0f)));
}
// the derivative (x dy - y dx) / (x^2 + y^2) scaled by max(|x|, |y|) first,
// so that a mode whose amplitude is near FP32's smallest normal (far in
// its ringdown) keeps a finite phase derivative; at the origin (the
// amplitude zero: its phase moves nothing) it is zero
__device__ inline DF datan2(const DF& y, const DF& x) {
    const float s = fmaxf(fabsf(x.v), fabsf(y.v));
    const float inv_s = s > 0.0f ? __builtin_amdgcn_rcpf(s) : 0.0f;
    const float xs = x.v * inv_s, ys = y.v * inv_s;
    const float r = s > 0.0f ? __builtin_amdgcn_rcpf(s * fmaf(xs, xs, ys * ys)) : 0.0f;
    DF out;
    out.v = atan2f(y.v, x.v);
    for (int i = 0; i < 4; ++i) out.d[i] = (xs * y.d[i] - ys * x.d[i]) * r;
    return out;
}
// (sin, cos) of an angle whose value is reduced first (the reduction is a
// constant shift: no derivative)
__device__ inline void dsincos(const DF& x, float reduced, DF& sn, DF& cs) {
    float s, c;
    __sincosf(reduced, &s, &c);
    sn = df_chain(s, x, c);
    cs = df_chain(c, x, -s);
}

// slot s of a record r with its derivatives t[4 s + k], k = 0..3 (a slot's
// four derivatives are contiguous: one 16-byte load)
__device__ inline DF ld(const float* __restrict__ r, const float* __restrict__ t, int s) {
    const float4 d = *reinterpret_cast<const float4*>(t + 4 * s);
    return DF{r[s], {d.x, d.y, d.z, d.w}};
}

// ----------------------------------------------------------- the (2,2)

// amp_phase_f32<kLeadF64> with derivatives: c the (2,2)'s split record, t
// its slots' derivatives (kSlots x 4: the post-inspiral constants' include
// -A's, T0's and A's FP64 copies' sit in the first word)
__device__ inline void amp_phase22_t(const float* __restrict__ c, const float* __restrict__ t, const PowF32& p,
                                     DF& amp, DF& phi) {
    auto C = [&](int s) { return ld(c, t, s); };
    const float f = p.f;
    const float tt = p.t;
    const float m1 = p.m1;
    const bool past = f >= c[kFPhaseIn];
    const bool ringdown = f > c[kFPhaseIm];
    DF rest;
    if (!past) {
        DF poly = C(kInsP2);
        poly = dfma(poly, tt, C(kInsP53));
        poly = dfma(poly, tt, C(kInsP43));
        poly = dfma(poly, tt, C(kInsP1));
        poly = dfma(poly, tt, C(kInsP23));
        poly = dfma(poly, tt, C(kInsP13));
        poly = dfma(poly, tt, C(kInsConst));
        poly = dfma(poly, tt, C(kInsM13));
        poly = dfma(poly, tt, C(kInsM23));
        poly = dfma(poly, tt, C(kInsM1));
        const DF lg = dfma(dfma(C(kInsP1L), tt * tt, C(kInsP13L)), tt, C(kInsLog));
        rest = dfma(poly, m1, lg * p.lf);
    } else if (!ringdown) {
        const DF inv = dfma(dfma(C(kIntM3), m1, C(kIntM2)), m1, C(kIntM1));
        rest = dfma(inv, m1, C(kIntConst));
        rest = dfma(C(kIntP1), f, rest);
        rest = dfma(C(kIntLog), p.lf, rest);
        rest = dfma(C(kAtanA1), datan(C(kAtanW1) * (df_const(f) - C(kShift))), rest);
    } else {
        const DF inv = dfma(C(kRdM3), m1 * m1, C(kRdM1));
        rest = dfma(inv, m1, C(kRdConst));
        rest = dfma(C(kRdP1), f, rest);
        rest = dfma(C(kRdP23), tt * tt, rest);
        rest = dfma(C(kAtanA2), datan(C(kAtanW2) * (df_const(f) - C(kShift))), rest);
    }
    if (past) {
        phi = rest;
        phi.v = reduce_f32(rest.v);  // the constants hold -A (fold_anchor)
    } else {
        const double full = fma(record_f64(c, kT0F64), p.m53, (double)rest.v - record_f64(c, kAnchorF64));
        const double turns = rint(full * kInvTwoPi);
        phi.v = (float)fma(-turns, kTwoPi, full);
        const float m53 = (float)p.m53;
        for (int i = 0; i < 4; ++i)
            phi.d[i] = fmaf(t[4 * kT0F64 + i], m53, rest.d[i] - t[4 * kAnchorF64 + i]);
    }
    if (f > c[kFAmpRd]) {
        const DF u = C(kAmpRdS) * (df_const(f) - C(kShift));
        amp = ddiv(C(kAmpRdN) * p.m76 * dexp(-C(kAmpRdGr) * u), dfma(u, u, C(kAmpRdG)));
    } else if (f >= c[kFAmpIn]) {
        DF den = C(kAmpIntD + 4);
        den = dfma(den, f, C(kAmpIntD + 3));
        den = dfma(den, f, C(kAmpIntD + 2));
        den = dfma(den, f, C(kAmpIntD + 1));
        den = dfma(den, f, C(kAmpIntD + 0));
        amp = ddiv(C(kAmpIntN), den);
    } else {
        DF poly = C(kAmpIns + 9);
        for (int k = 8; k >= 0; --k) poly = dfma(poly, tt, C(kAmpIns + k));
        amp = p.m76 * poly;
    }
    if (!(f <= c[kFCut])) amp = df_const(0.0f);
}

// --------------------------------------------------------- higher modes

// the powers of M f from those of f and the binary's M (their derivatives
// through M's slots)
struct MfPowDF {
    DF mf, t, m1, lg, m76;
};
__device__ inline MfPowDF mf_pow_t(const float* __restrict__ r, const float* __restrict__ t, const PowF32& p) {
    MfPowDF q;
    q.mf = ld(r, t, kHmMsec) * p.f;
    q.t = ld(r, t, kHmCbrtM) * p.t;
    q.m1 = ld(r, t, kHmInvM) * p.m1;
    q.lg = ld(r, t, kHmLogM) + p.lf;
    // m1 t^(-1/2), value as mf_pow_f32's
    const float rs = __builtin_amdgcn_rsqf(q.t.v);
    q.m76.v = q.m1.v * rs;
    const float drs = -0.5f * rs * rs * rs;
    for (int i = 0; i < 4; ++i) q.m76.d[i] = fmaf(q.m1.d[i], rs, q.m1.v * drs * q.t.d[i]);
    return q;
}

// hm_amp_phase_split with derivatives (the same operations on the values)
__device__ inline void hm_amp_phase_split_t(const float* __restrict__ r, const float* __restrict__ t,
                                            const MfPowDF& q, double m53, const DF& w_re, const DF& w_im, DF& a,
                                            DF& phi) {
    auto R = [&](int s) { return ld(r, t, s); };
    const bool mixing = r[kHmMixing] != 0.0f;
    const DF mf = q.mf, tt = q.t, m1 = q.m1;
    const DF f_ring = R(kHmFRing), f_damp = R(kHmFDamp);
    DF rot_re = df_const(0.0f), rot_im = df_const(0.0f);
    if (mixing && (mf.v >= r[kHmFAmpIm] || mf.v >= r[kHmFPhaseIm])) {
        DF amplm;
        if (mf.v < r[kHmFRdAux]) {
            amplm = dfma(dfma(dfma(R(kHmRdAux + 3), mf, R(kHmRdAux + 2)), mf, R(kHmRdAux + 1)), mf, R(kHmRdAux));
        } else if (mf.v >= r[kHmFalloff]) {
            amplm = R(kHmRd + 3) * dexp(-R(kHmRd + 4) * (mf - R(kHmFalloffAt)));
        } else {
            const DF dfr = mf - f_ring;
            amplm = ddiv(R(kHmRd), dexp(R(kHmRd + 1) * dfr) * dfma(dfr, dfr, R(kHmRd + 2)));
        }
        DF philm = dfma(R(kHmSph), mf, R(kHmSph + 4));
        philm = dfma(-R(kHmSph + 2), m1, philm);
        philm = dfma(-((1.0f / 3.0f) * R(kHmSph + 3)), m1 * m1 * m1, philm);
        philm = dfma(R(kHmSph + 1), datan(ddiv(mf - f_ring, f_damp)), philm);
        DF sn, cs;
        dsincos(philm, reduce_f32(philm.v), sn, cs);
        const DF m2r = R(kHmMix), m2i = R(kHmMix + 1), m3r = R(kHmMix + 2), m3i = R(kHmMix + 3);
        const DF b_re = amplm * m3r, b_im = -(m3i * amplm);
        rot_re = dfma(w_re, m2r, w_im * m2i) + dfma(b_re, cs, -b_im * sn);
        rot_im = dfma(w_im, m2r, -w_re * m2i) + dfma(b_re, sn, b_im * cs);
    }
    if (mf.v < r[kHmFAmpIn]) {
        DF re = R(kHmPnRe + 6), im = R(kHmPnIm + 6);
        for (int k = 5; k >= 0; --k) {
            re = dfma(re, tt, R(kHmPnRe + k));
            im = dfma(im, tt, R(kHmPnIm + k));
        }
        const DF t2 = tt * tt, t7 = t2 * t2 * t2 * tt;
        const DF pseudo = t7 * dfma(dfma(R(kHmIns + 2), tt, R(kHmIns + 1)), tt, R(kHmIns));
        a = q.m76 * dfma(dsqrt(dfma(re, re, im * im)), R(kHmPnGlobal) * R(kHmAmpNorm), pseudo * R(kHmPnDominant));
    } else if (mf.v >= r[kHmFAmpIm]) {
        if (mixing) {
            a = dsqrt(dfma(rot_re, rot_re, rot_im * rot_im));
        } else if (mf.v >= r[kHmFalloff]) {
            a = R(kHmRd + 3) * dexp(-R(kHmRd + 4) * (mf - R(kHmFalloffAt)));
        } else {
            const DF dfr = mf - f_ring;
            a = ddiv(R(kHmRd), dexp(R(kHmRd + 1) * dfr) * dfma(dfr, dfr, R(kHmRd + 2)));
        }
    } else {
        const DF u = (mf - R(kHmInterC)) * R(kHmInterH);
        const int n = (int)r[kHmNInter];
        DF result = R(kHmInter + n - 1);
        for (int j = n - 2; j >= 0; --j) result = dfma(result, u, R(kHmInter + j));
        a = result * q.m76;
    }
    if (a.v < 0.0f) a = df_const(1.0e-15f);
    if (mf.v < r[kHmFPhaseIn]) {
        DF poly = R(kHmPhi + 12);
        for (int i = 11; i >= 1; --i) poly = dfma(poly, tt, R(kHmPhi + i));
        const DF m43 = m1 * ddiv(df_const(1.0f), tt);
        DF rest = poly * m43;
        rest = dfma(q.lg, dfma(dfma(R(kHmPhiL + 8), tt * tt, df_const(0.0f)), tt, dfma(R(kHmPhiL + 6), tt, R(kHmPhiL + 5))),
                    rest);
        rest = dfma(R(kHmC1Insp), mf, rest);
        const double full = fma(record_f64(r, kHmT0F64), m53, record_f64(r, kHmKInsF64));
        const double turns = rint(full * kInvTwoPi);
        phi.v = (float)fma(-turns, kTwoPi, full) + rest.v;
        const float m53f = (float)m53;
        for (int i = 0; i < 4; ++i)
            phi.d[i] = fmaf(t[4 * kHmT0F64 + i], m53f, t[4 * kHmKInsF64 + i] + rest.d[i]);
    } else if (mf.v >= r[kHmFPhaseIm]) {
        DF rest;
        if (mixing) {
            rest = dfma(R(kHmC1Rd), mf, datan2(rot_im, rot_re));
        } else {
            rest = dfma(R(kHmC1Rd) + R(kHmAlpha0), mf, -(f_ring * f_ring * R(kHmAlpha2)) * m1);
            rest = dfma(R(kHmAlphaL), datan(ddiv(mf - f_ring, f_damp)), rest);
        }
        phi = rest + R(kHmCRd);
    } else {
        DF rest = dfma(R(kHmC0), mf, R(kHmC1) * q.lg);
        rest = dfma(-R(kHmC2), m1, rest);
        const DF m12 = m1 * m1;
        rest = dfma(-((1.0f / 3.0f) * R(kHmC4)), m12 * m1, rest);
        if (mixing) rest = dfma(-(0.5f * R(kHmC3)), m12, rest);
        rest = dfma(R(kHmCL), datan(ddiv(mf - f_ring, f_damp)), rest);
        phi = rest + R(kHmDeltaPhi);
    }
    phi.v = reduce_f32(phi.v);
}

// Mode k's contribution to the gradient's projections v (the slots of
// xhm_adjoint_split_kernel) from its value and derivatives h (re, im), its
// factors f = (F+, Fx) and their inclination derivatives df, the cotangents
// (cp, cc) and m its azimuthal number
__device__ inline void accumulate_mode(const DF& h_re, const DF& h_im, const float* f, const float* df, float emm,
                                       float2 cp, float2 cc, double* v) {
    const float g_re = f[0] * cp.x + f[1] * cp.y + f[2] * cc.x + f[3] * cc.y;
    const float g_im = f[0] * cp.y - f[1] * cp.x + f[2] * cc.y - f[3] * cc.x;
    for (int d = 0; d < 4; ++d) v[d] += (double)(g_re * h_re.d[d] + g_im * h_im.d[d]);
    const float re = h_re.v, im = h_im.v;
    v[4] += (double)(g_re * re + g_im * im);
    v[6] += (double)(emm * (g_im * re - g_re * im));
    const float dp_re = df[0] * re - df[1] * im, dp_im = df[0] * im + df[1] * re;
    const float dc_re = df[2] * re - df[3] * im, dc_im = df[2] * im + df[3] * re;
    v[5] += (double)(cp.x * dp_re + cp.y * dp_im + cc.x * dc_re + cc.y * dc_im);
}

// one bin's five modes, in forward mode one at a time, accumulated into v
// as they come (xhm_modes_split's values): rec22, t22 the (2,2)'s split
// record and derivatives (kSlots x 4), modes, tmodes the higher modes'
// (4 x kHmSlots, 4 x kHmSlots x 4), fac, dfac the head and its derivative
__device__ inline void xhm_bin_gradient(const float* __restrict__ rec22, const float* __restrict__ t22,
                                        const float* __restrict__ modes, const float* __restrict__ tmodes,
                                        const float* __restrict__ fac, const float* __restrict__ dfac, float f,
                                        float2 cp, float2 cc, double* v) {
    const PowF32 p = pow_f32(f);
    DF a22, phi22;
    amp_phase22_t(rec22, t22, p, a22, phi22);
    DF s22, c22;
    dsincos(phi22, phi22.v, s22, c22);
    const DF h0_re = a22 * c22, h0_im = a22 * s22;
    accumulate_mode(h0_re, h0_im, fac + kHeadFactors, dfac + kHeadFactors, 2.0f, cp, cc, v);
    const MfPowDF q = mf_pow_t(modes, tmodes, p);
#pragma unroll 1
    for (int k = 0; k < kHmModes; ++k) {
        const float* r = modes + k * kHmSlots;
        const float* t = tmodes + 4 * k * kHmSlots;
        if (r[kHmOn] == 0.0f) continue;
        const DF amp0 = ld(r, t, kHmAmp0);
        const float inv = __builtin_amdgcn_rcpf(fabsf(amp0.v));
        const DF inv_amp0 = df_chain(inv, amp0, -(amp0.v < 0.0f ? -1.0f : 1.0f) * inv * inv);
        DF a, phi;
        hm_amp_phase_split_t(r, t, q, p.m53, h0_re * inv_amp0, h0_im * inv_amp0, a, phi);
        const DF val = amp0 * a;
        DF sn, cs;
        dsincos(phi, phi.v, sn, cs);
        accumulate_mode(val * cs, val * sn, fac + kHeadFactors + 4 * (k + 1), dfac + kHeadFactors + 4 * (k + 1),
                        (float)kHmEmm[k], cp, cc, v);
    }
}

}  // namespace wfx
