// The adjoint of IMRPhenomXHM's split evaluation (xhm_point.h) with respect
// to a higher mode's record: given the cotangent g = dL/dRe h + i dL/dIm h
// of the mode h_{l,-m} at one bin, add dL/d(record slot) of that bin to the
// mode's accumulators (HmAcc: the slots that carry a derivative, numbered
// densely so that a wave keeps them in registers over its bins). The
// forward values are recomputed with the forward kernel's operations; the
// adjoint follows each region's formula back to its coefficients and to the
// powers of M f (through the binary's M slots), the (3,2)'s mixing back to
// the (2,2) mode it recycles (returned as that mode's cotangent). The
// region bounds carry no derivative, as the reference's jnp.where
// selections (xhm/reference.py:620-643).
//
// The reference differentiates this evaluation with jax's reverse mode
// through its kernels (src/waveformx/torch.py:108-119; the MLX port's
// hand-written cotangent_product, src/waveformx/models/xhm/mlx/_kernels.py:
// 300-441); here each mode's bins are reduced into its slots' totals per
// wave, and the totals contracted once with the records' derivatives.
#pragma once

#include "xas_adjoint.h"
#include "xhm_point.h"

namespace wfx {

// the accumulators of a higher mode's adjoint, one per record slot with a
// derivative (kHmAccSlot maps them to the slots)
enum HmAcc : int {
    aAmp0 = 0, aMsec = 1, aFRing = 2, aFDamp = 3, aFalloffAt = 4,
    aPnRe = 5,          // 7
    aPnIm = 12,         // 7
    aPnGlobal = 19, aAmpNorm = 20, aPnDominant = 21,
    aIns = 22,          // 3
    aRd = 25,           // 5
    aRdAux = 30,        // 4
    aInterC = 34, aInterH = 35,
    aInter = 36,        // 8
    aPhi = 44,          // 12: phi_1 .. phi_12
    aPhiL5 = 56, aPhiL6 = 57, aPhiL8 = 58,
    aAlpha0 = 59, aAlpha2 = 60, aAlphaL = 61, aC0 = 62, aC1 = 63, aC2 = 64, aC3 = 65, aC4 = 66, aCL = 67,
    aC1Insp = 68, aC1Rd = 69, aCRd = 70, aDeltaPhi = 71,
    aSph = 72,          // 5
    aMix = 77,          // 4
    aT0 = 81, aKIns = 82, aLogM = 83, aCbrtM = 84, aInvM = 85,
    kHmAcc = 86
};

WFX_CONST int kHmAccSlot[kHmAcc] = {
    kHmAmp0, kHmMsec, kHmFRing, kHmFDamp, kHmFalloffAt,
    kHmPnRe, kHmPnRe + 1, kHmPnRe + 2, kHmPnRe + 3, kHmPnRe + 4, kHmPnRe + 5, kHmPnRe + 6,
    kHmPnIm, kHmPnIm + 1, kHmPnIm + 2, kHmPnIm + 3, kHmPnIm + 4, kHmPnIm + 5, kHmPnIm + 6,
    kHmPnGlobal, kHmAmpNorm, kHmPnDominant,
    kHmIns, kHmIns + 1, kHmIns + 2,
    kHmRd, kHmRd + 1, kHmRd + 2, kHmRd + 3, kHmRd + 4,
    kHmRdAux, kHmRdAux + 1, kHmRdAux + 2, kHmRdAux + 3,
    kHmInterC, kHmInterH,
    kHmInter, kHmInter + 1, kHmInter + 2, kHmInter + 3, kHmInter + 4, kHmInter + 5, kHmInter + 6, kHmInter + 7,
    kHmPhi + 1, kHmPhi + 2, kHmPhi + 3, kHmPhi + 4, kHmPhi + 5, kHmPhi + 6, kHmPhi + 7, kHmPhi + 8, kHmPhi + 9,
    kHmPhi + 10, kHmPhi + 11, kHmPhi + 12,
    kHmPhiL + 5, kHmPhiL + 6, kHmPhiL + 8,
    kHmAlpha0, kHmAlpha2, kHmAlphaL, kHmC0, kHmC1, kHmC2, kHmC3, kHmC4, kHmCL,
    kHmC1Insp, kHmC1Rd, kHmCRd, kHmDeltaPhi,
    kHmSph, kHmSph + 1, kHmSph + 2, kHmSph + 3, kHmSph + 4,
    kHmMix, kHmMix + 1, kHmMix + 2, kHmMix + 3,
    kHmT0F64, kHmKInsF64, kHmLogM, kHmCbrtM, kHmInvM,
};

// the higher mode (index into the four records) whose ringdown recycles the
// (2,2): the (3,2), the only record with kHmMixing set
constexpr int kHmMixMode = 2;

// the cotangents of one bin's powers of M f: of M f, t = (M f)^(1/3),
// (M f)^-1, log(M f) and (M f)^(-7/6)
struct MfGrad {
    float mf, t, m1, lg, m76;
};

// z = (M f - f_ring) / f_damp's adjoint (the arctangents of the
// intermediate and ringdown phases), for its cotangent gz
__device__ __forceinline__ void atan_z_adjoint(float gz, float z, float inv_damp, float* acc, MfGrad& g) {
    g.mf += gz * inv_damp;
    acc[aFRing] -= gz * inv_damp;
    acc[aFDamp] -= gz * z * inv_damp;
}

// a Lorentzian ringdown amplitude rd0 / (exp(rd1 d) (d^2 + rd2)), d = M f -
// f_ring (its value a), for its cotangent ga
__device__ __forceinline__ void lorentzian_adjoint(const float* __restrict__ r, float mf, float a, float ga,
                                                   float* acc, MfGrad& g) {
    const float dfr = mf - r[kHmFRing];
    const float e = __expf(r[kHmRd + 1] * dfr);
    const float q = fmaf(dfr, dfr, r[kHmRd + 2]);
    const float inv_q = __builtin_amdgcn_rcpf(q);
    acc[aRd] += ga * __builtin_amdgcn_rcpf(e * q);
    acc[aRd + 1] -= ga * a * dfr;
    acc[aRd + 2] -= ga * a * inv_q;
    const float gd = -ga * a * fmaf(2.0f * dfr, inv_q, r[kHmRd + 1]);
    g.mf += gd;
    acc[aFRing] -= gd;
}

// the fall-off rd3 exp(-rd4 (M f - f_at)) (its value a), cotangent ga
__device__ __forceinline__ void falloff_adjoint(const float* __restrict__ r, float mf, float a, float ga, float* acc,
                                                MfGrad& g) {
    const float x = mf - r[kHmFalloffAt];
    acc[aRd + 3] += ga * __expf(-r[kHmRd + 4] * x);
    acc[aRd + 4] -= ga * a * x;
    const float gx = -ga * a * r[kHmRd + 4];
    g.mf += gx;
    acc[aFalloffAt] -= gx;
}

// One bin of higher mode r (its kHmOn set) at the powers p of f: the mode's
// value h = amp0 a e^{i phi} (the forward's operations), dL/d(slot) added to
// acc for the cotangent (g_re, g_im) of h, and the cotangent (gw_re, gw_im)
// of the mixing's input w = h_22 / |amp0| (zero where it does not mix).
// Inlined, so that every accumulator index is a constant and acc stays in
// registers; an empty volatile statement ends each region's branch, so that
// their last updates are not merged into one store through a selected
// accumulator (which would move acc to scratch memory).
__device__ __forceinline__ void hm_adjoint_split(const float* __restrict__ r, const BinPowers<float>& p, float w_re,
                                                 float w_im, float g_re, float g_im, float* acc, float& h_re,
                                                 float& h_im, float& gw_re, float& gw_im) {
    // the powers of M f (mf_pow_f32)
    const float mf = r[kHmMsec] * p.f;
    const float t = p.t * r[kHmCbrtM];
    const float m1 = p.m1 * r[kHmInvM];
    const float lg = p.lf + r[kHmLogM];
    const float rs = __builtin_amdgcn_rsqf(t);
    const float m76 = m1 * rs;
    const bool mixing = r[kHmMixing] != 0.0f;
    const float f_ring = r[kHmFRing], f_damp = r[kHmFDamp];
    // the mixing's rotation conj(-a322) w + conj(-a323) A_s e^{i phi_s}
    const bool rotating = mixing && (mf >= r[kHmFAmpIm] || mf >= r[kHmFPhaseIm]);
    // atan((M f - f_ring) / f_damp), of the post-inspiral phases and the
    // rotation only
    float z = 0.0f, atan_z = 0.0f, inv_damp = 0.0f, datan_z = 0.0f;
    if (rotating || !(mf < r[kHmFPhaseIn])) {
        z = fast_div(mf - f_ring, f_damp);
        atan_z = fast_atan(z);
        inv_damp = __builtin_amdgcn_rcpf(f_damp);
        datan_z = __builtin_amdgcn_rcpf(fmaf(z, z, 1.0f));
    }
    float rot_re = 0.0f, rot_im = 0.0f, amplm = 0.0f, sn_s = 0.0f, cs_s = 0.0f;
    if (rotating) {
        if (mf < r[kHmFRdAux]) {
            const float* x = r + kHmRdAux;
            amplm = fmaf(fmaf(fmaf(x[3], mf, x[2]), mf, x[1]), mf, x[0]);
        } else if (mf >= r[kHmFalloff]) {
            amplm = r[kHmRd + 3] * __expf(-r[kHmRd + 4] * (mf - r[kHmFalloffAt]));
        } else {
            const float dfr = mf - f_ring;
            amplm = fast_div(r[kHmRd], __expf(r[kHmRd + 1] * dfr) * fmaf(dfr, dfr, r[kHmRd + 2]));
        }
        const float* sp = r + kHmSph;
        float philm = fmaf(sp[0], mf, sp[4]);
        philm = fmaf(-sp[2], m1, philm);
        philm = fmaf(-(1.0f / 3.0f) * sp[3], m1 * m1 * m1, philm);
        philm = fmaf(sp[1], atan_z, philm);
        __sincosf(reduce_f32(philm), &sn_s, &cs_s);
        const float m2r = r[kHmMix], m2i = r[kHmMix + 1], m3r = r[kHmMix + 2], m3i = r[kHmMix + 3];
        const float b_re = amplm * m3r, b_im = -(m3i * amplm);
        rot_re = fmaf(w_re, m2r, w_im * m2i) + fmaf(b_re, cs_s, -b_im * sn_s);
        rot_im = fmaf(w_im, m2r, -w_re * m2i) + fmaf(b_re, sn_s, b_im * cs_s);
    }

    // the amplitude (hm_amp_phase_split's regions)
    const bool amp_ins = mf < r[kHmFAmpIn];
    const bool amp_rd = !amp_ins && mf >= r[kHmFAmpIm];
    const bool falloff = mf >= r[kHmFalloff];
    const int n_inter = (int)r[kHmNInter];
    float a;
    float pn_re = 0.0f, pn_im = 0.0f, mag = 0.0f, pseudo = 0.0f, ins_poly = 0.0f, t7 = 0.0f, u = 0.0f, inter = 0.0f;
    if (amp_ins) {
        pn_re = r[kHmPnRe + 6];
        pn_im = r[kHmPnIm + 6];
        for (int k = 5; k >= 0; --k) {
            pn_re = fmaf(pn_re, t, r[kHmPnRe + k]);
            pn_im = fmaf(pn_im, t, r[kHmPnIm + k]);
        }
        const float t2 = t * t;
        t7 = t2 * t2 * t2 * t;
        ins_poly = fmaf(fmaf(r[kHmIns + 2], t, r[kHmIns + 1]), t, r[kHmIns]);
        pseudo = t7 * ins_poly;
        mag = __builtin_sqrtf(fmaf(pn_re, pn_re, pn_im * pn_im));
        a = m76 * fmaf(mag, r[kHmPnGlobal] * r[kHmAmpNorm], pseudo * r[kHmPnDominant]);
    } else if (amp_rd) {
        if (mixing) {
            a = __builtin_sqrtf(fmaf(rot_re, rot_re, rot_im * rot_im));
        } else if (falloff) {
            a = r[kHmRd + 3] * __expf(-r[kHmRd + 4] * (mf - r[kHmFalloffAt]));
        } else {
            const float dfr = mf - f_ring;
            a = fast_div(r[kHmRd], __expf(r[kHmRd + 1] * dfr) * fmaf(dfr, dfr, r[kHmRd + 2]));
        }
    } else {
        u = (mf - r[kHmInterC]) * r[kHmInterH];
        // the forward's Horner sum from coefficient n - 1 down
#pragma unroll
        for (int j = 7; j >= 0; --j)
            if (j < n_inter) inter = fmaf(inter, u, r[kHmInter + j]);
        a = inter * m76;
    }
    const bool clamped = a < 0.0f;
    if (clamped) a = 1.0e-15f;

    // the phase
    const bool ph_ins = mf < r[kHmFPhaseIn];
    const bool ph_rd = !ph_ins && mf >= r[kHmFPhaseIm];
    float phi, poly = 0.0f, inv_t = 0.0f, m43 = 0.0f, logs = 0.0f;
    if (ph_ins) {
        poly = r[kHmPhi + 12];
        for (int i = 11; i >= 1; --i) poly = fmaf(poly, t, r[kHmPhi + i]);
        inv_t = fast_div(1.0f, t);
        m43 = m1 * inv_t;
        float rest = poly * m43;
        logs = fmaf(fmaf(r[kHmPhiL + 8], t * t, 0.0f), t, fmaf(r[kHmPhiL + 6], t, r[kHmPhiL + 5]));
        rest = fmaf(lg, logs, rest);
        rest = fmaf(r[kHmC1Insp], mf, rest);
        const double full = fma(record_f64(r, kHmT0F64), p.m53, record_f64(r, kHmKInsF64));
        const double turns = rint(full * kInvTwoPi);
        phi = (float)fma(-turns, kTwoPi, full) + rest;
    } else if (ph_rd) {
        float rest;
        if (mixing) {
            rest = fmaf(r[kHmC1Rd], mf, atan2f(rot_im, rot_re));
        } else {
            rest = fmaf(r[kHmC1Rd] + r[kHmAlpha0], mf, -(f_ring * f_ring * r[kHmAlpha2]) * m1);
            rest = fmaf(r[kHmAlphaL], atan_z, rest);
        }
        phi = rest + r[kHmCRd];
    } else {
        float rest = fmaf(r[kHmC0], mf, r[kHmC1] * lg);
        rest = fmaf(-r[kHmC2], m1, rest);
        const float m12 = m1 * m1;
        rest = fmaf(-(1.0f / 3.0f) * r[kHmC4], m12 * m1, rest);
        if (mixing) rest = fmaf(-0.5f * r[kHmC3], m12, rest);
        rest = fmaf(r[kHmCL], atan_z, rest);
        phi = rest + r[kHmDeltaPhi];
    }
    phi = reduce_f32(phi);
    float sn, cs;
    __sincosf(phi, &sn, &cs);
    const float amp0 = r[kHmAmp0];
    const float val = amp0 * a;
    h_re = val * cs;
    h_im = val * sn;

    // ---------------------------------------------------------- reverse
    const float g_val = fmaf(g_re, cs, g_im * sn);
    const float g_phi = val * fmaf(g_im, cs, -g_re * sn);
    acc[aAmp0] += g_val * a;
    const float g_a = clamped ? 0.0f : g_val * amp0;
    MfGrad g{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    float g_rot_re = 0.0f, g_rot_im = 0.0f;
    // where both the amplitude and the phase take the rotation, h = amp0
    // rot e^{i (c1_rd M f + c_rd)} exactly (a = |rot|, phi = arg rot +
    // c1_rd M f + c_rd): rot's cotangent then needs no division by |rot|,
    // which the deep ringdown takes below FP32's smallest normal
    const bool mix_rd = mixing && amp_rd && ph_rd;
    if (mix_rd) {
        float e_s, e_c;
        __sincosf(reduce_f32(fmaf(r[kHmC1Rd], mf, r[kHmCRd])), &e_s, &e_c);
        g_rot_re = amp0 * fmaf(g_re, e_c, g_im * e_s);
        g_rot_im = amp0 * fmaf(g_im, e_c, -g_re * e_s);
    }

    // the phase's coefficients
    if (ph_ins) {
        acc[aT0] += g_phi * (float)p.m53;
        acc[aKIns] += g_phi;
        acc[aC1Insp] += g_phi * mf;
        g.mf += g_phi * r[kHmC1Insp];
        g.lg += g_phi * logs;
        const float g_logs = g_phi * lg;
        const float t2 = t * t;
        acc[aPhiL5] += g_logs;
        acc[aPhiL6] += g_logs * t;
        acc[aPhiL8] += g_logs * t2 * t;
        g.t += g_logs * fmaf(3.0f * r[kHmPhiL + 8], t2, r[kHmPhiL + 6]);
        const float g_poly = g_phi * m43;
        const float g_m43 = g_phi * poly;
        g.m1 += g_m43 * inv_t;
        g.t -= g_m43 * m43 * inv_t;
        float pw = 1.0f, dpoly = 0.0f;
#pragma unroll
        for (int j = 0; j < 12; ++j) {
            acc[aPhi + j] += g_poly * pw;
            if (j < 11) dpoly = fmaf((float)(j + 1) * r[kHmPhi + j + 2], pw, dpoly);
            pw *= t;
        }
        g.t += g_poly * dpoly;
        __asm__ volatile("");
    } else if (ph_rd) {
        acc[aC1Rd] += g_phi * mf;
        acc[aCRd] += g_phi;
        if (mixing) {
            g.mf += g_phi * r[kHmC1Rd];
            if (!mix_rd) {
                // d atan2(y, x) = (x dy - y dx) / (x^2 + y^2), scaled by
                // max(|x|, |y|) first; zero below FP32's smallest normal
                const float s = fmaxf(fabsf(rot_re), fabsf(rot_im));
                const bool normal = s >= 1.17549435e-38f;
                const float inv_s = normal ? __builtin_amdgcn_rcpf(s) : 0.0f;
                const float xs = rot_re * inv_s, ys = rot_im * inv_s;
                const float rr = normal ? __builtin_amdgcn_rcpf(s * fmaf(xs, xs, ys * ys)) : 0.0f;
                g_rot_re -= g_phi * ys * rr;
                g_rot_im += g_phi * xs * rr;
            }
        } else {
            acc[aAlpha0] += g_phi * mf;
            g.mf += g_phi * (r[kHmC1Rd] + r[kHmAlpha0]);
            const float a2 = r[kHmAlpha2];
            acc[aAlpha2] -= g_phi * f_ring * f_ring * m1;
            acc[aFRing] -= g_phi * 2.0f * f_ring * a2 * m1;
            g.m1 -= g_phi * f_ring * f_ring * a2;
            acc[aAlphaL] += g_phi * atan_z;
            atan_z_adjoint(g_phi * r[kHmAlphaL] * datan_z, z, inv_damp, acc, g);
        }
        __asm__ volatile("");
    } else {
        acc[aC0] += g_phi * mf;
        g.mf += g_phi * r[kHmC0];
        acc[aC1] += g_phi * lg;
        g.lg += g_phi * r[kHmC1];
        acc[aC2] -= g_phi * m1;
        g.m1 -= g_phi * r[kHmC2];
        const float m12 = m1 * m1;
        acc[aC4] -= g_phi * m12 * m1 * (1.0f / 3.0f);
        g.m1 -= g_phi * r[kHmC4] * m12;
        if (mixing) {
            acc[aC3] -= 0.5f * g_phi * m12;
            g.m1 -= g_phi * r[kHmC3] * m1;
        }
        acc[aCL] += g_phi * atan_z;
        atan_z_adjoint(g_phi * r[kHmCL] * datan_z, z, inv_damp, acc, g);
        acc[aDeltaPhi] += g_phi;
    }

    // the amplitude's coefficients
    if (amp_ins) {
        const float glob = r[kHmPnGlobal] * r[kHmAmpNorm], dom = r[kHmPnDominant];
        g.m76 += g_a * fmaf(mag, glob, pseudo * dom);
        const float g_in = g_a * m76;
        acc[aPnGlobal] += g_in * mag * r[kHmAmpNorm];
        acc[aAmpNorm] += g_in * mag * r[kHmPnGlobal];
        acc[aPnDominant] += g_in * pseudo;
        const float g_mag = g_in * glob;
        const float inv_mag = mag > 0.0f ? __builtin_amdgcn_rcpf(mag) : 0.0f;
        const float g_pre = g_mag * pn_re * inv_mag, g_pim = g_mag * pn_im * inv_mag;
        float pw = 1.0f, d_re = 0.0f, d_im = 0.0f;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            acc[aPnRe + k] += g_pre * pw;
            acc[aPnIm + k] += g_pim * pw;
            if (k < 6) {
                d_re = fmaf((float)(k + 1) * r[kHmPnRe + k + 1], pw, d_re);
                d_im = fmaf((float)(k + 1) * r[kHmPnIm + k + 1], pw, d_im);
            }
            pw *= t;
        }
        g.t += fmaf(g_pre, d_re, g_pim * d_im);
        const float g_ps = g_in * dom;
        acc[aIns] += g_ps * t7;
        acc[aIns + 1] += g_ps * t7 * t;
        acc[aIns + 2] += g_ps * t7 * t * t;
        const float t6 = t7 * fast_div(1.0f, t);
        g.t += g_ps * fmaf(7.0f * t6, ins_poly, t7 * fmaf(2.0f * r[kHmIns + 2], t, r[kHmIns + 1]));
        __asm__ volatile("");
    } else if (amp_rd) {
        if (mixing) {
            if (!mix_rd) {
                const float inv_a = a >= 1.17549435e-38f ? __builtin_amdgcn_rcpf(a) : 0.0f;
                g_rot_re += g_a * rot_re * inv_a;
                g_rot_im += g_a * rot_im * inv_a;
            }
        } else if (falloff) {
            falloff_adjoint(r, mf, a, g_a, acc, g);
        } else {
            lorentzian_adjoint(r, mf, a, g_a, acc, g);
        }
        __asm__ volatile("");
    } else {
        g.m76 += g_a * inter;
        const float g_res = g_a * m76;
        float pw = 1.0f, d_u = 0.0f;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (j < n_inter) acc[aInter + j] += g_res * pw;
            if (j + 1 < n_inter) d_u = fmaf((float)(j + 1) * r[kHmInter + j + 1], pw, d_u);
            pw *= u;
        }
        const float g_u = g_res * d_u;
        acc[aInterH] += g_u * (mf - r[kHmInterC]);
        const float g_x = g_u * r[kHmInterH];
        g.mf += g_x;
        acc[aInterC] -= g_x;
    }

    // the mixing's rotation, back to w, the spheroidal phase and amplitude
    gw_re = 0.0f;
    gw_im = 0.0f;
    if (rotating) {
        const float m2r = r[kHmMix], m2i = r[kHmMix + 1], m3r = r[kHmMix + 2], m3i = r[kHmMix + 3];
        acc[aMix] += fmaf(g_rot_re, w_re, g_rot_im * w_im);
        acc[aMix + 1] += fmaf(g_rot_re, w_im, -g_rot_im * w_re);
        gw_re = fmaf(g_rot_re, m2r, -g_rot_im * m2i);
        gw_im = fmaf(g_rot_re, m2i, g_rot_im * m2r);
        const float b_re = amplm * m3r, b_im = -(m3i * amplm);
        const float g_bre = fmaf(g_rot_re, cs_s, g_rot_im * sn_s);
        const float g_bim = fmaf(-g_rot_re, sn_s, g_rot_im * cs_s);
        const float g_phl = g_rot_re * (-b_re * sn_s - b_im * cs_s) + g_rot_im * fmaf(b_re, cs_s, -b_im * sn_s);
        acc[aMix + 2] += g_bre * amplm;
        acc[aMix + 3] -= g_bim * amplm;
        const float g_aml = fmaf(g_bre, m3r, -g_bim * m3i);
        const float* sp = r + kHmSph;
        acc[aSph] += g_phl * mf;
        g.mf += g_phl * sp[0];
        acc[aSph + 4] += g_phl;
        acc[aSph + 2] -= g_phl * m1;
        g.m1 -= g_phl * sp[2];
        acc[aSph + 3] -= g_phl * m1 * m1 * m1 * (1.0f / 3.0f);
        g.m1 -= g_phl * sp[3] * m1 * m1;
        acc[aSph + 1] += g_phl * atan_z;
        atan_z_adjoint(g_phl * sp[1] * datan_z, z, inv_damp, acc, g);
        if (mf < r[kHmFRdAux]) {
            const float* x = r + kHmRdAux;
            acc[aRdAux] += g_aml;
            acc[aRdAux + 1] += g_aml * mf;
            acc[aRdAux + 2] += g_aml * mf * mf;
            acc[aRdAux + 3] += g_aml * mf * mf * mf;
            g.mf += g_aml * fmaf(fmaf(3.0f * x[3], mf, 2.0f * x[2]), mf, x[1]);
        } else if (mf >= r[kHmFalloff]) {
            falloff_adjoint(r, mf, amplm, g_aml, acc, g);
        } else {
            lorentzian_adjoint(r, mf, amplm, g_aml, acc, g);
        }
    }

    // the powers of M f, back to the binary's M slots
    g.m1 += g.m76 * rs;
    g.t -= 0.5f * g.m76 * m76 * fast_div(1.0f, t);
    acc[aMsec] += g.mf * p.f;
    acc[aCbrtM] += g.t * p.t;
    acc[aInvM] += g.m1 * p.m1;
    acc[aLogM] += g.lg;
}

}  // namespace wfx
