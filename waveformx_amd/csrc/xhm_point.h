// The per-(binary, frequency) formula of IMRPhenomXHM on the records of
// xhm_setup.h (FP64): each higher mode's amplitude and phase by region in
// Mf (xhm/reference.py:620-643, LAL's form), the (3,2) ringdown recycling
// the (2,2) mode through the spheroidal-to-spherical rotation, and the
// polarizations sum_lm (F+, Fx)_lm h_{l,-m} (reference.py:656-677).
#pragma once

#include "xas_point.h"
#include "xhm_setup.h"

namespace wfx {

// amplitude (without (-1)^l amp0) and phase of the higher mode of record r at
// mf; (w_re, w_im) the (2,2) mode divided by amp0, recycled by the (3,2)
WFX_HD inline void hm_amp_phase(const double* r, double mf, const Powers<double>& p, double w_re, double w_im,
                                double& a, double& phi) {
    const bool mixing = r[kHmMixing] != 0.0;
    const double f_ring = r[kHmFRing], f_damp = r[kHmFDamp];
    double rot_re = 0.0, rot_im = 0.0;
    if (mixing && (mf >= r[kHmFAmpIm] || mf >= r[kHmFPhaseIm])) {
        // conj(-a322) h22 + conj(-a323) A_s e^{i phi_s} (reference.py:366-378)
        const double amplm = rd_amp(mf, f_ring, f_damp, r[kHmFalloff], true, r + kHmRd, r + kHmRdAux, r[kHmFRdAux],
                                    r[kHmFalloffAt]);
        const double philm = rd_phase_int(mf, p, f_ring, f_damp, true, 0.0, 0.0, 0.0, r[kHmSph], r[kHmSph + 1],
                                          r[kHmSph + 2], r[kHmSph + 3], r[kHmSph + 4]);
        const double m2r = r[kHmMix], m2i = r[kHmMix + 1], m3r = r[kHmMix + 2], m3i = r[kHmMix + 3];
        const double a_re = w_re * m2r + w_im * m2i;
        const double a_im = w_im * m2r - w_re * m2i;
        const double b_re = amplm * m3r;
        const double b_im = -(m3i * amplm);
        double sn, cs;
        sincos(philm, &sn, &cs);
        rot_re = a_re + (b_re * cs - b_im * sn);
        rot_im = a_im + (b_re * sn + b_im * cs);
    }
    if (mf < r[kHmFAmpIn]) {
        a = inspiral_amp(p, r + kHmPnRe, r + kHmPnIm, r[kHmPnGlobal], r[kHmAmpNorm], r[kHmPnDominant], r[kHmFcut73],
                         r[kHmFcut83], r[kHmFcut3], r + kHmIns);
    } else if (mf >= r[kHmFAmpIm]) {
        a = mixing ? m_hypot(rot_re, rot_im)
                   : rd_amp(mf, f_ring, f_damp, r[kHmFalloff], false, r + kHmRd, r + kHmRdAux, r[kHmFRdAux],
                            r[kHmFalloffAt]);
    } else {
        // reference.py:337-345, in u = (f - f_c) / f_h
        const double u = (mf - r[kHmInterC]) / r[kHmInterH];
        const int n = (int)r[kHmNInter];
        double result = r[kHmInter + n - 1];
        for (int j = n - 2; j >= 0; --j) result = m_fma(result, u, r[kHmInter + j]);
        a = result * p.m_seven_sixths;
    }
    a = a < 0.0 ? 1.0e-15 : a;  // LAL's FALSE_ZERO
    if (mf < r[kHmFPhaseIn]) {
        phi = m_fma(r[kHmC1Insp], mf, inspiral_phase_int(mf, p, r + kHmPhi, r + kHmPhiL)) + r[kHmCInsp];
    } else if (mf >= r[kHmFPhaseIm]) {
        const double phi_mrd = mixing ? m_atan2(rot_im, rot_re)
                                      : rd_phase_int(mf, p, f_ring, f_damp, false, r[kHmAlpha0], r[kHmAlpha2],
                                                     r[kHmAlphaL], 0.0, 0.0, 0.0, 0.0, 0.0);
        phi = m_fma(r[kHmC1Rd], mf, phi_mrd) + r[kHmCRd];
    } else {
        double phase_ir = m_fma(r[kHmC0], mf, r[kHmC1] * p.log);
        phase_ir = m_fma(-r[kHmC2], p.m_one, phase_ir);
        phase_ir = m_fma(-(1.0 / 3.0 * r[kHmC4]), p.m_three, phase_ir);
        if (mixing) phase_ir = m_fma(-(0.5 * r[kHmC3]), p.m_two, phase_ir);
        phi = m_fma(r[kHmCL], atan((mf - f_ring) / f_damp), phase_ir);
    }
    phi = phi + r[kHmDeltaPhi];
}

// the five modes h_{l,-m} at f (Hz) of one binary: rec22 its IMRPhenomXAS
// FP64 record, modes its four higher-mode records (kHmSlots apart)
WFX_HD inline void xhm_modes_f64(const double* rec22, const double* modes, double f, double (&h)[5][2]) {
    const FreqPowers<double> fp = freq_powers(f);
    double a22, phi22;
    amp_phase_f64(rec22, fp, a22, phi22);
    double s22, c22;
    sincos(phi22, &s22, &c22);
    h[0][0] = a22 * c22;
    h[0][1] = a22 * s22;
    // the powers of M f from those of f and the binary's M
    const double mf = modes[kHmMsec] * f;
    const Powers<double> p = powers_from(mf, fp.t * modes[kHmCbrtM], fp.m1 * modes[kHmInvM], fp.log_f + modes[kHmLogM]);
    for (int k = 0; k < kHmModes; ++k) {
        const double* r = modes + k * kHmSlots;
        const double amp0 = fabs(r[kHmAmp0]);
        double a, phi;
        hm_amp_phase(r, mf, p, h[0][0] / amp0, h[0][1] / amp0, a, phi);
        const double v = r[kHmAmp0] * a;
        double sn, cs;
        sincos(phi, &sn, &cs);
        const bool on = r[kHmOn] != 0.0;
        h[k + 1][0] = on ? v * cs : 0.0;
        h[k + 1][1] = on ? v * sn : 0.0;
    }
}

// (h+, hx) = sum_k (F+, Fx)_k h_k (head: kHeadFactors, 4 per mode)
WFX_HD inline void xhm_sum(const double* head, const double (&h)[5][2], double& hp_re, double& hp_im, double& hc_re,
                           double& hc_im) {
    hp_re = hp_im = hc_re = hc_im = 0.0;
    for (int k = 0; k < 5; ++k) {
        const double* fac = head + kHeadFactors + 4 * k;
        hp_re = hp_re + (fac[0] * h[k][0] - fac[1] * h[k][1]);
        hp_im = hp_im + (fac[0] * h[k][1] + fac[1] * h[k][0]);
        hc_re = hc_re + (fac[2] * h[k][0] - fac[3] * h[k][1]);
        hc_im = hc_im + (fac[2] * h[k][1] + fac[3] * h[k][0]);
    }
}

// ------------------------------------------------------------- split

// the powers of M f in FP32 from those of f (pow_f32) and the binary's M
struct MfPowF32 {
    float mf, t, m1, lg, m76;
};
__device__ inline MfPowF32 mf_pow_f32(const float* r, const PowF32& p) {
    MfPowF32 q;
    q.mf = r[kHmMsec] * p.f;
    q.t = p.t * r[kHmCbrtM];
    q.m1 = p.m1 * r[kHmInvM];
    q.lg = p.lf + r[kHmLogM];
    q.m76 = q.m1 * __builtin_amdgcn_rsqf(q.t);
    return q;
}

// hm_amp_phase in FP32 on a split record (write_mode_record_split): the
// inspiral's leading term and constant in FP64, reduced modulo 2 pi, the
// rest in FP32; (w_re, w_im) the (2,2) mode over amp0
__device__ inline void hm_amp_phase_split(const float* __restrict__ r, const MfPowF32& q, double m53, float w_re,
                                          float w_im, float& a, float& phi) {
    const bool mixing = r[kHmMixing] != 0.0f;
    const float mf = q.mf, t = q.t, m1 = q.m1;
    const float f_ring = r[kHmFRing], f_damp = r[kHmFDamp];
    float rot_re = 0.0f, rot_im = 0.0f;
    if (mixing && (mf >= r[kHmFAmpIm] || mf >= r[kHmFPhaseIm])) {
        float amplm;
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
        philm = fmaf(sp[1], fast_atan(fast_div(mf - f_ring, f_damp)), philm);
        float sn, cs;
        __sincosf(reduce_f32(philm), &sn, &cs);
        const float m2r = r[kHmMix], m2i = r[kHmMix + 1], m3r = r[kHmMix + 2], m3i = r[kHmMix + 3];
        const float b_re = amplm * m3r, b_im = -(m3i * amplm);
        rot_re = fmaf(w_re, m2r, w_im * m2i) + fmaf(b_re, cs, -b_im * sn);
        rot_im = fmaf(w_im, m2r, -w_re * m2i) + fmaf(b_re, sn, b_im * cs);
    }
    if (mf < r[kHmFAmpIn]) {
        // |sum_k pn_k t^k| pn_global m76 amp_norm + t^7 (ins_0 + ins_1 t + ins_2 t^2) m76 pn_dominant
        float re = r[kHmPnRe + 6], im = r[kHmPnIm + 6];
        for (int k = 5; k >= 0; --k) {
            re = fmaf(re, t, r[kHmPnRe + k]);
            im = fmaf(im, t, r[kHmPnIm + k]);
        }
        const float t2 = t * t, t7 = t2 * t2 * t2 * t;
        const float pseudo = t7 * fmaf(fmaf(r[kHmIns + 2], t, r[kHmIns + 1]), t, r[kHmIns]);
        a = q.m76 * fmaf(__builtin_sqrtf(fmaf(re, re, im * im)), r[kHmPnGlobal] * r[kHmAmpNorm],
                         pseudo * r[kHmPnDominant]);
    } else if (mf >= r[kHmFAmpIm]) {
        if (mixing) {
            a = __builtin_sqrtf(fmaf(rot_re, rot_re, rot_im * rot_im));
        } else if (mf >= r[kHmFalloff]) {
            a = r[kHmRd + 3] * __expf(-r[kHmRd + 4] * (mf - r[kHmFalloffAt]));
        } else {
            const float dfr = mf - f_ring;
            a = fast_div(r[kHmRd], __expf(r[kHmRd + 1] * dfr) * fmaf(dfr, dfr, r[kHmRd + 2]));
        }
    } else {
        const float u = (mf - r[kHmInterC]) * r[kHmInterH];
        const int n = (int)r[kHmNInter];
        float result = r[kHmInter + n - 1];
        for (int j = n - 2; j >= 0; --j) result = fmaf(result, u, r[kHmInter + j]);
        a = result * q.m76;
    }
    a = a < 0.0f ? 1.0e-15f : a;
    if (mf < r[kHmFPhaseIn]) {
        // sum_{i >= 1} phi_i (M f)^((i-5)/3) as t^-4 times a polynomial in t, the log terms (i = 5, 6, 8)
        // and c1_insp M f in FP32; phi_0 (M f)^(-5/3) + c_insp + dphi_lm in FP64
        float poly = r[kHmPhi + 12];
        for (int i = 11; i >= 1; --i) poly = fmaf(poly, t, r[kHmPhi + i]);
        const float m43 = m1 * fast_div(1.0f, t);
        float rest = poly * m43;
        rest = fmaf(q.lg, fmaf(fmaf(r[kHmPhiL + 8], t * t, 0.0f), t, fmaf(r[kHmPhiL + 6], t, r[kHmPhiL + 5])), rest);
        rest = fmaf(r[kHmC1Insp], mf, rest);
        const double full = fma(record_f64(r, kHmT0F64), m53, record_f64(r, kHmKInsF64));
        const double turns = rint(full * kInvTwoPi);
        phi = (float)fma(-turns, kTwoPi, full) + rest;
    } else if (mf >= r[kHmFPhaseIm]) {
        float rest;
        if (mixing) {
            rest = fmaf(r[kHmC1Rd], mf, atan2f(rot_im, rot_re));
        } else {
            rest = fmaf(r[kHmC1Rd] + r[kHmAlpha0], mf, -(f_ring * f_ring * r[kHmAlpha2]) * m1);
            rest = fmaf(r[kHmAlphaL], fast_atan(fast_div(mf - f_ring, f_damp)), rest);
        }
        phi = rest + r[kHmCRd];
    } else {
        float rest = fmaf(r[kHmC0], mf, r[kHmC1] * q.lg);
        rest = fmaf(-r[kHmC2], m1, rest);
        const float m12 = m1 * m1;
        rest = fmaf(-(1.0f / 3.0f) * r[kHmC4], m12 * m1, rest);
        if (mixing) rest = fmaf(-0.5f * r[kHmC3], m12, rest);
        rest = fmaf(r[kHmCL], fast_atan(fast_div(mf - f_ring, f_damp)), rest);
        phi = rest + r[kHmDeltaPhi];
    }
    phi = reduce_f32(phi);
}

// the five modes in FP32 (split): rec22 the (2,2)'s split record, modes the
// four split higher-mode records
__device__ inline void xhm_modes_split(const float* __restrict__ rec22, const float* __restrict__ modes, float f,
                                       float (&h)[5][2]) {
    const PowF32 p = pow_f32(f);
    float a22, phi22;
    amp_phase_f32<kLeadF64>(rec22, p, a22, phi22);
    float s22, c22;
    __sincosf(phi22, &s22, &c22);
    h[0][0] = a22 * c22;
    h[0][1] = a22 * s22;
    const MfPowF32 q = mf_pow_f32(modes, p);
    for (int k = 0; k < kHmModes; ++k) {
        const float* r = modes + k * kHmSlots;
        const float inv_amp0 = __builtin_amdgcn_rcpf(fabsf(r[kHmAmp0]));
        float a, phi;
        hm_amp_phase_split(r, q, p.m53, h[0][0] * inv_amp0, h[0][1] * inv_amp0, a, phi);
        const float v = r[kHmAmp0] * a;
        float sn, cs;
        __sincosf(phi, &sn, &cs);
        const bool on = r[kHmOn] != 0.0f;
        h[k + 1][0] = on ? v * cs : 0.0f;
        h[k + 1][1] = on ? v * sn : 0.0f;
    }
}

__device__ inline void xhm_sum_f32(const float* head, const float (&h)[5][2], float2& hp, float2& hc) {
    float pr = 0.0f, pi = 0.0f, cr = 0.0f, ci = 0.0f;
    for (int k = 0; k < 5; ++k) {
        const float* fac = head + kHeadFactors + 4 * k;
        pr = fmaf(fac[0], h[k][0], fmaf(-fac[1], h[k][1], pr));
        pi = fmaf(fac[0], h[k][1], fmaf(fac[1], h[k][0], pi));
        cr = fmaf(fac[2], h[k][0], fmaf(-fac[3], h[k][1], cr));
        ci = fmaf(fac[2], h[k][1], fmaf(fac[3], h[k][0], ci));
    }
    hp = make_float2(pr, pi);
    hc = make_float2(cr, ci);
}

}  // namespace wfx
