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

// The split layout (write_mode_record_split): the records' slots as the
// split evaluation takes them, each mode's leading inspiral term and
// constant as an FP64 slot. Its formula is one template in the scalar T:
// FP32 for the split evaluation, FP64 for the exact FP64 gradient, which
// takes the same layout in doubles.

// the powers of M f from those of f (pow_f32 or freq_powers) and the
// binary's M slots
template <class T>
struct MfPow {
    T mf, t, m1, lg, m76;
};
using MfPowF32 = MfPow<float>;
__device__ inline float lf_of(const PowF32& p) { return p.lf; }
__device__ inline double lf_of(const FreqPowers<double>& p) { return p.log_f; }
// the FP64 leading term as the T sum takes it: FP32 reduced modulo 2 pi in
// FP64 first, FP64 as it is (the sum is reduced afterwards)
__device__ inline float lead_turned(float, double full) {
    const double turns = rint(full * kInvTwoPi);
    return (float)fma(-turns, kTwoPi, full);
}
__device__ inline double lead_turned(double, double full) { return full; }
template <class T>
__device__ inline MfPow<T> mf_pow(const T* r, T f, T t, T m1, T lf) {
    MfPow<T> q;
    q.mf = r[kHmMsec] * f;
    q.t = t * r[kHmCbrtM];
    q.m1 = m1 * r[kHmInvM];
    q.lg = lf + r[kHmLogM];
    q.m76 = q.m1 * d_rsq(q.t);
    return q;
}
__device__ inline MfPowF32 mf_pow_f32(const float* r, const PowF32& p) { return mf_pow(r, p.f, p.t, p.m1, p.lf); }

// hm_amp_phase in the split layout: the inspiral's leading term and
// constant in FP64 (reduced modulo 2 pi before the rest is added in FP32),
// the rest in T; (w_re, w_im) the (2,2) mode over amp0
template <class T>
__device__ inline void hm_amp_phase_split(const T* __restrict__ r, const MfPow<T>& q, double m53, T w_re, T w_im,
                                          T& a, T& phi) {
    const bool mixing = r[kHmMixing] != T(0);
    const T mf = q.mf, t = q.t, m1 = q.m1;
    const T f_ring = r[kHmFRing], f_damp = r[kHmFDamp];
    T rot_re = T(0), rot_im = T(0);
    if (mixing && (mf >= r[kHmFAmpIm] || mf >= r[kHmFPhaseIm])) {
        T amplm;
        if (mf < r[kHmFRdAux]) {
            const T* x = r + kHmRdAux;
            amplm = d_fma(d_fma(d_fma(x[3], mf, x[2]), mf, x[1]), mf, x[0]);
        } else if (mf >= r[kHmFalloff]) {
            amplm = r[kHmRd + 3] * d_exp(-r[kHmRd + 4] * (mf - r[kHmFalloffAt]));
        } else {
            const T dfr = mf - f_ring;
            amplm = d_div(r[kHmRd], d_exp(r[kHmRd + 1] * dfr) * d_fma(dfr, dfr, r[kHmRd + 2]));
        }
        const T* sp = r + kHmSph;
        T philm = d_fma(sp[0], mf, sp[4]);
        philm = d_fma(-sp[2], m1, philm);
        philm = d_fma(-(T(1) / T(3)) * sp[3], m1 * m1 * m1, philm);
        philm = d_fma(sp[1], d_atan(d_div(mf - f_ring, f_damp)), philm);
        T sn, cs;
        m_sincos(d_reduce(philm), &sn, &cs);
        const T m2r = r[kHmMix], m2i = r[kHmMix + 1], m3r = r[kHmMix + 2], m3i = r[kHmMix + 3];
        const T b_re = amplm * m3r, b_im = -(m3i * amplm);
        rot_re = d_fma(w_re, m2r, w_im * m2i) + d_fma(b_re, cs, -b_im * sn);
        rot_im = d_fma(w_im, m2r, -w_re * m2i) + d_fma(b_re, sn, b_im * cs);
    }
    if (mf < r[kHmFAmpIn]) {
        // |sum_k pn_k t^k| pn_global m76 amp_norm + t^7 (ins_0 + ins_1 t + ins_2 t^2) m76 pn_dominant
        T re = r[kHmPnRe + 6], im = r[kHmPnIm + 6];
        for (int k = 5; k >= 0; --k) {
            re = d_fma(re, t, r[kHmPnRe + k]);
            im = d_fma(im, t, r[kHmPnIm + k]);
        }
        const T t2 = t * t, t7 = t2 * t2 * t2 * t;
        const T pseudo = t7 * d_fma(d_fma(r[kHmIns + 2], t, r[kHmIns + 1]), t, r[kHmIns]);
        a = q.m76 * d_fma(d_sqrt(d_fma(re, re, im * im)), r[kHmPnGlobal] * r[kHmAmpNorm], pseudo * r[kHmPnDominant]);
    } else if (mf >= r[kHmFAmpIm]) {
        if (mixing) {
            a = d_sqrt(d_fma(rot_re, rot_re, rot_im * rot_im));
        } else if (mf >= r[kHmFalloff]) {
            a = r[kHmRd + 3] * d_exp(-r[kHmRd + 4] * (mf - r[kHmFalloffAt]));
        } else {
            const T dfr = mf - f_ring;
            a = d_div(r[kHmRd], d_exp(r[kHmRd + 1] * dfr) * d_fma(dfr, dfr, r[kHmRd + 2]));
        }
    } else {
        const T u = (mf - r[kHmInterC]) * r[kHmInterH];
        const int n = (int)r[kHmNInter];
        T result = r[kHmInter + n - 1];
        for (int j = n - 2; j >= 0; --j) result = d_fma(result, u, r[kHmInter + j]);
        a = result * q.m76;
    }
    a = a < T(0) ? T(1.0e-15) : a;
    if (mf < r[kHmFPhaseIn]) {
        // sum_{i >= 1} phi_i (M f)^((i-5)/3) as t^-4 times a polynomial in t, the log terms (i = 5, 6, 8)
        // and c1_insp M f in T; phi_0 (M f)^(-5/3) + c_insp + dphi_lm in FP64
        T poly = r[kHmPhi + 12];
        for (int i = 11; i >= 1; --i) poly = d_fma(poly, t, r[kHmPhi + i]);
        const T m43 = m1 * d_div(T(1), t);
        T rest = poly * m43;
        rest = d_fma(q.lg, d_fma(d_fma(r[kHmPhiL + 8], t * t, T(0)), t, d_fma(r[kHmPhiL + 6], t, r[kHmPhiL + 5])),
                     rest);
        rest = d_fma(r[kHmC1Insp], mf, rest);
        const double full = fma(wide_slot(r, kHmT0F64), m53, wide_slot(r, kHmKInsF64));
        phi = lead_turned(T(0), full) + rest;
    } else if (mf >= r[kHmFPhaseIm]) {
        T rest;
        if (mixing) {
            rest = d_fma(r[kHmC1Rd], mf, d_atan2(rot_im, rot_re));
        } else {
            rest = d_fma(r[kHmC1Rd] + r[kHmAlpha0], mf, -(f_ring * f_ring * r[kHmAlpha2]) * m1);
            rest = d_fma(r[kHmAlphaL], d_atan(d_div(mf - f_ring, f_damp)), rest);
        }
        phi = rest + r[kHmCRd];
    } else {
        T rest = d_fma(r[kHmC0], mf, r[kHmC1] * q.lg);
        rest = d_fma(-r[kHmC2], m1, rest);
        const T m12 = m1 * m1;
        rest = d_fma(-(T(1) / T(3)) * r[kHmC4], m12 * m1, rest);
        if (mixing) rest = d_fma(-T(0.5) * r[kHmC3], m12, rest);
        rest = d_fma(r[kHmCL], d_atan(d_div(mf - f_ring, f_damp)), rest);
        phi = rest + r[kHmDeltaPhi];
    }
    phi = d_reduce(phi);
}

// the (2,2) mode of a split-layout record at the bin's powers: FP32 from
// the FP32 record (amp_phase_f32 with its FP64 leading term), FP64 from the
// FP64 IMRPhenomXAS record
__device__ inline void mode22_split(const float* __restrict__ rec22, const PowF32& p, float& a, float& phi) {
    amp_phase_f32<kLeadF64>(rec22, p, a, phi);
}
__device__ inline void mode22_split(const double* __restrict__ rec22, const FreqPowers<double>& p, double& a,
                                    double& phi) {
    amp_phase_f64(rec22, p, a, phi);
}

// the five modes in the split layout at one bin: rec22 the (2,2)'s record,
// modes the four higher-mode records; p the frequency's powers (pow_f32 in
// FP32, freq_powers in FP64)
template <class T, class P>
__device__ inline void xhm_modes_split_at(const T* __restrict__ rec22, const T* __restrict__ modes, const P& p,
                                          T (&h)[5][2]) {
    T a22, phi22;
    mode22_split(rec22, p, a22, phi22);
    T s22, c22;
    m_sincos(phi22, &s22, &c22);
    h[0][0] = a22 * c22;
    h[0][1] = a22 * s22;
    const MfPow<T> q = mf_pow(modes, (T)p.f, (T)p.t, (T)p.m1, (T)lf_of(p));
    for (int k = 0; k < kHmModes; ++k) {
        const T* r = modes + k * kHmSlots;
        const T inv_amp0 = d_rcp(fabs(r[kHmAmp0]));
        T a, phi;
        hm_amp_phase_split(r, q, p.m53, h[0][0] * inv_amp0, h[0][1] * inv_amp0, a, phi);
        const T v = r[kHmAmp0] * a;
        T sn, cs;
        m_sincos(phi, &sn, &cs);
        const bool on = r[kHmOn] != T(0);
        h[k + 1][0] = on ? v * cs : T(0);
        h[k + 1][1] = on ? v * sn : T(0);
    }
}

// the five modes in FP32 (split) at f
__device__ inline void xhm_modes_split(const float* __restrict__ rec22, const float* __restrict__ modes, float f,
                                       float (&h)[5][2]) {
    xhm_modes_split_at(rec22, modes, pow_f32(f), h);
}

// (h+, hx) = sum_k (F+, Fx)_k h_k with fused products (the split layout's
// sum, in T)
template <class T, class T2>
__device__ inline void xhm_sum_split(const T* head, const T (&h)[5][2], T2& hp, T2& hc) {
    T pr = T(0), pi = T(0), cr = T(0), ci = T(0);
    for (int k = 0; k < 5; ++k) {
        const T* fac = head + kHeadFactors + 4 * k;
        pr = d_fma(fac[0], h[k][0], d_fma(-fac[1], h[k][1], pr));
        pi = d_fma(fac[0], h[k][1], d_fma(fac[1], h[k][0], pi));
        cr = d_fma(fac[2], h[k][0], d_fma(-fac[3], h[k][1], cr));
        ci = d_fma(fac[2], h[k][1], d_fma(fac[3], h[k][0], ci));
    }
    hp.x = pr;
    hp.y = pi;
    hc.x = cr;
    hc.y = ci;
}
__device__ inline void xhm_sum_f32(const float* head, const float (&h)[5][2], float2& hp, float2& hc) {
    xhm_sum_split(head, h, hp, hc);
}

}  // namespace wfx
