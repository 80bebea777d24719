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
        const double amplm = rd_amp(mf, f_ring, f_damp, r[kHmFalloff], true, r + kHmRd, r + kHmRdAux, r[kHmFRdAux]);
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
                   : rd_amp(mf, f_ring, f_damp, r[kHmFalloff], false, r + kHmRd, r + kHmRdAux, r[kHmFRdAux]);
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
    const double mf = modes[kHmMsec] * f;
    const Powers<double> p = useful_powers(mf);
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

}  // namespace wfx
