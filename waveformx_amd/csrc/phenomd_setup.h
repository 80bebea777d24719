// The per-binary setup of IMRPhenomD: LAL's coefficients in M f
// (phenomd/core.py:154-268, the fits of phenomd/fits.py, the TaylorF2
// phasing of taylorf2/pn.py) folded into the record of phenomd_layout.h in
// powers of f in Hz (core.py:274-377), with the reference phase and the
// leading term T0 f^(-5/3) - A as in IMRPhenomXAS.
//
// Every function is templated over the scalar R (double; a forward-mode
// dual for derivatives) and keeps the reference's order of operations:
// FP64 parity with the restatement the tests hold is at rounding level.
// Where the reference leaves the order to JAX, two choices are fixed:
// integer powers by binary exponentiation (ipow, the scheme of
// jax.lax.integer_pow) and the 5 x 5 amplitude system by pivoted_solve
// (LAPACK's pivots, unfused right-looking elimination).
#pragma once

#include "phenomd_layout.h"
#include "xas_setup.h"

#include "phenomd_fits.inc"

namespace wfx {

enum PhenomDFit : int {
    kFitRho1 = 0, kFitRho2, kFitRho3, kFitGamma1, kFitGamma2, kFitGamma3, kFitCollocation, kFitAlpha1, kFitAlpha2,
    kFitAlpha3, kFitAlpha4, kFitAlpha5, kFitBeta1, kFitBeta2, kFitBeta3, kFitSigma1, kFitSigma2, kFitSigma3,
    kFitSigma4, kFitCount
};

// the ends of the inspiral phase and amplitude in M f (core.py:44-46)
constexpr double kPdMfPhaseIns = 0.018;
constexpr double kPdMfAmpIns = 0.014;
constexpr double kPdLog6 = -6848.0 / 21.0;

// x^n, n a (constant) integer, by binary exponentiation; 1 / x^|n| for n < 0
template <class R>
WFX_HD inline R ipow(R x, int n) {
    if (n == 0) return R(1.0);
    const bool recip = n < 0;
    if (recip) n = -n;
    R acc = x;
    bool have = false;
    while (n > 0) {
        if (n & 1) {
            acc = have ? acc * x : x;
            have = true;
        }
        n >>= 1;
        if (n > 0) x = x * x;
    }
    return recip ? R(1.0) / acc : acc;
}

// ------------------------------------------------------------------ fits

// fits.py:25-29
template <class R>
WFX_HD inline R pd_chi_pn(const R& eta, const R& delta, const R& chi1, const R& chi2) {
    const R chi_s = R(0.5) * (chi1 + chi2);
    const R chi_a = R(0.5) * (chi1 - chi2);
    return chi_s * (R(1.0) - eta * R(76.0) / R(113.0)) + delta * chi_a;
}

// fits.py:32-37: sum_i xi^i sum_j c_ij eta^j
template <class R>
WFX_HD inline R pd_fit(const R& eta, const R& xi, int fit) {
    R total = R(0.0);
    for (int i = 3; i >= 0; --i)
        total = total * xi + (R(kPdFits[fit][i][0]) + eta * (R(kPdFits[fit][i][1]) + eta * R(kPdFits[fit][i][2])));
    return total;
}

// FinalSpin0815 (fits.py:131-147)
template <class R>
WFX_HD inline R pd_final_spin(const R& eta, const R& delta, const R& chi1, const R& chi2) {
    const R x1 = R(0.5) * (R(1.0) + delta), x2 = R(0.5) * (R(1.0) - delta);
    const R s = x1 * x1 * chi1 + x2 * x2 * chi2;
    R total = R(0.0);
    for (int i = 4; i >= 0; --i) {
        R inner = R(0.0);
        for (int j = 4; j >= 0; --j) inner = inner * eta + R(kPdFinalSpin[i][j]);
        total = total * s + inner;
    }
    return total;
}

// EradRational0815 (fits.py:150-161)
template <class R>
WFX_HD inline R pd_radiated_energy(const R& eta, const R& delta, const R& chi1, const R& chi2) {
    const R x1 = R(0.5) * (R(1.0) + delta), x2 = R(0.5) * (R(1.0) - delta);
    const R m1s = x1 * x1, m2s = x2 * x2;
    const R s = (m1s * chi1 + m2s * chi2) / (m1s + m2s);
    const R eta2 = eta * eta;
    const R eta3 = eta2 * eta;
    return (eta * (R(0.055974469826360077) + R(0.5809510763115132) * eta - R(0.9606726679372312) * eta2 +
                   R(3.352411249771192) * eta3) *
            (R(1.0) + (R(-0.0030302335878845507) - R(2.0066110851351073) * eta + R(7.7050567802399215) * eta2) * s)) /
           (R(1.0) + (R(-0.6714403054720589) - R(1.4756929437702908) * eta + R(7.304676214885011) * eta2) * s);
}

// Clenshaw's sum of a Chebyshev series (fits.py:164-169)
template <class R, int N>
WFX_HD inline R pd_clenshaw(const double (&c)[N], const R& t) {
    R b1 = R(0.0), b2 = R(0.0);
    for (int k = N - 1; k >= 1; --k) {
        const R next = R(2.0) * t * b1 - b2 + R(c[k]);
        b2 = b1;
        b1 = next;
    }
    return t * b1 - b2 + R(c[0]);
}

// ------------------------------------------------- TaylorF2 (taylorf2/pn.py)

template <class R>
WFX_HD inline R pd_phase_3(const R& x1, const R& x2, const R& chi1, const R& chi2) {
    R total = R(0.0);
    total = total + x1 * (R(25.0) + R(38.0 / 3.0) * x1) * chi1;
    total = total + x2 * (R(25.0) + R(38.0 / 3.0) * x2) * chi2;
    return R(-16.0 * kPi) + total;
}

template <class R>
WFX_HD inline R pd_phase_4(const R& eta, const R& x1, const R& x2, const R& chi1, const R& chi2) {
    R phase = R(15293365.0 / 508032.0) + R(27145.0 / 504.0) * eta + R(3085.0 / 72.0) * eta * eta;
    phase = phase - R(395.0 / 4.0) * eta * chi1 * chi2;
    phase = phase - R(50.0 + 5.0 / 8.0) * x1 * x1 * chi1 * chi1;
    phase = phase - R(50.0 + 5.0 / 8.0) * x2 * x2 * chi2 * chi2;
    return phase;
}

template <class R>
WFX_HD inline R pd_phase_5_term(const R& x, const R& chi) {
    return chi * x *
           (R(13915.0 / 84.0) - x * (R(1.0) - x) * R(10.0) / R(3.0) +
            x * (R(12760.0 / 81.0) + x * (R(1.0) - x) * R(170.0) / R(9.0)));
}

template <class R>
WFX_HD inline R pd_phase_5(const R& eta, const R& x1, const R& x2, const R& chi1, const R& chi2) {
    R phase = R(5.0 / 9.0) * (R(7729.0 / 84.0) - R(13.0) * eta) * R(kPi);
    phase = phase - pd_phase_5_term(x1, chi1);
    phase = phase - pd_phase_5_term(x2, chi2);
    return phase;
}

// Subtract3PNSS of IMRPhenomD
template <class R>
WFX_HD inline R pd_spin_spin_3pn(const R& eta, const R& x1, const R& x2, const R& chi1, const R& chi2) {
    R phase = (R(32675.0 / 112.0) + R(5575.0 / 18.0) * eta) * eta * chi1 * chi2;
    const R xs[2] = {x1, x2}, chis[2] = {chi1, chi2};
    for (int i = 0; i < 2; ++i) {
        const R x = xs[i], chi = chis[i];
        phase = phase + ((R(47035.0 / 84.0) + R(2935.0 / 6.0) * x - R(120.0) * x * x) +
                         (R(-410825.0 / 672.0) - R(1085.0 / 12.0) * x + R(1255.0 / 36.0) * x * x)) *
                            x * x * chi * chi;
    }
    return phase;
}

template <class R>
WFX_HD inline R pd_phase_6(const R& eta, const R& x1, const R& x2, const R& chi1, const R& chi2) {
    R phase = R(11583231236531.0 / 4694215680.0 - 640.0 / 3.0 * kPdPi2 - 6848.0 / 21.0 * kPdEulerGamma);
    phase = phase + eta * R(-15737765635.0 / 3048192.0 + 2255.0 / 12.0 * kPdPi2);
    phase = phase + eta * eta * R(76055.0) / R(1728.0) - eta * eta * eta * R(127825.0) / R(1296.0);
    phase = phase + R(kPdLog6 * kPdLn4) + pd_spin_spin_3pn(eta, x1, x2, chi1, chi2);
    phase = phase + R(kPi) * x1 * (R(1490.0 / 3.0) + x1 * R(260.0)) * chi1;
    phase = phase + R(kPi) * x2 * (R(1490.0 / 3.0) + x2 * R(260.0)) * chi2;
    return phase;
}

template <class R>
WFX_HD inline R pd_phase_7_term(const R& eta, const R& x, const R& chi) {
    return chi * x *
           (R(-170978035.0 / 48384.0) + eta * R(2876425.0) / R(672.0) + eta * eta * R(4735.0) / R(144.0) +
            x * (R(-7189233785.0 / 1524096.0) + eta * R(458555.0) / R(3024.0) - eta * eta * R(5345.0) / R(72.0)));
}

template <class R>
WFX_HD inline R pd_phase_7(const R& eta, const R& x1, const R& x2, const R& chi1, const R& chi2) {
    R phase = R(kPi) * (R(77096675.0 / 254016.0) + R(378515.0 / 1512.0) * eta - R(74045.0 / 756.0) * eta * eta);
    phase = phase + pd_phase_7_term(eta, x1, chi1);
    phase = phase + pd_phase_7_term(eta, x2, chi2);
    return phase;
}

// ----------------------------------------------- the model in M f

template <class R>
struct PdCoeffs {
    R eta, delta, final_spin, f_ring, f_damp;
    R v[12], l5, l6;  // v_k of v^(k-5), l_k of v^(k-5) log v (v = (pi M f)^(1/3))
    R beta[3], alpha[5];
    R c1_int, c2_int, c1_mrd, c2_mrd;
    R rho[10], gamma[3], f_peak, t0, sigma, collocation;
    R delta_int[5];
};

// the phase and amplitude pieces (core.py:178-245)
template <class R>
WFX_HD inline R pd_ins(const PdCoeffs<R>& c, const R& mf) {
    const R u = m_cbrt(R(kPi) * mf);
    const R log_u = m_log(u);
    R total = R(0.0);
    for (int k = 0; k < 12; ++k)
        if (k != 1) total = total + c.v[k] * ipow(u, k - 5);
    R extra = R(0.0);
    extra = extra + c.l5 * ipow(u, 0) * log_u;
    extra = extra + c.l6 * ipow(u, 1) * log_u;
    return total + extra;
}

template <class R>
WFX_HD inline R pd_dins(const PdCoeffs<R>& c, const R& mf) {
    const R u = m_cbrt(R(kPi) * mf);
    const R log_u = m_log(u);
    R total = R(0.0);
    for (int k = 0; k < 12; ++k)
        if (k != 1) total = total + R((k - 5) / 3.0) * c.v[k] * ipow(u, k - 5);
    R extra = R(0.0);
    extra = extra + c.l5 * ipow(u, 0) * (R(0.0 / 3.0) * log_u + R(1.0 / 3.0));
    extra = extra + c.l6 * ipow(u, 1) * (R(1.0 / 3.0) * log_u + R(1.0 / 3.0));
    return (total + extra) / mf;
}

template <class R>
WFX_HD inline R pd_inter(const PdCoeffs<R>& c, const R& mf) {
    return (c.beta[0] * mf - c.beta[2] / R(3.0) / ipow(mf, 3) + c.beta[1] * m_log(mf)) / c.eta;
}

template <class R>
WFX_HD inline R pd_dinter(const PdCoeffs<R>& c, const R& mf) {
    return (c.beta[0] + c.beta[2] / ipow(mf, 4) + c.beta[1] / mf) / c.eta;
}

template <class R>
WFX_HD inline R pd_mrd(const PdCoeffs<R>& c, const R& mf) {
    return (c.alpha[0] * mf - c.alpha[1] / mf + R(4.0 / 3.0) * c.alpha[2] * m_pow(mf, 0.75) +
            c.alpha[3] * m_atan((mf - c.alpha[4] * c.f_ring) / c.f_damp)) /
           c.eta;
}

template <class R>
WFX_HD inline R pd_dmrd(const PdCoeffs<R>& c, const R& mf) {
    const R y = (mf - c.alpha[4] * c.f_ring) / c.f_damp;
    return (c.alpha[0] + c.alpha[1] / ipow(mf, 2) + c.alpha[2] / m_pow(mf, 0.25) +
            c.alpha[3] / (c.f_damp * (R(1.0) + y * y))) /
           c.eta;
}

template <class R>
WFX_HD inline R pd_amp_ins(const PdCoeffs<R>& c, const R& mf) {
    const R u = m_cbrt(R(kPi) * mf);
    R total = R(0.0);
    for (int k = 0; k < 10; ++k)
        if (k != 1) total = total + c.rho[k] * ipow(u, k);
    return total;
}

template <class R>
WFX_HD inline R pd_damp_ins(const PdCoeffs<R>& c, const R& mf) {
    const R u = m_cbrt(R(kPi) * mf);
    R total = R(0.0);
    for (int k = 2; k < 10; ++k) total = total + R(k / 3.0) * c.rho[k] * ipow(u, k);
    return total / mf;
}

template <class R>
WFX_HD inline R pd_amp_mrd(const PdCoeffs<R>& c, const R& mf) {
    const R d = mf - c.f_ring;
    return c.gamma[0] * c.sigma / (d * d + c.sigma * c.sigma) * m_exp(-c.gamma[1] * d / c.sigma);
}

template <class R>
WFX_HD inline R pd_damp_mrd(const PdCoeffs<R>& c, const R& mf) {
    const R d = mf - c.f_ring;
    const R s2 = d * d + c.sigma * c.sigma;
    return (R(-2.0) * d * c.sigma * c.gamma[0] / s2 - c.gamma[1] * c.gamma[0]) /
           (m_exp(c.gamma[1] * d / c.sigma) * s2);
}

// -Psi + t0 M f but for its constant, without the inspiral's leading term
// unless leading (core.py:261-268)
template <class R>
WFX_HD inline R pd_phase_mf(const PdCoeffs<R>& c, const R& mf, bool leading) {
    R inspiral = pd_ins(c, mf);
    if (!leading) inspiral = inspiral - c.v[0] * ipow(m_cbrt(R(kPi) * mf), -5);
    const R f1 = R(kPdMfPhaseIns), f2 = R(0.5) * c.f_ring;
    R psi;
    if (mf < f1)
        psi = inspiral;
    else if (mf >= f2)
        psi = pd_mrd(c, mf) + c.c1_mrd + c.c2_mrd * mf;
    else
        psi = pd_inter(c, mf) + c.c1_int + c.c2_int * mf;
    return -psi + c.t0 * mf;
}

// The coefficients of masses x1, x2 relative to the total (core.py:154-258)
template <class R>
WFX_HD inline void pd_coefficients(const R& x1, const R& x2, const R& chi1, const R& chi2, PdCoeffs<R>& c) {
    const R eta = x1 * x2;
    const R delta = x1 - x2;
    c.eta = eta;
    c.delta = delta;
    const R xi = pd_chi_pn(eta, delta, chi1, chi2) - R(1.0);
    R fit[kFitCount];
    for (int k = 0; k < kFitCount; ++k) fit[k] = pd_fit(eta, xi, k);
    c.final_spin = pd_final_spin(eta, delta, chi1, chi2);
    const R t = R(2.0) * m_cbrt(R(0.5) * (R(1.0) - c.final_spin)) - R(1.0);
    const R ring = pd_clenshaw(kPdFring, t), damp = pd_clenshaw(kPdFdamp, t);
    const R remnant = R(1.0) - pd_radiated_energy(eta, delta, chi1, chi2);
    c.f_ring = ring / remnant;
    c.f_damp = damp / remnant;

    // TaylorF2 (pn.py:83-91) but for its 3PN spin-spin term, and the
    // pseudo-PN terms of the fits
    const R scale = R(3.0) / (R(128.0) * eta);
    const R v5 = pd_phase_5(eta, x1, x2, chi1, chi2);
    c.v[0] = scale * R(1.0);
    c.v[1] = scale * R(0.0);
    c.v[2] = scale * (R(55.0) * eta / R(9.0) + R(3715.0 / 756.0));
    c.v[3] = scale * pd_phase_3(x1, x2, chi1, chi2);
    c.v[4] = scale * pd_phase_4(eta, x1, x2, chi1, chi2);
    c.v[5] = scale * v5;
    c.v[6] = scale * pd_phase_6(eta, x1, x2, chi1, chi2);
    c.v[7] = scale * pd_phase_7(eta, x1, x2, chi1, chi2);
    c.l5 = scale * R(3.0) * v5;
    c.l6 = scale * R(kPdLog6);
    c.v[5] = c.v[5] - R(kPi / 4.0);
    c.v[6] = c.v[6] - pd_spin_spin_3pn(eta, x1, x2, chi1, chi2) * c.v[0];
    c.v[8] = fit[kFitSigma1] / eta / R(kPi);
    c.v[9] = fit[kFitSigma2] * R(0.75) / eta / R(kPdPi43);
    c.v[10] = fit[kFitSigma3] * R(0.6) / eta / R(kPdPi53);
    c.v[11] = fit[kFitSigma4] * R(0.5) / eta / R(kPdPi2);
    for (int k = 0; k < 3; ++k) c.beta[k] = fit[kFitBeta1 + k];
    for (int k = 0; k < 5; ++k) c.alpha[k] = fit[kFitAlpha1 + k];

    const R f1 = R(kPdMfPhaseIns), f2 = R(0.5) * c.f_ring;
    c.c2_int = pd_dins(c, f1) - pd_dinter(c, f1);
    c.c1_int = pd_ins(c, f1) - pd_inter(c, f1) - c.c2_int * f1;
    const R joined = pd_inter(c, f2) + c.c1_int + c.c2_int * f2;
    c.c2_mrd = c.c2_int + pd_dinter(c, f2) - pd_dmrd(c, f2);
    c.c1_mrd = joined - pd_mrd(c, f2) - c.c2_mrd * f2;

    // the amplitude (core.py:128-151, 214-258)
    {
        const R eta2 = eta * eta;
        const R eta3 = eta2 * eta;
        const R chi12 = chi1 * chi1, chi22 = chi2 * chi2;
        const R d1 = R(1.0) + delta;
        const R pi = R(kPi);
        c.rho[0] = R(1.0);
        c.rho[1] = R(0.0);
        c.rho[2] = (R(-969.0) + R(1804.0) * eta) / R(672.0);
        c.rho[3] = (chi1 * (R(81.0) * d1 - R(44.0) * eta) + chi2 * (R(81.0) - R(81.0) * delta - R(44.0) * eta)) /
                   R(48.0);
        c.rho[4] = (R(-27312085.0) - R(10287648.0) * chi22 - R(10287648.0) * chi12 * d1 +
                    R(10287648.0) * chi22 * delta +
                    R(24.0) * (R(-1975055.0) + R(857304.0) * chi12 - R(994896.0) * chi1 * chi2 + R(857304.0) * chi22) *
                        eta +
                    R(35371056.0) * eta2) /
                   R(8.128512e6);
        c.rho[5] = (chi2 * (R(-285197.0) * (R(-1.0) + delta) + R(4.0) * (R(-91902.0) + R(1579.0) * delta) * eta -
                            R(35632.0) * eta2) +
                    chi1 * (R(285197.0) * d1 - R(4.0) * (R(91902.0) + R(1579.0) * delta) * eta - R(35632.0) * eta2) +
                    R(42840.0) * (R(-1.0) + R(4.0) * eta) * pi) /
                   R(32256.0);
        c.rho[6] = -(R(-336.0) *
                         (R(-3248849057.0) + R(2943675504.0) * chi12 - R(3339284256.0) * chi1 * chi2 +
                          R(2943675504.0) * chi22) *
                         eta2 -
                     R(324322727232.0) * eta3 -
                     R(7.0) * (R(-177520268561.0) + R(107414046432.0) * chi22 + R(107414046432.0) * chi12 * d1 -
                               R(107414046432.0) * chi22 * delta +
                               R(11087290368.0) * (chi1 + chi2 + chi1 * delta - chi2 * delta) * pi) +
                     R(12.0) * eta *
                         (R(-545384828789.0) - R(176491177632.0) * chi1 * chi2 + R(202603761360.0) * chi22 +
                          R(77616.0) * chi12 * (R(2610335.0) + R(995766.0) * delta) -
                          R(77287373856.0) * chi22 * delta + R(5841690624.0) * (chi1 + chi2) * pi +
                          R(21384760320.0) * pi * pi)) /
                   R(6.0085960704e10);
    }
    c.rho[7] = fit[kFitRho1] / R(kPdPi73);
    c.rho[8] = fit[kFitRho2] / R(kPdPi83);
    c.rho[9] = fit[kFitRho3] / R(kPdPi3);
    for (int k = 0; k < 3; ++k) c.gamma[k] = fit[kFitGamma1 + k];
    c.collocation = fit[kFitCollocation];
    const R g1 = c.gamma[0], g2 = c.gamma[1], g3 = c.gamma[2];
    const bool below = value_of(g2) <= 1.0;
    // (LAL drops the square root for gamma2 > 1; its argument is kept
    // positive there, as the reference does, for finite derivatives)
    const R root = m_sqrt(below ? R(1.0) - g2 * g2 : R(1.0));
    c.f_peak = m_abs(below ? c.f_ring + c.f_damp * (root - R(1.0)) * g3 / g2 : c.f_ring - c.f_damp * g3 / g2);
    c.t0 = pd_dmrd(c, c.f_peak);
    c.sigma = c.f_damp * g3;

    // the intermediate amplitude sum_j delta_j y^j, y = M f / f_peak:
    // values at kPdMfAmpIns, halfway and f_peak, derivatives at the ends
    const R y1 = R(kPdMfAmpIns) / c.f_peak;
    const R y2 = R(0.5) * (R(1.0) + y1);
    R a[5][5] = {{R(1.0), y1, ipow(y1, 2), ipow(y1, 3), ipow(y1, 4)},
                 {R(1.0), y2, ipow(y2, 2), ipow(y2, 3), ipow(y2, 4)},
                 {R(1.0), R(1.0), R(1.0), R(1.0), R(1.0)},
                 {R(0.0), R(1.0), R(2.0) * y1, R(3.0) * ipow(y1, 2), R(4.0) * ipow(y1, 3)},
                 {R(0.0), R(1.0), R(2.0), R(3.0), R(4.0)}};
    R b[5] = {pd_amp_ins(c, R(kPdMfAmpIns)), c.collocation, pd_amp_mrd(c, c.f_peak),
              pd_damp_ins(c, R(kPdMfAmpIns)) * c.f_peak, pd_damp_mrd(c, c.f_peak) * c.f_peak};
    pivoted_solve(a, b);
    for (int j = 0; j < 5; ++j) c.delta_int[j] = b[j];
}

// ------------------------------------------------------ frequency powers

template <class R>
struct PdPowers {
    R f, t, m1, log_f, m76, f34;
    double m53;
};

// core.py:395-407
template <class R>
WFX_HD inline PdPowers<R> pd_powers(const R& f) {
    PdPowers<R> fp;
    fp.f = f;
    fp.m1 = R(1.0) / f;
    fp.log_f = lead_log(f);
    R t = lead_exp(fp.log_f / R(3.0));
    fp.t = t * (R(4.0) - t * t * t * fp.m1) / R(3.0);
    const R root = m_sqrt(f);
    fp.m76 = fp.m1 / m_sqrt(fp.t);
    fp.f34 = root * m_sqrt(root);
    fp.m53 = lead_exp(-5.0 / 3.0 * lead_log(value_of(f)));
    return fp;
}

// The phase but for the leading term and -A (core.py:425-440), the
// reference's order of terms; inspiral = f < f_phase_ins.
template <class R, class C>
WFX_HD inline R pd_phase_rest(const C* c, const PdPowers<R>& fp, bool& inspiral) {
    const R f = fp.f;
    inspiral = f < c[kPdFPhaseIns];
    const bool mrd = f >= c[kPdFPhaseMrd];
    const R t2 = fp.t * fp.t;
    const R mt1 = t2 * fp.m1;
    if (inspiral) {
        R total = R(c[kPdInsConst]);
        total = total + R(c[kPdInsM1]) * fp.m1;
        total = total + R(c[kPdInsM23]) * (mt1 * mt1);
        total = total + R(c[kPdInsM13]) * mt1;
        total = total + R(c[kPdInsLog]) * fp.log_f;
        total = total + R(c[kPdInsP13]) * fp.t;
        total = total + R(c[kPdInsP13L]) * (fp.t * fp.log_f);
        total = total + R(c[kPdInsP23]) * t2;
        total = total + R(c[kPdInsP1]) * f;
        total = total + R(c[kPdInsP43]) * (f * fp.t);
        total = total + R(c[kPdInsP53]) * (f * t2);
        total = total + R(c[kPdInsP2]) * (f * f);
        return total;
    }
    if (mrd) {
        R total = R(c[kPdMrdConst]);
        total = total + R(c[kPdMrdP1]) * f;
        total = total + R(c[kPdMrdM1]) * fp.m1;
        total = total + R(c[kPdMrdP34]) * fp.f34;
        return total + R(c[kPdAtanA]) * m_atan(R(c[kPdAtanW]) * (f - R(c[kPdAtanShift])));
    }
    R total = R(c[kPdIntConst]);
    total = total + R(c[kPdIntP1]) * f;
    total = total + R(c[kPdIntM3]) * (fp.m1 * fp.m1 * fp.m1);
    total = total + R(c[kPdIntLog]) * fp.log_f;
    return total;
}

// ----------------------------------------------------------------- record

// The record of one binary (core.py:274-377): params {m1, m2, chi1, chi2,
// d_L, inclination, phi_ref, f_ref}, the grid's first frequency f_min,
// reference_mf NaN for none; T0 and A land in rec[kT0], rec[kAnchor].
// Returns LAL's flags (bit 3: final spin above one).
template <class R>
WFX_HD inline unsigned pd_record(const R* params, const R& f_min, double amp_scale, double reference_mf, R* rec) {
    const R m1 = params[kMass1], m2 = params[kMass2], chi1 = params[kChi1], chi2 = params[kChi2];
    const R distance = params[kDistance], inclination = params[kInclination], phi_ref = params[kPhiRef];
    const R f_ref_in = params[kFRef];
    const R total = m1 + m2;
    PdCoeffs<R> c;
    pd_coefficients(m1 / total, m2 / total, chi1, chi2, c);
    const R eta = c.eta;
    const R s = total * R(kMtsunSI);
    const R log_s = m_log(s);

    // -Psi and the time shift t0 M f in powers of f (core.py:293-333):
    // rows[r][basis index of phenomd/core.py:50-53]
    R rows[3][14];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 14; ++k) rows[r][k] = R(0.0);
    const R log_pis = m_log(R(kPi) * s);
    const R u = m_cbrt(R(kPi) * s);
    R u_pow[10];
    u_pow[0] = R(1.0);
    u_pow[1] = u;
#pragma unroll
    for (int k = 2; k < 10; ++k) u_pow[k] = u_pow[k - 1] * u;
    const R u_inv = R(1.0) / u;
    const R u_pow_inv[4] = {R(1.0), u_inv, u_inv * u_inv, u_inv * u_inv * u_inv};
    // the basis index of f^((k-5)/3) and, for k = 5, 6, of its log term
    const int index[12] = {-1, -1, 0, 1, 2, 4, 5, 7, 8, 9, 10, 11};
#pragma unroll
    for (int k = 2; k < 12; ++k) {
        const R scale = k >= 5 ? u_pow[k - 5] : u_pow_inv[5 - k];
        R coefficient = c.v[k];
        if (k == 5 || k == 6) {
            const R l = k == 5 ? c.l5 : c.l6;
            coefficient = coefficient + l * log_pis / R(3.0);
            const int li = k == 5 ? 3 : 6;
            rows[0][li] = rows[0][li] + -l / R(3.0) * scale;
        }
        rows[0][index[k]] = rows[0][index[k]] + -coefficient * scale;
    }
    const R b1 = c.beta[0], b2 = c.beta[1], b3 = c.beta[2];
    rows[1][8] = rows[1][8] + -(b1 / eta + c.c2_int) * s;
    rows[1][12] = rows[1][12] + b3 / (R(3.0) * eta) / ipow(s, 3);
    rows[1][3] = rows[1][3] + -b2 / eta;
    rows[1][4] = rows[1][4] + (-b2 / eta * log_s - c.c1_int);
    const R a1 = c.alpha[0], a2 = c.alpha[1], a3 = c.alpha[2], a4 = c.alpha[3], a5 = c.alpha[4];
    rows[2][8] = rows[2][8] + -(a1 / eta + c.c2_mrd) * s;
    rows[2][0] = rows[2][0] + a2 / eta / s;
    rows[2][13] = rows[2][13] + R(-4.0 / 3.0) * a3 / eta * m_sqrt(s) * m_sqrt(m_sqrt(s));
    rows[2][4] = rows[2][4] + -c.c1_mrd;
#pragma unroll
    for (int r = 0; r < 3; ++r) rows[r][8] = rows[r][8] + c.t0 * s;

#pragma unroll
    for (int k = 0; k < kSlots; ++k) rec[k] = R(0.0);
    const int ins_basis[12] = {4, 0, 1, 2, 3, 5, 6, 7, 8, 9, 10, 11};
#pragma unroll
    for (int k = 0; k < 12; ++k) rec[kPdInsConst + k] = rows[0][ins_basis[k]];
    rec[kPdIntConst] = rows[1][4];
    rec[kPdIntP1] = rows[1][8];
    rec[kPdIntM3] = rows[1][12];
    rec[kPdIntLog] = rows[1][3];
    rec[kPdMrdConst] = rows[2][4];
    rec[kPdMrdP1] = rows[2][8];
    rec[kPdMrdM1] = rows[2][0];
    rec[kPdMrdP34] = rows[2][13];
    rec[kPdAtanA] = -a4 / eta;
    rec[kPdAtanW] = s / c.f_damp;
    rec[kPdAtanShift] = a5 * c.f_ring / s;
    rec[kPdFPhaseIns] = R(kPdMfPhaseIns) / s;
    rec[kPdFPhaseMrd] = R(0.5) * c.f_ring / s;

    // the reference phase: phi(f_ref) = 2 phi_ref (core.py:340-358)
    bool invalid = false, inspiral;
    R rest, f_ref64;
    if (reference_mf != reference_mf) {
        const R f_ref = value_of(f_ref_in) == 0.0 ? f_min : f_ref_in;
        rest = pd_phase_rest(rec, pd_powers(f_ref), inspiral);
        f_ref64 = f_ref;
    } else {
        invalid = value_of(f_ref_in) != 0.0;
        inspiral = reference_mf < kPdMfPhaseIns;
        rest = pd_phase_mf(c, R(reference_mf), false);
        f_ref64 = R(reference_mf) / (R(kMtsunSI) * (m1 + m2));
    }
    const R shift = R(2.0) * phi_ref - rest;
    rec[kPdInsConst] = rec[kPdInsConst] + shift;
    rec[kPdIntConst] = rec[kPdIntConst] + shift;
    rec[kPdMrdConst] = rec[kPdMrdConst] + shift;

    // the leading term (as IMRPhenomXAS's, xas/core.py:342-363)
    const R t0 = R(kPdK0) * lead_cbrt(m1 + m2) / (m1 * m2);
    rec[kT0] = t0;
    rec[kAnchor] = inspiral ? t0 * lead_exp(R(-5.0 / 3.0) * lead_log(f_ref64)) : R(0.0);

    // the amplitude
    const R prefactor = R(kPdKAmp) * R(amp_scale) * m_sqrt(m1 * m2) * m_exp(-m_log(total) / R(6.0)) / distance;
#pragma unroll
    for (int k = 0; k < 10; ++k) rec[kPdAmpIns + k] = prefactor * c.rho[k] * u_pow[k];
#pragma unroll
    for (int j = 0; j < 5; ++j) rec[kPdAmpInt + j] = prefactor * c.delta_int[j];
    const R g1 = c.gamma[0], g2 = c.gamma[1], g3 = c.gamma[2];
    const R width = g3 * c.f_damp;
    rec[kPdAmpIntS] = s / c.f_peak;
    rec[kPdAmpRdN] = prefactor * g1 / width;
    rec[kPdAmpRdGr] = g2;
    rec[kPdAmpRdS] = s / width;
    rec[kPdAmpRdShift] = c.f_ring / s;
    rec[kPdFAmpIns] = R(kPdMfAmpIns) / s;
    rec[kPdFAmpMrd] = c.f_peak / s;

    const unsigned flags = value_of(c.final_spin) > 1.0 ? 8u : 0u;
    const bool error = flags != 0 || invalid;
    const double nan = __builtin_nan("");
    const R cos_i = m_cos(inclination);
    rec[kPdKp] = error ? R(nan) : R(0.5) * (R(1.0) + cos_i * cos_i);
    rec[kPdKc] = error ? R(nan) : -cos_i;
    return flags;
}

// IMRPhenomD's phase -Psi + t0 M f at M f = mf but for its constant
// (reference.py:18-22), for the reference-convention converters
template <class R>
WFX_HD inline R pd_reference_phase(const R* params, const R& mf) {
    const R total = params[kMass1] + params[kMass2];
    PdCoeffs<R> c;
    pd_coefficients(params[kMass1] / total, params[kMass2] / total, params[kChi1], params[kChi2], c);
    return pd_phase_mf(c, mf, true);
}

}  // namespace wfx
