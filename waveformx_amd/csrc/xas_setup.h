// Per-binary setup of IMRPhenomXAS: the waveform variables, amplitude and
// phase coefficients (three pivoted collocation solves), the connection
// constants, the time shift and the reference-phase alignment, folded into
// the 64-slot record of xas_layout.h.
//
// One thread per binary (per binary and derivative direction for the dual
// scalar). Generic over the scalar R: double (FP64 and split precision),
// float (plain FP32) or Dual<N> (derivatives of the record).
//
// Follows src/waveformx/models/xpnr/internals.py:37-170 (waveform
// variables), xpnr/amp22.py:25-129, xpnr/phase22.py:39-326 and
// xas/core.py:221-379 (QNM in 1 - a_f, the table); the fits come from
// xas_fits.inc (transcribed from the reference's generated modules).
#pragma once

#include <type_traits>

#include "wfx_math.h"
#include "xas_layout.h"

namespace wfx {

// IMRPhenomX_UsefulPowers (xpnr/_math.py:190-246)
template <class R>
struct Powers {
    R one_sixth, m_one_sixth, itself, m_one, one_third, m_one_third, two_thirds, m_two_thirds, four_thirds,
        m_four_thirds, five_thirds, m_five_thirds, seven_thirds, m_seven_thirds, eight_thirds, m_eight_thirds, two,
        three, four, five, m_two, m_three, m_four, m_five, seven_sixths, m_seven_sixths, log, sqrt;
};

// the powers of x from x^(1/3), 1/x and log x
template <class R>
WFX_HD inline Powers<R> powers_from(const R& x, const R& one_third, const R& m_one, const R& log) {
    Powers<R> p;
    p.itself = x;
    p.one_third = one_third;
    p.one_sixth = m_sqrt(p.one_third);
    p.m_one = m_one;
    p.log = log;
    p.two_thirds = p.one_third * p.one_third;
    p.m_one_third = p.two_thirds * p.m_one;
    p.m_one_sixth = p.one_sixth * p.m_one_third;
    p.four_thirds = p.two_thirds * p.two_thirds;
    p.five_thirds = p.four_thirds * p.one_third;
    p.seven_thirds = p.four_thirds * x;
    p.eight_thirds = p.seven_thirds * p.one_third;
    p.m_two_thirds = p.m_one_third * p.m_one_third;
    p.m_four_thirds = p.m_two_thirds * p.m_two_thirds;
    p.m_five_thirds = p.m_four_thirds * p.m_one_third;
    p.m_seven_thirds = p.m_four_thirds * p.m_one;
    p.m_eight_thirds = p.m_seven_thirds * p.m_one_third;
    p.two = x * x;
    p.three = p.two * x;
    p.m_two = p.m_one * p.m_one;
    p.m_three = p.m_two * p.m_one;
    p.m_four = p.m_two * p.m_two;
    p.four = p.two * p.two;
    p.five = p.three * p.two;
    p.m_five = p.m_three * p.m_two;
    p.seven_sixths = p.one_sixth * x;
    p.m_seven_sixths = p.m_one_sixth * p.m_one;
    p.sqrt = p.one_sixth * p.one_third;
    return p;
}

template <class R>
WFX_HD inline Powers<R> useful_powers(const R& x) {
    return powers_from(x, m_cbrt(x), R(1.0) / x, m_log(x));
}

// IMRPhenomXSetWaveformVariables' fields used by the (2,2) mode
template <class R>
struct WaveformVars {
    R eta, delta, eta2, eta3, chi1l, chi2l, chi1l2, chi1l3, chi2l2, chi_eff, chi_pn_hat, s_tot_r, dchi, f_ref, phi0,
        m_sec, f_cut, amp0, amp_norm, dphase0, m_final, a_final, f_ring, f_damp, f_meco, f_isco, q, dchi_half;
    unsigned flags;
};

template <class R>
struct AmpCoeffs {
    R f_amp_match_in, f_amp_rd_min, pn_initial, pn_one_third, pn_two_thirds, pn_three_thirds, pn_four_thirds,
        pn_five_thirds, pn_six_thirds, rho1, rho2, rho3, gamma1, gamma2, gamma3, gamma_r, gamma_d2, gamma_d13, f_ring,
        f_damp, delta0, delta1, delta2, delta3, delta4;
};

template <class R>
struct PhaseCoeffs {
    R f_phase_match_in, f_phase_match_im, c0, c1, c2, c4, c_l, c_l_gr, c4_ov3, c_l_ov_fda, b0, b1, b2, b3, b4, a0,
        a1, a2, a3, sigma1, sigma2, sigma3, sigma4, phi0, phi2, phi3, phi4, phi5l, phi6, phi6l, phi7, phi8, phi8l,
        phi_norm, dphi0, dphi2, dphi3, dphi4, dphi5, dphi6, dphi6l, dphi7, dphi8, dphi8l, f_ring, f_damp, dphase_norm,
        c1_int, c2_int, c1_mrd, c2_mrd;
};

#include "xas_fits.inc"

// (xas/core.py:214-229) a fit of the final spin in powers of u = 1 - a_f
template <class R, int N>
WFX_HD inline R horner_u(const double (&c)[N], const R& u) {
    R total = R(c[N - 1]);
    for (int k = N - 2; k >= 0; --k) total = total * u + R(c[k]);
    return total;
}

// xpnr/internals.py:37-150 for a frequency sequence (f_max = 0), masses in
// solar masses, then the QNM frequencies of xas/core.py:270-278
template <class R>
WFX_HD inline WaveformVars<R> waveform_vars(const R* p, const R& f_min, double amp_scale) {
    WaveformVars<R> wf;
    R m1_in = p[kMass1], m2_in = p[kMass2];
    R mass_ratio = m1_in > m2_in ? m1_in / m2_in : m2_in / m1_in;
    bool swap = m1_in < m2_in;
    R m1 = swap ? m2_in : m1_in;
    R m2 = swap ? m1_in : m2_in;
    R chi1l = swap ? p[kChi2] : p[kChi1];
    R chi2l = swap ? p[kChi1] : p[kChi2];

    R delta = m_abs((m1 - m2) / (m1 + m2));
    R eta = m_abs(R(0.25) * m_fnms(delta, delta, R(1.0)));
    eta = m_min(eta, R(0.25));
    R m_tot = m1 + m2;
    R m_sec = m_tot * R(kMtsunSI);
    R chi_eff = ut_chi_eff(eta, chi1l, chi2l);

    R f_ref = p[kFRef] == R(0.0) ? f_min : p[kFRef];
    R f_cut_def = chi_eff > R(0.99) ? R(0.33) : R(0.3);
    R f_cut = f_cut_def / m_sec;

    wf.eta = eta;
    wf.delta = delta;
    wf.eta2 = eta * eta;
    wf.eta3 = eta * (eta * eta);
    wf.chi1l = chi1l;
    wf.chi2l = chi2l;
    wf.chi1l2 = chi1l * chi1l;
    wf.chi1l3 = chi1l * chi1l * chi1l;
    wf.chi2l2 = chi2l * chi2l;
    wf.chi_eff = chi_eff;
    wf.chi_pn_hat = ut_chi_pn_hat(eta, chi1l, chi2l);
    wf.s_tot_r = ut_s_tot_r(eta, chi1l, chi2l);
    wf.dchi = ut_dchi(chi1l, chi2l);
    wf.dchi_half = wf.dchi * R(0.5);
    wf.q = eta == R(0.25) ? R(1.0) : mass_ratio;
    wf.f_ref = f_ref;
    wf.phi0 = p[kPhiRef];
    wf.m_sec = m_sec;
    wf.f_cut = f_cut;  // f_max_prime for f_max = 0
    wf.amp0 = m_tot * R(kMrsunSI / kMpcSI) * (m_tot * R(kMtsunSI)) * R(amp_scale) / p[kDistance];
    wf.amp_norm = R(0.816496580927726) * m_sqrt(eta) * R(lalpi_m_one_sixth);  // sqrt(2/3)
    wf.dphase0 = R(0.005796647796902313);                                     // 5 / (128 pi^(5/3))
    wf.m_final = ut_final_mass_2017(eta, chi1l, chi2l);
    wf.a_final = ut_final_spin_2017(eta, chi1l, chi2l);
    wf.f_meco = ut_f_meco(eta, chi1l, chi2l);
    wf.f_isco = ut_f_isco(wf.a_final);
    R u = R(1.0) - wf.a_final;
    wf.f_ring = horner_u(kQnm_fring22_numerator, u) / horner_u(kQnm_fring22_denominator, u) / wf.m_final;
    wf.f_damp = horner_u(kQnm_fdamp22_numerator, u) / horner_u(kQnm_fdamp22_denominator, u) / wf.m_final;

    double q = value_of(mass_ratio);
    unsigned flags = 0;
    if (q > 1000.0 && fabs(q - 1000.0) > 1e-12) flags |= kFlagMassRatio;
    if (fabs(value_of(chi1l)) > 1.0 || fabs(value_of(chi2l)) > 1.0) flags |= kFlagSpin;
    if (value_of(wf.m_final) > 1.0) flags |= kFlagFinalMass;
    if (fabs(value_of(wf.a_final)) > 1.0) flags |= kFlagFinalSpin;
    if (value_of(wf.f_meco) > value_of(wf.f_isco)) flags |= kFlagMeco;
    wf.flags = flags;
    return wf;
}

// xpnr/amp22.py:25-129 (the co-precessing deviations are zero here)
template <class R>
WFX_HD inline AmpCoeffs<R> amp_coeffs(const WaveformVars<R>& wf) {
    AmpCoeffs<R> amp;
    const R eta = wf.eta, delta = wf.delta, f_ring = wf.f_ring, f_damp = wf.f_damp;
    const R s = wf.s_tot_r, dchi = wf.dchi;

    R gamma2 = rd_amp_22_gamma2(eta, s, dchi, delta);
    R gamma3 = rd_amp_22_gamma3(eta, s, dchi, delta);
    R f_rd = rd_amp_22_peak_frequency(gamma2, gamma3, f_ring, f_damp);
    R v1_rd = rd_amp_22_v1(eta, s, dchi, delta);
    R fdg3 = f_damp * gamma3;
    R gamma1 = (v1_rd / fdg3) *
               m_fma(f_damp * f_damp * gamma3, gamma3,
                     m_fma(f_ring, f_ring, m_fma(f_rd, f_rd, -(R(2.0) * f_rd * f_ring)))) *
               m_exp(((f_rd - f_ring) * gamma2) / fdg3);
    amp.gamma1 = gamma1;
    amp.gamma2 = gamma2;
    amp.gamma3 = gamma3;
    amp.gamma_r = gamma2 / fdg3;
    amp.gamma_d2 = (gamma3 * f_damp) * (gamma3 * f_damp);
    amp.gamma_d13 = f_damp * gamma1 * gamma3;
    amp.f_ring = f_ring;
    amp.f_damp = f_damp;
    amp.f_amp_rd_min = f_rd;
    R f_in = wf.f_meco + R(0.25) * (wf.f_isco - wf.f_meco);
    amp.f_amp_match_in = f_in;

    amp.pn_initial = R(1.0);
    amp.pn_one_third = R(0.0);
    amp_pn_coefficients(wf.chi1l, wf.chi1l2, wf.chi1l3, wf.chi2l, wf.chi2l2, delta, eta, wf.eta2, wf.eta3,
                        amp.pn_two_thirds, amp.pn_three_thirds, amp.pn_four_thirds, amp.pn_five_thirds,
                        amp.pn_six_thirds);

    // pseudo-PN coefficients from the collocation values at 1/2, 3/4 and 1
    // of the inspiral region
    R fc2 = R(0.50) * f_in, fc3 = R(0.75) * f_in, fc4 = R(1.00) * f_in;
    R vc2 = ins_amp_22_v2(eta, wf.chi_pn_hat, dchi, delta);
    R vc3 = ins_amp_22_v3(eta, wf.chi_pn_hat, dchi, delta);
    R vc4 = ins_amp_22_v4(eta, wf.chi_pn_hat, dchi, delta);
    amp.rho1 = ins_amp_22_rho1(vc2, vc3, vc4, fc2, fc3, fc4);
    amp.rho2 = ins_amp_22_rho2(vc2, vc3, vc4, fc2, fc3, fc4);
    amp.rho3 = ins_amp_22_rho3(vc2, vc3, vc4, fc2, fc3, fc4);

    // the intermediate from the values and slopes at both ends
    Powers<R> p1 = useful_powers(f_in);
    Powers<R> p4 = useful_powers(f_rd);
    R d1 = ins_amp_22_dansatz(f_in, wf, amp);
    R d4 = rd_amp_22_dansatz(f_rd, f_ring, f_damp, amp);
    R ins_f1 = ins_amp_22_ansatz_exact(f_in, p1, amp);
    R rd_f4 = rd_amp_22_ansatz_exact(f_rd, f_ring, f_damp, amp);
    d1 = ((R(7.0 / 6.0) * p1.one_sixth) / ins_f1) - ((p1.seven_sixths * d1) / (ins_f1 * ins_f1));
    d4 = ((R(7.0 / 6.0) * p4.one_sixth) / rd_f4) - ((p4.seven_sixths * d4) / (rd_f4 * rd_f4));
    R f2 = f_in + R(0.5) * (f_rd - f_in);
    R v1 = R(1.0) / (p1.m_seven_sixths * ins_f1);
    R v2 = R(1.0) / itm_amp_22_va(eta, s, dchi, delta);
    R v4 = R(1.0) / (p4.m_seven_sixths * rd_f4);
    amp.delta0 = itm_amp_22_delta0(d1, d4, v1, v2, v4, f_in, f2, f_rd);
    amp.delta1 = itm_amp_22_delta1(d1, d4, v1, v2, v4, f_in, f2, f_rd);
    amp.delta2 = itm_amp_22_delta2(d1, d4, v1, v2, v4, f_in, f2, f_rd);
    amp.delta3 = itm_amp_22_delta3(d1, d4, v1, v2, v4, f_in, f2, f_rd);
    amp.delta4 = itm_amp_22_delta4(d1, d4, v1, v2, v4, f_in, f2, f_rd);
    return amp;
}

// A x = b by Gaussian elimination with partial pivoting (largest |pivot|,
// first on ties), the algorithm behind phase22.py:44-49's
// jnp.linalg.solve; x overwrites b.
//
// For plain floating-point R the row exchange is written as selects over
// the candidate rows, so that every index is a compile-time constant once
// the loops unroll (an exchange through the runtime pivot index puts the
// arrays in scratch memory). The arithmetic, and its result, is that of the
// plain exchange. Dual numbers solve for their values the same way and for
// their tangents from the values' factors (pivoted_solve_implicit); other
// scalars (Taylor jets) keep the plain exchange.
template <class R, int N>
WFX_HD inline void pivoted_solve_exchange(R (&a)[N][N], R (&b)[N]) {
    for (int k = 0; k < N; ++k) {
        int piv = k;
        double big = fabs(value_of(a[k][k]));
        for (int i = k + 1; i < N; ++i) {
            double m = fabs(value_of(a[i][k]));
            if (m > big) {
                big = m;
                piv = i;
            }
        }
        if (piv != k) {
            for (int j = 0; j < N; ++j) {
                R t = a[k][j];
                a[k][j] = a[piv][j];
                a[piv][j] = t;
            }
            R t = b[k];
            b[k] = b[piv];
            b[piv] = t;
        }
        for (int i = k + 1; i < N; ++i) {
            R l = a[i][k] / a[k][k];
            for (int j = k + 1; j < N; ++j) a[i][j] = a[i][j] - l * a[k][j];
            b[i] = b[i] - l * b[k];
        }
    }
    for (int i = N - 1; i >= 0; --i) {
        R s = b[i];
        for (int j = i + 1; j < N; ++j) s = s - a[i][j] * b[j];
        b[i] = s / a[i][i];
    }
}

template <class R, int N>
WFX_HD inline void pivoted_solve_select(R (&a)[N][N], R (&b)[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
        int piv = k;
        double big = fabs(value_of(a[k][k]));
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            double m = fabs(value_of(a[i][k]));
            if (m > big) {
                big = m;
                piv = i;
            }
        }
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const bool swap = piv == i;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const R t = a[k][j];
                a[k][j] = swap ? a[i][j] : t;
                a[i][j] = swap ? t : a[i][j];
            }
            const R t = b[k];
            b[k] = swap ? b[i] : t;
            b[i] = swap ? t : b[i];
        }
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            R l = a[i][k] / a[k][k];
#pragma unroll
            for (int j = k + 1; j < N; ++j) a[i][j] = a[i][j] - l * a[k][j];
            b[i] = b[i] - l * b[k];
        }
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        R s = b[i];
#pragma unroll
        for (int j = i + 1; j < N; ++j) s = s - a[i][j] * b[j];
        b[i] = s / a[i][i];
    }
}

// Dual numbers: the values by pivoted_solve_select's arithmetic (so the
// same bits as the plain exchange), the multipliers kept in place of the
// eliminated entries and the row exchanges applied to whole rows (P A = L U),
// then each direction's tangent by the implicit-function identity
// A dx = db - dA x with the same factors: no elimination on the tangents,
// and every index a compile-time constant (the arrays stay in registers).
template <int M, int N>
WFX_HD inline void pivoted_solve_implicit(Dual<M> (&a)[N][N], Dual<M> (&b)[N]) {
    double lu[N][N], x[N];
    int pivots[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        x[i] = b[i].v;
#pragma unroll
        for (int j = 0; j < N; ++j) lu[i][j] = a[i][j].v;
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
        int piv = k;
        double big = fabs(lu[k][k]);
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const double m = fabs(lu[i][k]);
            if (m > big) {
                big = m;
                piv = i;
            }
        }
        pivots[k] = piv;
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const bool swap = piv == i;
#pragma unroll
            for (int j = 0; j < N; ++j) {
                const double t = lu[k][j];
                lu[k][j] = swap ? lu[i][j] : t;
                lu[i][j] = swap ? t : lu[i][j];
            }
            const double t = x[k];
            x[k] = swap ? x[i] : t;
            x[i] = swap ? t : x[i];
        }
#pragma unroll
        for (int i = k + 1; i < N; ++i) {
            const double l = lu[i][k] / lu[k][k];
#pragma unroll
            for (int j = k + 1; j < N; ++j) lu[i][j] = lu[i][j] - l * lu[k][j];
            x[i] = x[i] - l * x[k];
            lu[i][k] = l;
        }
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        double s = x[i];
#pragma unroll
        for (int j = i + 1; j < N; ++j) s = s - lu[i][j] * x[j];
        x[i] = s / lu[i][i];
    }
#pragma unroll
    for (int d = 0; d < M; ++d) {
        // r = db - dA x in the original row order, then P, L and U
        double r[N];
#pragma unroll
        for (int i = 0; i < N; ++i) {
            double s = b[i].d[d];
#pragma unroll
            for (int j = 0; j < N; ++j) s = s - a[i][j].d[d] * x[j];
            r[i] = s;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
#pragma unroll
            for (int i = k + 1; i < N; ++i) {
                const bool swap = pivots[k] == i;
                const double t = r[k];
                r[k] = swap ? r[i] : t;
                r[i] = swap ? t : r[i];
            }
        }
#pragma unroll
        for (int i = 1; i < N; ++i)
#pragma unroll
            for (int j = 0; j < i; ++j) r[i] = r[i] - lu[i][j] * r[j];
#pragma unroll
        for (int i = N - 1; i >= 0; --i) {
            double s = r[i];
#pragma unroll
            for (int j = i + 1; j < N; ++j) s = s - lu[i][j] * r[j];
            r[i] = s / lu[i][i];
        }
#pragma unroll
        for (int i = 0; i < N; ++i) b[i].d[d] = r[i];
    }
#pragma unroll
    for (int i = 0; i < N; ++i) b[i].v = x[i];
}

template <class R>
struct IsDualNumber : std::false_type {};
template <int M>
struct IsDualNumber<Dual<M>> : std::true_type {};

template <class R, int N>
WFX_HD inline void pivoted_solve(R (&a)[N][N], R (&b)[N]) {
    if constexpr (std::is_floating_point<R>::value)
        pivoted_solve_select(a, b);
    else if constexpr (IsDualNumber<R>::value)
        pivoted_solve_implicit(a, b);
    else
        pivoted_solve_exchange(a, b);
}

// phase22.py:52-263 then the connections of phase22.py:266-286
template <class R>
WFX_HD inline PhaseCoeffs<R> phase_coeffs(const WaveformVars<R>& wf) {
    PhaseCoeffs<R> ph;
    const R f_ring = wf.f_ring, f_damp = wf.f_damp, dphase0 = wf.dphase0;
    const R eta = wf.eta, delta = wf.delta, s = wf.s_tot_r, dchi = wf.dchi;

    R f_im_match = R(0.6) * (R(0.5) * f_ring + wf.f_isco);
    R f_in_match = wf.f_meco;
    R df = (f_im_match - f_in_match) * R(0.03);
    R f_match_in = f_in_match - R(1.0) * df;
    R f_match_im = f_im_match + R(0.5) * df;
    const R f_ins_min = R(0.0026);
    R f_ins_max = R(1.020) * wf.f_meco;
    R f_rd_min = f_im_match;
    R f_rd_max = m_fma(R(1.25), f_damp, f_ring);

    // Gauss-Chebyshev-Lobatto points on [0, 1]
    const double g5[5] = {0.0, 0.5 - 1.0 / (2.0 * 1.4142135623730951), 0.5, 0.5 + 1.0 / (2.0 * 1.4142135623730951),
                          1.0};
    const double g4[4] = {0.0, 0.25, 0.75, 1.0};

    // ringdown collocation, 5 x 5
    R width = f_rd_max - f_rd_min;
    R rd_v4 = rd_phase_22_v4(eta, s, dchi, delta);
    R vr[5] = {rd_phase_22_d12(eta, s, dchi, delta), rd_phase_22_d24(eta, s, dchi, delta),
               rd_phase_22_d34(eta, s, dchi, delta), rd_v4, rd_phase_22_d54(eta, s, dchi, delta)};
    vr[4] = vr[4] + vr[3];
    vr[2] = vr[2] + vr[3];
    vr[1] = vr[1] + vr[3];
    vr[0] = vr[0] + vr[1];
    R phase_rd = vr[0];
    R mr[5][5];
    for (int i = 0; i < 5; ++i) {
        R x = i == 3 ? f_ring : m_fma(R(g5[i]), width, f_rd_min);
        R inv = R(1.0) / x;
        R inv2 = inv * inv;
        mr[i][0] = R(1.0);
        mr[i][1] = m_cbrt(inv);
        mr[i][2] = inv2;
        mr[i][3] = inv2 * inv2;
        mr[i][4] = -dphase0 / m_fma(f_damp, f_damp, (x - f_ring) * (x - f_ring));
    }
    pivoted_solve(mr, vr);
    ph.c0 = vr[0];
    ph.c1 = vr[1];
    ph.c2 = vr[2];
    ph.c4 = vr[3];
    R c_l = -(dphase0 * vr[4]);

    // inspiral pseudo-PN collocation, 4 x 4
    width = f_ins_max - f_ins_min;
    R v3 = ins_phase_22_v3(eta, wf.chi_pn_hat, dchi, delta);
    R vi[4] = {ins_phase_22_d13(eta, wf.chi_pn_hat, dchi, delta) + v3,
               ins_phase_22_d23(eta, wf.chi_pn_hat, dchi, delta) + v3, v3,
               ins_phase_22_d43(eta, wf.chi_pn_hat, dchi, delta) + v3};
    R mi[4][4];
    for (int i = 0; i < 4; ++i) {
        R x = m_fma(R(g4[i]), width, f_ins_min);
        R x13 = m_cbrt(x);
        mi[i][0] = R(1.0);
        mi[i][1] = x13;
        mi[i][2] = x13 * x13;
        mi[i][3] = x;
    }
    pivoted_solve(mi, vi);
    ph.a0 = vi[0];
    ph.a1 = vi[1];
    ph.a2 = vi[2];
    ph.a3 = vi[3];
    ph.sigma1 = R(-5.0 / 3.0) * vi[0];
    ph.sigma2 = R(-5.0 / 4.0) * vi[1];
    ph.sigma3 = R(-5.0 / 5.0) * vi[2];
    ph.sigma4 = R(-5.0 / 6.0) * vi[3];

    // TaylorF2 3.5PN (+ spin) coefficients
    R chi1l = wf.chi1l, chi2l = wf.chi2l;
    R chi1l2l = chi1l * chi2l, chi1l2 = chi1l * chi1l, chi1l3 = chi1l * chi1l2;
    R chi2l2 = chi2l * chi2l, chi2l3 = chi2l * chi2l2;
    R eta2 = eta * eta, eta3 = eta * eta2;
    R phi2_ns, phi3_ns, phi3_s, phi4_ns, phi4_s, phi5l_ns, phi5l_s, phi6_ns, phi6_s, phi6l_ns, phi7_ns, phi7_s,
        phi8_s, phi8l_s;
    taylorf2_coefficients(chi1l, chi1l2, chi1l2l, chi1l3, chi2l, chi2l2, chi2l3, delta, eta, eta2, eta3, phi2_ns,
                          phi3_ns, phi3_s, phi4_ns, phi4_s, phi5l_ns, phi5l_s, phi6_ns, phi6_s, phi6l_ns, phi7_ns,
                          phi7_s, phi8_s, phi8l_s);
    ph.phi0 = R(1.0);
    ph.phi2 = phi2_ns;
    ph.phi3 = phi3_ns + phi3_s;
    ph.phi4 = phi4_ns + phi4_s;
    ph.phi5l = phi5l_ns + phi5l_s;
    ph.phi6 = phi6_ns + phi6_s;
    ph.phi6l = phi6l_ns;
    ph.phi7 = phi7_ns + phi7_s;
    ph.phi8 = phi8_s;
    ph.phi8l = phi8l_s;
    ph.phi_norm = R(-(3.0 * lalpi_m_five_thirds) / 128.0);
    ph.dphi0 = ph.phi0;
    ph.dphi2 = R(3.0 / 5.0) * ph.phi2;
    ph.dphi3 = R(2.0 / 5.0) * ph.phi3;
    ph.dphi4 = R(1.0 / 5.0) * ph.phi4;
    ph.dphi5 = R(-3.0 / 5.0) * ph.phi5l;
    ph.dphi6 = m_fma(R(-(1.0 / 5.0)), ph.phi6, -(R(3.0 / 5.0) * ph.phi6l));
    ph.dphi6l = R(-1.0 / 5.0) * ph.phi6l;
    ph.dphi7 = R(-2.0 / 5.0) * ph.phi7;
    ph.dphi8 = m_fma(R(-(3.0 / 5.0)), ph.phi8, -(R(3.0 / 5.0) * ph.phi8l));
    ph.dphi8l = R(-3.0 / 5.0) * ph.phi8l;

    // the inspiral's phase derivative where the intermediate starts
    Powers<R> p = useful_powers(f_match_in);
    R d_in = ph.dphi0;
    d_in = m_fma(ph.dphi2, p.two_thirds, d_in);
    d_in = m_fma(ph.dphi3, p.itself, d_in);
    d_in = m_fma(ph.dphi4, p.four_thirds, d_in);
    d_in = m_fma(ph.dphi5, p.five_thirds, d_in);
    d_in = m_fma(ph.dphi6, p.two, d_in);
    d_in = m_fma(ph.dphi6l * p.two, p.log, d_in);
    d_in = m_fma(ph.dphi7, p.seven_thirds, d_in);
    d_in = m_fma(ph.dphi8, p.eight_thirds, d_in);
    d_in = m_fma(ph.dphi8l * p.eight_thirds, p.log, d_in);
    R pseudo = m_fma(ph.a0, p.eight_thirds, ph.a1 * p.three);
    pseudo = m_fma(ph.a2 * p.eight_thirds, p.two_thirds, pseudo);
    pseudo = m_fma(ph.a3 * p.eight_thirds, p.itself, pseudo);
    d_in = d_in + pseudo;
    d_in = d_in * p.m_eight_thirds * dphase0;

    // intermediate collocation, 5 x 5
    width = f_match_im - f_match_in;
    R vm[5] = {d_in,
               m_fma(R(0.75), itm_phase_22_v2m_rdv4(eta, s, dchi, delta) + rd_v4,
                     R(0.25) * itm_phase_22_v2(eta, s, dchi, delta)),
               itm_phase_22_v3m_rdv4(eta, s, dchi, delta) + rd_v4, itm_phase_22_d43(eta, s, dchi, delta), phase_rd};
    vm[3] = vm[3] + vm[2];
    R mm[5][5];
    for (int i = 0; i < 5; ++i) {
        R x = m_fma(R(g5[i]), width, f_match_in);
        R r1 = f_ring / x;
        R r2 = r1 * r1;
        R lorentz = i == 0 ? (R(4.0) * c_l) / m_fma(R(2.0) * f_damp, R(2.0) * f_damp, (x - f_ring) * (x - f_ring))
                           : (R(4.0) * c_l) / m_fma(R(4.0) * f_damp, f_damp, (x - f_ring) * (x - f_ring));
        mm[i][0] = R(1.0);
        mm[i][1] = r1;
        mm[i][2] = r2;
        mm[i][3] = r1 * r2;
        mm[i][4] = r2 * r2;
        vm[i] = vm[i] - lorentz;
    }
    pivoted_solve(mm, vm);
    ph.b0 = vm[0];
    ph.b1 = vm[1] * f_ring;
    ph.b2 = vm[2] * f_ring * f_ring;
    ph.b3 = vm[3] * f_ring * f_ring * f_ring;
    ph.b4 = vm[4] * f_ring * f_ring * f_ring * f_ring;

    ph.f_phase_match_in = f_match_in;
    ph.f_phase_match_im = f_match_im;
    ph.c_l = c_l;
    ph.c_l_gr = c_l;
    ph.c4_ov3 = ph.c4 / R(3.0);
    ph.c_l_ov_fda = c_l / f_damp;
    ph.f_ring = f_ring;
    ph.f_damp = f_damp;
    ph.dphase_norm = R(5.0 / (128.0 * lalpi_five_thirds));

    // C1 continuity at both transitions
    Powers<R> pi = useful_powers(f_match_in);
    ph.c2_int = R(0.0);
    ph.c1_int = R(0.0);
    ph.c2_mrd = R(0.0);
    ph.c1_mrd = R(0.0);
    ph.c2_int = ins_phase_22_ansatz_exact(f_match_in, pi, ph) - itm_phase_22_ansatz_exact(f_match_in, pi, ph);
    ph.c1_int = m_fma(-ph.c2_int, f_match_in,
                      ins_phase_22_ansatz_int_exact(f_match_in, pi, ph) -
                          itm_phase_22_ansatz_int_exact(f_match_in, pi, ph));
    Powers<R> pm = useful_powers(f_match_im);
    R phi_imc = m_fma(ph.c2_int, f_match_im, itm_phase_22_ansatz_int_exact(f_match_im, pm, ph) + ph.c1_int);
    R phi_rd = rd_phase_22_ansatz_int_exact(f_match_im, pm, ph);
    R dphi_imc = itm_phase_22_ansatz_exact(f_match_im, pm, ph) + ph.c2_int;
    ph.c2_mrd = dphi_imc - rd_phase_22_ansatz_exact(f_match_im, pm, ph);
    ph.c1_mrd = m_fma(-ph.c2_mrd, f_match_im, phi_imc - phi_rd);
    return ph;
}

// phase22.py:319-326 (the derivative by region: phase22.py:310-316)
template <class R>
WFX_HDI inline R time_shift(const PhaseCoeffs<R>& ph, const WaveformVars<R>& wf) {
    R linb = ut_linb(wf.eta, wf.s_tot_r, wf.dchi, wf.delta);
    R x = wf.f_ring - wf.f_damp;
    Powers<R> p = useful_powers(x);
    R d;
    if (x < ph.f_phase_match_in)
        d = ins_phase_22_ansatz_exact(x, p, ph);
    else if (x >= ph.f_phase_match_im)
        d = rd_phase_22_ansatz_exact(x, p, ph) + ph.c2_mrd;
    else
        d = itm_phase_22_ansatz_exact(x, p, ph) + ph.c2_int;
    R dphi = R(1.0) / wf.eta * d;
    return m_fma(R(-2.0 * kPi), R(500.0) + ut_psi4_to_strain(wf.eta, wf.s_tot_r, wf.dchi), linb - dphi);
}

// LAL's (2,2) phase at M f = mf but for its constant, phase_22 / eta +
// linb mf (xas/reference.py:29-35, phase22.py:289-307): the quantity whose
// change between two reference frequencies moves phi_ref
template <class R>
WFX_HD inline R lal_phase_22(const R* params, const R& mf) {
    R p[kParams] = {params[kMass1], params[kMass2], params[kChi1], params[kChi2], R(1.0), R(0.0), R(0.0), R(1.0)};
    WaveformVars<R> wf = waveform_vars(p, R(1.0), 1.0);
    PhaseCoeffs<R> ph = phase_coeffs(wf);
    R linb = time_shift(ph, wf);
    Powers<R> pw = useful_powers(mf);
    R phi;
    if (mf < ph.f_phase_match_in)
        phi = ins_phase_22_ansatz_int_exact(mf, pw, ph);
    else if (mf >= ph.f_phase_match_im)
        phi = m_fma(ph.c2_mrd, mf, rd_phase_22_ansatz_int_exact(mf, pw, ph) + ph.c1_mrd);
    else
        phi = m_fma(ph.c2_int, mf, itm_phase_22_ansatz_int_exact(mf, pw, ph) + ph.c1_int);
    return phi / wf.eta + linb * mf;
}

// the powers of f the formula reads (xas/core.py:382-392); the leading
// term's f^(-5/3) is always from the FP64 value
template <class R>
struct FreqPowers {
    R f, t, m1, log_f, m76;
    double m53;
};

template <class R>
WFX_HD inline FreqPowers<R> freq_powers(const R& f) {
    FreqPowers<R> fp;
    fp.f = f;
    fp.m1 = R(1.0) / f;
    fp.log_f = lead_log(f);
    R t = lead_exp(fp.log_f / R(3.0));
    fp.t = t * (R(4.0) - t * t * t * fp.m1) / R(3.0);
    fp.m76 = fp.m1 / m_sqrt(fp.t);
    fp.m53 = lead_exp(-5.0 / 3.0 * lead_log(value_of(f)));
    return fp;
}

// The phase but for the leading term and -A (xas/core.py:423-442), in the
// reference's order of terms; c is a record, past = f >= f_phase_in.
template <class R, class C>
WFX_HD inline R phase_rest(const C* c, const FreqPowers<R>& fp, bool& past) {
    const R f = fp.f;
    past = f >= c[kFPhaseIn];
    const bool ringdown = f > c[kFPhaseIm];
    const R t = fp.t, m1 = fp.m1, lf = fp.log_f;
    const R t2 = t * t;
    const R mt1 = t2 * m1;
    const R mt2 = mt1 * mt1;
    const R m2 = m1 * m1;
    R total;
    if (ringdown) {
        total = R(c[kRdConst]);
        total = total + c[kRdP1] * f;
        total = total + c[kRdP23] * t2;
        total = total + c[kRdM1] * m1;
        total = total + c[kRdM3] * (m2 * m1);
    } else if (past) {
        total = R(c[kIntConst]);
        total = total + c[kIntP1] * f;
        total = total + c[kIntLog] * lf;
        total = total + c[kIntM1] * m1;
        total = total + c[kIntM2] * m2;
        total = total + c[kIntM3] * (m2 * m1);
    } else {
        total = R(c[kInsConst]);
        total = total + c[kInsM1] * m1;
        total = total + c[kInsM23] * mt2;
        total = total + c[kInsM13] * mt1;
        total = total + c[kInsLog] * lf;
        total = total + c[kInsP13] * t;
        total = total + c[kInsP13L] * (t * lf);
        total = total + c[kInsP23] * t2;
        total = total + c[kInsP1] * f;
        total = total + c[kInsP1L] * (f * lf);
        total = total + c[kInsP43] * (f * t);
        total = total + c[kInsP53] * (f * t2);
        total = total + c[kInsP2] * (f * f);
    }
    if (past) {
        // (the region's two slots selected by value: an index chosen at run
        // time would keep a record in registers in memory instead)
        const R a = ringdown ? R(c[kAtanA2]) : R(c[kAtanA1]);
        const R w = ringdown ? R(c[kAtanW2]) : R(c[kAtanW1]);
        total = total + a * m_atan(w * (f - c[kShift]));
    }
    return total;
}

// add c (x)^a (log x)^l, x = s f, to the f^a (and f^a log f) coefficients
// (xas/core.py:247-254)
template <class R>
WFX_HD inline void fold_term(R* rec, int slot, int log_slot, const R& c, const R& s_a, const R& log_s) {
    R scaled = c * s_a;
    if (log_slot >= 0) {
        rec[slot] = rec[slot] + scaled * log_s;
        rec[log_slot] = rec[log_slot] + scaled;
    } else {
        rec[slot] = rec[slot] + scaled;
    }
}

// The (2,2)'s coefficient structs of a set-up, which IMRPhenomXHM's
// higher-mode set-up takes from the (2,2)'s instead of recomputing them
template <class R>
struct XasCoeffs22 {
    WaveformVars<R> wf;
    AmpCoeffs<R> amp;
    PhaseCoeffs<R> ph;
};

// the amplitude's slots of the record (the rest are the phase's, the
// cutoff, the polarization factors and the leading term's)
constexpr bool is_amp_slot(int k) { return (k >= kAmpIns && k <= kAmpRdG) || k == kFAmpIn || k == kFAmpRd; }

// a record sink that fills an array
template <class R>
struct RecordArray {
    R* rec;
    WFX_HD void operator()(int k, const R& x) const { rec[k] = x; }
};

// The record of one binary (xas/core.py:257-379): params in the packed
// order of XasParam, f_min the first frequency of its grid; returns LAL's
// error flags. Each of the kSlots slots goes to out(k, value) once, when it
// is final: the amplitude's first, so that a sink that stores them (the
// set-up in dual numbers) holds none of them through the phase's solves.
// T0 and the anchor A are slots kT0, kAnchor in R (the FP32-split path runs
// this in double and splits them afterwards); with keep, the coefficient
// structs are also stored there, each as soon as it is final.
template <class R, class Out>
WFX_HD __attribute__((always_inline)) inline unsigned xas_record_to(const R* params, const R& f_min, double amp_scale, Out& out,
                                      XasCoeffs22<R>* keep = nullptr) {
    WaveformVars<R> wf = waveform_vars(params, f_min, amp_scale);
    if (keep) keep->wf = wf;
    R s = wf.m_sec;
    R one = R(1.0);
    // the phase's slots (and the rest), summed here; the amplitude's
    // entries stay unused
    R rec[kSlots];
#pragma unroll
    for (int k = 0; k < kSlots; ++k) rec[k] = R(0.0);

    // the amplitude first, so that its coefficients are dead before the
    // phase's solves (each slot's arithmetic is independent of the order)
    {
        AmpCoeffs<R> amp = amp_coeffs(wf);
        Powers<R> p = useful_powers(s);
        R prefactor = wf.amp0 * wf.amp_norm;
        R scale = prefactor * p.m_seven_sixths;
        const R third[10] = {one, p.one_third, p.two_thirds, s, p.four_thirds, p.five_thirds, p.two,
                             p.two * p.one_third, p.two * p.two_thirds, p.three};
        const R ins[10] = {one, R(0.0), amp.pn_two_thirds, amp.pn_three_thirds, amp.pn_four_thirds,
                           amp.pn_five_thirds, amp.pn_six_thirds, amp.rho1, amp.rho2, amp.rho3};
#pragma unroll
        for (int k = 0; k < 10; ++k) out(kAmpIns + k, scale * ins[k] * third[k]);
        out(kAmpIntN, prefactor);
        out(kAmpIntD + 0, amp.delta0 * one);
        out(kAmpIntD + 1, amp.delta1 * s);
        out(kAmpIntD + 2, amp.delta2 * p.two);
        out(kAmpIntD + 3, amp.delta3 * p.three);
        out(kAmpIntD + 4, amp.delta4 * p.four);
        out(kAmpRdN, scale * amp.gamma_d13);
        out(kAmpRdGr, amp.gamma_r);
        out(kAmpRdS, s);
        out(kAmpRdG, amp.gamma_d2);
        out(kFAmpIn, amp.f_amp_match_in / s);
        out(kFAmpRd, amp.f_amp_rd_min / s);
        if (keep) keep->amp = amp;
    }

    PhaseCoeffs<R> ph = phase_coeffs(wf);
    if (keep) keep->ph = ph;
    R linb = time_shift(ph, wf);
    R inveta = R(1.0) / wf.eta;
    Powers<R> p = useful_powers(s);
    R norm = inveta * ph.phi_norm;
    // inspiral (the order of xas/core.py:306-310, the time shift last)
    fold_term(rec, kInsM1, -1, norm * ph.phi2, p.m_one, p.log);
    fold_term(rec, kInsM23, -1, norm * ph.phi3, p.m_two_thirds, p.log);
    fold_term(rec, kInsM13, -1, norm * ph.phi4, p.m_one_third, p.log);
    fold_term(rec, kInsConst, kInsLog, norm * ph.phi5l, one, p.log);
    fold_term(rec, kInsP13, -1, norm * ph.phi6, p.one_third, p.log);
    fold_term(rec, kInsP13, kInsP13L, norm * ph.phi6l, p.one_third, p.log);
    fold_term(rec, kInsP23, -1, norm * ph.phi7, p.two_thirds, p.log);
    fold_term(rec, kInsP1, -1, norm * (ph.phi8 + ph.sigma1), s, p.log);
    fold_term(rec, kInsP1, kInsP1L, norm * ph.phi8l, s, p.log);
    fold_term(rec, kInsP43, -1, norm * ph.sigma2, p.four_thirds, p.log);
    fold_term(rec, kInsP53, -1, norm * ph.sigma3, p.five_thirds, p.log);
    fold_term(rec, kInsP2, -1, norm * ph.sigma4, p.two, p.log);
    fold_term(rec, kInsP1, -1, linb, s, p.log);
    // intermediate (xas/core.py:311-312)
    fold_term(rec, kIntP1, -1, inveta * (ph.b0 + ph.c2_int), s, p.log);
    fold_term(rec, kIntConst, kIntLog, inveta * ph.b1, one, p.log);
    fold_term(rec, kIntM1, -1, -inveta * ph.b2, p.m_one, p.log);
    fold_term(rec, kIntM2, -1, -inveta * ph.b3 / R(2.0), p.m_two, p.log);
    fold_term(rec, kIntM3, -1, -inveta * ph.b4 / R(3.0), p.m_three, p.log);
    fold_term(rec, kIntConst, -1, inveta * ph.c1_int, one, p.log);
    fold_term(rec, kIntP1, -1, linb, s, p.log);
    // ringdown (xas/core.py:313-314)
    fold_term(rec, kRdP1, -1, inveta * (ph.c0 + ph.c2_mrd), s, p.log);
    fold_term(rec, kRdP23, -1, inveta * R(1.5) * ph.c1, p.two_thirds, p.log);
    fold_term(rec, kRdM1, -1, -inveta * ph.c2, p.m_one, p.log);
    fold_term(rec, kRdM3, -1, -inveta * ph.c4_ov3, p.m_three, p.log);
    fold_term(rec, kRdConst, -1, inveta * ph.c1_mrd, one, p.log);
    fold_term(rec, kRdP1, -1, linb, s, p.log);

    rec[kAtanA1] = inveta * R(2.0) * ph.c_l_gr / wf.f_damp;
    rec[kAtanW1] = s / (R(2.0) * wf.f_damp);
    rec[kAtanA2] = inveta * ph.c_l_ov_fda;
    rec[kAtanW2] = s / wf.f_damp;
    rec[kShift] = wf.f_ring / s;
    rec[kFPhaseIn] = ph.f_phase_match_in / s;
    rec[kFPhaseIm] = ph.f_phase_match_im / s;

    rec[kFCut] = wf.f_cut;

    const double nan = __builtin_nan("");
    R ci = m_cos(params[kInclination]);
    R pfac = R(0.5) * (R(1.0) + ci * ci);
    rec[kKp] = wf.flags ? R(nan) : R(-kYlm) * pfac;
    rec[kKc] = wf.flags ? R(nan) : R(kYlm) * ci;

    // the reference phase: phi(f_ref) = 2 phi_ref + pi / 4 (xas/core.py:351-358)
    bool past;
    R rest = phase_rest(rec, freq_powers(wf.f_ref), past);
    R shift_constant = R(2.0) * wf.phi0 + R(kPi / 4.0) - rest;
    rec[kInsConst] = rec[kInsConst] + shift_constant;
    rec[kIntConst] = rec[kIntConst] + shift_constant;
    rec[kRdConst] = rec[kRdConst] + shift_constant;

    // T0 = K0 M^(1/3) / (m1 m2) and A = T0 f_ref^(-5/3), zero past the
    // inspiral (xas/core.py:342-344,359-363)
    const double K0 = -2439189.9493546346;  // -3/128 (pi MTSUN_SI)^(-5/3)
    R t0 = R(K0) * lead_cbrt(params[kMass1] + params[kMass2]) / (params[kMass1] * params[kMass2]);
    rec[kT0] = t0;
    rec[kAnchor] = past ? R(0.0) : t0 * lead_exp(R(-5.0 / 3.0) * lead_log(wf.f_ref));
#pragma unroll
    for (int k = 0; k < kSlots; ++k)
        if (!is_amp_slot(k)) out(k, rec[k]);
    return wf.flags;
}

// xas_record_to into rec[0 .. kSlots)
template <class R>
WFX_HD inline unsigned xas_record(const R* params, const R& f_min, double amp_scale, R* rec,
                                   XasCoeffs22<R>* keep = nullptr) {
    RecordArray<R> out{rec};
    return xas_record_to(params, f_min, amp_scale, out, keep);
}

}  // namespace wfx
