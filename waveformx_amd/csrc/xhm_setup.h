// Per-(binary, higher mode) setup of IMRPhenomXHM (release 122022): the
// mode's variables, amplitude coefficients (PN series, pseudo-PN
// collocation, ringdown Lorentzian and fall-off, the (3,2)'s auxiliary
// cubic, the intermediate polynomial), phase coefficients (the rescaled
// (2,2) inspiral, intermediate collocation, ringdown, connections,
// alignment with the (2,2)) and, for the (3,2), the spheroidal ringdown and
// the spheroidal-to-spherical rotation, written into one record per mode.
//
// One FP64 thread per binary and mode. Where LAL differentiates by finite
// differences the reference takes exact derivatives; here they come from
// a second-order Taylor scalar (Jet2) through the same code.
//
// Follows src/waveformx/models/xhm/reference.py:86-608 (LAL's form of the
// set-up), xpnr/hm/variables.py:118-169, xpnr/hm/inspiral.py:49-215 and
// 319-352, xpnr/hm/phase.py:58-229; the fits come from xhm_fits.inc
// (transcribed from the reference's generated modules).
#pragma once

#include "xas_setup.h"

namespace wfx {

// ------------------------------------------------------- Taylor scalar

// value, first and second derivative in one variable, over a scalar R
// (double, or a dual number for the set-up's derivatives in the
// parameters)
template <class R>
struct Jet2 {
    R v, d, dd;
    Jet2() = default;  // trivial, as Dual's (Jet2() is zero)
    WFX_HD Jet2(const R& x) : v(x), d(0.0), dd(0.0) {}
    WFX_HD Jet2(const R& x, const R& dx, const R& ddx) : v(x), d(dx), dd(ddx) {}
};
template <class T>
struct IsJet {
    static constexpr bool value = false;
};
template <class R>
struct IsJet<Jet2<R>> {
    static constexpr bool value = true;
};
template <class R>
WFX_HD inline Jet2<R> jet_chain(const Jet2<R>& x, const R& f, const R& f1, const R& f2) {
    return Jet2<R>(f, f1 * x.d, f2 * x.d * x.d + f1 * x.dd);
}
template <class R>
WFX_HD inline Jet2<R> operator+(const Jet2<R>& a, const Jet2<R>& b) {
    return Jet2<R>(a.v + b.v, a.d + b.d, a.dd + b.dd);
}
template <class R>
WFX_HD inline Jet2<R> operator-(const Jet2<R>& a, const Jet2<R>& b) {
    return Jet2<R>(a.v - b.v, a.d - b.d, a.dd - b.dd);
}
template <class R>
WFX_HD inline Jet2<R> operator-(const Jet2<R>& a) {
    return Jet2<R>(-a.v, -a.d, -a.dd);
}
template <class R>
WFX_HD inline Jet2<R> operator*(const Jet2<R>& a, const Jet2<R>& b) {
    return Jet2<R>(a.v * b.v, a.d * b.v + a.v * b.d, a.dd * b.v + 2.0 * a.d * b.d + a.v * b.dd);
}
template <class R>
WFX_HD inline Jet2<R> operator/(const Jet2<R>& a, const Jet2<R>& b) {
    const R q = a.v / b.v;
    const R dq = (a.d - q * b.d) / b.v;
    return Jet2<R>(q, dq, (a.dd - 2.0 * dq * b.d - q * b.dd) / b.v);
}
// a Jet2 and a constant (a scalar without Taylor parts): the operations of
// the Jet2 forms above with the constant's zero parts left out (equal
// values for finite operands), so that a struct of constants enters a Jet2
// evaluation without a lifted copy
template <class S>
using IfNotJet = typename std::enable_if<!IsJet<S>::value, int>::type;
template <class R, class S, IfNotJet<S> = 0>
WFX_HD inline Jet2<R> operator+(const Jet2<R>& a, const S& b) {
    return Jet2<R>(a.v + b, a.d, a.dd);
}
template <class R, class S, IfNotJet<S> = 0>
WFX_HD inline Jet2<R> operator+(const S& a, const Jet2<R>& b) {
    return Jet2<R>(a + b.v, b.d, b.dd);
}
template <class R, class S, IfNotJet<S> = 0>
WFX_HD inline Jet2<R> operator-(const Jet2<R>& a, const S& b) {
    return Jet2<R>(a.v - b, a.d, a.dd);
}
template <class R, class S, IfNotJet<S> = 0>
WFX_HD inline Jet2<R> operator-(const S& a, const Jet2<R>& b) {
    return Jet2<R>(a - b.v, -b.d, -b.dd);
}
template <class R, class S, IfNotJet<S> = 0>
WFX_HD inline Jet2<R> operator*(const Jet2<R>& a, const S& b) {
    return Jet2<R>(a.v * b, a.d * b, a.dd * b);
}
template <class R, class S, IfNotJet<S> = 0>
WFX_HD inline Jet2<R> operator*(const S& a, const Jet2<R>& b) {
    return Jet2<R>(a * b.v, a * b.d, a * b.dd);
}
template <class R, class S, IfNotJet<S> = 0>
WFX_HD inline Jet2<R> operator/(const Jet2<R>& a, const S& b) {
    return Jet2<R>(a.v / b, a.d / b, a.dd / b);
}
template <class R, class S, IfNotJet<S> = 0>
WFX_HD inline Jet2<R> operator/(const S& a, const Jet2<R>& b) {
    const R q = a / b.v;
    const R dq = -(q * b.d) / b.v;
    return Jet2<R>(q, dq, (-(2.0 * dq * b.d) - q * b.dd) / b.v);
}
#define WFX_JET_CMP(op)                                                                                \
    template <class R>                                                                                 \
    WFX_HD inline bool operator op(const Jet2<R>& a, const Jet2<R>& b) {                               \
        return a.v op b.v;                                                                             \
    }                                                                                                  \
    template <class R, class S, typename std::enable_if<!IsJet<S>::value, int>::type = 0>              \
    WFX_HD inline bool operator op(const Jet2<R>& a, const S& b) {                                     \
        return a.v op b;                                                                               \
    }                                                                                                  \
    template <class R, class S, typename std::enable_if<!IsJet<S>::value, int>::type = 0>              \
    WFX_HD inline bool operator op(const S& a, const Jet2<R>& b) {                                     \
        return a op b.v;                                                                               \
    }
WFX_JET_CMP(<)
WFX_JET_CMP(>)
WFX_JET_CMP(<=)
WFX_JET_CMP(>=)
WFX_JET_CMP(==)
WFX_JET_CMP(!=)
#undef WFX_JET_CMP
template <class R>
WFX_HD inline double value_of(const Jet2<R>& x) {
    return value_of(x.v);
}
template <class R>
WFX_HD inline Jet2<R> m_exp(const Jet2<R>& x) {
    const R e = m_exp(x.v);
    return jet_chain(x, e, e, e);
}
template <class R>
WFX_HD inline Jet2<R> m_log(const Jet2<R>& x) {
    return jet_chain(x, m_log(x.v), R(1.0) / x.v, R(-1.0) / (x.v * x.v));
}
template <class R>
WFX_HD inline Jet2<R> m_sqrt(const Jet2<R>& x) {
    const R s = m_sqrt(x.v);
    return jet_chain(x, s, R(0.5) / s, R(-0.25) / (x.v * s));
}
template <class R>
WFX_HD inline Jet2<R> m_cbrt(const Jet2<R>& x) {
    const R c = m_cbrt(x.v);
    return jet_chain(x, c, c / (3.0 * x.v), -2.0 * c / (9.0 * x.v * x.v));
}
template <class R>
WFX_HD inline Jet2<R> m_atan(const Jet2<R>& x) {
    const R r = R(1.0) / (R(1.0) + x.v * x.v);
    return jet_chain(x, m_atan(x.v), r, -2.0 * x.v * r * r);
}
template <class R>
WFX_HD inline Jet2<R> m_cos(const Jet2<R>& x) {
    return jet_chain(x, m_cos(x.v), -m_sin(x.v), -m_cos(x.v));
}
template <class R>
WFX_HD inline Jet2<R> m_sin(const Jet2<R>& x) {
    return jet_chain(x, m_sin(x.v), m_cos(x.v), -m_sin(x.v));
}
template <class R>
WFX_HD inline Jet2<R> m_abs(const Jet2<R>& x) {
    return x.v < 0.0 ? -x : x;
}
template <class R>
WFX_HD inline Jet2<R> m_pow(const Jet2<R>& x, double y) {
    return jet_chain(x, m_pow(x.v, y), y * m_pow(x.v, y - 1.0), y * (y - 1.0) * m_pow(x.v, y - 2.0));
}

// jnp.hypot: the larger modulus times sqrt(1 + (smaller / larger)^2)
WFX_HD inline double m_hypot(double x, double y) {
    const double a = fabs(x), b = fabs(y);
    const double hi = a > b ? a : b, lo = a > b ? b : a;
    if (hi == 0.0) return hi;
    const double r = lo / hi;
    return hi * sqrt(1.0 + r * r);
}
template <int N>
WFX_HD inline Dual<N> m_hypot(const Dual<N>& x, const Dual<N>& y) {
    const Dual<N> a = m_abs(x), b = m_abs(y);
    const Dual<N> hi = a > b ? a : b, lo = a > b ? b : a;
    if (hi == 0.0) return hi;
    const Dual<N> r = lo / hi;
    return hi * m_sqrt(1.0 + r * r);
}
template <class R>
WFX_HD inline Jet2<R> m_hypot(const Jet2<R>& x, const Jet2<R>& y) {
    const R h = m_hypot(x.v, y.v);
    const R s = x.v * x.d + y.v * y.d;
    return Jet2<R>(h, s / h, (x.d * x.d + x.v * x.dd + y.d * y.d + y.v * y.dd) / h - s * s / (h * h * h));
}
WFX_HD inline double m_atan2(double y, double x) { return atan2(y, x); }
template <class R>
WFX_HD inline Jet2<R> m_atan2(const Jet2<R>& y, const Jet2<R>& x) {
    const R r2 = x.v * x.v + y.v * y.v;
    const R n = x.v * y.d - y.v * x.d;
    const R n1 = x.v * y.dd - y.v * x.dd;
    const R r21 = 2.0 * (x.v * x.d + y.v * y.d);
    return Jet2<R>(m_atan2(y.v, x.v), n / r2, (n1 * r2 - n * r21) / (r2 * r2));
}

// lax.integer_pow (the reference's `x ** n` for a traced x, an int n):
// binary exponentiation
template <class R>
WFX_HD inline R int_pow(const R& x, int n) {
    if (n == 0) return R(1.0);
    R acc(0.0), base = x;
    bool have = false;
    while (n > 0) {
        if (n & 1) {
            acc = have ? acc * base : base;
            have = true;
        }
        n >>= 1;
        if (n > 0) base = base * base;
    }
    return acc;
}

#include "xhm_fits.inc"

// --------------------------------------------------------------- records

// One record per higher mode (kHmSlots doubles) and a head per binary
// (kHeadSlots doubles): the coefficients of xhm/reference.py's LAL form,
// read by the evaluation per frequency in Mf.
enum HmSlot : int {
    kHmTag = 0, kHmOn, kHmAmp0, kHmMsec, kHmFRing, kHmFDamp, kHmFalloff, kHmMixing,
    kHmFAmpIn = 8, kHmFAmpIm, kHmFPhaseIn, kHmFPhaseIm,
    kHmPnRe = 12,  // 7
    kHmPnIm = 19,  // 7
    kHmPnGlobal = 26, kHmAmpNorm, kHmPnDominant, kHmFcut73, kHmFcut83, kHmFcut3,
    kHmIns = 32,    // 3 pseudo-PN coefficients
    kHmRd = 35,     // 5: Lorentzian (3), fall-off value and rate
    kHmRdAux = 40,  // 4: the (3,2)'s cubic
    kHmFRdAux = 44, kHmInterC, kHmInterH,
    kHmInter = 47,  // 8: the intermediate polynomial in u
    kHmNInter = 55,
    kHmPhi = 56,    // 13
    kHmPhiL = 69,   // 13
    kHmAlpha0 = 82, kHmAlpha2, kHmAlphaL, kHmC0, kHmC1, kHmC2, kHmC3, kHmC4, kHmCL, kHmC1Insp, kHmCInsp, kHmC1Rd,
    kHmCRd, kHmDeltaPhi,
    kHmSph = 96,    // 5: alpha0_s, alphaL_s, alpha2_s, alpha4_s, phi0_s
    kHmMix = 101,   // 4: -a322 (re, im), -a323 (re, im)
    kHmFalloffAt = 119,  // the fall-off's position in its formula (kHmFalloff: its region's bound)
    kHmSlots = 128
};
// the head: (F+, Fx) of each mode (re, im), mode order (2,2), (2,1), (3,3),
// (3,2), (4,4); NaN where LAL fails
enum HeadSlot : int { kHeadFactors = 0, kHeadSlots = 32 };
constexpr int kHmModes = 4;
WFX_CONST int kHmEll[kHmModes] = {2, 3, 3, 4};
WFX_CONST int kHmEmm[kHmModes] = {1, 3, 2, 4};

// ------------------------------------------------------------ the (2,2)

// phase22.py:304-307 IMRPhenomX_Phase_22 (Mf >= fPhaseMatchIM: ringdown)
template <class R, class C = R>
WFX_HD inline R phase_22(const R& mf, const Powers<R>& p, const PhaseCoeffs<C>& ph) {
    if (mf < ph.f_phase_match_in) return ins_phase_22_ansatz_int_exact(mf, p, ph);
    if (mf >= ph.f_phase_match_im)
        return m_fma(ph.c2_mrd, mf, rd_phase_22_ansatz_int_exact(mf, p, ph) + ph.c1_mrd);
    return m_fma(ph.c2_int, mf, itm_phase_22_ansatz_int_exact(mf, p, ph) + ph.c1_int);
}
// phase22.py:310-316 IMRPhenomX_dPhase_22
template <class R, class C = R>
WFX_HD inline R dphase_22(const R& mf, const Powers<R>& p, const PhaseCoeffs<C>& ph) {
    if (mf < ph.f_phase_match_in) return ins_phase_22_ansatz_exact(mf, p, ph);
    if (mf >= ph.f_phase_match_im) return rd_phase_22_ansatz_exact(mf, p, ph) + ph.c2_mrd;
    return itm_phase_22_ansatz_exact(mf, p, ph) + ph.c2_int;
}

// ------------------------------------------------------------ the mode

template <class R>
struct HmCoeffs {
    int ell, emm, tag, mode_int;
    bool mixing, amp_zero;
    R chi_s, chi_a, f_meco_lm, f_ring, f_damp, delta_t, f_amp_rd_falloff;
    R mix2_re, mix2_im, mix3_re, mix3_im;
    // amplitude
    R amp_norm, pn_dominant, f_rd_aux, f_amp_match_in, f_amp_match_im, pn_re[7], pn_im[7], pn_global, fcut73,
        fcut83, fcut3, ins[3], rd[5], rd_aux[4], inter_center, inter_half, inter[8];
    int n_inter;
    // phase
    R phi[13], phi_l[13], alpha0, alpha2, alpha_l, c0, c1, c2, c3, c4, c_l, c1_insp, c_insp, c1_rd, c_rd, deltaphi,
        f_phase_match_in, f_phase_match_im;
    R alpha0_s, alpha_l_s, alpha2_s, alpha4_s, phi0_s;
    // the (2,2) data of the (3,2) rotation (reference.py:625-633)
    R f_ring22, f_damp22, inveta, timeshift, phiref22;
};

template <int N, class R>
WFX_HD inline R hm_horner(const double (&c)[N], const R& u) {
    R total = R(c[N - 1]);
    for (int k = N - 2; k >= 0; --k) total = total * u + c[k];
    return total;
}
#define WFX_HMQ(name, af) (hm_horner(kHmq_##name##_num, 1.0 - (af)) / hm_horner(kHmq_##name##_den, 1.0 - (af)))

// variables.py:118-169 with reference.py:102-122 (QNM in 1 - a_f, the
// 122022 flags)
template <class R>
WFX_HDI inline void mode_vars(const WaveformVars<R>& wf, int k, HmCoeffs<R>& c) {
    c.ell = kHmEll[k];
    c.emm = kHmEmm[k];
    c.tag = 10 * c.ell + c.emm;
    c.mode_int = k;
    c.mixing = c.tag == 32;
    c.amp_zero = (c.tag == 21 || c.tag == 33) && wf.q == 1.0 && wf.chi1l == wf.chi2l;
    c.chi_s = (wf.chi1l + wf.chi2l) * 0.5;
    c.chi_a = (wf.chi1l - wf.chi2l) * 0.5;
    c.f_meco_lm = wf.f_meco * (double)c.emm * 0.5;
    const R af = wf.a_final;
    R fr, fd;
    switch (c.tag) {
        case 21: fr = WFX_HMQ(fring21, af), fd = WFX_HMQ(fdamp21, af); break;
        case 33: fr = WFX_HMQ(fring33, af), fd = WFX_HMQ(fdamp33, af); break;
        case 32: fr = WFX_HMQ(fring32, af), fd = WFX_HMQ(fdamp32, af); break;
        default: fr = WFX_HMQ(fring44, af), fd = WFX_HMQ(fdamp44, af); break;
    }
    c.f_ring = fr / wf.m_final;
    c.f_damp = fd / wf.m_final;
    c.delta_t = -2.0 * kPi * (500.0 + ut_psi4_to_strain(wf.eta, wf.s_tot_r, wf.dchi));
    c.f_amp_rd_falloff = m_fma(2.0, c.f_damp, c.f_ring);
    c.mix2_re = c.mix2_im = c.mix3_re = c.mix3_im = R(0.0);
    if (c.mixing) {
        c.mix2_re = -WFX_HMQ(re_l3m2lp2, af);
        c.mix2_im = -WFX_HMQ(im_l3m2lp2, af);
        c.mix3_re = -WFX_HMQ(re_l3m2lp3, af);
        c.mix3_im = -WFX_HMQ(im_l3m2lp3, af);
    }
}

// inspiral.py:49-215 IMRPhenomXHM_GetPNAmplitudeCoefficients (LAL's
// complex PN coefficients as (re, im) pairs, the reference's order)
template <class R>
WFX_HDI inline void pn_amp_coeffs(const WaveformVars<R>& wf, HmCoeffs<R>& c) {
    const R chi_a = c.chi_a, chi_s = c.chi_s, eta = wf.eta, delta = wf.delta;
    const double pi = lalpi_itself;
    const double sqrt_5_7 = sqrt(5.0 / 7.0);
    double prefactor;
    switch (c.mode_int) {
        case 0: prefactor = sqrt(2.0) / 3.0; break;
        case 1: prefactor = 0.75 * sqrt_5_7; break;
        case 2: prefactor = sqrt_5_7 / 3.0; break;
        default: prefactor = 4.0 * sqrt(2.0) / 9.0 * sqrt_5_7; break;
    }
    c.pn_global = kHmTwoOverMPow[c.mode_int] * prefactor;
    const R zero = 0.0 * eta;
    R* re = c.pn_re;
    R* im = c.pn_im;
#pragma unroll
    for (int k = 0; k < 7; ++k) re[k] = im[k] = zero;
    if (c.tag == 21) {
        const Powers<double> p2 = useful_powers(2.0);
        R r = 3427.0 * chi_a;
        r = r + 3427.0 * chi_s * delta;
        r = r - 8404.0 * chi_a * eta;
        r = r - 3860.0 * chi_s * delta * eta;
        r = r - 1344.0 * delta * pi;
        R i = -(672.0 * delta);
        i = i - 672.0 * delta * log(16.0);
        double scale = lalpi_four_thirds * p2.four_thirds;
        re[4] = r / 1344.0 * scale;
        im[4] = i / 1344.0 * scale;
        r = 143063173.0 * chi_a;
        r = r + 143063173.0 * chi_s * delta;
        r = r - 546199608.0 * chi_a * eta;
        r = r - 169191096.0 * chi_s * delta * eta;
        r = r - 9898560.0 * delta * pi;
        r = r + 20176128.0 * delta * eta * pi;
        r = r + 61725888.0 * chi_s * delta * chi_a * chi_a;
        r = r - 81285120.0 * chi_s * delta * eta * chi_a * chi_a;
        r = r + 20575296.0 * m_pow(chi_a, 3.0);
        r = r - 81285120.0 * eta * m_pow(chi_a, 3.0);
        r = r + 61725888.0 * chi_a * chi_s * chi_s;
        r = r - 165618432.0 * chi_a * eta * chi_s * chi_s;
        r = r + 20575296.0 * delta * m_pow(chi_s, 3.0);
        r = r - 1016064.0 * delta * eta * chi_s * chi_s * chi_s;
        r = r + 128873808.0 * chi_a * eta * eta;
        r = r - 3859632.0 * chi_s * delta * eta * eta;
        i = -(1350720.0 * delta);
        i = i - 72043776.0 * delta * eta;
        i = i - 5402880.0 * delta * kLn2;
        i = i - 17224704.0 * delta * eta * kLn2;
        scale = lalpi_two * p2.two;
        re[6] = r / 5.419008e6 * scale;
        im[6] = i / 5.419008e6 * scale;
        const R five = m_fma(21535920.0 * delta * eta, eta, m_fma(39626496.0 * delta * eta * chi_s, chi_s, m_fma(
            -(77982912.0 * delta * chi_s), chi_s, m_fma(81285120.0 * delta * eta * chi_a, chi_a, m_fma(
                -(77982912.0 * delta * chi_a), chi_a, m_fma(24385536.0 * chi_s * delta, pi, m_fma(
                    24385536.0 * chi_a, pi, m_fma(-(23670792.0 * delta), eta, m_fma(
                        432843264.0 * chi_a * chi_s, eta, m_fma(-155965824.0 * chi_a, chi_s, -(964357.0 * delta)))))))))));
        re[1] = delta * lalpi_one_third * p2.one_third;
        re[2] = -3.0 * m_fma(chi_s, delta, chi_a) / 2.0 * lalpi_two_thirds * p2.two_thirds;
        re[3] = m_fma(335.0, delta, 1404.0 * delta * eta) / 672.0 * lalpi_itself * p2.itself;
        re[5] = five / 8.128512e6 * lalpi_five_thirds * p2.five_thirds;
    } else if (c.tag == 33) {
        const Powers<double> p2 = useful_powers(2.0 / 3.0);
        const double ln15 = log(1.5);
        R r = 325.0 * chi_a;
        r = r + 325.0 * chi_s * delta;
        r = r - 1120.0 * chi_a * eta;
        r = r - 80.0 * chi_s * delta * eta;
        r = r + 120.0 * delta * pi;
        R i = -(504.0 * delta);
        i = i + 720.0 * delta * ln15;
        double scale = lalpi_four_thirds * p2.four_thirds;
        re[4] = r / 120.0 * scale;
        im[4] = i / 120.0 * scale;
        r = 22007835.0 * chi_a;
        r = r + 22007835.0 * chi_s * delta;
        r = r - 80190540.0 * chi_a * eta;
        r = r - 31722300.0 * chi_s * delta * eta;
        r = r - 9193500.0 * delta * pi;
        r = r + 17826480.0 * delta * eta * pi;
        r = r - 12428640.0 * chi_a * eta * eta;
        r = r - 6078240.0 * chi_s * delta * eta * eta;
        i = 26467560.0 * delta;
        i = i - 98774368.0 * delta * eta;
        i = i - 37810800.0 * delta * ln15;
        i = i + 37558080.0 * delta * eta * ln15;
        scale = lalpi_two * p2.two;
        re[6] = r / 2.17728e6 * scale;
        im[6] = i / 2.17728e6 * scale;
        const R five = m_fma(2966264784.0 * delta * eta, eta, m_fma(55883520.0 * delta * eta * chi_s, chi_s, m_fma(
            -(1131641280.0 * delta * chi_s), chi_s, m_fma(4470681600.0 * delta * eta * chi_a, chi_a, m_fma(
                -(1131641280.0 * delta * chi_a), chi_a, m_fma(-(5926068792.0 * delta), eta, m_fma(
                    9053130240.0 * chi_a * chi_s, eta, m_fma(-2263282560.0 * chi_a, chi_s, -(1077664867.0 * delta)))))))));
        re[1] = delta * lalpi_one_third * p2.one_third;
        re[3] = m_fma(-1945.0, delta, 2268.0 * delta * eta) / 672.0 * lalpi_itself * p2.itself;
        re[5] = five / 4.4706816e8 * lalpi_five_thirds * p2.five_thirds;
    } else if (c.tag == 32) {
        R r = 0.0 - 3955.0 * chi_s;
        r = r - 3955.0 * chi_a * delta;
        r = r + 10810.0 * chi_s * eta;
        r = r + 11865.0 * chi_a * delta * eta;
        r = r - 12600.0 * chi_s * eta * eta;
        R i = 2520.0 + zero;
        i = i - 11088.0 * eta;
        const double scale = lalpi_five_thirds;
        re[5] = r / 840.0 * scale;
        im[5] = i / 840.0 * scale;
        const R six = m_fma(5907445488.0, m_pow(eta, 3.0), m_fma(5830513920.0 * chi_s * chi_s * eta, eta, m_fma(
            13412044800.0 * chi_a * chi_a * eta, eta, m_fma(-(13202119896.0 * eta), eta, m_fma(
                -(11870591040.0 * eta * chi_s), chi_s, m_fma(1131641280.0 * chi_s, chi_s, m_fma(
                    -(7865605440.0 * eta * chi_a), chi_a, m_fma(1131641280.0 * chi_a, chi_a, m_fma(
                        3576545280.0 * chi_s * eta, pi, m_fma(-(15209631360.0 * chi_a * chi_s * delta), eta, m_fma(
                            -26069649.0, eta, m_fma(2263282560.0 * chi_a * chi_s, delta, 824173699.0))))))))))));
        re[2] = m_fma(3.0, eta, -1.0) * lalpi_two_thirds;
        re[3] = -4.0 * chi_s * eta * lalpi_itself;
        re[4] = m_fma(82460.0 * eta, eta, m_fma(-61625.0, eta, 10471.0)) / 10080.0 * lalpi_four_thirds;
        re[6] = six / 4.4706816e8 * lalpi_two;
    } else {
        const Powers<double> p2 = useful_powers(0.5);
        R r = -0.0 + 565.0 * chi_s;
        r = r + 565.0 * chi_a * delta;
        r = r - 2075.0 * chi_s * eta;
        r = r - 1695.0 * chi_a * delta * eta;
        r = r + 240.0 * pi;
        r = r - 720.0 * eta * pi;
        r = r + 1140.0 * chi_s * eta * eta;
        R i = -1008.0 + zero;
        i = i + 3579.0 * eta;
        i = i + 960.0 * kLn2;
        i = i - 2880.0 * eta * kLn2;
        const double scale = lalpi_five_thirds * p2.five_thirds;
        re[5] = r / 120.0 * scale;
        im[5] = i / 120.0 * scale;
        const R six = m_fma(-805075876080.0, m_pow(eta, 3.0), m_fma(-(10897286400.0 * chi_s * chi_s * eta), eta, m_fma(
            -(871782912000.0 * chi_a * chi_a * eta), eta, m_fma(2271682065240.0 * eta, eta, m_fma(
                224302478400.0 * eta * chi_s, chi_s, m_fma(-(73556683200.0 * chi_s), chi_s, m_fma(
                    511264353600.0 * eta * chi_a, chi_a, m_fma(-(73556683200.0 * chi_a), chi_a, m_fma(
                        441340099200.0 * chi_a * chi_s * delta, eta, m_fma(-745140957231.0, eta, m_fma(
                            -(147113366400.0 * chi_a * chi_s), delta, 7888301437.0)))))))))));
        re[2] = m_fma(-3.0, eta, 1.0) * lalpi_two_thirds * p2.two_thirds;
        re[4] = m_fma(-(446460.0 * eta), eta, m_fma(641105.0, eta, -158383.0)) / 36960.0 * lalpi_four_thirds *
                p2.four_thirds;
        re[6] = six / 2.90594304e10 * lalpi_two * p2.two;
    }
}

// ------------------------------------------------------------ amplitude

// reference.py:157-176 the PN and inspiral amplitude ansatzes; R the
// frequency's scalar, C the coefficients'
template <class R, class C>
WFX_HD inline R pn_amp(const Powers<R>& p, const C* pn_re, const C* pn_im, const typename NoDeduce<C>::type& pn_global,
                       const typename NoDeduce<C>::type& amp_norm) {
    const R terms[6] = {p.one_third, p.two_thirds, p.itself, p.four_thirds, p.five_thirds, p.two};
    R re = R(pn_re[0]), im = R(pn_im[0]);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        re = re + terms[k] * pn_re[k + 1];
        im = im + terms[k] * pn_im[k + 1];
    }
    return m_hypot(re, im) * (pn_global * p.m_seven_sixths * amp_norm);
}
template <class R, class C>
WFX_HDI inline R inspiral_amp(const Powers<R>& p, const C* pn_re, const C* pn_im,
                             const typename NoDeduce<C>::type& pn_global, const typename NoDeduce<C>::type& amp_norm,
                             const typename NoDeduce<C>::type& pn_dominant, const typename NoDeduce<C>::type& fcut73,
                             const typename NoDeduce<C>::type& fcut83, const typename NoDeduce<C>::type& fcut3,
                             const C* ins) {
    R pseudo = m_fma(p.seven_thirds / fcut73, R(ins[0]), p.eight_thirds / fcut83 * ins[1]);
    pseudo = m_fma(p.three / fcut3, R(ins[2]), pseudo);
    return pn_amp(p, pn_re, pn_im, pn_global, amp_norm) + pseudo * p.m_seven_sixths * pn_dominant;
}

// reference.py:203-220 the Lorentzian and its derivative
template <class R, class C>
WFX_HD inline R hm_lorentzian(const R& mf, const C& f_ring, const typename NoDeduce<C>::type& f_damp, const C* rd) {
    const R dfr = mf - f_ring;
    const C dfd = f_damp * rd[2];
    return rd[0] * f_damp / (m_exp(rd[1] / dfd * dfr) * m_fma(dfr, dfr, R(dfd * dfd)));
}
template <class R, class C>
WFX_HD inline R hm_dlorentzian(const R& mf, const C& f_ring, const typename NoDeduce<C>::type& fda, const C* rd) {
    const R dfr = mf - f_ring;
    const R numerator =
        rd[0] * m_fma(R(fda * fda * rd[1] * rd[2]), R(rd[2]), m_fma(dfr * dfr, R(rd[1]), 2.0 * fda * dfr * rd[2]));
    R denom = m_fma(dfr, dfr, R(fda * fda * rd[2] * rd[2]));
    denom = rd[2] * denom * denom * m_exp(dfr * rd[1] / (fda * rd[2]));
    return -numerator / denom;
}
// reference.py:223-236 the ringdown amplitude (fall-off; the (3,2)'s cubic);
// falloff_at is the fall-off's position in its formula, falloff the bound of
// its region (the same value: the records keep them apart so that the bound
// can be held fixed while the position varies, as a jnp.where selection)
template <class R, class C>
WFX_HD inline R rd_amp(const R& mf, const C& f_ring, const typename NoDeduce<C>::type& f_damp,
                       const typename NoDeduce<C>::type& falloff, bool mixing, const C* rd, const C* aux,
                       const typename NoDeduce<C>::type& f_rd_aux, const typename NoDeduce<C>::type& falloff_at) {
    R out = mf >= falloff ? R(rd[3] * m_exp(-rd[4] * (mf - falloff_at))) : hm_lorentzian(mf, f_ring, f_damp, rd);
    if (mixing) {
        const R cubic = m_fma(m_fma(m_fma(R(aux[3]), mf, R(aux[2])), mf, R(aux[1])), mf, R(aux[0]));
        out = mf < f_rd_aux ? cubic : out;
    }
    return out;
}

// ------------------------------------------------------ (3,2) ringdown

// phase.py:58-80 the ringdown phase ansatz and its integral: the rescaled
// (2,2) one or, with mixing, the spheroidal one
template <class R, class C>
WFX_HD inline R rd_phase(const R& ff, const Powers<R>& p, const C& frd, const typename NoDeduce<C>::type& fda,
                         bool mixing, const typename NoDeduce<C>::type& alpha0,
                         const typename NoDeduce<C>::type& alpha2, const typename NoDeduce<C>::type& alpha_l,
                         const typename NoDeduce<C>::type& alpha0_s, const typename NoDeduce<C>::type& alpha_l_s,
                         const typename NoDeduce<C>::type& alpha2_s, const typename NoDeduce<C>::type& alpha4_s) {
    const R den = m_fma(R(fda), R(fda), (ff - frd) * (ff - frd));
    if (!mixing) return m_fma(R(frd * frd * alpha2), p.m_two, R(alpha0)) + (alpha_l * fda) / den;
    const R d = m_fma(R(alpha2_s), p.m_two, R(alpha0_s));
    return m_fma(R(alpha4_s), p.m_four, d) + (alpha_l_s * fda) / den;
}
template <class R, class C>
WFX_HD inline R rd_phase_int(const R& ff, const Powers<R>& p, const C& frd, const typename NoDeduce<C>::type& fda,
                             bool mixing, const typename NoDeduce<C>::type& alpha0,
                             const typename NoDeduce<C>::type& alpha2, const typename NoDeduce<C>::type& alpha_l,
                             const typename NoDeduce<C>::type& alpha0_s, const typename NoDeduce<C>::type& alpha_l_s,
                             const typename NoDeduce<C>::type& alpha2_s, const typename NoDeduce<C>::type& alpha4_s,
                             const typename NoDeduce<C>::type& phi0_s) {
    const R invf = p.m_one;
    const R arctan = m_atan((ff - frd) / fda);
    if (!mixing) {
        const R phase_rd = m_fma(R(alpha0), ff, -(frd * frd * alpha2 * invf));
        return m_fma(R(alpha_l), arctan, phase_rd);
    }
    R phase_rd = m_fma(R(alpha0_s), ff, R(phi0_s));
    phase_rd = m_fma(R(-alpha2_s), invf, phase_rd);
    phase_rd = m_fma(R(-(1.0 / 3.0 * alpha4_s)), p.m_three, phase_rd);
    return m_fma(R(alpha_l_s), arctan, phase_rd);
}
// reference.py:366-378 conj(a322) w22 + conj(a323) A_s e^{i phi_s}
template <class R, class C>
WFX_HD inline void rotate(const R& mf, const Powers<R>& p, const HmCoeffs<C>& c, const R& w_re, const R& w_im,
                          R& out_re, R& out_im) {
    const R amplm = rd_amp(mf, c.f_ring, c.f_damp, c.f_amp_rd_falloff, true, c.rd, c.rd_aux, c.f_rd_aux,
                           c.f_amp_rd_falloff);
    const R philm = rd_phase_int(mf, p, c.f_ring, c.f_damp, true, C(0.0), C(0.0), C(0.0), c.alpha0_s, c.alpha_l_s,
                                 c.alpha2_s, c.alpha4_s, c.phi0_s);
    const R a_re = w_re * c.mix2_re + w_im * c.mix2_im;
    const R a_im = w_im * c.mix2_re - w_re * c.mix2_im;
    const R b_re = amplm * c.mix3_re;
    const R b_im = -(c.mix3_im * amplm);
    const R cs = m_cos(philm), sn = m_sin(philm);
    out_re = a_re + (b_re * cs - b_im * sn);
    out_im = a_im + (b_re * sn + b_im * cs);
}
// reference.py:381-388 SpheroidalToSpherical with the (2,2) ringdown ansatz
// (mf a Taylor scalar over the set-up's C, the (2,2)'s coefficients of C)
template <class R, class C>
WFX_HDI inline void spheroidal_to_spherical(const R& mf, const HmCoeffs<C>& c, const AmpCoeffs<C>& amp22,
                                           const PhaseCoeffs<C>& ph22, R& re, R& im) {
    const Powers<R> p = useful_powers(mf);
    const R a22 = rd_amp_22_ansatz_exact(mf, R(c.f_ring22), R(c.f_damp22), amp22) * c.amp_norm * p.m_seven_sixths;
    const R phi22 = m_fma(R(c.inveta), phase_22(mf, p, ph22), c.timeshift * mf) + c.phiref22;
    rotate(mf, p, c, a22 * m_cos(phi22), a22 * m_sin(phi22), re, im);
}

// phase.py:155-193 GetSpheroidalCoefficients (and the (2,2) time shift and
// phase offset of the (3,2))
template <class R>
WFX_HDI inline void spheroidal_coeffs(const WaveformVars<R>& wf, const PhaseCoeffs<R>& ph22, HmCoeffs<R>& c) {
    const R fringlm = c.f_ring, fdamplm = c.f_damp;
    const R freqs[4] = {wf.f_ring, m_fma(-1.5, fdamplm, fringlm), m_fma(-0.5, fdamplm, fringlm),
                        m_fma(0.5, fdamplm, fringlm)};
    R a[4][4], x[4] = {hpr_phase_32_p1(wf), hpr_phase_32_p2(wf), hpr_phase_32_p3(wf), hpr_phase_32_p4(wf)};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const R ff = freqs[i];
        const R ffm1 = 1.0 / ff, ffm2 = ffm1 * ffm1;
        a[i][0] = 1.0;
        a[i][1] = fdamplm / (fdamplm * fdamplm + (ff - fringlm) * (ff - fringlm));
        a[i][2] = ffm2;
        a[i][3] = ffm2 * ffm2;
    }
    pivoted_solve<R, 4>(a, x);
    c.alpha0_s = x[0], c.alpha_l_s = x[1], c.alpha2_s = x[2], c.alpha4_s = x[3], c.phi0_s = R(0.0);
    R fref_rd = wf.f_ring + wf.f_damp;
    Powers<R> p_ref = useful_powers(fref_rd);
    const R tshift = hpr_phase_32_spheroidal_time_shift(wf);
    const R timeshift = time_shift(ph22, wf);
    const R dphi22_ref = m_fma(1.0 / wf.eta, dphase_22(fref_rd, p_ref, ph22), timeshift);
    c.alpha0_s = c.alpha0_s + dphi22_ref + tshift -
                 rd_phase(fref_rd, p_ref, fringlm, fdamplm, true, R(0.0), R(0.0), R(0.0), c.alpha0_s, c.alpha_l_s,
                          c.alpha2_s, c.alpha4_s);
    const double phaseshift = 0.0;
    fref_rd = wf.f_ring;
    p_ref = useful_powers(fref_rd);
    const R mf_ref = wf.f_ref * wf.m_sec;
    const R phase_22_mf_ref = phase_22(mf_ref, useful_powers(mf_ref), ph22);
    const R phase_22_fref = phase_22(fref_rd, p_ref, ph22);
    const R phiref22 =
        m_fma(-1.0 / wf.eta, phase_22_mf_ref, -(timeshift * mf_ref)) - phaseshift + 2.0 * wf.phi0 + kPi / 4.0;
    const R phi22_ref = m_fma(1.0 / wf.eta, phase_22_fref, timeshift * fref_rd) + phaseshift + phiref22;
    const R phishift = hpr_phase_32_spheroidal_phase_shift(wf);
    c.phi0_s = phi22_ref -
               rd_phase_int(fref_rd, p_ref, fringlm, fdamplm, true, R(0.0), R(0.0), R(0.0), c.alpha0_s, c.alpha_l_s,
                            c.alpha2_s, c.alpha4_s, R(0.0)) +
               phishift;
    c.timeshift = timeshift;
    c.phiref22 = phiref22;
}

// reference.py:179-197 the pseudo-PN coefficients (three collocation points)
template <class R>
WFX_HDI inline void inspiral_amp_coeffs(const WaveformVars<R>& wf, HmCoeffs<R>& c) {
    const R fcut = c.f_amp_match_in;
    const R freqs[3] = {0.5 * fcut, 0.75 * fcut, fcut};
    R values[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        R v;
        switch (c.tag * 10 + i) {
            case 210: v = hai_amp_21_iv1(wf); break;
            case 211: v = hai_amp_21_iv2(wf); break;
            case 212: v = hai_amp_21_iv3(wf); break;
            case 330: v = hai_amp_33_iv1(wf); break;
            case 331: v = hai_amp_33_iv2(wf); break;
            case 332: v = hai_amp_33_iv3(wf); break;
            case 320: v = hai_amp_32_iv1(wf); break;
            case 321: v = hai_amp_32_iv2(wf); break;
            case 322: v = hai_amp_32_iv3(wf); break;
            case 440: v = hai_amp_44_iv1(wf); break;
            case 441: v = hai_amp_44_iv2(wf); break;
            default: v = hai_amp_44_iv3(wf); break;
        }
        values[i] = m_abs(v);
    }
    const Powers<R> finsp = useful_powers(fcut);
    c.fcut73 = finsp.seven_thirds;
    c.fcut83 = finsp.eight_thirds;
    c.fcut3 = finsp.three;
    R v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const Powers<R> f = useful_powers(freqs[i]);
        const R pn = pn_amp(f, c.pn_re, c.pn_im, c.pn_global, c.amp_norm);
        v[i] = (values[i] - pn) / c.pn_dominant / f.m_seven_sixths;
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        R s = R(0.0);
#pragma unroll
        for (int i = 0; i < 3; ++i) s = s + kHmInspiralInverse[j][i] * v[i];
        c.ins[j] = s;
    }
}

// reference.py:239-272 the ringdown amplitude coefficients (and the (3,2)'s
// auxiliary cubic)
template <class R>
WFX_HDI inline void rd_amp_coeffs(const WaveformVars<R>& wf, HmCoeffs<R>& c) {
    const R fdamp = c.f_damp;
    R rd[3];
    if (c.tag == 32) {
        rd[0] = m_abs(har_amp_32_alambda(wf));
        rd[1] = har_amp_32_lambda(wf);
        rd[2] = har_amp_32_sigma(wf);
    } else {
        R c1, c2, c3;
        if (c.tag == 21) {
            c1 = har_amp_21_rdcp1(wf), c2 = har_amp_21_rdcp2(wf), c3 = har_amp_21_rdcp3(wf);
        } else if (c.tag == 33) {
            c1 = har_amp_33_rdcp1(wf), c2 = har_amp_33_rdcp2(wf), c3 = har_amp_33_rdcp3(wf);
        } else {
            c1 = har_amp_44_rdcp1(wf), c2 = har_amp_44_rdcp2(wf), c3 = har_amp_44_rdcp3(wf);
        }
        const R rdcp1 = m_abs(c1), rdcp2 = m_abs(c2);
        R rdcp3 = m_abs(c3);
        rdcp3 = rdcp3 >= rdcp2 * rdcp2 / rdcp1 ? R(0.5 * rdcp2 * rdcp2 / rdcp1) : rdcp3;
        rdcp3 = rdcp3 > rdcp2 ? R(0.5 * rdcp2) : rdcp3;
        rdcp3 = (rdcp1 < rdcp2 && rdcp3 > rdcp1) ? rdcp1 : rdcp3;
        R deno = m_sqrt(rdcp1 / rdcp3) - rdcp1 / rdcp2;
        deno = deno <= 0.0 ? R(1e-16) : deno;
        const R a_lambda = rdcp1 * fdamp / deno;
        const R sigma = m_sqrt(a_lambda / (rdcp2 * fdamp));
        rd[0] = a_lambda;
        rd[1] = 0.5 * sigma * m_log(rdcp1 / rdcp3);
        rd[2] = sigma;
    }
    const R falloff = c.f_amp_rd_falloff;
    const R value = hm_lorentzian(falloff, c.f_ring, c.f_damp, rd);
    c.rd[0] = rd[0];
    c.rd[1] = rd[1];
    c.rd[2] = rd[2];
    c.rd[3] = value;
    c.rd[4] = -1.0 * hm_dlorentzian(falloff, c.f_ring, c.f_damp, rd) / value;
#pragma unroll
    for (int j = 0; j < 4; ++j) c.rd_aux[j] = R(0.0);
    if (c.mixing) {
        R values[4] = {m_abs(har_amp_32_rdaux1(wf)), m_abs(har_amp_32_rdaux2(wf)),
                       hm_lorentzian(c.f_rd_aux, c.f_ring, c.f_damp, c.rd),
                       hm_dlorentzian(c.f_rd_aux, c.f_ring, c.f_damp, c.rd)};
        const R freqs[3] = {c.f_amp_match_im, 0.5 * (c.f_amp_match_im + c.f_rd_aux), c.f_rd_aux};
        R a[4][4];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const R f = freqs[i];
            a[i][0] = 1.0, a[i][1] = f, a[i][2] = f * f, a[i][3] = f * f * f;
        }
        const R f = c.f_rd_aux;
        a[3][0] = 0.0, a[3][1] = 1.0, a[3][2] = 2.0 * f, a[3][3] = 3.0 * f * f;
        pivoted_solve<R, 4>(a, values);
#pragma unroll
        for (int j = 0; j < 4; ++j) c.rd_aux[j] = values[j];
    }
}

// reference.py:275-334 the intermediate amplitude (a polynomial in u times
// f^(-7/6) through the collocation points)
template <class R>
WFX_HDI inline void intermediate_amp_coeffs(const WaveformVars<R>& wf, const AmpCoeffs<R>& amp22,
                                           const PhaseCoeffs<R>& ph22, HmCoeffs<R>& c) {
    // IMRPhenomXHMIntermediateAmpVersion (reference.py:59-60): 0 skipped,
    // 1 value, 2 value and derivative
    const int v21[6] = {1, 1, 0, 1, 0, 2}, vhm[6] = {2, 1, 1, 1, 1, 2};
    const int* version = c.tag == 21 ? v21 : vhm;
    const int n_colloc = 6;
    const R deltaf = (c.f_amp_match_im - c.f_amp_match_in) / (double)(n_colloc - 1);
    const int n = c.tag == 21 ? 5 : 8;
    const double seven_sixths = 7.0 / 6.0;
    const R center = 0.5 * (c.f_amp_match_im + c.f_amp_match_in);
    const R half = 0.5 * (c.f_amp_match_im - c.f_amp_match_in);
    c.inter_center = center;
    c.inter_half = half;
    R a[8][8], values[8];
    int row = 0;
    auto value_row = [&](const R& f) {
        const R m76 = m_pow(f, -seven_sixths);
        const R u = (f - center) / half;
#pragma unroll
        for (int j = 0; j < n; ++j) a[row][j] = m76 * int_pow(u, j);
    };
    auto derivative_row = [&](const R& f) {
        const R m76 = m_pow(f, -seven_sixths);
        const R u = (f - center) / half;
#pragma unroll
        for (int j = 0; j < n; ++j)
            a[row][j] =
                m76 * (-seven_sixths / f * int_pow(u, j) + (j ? R((double)j * int_pow(u, j - 1) / half) : R(0.0)));
    };
    const R f_in = c.f_amp_match_in;
    const Jet2<R> ins = inspiral_amp(useful_powers(Jet2<R>(f_in, R(1.0), R(0.0))), c.pn_re, c.pn_im, c.pn_global,
                                     c.amp_norm, c.pn_dominant, c.fcut73, c.fcut83, c.fcut3, c.ins);
    const R f_rd = c.f_amp_match_im;
    R rd_value, rd_derivative;
    if (c.mixing) {
        Jet2<R> re, im;
        spheroidal_to_spherical(Jet2<R>(f_rd, R(1.0), R(0.0)), c, amp22, ph22, re, im);
        const Jet2<R> h = m_hypot(re, im);
        rd_value = h.v;
        rd_derivative = h.d;
    } else {
        rd_value = rd_amp(f_rd, c.f_ring, c.f_damp, c.f_amp_rd_falloff, false, c.rd, c.rd_aux, c.f_rd_aux,
                          c.f_amp_rd_falloff);
        rd_derivative = hm_dlorentzian(f_rd, c.f_ring, c.f_damp, c.rd);
    }
#pragma unroll
    for (int i = 0; i < n_colloc; ++i) {
        const int digit = version[i];
        const R f = m_fma(deltaf, (double)i, f_in);
        if (digit == 0) continue;
        if (i == 0) {
            value_row(f);
            values[row++] = ins.v;
            if (digit == 2) {
                derivative_row(f);
                values[row++] = ins.d;
            }
        } else if (i == n_colloc - 1) {
            value_row(f);
            values[row++] = rd_value;
            if (digit == 2) {
                derivative_row(f);
                values[row++] = rd_derivative;
            }
        } else {
            value_row(f);
            R fit;
            switch (c.tag * 10 + (i - 1)) {
                case 210: fit = ham_amp_21_int1(wf); break;
                case 211: fit = ham_amp_21_int2(wf); break;
                case 212: fit = ham_amp_21_int3(wf); break;
                case 213: fit = ham_amp_21_int4(wf); break;
                case 330: fit = ham_amp_33_int1(wf); break;
                case 331: fit = ham_amp_33_int2(wf); break;
                case 332: fit = ham_amp_33_int3(wf); break;
                case 333: fit = ham_amp_33_int4(wf); break;
                case 320: fit = ham_amp_32_int1(wf); break;
                case 321: fit = ham_amp_32_int2(wf); break;
                case 322: fit = ham_amp_32_int3(wf); break;
                case 323: fit = ham_amp_32_int4(wf); break;
                case 440: fit = ham_amp_44_int1(wf); break;
                case 441: fit = ham_amp_44_int2(wf); break;
                case 442: fit = ham_amp_44_int3(wf); break;
                default: fit = ham_amp_44_int4(wf); break;
            }
            values[row++] = m_abs(fit);
        }
    }
    if (n == 5) {
        R a5[5][5], b5[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            b5[i] = values[i];
#pragma unroll
            for (int j = 0; j < 5; ++j) a5[i][j] = a[i][j];
        }
        pivoted_solve<R, 5>(a5, b5);
#pragma unroll
        for (int j = 0; j < 5; ++j) c.inter[j] = b5[j];
    } else {
        pivoted_solve<R, 8>(a, values);
#pragma unroll
        for (int j = 0; j < 8; ++j) c.inter[j] = values[j];
    }
    c.n_inter = n;
#pragma unroll
    for (int j = n; j < 8; ++j) c.inter[j] = R(0.0);
}

// -------------------------------------------------------------- phase

// inspiral.py:319-352 the inspiral phase (the (2,2)'s rescaled) and its
// derivative
WFX_CONST double kHmInspiralExponents[13] = {-5.0 / 3.0, -4.0 / 3.0, -1.0, -2.0 / 3.0, -1.0 / 3.0, 0.0, 1.0 / 3.0,
                                             2.0 / 3.0, 1.0, 4.0 / 3.0, 5.0 / 3.0, 2.0, 7.0 / 3.0};
template <class R, class C>
WFX_HD inline R inspiral_phase_int(const R& mf, const Powers<R>& p, const C* phi, const C* phi_l) {
    const R freqs[13] = {p.m_five_thirds, p.m_four_thirds, p.m_one, p.m_two_thirds, p.m_one_third, R(1.0),
                         p.one_third, p.two_thirds, mf, p.four_thirds, p.five_thirds, p.two, p.seven_thirds};
    R philm = R(0.0);
#pragma unroll
    for (int i = 0; i < 13; ++i) philm = m_fma(m_fma(phi_l[i], p.log, phi[i]), freqs[i], philm);
    return philm;
}
template <class R, class C>
WFX_HD inline R inspiral_phase(const R& mf, const Powers<R>& p, const C* phi, const C* phi_l) {
    const R freqs[13] = {p.m_eight_thirds, p.m_seven_thirds, p.m_two, p.m_five_thirds, p.m_four_thirds,
                         p.m_one, p.m_two_thirds, p.m_one_third, R(1.0), p.one_third, p.two_thirds, mf,
                         p.four_thirds};
    R dphilm = R(0.0);
#pragma unroll
    for (int i = 0; i < 13; ++i) {
        const R coefficient = m_fma(phi_l[i], p.log, phi[i]);
        dphilm = m_fma(m_fma(coefficient, kHmInspiralExponents[i], phi_l[i]), freqs[i], dphilm);
    }
    return dphilm;
}
// phase.py:83-111 the intermediate phase derivative and phase
template <class R>
WFX_HD inline R inter_phase(const R& ff, const Powers<R>& p, const HmCoeffs<R>& c) {
    const R lorentzian = (c.c_l * c.f_damp) / m_fma(c.f_damp, c.f_damp, (ff - c.f_ring) * (ff - c.f_ring));
    R dphase = m_fma(c.c1, p.m_one, c.c0);
    dphase = m_fma(c.c2, p.m_two, dphase);
    dphase = m_fma(c.c4, p.m_four, dphase);
    if (c.tag == 32) dphase = m_fma(c.c3, p.m_three, dphase);
    return dphase + lorentzian;
}
template <class R>
WFX_HD inline R inter_phase_int(const R& ff, const Powers<R>& p, const HmCoeffs<R>& c) {
    R phase_ir = m_fma(c.c0, ff, c.c1 * p.log);
    phase_ir = m_fma(-c.c2, p.m_one, phase_ir);
    phase_ir = m_fma(-(1.0 / 3.0 * c.c4), p.m_three, phase_ir);
    if (c.tag == 32) phase_ir = m_fma(-(0.5 * c.c3), p.m_two, phase_ir);
    return m_fma(c.c_l, m_atan((ff - c.f_ring) / c.f_damp), phase_ir);
}

// phase.py:205-219 IMRPhenomXHM_PN21AmpSign (a selection: on the values)
template <class R>
WFX_HD inline bool pn21_amp_sign(double ff, const WaveformVars<R>& wf) {
    const double eta = value_of(wf.eta), chi1 = value_of(wf.chi1l), chi2 = value_of(wf.chi2l);
    const double delta = sqrt(m_fma(-4.0, eta, 1.0));
    const double sqrt5 = sqrt(5.0);
    double output = (-16.0 * delta * eta * ff * pow(kPi, 1.5)) / (3.0 * sqrt5);
    output = output + (4.0 * pow(2.0, 0.3333333333333333) * (chi1 - chi2 + delta * (chi1 + chi2)) * eta *
                       pow(ff, 1.3333333333333333) * pow(kPi, 1.8333333333333333)) / sqrt5;
    output = output + (2.0 * pow(2.0, 0.6666666666666666) * eta * (306.0 * delta - 360.0 * delta * eta) *
                       pow(ff, 1.6666666666666667) * pow(kPi, 2.1666666666666665)) / (189.0 * sqrt5);
    return output >= 0.0;
}

// phase.py:222-229 the intermediate collocation values and lambda
template <class R>
WFX_HD inline R phase_fit(const WaveformVars<R>& wf, int tag, int i) {
    switch (tag * 10 + i) {
        case 210: return hpi_phase_21_p1(wf);
        case 211: return hpi_phase_21_p2(wf);
        case 212: return hpi_phase_21_p3(wf);
        case 213: return hpi_phase_21_p4(wf);
        case 214: return hpi_phase_21_p5(wf);
        case 215: return hpi_phase_21_p6(wf);
        case 330: return hpi_phase_33_p1(wf);
        case 331: return hpi_phase_33_p2(wf);
        case 332: return hpi_phase_33_p3(wf);
        case 333: return hpi_phase_33_p4(wf);
        case 334: return hpi_phase_33_p5(wf);
        case 335: return hpi_phase_33_p6(wf);
        case 320: return hpi_phase_32_p1(wf);
        case 321: return hpi_phase_32_p2(wf);
        case 322: return hpi_phase_32_p3(wf);
        case 323: return hpi_phase_32_p4(wf);
        case 324: return hpi_phase_32_p5(wf);
        case 325: return hpi_phase_32_p6(wf);
        case 440: return hpi_phase_44_p1(wf);
        case 441: return hpi_phase_44_p2(wf);
        case 442: return hpi_phase_44_p3(wf);
        case 443: return hpi_phase_44_p4(wf);
        case 444: return hpi_phase_44_p5(wf);
        default: return hpi_phase_44_p6(wf);
    }
}

// reference.py:397-563 IMRPhenomXHM_GetPhaseCoefficients (122022, exact
// derivatives of the rotated (3,2) ringdown phase)
template <class R>
WFX_HDI inline void phase_coeffs_hm(const WaveformVars<R>& wf, const PhaseCoeffs<R>& ph22,
                                   const AmpCoeffs<R>& amp22,
                                   HmCoeffs<R>& c) {
    const int emm = c.emm, tag = c.tag;
    const R eta = wf.eta, eta_m1 = 1.0 / eta;
    const R fring = c.f_ring, fdamp = c.f_damp;
    const R fcut = m_fma(0.001, 0.25 / wf.eta - 1.0, 1.0) * (emm * 1.0) * 0.5 * wf.f_meco;  // GetfcutInsp
    const double sqrt3 = sqrt(3.0);
    R freqs[6], f_phase_match_im;
    if (tag == 32) {
        const R f_end = m_fma(-0.5, wf.f_damp, wf.f_ring);
        freqs[0] = fcut;
        freqs[1] = m_fma(sqrt3, fcut - f_end, 2.0 * (fcut + f_end)) / 4.0;
        freqs[2] = m_fma(3.0, fcut, f_end) / 4.0;
        freqs[3] = (fcut + f_end) / 2.0;
        freqs[4] = f_end;
        freqs[5] = f_end;
        f_phase_match_im = (eta < 0.01 && wf.chi1l < 0.0) ? R(f_end * m_fma(-0.25, wf.chi1l, 1.2)) : f_end;
    } else {
        freqs[0] = fcut;
        freqs[1] = m_fma(sqrt3, fcut - fring, 2.0 * (fcut + fring)) / 4.0;
        freqs[2] = m_fma(3.0, fcut, fring) / 4.0;
        freqs[3] = (fcut + fring) / 2.0;
        freqs[4] = m_fma(3.0, fring, fcut) / 4.0;
        freqs[5] = m_fma(7.0, fring, fcut) / 8.0;
        f_phase_match_im = fring - fdamp;
    }
    const R f_phase_match_in = c.f_meco_lm;
    R values[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) values[i] = phase_fit(wf, tag, i) + c.delta_t;
    R lambda_fit;
    switch (tag) {
        case 21: lambda_fit = hpn_phase_21_lambda(wf); break;
        case 33: lambda_fit = hpn_phase_33_lambda(wf); break;
        case 32: lambda_fit = hpn_phase_32_lambda(wf); break;
        default: lambda_fit = hpn_phase_44_lambda(wf); break;
    }
    const R lambda_pn = wf.eta > 0.01 ? hpn_lambda_pn(wf.eta, tag) : lambda_fit;
    const R fcut_rd = f_phase_match_im, fcut_insp = f_phase_match_in;

    // inspiral: the (2,2) coefficients rescaled, phi_lm(f) ~ m/2 phi_22(2 f / m)
    const R zero = R(0.0);
    const R non_log[13] = {ph22.phi0, zero, ph22.phi2, ph22.phi3, ph22.phi4, zero, ph22.phi6,
                           ph22.phi7, ph22.phi8, zero, zero, zero, zero};
    const R log_terms[13] = {zero, zero, zero, zero, zero, ph22.phi5l, ph22.phi6l, zero, ph22.phi8l, zero,
                             zero, zero, zero};
    const R pseudo_pn[13] = {zero, zero, zero, zero, zero, zero, zero, zero, ph22.sigma1, ph22.sigma2,
                             ph22.sigma3, ph22.sigma4, zero};
    const double m_over_2 = emm * 0.5, two_over_m = 1.0 / m_over_2;
    const R fact = ph22.phi_norm / eta;
    const double log_m_over_2 = kHmLogMOver2[c.mode_int];
#pragma unroll
    for (int i = 0; i < 13; ++i) {
        const R nl = non_log[i] * fact, lg = log_terms[i] * fact, ps = pseudo_pn[i] * fact;
        const double scale = kHmInspiralScale[c.mode_int][i];
        c.phi[i] = m_fma(-lg, log_m_over_2, nl + ps) * scale;
        c.phi_l[i] = lg * scale;
    }
    c.phi[8] = c.phi[8] + lambda_pn;
    c.alpha0 = c.alpha2 = c.alpha_l = R(0.0);
    c.c0 = c.c1 = c.c2 = c.c3 = c.c4 = c.c_l = R(0.0);

    if (tag == 21) {
        // for high spins the (2,1) inspiral extends into the intermediate region
        PhaseCoeffs<R> unconnected = ph22;
        unconnected.c1_int = unconnected.c2_int = unconnected.c1_mrd = unconnected.c2_mrd = R(0.0);
        R insp[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const R ff = two_over_m * freqs[k];
            insp[k] = 1.0 / eta * dphase_22(ff, useful_powers(ff), unconnected);
        }
        const R diff12 = insp[0] - insp[1], diff23 = insp[1] - insp[2];
        const bool high = wf.s_tot_r >= 0.8;
        const R v1 = values[2] + diff23;
        values[1] = high ? v1 : values[1];
        values[0] = high ? R(v1 + diff12) : values[0];
    }
    R x[6], phi0_rd, dphi0_rd;
    if (tag != 32) {
        bool cond_a = eta < 0.05;
        if (emm == c.ell) cond_a = cond_a || wf.s_tot_r >= 0.8;
        if (tag == 33) cond_a = cond_a || wf.s_tot_r < 0.0;
        const bool cond_b = tag == 21 && wf.s_tot_r >= 0.8;
        const int idx2 = cond_a ? 3 : 2, idx3 = cond_a ? 4 : (cond_b ? 4 : 3);
        const int chosen[5] = {0, 1, idx2, idx3, 5};
        const double wlm = emm == c.ell ? 2.0 : emm / 3.0;
        c.alpha2 = 1.0 / (fring * fring) * wlm * hpr_phase_22_alpha2(wf);
        c.alpha_l = eta_m1 * hpr_phase_22_alpha_l(wf);
        const Powers<R> p_rd = useful_powers(fcut_rd);
        phi0_rd = rd_phase_int(fcut_rd, p_rd, fring, fdamp, false, c.alpha0, c.alpha2, c.alpha_l, R(0.0), R(0.0),
                               R(0.0), R(0.0), R(0.0));
        dphi0_rd = rd_phase(fcut_rd, p_rd, fring, fdamp, false, c.alpha0, c.alpha2, c.alpha_l, R(0.0), R(0.0), R(0.0),
                            R(0.0));
        R a[5][5], b[5];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const R ff = freqs[chosen[i]];
            const R ffm1 = 1.0 / ff, ffm2 = ffm1 * ffm1;
            a[i][0] = 1.0;
            a[i][1] = fdamp / (fdamp * fdamp + (ff - fring) * (ff - fring));
            a[i][2] = ffm1;
            a[i][3] = ffm2;
            a[i][4] = ffm2 * ffm2;
            b[i] = values[chosen[i]];
        }
        pivoted_solve<R, 5>(a, b);
#pragma unroll
        for (int j = 0; j < 5; ++j) x[j] = b[j];
        x[5] = R(0.0);
    } else {
        // the rotated ringdown phase with its first and second derivatives
        Jet2<R> re, im;
        spheroidal_to_spherical(Jet2<R>(fcut_rd, R(1.0), R(0.0)), c, amp22, ph22, re, im);
        const Jet2<R> rotated = m_atan2(im, re);
        phi0_rd = rotated.v;
        dphi0_rd = rotated.d;
        const bool comparable = eta > 0.05;
        freqs[4] = comparable ? fcut_rd : freqs[4];
        values[4] = comparable ? dphi0_rd : values[4];
        freqs[5] = fcut_rd;
        values[5] = rotated.dd;
        // the quartic in v = (1/f - g_c) / g_h, converted to powers of 1/f
        const R g_c = 0.5 * (1.0 / fcut_insp + 1.0 / fcut_rd);
        const R g_h = 0.5 * (1.0 / fcut_insp - 1.0 / fcut_rd);
        R a[6][6], b[6];
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            const R ff = freqs[i];
            const R v = (1.0 / ff - g_c) / g_h;
            a[i][0] = 1.0;
            a[i][1] = fdamp / (fdamp * fdamp + (ff - fring) * (ff - fring));
            a[i][2] = v;
            a[i][3] = v * v;
            a[i][4] = v * v * v;
            a[i][5] = v * v * v * v;
            b[i] = values[i];
        }
        const R ff = freqs[5];
        const R v = (1.0 / ff - g_c) / g_h;
        const R dv = -1.0 / (ff * ff * g_h);
        const R den = fdamp * fdamp + (ff - fring) * (ff - fring);
        a[5][0] = 0.0;
        a[5][1] = -2.0 * fdamp * (ff - fring) / (den * den);
        a[5][2] = dv;
        a[5][3] = 2.0 * v * dv;
        a[5][4] = 3.0 * v * v * dv;
        a[5][5] = 4.0 * v * v * v * dv;
        b[5] = values[5];
        pivoted_solve<R, 6>(a, b);
        const R quartic[5] = {b[0], b[2], b[3], b[4], b[5]};
        const double binom[5][5] = {{1, 0, 0, 0, 0}, {1, 1, 0, 0, 0}, {1, 2, 1, 0, 0}, {1, 3, 3, 1, 0},
                                    {1, 4, 6, 4, 1}};
        R pc[5] = {R(0.0), R(0.0), R(0.0), R(0.0), R(0.0)};
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const R scale = quartic[k] / int_pow(g_h, k);
#pragma unroll
            for (int j = 0; j <= k; ++j) pc[j] = pc[j] + scale * binom[k][j] * int_pow(-g_c, k - j);
        }
        x[0] = pc[0], x[1] = b[1], x[2] = pc[1], x[3] = pc[2], x[4] = pc[4], x[5] = pc[3];
    }
    c.c0 = x[0], c.c_l = x[1], c.c1 = x[2], c.c2 = x[3], c.c4 = x[4];
    if (tag == 32) {
        c.c3 = x[5];
        const Powers<R> p_rd = useful_powers(fcut_rd);
        const R glued = c.c0 + dphi0_rd - inter_phase(fcut_rd, p_rd, c);
        c.c0 = eta < 0.05 ? glued : c.c0;
    }
    const Powers<R> p_in = useful_powers(fcut_insp);
    c.c1_insp = inter_phase(fcut_insp, p_in, c) - inspiral_phase(fcut_insp, p_in, c.phi, c.phi_l);
    c.c_insp = m_fma(-c.c1_insp, fcut_insp, inter_phase_int(fcut_insp, p_in, c)) -
               inspiral_phase_int(fcut_insp, p_in, c.phi, c.phi_l);
    const Powers<R> p_rd = useful_powers(fcut_rd);
    c.c1_rd = inter_phase(fcut_rd, p_rd, c) - dphi0_rd;
    c.c_rd = m_fma(-c.c1_rd, fcut_rd, inter_phase_int(fcut_rd, p_rd, c)) - phi0_rd;

    // align the mode with the (2,2) at low frequency (Eq. 4.13 of arXiv:2001.10914)
    const R falign = eta > 0.05 ? R(0.6 * m_over_2 * wf.f_meco) : R(m_over_2 * wf.f_meco);
    const Powers<R> p_falign = useful_powers(falign);
    const R x22 = two_over_m * falign;
    const R phi22 = phase_22(x22, useful_powers(x22), ph22) / eta + c.phiref22;
    const R phi_lm = m_fma(c.c1_insp, falign, inspiral_phase_int(falign, p_falign, c.phi, c.phi_l)) + c.c_insp;
    R deltaphi = m_fma(-(3.0 * (kPi / 4.0)), 1.0 - m_over_2, m_fma(m_over_2, phi22, c.timeshift * falign)) - phi_lm;
    if (tag == 21) deltaphi = deltaphi + (pn21_amp_sign(0.008, wf) ? kPi : 0.0);
    c.deltaphi = deltaphi;
    c.f_phase_match_in = f_phase_match_in;
    c.f_phase_match_im = f_phase_match_im;
}

// reference.py:580-608 mode_setup: the coefficients of higher mode k
template <class R>
WFX_HDI inline void hm_setup(const WaveformVars<R>& wf, const AmpCoeffs<R>& amp22, const PhaseCoeffs<R>& ph22,
                            int k, HmCoeffs<R>& c) {
    mode_vars(wf, k, c);
    c.f_ring22 = wf.f_ring;
    c.f_damp22 = wf.f_damp;
    c.inveta = 1.0 / wf.eta;
    const R timeshift = time_shift(ph22, wf);
    const R mf_ref = wf.f_ref * wf.m_sec;
    c.timeshift = timeshift;
    c.phiref22 = m_fma(-1.0 / wf.eta, phase_22(mf_ref, useful_powers(mf_ref), ph22), -(timeshift * mf_ref)) +
                 2.0 * wf.phi0 + kPi / 4.0;
    c.alpha0_s = c.alpha_l_s = c.alpha2_s = c.alpha4_s = c.phi0_s = R(0.0);
    if (c.mixing) spheroidal_coeffs(wf, ph22, c);
    // amplitude_coefficients (reference.py:348-360)
    c.amp_norm = wf.amp_norm;
    c.pn_dominant = wf.amp_norm * kHmTwoOverMPow[c.mode_int];
    c.f_rd_aux = c.mixing ? R(c.f_ring - c.f_damp) : R(0.0);
    // IMRPhenomXHM_Amplitude_fcutInsp and _fcutRD (reference.py:130-154)
    {
        const double emm = (double)c.emm;
        const R eta = wf.eta, chi1 = wf.chi1l;
        const R fcut_emr =
            1.25 * emm *
            (m_fma(0.000316707064291237, chi1 * chi1, m_fma(-0.0000858396080377194, chi1, 0.011671068725758493)) *
             m_fma(6.2873167352395125, eta, 0.8447212540381764)) /
            m_fma(-0.9977728883419751, chi1, 1.2857082764038923);
        const double transition_eta = 0.0192234, sharpness = 0.004;
        const R funcs = m_fma(0.5, m_tanh((eta - transition_eta) / sharpness), 0.5);
        c.f_amp_match_in = wf.q < 20.0 ? c.f_meco_lm : m_fma(funcs, c.f_meco_lm, (1.0 - funcs) * fcut_emr);
        c.f_amp_match_im = c.mixing ? m_fma(-0.5, wf.f_damp, wf.f_ring) : R(c.f_ring - c.f_damp);
    }
    pn_amp_coeffs(wf, c);
    inspiral_amp_coeffs(wf, c);
    rd_amp_coeffs(wf, c);
    intermediate_amp_coeffs(wf, amp22, ph22, c);
    phase_coeffs_hm(wf, ph22, amp22, c);
}

// the record of a mode
// log M, M^(1/3) and 1 / M (both records), which give the powers of M f
// from those of f; the split record's FP64 leading term and inspiral
// constant (two floats each)
enum HmMassSlot : int { kHmT0F64 = 112, kHmKInsF64 = 114, kHmLogM = 116, kHmCbrtM = 117, kHmInvM = 118 };

template <class R>
WFX_HDI inline void write_mode_record(const HmCoeffs<R>& c, const WaveformVars<R>& wf, R* r) {
#pragma unroll
    for (int s = 0; s < kHmSlots; ++s) r[s] = R(0.0);
    r[kHmTag] = R((double)c.tag);
    r[kHmOn] = R(c.amp_zero ? 0.0 : 1.0);
    r[kHmAmp0] = (c.ell % 2 ? -1.0 : 1.0) * wf.amp0;
    r[kHmMsec] = wf.m_sec;
    r[kHmFRing] = c.f_ring;
    r[kHmFDamp] = c.f_damp;
    r[kHmFalloff] = c.f_amp_rd_falloff;
    r[kHmFalloffAt] = c.f_amp_rd_falloff;
    const Powers<R> pm = useful_powers(wf.m_sec);
    r[kHmLogM] = pm.log;
    r[kHmCbrtM] = pm.one_third;
    r[kHmInvM] = pm.m_one;
    r[kHmMixing] = R(c.mixing ? 1.0 : 0.0);
    r[kHmFAmpIn] = c.f_amp_match_in;
    r[kHmFAmpIm] = c.f_amp_match_im;
    r[kHmFPhaseIn] = c.f_phase_match_in;
    r[kHmFPhaseIm] = c.f_phase_match_im;
#pragma unroll
    for (int k = 0; k < 7; ++k) r[kHmPnRe + k] = c.pn_re[k], r[kHmPnIm + k] = c.pn_im[k];
    r[kHmPnGlobal] = c.pn_global;
    r[kHmAmpNorm] = c.amp_norm;
    r[kHmPnDominant] = c.pn_dominant;
    r[kHmFcut73] = c.fcut73;
    r[kHmFcut83] = c.fcut83;
    r[kHmFcut3] = c.fcut3;
#pragma unroll
    for (int k = 0; k < 3; ++k) r[kHmIns + k] = c.ins[k];
#pragma unroll
    for (int k = 0; k < 5; ++k) r[kHmRd + k] = c.rd[k];
#pragma unroll
    for (int k = 0; k < 4; ++k) r[kHmRdAux + k] = c.rd_aux[k];
    r[kHmFRdAux] = c.f_rd_aux;
    r[kHmInterC] = c.inter_center;
    r[kHmInterH] = c.inter_half;
#pragma unroll
    for (int k = 0; k < 8; ++k) r[kHmInter + k] = c.inter[k];
    r[kHmNInter] = R((double)c.n_inter);
#pragma unroll
    for (int k = 0; k < 13; ++k) r[kHmPhi + k] = c.phi[k], r[kHmPhiL + k] = c.phi_l[k];
    r[kHmAlpha0] = c.alpha0;
    r[kHmAlpha2] = c.alpha2;
    r[kHmAlphaL] = c.alpha_l;
    r[kHmC0] = c.c0;
    r[kHmC1] = c.c1;
    r[kHmC2] = c.c2;
    r[kHmC3] = c.c3;
    r[kHmC4] = c.c4;
    r[kHmCL] = c.c_l;
    r[kHmC1Insp] = c.c1_insp;
    r[kHmCInsp] = c.c_insp;
    r[kHmC1Rd] = c.c1_rd;
    r[kHmCRd] = c.c_rd;
    r[kHmDeltaPhi] = c.deltaphi;
    r[kHmSph + 0] = c.alpha0_s;
    r[kHmSph + 1] = c.alpha_l_s;
    r[kHmSph + 2] = c.alpha2_s;
    r[kHmSph + 3] = c.alpha4_s;
    r[kHmSph + 4] = c.phi0_s;
    r[kHmMix + 0] = c.mix2_re;
    r[kHmMix + 1] = c.mix2_im;
    r[kHmMix + 2] = c.mix3_re;
    r[kHmMix + 3] = c.mix3_im;
}

// --------------------------------------------------------- split record

// The FP32 record of a mode for the split evaluation: the FP64 record's
// values rounded, except
//  * the leading inspiral term phi_0 (M f)^(-5/3) = T0_lm f^(-5/3) and the
//    inspiral's constant c_insp + dphi_lm, kept in FP64 (two float slots
//    each): the evaluation sums them in FP64 and reduces modulo 2 pi before
//    adding the rest of the inspiral phase in FP32;
//  * the intermediate and ringdown constants (dphi_lm, c_rd + dphi_lm) and
//    the spheroidal phase constant phi0_s reduced modulo 2 pi in FP64: past
//    the inspiral the phase is an FP32 sum of terms of <~1e3 rad;
//  * quotients folded as the evaluation uses them (the pseudo-PN
//    coefficients over their cut-off powers, 1 / f_h, the Lorentzian's
//    rd_0 f_damp, rd_1 / (f_damp rd_2), (f_damp rd_2)^2).
// With dual numbers, tangent (if not null) receives each slot's derivative
// in the dual's first direction, as float (the FP64 pairs' in their first
// word, zero in the second; the reductions modulo 2 pi carry none).

template <class R>
WFX_HD inline R fold_two_pi(const R& x) {
    return x - kTwoPi * rint(value_of(x) * (1.0 / kTwoPi));
}
WFX_HD inline void store_f64(float* r, int slot, double v) { __builtin_memcpy(r + slot, &v, sizeof(double)); }

WFX_HD inline void store_f64(double* r, int slot, double v) {
    r[slot] = v;
    r[slot + 1] = 0.0;
}

// O = float: the split evaluation's record; O = double: the same layout in
// FP64 (the FP64 gradient's), the FP64 slots as one double and a zero
template <class R, class O>
WFX_HDI inline void write_mode_record_split(const HmCoeffs<R>& c, const WaveformVars<R>& wf, O* r,
                                           O* tangent = nullptr) {
    R d[kHmSlots];
    write_mode_record(c, wf, d);
    d[kHmPhi] = R(0.0);  // (the leading term is the FP64 one)
    d[kHmIns + 0] = c.ins[0] / c.fcut73;
    d[kHmIns + 1] = c.ins[1] / c.fcut83;
    d[kHmIns + 2] = c.ins[2] / c.fcut3;
    d[kHmInterH] = 1.0 / c.inter_half;
    const R dfd = c.f_damp * c.rd[2];
    d[kHmRd + 0] = c.rd[0] * c.f_damp;
    d[kHmRd + 1] = c.rd[1] / dfd;
    d[kHmRd + 2] = dfd * dfd;
    d[kHmDeltaPhi] = fold_two_pi(c.deltaphi);
    d[kHmCRd] = fold_two_pi(c.c_rd + c.deltaphi);
    d[kHmSph + 4] = fold_two_pi(c.phi0_s);
    const Powers<R> pm = useful_powers(wf.m_sec);
    const R t0 = c.phi[0] * pm.m_five_thirds;
    const R k_ins = c.c_insp + c.deltaphi;
#pragma unroll
    for (int s = 0; s < kHmSlots; ++s) r[s] = (O)value_of(d[s]);
    store_f64(r, kHmT0F64, value_of(t0));
    store_f64(r, kHmKInsF64, value_of(k_ins));
    if (tangent) {
#pragma unroll
        for (int s = 0; s < kHmSlots; ++s) tangent[s] = (O)tangent_of(d[s]);
        tangent[kHmT0F64] = (O)tangent_of(t0);
        tangent[kHmT0F64 + 1] = O(0);
        tangent[kHmKInsF64] = (O)tangent_of(k_ins);
        tangent[kHmKInsF64 + 1] = O(0);
    }
}

// ------------------------------------------------------------ harmonics

// _harmonics.py:17-38 Wigner's d^l_{m' m}(beta) and {}_{-2}Y_{lm} (generic
// in beta's scalar, for the inclination's derivative)
WFX_HD inline double factorial_d(int n) {
    double f = 1.0;
    for (int k = 2; k <= n; ++k) f *= k;
    return f;
}
template <class R>
WFX_HD inline R wigner_d(int ell, int mp, int m, const R& beta) {
    const R cos_half = m_cos(beta / 2), sin_half = m_sin(beta / 2);
    R total = R(0.0);
    const double prefactor = sqrt(factorial_d(ell + mp) * factorial_d(ell - mp) * factorial_d(ell + m) *
                                  factorial_d(ell - m));
    const int k0 = m - mp > 0 ? m - mp : 0;
    const int k1 = ell + m < ell - mp ? ell + m : ell - mp;
    for (int k = k0; k <= k1; ++k) {
        const double sign = ((k - m + mp) % 2 == 0) ? 1.0 : -1.0;
        const double coefficient = sign * prefactor / (factorial_d(ell + m - k) * factorial_d(k) *
                                                       factorial_d(ell - k - mp) * factorial_d(k - m + mp));
        total = total + coefficient * int_pow(cos_half, 2 * ell - 2 * k + m - mp) * int_pow(sin_half, 2 * k - m + mp);
    }
    return total;
}
template <class R>
WFX_HD inline void swsh(const R& theta, double phi, int ell, int m, R& re, R& im) {
    const R d = sqrt((2 * ell + 1) / (4 * kPi)) * wigner_d(ell, m, 2, theta);
    const double arg = m * phi;
    re = d * cos(arg);
    im = d * sin(arg);
}
// reference.py:646-653 (F+, Fx) of h_{l,-m}: 0.5 (Y_{l,-m} + (-1)^l conj Y_{lm}),
// 0.5 i (Y_{l,-m} - (-1)^l conj Y_{lm}), at phi = pi / 2
template <class R>
WFX_HD inline void mode_factors(const R& inclination, int ell, int emm, R* out) {
    const double sign = ell % 2 ? -1.0 : 1.0;
    R ym_re, ym_im, y_re, y_im;
    swsh(inclination, kPi / 2.0, ell, -emm, ym_re, ym_im);
    swsh(inclination, kPi / 2.0, ell, emm, y_re, y_im);
    const R ys_re = y_re, ys_im = -y_im;  // conj
    const R sum_re = ym_re + sign * ys_re, sum_im = ym_im + sign * ys_im;
    const R dif_re = ym_re - sign * ys_re, dif_im = ym_im - sign * ys_im;
    out[0] = 0.5 * sum_re;
    out[1] = 0.5 * sum_im;
    // (0 + 0.5 i) (dif_re + i dif_im), complex multiplication
    out[2] = 0.0 * dif_re - 0.5 * dif_im;
    out[3] = 0.0 * dif_im + 0.5 * dif_re;
}

}  // namespace wfx
