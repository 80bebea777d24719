// IMRPhenomXHM (release 122022) as LALSimulation evaluates it, in Mf: the
// CPU oracle of the higher modes. Test infrastructure, included at the end
// of xas_oracle.cpp (it uses that file's waveform variables, (2,2)
// coefficients, LU solve and scalar types); nothing here is product code.
//
// It follows the reference's LAL form, op for op:
//   src/waveformx/models/xhm/reference.py (set-up, amplitude 122022, phase,
//     the (3,2) rotation, evaluation, mode factors),
//   src/waveformx/models/xpnr/hm/variables.py:118-169 (mode variables),
//   src/waveformx/models/xpnr/hm/inspiral.py:49-215,319-361 (PN amplitude
//     coefficients, inspiral phase ansatzes),
//   src/waveformx/models/xpnr/hm/phase.py:58-174,214-246 (phase ansatzes,
//     spheroidal coefficients, fcut, PN21 sign, fits),
//   src/waveformx/models/xpnr/xas.py:20-47 (the (2,2) mode recycled by the
//     (3,2) ringdown), src/waveformx/_harmonics.py:17-38 (harmonics),
// with the fits of oracle/xhm_fits.h. Where LAL differentiates by finite
// differences, the reference (and this oracle) takes exact derivatives:
// here with a second-order Taylor scalar (Jet2).
#pragma once

#include <complex>

#include "xhm_fits.h"

// ------------------------------------------------------------ Taylor scalar

// value, first and second derivative in one variable; the value is computed
// exactly as the double arithmetic computes it
struct Jet2 {
    double v, d, dd;
    Jet2() : v(0.0), d(0.0), dd(0.0) {}
    Jet2(double x) : v(x), d(0.0), dd(0.0) {}
    Jet2(double x, double dx, double ddx) : v(x), d(dx), dd(ddx) {}
};
static inline Jet2 chain(const Jet2& x, double f, double f1, double f2) {
    return Jet2(f, f1 * x.d, f2 * x.d * x.d + f1 * x.dd);
}
static inline Jet2 operator+(const Jet2& a, const Jet2& b) { return Jet2(a.v + b.v, a.d + b.d, a.dd + b.dd); }
static inline Jet2 operator-(const Jet2& a, const Jet2& b) { return Jet2(a.v - b.v, a.d - b.d, a.dd - b.dd); }
static inline Jet2 operator-(const Jet2& a) { return Jet2(-a.v, -a.d, -a.dd); }
static inline Jet2 operator*(const Jet2& a, const Jet2& b) {
    return Jet2(a.v * b.v, a.d * b.v + a.v * b.d, a.dd * b.v + 2.0 * a.d * b.d + a.v * b.dd);
}
static inline Jet2 operator/(const Jet2& a, const Jet2& b) {
    double q = a.v / b.v;
    double dq = (a.d - q * b.d) / b.v;
    return Jet2(q, dq, (a.dd - 2.0 * dq * b.d - q * b.dd) / b.v);
}
#define JMIX(op)                                                                              \
    static inline Jet2 operator op(const Jet2& a, double b) { return a op Jet2(b); }         \
    static inline Jet2 operator op(double a, const Jet2& b) { return Jet2(a) op b; }
JMIX(+) JMIX(-) JMIX(*) JMIX(/)
#undef JMIX
#define JCMP(op)                                                                              \
    static inline bool operator op(const Jet2& a, const Jet2& b) { return a.v op b.v; }      \
    static inline bool operator op(const Jet2& a, double b) { return a.v op b; }             \
    static inline bool operator op(double a, const Jet2& b) { return a op b.v; }
JCMP(<) JCMP(>) JCMP(<=) JCMP(>=) JCMP(==) JCMP(!=)
#undef JCMP
static inline double val(const Jet2& x) { return x.v; }
static inline Jet2 o_exp(const Jet2& x) { double e = std::exp(x.v); return chain(x, e, e, e); }
static inline Jet2 o_log(const Jet2& x) { return chain(x, std::log(x.v), 1.0 / x.v, -1.0 / (x.v * x.v)); }
static inline Jet2 o_sqrt(const Jet2& x) {
    double s = std::sqrt(x.v);
    return chain(x, s, 0.5 / s, -0.25 / (x.v * s));
}
static inline Jet2 o_cbrt(const Jet2& x) {
    double c = std::cbrt(x.v);
    return chain(x, c, c / (3.0 * x.v), -2.0 * c / (9.0 * x.v * x.v));
}
static inline Jet2 o_atan(const Jet2& x) {
    double r = 1.0 / (1.0 + x.v * x.v);
    return chain(x, std::atan(x.v), r, -2.0 * x.v * r * r);
}
static inline Jet2 o_cos(const Jet2& x) { return chain(x, std::cos(x.v), -std::sin(x.v), -std::cos(x.v)); }
static inline Jet2 o_sin(const Jet2& x) { return chain(x, std::sin(x.v), std::cos(x.v), -std::sin(x.v)); }
static inline Jet2 o_abs(const Jet2& x) { return x.v < 0.0 ? -x : x; }
static inline Jet2 o_pow(const Jet2& x, double y) {
    return chain(x, std::pow(x.v, y), y * std::pow(x.v, y - 1.0), y * (y - 1.0) * std::pow(x.v, y - 2.0));
}

// jnp.hypot (max |x|, |y| times sqrt(1 + (min / max)^2)) and its derivatives
static inline double o_hypot(double x, double y) {
    double a = std::fabs(x), b = std::fabs(y);
    double hi = a > b ? a : b, lo = a > b ? b : a;
    if (hi == 0.0) return hi;
    double r = lo / hi;
    return hi * std::sqrt(1.0 + r * r);
}
static inline Jet2 o_hypot(const Jet2& x, const Jet2& y) {
    double h = o_hypot(x.v, y.v);
    double s = x.v * x.d + y.v * y.d;
    double d = s / h;
    double dd = (x.d * x.d + x.v * x.dd + y.d * y.d + y.v * y.dd) / h - s * s / (h * h * h);
    return Jet2(h, d, dd);
}
static inline double o_atan2(double y, double x) { return std::atan2(y, x); }
static inline Jet2 o_atan2(const Jet2& y, const Jet2& x) {
    double r2 = x.v * x.v + y.v * y.v;
    double n = x.v * y.d - y.v * x.d;
    double n1 = x.v * y.dd - y.v * x.dd;
    double r21 = 2.0 * (x.v * x.d + y.v * y.d);
    return Jet2(std::atan2(y.v, x.v), n / r2, (n1 * r2 - n * r21) / (r2 * r2));
}

// lax.integer_pow: binary exponentiation (the reference's `x ** n` for a
// traced x and a Python int n)
template <class R>
static R ipow(const R& x, int n) {
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

// a coefficient struct of doubles as the same struct of Jet2 constants
// (the structs hold nothing but their scalar fields)
template <template <class> class S>
static S<Jet2> lift(const S<double>& s) {
    static_assert(sizeof(S<double>) % sizeof(double) == 0, "a struct of doubles");
    constexpr int n = sizeof(S<double>) / sizeof(double);
    static_assert(sizeof(S<Jet2>) == n * sizeof(Jet2), "a struct of Jet2");
    S<Jet2> out;
    const double* a = reinterpret_cast<const double*>(&s);
    Jet2* b = reinterpret_cast<Jet2*>(&out);
    for (int i = 0; i < n; ++i) b[i] = Jet2(a[i]);
    return out;
}

// --------------------------------------------------------------- the (2,2)

// phase22.py:304-307 (IMRPhenomX_Phase_22: Mf >= fPhaseMatchIM is the ringdown)
template <class R>
static R phase_22(const R& mf, const Powers<R>& p, const PhaseCoeffs<R>& ph) {
    R ins = ins_phase_22_ansatz_int_exact(mf, p, ph);
    R rd = o_fma(ph.c2_mrd, mf, rd_phase_22_ansatz_int_exact(mf, p, ph) + ph.c1_mrd);
    R itm = o_fma(ph.c2_int, mf, itm_phase_22_ansatz_int_exact(mf, p, ph) + ph.c1_int);
    return mf < ph.f_phase_match_in ? ins : (mf >= ph.f_phase_match_im ? rd : itm);
}

// ------------------------------------------------------------ the modes

static const int HM_ELL[4] = {2, 3, 3, 4};
static const int HM_EMM[4] = {1, 3, 2, 4};
// reference.py:59-60 IMRPhenomXHMIntermediateAmpVersion per collocation point
static const int HM_INTER_VERSION[4][6] = {{1, 1, 0, 1, 0, 2}, {2, 1, 1, 1, 1, 2}, {2, 1, 1, 1, 1, 2}, {2, 1, 1, 1, 1, 2}};
static const double HM_ETA_EMR = 0.05;  // variables.py:32
static const double HM_FALSE_ZERO = 1.0e-15;  // reference.py:63

struct HmVars {
    int ell, emm, tag, mode_int;
    bool mixing, amp_zero;
    double chi_s, chi_a, f_meco_lm, f_ring, f_damp, delta_t, f_amp_rd_falloff;
    double mix[4][2];  // a222, a223, -a322, -a323 as (re, im)
};
struct HmAmp {
    double amp_norm, pn_dominant, f_rd_aux, f_amp_match_in, f_amp_match_im;
    double pn_re[7], pn_im[7], pn_global_factor;
    double fcut73, fcut83, fcut3, ins[3];
    double rd[5], rd_aux[4];
    double inter_center, inter_half, inter[8];
    int n_inter;
};
struct HmSpheroidal {
    double alpha0_s, alpha_l_s, alpha2_s, alpha4_s, phi0_s;
};
struct HmPhase {
    double phi[13], phi_l[13], alpha0, alpha2, alpha_l, c0, c1, c2, c3, c4, c_l, c1_insp, c_insp, c1_rd, c_rd,
        deltaphi, f_phase_match_in, f_phase_match_im;
};
// the (2,2) data the (3,2) rotation uses (reference.py:625-633)
struct HmMix22 {
    AmpCoeffs<double> amp22;
    PhaseCoeffs<double> ph22;
    AmpCoeffs<Jet2> amp22j;
    PhaseCoeffs<Jet2> ph22j;
    double f_ring22, f_damp22, inveta, amp_norm, timeshift, phiref22;
};
static const AmpCoeffs<double>& amp22_of(const HmMix22& m, double) { return m.amp22; }
static const AmpCoeffs<Jet2>& amp22_of(const HmMix22& m, const Jet2&) { return m.amp22j; }
static const PhaseCoeffs<double>& ph22_of(const HmMix22& m, double) { return m.ph22; }
static const PhaseCoeffs<Jet2>& ph22_of(const HmMix22& m, const Jet2&) { return m.ph22j; }

struct HmMode {
    HmVars hm;
    HmAmp amp;
    HmPhase ph;
    HmSpheroidal sph;
    HmMix22 mix22;
    double alpha2_fit, alpha_l_fit, amp0;
    // the collocation conditions the solves met (for the checks): the phase
    // derivative's points and values, the intermediate amplitude's fitted points
    int n_colloc_phase, n_colloc_amp;
    double colloc_phase_f[6], colloc_phase_v[6], colloc_amp_f[6], colloc_amp_v[6];
};

// reference.py:96-99 QNM: the fits of hm_qnm.py in powers of 1 - af
static double hm_qnm(const double* num, int nn, const double* den, int nd, double af) {
    AboutOne a = about_one(num, nn), b = about_one(den, nd);
    double u = 1.0 - af;
    return horner(a.c, u) / horner(b.c, u);
}
#define HMQ(name, af) hm_qnm(hmq_##name##_num, (int)(sizeof(hmq_##name##_num) / sizeof(double)), hmq_##name##_den, \
                             (int)(sizeof(hmq_##name##_den) / sizeof(double)), af)

// variables.py:118-169 and reference.py:102-122 (mode_struct, release 122022)
static HmVars mode_struct(const WfVars<double>& wf, int k) {
    HmVars hm;
    hm.ell = HM_ELL[k];
    hm.emm = HM_EMM[k];
    hm.tag = 10 * hm.ell + hm.emm;
    hm.mode_int = k;
    hm.mixing = hm.tag == 32;
    hm.amp_zero = (hm.tag == 21 || hm.tag == 33) && wf.q == 1.0 && wf.chi1l == wf.chi2l;
    hm.chi_s = (wf.chi1l + wf.chi2l) * 0.5;
    hm.chi_a = (wf.chi1l - wf.chi2l) * 0.5;
    hm.f_meco_lm = wf.f_meco * (double)hm.emm * 0.5;
    double af = wf.a_final;
    double fr = hm.tag == 21 ? HMQ(fring21, af) : hm.tag == 33 ? HMQ(fring33, af) : hm.tag == 32 ? HMQ(fring32, af)
                                                                                                   : HMQ(fring44, af);
    double fd = hm.tag == 21 ? HMQ(fdamp21, af) : hm.tag == 33 ? HMQ(fdamp33, af) : hm.tag == 32 ? HMQ(fdamp32, af)
                                                                                                   : HMQ(fdamp44, af);
    hm.f_ring = fr / wf.m_final;
    hm.f_damp = fd / wf.m_final;
    hm.delta_t = -2.0 * M_PI_ * (500.0 + ut_psi4_to_strain(wf.eta, wf.s_tot_r, wf.dchi));
    hm.f_amp_rd_falloff = o_fma(2.0, hm.f_damp, hm.f_ring);
    for (auto& row : hm.mix) row[0] = row[1] = 0.0;
    if (hm.mixing) {
        hm.mix[0][0] = HMQ(re_l2m2lp2, af);
        hm.mix[0][1] = HMQ(im_l2m2lp2, af);
        hm.mix[1][0] = HMQ(re_l2m2lp3, af);
        hm.mix[1][1] = HMQ(im_l2m2lp3, af);
        hm.mix[2][0] = -HMQ(re_l3m2lp2, af);
        hm.mix[2][1] = -HMQ(im_l3m2lp2, af);
        hm.mix[3][0] = -HMQ(re_l3m2lp3, af);
        hm.mix[3][1] = -HMQ(im_l3m2lp3, af);
    }
    return hm;
}

// --------------------------------------------------------- PN amplitude

// inspiral.py:49-215 IMRPhenomXHM_GetPNAmplitudeCoefficients
static void pn_amplitude_coefficients(const WfVars<double>& wf, const HmVars& hm, HmAmp& amp) {
    const double chi_a = hm.chi_a, chi_s = hm.chi_s, eta = wf.eta, delta = wf.delta;
    const Powers<double>& lp = lalpi();
    const double pi = lp.itself;
    const double sqrt_5_7 = std::sqrt(5.0 / 7.0);
    const double prefactors[4] = {std::sqrt(2.0) / 3.0, 0.75 * sqrt_5_7, sqrt_5_7 / 3.0,
                                  4.0 * std::sqrt(2.0) / 9.0 * sqrt_5_7};
    amp.pn_global_factor = std::pow(2.0 / hm.emm, -7.0 / 6.0) * prefactors[hm.mode_int];
    const double zero = 0.0 * eta;
    double re[7], im[7];
    for (int k = 0; k < 7; ++k) re[k] = im[k] = zero;
    if (hm.tag == 21) {
        Powers<double> p2 = powers(2.0);
        double r = 3427.0 * chi_a;
        r = r + 3427.0 * chi_s * delta;
        r = r - 8404.0 * chi_a * eta;
        r = r - 3860.0 * chi_s * delta * eta;
        r = r - 1344.0 * delta * pi;
        double i = -(672.0 * delta);
        i = i - 672.0 * delta * std::log(16.0);
        double scale = lp.four_thirds * p2.four_thirds;
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
        r = r + 20575296.0 * std::pow(chi_a, 3.0);
        r = r - 81285120.0 * eta * std::pow(chi_a, 3.0);
        r = r + 61725888.0 * chi_a * chi_s * chi_s;
        r = r - 165618432.0 * chi_a * eta * chi_s * chi_s;
        r = r + 20575296.0 * delta * std::pow(chi_s, 3.0);
        r = r - 1016064.0 * delta * eta * chi_s * chi_s * chi_s;
        r = r + 128873808.0 * chi_a * eta * eta;
        r = r - 3859632.0 * chi_s * delta * eta * eta;
        i = -(1350720.0 * delta);
        i = i - 72043776.0 * delta * eta;
        i = i - 5402880.0 * delta * LN2_;
        i = i - 17224704.0 * delta * eta * LN2_;
        scale = lp.two * p2.two;
        re[6] = r / 5.419008e6 * scale;
        im[6] = i / 5.419008e6 * scale;
        double five = o_fma(21535920.0 * delta * eta, eta, o_fma(39626496.0 * delta * eta * chi_s, chi_s, o_fma(
            -(77982912.0 * delta * chi_s), chi_s, o_fma(81285120.0 * delta * eta * chi_a, chi_a, o_fma(
                -(77982912.0 * delta * chi_a), chi_a, o_fma(24385536.0 * chi_s * delta, pi, o_fma(
                    24385536.0 * chi_a, pi, o_fma(-(23670792.0 * delta), eta, o_fma(
                        432843264.0 * chi_a * chi_s, eta, o_fma(-155965824.0 * chi_a, chi_s, -(964357.0 * delta)))))))))));
        re[1] = delta * lp.one_third * p2.one_third;
        re[2] = -3.0 * o_fma(chi_s, delta, chi_a) / 2.0 * lp.two_thirds * p2.two_thirds;
        re[3] = o_fma(335.0, delta, 1404.0 * delta * eta) / 672.0 * lp.itself * p2.itself;
        re[5] = five / 8.128512e6 * lp.five_thirds * p2.five_thirds;
    } else if (hm.tag == 33) {
        Powers<double> p2 = powers(2.0 / 3.0);
        const double ln15 = std::log(1.5);
        double r = 325.0 * chi_a;
        r = r + 325.0 * chi_s * delta;
        r = r - 1120.0 * chi_a * eta;
        r = r - 80.0 * chi_s * delta * eta;
        r = r + 120.0 * delta * pi;
        double i = -(504.0 * delta);
        i = i + 720.0 * delta * ln15;
        double scale = lp.four_thirds * p2.four_thirds;
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
        scale = lp.two * p2.two;
        re[6] = r / 2.17728e6 * scale;
        im[6] = i / 2.17728e6 * scale;
        double five = o_fma(2966264784.0 * delta * eta, eta, o_fma(55883520.0 * delta * eta * chi_s, chi_s, o_fma(
            -(1131641280.0 * delta * chi_s), chi_s, o_fma(4470681600.0 * delta * eta * chi_a, chi_a, o_fma(
                -(1131641280.0 * delta * chi_a), chi_a, o_fma(-(5926068792.0 * delta), eta, o_fma(
                    9053130240.0 * chi_a * chi_s, eta, o_fma(-2263282560.0 * chi_a, chi_s, -(1077664867.0 * delta)))))))));
        re[1] = delta * lp.one_third * p2.one_third;
        re[3] = o_fma(-1945.0, delta, 2268.0 * delta * eta) / 672.0 * lp.itself * p2.itself;
        re[5] = five / 4.4706816e8 * lp.five_thirds * p2.five_thirds;
    } else if (hm.tag == 32) {
        double r = 0.0 - 3955.0 * chi_s;
        r = r - 3955.0 * chi_a * delta;
        r = r + 10810.0 * chi_s * eta;
        r = r + 11865.0 * chi_a * delta * eta;
        r = r - 12600.0 * chi_s * eta * eta;
        double i = 2520.0 + zero;
        i = i - 11088.0 * eta;
        double scale = lp.five_thirds;
        re[5] = r / 840.0 * scale;
        im[5] = i / 840.0 * scale;
        double six = o_fma(5907445488.0, std::pow(eta, 3.0), o_fma(5830513920.0 * chi_s * chi_s * eta, eta, o_fma(
            13412044800.0 * chi_a * chi_a * eta, eta, o_fma(-(13202119896.0 * eta), eta, o_fma(
                -(11870591040.0 * eta * chi_s), chi_s, o_fma(1131641280.0 * chi_s, chi_s, o_fma(
                    -(7865605440.0 * eta * chi_a), chi_a, o_fma(1131641280.0 * chi_a, chi_a, o_fma(
                        3576545280.0 * chi_s * eta, pi, o_fma(-(15209631360.0 * chi_a * chi_s * delta), eta, o_fma(
                            -26069649.0, eta, o_fma(2263282560.0 * chi_a * chi_s, delta, 824173699.0))))))))))));
        re[2] = o_fma(3.0, eta, -1.0) * lp.two_thirds;
        re[3] = -4.0 * chi_s * eta * lp.itself;
        re[4] = o_fma(82460.0 * eta, eta, o_fma(-61625.0, eta, 10471.0)) / 10080.0 * lp.four_thirds;
        re[6] = six / 4.4706816e8 * lp.two;
    } else {
        Powers<double> p2 = powers(0.5);
        double r = -0.0 + 565.0 * chi_s;
        r = r + 565.0 * chi_a * delta;
        r = r - 2075.0 * chi_s * eta;
        r = r - 1695.0 * chi_a * delta * eta;
        r = r + 240.0 * pi;
        r = r - 720.0 * eta * pi;
        r = r + 1140.0 * chi_s * eta * eta;
        double i = -1008.0 + zero;
        i = i + 3579.0 * eta;
        i = i + 960.0 * LN2_;
        i = i - 2880.0 * eta * LN2_;
        double scale = lp.five_thirds * p2.five_thirds;
        re[5] = r / 120.0 * scale;
        im[5] = i / 120.0 * scale;
        double six = o_fma(-805075876080.0, std::pow(eta, 3.0), o_fma(-(10897286400.0 * chi_s * chi_s * eta), eta, o_fma(
            -(871782912000.0 * chi_a * chi_a * eta), eta, o_fma(2271682065240.0 * eta, eta, o_fma(
                224302478400.0 * eta * chi_s, chi_s, o_fma(-(73556683200.0 * chi_s), chi_s, o_fma(
                    511264353600.0 * eta * chi_a, chi_a, o_fma(-(73556683200.0 * chi_a), chi_a, o_fma(
                        441340099200.0 * chi_a * chi_s * delta, eta, o_fma(-745140957231.0, eta, o_fma(
                            -(147113366400.0 * chi_a * chi_s), delta, 7888301437.0)))))))))));
        re[2] = o_fma(-3.0, eta, 1.0) * lp.two_thirds * p2.two_thirds;
        re[4] = o_fma(-(446460.0 * eta), eta, o_fma(641105.0, eta, -158383.0)) / 36960.0 * lp.four_thirds * p2.four_thirds;
        re[6] = six / 2.90594304e10 * lp.two * p2.two;
    }
    for (int k = 0; k < 7; ++k) {
        amp.pn_re[k] = re[k];
        amp.pn_im[k] = im[k];
    }
}

// ------------------------------------------------------------- amplitude

// reference.py:130-147 IMRPhenomXHM_Amplitude_fcutInsp (122022)
static double amplitude_fcut_insp(const HmVars& hm, const WfVars<double>& wf) {
    double emm = (double)hm.emm, eta = wf.eta, chi1 = wf.chi1l;
    double fcut_emr = 1.25 * emm *
                      (o_fma(0.000316707064291237, chi1 * chi1, o_fma(-0.0000858396080377194, chi1, 0.011671068725758493)) *
                       o_fma(6.2873167352395125, eta, 0.8447212540381764)) /
                      o_fma(-0.9977728883419751, chi1, 1.2857082764038923);
    const double transition_eta = 0.0192234, sharpness = 0.004;
    double funcs = o_fma(0.5, std::tanh((eta - transition_eta) / sharpness), 0.5);
    return wf.q < 20.0 ? hm.f_meco_lm : o_fma(funcs, hm.f_meco_lm, (1.0 - funcs) * fcut_emr);
}

// reference.py:150-154 IMRPhenomXHM_Amplitude_fcutRD (122022)
static double amplitude_fcut_rd(const HmVars& hm, const WfVars<double>& wf) {
    return hm.mixing ? o_fma(-0.5, wf.f_damp, wf.f_ring) : hm.f_ring - hm.f_damp;
}

// reference.py:157-166 IMRPhenomXHM_Inspiral_PNAmp_Ansatz
template <class R>
static R pn_amp_ansatz(const Powers<R>& p, const HmAmp& amp) {
    const R terms[6] = {p.one_third, p.two_thirds, p.itself, p.four_thirds, p.five_thirds, p.two};
    R re = R(amp.pn_re[0]), im = R(amp.pn_im[0]);
    for (int k = 0; k < 6; ++k) {
        re = re + terms[k] * amp.pn_re[k + 1];
        im = im + terms[k] * amp.pn_im[k + 1];
    }
    return o_hypot(re, im) * (amp.pn_global_factor * p.m_seven_sixths * amp.amp_norm);
}

// reference.py:169-176 IMRPhenomXHM_Inspiral_Amp_Ansatz (122022)
template <class R>
static R inspiral_amp_ansatz(const Powers<R>& p, const HmAmp& amp) {
    R pseudo = o_fma(p.seven_thirds / amp.fcut73, amp.ins[0], p.eight_thirds / amp.fcut83 * amp.ins[1]);
    pseudo = o_fma(p.three / amp.fcut3, amp.ins[2], pseudo);
    return pn_amp_ansatz(p, amp) + pseudo * p.m_seven_sixths * amp.pn_dominant;
}

// reference.py:179-197 (three collocation points, the constant inverse)
static void inspiral_amp_coefficients(HmAmp& amp, const HmVars& hm, const WfVars<double>& wf) {
    double fcut = amp.f_amp_match_in;
    double freqs[3] = {0.5 * fcut, 0.75 * fcut, fcut};
    double values[3];
    for (int i = 0; i < 3; ++i) {
        double v;
        switch (hm.tag * 10 + i) {
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
        values[i] = std::fabs(v);
    }
    Powers<double> finsp = powers(fcut);
    amp.fcut73 = finsp.seven_thirds;
    amp.fcut83 = finsp.eight_thirds;
    amp.fcut3 = finsp.three;
    double v[3];
    for (int i = 0; i < 3; ++i) {
        Powers<double> f = powers(freqs[i]);
        double pn = pn_amp_ansatz(f, amp);
        v[i] = (values[i] - pn) / amp.pn_dominant / f.m_seven_sixths;
    }
    for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int i = 0; i < 3; ++i) s = s + hm_inspiral_inverse[j][i] * v[i];
        amp.ins[j] = s;
    }
}

// reference.py:203-208
template <class R>
static R hm_lorentzian(const R& mf, const HmVars& hm, const double* rd) {
    double fda = hm.f_damp;
    R dfr = mf - hm.f_ring;
    double dfd = fda * rd[2];
    return rd[0] * fda / (o_exp(rd[1] / dfd * dfr) * o_fma(dfr, dfr, dfd * dfd));
}

// reference.py:211-220 IMRPhenomXHM_RD_Amp_DAnsatz
template <class R>
static R rd_amp_dansatz(const R& mf, const HmVars& hm, const double* rd) {
    double fda = hm.f_damp;
    R dfr = mf - hm.f_ring;
    R numerator = rd[0] * o_fma(fda * fda * rd[1] * rd[2], rd[2], o_fma(dfr * dfr, rd[1], 2.0 * fda * dfr * rd[2]));
    R denom = o_fma(dfr, dfr, fda * fda * rd[2] * rd[2]);
    denom = rd[2] * denom * denom * o_exp(dfr * rd[1] / (fda * rd[2]));
    return -numerator / denom;
}

// reference.py:223-236 IMRPhenomXHM_RD_Amp_Ansatz
template <class R>
static R rd_amp_ansatz(const R& mf, const HmVars& hm, const HmAmp& amp) {
    const double* rd = amp.rd;
    R out = mf >= hm.f_amp_rd_falloff ? rd[3] * o_exp(-rd[4] * (mf - hm.f_amp_rd_falloff)) : hm_lorentzian(mf, hm, rd);
    if (hm.mixing) {
        const double* a = amp.rd_aux;
        R aux = o_fma(o_fma(o_fma(R(a[3]), mf, R(a[2])), mf, R(a[1])), mf, R(a[0]));
        out = mf < amp.f_rd_aux ? aux : out;
    }
    return out;
}

// reference.py:239-272 IMRPhenomXHM_RD_Amp_Coefficients (+ RDAux for the (3,2))
static void rd_amp_coefficients(HmAmp& amp, const HmVars& hm, const WfVars<double>& wf) {
    double fdamp = hm.f_damp;
    double rd[3];
    if (hm.tag == 32) {
        rd[0] = std::fabs(har_amp_32_alambda(wf));
        rd[1] = har_amp_32_lambda(wf);
        rd[2] = har_amp_32_sigma(wf);
    } else {
        double c1, c2, c3;
        if (hm.tag == 21) {
            c1 = har_amp_21_rdcp1(wf), c2 = har_amp_21_rdcp2(wf), c3 = har_amp_21_rdcp3(wf);
        } else if (hm.tag == 33) {
            c1 = har_amp_33_rdcp1(wf), c2 = har_amp_33_rdcp2(wf), c3 = har_amp_33_rdcp3(wf);
        } else {
            c1 = har_amp_44_rdcp1(wf), c2 = har_amp_44_rdcp2(wf), c3 = har_amp_44_rdcp3(wf);
        }
        double rdcp1 = std::fabs(c1), rdcp2 = std::fabs(c2), rdcp3 = std::fabs(c3);
        rdcp3 = rdcp3 >= rdcp2 * rdcp2 / rdcp1 ? 0.5 * rdcp2 * rdcp2 / rdcp1 : rdcp3;
        rdcp3 = rdcp3 > rdcp2 ? 0.5 * rdcp2 : rdcp3;
        rdcp3 = (rdcp1 < rdcp2 && rdcp3 > rdcp1) ? rdcp1 : rdcp3;
        double deno = std::sqrt(rdcp1 / rdcp3) - rdcp1 / rdcp2;
        deno = deno <= 0.0 ? 1e-16 : deno;
        double a_lambda = rdcp1 * fdamp / deno;
        double sigma = std::sqrt(a_lambda / (rdcp2 * fdamp));
        rd[0] = a_lambda;
        rd[1] = 0.5 * sigma * std::log(rdcp1 / rdcp3);
        rd[2] = sigma;
    }
    double falloff = hm.f_amp_rd_falloff;
    double value = hm_lorentzian(falloff, hm, rd);
    amp.rd[0] = rd[0];
    amp.rd[1] = rd[1];
    amp.rd[2] = rd[2];
    amp.rd[3] = value;
    amp.rd[4] = -1.0 * rd_amp_dansatz(falloff, hm, rd) / value;
    for (double& a : amp.rd_aux) a = 0.0;
    if (hm.mixing) {
        double values[4] = {std::fabs(har_amp_32_rdaux1(wf)), std::fabs(har_amp_32_rdaux2(wf)),
                            hm_lorentzian(amp.f_rd_aux, hm, amp.rd), rd_amp_dansatz(amp.f_rd_aux, hm, amp.rd)};
        double freqs[3] = {amp.f_amp_match_im, 0.5 * (amp.f_amp_match_im + amp.f_rd_aux), amp.f_rd_aux};
        double m[4][4];
        for (int i = 0; i < 3; ++i) {
            double f = freqs[i];
            m[i][0] = 1.0, m[i][1] = f, m[i][2] = f * f, m[i][3] = f * f * f;
        }
        double f = amp.f_rd_aux;
        m[3][0] = 0.0, m[3][1] = 1.0, m[3][2] = 2.0 * f, m[3][3] = 3.0 * f * f;
        lu_solve<double, 4>(m, values);
        for (int j = 0; j < 4; ++j) amp.rd_aux[j] = values[j];
    }
}

// ---------------------------------------------------- the (3,2) ringdown

// phase.py:58-80 IMRPhenomXHM_RD_Phase_Ansatz / AnsatzInt: the rescaled
// (2,2) ringdown (coefficients alpha0, alpha2, alphaL) or, with mixing,
// the spheroidal one
template <class R>
static R ringdown_phase_ansatz(const R& ff, const Powers<R>& p, double frd, double fda, bool mixing,
                               const HmPhase& ph, const HmSpheroidal& s) {
    R den = o_fma(R(fda), R(fda), (ff - frd) * (ff - frd));
    if (!mixing) return o_fma(R(frd * frd * ph.alpha2), p.m_two, R(ph.alpha0)) + (ph.alpha_l * fda) / den;
    R d = o_fma(R(s.alpha2_s), p.m_two, R(s.alpha0_s));
    return o_fma(R(s.alpha4_s), p.m_four, d) + (s.alpha_l_s * fda) / den;
}
template <class R>
static R ringdown_phase_ansatz_int(const R& ff, const Powers<R>& p, double frd, double fda, bool mixing,
                                   const HmPhase& ph, const HmSpheroidal& s) {
    R invf = p.m_one;
    R arctan = o_atan((ff - frd) / fda);
    if (!mixing) {
        R phase_rd = o_fma(R(ph.alpha0), ff, -(frd * frd * ph.alpha2 * invf));
        return o_fma(R(ph.alpha_l), arctan, phase_rd);
    }
    R phase_rd = o_fma(R(s.alpha0_s), ff, R(s.phi0_s));
    phase_rd = o_fma(R(-s.alpha2_s), invf, phase_rd);
    phase_rd = o_fma(R(-(1.0 / 3.0 * s.alpha4_s)), p.m_three, phase_rd);
    return o_fma(R(s.alpha_l_s), arctan, phase_rd);
}

// reference.py:366-378 rotate: conj(a322) h22 + conj(a323) h32_spheroidal
template <class R>
static void rotate(const R& mf, const Powers<R>& p, const HmMode& c, const R& w_re, const R& w_im, R& out_re,
                   R& out_im) {
    const HmVars& hm = c.hm;
    R amplm = rd_amp_ansatz(mf, hm, c.amp);
    R philm = ringdown_phase_ansatz_int(mf, p, hm.f_ring, hm.f_damp, true, c.ph, c.sph);
    double m2r = hm.mix[2][0], m2i = hm.mix[2][1], m3r = hm.mix[3][0], m3i = hm.mix[3][1];
    R a_re = w_re * m2r + w_im * m2i;
    R a_im = w_im * m2r - w_re * m2i;
    R b_re = amplm * m3r;
    R b_im = -(m3i * amplm);
    R cs = o_cos(philm), sn = o_sin(philm);
    out_re = a_re + (b_re * cs - b_im * sn);
    out_im = a_im + (b_re * sn + b_im * cs);
}

// reference.py:381-388 SpheroidalToSpherical (122022)
template <class R>
static void spheroidal_to_spherical(const R& mf, const HmMode& c, R& re, R& im) {
    const HmMix22& m = c.mix22;
    Powers<R> p = powers(mf);
    R amp22 = rd_amp_22_ansatz_exact(mf, R(m.f_ring22), R(m.f_damp22), amp22_of(m, mf)) * m.amp_norm *
              p.m_seven_sixths;
    R phi22 = o_fma(R(m.inveta), phase_22(mf, p, ph22_of(m, mf)), m.timeshift * mf) + m.phiref22;
    rotate(mf, p, c, amp22 * o_cos(phi22), amp22 * o_sin(phi22), re, im);
}

// phase.py:155-193 GetSpheroidalCoefficients (with the (2,2) time shift
// and phase offset)
static void spheroidal_coefficients(const WfVars<double>& wf, HmMode& c) {
    const HmVars& hm = c.hm;
    const PhaseCoeffs<double>& ph22 = c.mix22.ph22;
    double fringlm = hm.f_ring, fdamplm = hm.f_damp;
    double freqs[4] = {wf.f_ring, o_fma(-1.5, fdamplm, fringlm), o_fma(-0.5, fdamplm, fringlm),
                       o_fma(0.5, fdamplm, fringlm)};
    double m[4][4], x[4] = {hpr_phase_32_p1(wf), hpr_phase_32_p2(wf), hpr_phase_32_p3(wf), hpr_phase_32_p4(wf)};
    for (int i = 0; i < 4; ++i) {
        double ff = freqs[i];
        double ffm1 = 1.0 / ff, ffm2 = ffm1 * ffm1;
        double lorentzian = fdamplm / (fdamplm * fdamplm + (ff - fringlm) * (ff - fringlm));
        m[i][0] = 1.0, m[i][1] = lorentzian, m[i][2] = ffm2, m[i][3] = ffm2 * ffm2;
    }
    lu_solve<double, 4>(m, x);
    HmSpheroidal& s = c.sph;
    s.alpha0_s = x[0], s.alpha_l_s = x[1], s.alpha2_s = x[2], s.alpha4_s = x[3], s.phi0_s = 0.0;
    // time shift
    double fref_rd = wf.f_ring + wf.f_damp;
    Powers<double> p_ref = powers(fref_rd);
    double tshift = hpr_phase_32_spheroidal_time_shift(wf);
    double timeshift = time_shift_22(ph22, wf);
    double dphi22_ref = o_fma(1.0 / wf.eta, dphase_22(fref_rd, p_ref, ph22), timeshift);
    s.alpha0_s = s.alpha0_s + dphi22_ref + tshift -
                 ringdown_phase_ansatz(fref_rd, p_ref, fringlm, fdamplm, true, c.ph, s);
    // phase shift
    double phaseshift = 0.0;
    fref_rd = wf.f_ring;
    p_ref = powers(fref_rd);
    double phase_22_mf_ref = phase_22(wf.mf_ref, powers(wf.mf_ref), ph22);
    double phase_22_fref = phase_22(fref_rd, p_ref, ph22);
    double phiref22 = o_fma(-1.0 / wf.eta, phase_22_mf_ref, -(timeshift * wf.mf_ref)) - phaseshift +
                      2.0 * wf.phi0 + M_PI_ / 4.0;
    double phi22_ref = o_fma(1.0 / wf.eta, phase_22_fref, timeshift * fref_rd) + phaseshift + phiref22;
    double phishift = hpr_phase_32_spheroidal_phase_shift(wf);
    s.phi0_s = phi22_ref - ringdown_phase_ansatz_int(fref_rd, p_ref, fringlm, fdamplm, true, c.ph, s) + phishift;
    c.mix22.timeshift = timeshift;
    c.mix22.phiref22 = phiref22;
}

// reference.py:275-334 intermediate amplitude (in u = (f - f_c) / f_h)
static void intermediate_amp_coefficients(HmMode& c, const WfVars<double>& wf) {
    HmAmp& amp = c.amp;
    const HmVars& hm = c.hm;
    const int* version = HM_INTER_VERSION[hm.mode_int];
    const int n_colloc = 6;
    double deltaf = (amp.f_amp_match_im - amp.f_amp_match_in) / (n_colloc - 1);
    int n = 0;
    for (int i = 0; i < n_colloc; ++i) n += version[i];
    const double seven_sixths = 7.0 / 6.0;
    double center = 0.5 * (amp.f_amp_match_im + amp.f_amp_match_in);
    double half = 0.5 * (amp.f_amp_match_im - amp.f_amp_match_in);
    amp.inter_center = center;
    amp.inter_half = half;
    double m[8][8], values[8];
    int row = 0;
    auto value_row = [&](double f) {
        double m76 = std::pow(f, -seven_sixths);
        double u = (f - center) / half;
        for (int j = 0; j < n; ++j) m[row][j] = m76 * ipow(u, j);
    };
    auto derivative_row = [&](double f) {
        double m76 = std::pow(f, -seven_sixths);
        double u = (f - center) / half;
        for (int j = 0; j < n; ++j)
            m[row][j] = m76 * (-seven_sixths / f * ipow(u, j) + (j ? (double)j * ipow(u, j - 1) / half : 0.0));
    };
    double f_in = amp.f_amp_match_in;
    Jet2 xin(f_in, 1.0, 0.0);
    Jet2 ins = inspiral_amp_ansatz(powers(xin), amp);
    double f_rd = amp.f_amp_match_im;
    double rd_value, rd_derivative;
    if (hm.mixing) {
        Jet2 xr(f_rd, 1.0, 0.0), re, im;
        spheroidal_to_spherical(xr, c, re, im);
        Jet2 h = o_hypot(re, im);
        rd_value = h.v;
        rd_derivative = h.d;
    } else {
        rd_value = rd_amp_ansatz(f_rd, hm, amp);
        rd_derivative = rd_amp_dansatz(f_rd, hm, amp.rd);
    }
    for (int i = 0; i < n_colloc; ++i) {
        int digit = version[i];
        double f = o_fma(deltaf, (double)i, f_in);
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
            double fit;
            switch (hm.tag * 10 + (i - 1)) {
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
            values[row++] = std::fabs(fit);
            c.colloc_amp_f[c.n_colloc_amp] = f;
            c.colloc_amp_v[c.n_colloc_amp++] = std::fabs(fit);
        }
    }
    if (n == 5) {
        double a[5][5], b[5];
        for (int i = 0; i < 5; ++i) {
            b[i] = values[i];
            for (int j = 0; j < 5; ++j) a[i][j] = m[i][j];
        }
        lu_solve<double, 5>(a, b);
        for (int j = 0; j < 5; ++j) amp.inter[j] = b[j];
    } else {
        lu_solve<double, 8>(m, values);
        for (int j = 0; j < 8; ++j) amp.inter[j] = values[j];
    }
    amp.n_inter = n;
    for (int j = n; j < 8; ++j) amp.inter[j] = 0.0;
}

// reference.py:337-345 IMRPhenomXHM_Intermediate_Amp_Ansatz (in u)
template <class R>
static R inter_amp_ansatz(const R& mf, const Powers<R>& p, const HmAmp& amp) {
    R u = (mf - amp.inter_center) / amp.inter_half;
    R result = R(amp.inter[amp.n_inter - 1]);
    for (int j = amp.n_inter - 2; j >= 0; --j) result = o_fma(result, u, R(amp.inter[j]));
    return result * p.m_seven_sixths;
}

// ------------------------------------------------------------------ phase

// inspiral.py:319-330 IMRPhenomXHM_Inspiral_Phase_AnsatzInt
template <class R>
static R inspiral_phase_ansatz_int(const R& mf, const Powers<R>& p, const HmPhase& ph) {
    const R freqs[13] = {p.m_five_thirds, p.m_four_thirds, p.m_one, p.m_two_thirds, p.m_one_third, R(1.0),
                         p.one_third, p.two_thirds, mf, p.four_thirds, p.five_thirds, p.two, p.seven_thirds};
    R log_mf = p.log;
    R philm = R(0.0);
    for (int i = 0; i < 13; ++i) philm = o_fma(o_fma(R(ph.phi_l[i]), log_mf, R(ph.phi[i])), freqs[i], philm);
    return philm;
}
// inspiral.py:337-352 IMRPhenomXHM_Inspiral_Phase_Ansatz
static const double HM_INSPIRAL_EXPONENTS[13] = {-5.0 / 3.0, -4.0 / 3.0, -1.0, -2.0 / 3.0, -1.0 / 3.0, 0.0,
                                                 1.0 / 3.0, 2.0 / 3.0, 1.0, 4.0 / 3.0, 5.0 / 3.0, 2.0, 7.0 / 3.0};
template <class R>
static R inspiral_phase_ansatz(const R& mf, const Powers<R>& p, const HmPhase& ph) {
    const R freqs[13] = {p.m_eight_thirds, p.m_seven_thirds, p.m_two, p.m_five_thirds, p.m_four_thirds, p.m_one,
                         p.m_two_thirds, p.m_one_third, R(1.0), p.one_third, p.two_thirds, mf, p.four_thirds};
    R log_mf = p.log;
    R dphilm = R(0.0);
    for (int i = 0; i < 13; ++i) {
        R coefficient = o_fma(R(ph.phi_l[i]), log_mf, R(ph.phi[i]));
        dphilm = o_fma(o_fma(coefficient, R(HM_INSPIRAL_EXPONENTS[i]), R(ph.phi_l[i])), freqs[i], dphilm);
    }
    return dphilm;
}

// phase.py:83-111 IMRPhenomXHM_Inter_Phase_Ansatz / AnsatzInt
static double intermediate_phase_ansatz(double ff, const Powers<double>& p, const HmVars& hm, const HmPhase& ph) {
    double fda = hm.f_damp, frd = hm.f_ring;
    double lorentzian = (ph.c_l * fda) / o_fma(fda, fda, (ff - frd) * (ff - frd));
    double dphase = o_fma(ph.c1, p.m_one, ph.c0);
    dphase = o_fma(ph.c2, p.m_two, dphase);
    dphase = o_fma(ph.c4, p.m_four, dphase);
    if (hm.tag == 32) dphase = o_fma(ph.c3, p.m_three, dphase);
    return dphase + lorentzian;
}
template <class R>
static R intermediate_phase_ansatz_int(const R& ff, const Powers<R>& p, const HmVars& hm, const HmPhase& ph) {
    double fda = hm.f_damp, frd = hm.f_ring;
    R phase_ir = o_fma(ph.c0, ff, ph.c1 * p.log);
    phase_ir = o_fma(-ph.c2, p.m_one, phase_ir);
    phase_ir = o_fma(-(1.0 / 3.0 * ph.c4), p.m_three, phase_ir);
    if (hm.tag == 32) phase_ir = o_fma(-(0.5 * ph.c3), p.m_two, phase_ir);
    return o_fma(ph.c_l, o_atan((ff - frd) / fda), phase_ir);
}

// phase.py:199-202 GetfcutInsp
static double get_fcut_insp(const WfVars<double>& wf, const HmVars& hm) {
    double eta_factor = o_fma(0.001, 0.25 / wf.eta - 1.0, 1.0);
    return eta_factor * (hm.emm * 1.0) * 0.5 * wf.f_meco;
}

// phase.py:205-219 IMRPhenomXHM_PN21AmpSign
static bool pn21_amp_sign(double ff, const WfVars<double>& wf) {
    double eta = wf.eta, chi1 = wf.chi1l, chi2 = wf.chi2l;
    double delta = std::sqrt(o_fma(-4.0, eta, 1.0));
    double sqrt5 = std::sqrt(5.0);
    double output = (-16.0 * delta * eta * ff * std::pow(M_PI_, 1.5)) / (3.0 * sqrt5);
    output = output + (4.0 * std::pow(2.0, 0.3333333333333333) * (chi1 - chi2 + delta * (chi1 + chi2)) * eta *
                       std::pow(ff, 1.3333333333333333) * std::pow(M_PI_, 1.8333333333333333)) / sqrt5;
    output = output + (2.0 * std::pow(2.0, 0.6666666666666666) * eta * (306.0 * delta - 360.0 * delta * eta) *
                       std::pow(ff, 1.6666666666666667) * std::pow(M_PI_, 2.1666666666666665)) / (189.0 * sqrt5);
    return output >= 0.0;
}

// phase.py:222-229 phase_fits: the intermediate collocation values and lambda
static void phase_fits(const WfVars<double>& wf, const HmVars& hm, double values[6], double& lambda_pn) {
    for (int i = 0; i < 6; ++i) {
        double v;
        switch (hm.tag * 10 + i) {
            case 210: v = hpi_phase_21_p1(wf); break;
            case 211: v = hpi_phase_21_p2(wf); break;
            case 212: v = hpi_phase_21_p3(wf); break;
            case 213: v = hpi_phase_21_p4(wf); break;
            case 214: v = hpi_phase_21_p5(wf); break;
            case 215: v = hpi_phase_21_p6(wf); break;
            case 330: v = hpi_phase_33_p1(wf); break;
            case 331: v = hpi_phase_33_p2(wf); break;
            case 332: v = hpi_phase_33_p3(wf); break;
            case 333: v = hpi_phase_33_p4(wf); break;
            case 334: v = hpi_phase_33_p5(wf); break;
            case 335: v = hpi_phase_33_p6(wf); break;
            case 320: v = hpi_phase_32_p1(wf); break;
            case 321: v = hpi_phase_32_p2(wf); break;
            case 322: v = hpi_phase_32_p3(wf); break;
            case 323: v = hpi_phase_32_p4(wf); break;
            case 324: v = hpi_phase_32_p5(wf); break;
            case 325: v = hpi_phase_32_p6(wf); break;
            case 440: v = hpi_phase_44_p1(wf); break;
            case 441: v = hpi_phase_44_p2(wf); break;
            case 442: v = hpi_phase_44_p3(wf); break;
            case 443: v = hpi_phase_44_p4(wf); break;
            case 444: v = hpi_phase_44_p5(wf); break;
            default: v = hpi_phase_44_p6(wf); break;
        }
        values[i] = v;
    }
    double fit = hm.tag == 21 ? hpn_phase_21_lambda(wf) : hm.tag == 33 ? hpn_phase_33_lambda(wf)
               : hm.tag == 32 ? hpn_phase_32_lambda(wf) : hpn_phase_44_lambda(wf);
    lambda_pn = wf.eta > 0.01 ? hpn_lambda_pn(wf.eta, hm.tag) : fit;
}

// reference.py:397-569 IMRPhenomXHM_GetPhaseCoefficients (122022; exact
// derivatives of the rotated (3,2) ringdown phase)
static void phase_coefficients(const WfVars<double>& wf, HmMode& c) {
    const HmVars& hm = c.hm;
    const PhaseCoeffs<double>& ph22 = c.mix22.ph22;
    HmPhase& ph = c.ph;
    const int emm = hm.emm, tag = hm.tag;
    const double eta = wf.eta, eta_m1 = 1.0 / eta;
    const double fring = hm.f_ring, fdamp = hm.f_damp;
    const double fcut = get_fcut_insp(wf, hm);
    const double sqrt3 = std::sqrt(3.0);
    double freqs[6], f_phase_match_im;
    if (tag == 32) {
        double f_end = o_fma(-0.5, wf.f_damp, wf.f_ring);
        freqs[0] = fcut;
        freqs[1] = o_fma(sqrt3, fcut - f_end, 2.0 * (fcut + f_end)) / 4.0;
        freqs[2] = o_fma(3.0, fcut, f_end) / 4.0;
        freqs[3] = (fcut + f_end) / 2.0;
        freqs[4] = f_end;
        freqs[5] = f_end;
        f_phase_match_im = (eta < 0.01 && wf.chi1l < 0.0) ? f_end * o_fma(-0.25, wf.chi1l, 1.2) : f_end;
    } else {
        freqs[0] = fcut;
        freqs[1] = o_fma(sqrt3, fcut - fring, 2.0 * (fcut + fring)) / 4.0;
        freqs[2] = o_fma(3.0, fcut, fring) / 4.0;
        freqs[3] = (fcut + fring) / 2.0;
        freqs[4] = o_fma(3.0, fring, fcut) / 4.0;
        freqs[5] = o_fma(7.0, fring, fcut) / 8.0;
        f_phase_match_im = fring - fdamp;
    }
    const double f_phase_match_in = hm.f_meco_lm;
    double values[6], lambda_pn;
    phase_fits(wf, hm, values, lambda_pn);
    for (double& v : values) v = v + hm.delta_t;
    const double fcut_rd = f_phase_match_im, fcut_insp = f_phase_match_in;

    // inspiral: the (2,2) coefficients rescaled
    const double zero = 0.0;
    const double non_log[13] = {ph22.phi0, zero, ph22.phi2, ph22.phi3, ph22.phi4, zero, ph22.phi6,
                                ph22.phi7, ph22.phi8, zero, zero, zero, zero};
    const double log_terms[13] = {zero, zero, zero, zero, zero, ph22.phi5l, ph22.phi6l, zero, ph22.phi8l, zero,
                                  zero, zero, zero};
    const double pseudo_pn[13] = {zero, zero, zero, zero, zero, zero, zero, zero, ph22.sigma1, ph22.sigma2,
                                  ph22.sigma3, ph22.sigma4, zero};
    const double m_over_2 = emm * 0.5, two_over_m = 1.0 / m_over_2;
    const double fact = ph22.phi_norm / eta;
    const double log_m_over_2 = std::log(m_over_2);
    for (int i = 0; i < 13; ++i) {
        double nl = non_log[i] * fact, lg = log_terms[i] * fact, ps = pseudo_pn[i] * fact;
        double scale = std::pow(m_over_2, (8 - i) / 3.0);
        ph.phi[i] = o_fma(-lg, log_m_over_2, nl + ps) * scale;
        ph.phi_l[i] = lg * scale;
    }
    ph.phi[8] = ph.phi[8] + lambda_pn;
    ph.alpha0 = ph.alpha2 = ph.alpha_l = 0.0;
    ph.c0 = ph.c1 = ph.c2 = ph.c3 = ph.c4 = ph.c_l = 0.0;

    if (tag == 21) {
        PhaseCoeffs<double> unconnected = ph22;
        unconnected.c1_int = unconnected.c2_int = unconnected.c1_mrd = unconnected.c2_mrd = 0.0;
        double insp[3];
        for (int k = 0; k < 3; ++k) {
            double ff = two_over_m * freqs[k];
            insp[k] = 1.0 / eta * dphase_22(ff, powers(ff), unconnected);
        }
        double diff12 = insp[0] - insp[1], diff23 = insp[1] - insp[2];
        bool high = wf.s_tot_r >= 0.8;
        double v1 = values[2] + diff23;
        values[1] = high ? v1 : values[1];
        values[0] = high ? v1 + diff12 : values[0];
    }
    auto row = [&](double ff, double* r) {
        double ffm1 = 1.0 / ff, ffm2 = ffm1 * ffm1;
        double lorentzian = fdamp / (fdamp * fdamp + (ff - fring) * (ff - fring));
        r[0] = 1.0, r[1] = lorentzian, r[2] = ffm1, r[3] = ffm2, r[4] = ffm2 * ffm2, r[5] = ffm1 * ffm2;
    };
    double x[6], phi0_rd, dphi0_rd;
    if (tag != 32) {
        bool cond_a = eta < HM_ETA_EMR;
        if (emm == hm.ell) cond_a = cond_a || wf.s_tot_r >= 0.8;
        if (tag == 33) cond_a = cond_a || wf.s_tot_r < 0.0;
        bool cond_b = tag == 21 && wf.s_tot_r >= 0.8;
        int idx2 = cond_a ? 3 : 2, idx3 = cond_a ? 4 : (cond_b ? 4 : 3);
        const int chosen[5] = {0, 1, idx2, idx3, 5};
        double wlm = emm == hm.ell ? 2.0 : emm / 3.0;
        ph.alpha2 = 1.0 / (fring * fring) * wlm * c.alpha2_fit;
        ph.alpha_l = eta_m1 * c.alpha_l_fit;
        Powers<double> p_rd = powers(fcut_rd);
        phi0_rd = ringdown_phase_ansatz_int(fcut_rd, p_rd, fring, fdamp, false, ph, c.sph);
        dphi0_rd = ringdown_phase_ansatz(fcut_rd, p_rd, fring, fdamp, false, ph, c.sph);
        double a[5][5], b[5];
        for (int i = 0; i < 5; ++i) {
            double r[6];
            row(freqs[chosen[i]], r);
            for (int j = 0; j < 5; ++j) a[i][j] = r[j];
            b[i] = values[chosen[i]];
            c.colloc_phase_f[i] = freqs[chosen[i]];
            c.colloc_phase_v[i] = values[chosen[i]];
        }
        c.n_colloc_phase = 5;
        lu_solve<double, 5>(a, b);
        for (int j = 0; j < 5; ++j) x[j] = b[j];
        x[5] = 0.0;
    } else {
        // the rotated ringdown phase and its first two derivatives at fcut_rd
        Jet2 xr(fcut_rd, 1.0, 0.0), re, im;
        spheroidal_to_spherical(xr, c, re, im);
        Jet2 rotated = o_atan2(im, re);
        phi0_rd = rotated.v;
        dphi0_rd = rotated.d;
        double d2phi0_rd = rotated.dd;
        bool comparable = eta > HM_ETA_EMR;
        freqs[4] = comparable ? fcut_rd : freqs[4];
        values[4] = comparable ? dphi0_rd : values[4];
        freqs[5] = fcut_rd;
        values[5] = d2phi0_rd;
        double g_c = 0.5 * (1.0 / fcut_insp + 1.0 / fcut_rd);
        double g_h = 0.5 * (1.0 / fcut_insp - 1.0 / fcut_rd);
        auto lorentzian = [&](double ff) { return fdamp / (fdamp * fdamp + (ff - fring) * (ff - fring)); };
        double a[6][6], b[6];
        for (int i = 0; i < 5; ++i) {
            double v = (1.0 / freqs[i] - g_c) / g_h;
            a[i][0] = 1.0, a[i][1] = lorentzian(freqs[i]), a[i][2] = v, a[i][3] = v * v, a[i][4] = v * v * v,
            a[i][5] = v * v * v * v;
            b[i] = values[i];
            c.colloc_phase_f[i] = freqs[i];
            c.colloc_phase_v[i] = values[i];
        }
        c.n_colloc_phase = 5;
        double ff = freqs[5];
        double v = (1.0 / ff - g_c) / g_h;
        double dv = -1.0 / (ff * ff * g_h);
        double den = fdamp * fdamp + (ff - fring) * (ff - fring);
        a[5][0] = 0.0, a[5][1] = -2.0 * fdamp * (ff - fring) / (den * den), a[5][2] = dv, a[5][3] = 2.0 * v * dv,
        a[5][4] = 3.0 * v * v * dv, a[5][5] = 4.0 * v * v * v * dv;
        b[5] = values[5];
        lu_solve<double, 6>(a, b);
        const double quartic[5] = {b[0], b[2], b[3], b[4], b[5]};
        static const double comb[5][5] = {{1, 0, 0, 0, 0}, {1, 1, 0, 0, 0}, {1, 2, 1, 0, 0}, {1, 3, 3, 1, 0},
                                          {1, 4, 6, 4, 1}};
        double pc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < 5; ++k) {
            double scale = quartic[k] / ipow(g_h, k);
            for (int j = 0; j <= k; ++j) pc[j] = pc[j] + scale * comb[k][j] * ipow(-g_c, k - j);
        }
        x[0] = pc[0], x[1] = b[1], x[2] = pc[1], x[3] = pc[2], x[4] = pc[4], x[5] = pc[3];
    }
    ph.c0 = x[0], ph.c_l = x[1], ph.c1 = x[2], ph.c2 = x[3], ph.c4 = x[4];
    if (tag == 32) {
        ph.c3 = x[5];
        Powers<double> p_rd = powers(fcut_rd);
        double glued = ph.c0 + dphi0_rd - intermediate_phase_ansatz(fcut_rd, p_rd, hm, ph);
        ph.c0 = eta < HM_ETA_EMR ? glued : ph.c0;
    }
    Powers<double> p_in = powers(fcut_insp);
    ph.c1_insp = intermediate_phase_ansatz(fcut_insp, p_in, hm, ph) - inspiral_phase_ansatz(fcut_insp, p_in, ph);
    ph.c_insp = o_fma(-ph.c1_insp, fcut_insp, intermediate_phase_ansatz_int(fcut_insp, p_in, hm, ph)) -
                inspiral_phase_ansatz_int(fcut_insp, p_in, ph);
    Powers<double> p_rd = powers(fcut_rd);
    ph.c1_rd = intermediate_phase_ansatz(fcut_rd, p_rd, hm, ph) - dphi0_rd;
    ph.c_rd = o_fma(-ph.c1_rd, fcut_rd, intermediate_phase_ansatz_int(fcut_rd, p_rd, hm, ph)) - phi0_rd;

    // align the mode with the (2,2) at low frequency
    double falign = eta > HM_ETA_EMR ? 0.6 * m_over_2 * wf.f_meco : m_over_2 * wf.f_meco;
    Powers<double> p_falign = powers(falign);
    double x22 = two_over_m * falign;
    double phase_22_falign = phase_22(x22, powers(x22), ph22);
    double phi22 = phase_22_falign / eta + c.mix22.phiref22;
    double phi_lm = o_fma(ph.c1_insp, falign, inspiral_phase_ansatz_int(falign, p_falign, ph)) + ph.c_insp;
    double deltaphi = o_fma(-(3.0 * (M_PI_ / 4.0)), 1.0 - m_over_2,
                            o_fma(m_over_2, phi22, c.mix22.timeshift * falign)) - phi_lm;
    if (tag == 21) deltaphi = deltaphi + (pn21_amp_sign(0.008, wf) ? M_PI_ : 0.0);
    ph.deltaphi = deltaphi;
    ph.f_phase_match_in = f_phase_match_in;
    ph.f_phase_match_im = f_phase_match_im;
}

// reference.py:580-608 mode_setup
static void mode_setup(const WfVars<double>& wf, const AmpCoeffs<double>& amp22, const PhaseCoeffs<double>& ph22,
                       int k, HmMode& c) {
    c.hm = mode_struct(wf, k);
    HmVars& hm = c.hm;
    HmMix22& m = c.mix22;
    m.amp22 = amp22;
    m.ph22 = ph22;
    m.amp22j = lift(amp22);
    m.ph22j = lift(ph22);
    m.f_ring22 = wf.f_ring;
    m.f_damp22 = wf.f_damp;
    m.inveta = 1.0 / wf.eta;
    m.amp_norm = wf.amp_norm;
    double timeshift = time_shift_22(ph22, wf);
    double mf_ref = wf.mf_ref;
    m.timeshift = timeshift;
    m.phiref22 = o_fma(-1.0 / wf.eta, phase_22(mf_ref, powers(mf_ref), ph22), -(timeshift * wf.mf_ref)) +
                 2.0 * wf.phi0 + M_PI_ / 4.0;
    c.alpha2_fit = hpr_phase_22_alpha2(wf);
    c.alpha_l_fit = hpr_phase_22_alpha_l(wf);
    c.n_colloc_phase = c.n_colloc_amp = 0;
    c.sph = HmSpheroidal{0.0, 0.0, 0.0, 0.0, 0.0};
    std::memset(&c.ph, 0, sizeof(c.ph));
    if (hm.mixing) spheroidal_coefficients(wf, c);  // (sets the (3,2)'s time shift and phase offset)
    // amplitude_coefficients (reference.py:348-360)
    HmAmp& amp = c.amp;
    std::memset(&amp, 0, sizeof(amp));
    amp.amp_norm = wf.amp_norm;
    amp.pn_dominant = wf.amp_norm * std::pow(2.0 / hm.emm, -7.0 / 6.0);
    if (hm.mixing) amp.f_rd_aux = hm.f_ring - hm.f_damp;
    amp.f_amp_match_in = amplitude_fcut_insp(hm, wf);
    amp.f_amp_match_im = amplitude_fcut_rd(hm, wf);
    pn_amplitude_coefficients(wf, hm, amp);
    inspiral_amp_coefficients(amp, hm, wf);
    rd_amp_coefficients(amp, hm, wf);
    intermediate_amp_coefficients(c, wf);
    phase_coefficients(wf, c);
    c.amp0 = (hm.ell % 2 ? -1.0 : 1.0) * wf.amp0;
}

// reference.py:620-643 mode_amplitude_phase: amplitude (without (-1)^l amp0)
// and phase at mf; (h22_re, h22_im) the (2,2) mode recycled by the (3,2)
static void mode_amplitude_phase(double mf, const HmMode& c, double h22_re, double h22_im, double& a, double& phi) {
    const HmVars& hm = c.hm;
    const HmAmp& amp = c.amp;
    const HmPhase& ph = c.ph;
    Powers<double> p = powers(mf);
    double rot_re = 0.0, rot_im = 0.0;
    if (hm.mixing) {
        double amp0 = std::fabs(c.amp0);
        rotate(mf, p, c, h22_re / amp0, h22_im / amp0, rot_re, rot_im);
    }
    double ins = inspiral_amp_ansatz(p, amp);
    double inter = inter_amp_ansatz(mf, p, amp);
    double rd = hm.mixing ? o_hypot(rot_re, rot_im) : rd_amp_ansatz(mf, hm, amp);
    a = mf < amp.f_amp_match_in ? ins : (mf >= amp.f_amp_match_im ? rd : inter);
    a = a < 0.0 ? HM_FALSE_ZERO : a;
    double phi_ins = o_fma(ph.c1_insp, mf, inspiral_phase_ansatz_int(mf, p, ph)) + ph.c_insp;
    double phi_mrd = hm.mixing ? o_atan2(rot_im, rot_re)
                               : ringdown_phase_ansatz_int(mf, p, hm.f_ring, hm.f_damp, false, ph, c.sph);
    double phi_rd = o_fma(ph.c1_rd, mf, phi_mrd) + ph.c_rd;
    double phi_int = intermediate_phase_ansatz_int(mf, p, hm, ph);
    phi = mf < ph.f_phase_match_in ? phi_ins : (mf >= ph.f_phase_match_im ? phi_rd : phi_int);
    phi = phi + ph.deltaphi;
}

// ------------------------------------------------------------- harmonics

static double factorial_(int n) {
    double f = 1.0;
    for (int k = 2; k <= n; ++k) f *= k;
    return f;
}
// _harmonics.py:17-30 Wigner's d^l_{m' m}(beta) (XLALWignerdMatrix's convention)
static double wigner_small_d(int ell, int mp, int m, double beta) {
    double cos_half = std::cos(beta / 2), sin_half = std::sin(beta / 2);
    double total = 0.0;
    double prefactor =
        std::sqrt(factorial_(ell + mp) * factorial_(ell - mp) * factorial_(ell + m) * factorial_(ell - m));
    int k0 = std::max(0, m - mp), k1 = std::min(ell + m, ell - mp);
    for (int k = k0; k <= k1; ++k) {
        double sign = ((k - m + mp) % 2 == 0) ? 1.0 : -1.0;
        double coefficient = sign * prefactor / (factorial_(ell + m - k) * factorial_(k) * factorial_(ell - k - mp) *
                                                 factorial_(k - m + mp));
        total = total + coefficient * ipow(cos_half, 2 * ell - 2 * k + m - mp) * ipow(sin_half, 2 * k - m + mp);
    }
    return total;
}
// _harmonics.py:33-38 {}_{-2}Y_{lm}(theta, phi)
static std::complex<double> swsh(double theta, double phi, int ell, int m) {
    const int s = -2;
    double norm = std::sqrt((2 * ell + 1) / (4 * M_PI_));
    double d = norm * wigner_small_d(ell, m, -s, theta);
    double arg = m * phi;
    return std::complex<double>(d * std::cos(arg), d * std::sin(arg));
}
// reference.py:646-653 mode_factors: (F+, Fx) of h_{l,-m}
static void mode_factors(double inclination, int ell, int emm, std::complex<double>& fp, std::complex<double>& fc) {
    double sign = ell % 2 ? -1.0 : 1.0;
    std::complex<double> ym = swsh(inclination, M_PI_ / 2.0, ell, -emm);
    std::complex<double> ystar = std::conj(swsh(inclination, M_PI_ / 2.0, ell, emm));
    fp = 0.5 * (ym + sign * ystar);
    fc = std::complex<double>(0.0, 0.5) * (ym - sign * ystar);
}

// ------------------------------------------------------------- evaluation

struct XhmSetup {
    WfVars<double> wf;
    AmpCoeffs<double> amp22;
    PhaseCoeffs<double> ph22;
    double linb, inveta, amp_prefactor, phif_ref, mf_cut;
    HmMode modes[4];
};

// reference.py:566-577 setup (and xas.py:20-36 for the (2,2))
static void xhm_setup(const double* params, double f_min, double amp_scale, XhmSetup& s) {
    double pr[8];
    for (int i = 0; i < 8; ++i) pr[i] = params[i];
    s.wf = waveform_variables<double>(pr, f_min, amp_scale);
    const WfVars<double>& wf = s.wf;
    s.amp22 = amp_coefficients(wf);
    s.ph22 = phase_coefficients(wf);
    s.linb = time_shift_22(s.ph22, wf);
    s.inveta = 1.0 / wf.eta;
    s.amp_prefactor = wf.amp0 * wf.amp_norm;
    double mf_ref = wf.mf_ref;
    double phi = s.inveta * phase_22(mf_ref, powers(mf_ref), s.ph22);
    s.phif_ref = -(o_fma(s.linb, mf_ref, phi) + 0.0) + 2.0 * wf.phi0 + M_PI_ / 4.0;
    s.mf_cut = wf.f_max_prime * wf.m_sec;
    for (int k = 0; k < 4; ++k) mode_setup(wf, s.amp22, s.ph22, k, s.modes[k]);
}

// xas.py:39-47 h22(Mf), zero above the cutoff
static void xhm_h22(const XhmSetup& s, double mf, double& re, double& im) {
    const PhaseCoeffs<double>& ph = s.ph22;
    const AmpCoeffs<double>& amp = s.amp22;
    Powers<double> p = powers(mf);
    double pi_ = ins_phase_22_ansatz_int_exact(mf, p, ph);
    double pr_ = o_fma(ph.c2_mrd, mf, rd_phase_22_ansatz_int_exact(mf, p, ph) + ph.c1_mrd);
    double pm_ = o_fma(ph.c2_int, mf, itm_phase_22_ansatz_int_exact(mf, p, ph) + ph.c1_int);
    double phi = (mf < ph.f_phase_match_in ? pi_ : (mf > ph.f_phase_match_im ? pr_ : pm_)) * s.inveta;
    phi = phi + (o_fma(s.linb, mf, 0.0) + s.phif_ref);
    double ai = ins_amp_22_ansatz(mf, p, amp);
    double ar = rd_amp_22_ansatz(mf, amp);
    double am = itm_amp_22_ansatz(mf, p, amp);
    double a = s.amp_prefactor * p.m_seven_sixths * (mf < amp.f_amp_match_in ? ai : (mf > amp.f_amp_rd_min ? ar : am));
    re = mf <= s.mf_cut ? a * std::cos(phi) : 0.0;
    im = mf <= s.mf_cut ? a * std::sin(phi) : 0.0;
}

// reference.py:656-677 evaluate: the modes h_{l,-m} (2,2), (2,1), (3,3),
// (3,2), (4,4) and the polarizations; modes may be null
static void xhm_evaluate(const XhmSetup& s, double inclination, const double* freqs, long n, double* hp, double* hc,
                         double* modes) {
    const bool nan = any_error(s.wf);
    std::complex<double> fp[5], fc[5];
    mode_factors(inclination, 2, 2, fp[0], fc[0]);
    for (int k = 0; k < 4; ++k) mode_factors(inclination, HM_ELL[k], HM_EMM[k], fp[k + 1], fc[k + 1]);
    for (long i = 0; i < n; ++i) {
        double mf = s.wf.m_sec * freqs[i];
        double re22, im22;
        xhm_h22(s, mf, re22, im22);
        std::complex<double> h[5];
        h[0] = std::complex<double>(re22, im22);
        for (int k = 0; k < 4; ++k) {
            const HmMode& c = s.modes[k];
            double a, phi;
            mode_amplitude_phase(mf, c, re22, im22, a, phi);
            double r = c.amp0 * a;
            h[k + 1] = c.hm.amp_zero ? std::complex<double>(0.0, 0.0)
                                     : std::complex<double>(r * std::cos(phi), r * std::sin(phi));
        }
        std::complex<double> p(0.0, 0.0), x(0.0, 0.0);
        for (int k = 0; k < 5; ++k) {
            p = p + fp[k] * h[k];
            x = x + fc[k] * h[k];
        }
        if (nan) p = x = std::complex<double>(NAN, NAN);
        hp[2 * i] = p.real(), hp[2 * i + 1] = p.imag();
        hc[2 * i] = x.real(), hc[2 * i + 1] = x.imag();
        if (modes) {
            for (int k = 0; k < 5; ++k) {
                modes[(k * n + i) * 2] = nan ? NAN : h[k].real();
                modes[(k * n + i) * 2 + 1] = nan ? NAN : h[k].imag();
            }
        }
    }
}

// ------------------------------------------------------------- op count

// The algorithmic work of one bin, tallied on the Counted scalar: each
// higher mode's amplitude and phase in the region the bin falls in (the
// branches the device takes, where the reference's jnp.where evaluates
// every region), the (2,2) by IMRPhenomXAS's table form (the device's
// (2,2) is that record), each mode's sine and cosine, and the sum with the
// factors. The powers of M f come from the frequency's (counted once per
// frequency, as the device shares them over binaries) through the
// binary's M: M f, (M f)^(1/3), 1/(M f), log(M f) and (M f)^(-7/6), five
// operations per bin; the other powers the formula reads are given.
static inline Counted o_hypot(const Counted& x, const Counted& y) {
    g_tally.mul += 2;
    g_tally.add += 1;
    g_tally.sqrt_ += 1;
    return Counted(o_hypot(x.v, y.v));
}
static inline Counted o_atan2(const Counted& y, const Counted& x) {
    g_tally.atan_ += 1;
    g_tally.div += 1;
    return Counted(std::atan2(y.v, x.v));
}

template <template <class> class S>
static S<Counted> counted(const S<double>& s) {
    static_assert(sizeof(S<Counted>) == sizeof(S<double>), "a struct of doubles");
    S<Counted> out;
    std::memcpy(static_cast<void*>(&out), static_cast<const void*>(&s), sizeof(out));
    return out;
}

template <class R>
static void mode_point_counted(const R& mf, const Powers<R>& p, const HmMode& c, const R& h22_re, const R& h22_im,
                               R& a, R& phi) {
    const HmVars& hm = c.hm;
    const HmAmp& amp = c.amp;
    const HmPhase& ph = c.ph;
    const bool amp_ins = mf < amp.f_amp_match_in, amp_rd = !amp_ins && mf >= amp.f_amp_match_im;
    const bool ph_ins = mf < ph.f_phase_match_in, ph_rd = !ph_ins && mf >= ph.f_phase_match_im;
    R rot_re = R(0.0), rot_im = R(0.0);
    if (hm.mixing && (amp_rd || ph_rd)) {
        double amp0 = std::fabs(c.amp0);
        rotate(mf, p, c, h22_re / amp0, h22_im / amp0, rot_re, rot_im);
    }
    if (amp_ins)
        a = inspiral_amp_ansatz(p, amp);
    else if (amp_rd)
        a = hm.mixing ? o_hypot(rot_re, rot_im) : rd_amp_ansatz(mf, hm, amp);
    else
        a = inter_amp_ansatz(mf, p, amp);
    if (ph_ins)
        phi = o_fma(ph.c1_insp, mf, inspiral_phase_ansatz_int(mf, p, ph)) + ph.c_insp;
    else if (ph_rd)
        phi = o_fma(ph.c1_rd, mf,
                    hm.mixing ? o_atan2(rot_im, rot_re)
                              : ringdown_phase_ansatz_int(mf, p, hm.f_ring, hm.f_damp, false, ph, c.sph)) +
              ph.c_rd;
    else
        phi = intermediate_phase_ansatz_int(mf, p, hm, ph);
    phi = phi + ph.deltaphi;
}

extern "C" {

// The polarizations (and, with modes non-null, the five modes h_{l,-m}
// as (5, nb, n) complex) of nb binaries (threads: OpenMP threads, 0 for
// the default), params (nb, 8) = (m1, m2, chi1z,
// chi2z, d_L, inclination, phi_ref, f_ref), on the grid freqs (n); f_min
// = freqs[0]. hp, hc: (nb, n) complex as interleaved doubles.
void oracle_xhm_polarizations(const double* params, long nb, const double* freqs, long n, double amp_scale,
                              double* hp, double* hc, double* modes, int threads) {
    const int nt = threads > 0 ? threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (long b = 0; b < nb; ++b) {
        XhmSetup s;
        xhm_setup(params + 8 * b, freqs[0], amp_scale, s);
        std::vector<double> m(modes ? 10 * n : 0);
        xhm_evaluate(s, params[8 * b + 5], freqs, n, hp + 2 * n * b, hc + 2 * n * b, modes ? m.data() : nullptr);
        if (modes)
            for (int k = 0; k < 5; ++k)
                std::memcpy(modes + 2 * ((k * nb + b) * n), m.data() + 2 * k * n, 2 * n * sizeof(double));
    }
}

// The internals of one binary's mode k (0..3 for (2,1), (3,3), (3,2), (4,4)):
// out[0..17] = f_ring, f_damp, f_amp_match_in, f_amp_match_im,
// f_phase_match_in, f_phase_match_im, c1_insp, c_insp, c1_rd, c_rd,
// deltaphi, c0, c1, c2, c3, c4, c_l, amp0
void oracle_xhm_mode_internals(const double* params, double f_min, int k, double* out) {
    XhmSetup s;
    xhm_setup(params, f_min, 1.0, s);
    const HmMode& c = s.modes[k];
    const double v[18] = {c.hm.f_ring, c.hm.f_damp, c.amp.f_amp_match_in, c.amp.f_amp_match_im,
                          c.ph.f_phase_match_in, c.ph.f_phase_match_im, c.ph.c1_insp, c.ph.c_insp, c.ph.c1_rd,
                          c.ph.c_rd, c.ph.deltaphi, c.ph.c0, c.ph.c1, c.ph.c2, c.ph.c3, c.ph.c4, c.ph.c_l, c.amp0};
    std::memcpy(out, v, sizeof(v));
}

// Amplitude and phase of mode k at the mass-scaled frequencies mf (n), and
// (for the (3,2)) the rotated ringdown phase with its first and second
// derivatives at each mf in rot (3 n) when rot is non-null.
void oracle_xhm_mode_amplitude_phase(const double* params, double f_min, int k, const double* mf, long n,
                                     double* amp, double* phase, double* rot) {
    XhmSetup s;
    xhm_setup(params, f_min, 1.0, s);
    const HmMode& c = s.modes[k];
    for (long i = 0; i < n; ++i) {
        double re22, im22;
        xhm_h22(s, mf[i], re22, im22);
        mode_amplitude_phase(mf[i], c, re22, im22, amp[i], phase[i]);
        if (rot && c.hm.mixing) {
            Jet2 x(mf[i], 1.0, 0.0), re, im;
            spheroidal_to_spherical(x, c, re, im);
            Jet2 r = o_atan2(im, re);
            rot[3 * i] = r.v, rot[3 * i + 1] = r.d, rot[3 * i + 2] = r.dd;
        }
    }
}

// The collocation conditions of mode k: out[0] = number of phase-derivative
// points, out[1..6] their M f, out[7..12] the values; out[13] = number of the
// intermediate amplitude's fitted points, out[14..19] their M f, out[20..25]
// the fitted amplitudes (without amp0).
void oracle_xhm_collocation(const double* params, double f_min, int k, double* out) {
    XhmSetup s;
    xhm_setup(params, f_min, 1.0, s);
    const HmMode& c = s.modes[k];
    for (int i = 0; i < 26; ++i) out[i] = 0.0;
    out[0] = c.n_colloc_phase;
    for (int i = 0; i < c.n_colloc_phase; ++i) out[1 + i] = c.colloc_phase_f[i], out[7 + i] = c.colloc_phase_v[i];
    out[13] = c.n_colloc_amp;
    for (int i = 0; i < c.n_colloc_amp; ++i) out[14 + i] = c.colloc_amp_f[i], out[20 + i] = c.colloc_amp_v[i];
}

// The algorithmic work of the polarizations of nb binaries on a shared grid
// (mode_point_counted): out[0..10] = {add, mul, div, sqrt, exp, log,
// sincos, atan, cmp, cbrt, pow} of the per-frequency powers, out[11..21]
// those of the bins.
void oracle_xhm_count(const double* params, long nb, const double* freqs, long n, long long* out) {
    std::memset(&g_tally, 0, sizeof g_tally);
    std::vector<FreqPowers<Counted>> fps(n);
    for (long i = 0; i < n; ++i) fps[i] = frequency_powers(Counted(freqs[i]));
    OpTally freq = g_tally;
    std::memset(&g_tally, 0, sizeof g_tally);
    for (long b = 0; b < nb; ++b) {
        XhmSetup s;
        xhm_setup(params + 8 * b, freqs[0], 1e21, s);
        double pb[8];
        for (int k = 0; k < 8; ++k) pb[k] = params[8 * b + k];
        std::vector<double> tab(TABLE_SIZE);
        table_of<double>(pb, freqs[0], 1e21, tab.data());
        std::complex<double> fp[5], fc[5];
        mode_factors(params[8 * b + 5], 2, 2, fp[0], fc[0]);
        for (int k = 0; k < 4; ++k) mode_factors(params[8 * b + 5], HM_ELL[k], HM_EMM[k], fp[k + 1], fc[k + 1]);
        for (long i = 0; i < n; ++i) {
            // the (2,2): IMRPhenomXAS's table form with kp = 1, kc = 0 (its polarization factors' products
            // counted below with the other modes')
            Counted hp22[2], hc22[2];
            OpTally before = g_tally;
            evaluate_point(tab.data(), fps[i], hp22, hc22);
            // evaluate_point's products by kp and kc (four multiplications) are not the mode's
            g_tally.mul = before.mul + (g_tally.mul - before.mul) - 4;
            // the powers of M f from the frequency's and the binary's M (five operations)
            const double mf_v = s.wf.m_sec * freqs[i];
            Counted mf = Counted(s.wf.m_sec) * Counted(freqs[i]);
            g_tally.mul += 3;
            g_tally.add += 1;
            g_tally.sqrt_ += 1;
            const Powers<double> pd = powers(mf_v);
            const Powers<Counted> p = counted(pd);
            Counted h[5][2];
            h[0][0] = hp22[0];
            h[0][1] = hp22[1];
            for (int k = 0; k < 4; ++k) {
                const HmMode& c = s.modes[k];
                Counted a, phi;
                mode_point_counted(mf, p, c, h[0][0], h[0][1], a, phi);
                Counted r = Counted(c.amp0) * a;
                h[k + 1][0] = r * o_cos(phi);
                h[k + 1][1] = r * o_sin(phi);
            }
            // (h+, hx) = sum_k (F+, Fx)_k h_k: four complex multiply-adds per mode
            Counted pr, pi, cr, ci;
            for (int k = 0; k < 5; ++k) {
                pr = pr + (Counted(fp[k].real()) * h[k][0] - Counted(fp[k].imag()) * h[k][1]);
                pi = pi + (Counted(fp[k].real()) * h[k][1] + Counted(fp[k].imag()) * h[k][0]);
                cr = cr + (Counted(fc[k].real()) * h[k][0] - Counted(fc[k].imag()) * h[k][1]);
                ci = ci + (Counted(fc[k].real()) * h[k][1] + Counted(fc[k].imag()) * h[k][0]);
            }
        }
    }
    OpTally bins = g_tally;
    long long* f = reinterpret_cast<long long*>(&freq);
    long long* g = reinterpret_cast<long long*>(&bins);
    for (int k = 0; k < 11; ++k) {
        out[k] = f[k];
        out[11 + k] = g[k];
    }
}

// {}_{-2}Y_{lm}(theta, phi) as (re, im)
void oracle_swsh(double theta, double phi, int ell, int m, double* out) {
    std::complex<double> y = swsh(theta, phi, ell, m);
    out[0] = y.real(), out[1] = y.imag();
}

}  // extern "C"
