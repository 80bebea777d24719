// IMRPhenomXAS oracle: a CPU FP64 restatement of the reference's table +
// formula model, for checking the HIP product. TEST INFRASTRUCTURE ONLY:
// only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
// it (as liboracle_xas.so through oracle/oracle.py); the product never does.
//
// It follows, operation by operation and with contraction off
// (-ffp-contract=off, fma(a, b, c) as the reference's unfused a * b + c):
//   src/waveformx/models/xpnr/_math.py:36-51,125-137,190-246,522-526
//   src/waveformx/models/xpnr/internals.py:37-170   (waveform variables)
//   src/waveformx/models/xpnr/amp22.py:25-129        (amplitude coefficients)
//   src/waveformx/models/xpnr/phase22.py:39-326      (phase coefficients, solves,
//                                                     connections, time shift)
//   src/waveformx/models/xas/core.py:183-507         (QNM in 1 - a, table, formula)
//   src/waveformx/models/xas/model.py:35-60          (reference_mf, f_min)
//   src/waveformx/models/xpnr/xas.py:20-61           (LAL-form cross-check)
// and the fits transcribed by oracle/gen_fits.py into xas_fits.h.
//
// Everything is generic over the scalar R: double (the oracle), Dual (eight
// forward-mode directions, for the analytic FP64 gradient the adjoint is
// checked against) and Counted (an op counter for the algorithmic work per
// bin). The 86-entry table keeps the reference's LAYOUT (core.py:68-101).
//
// Parity status of each piece (what pins it) is stated in the tests that
// exercise it; the transcribed fit literals and XLA's own transcendental
// rounding are "parity unpinned" against the reference's JAX outputs.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>

#include "../waveformx_amd/csrc/wfx_libm.h"

#ifdef _OPENMP
#include <omp.h>
#endif

// ---------------------------------------------------------------- scalars

static const double M_PI_ = 3.141592653589793;
static const double EGAMMA_ = 0.5772156649015329;
static const double LN2_ = 0.6931471805599453;
// LAL's constants (xpnr/_math.py:46-51)
static const double PC_SI_ = 3.085677581491367278913937957796471611e16;
static const double MPC_SI_ = 1.0e6 * PC_SI_;
static const double MRSUN_SI_ = 1.476625038050124729627979840144936351e3;
static const double MTSUN_SI_ = 4.925490947641266978197229498498379006e-6;

static const int ND = 8;  // forward-mode directions

struct Dual {
    double v;
    double d[ND];
    Dual() : v(0.0) { std::memset(d, 0, sizeof d); }
    Dual(double x) : v(x) { std::memset(d, 0, sizeof d); }
};
static inline Dual mk(double v, const Dual& a, double da) {
    Dual r(v);
    for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] * da;
    return r;
}
static inline Dual operator+(const Dual& a, const Dual& b) {
    Dual r(a.v + b.v);
    for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
static inline Dual operator-(const Dual& a, const Dual& b) {
    Dual r(a.v - b.v);
    for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
static inline Dual operator-(const Dual& a) {
    Dual r(-a.v);
    for (int i = 0; i < ND; ++i) r.d[i] = -a.d[i];
    return r;
}
static inline Dual operator*(const Dual& a, const Dual& b) {
    Dual r(a.v * b.v);
    for (int i = 0; i < ND; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
static inline Dual operator/(const Dual& a, const Dual& b) {
    double q = a.v / b.v;
    Dual r(q);
    for (int i = 0; i < ND; ++i) r.d[i] = (a.d[i] - q * b.d[i]) / b.v;
    return r;
}
static inline Dual operator+(const Dual& a, double b) { return a + Dual(b); }
static inline Dual operator+(double a, const Dual& b) { return Dual(a) + b; }
static inline Dual operator-(const Dual& a, double b) { return a - Dual(b); }
static inline Dual operator-(double a, const Dual& b) { return Dual(a) - b; }
static inline Dual operator*(const Dual& a, double b) { return a * Dual(b); }
static inline Dual operator*(double a, const Dual& b) { return Dual(a) * b; }
static inline Dual operator/(const Dual& a, double b) { return a / Dual(b); }
static inline Dual operator/(double a, const Dual& b) { return Dual(a) / b; }
#define DCMP(op)                                                                       \
    static inline bool operator op(const Dual& a, const Dual& b) { return a.v op b.v; } \
    static inline bool operator op(const Dual& a, double b) { return a.v op b; }        \
    static inline bool operator op(double a, const Dual& b) { return a op b.v; }
DCMP(<) DCMP(>) DCMP(<=) DCMP(>=) DCMP(==) DCMP(!=)
#undef DCMP

static inline double val(double x) { return x; }
static inline double val(const Dual& x) { return x.v; }

// Op counter: the algorithmic work of one bin (VALU-class ops by kind).
struct OpTally {
    long long add, mul, div, sqrt_, exp_, log_, sincos, atan_, cmp, cbrt_, pow_;
};
static OpTally g_tally;
struct Counted {
    double v;
    Counted() : v(0.0) {}
    Counted(double x) : v(x) {}
};
#define CBIN(op, field)                                                                                   \
    static inline Counted operator op(const Counted& a, const Counted& b) { g_tally.field++; return Counted(a.v op b.v); } \
    static inline Counted operator op(const Counted& a, double b) { g_tally.field++; return Counted(a.v op b); }           \
    static inline Counted operator op(double a, const Counted& b) { g_tally.field++; return Counted(a op b.v); }
CBIN(+, add) CBIN(-, add) CBIN(*, mul) CBIN(/, div)
#undef CBIN
static inline Counted operator-(const Counted& a) { return Counted(-a.v); }
#define CCMP(op)                                                                                           \
    static inline bool operator op(const Counted& a, const Counted& b) { g_tally.cmp++; return a.v op b.v; } \
    static inline bool operator op(const Counted& a, double b) { g_tally.cmp++; return a.v op b; }
CCMP(<) CCMP(>) CCMP(<=) CCMP(>=)
#undef CCMP
static inline double val(const Counted& x) { return x.v; }

// elementary functions (the reference's jnp calls), for each scalar
static inline double o_cbrt(double x) { return std::cbrt(x); }
static inline double o_sqrt(double x) { return std::sqrt(x); }
static inline double o_exp(double x) { return std::exp(x); }
static inline double o_log(double x) { return std::log(x); }
static inline double o_atan(double x) { return std::atan(x); }
static inline double o_cos(double x) { return std::cos(x); }
static inline double o_sin(double x) { return std::sin(x); }
static inline double o_abs(double x) { return std::fabs(x); }
static inline double o_pow(double x, double y) { return std::pow(x, y); }

static inline Dual o_cbrt(const Dual& x) { double c = std::cbrt(x.v); return mk(c, x, c / (3.0 * x.v)); }
static inline Dual o_sqrt(const Dual& x) { double s = std::sqrt(x.v); return mk(s, x, 0.5 / s); }
static inline Dual o_exp(const Dual& x) { double e = std::exp(x.v); return mk(e, x, e); }
static inline Dual o_log(const Dual& x) { return mk(std::log(x.v), x, 1.0 / x.v); }
static inline Dual o_atan(const Dual& x) { return mk(std::atan(x.v), x, 1.0 / (1.0 + x.v * x.v)); }
static inline Dual o_cos(const Dual& x) { return mk(std::cos(x.v), x, -std::sin(x.v)); }
static inline Dual o_sin(const Dual& x) { return mk(std::sin(x.v), x, std::cos(x.v)); }
// d|x| = sign(x) dx, zero at zero (JAX's convention)
static inline Dual o_abs(const Dual& x) {
    return mk(std::fabs(x.v), x, x.v > 0.0 ? 1.0 : (x.v < 0.0 ? -1.0 : 0.0));
}
static inline Dual o_pow(const Dual& x, double y) {
    double p = std::pow(x.v, y);
    return mk(p, x, y * std::pow(x.v, y - 1.0));
}

static inline Counted o_cbrt(const Counted& x) { g_tally.cbrt_++; return Counted(std::cbrt(x.v)); }
static inline Counted o_sqrt(const Counted& x) { g_tally.sqrt_++; return Counted(std::sqrt(x.v)); }
static inline Counted o_exp(const Counted& x) { g_tally.exp_++; return Counted(std::exp(x.v)); }
static inline Counted o_log(const Counted& x) { g_tally.log_++; return Counted(std::log(x.v)); }
static inline Counted o_atan(const Counted& x) { g_tally.atan_++; return Counted(std::atan(x.v)); }
// sin and cos of one angle count once (one sincos)
static inline Counted o_cos(const Counted& x) { g_tally.sincos++; return Counted(std::cos(x.v)); }
static inline Counted o_sin(const Counted& x) { return Counted(std::sin(x.v)); }

// The leading phase term T0 f^(-5/3) - A and the frequency powers take exp,
// log and cbrt from the product's shared-source elementary functions
// (wfx_libm.h, built here by g++ as on the device by hipcc, both with
// contraction off): the reference's leading term reaches ~1e4 rad, where the
// host's and the device's own libraries differ by ~1e-12 rad, so only a
// common source gives the FP64 kernels bit-comparable inputs. Those three
// functions are pinned against long double in tests/test_libm.py.
static inline double o_lead_exp(double x) { return wfx_libm::exp(x); }
static inline double o_lead_log(double x) { return wfx_libm::log(x); }
static inline double o_lead_cbrt(double x) { return wfx_libm::cbrt(x); }
static inline Dual o_lead_exp(const Dual& x) { double e = wfx_libm::exp(x.v); return mk(e, x, e); }
static inline Dual o_lead_log(const Dual& x) { return mk(wfx_libm::log(x.v), x, 1.0 / x.v); }
static inline Dual o_lead_cbrt(const Dual& x) {
    double c = wfx_libm::cbrt(x.v);
    return mk(c, x, c / (3.0 * x.v));
}
static inline Counted o_lead_exp(const Counted& x) { g_tally.exp_++; return Counted(wfx_libm::exp(x.v)); }
static inline Counted o_lead_log(const Counted& x) { g_tally.log_++; return Counted(wfx_libm::log(x.v)); }
static inline Counted o_lead_cbrt(const Counted& x) { g_tally.cbrt_++; return Counted(wfx_libm::cbrt(x.v)); }

template <class A, class B, class C>
static inline auto o_fma(const A& a, const B& b, const C& c) { return a * b + c; }
template <class A, class B, class C>
static inline auto o_fms(const A& a, const B& b, const C& c) { return a * b - c; }
template <class A, class B, class C>
static inline auto o_fnms(const A& a, const B& b, const C& c) { return c - a * b; }
template <class R>
static inline R o_where(bool c, const R& a, const R& b) { return c ? a : b; }
// jnp.minimum: NaN if either is NaN; at a tie JAX splits the derivative
template <class R>
static inline R o_min(const R& a, const R& b) {
    if (std::isnan(val(a)) || std::isnan(val(b))) return R(std::numeric_limits<double>::quiet_NaN());
    if (val(a) == val(b)) return 0.5 * (a + b);
    return val(a) < val(b) ? a : b;
}
// _math.py:522-526
template <class R>
static inline R o_sqrt_nonneg(const R& x) { return x > 0.0 ? o_sqrt(x) : R(0.0); }
// round half to even (jnp.round), on the value
static inline double o_rint(double x) { return std::nearbyint(x); }

// ---------------------------------------------------------------- structs

// IMRPhenomX_UsefulPowers (_math.py:190-246)
template <class R>
struct Powers {
    R one_sixth, m_one_sixth, itself, m_one, one_third, m_one_third, two_thirds, m_two_thirds, four_thirds,
        m_four_thirds, five_thirds, m_five_thirds, seven_thirds, m_seven_thirds, eight_thirds, m_eight_thirds, two,
        three, four, five, m_two, m_three, m_four, m_five, seven_sixths, m_seven_sixths, log, sqrt;
};
template <class R>
static Powers<R> powers(const R& number) {
    Powers<R> p;
    R one_third = o_cbrt(number);
    R sixth = o_sqrt(one_third);
    R m_one = 1.0 / number;
    R log = o_log(number);
    R two_thirds = one_third * one_third;
    R m_one_third = two_thirds * m_one;
    R m_one_sixth = sixth * m_one_third;
    R four_thirds = two_thirds * two_thirds;
    R five_thirds = four_thirds * one_third;
    R seven_thirds = four_thirds * number;
    R eight_thirds = seven_thirds * one_third;
    R m_two_thirds = m_one_third * m_one_third;
    R m_four_thirds = m_two_thirds * m_two_thirds;
    R m_five_thirds = m_four_thirds * m_one_third;
    R m_seven_thirds = m_four_thirds * m_one;
    R m_eight_thirds = m_seven_thirds * m_one_third;
    R two = number * number;
    R three = two * number;
    R m_two = m_one * m_one;
    R m_three = m_two * m_one;
    R m_four = m_two * m_two;
    p.one_sixth = sixth;
    p.m_one_sixth = m_one_sixth;
    p.itself = number;
    p.m_one = m_one;
    p.one_third = one_third;
    p.m_one_third = m_one_third;
    p.two_thirds = two_thirds;
    p.m_two_thirds = m_two_thirds;
    p.four_thirds = four_thirds;
    p.m_four_thirds = m_four_thirds;
    p.five_thirds = five_thirds;
    p.m_five_thirds = m_five_thirds;
    p.seven_thirds = seven_thirds;
    p.m_seven_thirds = m_seven_thirds;
    p.eight_thirds = eight_thirds;
    p.m_eight_thirds = m_eight_thirds;
    p.two = two;
    p.three = three;
    p.four = two * two;
    p.five = three * two;
    p.m_two = m_two;
    p.m_three = m_three;
    p.m_four = m_four;
    p.m_five = m_three * m_two;
    p.seven_sixths = sixth * number;
    p.m_seven_sixths = m_one_sixth * m_one;
    p.log = log;
    p.sqrt = sixth * one_third;
    return p;
}
// POWERS_OF_LALPI (_math.py:249-260)
static const Powers<double>& lalpi() {
    static const Powers<double> p = powers<double>(M_PI_);
    return p;
}

template <class R>
struct WfVars {
    R m1, m2, m_tot, q, eta, delta, eta2, eta3, chi1l, chi2l, chi1l2l, chi1l2, chi1l3, chi2l2, chi2l3, chi_eff,
        chi_pn_hat, s_tot_r, dchi, dchi_half, f_ref, phi0, mf_ref, m_sec, f_min, f_max, f_cut, f_max_prime, distance, amp0,
        amp_norm, dphase0, m_final, a_final, f_ring, f_damp, f_meco, f_isco;
    bool error_mass_ratio, error_spin, error_final_mass, error_final_spin, error_meco;
};

template <class R>
struct AmpCoeffs {
    R f_amp_match_in, f_amp_rd_min, pn_initial, pn_one_third, pn_two_thirds, pn_three_thirds, pn_four_thirds,
        pn_five_thirds, pn_six_thirds, rho1, rho2, rho3, v1_rd, gamma1, gamma2, gamma3, gamma_r, gamma_d2, gamma_d13,
        f_ring, f_damp, delta0, delta1, delta2, delta3, delta4;
};

template <class R>
struct PhaseCoeffs {
    R f_phase_match_in, f_phase_match_im, c0, c1, c2, c4, c_rd, c_l, c_l_gr, c4_ov3, c_l_ov_fda, b0, b1, b2, b3, b4,
        a0, a1, a2, a3, sigma1, sigma2, sigma3, sigma4, phi0, phi2, phi3, phi4, phi5l, phi6, phi6l, phi7, phi8, phi8l,
        phi_norm, dphi0, dphi2, dphi3, dphi4, dphi5, dphi6, dphi6l, dphi7, dphi8, dphi8l, f_ring, f_damp,
        dphase_norm, c1_int, c2_int, c1_mrd, c2_mrd;
};

#include "xas_fits.h"

// ------------------------------------------------------- waveform variables

// internals.py:37-150 (f_max = 0, delta_f = 0: a frequency sequence)
template <class R>
static WfVars<R> waveform_struct(R m1_in, R m2_in, R chi1l_in, R chi2l_in, R f_ref, R phi0, R f_min, R f_max,
                                 R distance, double amp_scale) {
    WfVars<R> wf;
    R mass_ratio = m1_in > m2_in ? m1_in / m2_in : m2_in / m1_in;
    bool swap = m1_in < m2_in;
    R m1 = swap ? m2_in : m1_in;
    R m2 = swap ? m1_in : m2_in;
    R chi1l = swap ? chi2l_in : chi1l_in;
    R chi2l = swap ? chi1l_in : chi2l_in;

    R delta = o_abs((m1 - m2) / (m1 + m2));
    R eta = o_abs(0.25 * o_fnms(delta, delta, 1.0));
    eta = o_min(eta, R(0.25));
    R q = m1 > m2 ? m1 / m2 : m2 / m1;
    q = eta == 0.25 ? R(1.0) : q;

    R m_tot = m1 + m2;
    R m_sec = m_tot * MTSUN_SI_;
    R chi_eff = ut_chi_eff(eta, chi1l, chi2l);
    R dchi = ut_dchi(chi1l, chi2l);
    m1 = m1 / m_tot;
    m2 = m2 / m_tot;

    f_ref = f_ref == 0.0 ? f_min : f_ref;
    R f_cut_def = chi_eff > 0.99 ? R(0.33) : R(0.3);
    R f_cut = f_cut_def / m_sec;
    R fmp = f_max != 0.0 ? f_max : f_cut;
    fmp = o_min(fmp, f_cut);

    wf.m1 = m1;
    wf.m2 = m2;
    wf.m_tot = m_tot;
    wf.q = q;
    wf.eta = eta;
    wf.delta = delta;
    wf.eta2 = eta * eta;
    wf.eta3 = eta * (eta * eta);
    wf.chi1l = chi1l;
    wf.chi2l = chi2l;
    wf.chi1l2l = chi1l * chi2l;
    wf.chi1l2 = chi1l * chi1l;
    wf.chi1l3 = chi1l * chi1l * chi1l;
    wf.chi2l2 = chi2l * chi2l;
    wf.chi2l3 = chi2l * chi2l * chi2l;
    wf.chi_eff = chi_eff;
    wf.chi_pn_hat = ut_chi_pn_hat(eta, chi1l, chi2l);
    wf.s_tot_r = ut_s_tot_r(eta, chi1l, chi2l);
    wf.dchi = dchi;
    wf.dchi_half = dchi * 0.5;  // internals.py:105
    wf.f_ref = f_ref;
    wf.phi0 = phi0;
    wf.mf_ref = f_ref * (MTSUN_SI_ * m_tot);
    wf.m_sec = m_sec;
    wf.f_min = f_min;
    wf.f_max = f_max;
    wf.f_cut = f_cut;
    wf.f_max_prime = fmp;
    wf.distance = distance;
    wf.amp0 = m_tot * (MRSUN_SI_ / MPC_SI_) * (m_tot * MTSUN_SI_) * amp_scale / distance;
    wf.amp_norm = std::sqrt(2.0 / 3.0) * o_sqrt(eta) * lalpi().m_one_sixth;
    wf.dphase0 = 5.0 / (128.0 * std::pow(M_PI_, 5.0 / 3.0));
    wf.error_mass_ratio = (val(mass_ratio) > 1000.0) && (std::fabs(val(mass_ratio) - 1000.0) > 1e-12);
    wf.error_spin = (std::fabs(val(chi1l)) > 1.0) || (std::fabs(val(chi2l)) > 1.0);
    wf.m_final = ut_final_mass_2017(eta, chi1l, chi2l);
    wf.a_final = ut_final_spin_2017(eta, chi1l, chi2l);
    // set_ringdown_frequencies (internals.py:153-158)
    wf.f_ring = qnm_fring22(wf.a_final) / wf.m_final;
    wf.f_damp = qnm_fdamp22(wf.a_final) / wf.m_final;
    wf.f_meco = ut_f_meco(eta, chi1l, chi2l);
    wf.f_isco = ut_f_isco(wf.a_final);
    wf.error_final_mass = val(wf.m_final) > 1.0;
    wf.error_final_spin = std::fabs(val(wf.a_final)) > 1.0;
    wf.error_meco = val(wf.f_meco) > val(wf.f_isco);
    return wf;
}

// internals.py:168-170
template <class R>
static bool any_error(const WfVars<R>& wf) {
    return wf.error_mass_ratio || wf.error_spin || wf.error_final_mass || wf.error_final_spin || wf.error_meco;
}

// --------------------------------------------------- QNM fits in powers of 1 - a

// core.py:183-211: the numerator and denominator of a fit in powers of af,
// re-expanded in powers of u = 1 - af (float64, accumulated as numpy does)
struct AboutOne {
    std::vector<double> c;
};
static AboutOne about_one(const double* coeffs, int n) {
    int len = n;
    while (len > 1 && coeffs[len - 1] == 0.0) --len;  // np.trim_zeros(..., "b")
    std::vector<double> out(len, 0.0);
    std::vector<double> binom(1, 1.0);  // (1 - u)^k in powers of u
    for (int k = 0; k < len; ++k) {
        if (k > 0) {
            std::vector<double> next(k + 1, 0.0);
            for (int j = 0; j < k; ++j) {
                next[j] += binom[j];
                next[j + 1] += -binom[j];
            }
            binom = next;
        }
        for (int j = 0; j <= k; ++j) out[j] = out[j] + coeffs[k] * binom[j];
    }
    return AboutOne{out};
}
// core.py:214-218
template <class R>
static R horner(const std::vector<double>& c, const R& x) {
    R total = R(c.back());
    for (int k = (int)c.size() - 2; k >= 0; --k) total = total * x + c[k];
    return total;
}
struct Qnm1a {
    AboutOne ring_num, ring_den, damp_num, damp_den;
};
static const Qnm1a& qnm_tables() {
    static const Qnm1a t = {about_one(qnm_fring22_numerator, 8), about_one(qnm_fring22_denominator, 7),
                            about_one(qnm_fdamp22_numerator, 7), about_one(qnm_fdamp22_denominator, 7)};
    return t;
}
// core.py:221-229
template <class R>
static void qnm_22(const R& a_final, R& f_ring, R& f_damp) {
    const Qnm1a& t = qnm_tables();
    R u = 1.0 - a_final;
    f_ring = horner(t.ring_num.c, u) / horner(t.ring_den.c, u);
    f_damp = horner(t.damp_num.c, u) / horner(t.damp_den.c, u);
}

// ---------------------------------------------------------------- LU solve

// Partial-pivot LU (the largest |pivot|, first on ties) and substitution:
// the algorithm of gsl_linalg_LU_solve / LAPACK getrf+getrs that
// phase22.py:44-49 calls through jnp.linalg.solve.
template <class R, int N>
static void lu_solve(R a[N][N], R b[N]) {
    for (int k = 0; k < N; ++k) {
        int p = k;
        double best = std::fabs(val(a[k][k]));
        for (int i = k + 1; i < N; ++i) {
            double m = std::fabs(val(a[i][k]));
            if (m > best) {
                best = m;
                p = i;
            }
        }
        if (p != k) {
            for (int j = 0; j < N; ++j) std::swap(a[k][j], a[p][j]);
            std::swap(b[k], b[p]);
        }
        for (int i = k + 1; i < N; ++i) {
            R l = a[i][k] / a[k][k];
            a[i][k] = l;
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

// ------------------------------------------------------ amplitude coefficients

// amp22.py:25-129 (no co-precessing deviations: they are zero for XAS)
template <class R>
static AmpCoeffs<R> amp_coefficients(const WfVars<R>& wf) {
    AmpCoeffs<R> amp;
    R chi1l = wf.chi1l, chi2l = wf.chi2l, chi1l2 = wf.chi1l2, chi1l3 = wf.chi1l3, chi2l2 = wf.chi2l2;
    R delta = wf.delta, eta = wf.eta, eta2 = wf.eta2, eta3 = wf.eta3;
    R f_ring = wf.f_ring, f_damp = wf.f_damp;

    R gamma2 = rd_amp_22_gamma2(eta, wf.s_tot_r, wf.dchi, delta);
    R gamma3 = rd_amp_22_gamma3(eta, wf.s_tot_r, wf.dchi, delta);
    R f_amp_rd_min = rd_amp_22_peak_frequency(gamma2, gamma3, f_ring, f_damp);
    R v1_rd = rd_amp_22_v1(eta, wf.s_tot_r, wf.dchi, delta);
    R f1 = f_amp_rd_min;
    R gamma1 = (v1_rd / (f_damp * gamma3)) *
               o_fma(f_damp * f_damp * gamma3, gamma3, o_fma(f_ring, f_ring, o_fma(f1, f1, -(2.0 * f1 * f_ring)))) *
               o_exp(((f1 - f_ring) * gamma2) / (f_damp * gamma3));
    R gamma_r = gamma2 / (f_damp * gamma3);
    R gamma_d2 = (gamma3 * f_damp) * (gamma3 * f_damp);
    R gamma_d13 = f_damp * gamma1 * gamma3;

    R f_amp_ins_max = wf.f_meco + 0.25 * (wf.f_isco - wf.f_meco);
    R f_amp_match_in = f_amp_ins_max;

    amp.pn_initial = 1.0;
    amp.pn_one_third = 0.0;
    amp22_pn_terms(chi1l, chi1l2, chi1l3, chi2l, chi2l2, delta, eta, eta2, eta3, amp.pn_two_thirds,
                   amp.pn_three_thirds, amp.pn_four_thirds, amp.pn_five_thirds, amp.pn_six_thirds);

    R fi2 = 0.50 * f_amp_match_in, fi3 = 0.75 * f_amp_match_in, fi4 = 1.00 * f_amp_match_in;
    R vi2 = ins_amp_22_v2(eta, wf.chi_pn_hat, wf.dchi, delta);
    R vi3 = ins_amp_22_v3(eta, wf.chi_pn_hat, wf.dchi, delta);
    R vi4 = ins_amp_22_v4(eta, wf.chi_pn_hat, wf.dchi, delta);
    amp.rho1 = ins_amp_22_rho1(vi2, vi3, vi4, fi2, fi3, fi4);
    amp.rho2 = ins_amp_22_rho2(vi2, vi3, vi4, fi2, fi3, fi4);
    amp.rho3 = ins_amp_22_rho3(vi2, vi3, vi4, fi2, fi3, fi4);

    amp.f_amp_match_in = f_amp_match_in;
    amp.f_amp_rd_min = f_amp_rd_min;
    amp.v1_rd = v1_rd;
    amp.gamma1 = gamma1;
    amp.gamma2 = gamma2;
    amp.gamma3 = gamma3;
    amp.gamma_r = gamma_r;
    amp.gamma_d2 = gamma_d2;
    amp.gamma_d13 = gamma_d13;
    amp.f_ring = f_ring;
    amp.f_damp = f_damp;

    R fa = amp.f_amp_match_in;
    R fd = amp.f_amp_rd_min;
    Powers<R> p1 = powers(fa);
    Powers<R> p4 = powers(fd);
    R d1 = ins_amp_22_dansatz(fa, wf, amp);
    R d4 = rd_amp_22_dansatz(fd, f_ring, f_damp, amp);
    R insp_f1 = ins_amp_22_ansatz_exact(fa, p1, amp);
    R rd_f4 = rd_amp_22_ansatz_exact(fd, f_ring, f_damp, amp);
    d1 = ((7.0 / 6.0) * p1.one_sixth / insp_f1) - (p1.seven_sixths * d1 / (insp_f1 * insp_f1));
    d4 = ((7.0 / 6.0) * p4.one_sixth / rd_f4) - (p4.seven_sixths * d4 / (rd_f4 * rd_f4));
    R f2 = fa + (1.0 / 2.0) * (fd - fa);
    R v1 = 1.0 / (p1.m_seven_sixths * insp_f1);
    R v2 = 1.0 / itm_amp_22_va(eta, wf.s_tot_r, wf.dchi, delta);
    R v4 = 1.0 / (p4.m_seven_sixths * rd_f4);
    amp.delta0 = itm_amp_22_delta0(d1, d4, v1, v2, v4, fa, f2, fd);
    amp.delta1 = itm_amp_22_delta1(d1, d4, v1, v2, v4, fa, f2, fd);
    amp.delta2 = itm_amp_22_delta2(d1, d4, v1, v2, v4, fa, f2, fd);
    amp.delta3 = itm_amp_22_delta3(d1, d4, v1, v2, v4, fa, f2, fd);
    amp.delta4 = itm_amp_22_delta4(d1, d4, v1, v2, v4, fa, f2, fd);
    return amp;
}

// ---------------------------------------------------------- phase coefficients

// phase22.py:266-286
template <class R>
static void connection_coefficients(PhaseCoeffs<R>& ph) {
    R f_ins = ph.f_phase_match_in;
    R f_int = ph.f_phase_match_im;
    Powers<R> p_ins = powers(f_ins);
    R dphi_ins = ins_phase_22_ansatz_exact(f_ins, p_ins, ph);
    R dphi_int = itm_phase_22_ansatz_exact(f_ins, p_ins, ph);
    ph.c2_int = dphi_ins - dphi_int;
    R phi_in = ins_phase_22_ansatz_int_exact(f_ins, p_ins, ph);
    R phi_im = itm_phase_22_ansatz_int_exact(f_ins, p_ins, ph);
    ph.c1_int = o_fma(-ph.c2_int, f_ins, phi_in - phi_im);

    Powers<R> p_int = powers(f_int);
    R phi_imc = o_fma(ph.c2_int, f_int, itm_phase_22_ansatz_int_exact(f_int, p_int, ph) + ph.c1_int);
    R phi_rd = rd_phase_22_ansatz_int_exact(f_int, p_int, ph);
    R dphi_intc = itm_phase_22_ansatz_exact(f_int, p_int, ph) + ph.c2_int;
    R dphi_rd = rd_phase_22_ansatz_exact(f_int, p_int, ph);
    ph.c2_mrd = dphi_intc - dphi_rd;
    ph.c1_mrd = o_fma(-ph.c2_mrd, f_int, phi_imc - phi_rd);
}

// phase22.py:52-263
template <class R>
static PhaseCoeffs<R> phase_coefficients(const WfVars<R>& wf) {
    PhaseCoeffs<R> ph;
    R f_ring = wf.f_ring, f_damp = wf.f_damp, dphase0 = wf.dphase0;

    R f_im_match = 0.6 * (0.5 * f_ring + wf.f_isco);
    R f_in_match = wf.f_meco;
    R deltaf = (f_im_match - f_in_match) * 0.03;
    R f_phase_match_in = f_in_match - 1.0 * deltaf;
    R f_phase_match_im = f_im_match + 0.5 * deltaf;
    R f_phase_ins_min = 0.0026;
    R f_phase_ins_max = 1.020 * wf.f_meco;
    R f_phase_rd_min = f_im_match;
    R f_phase_rd_max = o_fma(1.25, f_damp, f_ring);
    R phi_norm = -(3.0 * lalpi().m_five_thirds) / 128.0;

    R chi1l = wf.chi1l, chi2l = wf.chi2l;
    R chi1l2l = chi1l * chi2l, chi1l2 = chi1l * chi1l, chi1l3 = chi1l * chi1l2;
    R chi2l2 = chi2l * chi2l, chi2l3 = chi2l * chi2l2;
    R eta = wf.eta, eta2 = eta * eta, eta3 = eta * eta2, delta = wf.delta;
    R s = wf.s_tot_r, dchi = wf.dchi;

    // ringdown: 5 x 5 collocation (phase22.py:89-118)
    const double g5[5] = {0.0, 0.5 - 1.0 / (2.0 * std::pow(2.0, 0.5)), 0.5, 0.5 + 1.0 / (2.0 * std::pow(2.0, 0.5)),
                          1.0};
    R deltax = f_phase_rd_max - f_phase_rd_min;
    R points_rd[5];
    for (int i = 0; i < 5; ++i) points_rd[i] = o_fma(g5[i], deltax, f_phase_rd_min);
    points_rd[3] = f_ring;
    R rd_v4 = rd_phase_22_v4(eta, s, dchi, delta);
    R vrd[5] = {rd_phase_22_d12(eta, s, dchi, delta), rd_phase_22_d24(eta, s, dchi, delta),
                rd_phase_22_d34(eta, s, dchi, delta), rd_v4, rd_phase_22_d54(eta, s, dchi, delta)};
    vrd[4] = vrd[4] + vrd[3];
    vrd[2] = vrd[2] + vrd[3];
    vrd[1] = vrd[1] + vrd[3];
    vrd[0] = vrd[0] + vrd[1];
    R phase_rd = vrd[0];
    R a5[5][5];
    for (int i = 0; i < 5; ++i) {
        R ff = points_rd[i];
        R invff = 1.0 / ff;
        R ff1 = o_cbrt(invff);
        R ff2 = invff * invff;
        R ff3 = ff2 * ff2;
        R ff4 = -dphase0 / o_fma(f_damp, f_damp, (ff - f_ring) * (ff - f_ring));
        a5[i][0] = 1.0;
        a5[i][1] = ff1;
        a5[i][2] = ff2;
        a5[i][3] = ff3;
        a5[i][4] = ff4;
    }
    lu_solve<R, 5>(a5, vrd);
    R c0 = vrd[0], c1 = vrd[1], c2 = vrd[2], c4 = vrd[3], c_rd = vrd[4];
    R c_l = -(dphase0 * c_rd);

    // inspiral: 4 x 4 collocation (phase22.py:121-141)
    const double g4[4] = {0.0, 1.0 / 4.0, 3.0 / 4.0, 1.0};
    deltax = f_phase_ins_max - f_phase_ins_min;
    R s_ins = wf.chi_pn_hat;
    R v3 = ins_phase_22_v3(eta, s_ins, dchi, delta);
    R vins[4] = {ins_phase_22_d13(eta, s_ins, dchi, delta) + v3, ins_phase_22_d23(eta, s_ins, dchi, delta) + v3, v3,
                 ins_phase_22_d43(eta, s_ins, dchi, delta) + v3};
    R a4[4][4];
    for (int i = 0; i < 4; ++i) {
        R ff = o_fma(g4[i], deltax, f_phase_ins_min);
        R ff1 = o_cbrt(ff);
        R ff2 = ff1 * ff1;
        a4[i][0] = 1.0;
        a4[i][1] = ff1;
        a4[i][2] = ff2;
        a4[i][3] = ff;
    }
    lu_solve<R, 4>(a4, vins);
    R a0 = vins[0], a1 = vins[1], a2 = vins[2], a3 = vins[3];
    R sigma1 = (-5.0 / 3.0) * a0;
    R sigma2 = (-5.0 / 4.0) * a1;
    R sigma3 = (-5.0 / 5.0) * a2;
    R sigma4 = (-5.0 / 6.0) * a3;

    // TaylorF2 (phase22.py:143-173)
    R phi2_ns, phi3_ns, phi3_s, phi4_ns, phi4_s, phi5l_ns, phi5l_s, phi6_ns, phi6_s, phi6l_ns, phi7_ns, phi7_s, phi8_s,
        phi8l_s;
    phase22_pn_terms(chi1l, chi1l2, chi1l2l, chi1l3, chi2l, chi2l2, chi2l3, delta, eta, eta2, eta3, phi2_ns, phi3_ns,
                     phi3_s, phi4_ns, phi4_s, phi5l_ns, phi5l_s, phi6_ns, phi6_s, phi6l_ns, phi7_ns, phi7_s, phi8_s,
                     phi8l_s);
    R phi0 = 1.0;
    R phi2 = phi2_ns;
    R phi3 = phi3_ns + phi3_s;
    R phi4 = phi4_ns + phi4_s;
    R phi5l = phi5l_ns + phi5l_s;
    R phi6 = phi6_ns + phi6_s;
    R phi6l = phi6l_ns;
    R phi7 = phi7_ns + phi7_s;
    R phi8 = phi8_s;
    R phi8l = phi8l_s;

    R dphi0 = phi0;
    R dphi2 = 3.0 / 5.0 * phi2;
    R dphi3 = 2.0 / 5.0 * phi3;
    R dphi4 = 1.0 / 5.0 * phi4;
    R dphi5 = -3.0 / 5.0 * phi5l;
    R dphi6 = o_fma(-(1.0 / 5.0), phi6, -(3.0 / 5.0 * phi6l));
    R dphi6l = -1.0 / 5.0 * phi6l;
    R dphi7 = -2.0 / 5.0 * phi7;
    R dphi8 = o_fma(-(3.0 / 5.0), phi8, -(3.0 / 5.0 * phi8l));
    R dphi8l = -3.0 / 5.0 * phi8l;

    // the derivative at the inspiral-intermediate transition (phase22.py:186-202)
    Powers<R> p = powers(f_phase_match_in);
    R phase_in = dphi0;
    phase_in = o_fma(dphi2, p.two_thirds, phase_in);
    phase_in = o_fma(dphi3, p.itself, phase_in);
    phase_in = o_fma(dphi4, p.four_thirds, phase_in);
    phase_in = o_fma(dphi5, p.five_thirds, phase_in);
    phase_in = o_fma(dphi6, p.two, phase_in);
    phase_in = o_fma(dphi6l * p.two, p.log, phase_in);
    phase_in = o_fma(dphi7, p.seven_thirds, phase_in);
    phase_in = o_fma(dphi8, p.eight_thirds, phase_in);
    phase_in = o_fma(dphi8l * p.eight_thirds, p.log, phase_in);
    R pseudo = o_fma(a0, p.eight_thirds, a1 * p.three);
    pseudo = o_fma(a2 * p.eight_thirds, p.two_thirds, pseudo);
    pseudo = o_fma(a3 * p.eight_thirds, p.itself, pseudo);
    phase_in = phase_in + pseudo;
    phase_in = phase_in * p.m_eight_thirds * dphase0;

    // intermediate: 5 x 5 collocation (phase22.py:204-239)
    deltax = f_phase_match_im - f_phase_match_in;
    R v2_im_m_rd_v4 = itm_phase_22_v2m_rdv4(eta, s, dchi, delta);
    R v3_im_m_rd_v4 = itm_phase_22_v3m_rdv4(eta, s, dchi, delta);
    R v2_im = itm_phase_22_v2(eta, s, dchi, delta);
    R vint[5] = {phase_in, o_fma(0.75, v2_im_m_rd_v4 + rd_v4, 0.25 * v2_im), v3_im_m_rd_v4 + rd_v4,
                 itm_phase_22_d43(eta, s, dchi, delta), phase_rd};
    vint[3] = vint[3] + vint[2];
    R b5[5][5], rhs[5];
    for (int i = 0; i < 5; ++i) {
        R ff = o_fma(g5[i], deltax, f_phase_match_in);
        R ff1 = f_ring / ff;
        R ff2 = ff1 * ff1;
        R ff3 = ff1 * ff2;
        R ff4 = ff2 * ff2;
        R ff0 = i == 0 ? (4.0 * c_l) / o_fma(2.0 * f_damp, 2.0 * f_damp, (ff - f_ring) * (ff - f_ring))
                       : (4.0 * c_l) / o_fma(4.0 * f_damp, f_damp, (ff - f_ring) * (ff - f_ring));
        b5[i][0] = 1.0;
        b5[i][1] = ff1;
        b5[i][2] = ff2;
        b5[i][3] = ff3;
        b5[i][4] = ff4;
        rhs[i] = vint[i] - ff0;
    }
    lu_solve<R, 5>(b5, rhs);
    R b0 = rhs[0];
    R b1 = rhs[1] * f_ring;
    R b2 = rhs[2] * f_ring * f_ring;
    R b3 = rhs[3] * f_ring * f_ring * f_ring;
    R b4 = rhs[4] * f_ring * f_ring * f_ring * f_ring;

    ph.f_phase_match_in = f_phase_match_in;
    ph.f_phase_match_im = f_phase_match_im;
    ph.c0 = c0;
    ph.c1 = c1;
    ph.c2 = c2;
    ph.c4 = c4;
    ph.c_rd = c_rd;
    ph.c_l = c_l;
    ph.c_l_gr = c_l;
    ph.c4_ov3 = c4 / 3.0;
    ph.c_l_ov_fda = c_l / f_damp;
    ph.b0 = b0;
    ph.b1 = b1;
    ph.b2 = b2;
    ph.b3 = b3;
    ph.b4 = b4;
    ph.a0 = a0;
    ph.a1 = a1;
    ph.a2 = a2;
    ph.a3 = a3;
    ph.sigma1 = sigma1;
    ph.sigma2 = sigma2;
    ph.sigma3 = sigma3;
    ph.sigma4 = sigma4;
    ph.phi0 = phi0;
    ph.phi2 = phi2;
    ph.phi3 = phi3;
    ph.phi4 = phi4;
    ph.phi5l = phi5l;
    ph.phi6 = phi6;
    ph.phi6l = phi6l;
    ph.phi7 = phi7;
    ph.phi8 = phi8;
    ph.phi8l = phi8l;
    ph.phi_norm = phi_norm;
    ph.dphi0 = dphi0;
    ph.dphi2 = dphi2;
    ph.dphi3 = dphi3;
    ph.dphi4 = dphi4;
    ph.dphi5 = dphi5;
    ph.dphi6 = dphi6;
    ph.dphi6l = dphi6l;
    ph.dphi7 = dphi7;
    ph.dphi8 = dphi8;
    ph.dphi8l = dphi8l;
    ph.f_ring = f_ring;
    ph.f_damp = f_damp;
    ph.dphase_norm = ins_inspiral_dphase_norm();
    ph.c1_int = 0.0;
    ph.c2_int = 0.0;
    ph.c1_mrd = 0.0;
    ph.c2_mrd = 0.0;
    connection_coefficients(ph);
    return ph;
}

// phase22.py:310-316
template <class R>
static R dphase_22(const R& mf, const Powers<R>& p, const PhaseCoeffs<R>& ph) {
    R ins = ins_phase_22_ansatz_exact(mf, p, ph);
    R rd = rd_phase_22_ansatz_exact(mf, p, ph) + ph.c2_mrd;
    R itm = itm_phase_22_ansatz_exact(mf, p, ph) + ph.c2_int;
    return mf < ph.f_phase_match_in ? ins : (mf >= ph.f_phase_match_im ? rd : itm);
}

// phase22.py:319-326
template <class R>
static R time_shift_22(const PhaseCoeffs<R>& ph, const WfVars<R>& wf) {
    R linb = ut_linb(wf.eta, wf.s_tot_r, wf.dchi, wf.delta);
    R f_ref_fit = wf.f_ring - wf.f_damp;
    R dphi22_ref = 1.0 / wf.eta * dphase_22(f_ref_fit, powers(f_ref_fit), ph);
    R psi4_to_strain = ut_psi4_to_strain(wf.eta, wf.s_tot_r, wf.dchi);
    return o_fma(-(2.0 * M_PI_), 500.0 + psi4_to_strain, linb - dphi22_ref);
}

// ----------------------------------------------------------------- the table

// core.py:49-53 basis indices and core.py:68-101 layout
enum {
    B_M53 = 0, B_M1, B_M23, B_M13, B_LOG, B_ONE, B_P13, B_P13L, B_P23, B_P1, B_P1L, B_P43, B_P53, B_P2, B_M2, B_M3,
    NBASIS
};
enum {
    T_PHASE = 0,          // 3 x 16
    T_BIG = 48,           // T0, A
    T_BIG_LO = 50,
    T_ATAN_A = 52,        // 3
    T_ATAN_W = 55,        // 3
    T_SHIFT = 58,
    T_AMP_INS = 59,       // 10
    T_AMP_INT_N = 69,
    T_AMP_INT_D = 70,     // 5
    T_AMP_RD_N = 75,
    T_AMP_RD_GR = 76,
    T_AMP_RD_S = 77,
    T_AMP_RD_G = 78,
    T_F_PHASE_IN = 79,
    T_F_PHASE_IM = 80,
    T_F_AMP_IN = 81,
    T_F_AMP_RD = 82,
    T_F_CUT = 83,
    T_KP = 84,
    T_KC = 85,
    TABLE_SIZE = 86
};

// the powers of f (core.py:382-392): f, t = f^(1/3), 1/f, log f, f^(-7/6),
// f^(-5/3)
template <class R>
struct FreqPowers {
    R f, t, m1, log_f, m76, m53;
};
template <class R>
static FreqPowers<R> frequency_powers(const R& f) {
    FreqPowers<R> fp;
    R m1 = 1.0 / f;
    R log_f = o_lead_log(f);
    R t = o_lead_exp(log_f / 3.0);
    t = t * (4.0 - t * t * t * m1) / 3.0;
    fp.f = f;
    fp.t = t;
    fp.m1 = m1;
    fp.log_f = log_f;
    fp.m76 = m1 / o_sqrt(t);
    fp.m53 = o_lead_exp(-5.0 / 3.0 * o_lead_log(f));
    return fp;
}

// core.py:395-415
template <class R>
static void phase_basis(const FreqPowers<R>& fp, R basis[NBASIS]) {
    R t2 = fp.t * fp.t;
    R mt1 = t2 * fp.m1;
    R mt2 = mt1 * mt1;
    R m2 = fp.m1 * fp.m1;
    basis[B_M53] = fp.m1 * mt2;
    basis[B_M1] = fp.m1;
    basis[B_M23] = mt2;
    basis[B_M13] = mt1;
    basis[B_LOG] = fp.log_f;
    basis[B_ONE] = R(1.0);  // None in the reference: the constant is the start of each sum
    basis[B_P13] = fp.t;
    basis[B_P13L] = fp.t * fp.log_f;
    basis[B_P23] = t2;
    basis[B_P1] = fp.f;
    basis[B_P1L] = fp.f * fp.log_f;
    basis[B_P43] = fp.f * fp.t;
    basis[B_P53] = fp.f * t2;
    basis[B_P2] = fp.f * fp.f;
    basis[B_M2] = m2;
    basis[B_M3] = m2 * fp.m1;
}
static const int REGION0[] = {B_M1, B_M23, B_M13, B_LOG, B_P13, B_P13L, B_P23, B_P1, B_P1L, B_P43, B_P53, B_P2};
static const int REGION1[] = {B_P1, B_LOG, B_M1, B_M2, B_M3};
static const int REGION2[] = {B_P1, B_P23, B_M1, B_M3};

// core.py:423-442: the phase but for the leading term (and -A); ``past``
// whether f >= f_phase_in
template <class R, class T>
static R phase_rest(const T* tab, const FreqPowers<R>& fp, bool& past) {
    R f = fp.f;
    past = f >= tab[T_F_PHASE_IN];
    bool ringdown = f > tab[T_F_PHASE_IM];
    R basis[NBASIS];
    phase_basis(fp, basis);
    int r = ringdown ? 2 : (past ? 1 : 0);
    const int* terms = r == 0 ? REGION0 : (r == 1 ? REGION1 : REGION2);
    int nterms = r == 0 ? 12 : (r == 1 ? 5 : 4);
    const T* row = tab + T_PHASE + 16 * r;
    R total = R(row[B_ONE]);
    for (int i = 0; i < nterms; ++i) total = total + row[terms[i]] * basis[terms[i]];
    if (past) {
        R y = f - tab[T_SHIFT];
        R a = R(tab[T_ATAN_A + (ringdown ? 2 : 1)]);
        R w = R(tab[T_ATAN_W + (ringdown ? 2 : 1)]);
        total = total + a * o_atan(w * y);
    }
    return total;
}

// core.py:465-468 (turns carry no derivative)
template <class R>
static R reduce_phase(const R& phi) {
    R q = phi / (2.0 * M_PI_);
    double turns = o_rint(val(q));
    return phi - turns * (2.0 * M_PI_);
}

// core.py:445-462 and 471-507 for one bin: h+ = (hp[0], hp[1]), hx likewise
template <class R, class T>
static void evaluate_point(const T* tab, const FreqPowers<R>& fp, R hp[2], R hc[2]) {
    bool past;
    R rest = phase_rest(tab, fp, past);
    R phi = (past ? R(0.0) : (tab[T_BIG] + tab[T_BIG_LO]) * fp.m53) + rest;
    phi = phi - (tab[T_BIG + 1] + tab[T_BIG_LO + 1]);
    phi = reduce_phase(phi);

    R f = fp.f;
    bool in_amp = f >= tab[T_F_AMP_IN];
    bool rd_amp = f > tab[T_F_AMP_RD];
    R amp;
    if (rd_amp) {
        R u = tab[T_AMP_RD_S] * (f - tab[T_SHIFT]);
        amp = tab[T_AMP_RD_N] * fp.m76 * o_exp(-tab[T_AMP_RD_GR] * u) / (u * u + tab[T_AMP_RD_G]);
    } else if (in_amp) {
        R den = R(tab[T_AMP_INT_D + 4]);
        for (int j = 3; j >= 0; --j) den = den * f + tab[T_AMP_INT_D + j];
        amp = tab[T_AMP_INT_N] / den;
    } else {
        R poly = R(tab[T_AMP_INS + 9]);
        for (int k = 8; k >= 0; --k) poly = poly * fp.t + tab[T_AMP_INS + k];
        amp = fp.m76 * poly;
    }
    amp = f <= tab[T_F_CUT] ? amp : R(0.0);
    R c = o_cos(phi), s = o_sin(phi);
    R ap = tab[T_KP] * amp;
    R ac = tab[T_KC] * amp;
    hp[0] = ap * c;
    hp[1] = ap * s;
    hc[0] = -ac * s;
    hc[1] = ac * c;
}

// core.py:247-254
template <class R>
static void add_term(R row[NBASIS], const R& c, int a_index, int a_log_index, bool with_log, const R& s_a,
                     const R& log_s) {
    R scaled = c * s_a;
    row[a_index] = row[a_index] + (with_log ? scaled * log_s : scaled);
    if (with_log) row[a_log_index] = row[a_log_index] + scaled;
}

// core.py:270-278
template <class R>
static WfVars<R> waveform_variables(const R* params, const R& f_min, double amp_scale) {
    WfVars<R> wf = waveform_struct(params[0], params[1], params[2], params[3], params[7], params[6], f_min, R(0.0),
                                   params[4], amp_scale);
    R fr, fd;
    qnm_22(wf.a_final, fr, fd);
    wf.f_ring = fr / wf.m_final;
    wf.f_damp = fd / wf.m_final;
    return wf;
}

// core.py:286-379: params = (m1, m2, chi1, chi2, distance, inclination,
// phi_ref, f_ref); writes the 86 entries, returns LAL's error flags
template <class R>
static unsigned table_of(const R* params, const R& f_min, double amp_scale, R* tab) {
    WfVars<R> wf = waveform_variables(params, f_min, amp_scale);
    AmpCoeffs<R> amp = amp_coefficients(wf);
    PhaseCoeffs<R> ph = phase_coefficients(wf);
    R linb = time_shift_22(ph, wf);
    R inveta = 1.0 / wf.eta;

    R s = wf.m_sec;
    Powers<R> p = powers(s);
    R s_pow[NBASIS];
    s_pow[B_M53] = p.m_five_thirds;
    s_pow[B_M1] = p.m_one;
    s_pow[B_M23] = p.m_two_thirds;
    s_pow[B_M13] = p.m_one_third;
    s_pow[B_ONE] = R(1.0);
    s_pow[B_P13] = p.one_third;
    s_pow[B_P23] = p.two_thirds;
    s_pow[B_P1] = s;
    s_pow[B_P43] = p.four_thirds;
    s_pow[B_P53] = p.five_thirds;
    s_pow[B_P2] = p.two;
    s_pow[B_M2] = p.m_two;
    s_pow[B_M3] = p.m_three;
    R norm = inveta * ph.phi_norm;

    R rows[3][NBASIS];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < NBASIS; ++k) rows[r][k] = R(0.0);
    // (coefficient, basis index of f^a, of f^a log f, with the log) in the
    // reference's order, the time shift last
    auto term = [&](int r, const R& c, int ia, int il, bool lg) { add_term(rows[r], c, ia, il, lg, s_pow[ia], p.log); };
    term(0, norm * ph.phi2, B_M1, 0, false);
    term(0, norm * ph.phi3, B_M23, 0, false);
    term(0, norm * ph.phi4, B_M13, 0, false);
    term(0, norm * ph.phi5l, B_ONE, B_LOG, true);
    term(0, norm * ph.phi6, B_P13, 0, false);
    term(0, norm * ph.phi6l, B_P13, B_P13L, true);
    term(0, norm * ph.phi7, B_P23, 0, false);
    term(0, norm * (ph.phi8 + ph.sigma1), B_P1, 0, false);
    term(0, norm * ph.phi8l, B_P1, B_P1L, true);
    term(0, norm * ph.sigma2, B_P43, 0, false);
    term(0, norm * ph.sigma3, B_P53, 0, false);
    term(0, norm * ph.sigma4, B_P2, 0, false);
    term(0, linb, B_P1, 0, false);

    term(1, inveta * (ph.b0 + ph.c2_int), B_P1, 0, false);
    term(1, inveta * ph.b1, B_ONE, B_LOG, true);
    term(1, -inveta * ph.b2, B_M1, 0, false);
    term(1, -inveta * ph.b3 / 2.0, B_M2, 0, false);
    term(1, -inveta * ph.b4 / 3.0, B_M3, 0, false);
    term(1, inveta * ph.c1_int, B_ONE, 0, false);
    term(1, linb, B_P1, 0, false);

    term(2, inveta * (ph.c0 + ph.c2_mrd), B_P1, 0, false);
    term(2, inveta * 1.5 * ph.c1, B_P23, 0, false);
    term(2, -inveta * ph.c2, B_M1, 0, false);
    term(2, -inveta * ph.c4_ov3, B_M3, 0, false);
    term(2, inveta * ph.c1_mrd, B_ONE, 0, false);
    term(2, linb, B_P1, 0, false);

    R f_damp = wf.f_damp, f_ring = wf.f_ring;
    for (int k = 0; k < TABLE_SIZE; ++k) tab[k] = R(0.0);
    tab[T_ATAN_A + 0] = R(0.0);
    tab[T_ATAN_A + 1] = inveta * 2.0 * ph.c_l_gr / f_damp;
    tab[T_ATAN_A + 2] = inveta * ph.c_l_ov_fda;
    tab[T_ATAN_W + 0] = R(0.0);
    tab[T_ATAN_W + 1] = s / (2.0 * f_damp);
    tab[T_ATAN_W + 2] = s / f_damp;
    tab[T_SHIFT] = f_ring / s;
    tab[T_F_PHASE_IN] = ph.f_phase_match_in / s;
    tab[T_F_PHASE_IM] = ph.f_phase_match_im / s;

    R prefactor = wf.amp0 * wf.amp_norm;
    R scale = prefactor * p.m_seven_sixths;
    R third[10] = {R(1.0), p.one_third, p.two_thirds, s, p.four_thirds, p.five_thirds, p.two, p.two * p.one_third,
                   p.two * p.two_thirds, p.three};
    R ins[10] = {R(1.0), R(0.0), amp.pn_two_thirds, amp.pn_three_thirds, amp.pn_four_thirds, amp.pn_five_thirds,
                 amp.pn_six_thirds, amp.rho1, amp.rho2, amp.rho3};
    for (int k = 0; k < 10; ++k) tab[T_AMP_INS + k] = scale * ins[k] * third[k];
    R s_int[5] = {R(1.0), s, p.two, p.three, p.four};
    R deltas[5] = {amp.delta0, amp.delta1, amp.delta2, amp.delta3, amp.delta4};
    for (int j = 0; j < 5; ++j) tab[T_AMP_INT_D + j] = deltas[j] * s_int[j];
    tab[T_AMP_INT_N] = prefactor;
    tab[T_AMP_RD_N] = scale * amp.gamma_d13;
    tab[T_AMP_RD_GR] = amp.gamma_r;
    tab[T_AMP_RD_S] = s;
    tab[T_AMP_RD_G] = amp.gamma_d2;
    tab[T_F_AMP_IN] = amp.f_amp_match_in / s;
    tab[T_F_AMP_RD] = amp.f_amp_rd_min / s;
    tab[T_F_CUT] = wf.f_max_prime;

    unsigned flags = (wf.error_mass_ratio ? 1u : 0u) | (wf.error_spin ? 2u : 0u) | (wf.error_final_mass ? 4u : 0u) |
                     (wf.error_final_spin ? 8u : 0u) | (wf.error_meco ? 16u : 0u);
    R inclination = params[5];
    R cfac = o_cos(inclination);
    R pfac = 0.5 * (1.0 + cfac * cfac);
    const double YLM = 2.0 * std::sqrt(5.0 / (64.0 * M_PI_));
    double nan = std::numeric_limits<double>::quiet_NaN();
    tab[T_KP] = flags ? R(nan) : -YLM * pfac;
    tab[T_KC] = flags ? R(nan) : YLM * cfac;

    // the leading term in float64 (core.py:342-344)
    R m1d = params[0], m2d = params[1];
    const double K0 = -3.0 / 128.0 * std::pow(M_PI_ * MTSUN_SI_, -5.0 / 3.0);
    R t0 = K0 * o_lead_cbrt(m1d + m2d) / (m1d * m2d);
    // the reference phase (core.py:351-358): phi(f_ref) = 2 phi0 + pi / 4
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < NBASIS; ++k) tab[T_PHASE + 16 * r + k] = rows[r][k];
    R f_ref = wf.f_ref;
    bool past;
    R rest = phase_rest(tab, frequency_powers(f_ref), past);
    R shift_constant = 2.0 * wf.phi0 + M_PI_ / 4.0 - rest;
    for (int r = 0; r < 3; ++r) tab[T_PHASE + 16 * r + B_ONE] = tab[T_PHASE + 16 * r + B_ONE] + shift_constant;
    R anchor = past ? R(0.0) : t0 * o_lead_exp(-5.0 / 3.0 * o_lead_log(f_ref));
    tab[T_BIG] = t0;
    tab[T_BIG + 1] = anchor;
    tab[T_BIG_LO] = R(0.0);
    tab[T_BIG_LO + 1] = R(0.0);
    return flags;
}

// model.py:35-45: f_ref <- reference_mf / M where f_ref was zero, NaN
// otherwise (no derivative with respect to the given f_ref)
template <class R>
static void with_reference(R* params, double reference_mf) {
    if (std::isnan(reference_mf)) return;
    R scaled = reference_mf / (MTSUN_SI_ * (params[0] + params[1]));
    params[7] = val(params[7]) == 0.0 ? scaled : R(std::numeric_limits<double>::quiet_NaN());
}

// ------------------------------------------------------------- LAL form

// xpnr/xas.py:20-61: frequency_sequence, then the XASModel polarization
// factors; an independent arrangement of the same model (per Mf, the QNM
// fits in powers of a_f, the reference phase through phase_22).
static void lal_form(const double* params, const double* freqs, long n, double amp_scale, double* hp, double* hc) {
    double m1 = params[0], m2 = params[1], chi1 = params[2], chi2 = params[3], distance = params[4],
           inclination = params[5], phi0 = params[6], f_ref = params[7];
    WfVars<double> wf =
        waveform_struct<double>(m1, m2, chi1, chi2, f_ref, phi0, freqs[0], freqs[n - 1], distance, amp_scale);
    AmpCoeffs<double> amp = amp_coefficients(wf);
    PhaseCoeffs<double> ph = phase_coefficients(wf);
    double linb = time_shift_22(ph, wf);
    double lina = 0.0;
    double inveta = 1.0 / wf.eta;
    double amp_prefactor = wf.amp0 * wf.amp_norm;
    // phifRef (xas.py:31-36) with phase_22 (phase22.py:304-307)
    double mf_ref = wf.mf_ref;
    Powers<double> pr = powers(mf_ref);
    double ins = ins_phase_22_ansatz_int_exact(mf_ref, pr, ph);
    double rd = o_fma(ph.c2_mrd, mf_ref, rd_phase_22_ansatz_int_exact(mf_ref, pr, ph) + ph.c1_mrd);
    double itm = o_fma(ph.c2_int, mf_ref, itm_phase_22_ansatz_int_exact(mf_ref, pr, ph) + ph.c1_int);
    double phi_ref22 = mf_ref < ph.f_phase_match_in ? ins : (mf_ref >= ph.f_phase_match_im ? rd : itm);
    double phif_ref = -(o_fma(linb, mf_ref, inveta * phi_ref22) + lina) + 2.0 * wf.phi0 + M_PI_ / 4.0;
    double mf_cut = wf.f_max_prime * wf.m_sec;
    const double YLM = 2.0 * std::sqrt(5.0 / (64.0 * M_PI_));
    double ci = std::cos(inclination);
    double kp = -YLM * 0.5 * (1.0 + ci * ci), kc = YLM * ci;
    if (any_error(wf)) kp = kc = std::numeric_limits<double>::quiet_NaN();
    for (long i = 0; i < n; ++i) {
        double mf = wf.m_sec * freqs[i];
        Powers<double> p = powers(mf);
        // phase_xas (phase22.py:296-301)
        double pi_ = ins_phase_22_ansatz_int_exact(mf, p, ph);
        double pr_ = o_fma(ph.c2_mrd, mf, rd_phase_22_ansatz_int_exact(mf, p, ph) + ph.c1_mrd);
        double pm_ = o_fma(ph.c2_int, mf, itm_phase_22_ansatz_int_exact(mf, p, ph) + ph.c1_int);
        double phi = (mf < ph.f_phase_match_in ? pi_ : (mf > ph.f_phase_match_im ? pr_ : pm_)) * inveta;
        phi = phi + (o_fma(linb, mf, lina) + phif_ref);
        // amp_ansatz (amp22.py:132-139)
        double ai = ins_amp_22_ansatz(mf, p, amp);
        double ar = rd_amp_22_ansatz(mf, amp);
        double am = itm_amp_22_ansatz(mf, p, amp);
        double a = amp_prefactor * p.m_seven_sixths *
                   (mf < amp.f_amp_match_in ? ai : (mf > amp.f_amp_rd_min ? ar : am));
        double re = mf <= mf_cut ? a * std::cos(phi) : 0.0;
        double im = mf <= mf_cut ? a * std::sin(phi) : 0.0;
        hp[2 * i] = kp * re;
        hp[2 * i + 1] = kp * im;
        hc[2 * i] = -kc * im;
        hc[2 * i + 1] = kc * re;
    }
}

// ------------------------------------------------------------------ C ABI

extern "C" {

int oracle_table_size() { return TABLE_SIZE; }

// the shared-source elementary functions on an array: which 0 = exp, 1 = log,
// 2 = cbrt (for their accuracy tests and the numpy IMRPhenomD oracle's
// leading term)
void oracle_libm(int which, const double* x, double* y, long n) {
    for (long i = 0; i < n; ++i)
        y[i] = which == 0 ? wfx_libm::exp(x[i]) : which == 1 ? wfx_libm::log(x[i]) : wfx_libm::cbrt(x[i]);
}

// tables (B, 86) and LAL flags (B,) of params (B, 8), f_min (B,)
void oracle_xas_tables(const double* params, const double* f_min, long nb, double amp_scale, double reference_mf,
                       double* tables, unsigned* flags) {
#pragma omp parallel for schedule(dynamic, 16)
    for (long b = 0; b < nb; ++b) {
        double p[8];
        for (int k = 0; k < 8; ++k) p[k] = params[8 * b + k];
        with_reference(p, reference_mf);
        flags[b] = table_of<double>(p, f_min[b], amp_scale, tables + TABLE_SIZE * b);
    }
}

// (h+, hx) of tables (B, 86) on a shared grid freqs (N,): hp, hc (B, N, 2)
void oracle_xas_eval(const double* tables, long nb, const double* freqs, long n, double* hp, double* hc) {
#pragma omp parallel for schedule(dynamic, 4)
    for (long b = 0; b < nb; ++b) {
        const double* tab = tables + TABLE_SIZE * b;
        for (long i = 0; i < n; ++i) {
            FreqPowers<double> fp = frequency_powers(freqs[i]);
            double p[2], c[2];
            evaluate_point(tab, fp, p, c);
            hp[2 * (b * n + i)] = p[0];
            hp[2 * (b * n + i) + 1] = p[1];
            hc[2 * (b * n + i)] = c[0];
            hc[2 * (b * n + i) + 1] = c[1];
        }
    }
}

// model.py:48-60 batched: params (B, 8) on ragged grids (freqs at offsets
// [off[b], off[b+1])), f_min = each binary's first frequency; offsets NULL
// means one shared grid of n frequencies
void oracle_xas_polarizations(const double* params, long nb, const double* freqs, const long* offsets, long n,
                              double amp_scale, double reference_mf, double* hp, double* hc, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
#pragma omp parallel for schedule(dynamic, 4)
    for (long b = 0; b < nb; ++b) {
        long start = offsets ? offsets[b] : 0;
        long stop = offsets ? offsets[b + 1] : n;
        long out = offsets ? offsets[b] : b * n;
        if (stop <= start) continue;
        double p[8], tab[TABLE_SIZE];
        for (int k = 0; k < 8; ++k) p[k] = params[8 * b + k];
        with_reference(p, reference_mf);
        table_of<double>(p, freqs[start], amp_scale, tab);
        for (long i = start; i < stop; ++i) {
            FreqPowers<double> fp = frequency_powers(freqs[i]);
            double a[2], c[2];
            evaluate_point(tab, fp, a, c);
            long o = out + (i - start);
            hp[2 * o] = a[0];
            hp[2 * o + 1] = a[1];
            hc[2 * o] = c[0];
            hc[2 * o + 1] = c[1];
        }
    }
}

// forward-mode derivatives of (h+, hx) with respect to the 8 parameters
// (m1, m2, chi1, chi2, distance, inclination, phi_ref, f_ref), shared grid:
// dhp, dhc (B, N, 8, 2); the grid's first frequency carries no derivative
void oracle_xas_jacobian(const double* params, long nb, const double* freqs, long n, double amp_scale,
                         double* dhp, double* dhc) {
#pragma omp parallel for schedule(dynamic, 1)
    for (long b = 0; b < nb; ++b) {
        Dual p[8];
        for (int k = 0; k < 8; ++k) {
            p[k] = Dual(params[8 * b + k]);
            p[k].d[k] = 1.0;
        }
        Dual tab[TABLE_SIZE];
        table_of<Dual>(p, Dual(freqs[0]), amp_scale, tab);
        for (long i = 0; i < n; ++i) {
            FreqPowers<Dual> fp = frequency_powers(Dual(freqs[i]));
            Dual a[2], c[2];
            evaluate_point(tab, fp, a, c);
            for (int j = 0; j < ND; ++j)
                for (int r = 0; r < 2; ++r) {
                    dhp[((b * n + i) * ND + j) * 2 + r] = a[r].d[j];
                    dhc[((b * n + i) * ND + j) * 2 + r] = c[r].d[j];
                }
        }
    }
}

// table derivatives (B, 86, 8) from forward mode through the table build
void oracle_xas_table_jacobian(const double* params, const double* f_min, long nb, double amp_scale, double* tables,
                               double* dtables) {
#pragma omp parallel for schedule(dynamic, 4)
    for (long b = 0; b < nb; ++b) {
        Dual p[8];
        for (int k = 0; k < 8; ++k) {
            p[k] = Dual(params[8 * b + k]);
            p[k].d[k] = 1.0;
        }
        Dual tab[TABLE_SIZE];
        table_of<Dual>(p, Dual(f_min[b]), amp_scale, tab);
        for (int k = 0; k < TABLE_SIZE; ++k) {
            tables[TABLE_SIZE * b + k] = tab[k].v;
            for (int j = 0; j < ND; ++j) dtables[(TABLE_SIZE * b + k) * ND + j] = tab[k].d[j];
        }
    }
}

void oracle_xas_lal_form(const double* params, const double* freqs, long n, double amp_scale, double* hp,
                         double* hc) {
    lal_form(params, freqs, n, amp_scale, hp, hc);
}

// the QNM fits at a_final: in powers of 1 - a (core.py) and of a (qnm.py)
void oracle_qnm22(double a, double* out4) {
    qnm_22<double>(a, out4[0], out4[1]);
    out4[2] = qnm_fring22<double>(a);
    out4[3] = qnm_fdamp22<double>(a);
}

// f_ISCO of a Kerr hole of spin a (utilities.py:16-26)
double oracle_f_isco(double a) { return ut_f_isco<double>(a); }

// x of A x = b by the oracle's pivoted LU (n = 4 or 5), A row-major
int oracle_solve(int n, const double* a, const double* b, double* x) {
    if (n == 4) {
        double m[4][4], v[4];
        for (int i = 0; i < 4; ++i) {
            v[i] = b[i];
            for (int j = 0; j < 4; ++j) m[i][j] = a[4 * i + j];
        }
        lu_solve<double, 4>(m, v);
        for (int i = 0; i < 4; ++i) x[i] = v[i];
        return 0;
    }
    if (n == 5) {
        double m[5][5], v[5];
        for (int i = 0; i < 5; ++i) {
            v[i] = b[i];
            for (int j = 0; j < 5; ++j) m[i][j] = a[5 * i + j];
        }
        lu_solve<double, 5>(m, v);
        for (int i = 0; i < 5; ++i) x[i] = v[i];
        return 0;
    }
    return -1;
}

// named intermediate values of one binary, for the closed-form checks:
// out = [eta, delta, chi_eff, m_final, a_final, f_ring, f_damp (1 - a form,
// Mf), f_meco, f_isco, m_sec, T0, f_phase_in (Hz), f_phase_im, f_amp_in,
// f_amp_rd, f_cut, phi2..phi7 (the TaylorF2 coefficients, 6), linb, flags]
int oracle_xas_internals(const double* params, double f_min, double amp_scale, double* out) {
    double p[8];
    for (int k = 0; k < 8; ++k) p[k] = params[k];
    WfVars<double> wf = waveform_variables<double>(p, f_min, amp_scale);
    PhaseCoeffs<double> ph = phase_coefficients(wf);
    AmpCoeffs<double> amp = amp_coefficients(wf);
    double tab[TABLE_SIZE];
    unsigned flags = table_of<double>(p, f_min, amp_scale, tab);
    int i = 0;
    out[i++] = wf.eta;
    out[i++] = wf.delta;
    out[i++] = wf.chi_eff;
    out[i++] = wf.m_final;
    out[i++] = wf.a_final;
    out[i++] = wf.f_ring;
    out[i++] = wf.f_damp;
    out[i++] = wf.f_meco;
    out[i++] = wf.f_isco;
    out[i++] = wf.m_sec;
    out[i++] = tab[T_BIG];
    out[i++] = tab[T_F_PHASE_IN];
    out[i++] = tab[T_F_PHASE_IM];
    out[i++] = tab[T_F_AMP_IN];
    out[i++] = tab[T_F_AMP_RD];
    out[i++] = tab[T_F_CUT];
    out[i++] = ph.phi2;
    out[i++] = ph.phi3;
    out[i++] = ph.phi4;
    out[i++] = ph.phi5l;
    out[i++] = ph.phi6;
    out[i++] = ph.phi7;
    out[i++] = time_shift_22(ph, wf);
    out[i++] = (double)flags;
    (void)amp;
    return i;
}

// algorithmic work of the per-bin formula (core.py:382-507), tallied on the
// given tables and shared grid: the per-frequency powers once per
// frequency, the rest once per bin. out: {add, mul, div, sqrt, exp, log,
// sincos, atan, cmp, cbrt, pow} for the frequencies, then for the bins
void oracle_xas_count(const double* tables, long nb, const double* freqs, long n, long long* out) {
    std::memset(&g_tally, 0, sizeof g_tally);
    std::vector<FreqPowers<Counted>> fps(n);
    for (long i = 0; i < n; ++i) fps[i] = frequency_powers(Counted(freqs[i]));
    OpTally freq = g_tally;
    std::memset(&g_tally, 0, sizeof g_tally);
    for (long b = 0; b < nb; ++b)
        for (long i = 0; i < n; ++i) {
            Counted a[2], c[2];
            evaluate_point(tables + TABLE_SIZE * b, fps[i], a, c);
        }
    OpTally bins = g_tally;
    long long* f = reinterpret_cast<long long*>(&freq);
    long long* g = reinterpret_cast<long long*>(&bins);
    for (int k = 0; k < 11; ++k) {
        out[k] = f[k];
        out[11 + k] = g[k];
    }
}

}  // extern "C"

// IMRPhenomXHM (test infrastructure as the rest of this file)
#include "xhm_oracle.h"
