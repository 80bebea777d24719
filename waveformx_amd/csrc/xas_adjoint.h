// The adjoint of the per-bin formula of IMRPhenomXAS with respect to the
// binary's record: given the cotangents g+ = dL/dRe h+ + i dL/dIm h+ (and
// gx), add dL/d(record slot) of one bin to acc[slot].
//
// With E = e^{i phi}, h+ = kp A E and hx = i kc A E:
//   dL/dA   = Re P,  dL/dphi = -A Im P,
//   P = conj(g+) kp E + conj(gx) i kc E,
//   dL/dkp  = A Re(conj(g+) E),  dL/dkc = -A Im(conj(gx) E);
// then through the phase sums (the basis values of each region), the
// arctangent a atan(w (f - shift)), the leading term T0 f^(-5/3) - A, and
// the three amplitude forms. The region bounds and the cutoff carry no
// derivative (the reference's jnp.where selections). Intermediates are
// recomputed, none stored.
#pragma once

#include <type_traits>

#include "xas_point.h"

namespace wfx {

// (the FP32 adjoint uses the forward kernel's fast forms through m_sincos,
// d_atan, d_exp, d_rcp: |phi| <= pi, so the hardware sine and cosine need
// no range reduction; FP64 the library's)

template <class T>
struct BinPowers {
    T f, t, m1, lf, m76;
    double m53;
};

// one frequency's powers as stored in a grid's table (32 B for float, 48 B
// for double): the adjoint kernels, which walk every frequency once per
// binary, read them instead of recomputing them per bin
template <class T>
struct PowerRow {
    T f, t, m1, lf, m76;
    double m53;
};

__device__ inline BinPowers<float> bin_powers(const PowF32& p) {
    return BinPowers<float>{p.f, p.t, p.m1, p.lf, p.m76, p.m53};
}
__device__ inline BinPowers<double> bin_powers(const FreqPowers<double>& p) {
    return BinPowers<double>{p.f, p.t, p.m1, p.log_f, p.m76, p.m53};
}

// h+, hx of one bin (the forward formula in T; the leading term and its
// reduction in FP64)
template <class T>
__device__ inline T phase_and_amp(const T* __restrict__ c, const BinPowers<T>& p, bool& past, bool& ringdown,
                                  T& z, T& amp) {
    const T f = p.f, t = p.t, m1 = p.m1, lf = p.lf;
    past = f >= c[kFPhaseIn];
    ringdown = f > c[kFPhaseIm];
    T rest;
    z = T(0);
    if (!past) {
        const T t2 = t * t, mt1 = t2 * m1, mt2 = mt1 * mt1;
        rest = c[kInsConst] + c[kInsM1] * m1 + c[kInsM23] * mt2 + c[kInsM13] * mt1 + c[kInsLog] * lf +
               c[kInsP13] * t + c[kInsP13L] * (t * lf) + c[kInsP23] * t2 + c[kInsP1] * f + c[kInsP1L] * (f * lf) +
               c[kInsP43] * (f * t) + c[kInsP53] * (f * t2) + c[kInsP2] * (f * f);
    } else if (!ringdown) {
        const T m2 = m1 * m1;
        z = d_atan(c[kAtanW1] * (f - c[kShift]));
        rest = c[kIntConst] + c[kIntP1] * f + c[kIntLog] * lf + c[kIntM1] * m1 + c[kIntM2] * m2 +
               c[kIntM3] * (m2 * m1) + c[kAtanA1] * z;
    } else {
        const T m2 = m1 * m1;
        z = d_atan(c[kAtanW2] * (f - c[kShift]));
        rest = c[kRdConst] + c[kRdP1] * f + c[kRdP23] * (t * t) + c[kRdM1] * m1 + c[kRdM3] * (m2 * m1) +
               c[kAtanA2] * z;
    }
    T phi;
    if (std::is_same<T, float>::value && past) {
        phi = (T)reduce_f32((float)rest);  // FP32 records hold -A in these constants (fold_anchor)
    } else {
        double full = (double)rest - ((double)c[kAnchor] + (double)c[kAnchorLo]);
        if (!past) full = fma((double)c[kT0] + (double)c[kT0Lo], p.m53, full);
        const double turns = rint(full * (1.0 / kTwoPi));
        phi = (T)fma(-turns, kTwoPi, full);
    }

    if (f > c[kFAmpRd]) {
        const T u = c[kAmpRdS] * (f - c[kShift]);
        amp = c[kAmpRdN] * p.m76 * d_exp(-c[kAmpRdGr] * u) * d_rcp(u * u + c[kAmpRdG]);
    } else if (f >= c[kFAmpIn]) {
        T den = c[kAmpIntD + 4];
        for (int j = 3; j >= 0; --j) den = den * f + c[kAmpIntD + j];
        amp = c[kAmpIntN] * d_rcp(den);
    } else {
        T poly = c[kAmpIns + 9];
        for (int k = 8; k >= 0; --k) poly = poly * t + c[kAmpIns + k];
        amp = p.m76 * poly;
    }
    if (!(f <= c[kFCut])) amp = T(0);
    return phi;
}

// One bin's dL/d(slot), added to acc, with the polarization factors kp, kc
// given (adjoint_point: the record's; IMRPhenomXHM's (2,2) takes kp = 1,
// kc = 0, the cotangent of the mode itself), the bin's amplitude and its
// phase's sine and cosine returned (zero above the cutoff). Inlined before
// the optimiser simplifies it on its own: compiled as a function, acc is
// opaque memory, and the compiler merges the regions' last updates into one
// store through a selected address, which later keeps the caller's
// accumulator array out of registers.
template <class T>
__device__ __forceinline__ void adjoint_point_k(const T* __restrict__ c, const BinPowers<T>& p, T kp, T kc, T gp_re,
                                                T gp_im, T gx_re, T gx_im, T* acc, T& amp, T& s, T& co) {
    bool past, ringdown;
    T z;
    const T phi = phase_and_amp(c, p, past, ringdown, z, amp);
    if (!(p.f <= c[kFCut])) {  // zero above the cutoff, whatever the record
        amp = s = co = T(0);
        return;
    }
    m_sincos(phi, &s, &co);
    const T re_p = kp * (gp_re * co + gp_im * s) + kc * (gx_im * co - gx_re * s);
    const T im_p = kp * (gp_re * s - gp_im * co) + kc * (gx_re * co + gx_im * s);
    const T g_amp = re_p;
    const T g_phi = -amp * im_p;
    acc[kKp] += amp * (gp_re * co + gp_im * s);
    acc[kKc] += amp * (gx_im * co - gx_re * s);

    const T f = p.f, t = p.t, m1 = p.m1, lf = p.lf;
    // the phase
    acc[kAnchor] -= g_phi;
    if (!past) {
        const T t2 = t * t, mt1 = t2 * m1, mt2 = mt1 * mt1;
        acc[kT0] += g_phi * (T)p.m53;
        acc[kInsConst] += g_phi;
        acc[kInsM1] += g_phi * m1;
        acc[kInsM23] += g_phi * mt2;
        acc[kInsM13] += g_phi * mt1;
        acc[kInsLog] += g_phi * lf;
        acc[kInsP13] += g_phi * t;
        acc[kInsP13L] += g_phi * (t * lf);
        acc[kInsP23] += g_phi * t2;
        acc[kInsP1] += g_phi * f;
        acc[kInsP1L] += g_phi * (f * lf);
        acc[kInsP43] += g_phi * (f * t);
        acc[kInsP53] += g_phi * (f * t2);
        acc[kInsP2] += g_phi * (f * f);
        // an empty statement that the compiler may not move or merge: the
        // branches then end differently, and their last updates are not
        // merged into one store through a selected slot address (which
        // would keep the accumulators in scratch memory)
        __asm__ volatile("");
    } else {
        const T y = f - c[kShift];
        const T a = ringdown ? c[kAtanA2] : c[kAtanA1];
        const T w = ringdown ? c[kAtanW2] : c[kAtanW1];
        const T wy = w * y;
        const T datan = d_rcp(T(1) + wy * wy);
        acc[kShift] -= g_phi * a * w * datan;
        const T m2 = m1 * m1, m3 = m2 * m1;
        // both regions' slots, the other one's with a zero cotangent (no
        // pair of same-shaped branches for the compiler to merge)
        const T g_rd = ringdown ? g_phi : T(0), g_int = ringdown ? T(0) : g_phi;
        acc[kAtanA2] += g_rd * z;
        acc[kAtanW2] += g_rd * a * y * datan;
        acc[kRdConst] += g_rd;
        acc[kRdP1] += g_rd * f;
        acc[kRdP23] += g_rd * (t * t);
        acc[kRdM1] += g_rd * m1;
        acc[kRdM3] += g_rd * m3;
        acc[kAtanA1] += g_int * z;
        acc[kAtanW1] += g_int * a * y * datan;
        acc[kIntConst] += g_int;
        acc[kIntP1] += g_int * f;
        acc[kIntLog] += g_int * lf;
        acc[kIntM1] += g_int * m1;
        acc[kIntM2] += g_int * m2;
        acc[kIntM3] += g_int * m3;
    }
    // the amplitude
    if (f > c[kFAmpRd]) {
        const T y = f - c[kShift];
        const T u = c[kAmpRdS] * y;
        const T inv_q = d_rcp(u * u + c[kAmpRdG]);
        const T e = p.m76 * d_exp(-c[kAmpRdGr] * u) * inv_q;  // dA/dn
        acc[kAmpRdN] += g_amp * e;
        acc[kAmpRdGr] -= g_amp * u * amp;
        acc[kAmpRdG] -= g_amp * amp * inv_q;
        const T dadu = amp * (-c[kAmpRdGr] - T(2) * u * inv_q);
        acc[kAmpRdS] += g_amp * dadu * y;
        acc[kShift] -= g_amp * dadu * c[kAmpRdS];
    } else if (f >= c[kFAmpIn]) {
        T den = c[kAmpIntD + 4];
        for (int j = 3; j >= 0; --j) den = den * f + c[kAmpIntD + j];
        const T inv = d_rcp(den);
        acc[kAmpIntN] += g_amp * inv;
        T q = -g_amp * amp * inv;
        for (int j = 0; j < 5; ++j) {
            acc[kAmpIntD + j] += q;
            q *= f;
        }
    } else {
        T q = g_amp * p.m76;
        for (int k = 0; k < 10; ++k) {
            acc[kAmpIns + k] += q;
            q *= t;
        }
    }
}

template <class T>
__device__ __forceinline__ void adjoint_point(const T* __restrict__ c, const BinPowers<T>& p, T gp_re, T gp_im, T gx_re,
                                              T gx_im, T* acc) {
    T amp, s, co;
    adjoint_point_k(c, p, c[kKp], c[kKc], gp_re, gp_im, gx_re, gx_im, acc, amp, s, co);
}

// dL/df of one bin (the frequency itself moving): dL/dA A'(f) + dL/dphi
// phi'(f), with phi' the derivative of the region's phase (basis terms,
// arctangent, leading term -5/3 T0 f^(-8/3)) and A' that of the region's
// amplitude; zero above the cutoff. The polarization factors kp, kc given
// (grid_adjoint_point: the record's; IMRPhenomXHM's (2,2) takes kp = 1,
// kc = 0, the cotangent of the mode itself).
template <class T>
__device__ inline T grid_adjoint_point_k(const T* __restrict__ c, const BinPowers<T>& p, T kp, T kc, T gp_re, T gp_im,
                                         T gx_re, T gx_im) {
    bool past, ringdown;
    T z, amp;
    const T phi = phase_and_amp(c, p, past, ringdown, z, amp);
    if (!(p.f <= c[kFCut])) return T(0);
    T s, co;
    m_sincos(phi, &s, &co);
    const T g_amp = kp * (gp_re * co + gp_im * s) + kc * (gx_im * co - gx_re * s);
    const T g_phi = -amp * (kp * (gp_re * s - gp_im * co) + kc * (gx_re * co + gx_im * s));
    const T f = p.f, t = p.t, m1 = p.m1, lf = p.lf;
    // d/df of f^a is a f^(a-1); of f^a log f, f^(a-1) (a log f + 1)
    T dphi;
    if (!past) {
        const T t2 = t * t, mt1 = t2 * m1, mt2 = mt1 * mt1, m2 = m1 * m1;
        const T third = T(1.0 / 3.0);
        dphi = -c[kInsM1] * m2 - T(2) * third * c[kInsM23] * mt2 * m1 - third * c[kInsM13] * mt1 * m1 +
               c[kInsLog] * m1 + third * c[kInsP13] * t * m1 + c[kInsP13L] * t * m1 * (third * lf + T(1)) +
               T(2) * third * c[kInsP23] * t2 * m1 + c[kInsP1] + c[kInsP1L] * (lf + T(1)) +
               T(4) * third * c[kInsP43] * t + T(5) * third * c[kInsP53] * t2 + T(2) * c[kInsP2] * f;
        const double t0 = (double)c[kT0] + (double)c[kT0Lo];
        dphi += (T)(-5.0 / 3.0 * t0 * p.m53 / (double)f);
    } else {
        const T y = f - c[kShift];
        const T a = ringdown ? c[kAtanA2] : c[kAtanA1];
        const T w = ringdown ? c[kAtanW2] : c[kAtanW1];
        const T wy = w * y;
        dphi = a * w / (T(1) + wy * wy);
        const T m2 = m1 * m1;
        if (ringdown)
            dphi += c[kRdP1] + T(2.0 / 3.0) * c[kRdP23] * (t * t) * m1 - c[kRdM1] * m2 - T(3) * c[kRdM3] * m2 * m2;
        else
            dphi += c[kIntP1] + c[kIntLog] * m1 - c[kIntM1] * m2 - T(2) * c[kIntM2] * m2 * m1 -
                    T(3) * c[kIntM3] * m2 * m2;
    }
    T damp;
    if (f > c[kFAmpRd]) {
        const T u = c[kAmpRdS] * (f - c[kShift]);
        damp = amp * (T(-7.0 / 6.0) * m1 - c[kAmpRdGr] * c[kAmpRdS] -
                      T(2) * u * c[kAmpRdS] / (u * u + c[kAmpRdG]));
    } else if (f >= c[kFAmpIn]) {
        T den = c[kAmpIntD + 4], dden = T(4) * c[kAmpIntD + 4];
        for (int j = 3; j >= 0; --j) {
            den = den * f + c[kAmpIntD + j];
            if (j > 0) dden = dden * f + T(j) * c[kAmpIntD + j];
        }
        damp = -amp * dden / den;
    } else {
        // A = f^(-7/6) sum_k c_k t^k: A' = f^(-13/6) sum_k (k/3 - 7/6) c_k t^k
        T poly = T(9.0 / 3.0 - 7.0 / 6.0) * c[kAmpIns + 9];
        for (int k = 8; k >= 0; --k) poly = poly * t + T(k / 3.0 - 7.0 / 6.0) * c[kAmpIns + k];
        damp = p.m76 * m1 * poly;
    }
    return g_amp * damp + g_phi * dphi;
}

template <class T>
__device__ inline T grid_adjoint_point(const T* __restrict__ c, const BinPowers<T>& p, T gp_re, T gp_im, T gx_re,
                                       T gx_im) {
    return grid_adjoint_point_k(c, p, c[kKp], c[kKc], gp_re, gp_im, gx_re, gx_im);
}

}  // namespace wfx
