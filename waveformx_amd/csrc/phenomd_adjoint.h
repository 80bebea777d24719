// The adjoint of IMRPhenomD's per-bin formula (phenomd_point.h) with
// respect to its record: given the cotangents g+ = dL/dRe h+ + i dL/dIm h+
// (and gx), add dL/d(record slot) of one bin to acc[slot], as the
// IMRPhenomXAS adjoint does (xas_adjoint.h): dL/dA = Re P, dL/dphi =
// -A Im P, then through the region's phase sum (the basis values), the
// merger-ringdown's a atan(w (f - shift)), the leading term T0 f^(-5/3) - A,
// and the three amplitude forms. The region bounds carry no derivative (the
// reference's jnp.where selections); there is no cutoff.
#pragma once

#include "phenomd_point.h"
#include "xas_adjoint.h"

namespace wfx {

template <class T>
__device__ inline T pd_f34(T f) {
    const T root = m_sqrt(f);
    return root * m_sqrt(root);
}
__device__ inline float pd_f34(float f) {
    const float root = sqrtf(f);
    return root * sqrtf(root);
}

// phi and A of one bin (the forward formula in T; the leading term and its
// reduction in FP64; FP32 records hold -A in the post-inspiral constants)
template <class T>
__device__ inline T pd_phase_and_amp(const T* __restrict__ c, const BinPowers<T>& p, T f34, bool& inspiral,
                                     bool& mrd, T& z, T& amp) {
    const T f = p.f, t = p.t, m1 = p.m1, lf = p.lf;
    inspiral = f < c[kPdFPhaseIns];
    mrd = f >= c[kPdFPhaseMrd];
    z = T(0);
    T rest;
    if (inspiral) {
        const T t2 = t * t, mt1 = t2 * m1;
        rest = c[kPdInsConst] + c[kPdInsM1] * m1 + c[kPdInsM23] * (mt1 * mt1) + c[kPdInsM13] * mt1 +
               c[kPdInsLog] * lf + c[kPdInsP13] * t + c[kPdInsP13L] * (t * lf) + c[kPdInsP23] * t2 +
               c[kPdInsP1] * f + c[kPdInsP43] * (f * t) + c[kPdInsP53] * (f * t2) + c[kPdInsP2] * (f * f);
    } else if (mrd) {
        z = d_atan(c[kPdAtanW] * (f - c[kPdAtanShift]));
        rest = c[kPdMrdConst] + c[kPdMrdP1] * f + c[kPdMrdM1] * m1 + c[kPdMrdP34] * f34 + c[kPdAtanA] * z;
    } else {
        rest = c[kPdIntConst] + c[kPdIntP1] * f + c[kPdIntM3] * (m1 * m1 * m1) + c[kPdIntLog] * lf;
    }
    T phi;
    if (std::is_same<T, float>::value && !inspiral) {
        phi = (T)reduce_f32((float)rest);
    } else {
        double full = (double)rest - ((double)c[kAnchor] + (double)c[kAnchorLo]);
        if (inspiral) full = fma((double)c[kT0] + (double)c[kT0Lo], p.m53, full);
        const double turns = rint(full * (1.0 / kTwoPi));
        phi = (T)fma(-turns, kTwoPi, full);
    }

    if (f < c[kPdFAmpIns]) {
        T poly = c[kPdAmpIns + 9];
        for (int k = 8; k >= 0; --k) poly = poly * t + c[kPdAmpIns + k];
        amp = p.m76 * poly;
    } else if (f >= c[kPdFAmpMrd]) {
        const T u = c[kPdAmpRdS] * (f - c[kPdAmpRdShift]);
        amp = c[kPdAmpRdN] * p.m76 * d_exp(-c[kPdAmpRdGr] * u) * d_rcp(u * u + T(1));
    } else {
        const T y = c[kPdAmpIntS] * f;
        T poly = c[kPdAmpInt + 4];
        for (int j = 3; j >= 0; --j) poly = poly * y + c[kPdAmpInt + j];
        amp = p.m76 * poly;
    }
    return phi;
}

// (forced inline, and the regions' updates written so that no two branches
// end alike: see adjoint_point)
template <class T>
__device__ __forceinline__ void pd_adjoint_point(const T* __restrict__ c, const BinPowers<T>& p, T gp_re, T gp_im,
                                                 T gx_re, T gx_im, T* acc) {
    const T f34 = pd_f34(p.f);
    bool inspiral, mrd;
    T z, amp;
    const T phi = pd_phase_and_amp(c, p, f34, inspiral, mrd, z, amp);
    T s, co;
    m_sincos(phi, &s, &co);
    const T kp = c[kPdKp], kc = c[kPdKc];
    const T g_amp = kp * (gp_re * co + gp_im * s) + kc * (gx_im * co - gx_re * s);
    const T g_phi = -amp * (kp * (gp_re * s - gp_im * co) + kc * (gx_re * co + gx_im * s));
    acc[kPdKp] += amp * (gp_re * co + gp_im * s);
    acc[kPdKc] += amp * (gx_im * co - gx_re * s);

    const T f = p.f, t = p.t, m1 = p.m1, lf = p.lf;
    acc[kAnchor] -= g_phi;
    if (inspiral) {
        const T t2 = t * t, mt1 = t2 * m1;
        acc[kT0] += g_phi * (T)p.m53;
        acc[kPdInsConst] += g_phi;
        acc[kPdInsM1] += g_phi * m1;
        acc[kPdInsM23] += g_phi * (mt1 * mt1);
        acc[kPdInsM13] += g_phi * mt1;
        acc[kPdInsLog] += g_phi * lf;
        acc[kPdInsP13] += g_phi * t;
        acc[kPdInsP13L] += g_phi * (t * lf);
        acc[kPdInsP23] += g_phi * t2;
        acc[kPdInsP1] += g_phi * f;
        acc[kPdInsP43] += g_phi * (f * t);
        acc[kPdInsP53] += g_phi * (f * t2);
        acc[kPdInsP2] += g_phi * (f * f);
        __asm__ volatile("");
    } else {
        const T g_mrd = mrd ? g_phi : T(0), g_int = mrd ? T(0) : g_phi;
        const T a = c[kPdAtanA], w = c[kPdAtanW];
        const T y = f - c[kPdAtanShift];
        const T wy = w * y;
        const T datan = d_rcp(T(1) + wy * wy);
        acc[kPdMrdConst] += g_mrd;
        acc[kPdMrdP1] += g_mrd * f;
        acc[kPdMrdM1] += g_mrd * m1;
        acc[kPdMrdP34] += g_mrd * f34;
        acc[kPdAtanA] += g_mrd * z;
        acc[kPdAtanW] += g_mrd * a * y * datan;
        acc[kPdAtanShift] -= g_mrd * a * w * datan;
        acc[kPdIntConst] += g_int;
        acc[kPdIntP1] += g_int * f;
        acc[kPdIntM3] += g_int * (m1 * m1 * m1);
        acc[kPdIntLog] += g_int * lf;
    }

    if (f < c[kPdFAmpIns]) {
        T q = g_amp * p.m76;
        for (int k = 0; k < 10; ++k) {
            acc[kPdAmpIns + k] += q;
            q *= t;
        }
    } else if (f >= c[kPdFAmpMrd]) {
        const T d = f - c[kPdAmpRdShift];
        const T u = c[kPdAmpRdS] * d;
        const T inv_q = d_rcp(u * u + T(1));
        const T e = p.m76 * d_exp(-c[kPdAmpRdGr] * u) * inv_q;  // dA/dn
        acc[kPdAmpRdN] += g_amp * e;
        acc[kPdAmpRdGr] -= g_amp * u * amp;
        const T dadu = amp * (-c[kPdAmpRdGr] - T(2) * u * inv_q);
        acc[kPdAmpRdS] += g_amp * dadu * d;
        acc[kPdAmpRdShift] -= g_amp * dadu * c[kPdAmpRdS];
    } else {
        const T y = c[kPdAmpIntS] * f;
        T dpoly = T(4) * c[kPdAmpInt + 4];
        for (int j = 3; j >= 1; --j) dpoly = dpoly * y + T(j) * c[kPdAmpInt + j];
        acc[kPdAmpIntS] += g_amp * p.m76 * dpoly * f;
        T q = g_amp * p.m76;
        for (int j = 0; j < 5; ++j) {
            acc[kPdAmpInt + j] += q;
            q *= y;
        }
    }
}

// dL/df of one bin (the frequency itself moving; the tables fixed): dL/dA
// A'(f) + dL/dphi phi'(f), with phi' the derivative of the region's phase
// (basis terms, the merger-ringdown's arctangent, the inspiral's leading
// term -5/3 T0 f^(-8/3)) and A' that of the region's amplitude form.
template <class T>
__device__ inline T pd_grid_adjoint_point(const T* __restrict__ c, const BinPowers<T>& p, T gp_re, T gp_im, T gx_re,
                                          T gx_im) {
    const T f34 = pd_f34(p.f);
    bool inspiral, mrd;
    T z, amp;
    const T phi = pd_phase_and_amp(c, p, f34, inspiral, mrd, z, amp);
    T s, co;
    m_sincos(phi, &s, &co);
    const T kp = c[kPdKp], kc = c[kPdKc];
    const T g_amp = kp * (gp_re * co + gp_im * s) + kc * (gx_im * co - gx_re * s);
    const T g_phi = -amp * (kp * (gp_re * s - gp_im * co) + kc * (gx_re * co + gx_im * s));
    const T f = p.f, t = p.t, m1 = p.m1, lf = p.lf;
    const T m2 = m1 * m1;
    // d/df of f^a is a f^(a-1); of f^a log f, f^(a-1) (a log f + 1)
    T dphi;
    if (inspiral) {
        const T t2 = t * t, mt1 = t2 * m1;
        const T third = T(1.0 / 3.0);
        dphi = -c[kPdInsM1] * m2 - T(2) * third * c[kPdInsM23] * (mt1 * mt1) * m1 - third * c[kPdInsM13] * mt1 * m1 +
               c[kPdInsLog] * m1 + third * c[kPdInsP13] * t * m1 + c[kPdInsP13L] * t * m1 * (third * lf + T(1)) +
               T(2) * third * c[kPdInsP23] * t2 * m1 + c[kPdInsP1] + T(4) * third * c[kPdInsP43] * t +
               T(5) * third * c[kPdInsP53] * t2 + T(2) * c[kPdInsP2] * f;
        const double t0 = (double)c[kT0] + (double)c[kT0Lo];
        dphi += (T)(-5.0 / 3.0 * t0 * p.m53 / (double)f);
    } else if (mrd) {
        const T wy = c[kPdAtanW] * (f - c[kPdAtanShift]);
        dphi = c[kPdMrdP1] - c[kPdMrdM1] * m2 + T(0.75) * c[kPdMrdP34] * f34 * m1 +
               c[kPdAtanA] * c[kPdAtanW] / (T(1) + wy * wy);
    } else {
        dphi = c[kPdIntP1] - T(3) * c[kPdIntM3] * m2 * m2 + c[kPdIntLog] * m1;
    }
    T damp;
    if (f < c[kPdFAmpIns]) {
        // A = f^(-7/6) sum_k c_k t^k: A' = f^(-13/6) sum_k (k/3 - 7/6) c_k t^k
        T poly = T(9.0 / 3.0 - 7.0 / 6.0) * c[kPdAmpIns + 9];
        for (int k = 8; k >= 0; --k) poly = poly * t + T(k / 3.0 - 7.0 / 6.0) * c[kPdAmpIns + k];
        damp = p.m76 * m1 * poly;
    } else if (f >= c[kPdFAmpMrd]) {
        const T u = c[kPdAmpRdS] * (f - c[kPdAmpRdShift]);
        damp = amp * (T(-7.0 / 6.0) * m1 - c[kPdAmpRdGr] * c[kPdAmpRdS] - T(2) * u * c[kPdAmpRdS] / (u * u + T(1)));
    } else {
        // A = f^(-7/6) sum_j d_j y^j, y = S f
        const T y = c[kPdAmpIntS] * f;
        T dpoly = T(4) * c[kPdAmpInt + 4];
        for (int j = 3; j >= 1; --j) dpoly = dpoly * y + T(j) * c[kPdAmpInt + j];
        damp = T(-7.0 / 6.0) * m1 * amp + p.m76 * c[kPdAmpIntS] * dpoly;
    }
    return g_amp * damp + g_phi * dphi;
}

}  // namespace wfx
