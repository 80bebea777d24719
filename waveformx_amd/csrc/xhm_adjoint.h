// The adjoint of IMRPhenomXHM's split evaluation (xhm_point.h) with respect
// to a higher mode's record: given the cotangent g = dL/dRe h + i dL/dIm h
// of the mode h_{l,-m} at one bin, add dL/d(record slot) of that bin to the
// mode's accumulators (HmAcc: the slots that carry a derivative, numbered
// densely so that a wave keeps them in registers over its bins). The
// forward values are recomputed with the forward kernel's operations; the
// adjoint follows each region's formula back to its coefficients and to the
// powers of M f (through the binary's M slots), the (3,2)'s mixing back to
// the (2,2) mode it recycles (returned as that mode's cotangent). The
// region bounds carry no derivative, as the reference's jnp.where
// selections (xhm/reference.py:620-643).
//
// The reference differentiates this evaluation with jax's reverse mode
// through its kernels (src/waveformx/torch.py:108-119; the MLX port's
// hand-written cotangent_product, src/waveformx/models/xhm/mlx/_kernels.py:
// 300-441); here each mode's bins are reduced into its slots' totals per
// wave, and the totals contracted once with the records' derivatives.
#pragma once

#include "xas_adjoint.h"
#include "xhm_point.h"

namespace wfx {

// the accumulators of a higher mode's adjoint, one per record slot with a
// derivative (kHmAccSlot maps them to the slots)
enum HmAcc : int {
    aAmp0 = 0, aMsec = 1, aFRing = 2, aFDamp = 3, aFalloffAt = 4,
    aPnRe = 5,          // 7
    aPnIm = 12,         // 7
    aPnGlobal = 19, aAmpNorm = 20, aPnDominant = 21,
    aIns = 22,          // 3
    aRd = 25,           // 5
    aRdAux = 30,        // 4
    aInterC = 34, aInterH = 35,
    aInter = 36,        // 8
    aPhi = 44,          // 12: phi_1 .. phi_12
    aPhiL5 = 56, aPhiL6 = 57, aPhiL8 = 58,
    aAlpha0 = 59, aAlpha2 = 60, aAlphaL = 61, aC0 = 62, aC1 = 63, aC2 = 64, aC3 = 65, aC4 = 66, aCL = 67,
    aC1Insp = 68, aC1Rd = 69, aCRd = 70, aDeltaPhi = 71,
    aSph = 72,          // 5
    aMix = 77,          // 4
    aT0 = 81, aKIns = 82, aLogM = 83, aCbrtM = 84, aInvM = 85,
    kHmAcc = 86
};

WFX_CONST int kHmAccSlot[kHmAcc] = {
    kHmAmp0, kHmMsec, kHmFRing, kHmFDamp, kHmFalloffAt,
    kHmPnRe, kHmPnRe + 1, kHmPnRe + 2, kHmPnRe + 3, kHmPnRe + 4, kHmPnRe + 5, kHmPnRe + 6,
    kHmPnIm, kHmPnIm + 1, kHmPnIm + 2, kHmPnIm + 3, kHmPnIm + 4, kHmPnIm + 5, kHmPnIm + 6,
    kHmPnGlobal, kHmAmpNorm, kHmPnDominant,
    kHmIns, kHmIns + 1, kHmIns + 2,
    kHmRd, kHmRd + 1, kHmRd + 2, kHmRd + 3, kHmRd + 4,
    kHmRdAux, kHmRdAux + 1, kHmRdAux + 2, kHmRdAux + 3,
    kHmInterC, kHmInterH,
    kHmInter, kHmInter + 1, kHmInter + 2, kHmInter + 3, kHmInter + 4, kHmInter + 5, kHmInter + 6, kHmInter + 7,
    kHmPhi + 1, kHmPhi + 2, kHmPhi + 3, kHmPhi + 4, kHmPhi + 5, kHmPhi + 6, kHmPhi + 7, kHmPhi + 8, kHmPhi + 9,
    kHmPhi + 10, kHmPhi + 11, kHmPhi + 12,
    kHmPhiL + 5, kHmPhiL + 6, kHmPhiL + 8,
    kHmAlpha0, kHmAlpha2, kHmAlphaL, kHmC0, kHmC1, kHmC2, kHmC3, kHmC4, kHmCL,
    kHmC1Insp, kHmC1Rd, kHmCRd, kHmDeltaPhi,
    kHmSph, kHmSph + 1, kHmSph + 2, kHmSph + 3, kHmSph + 4,
    kHmMix, kHmMix + 1, kHmMix + 2, kHmMix + 3,
    kHmT0F64, kHmKInsF64, kHmLogM, kHmCbrtM, kHmInvM,
};

// the higher mode (index into the four records) whose ringdown recycles the
// (2,2): the (3,2), the only record with kHmMixing set
constexpr int kHmMixMode = 2;

// the cotangents of one bin's powers of M f: of M f, t = (M f)^(1/3),
// (M f)^-1, log(M f) and (M f)^(-7/6)
template <class T>
struct MfGrad {
    T mf, t, m1, lg, m76;
};

// where the helpers below add: the accumulator indices of the slots they
// touch (HmIx: the full set of HmAcc; the ringdown kernels' compact sets
// define their own)
struct HmIx {
    static constexpr int FRing = aFRing, FDamp = aFDamp, Rd = aRd, FalloffAt = aFalloffAt;
};

// z = (M f - f_ring) / f_damp's adjoint (the arctangents of the
// intermediate and ringdown phases), for its cotangent gz
template <class Ix = HmIx, class T>
__device__ __forceinline__ void atan_z_adjoint(T gz, T z, T inv_damp, T* acc, MfGrad<T>& g) {
    g.mf += gz * inv_damp;
    acc[Ix::FRing] -= gz * inv_damp;
    acc[Ix::FDamp] -= gz * z * inv_damp;
}

// a Lorentzian ringdown amplitude rd0 / (exp(rd1 d) (d^2 + rd2)), d = M f -
// f_ring (its value a), for its cotangent ga
template <class Ix = HmIx, class T>
__device__ __forceinline__ void lorentzian_adjoint(const T* __restrict__ r, T mf, T a, T ga, T* acc, MfGrad<T>& g) {
    const T dfr = mf - r[kHmFRing];
    const T e = d_exp(r[kHmRd + 1] * dfr);
    const T q = d_fma(dfr, dfr, r[kHmRd + 2]);
    const T inv_q = d_rcp(q);
    acc[Ix::Rd] += ga * d_rcp(e * q);
    acc[Ix::Rd + 1] -= ga * a * dfr;
    acc[Ix::Rd + 2] -= ga * a * inv_q;
    const T gd = -ga * a * d_fma(T(2) * dfr, inv_q, r[kHmRd + 1]);
    g.mf += gd;
    acc[Ix::FRing] -= gd;
}

// the fall-off rd3 exp(-rd4 (M f - f_at)) (its value a), cotangent ga
template <class Ix = HmIx, class T>
__device__ __forceinline__ void falloff_adjoint(const T* __restrict__ r, T mf, T a, T ga, T* acc, MfGrad<T>& g) {
    const T x = mf - r[kHmFalloffAt];
    acc[Ix::Rd + 3] += ga * d_exp(-r[kHmRd + 4] * x);
    acc[Ix::Rd + 4] -= ga * a * x;
    const T gx = -ga * a * r[kHmRd + 4];
    g.mf += gx;
    acc[Ix::FalloffAt] -= gx;
}

// A bin in the ringdown of both the amplitude and the phase (most of a
// mode's bins on a linear grid): the region predicate that the general
// adjoint and the ringdown adjoints below split the bins by, with the same
// arithmetic on both sides, so that every bin is taken exactly once
template <class T>
__device__ __forceinline__ bool hm_ringdown_bin(const T* __restrict__ r, T mf) {
    return !(mf < r[kHmFAmpIn]) && mf >= r[kHmFAmpIm] && !(mf < r[kHmFPhaseIn]) && mf >= r[kHmFPhaseIm];
}

// One bin of higher mode r (its kHmOn set) at the powers p of f, in the
// split layout's formula in T (FP32 for the split gradient, FP64 for the
// FP64 one): the mode's value h = amp0 a e^{i phi} (the forward's
// operations), dL/d(slot) added to acc for the cotangent (g_re, g_im) of h,
// the cotangent (gw_re, gw_im) of the mixing's input w = h_22 / |amp0|
// (zero where it does not mix) and, with Grid, gf = dL/df of the bin's
// frequency f (Hz) moving, through M f, its powers and the leading term.
// Inlined, so that every accumulator index is a constant and acc stays in
// registers; an empty volatile statement ends each region's branch, so that
// their last updates are not merged into one store through a selected
// accumulator (which would move acc to scratch memory).
template <bool Grid, class T>
__device__ __forceinline__ void hm_adjoint_split(const T* __restrict__ r, const BinPowers<T>& p, T w_re, T w_im, T g_re,
                                                 T g_im, T* acc, T& h_re, T& h_im, T& gw_re, T& gw_im, T& gf) {
    // the powers of M f (mf_pow)
    const T mf = r[kHmMsec] * p.f;
    const T t = p.t * r[kHmCbrtM];
    const T m1 = p.m1 * r[kHmInvM];
    const T lg = p.lf + r[kHmLogM];
    const T rs = d_rsq(t);
    const T m76 = m1 * rs;
    const bool mixing = r[kHmMixing] != T(0);
    const T f_ring = r[kHmFRing], f_damp = r[kHmFDamp];
    // the mixing's rotation conj(-a322) w + conj(-a323) A_s e^{i phi_s}
    const bool rotating = mixing && (mf >= r[kHmFAmpIm] || mf >= r[kHmFPhaseIm]);
    // atan((M f - f_ring) / f_damp), of the post-inspiral phases and the
    // rotation only
    T z = T(0), atan_z = T(0), inv_damp = T(0), datan_z = T(0);
    if (rotating || !(mf < r[kHmFPhaseIn])) {
        z = d_div(mf - f_ring, f_damp);
        atan_z = d_atan(z);
        inv_damp = d_rcp(f_damp);
        datan_z = d_rcp(d_fma(z, z, T(1)));
    }
    T rot_re = T(0), rot_im = T(0), amplm = T(0), sn_s = T(0), cs_s = T(0);
    if (rotating) {
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
        philm = d_fma(sp[1], atan_z, philm);
        m_sincos(d_reduce(philm), &sn_s, &cs_s);
        const T m2r = r[kHmMix], m2i = r[kHmMix + 1], m3r = r[kHmMix + 2], m3i = r[kHmMix + 3];
        const T b_re = amplm * m3r, b_im = -(m3i * amplm);
        rot_re = d_fma(w_re, m2r, w_im * m2i) + d_fma(b_re, cs_s, -b_im * sn_s);
        rot_im = d_fma(w_im, m2r, -w_re * m2i) + d_fma(b_re, sn_s, b_im * cs_s);
    }

    // the amplitude (hm_amp_phase_split's regions)
    const bool amp_ins = mf < r[kHmFAmpIn];
    const bool amp_rd = !amp_ins && mf >= r[kHmFAmpIm];
    const bool falloff = mf >= r[kHmFalloff];
    const int n_inter = (int)r[kHmNInter];
    T a;
    T pn_re = T(0), pn_im = T(0), mag = T(0), pseudo = T(0), ins_poly = T(0), t7 = T(0), u = T(0), inter = T(0);
    if (amp_ins) {
        pn_re = r[kHmPnRe + 6];
        pn_im = r[kHmPnIm + 6];
        for (int k = 5; k >= 0; --k) {
            pn_re = d_fma(pn_re, t, r[kHmPnRe + k]);
            pn_im = d_fma(pn_im, t, r[kHmPnIm + k]);
        }
        const T t2 = t * t;
        t7 = t2 * t2 * t2 * t;
        ins_poly = d_fma(d_fma(r[kHmIns + 2], t, r[kHmIns + 1]), t, r[kHmIns]);
        pseudo = t7 * ins_poly;
        mag = d_sqrt(d_fma(pn_re, pn_re, pn_im * pn_im));
        a = m76 * d_fma(mag, r[kHmPnGlobal] * r[kHmAmpNorm], pseudo * r[kHmPnDominant]);
    } else if (amp_rd) {
        if (mixing) {
            a = d_sqrt(d_fma(rot_re, rot_re, rot_im * rot_im));
        } else if (falloff) {
            a = r[kHmRd + 3] * d_exp(-r[kHmRd + 4] * (mf - r[kHmFalloffAt]));
        } else {
            const T dfr = mf - f_ring;
            a = d_div(r[kHmRd], d_exp(r[kHmRd + 1] * dfr) * d_fma(dfr, dfr, r[kHmRd + 2]));
        }
    } else {
        u = (mf - r[kHmInterC]) * r[kHmInterH];
        // the forward's Horner sum from coefficient n - 1 down
#pragma unroll
        for (int j = 7; j >= 0; --j)
            if (j < n_inter) inter = d_fma(inter, u, r[kHmInter + j]);
        a = inter * m76;
    }
    const bool clamped = a < T(0);
    if (clamped) a = T(1.0e-15);

    // the phase
    const bool ph_ins = mf < r[kHmFPhaseIn];
    const bool ph_rd = !ph_ins && mf >= r[kHmFPhaseIm];
    T phi, poly = T(0), inv_t = T(0), m43 = T(0), logs = T(0);
    if (ph_ins) {
        poly = r[kHmPhi + 12];
        for (int i = 11; i >= 1; --i) poly = d_fma(poly, t, r[kHmPhi + i]);
        inv_t = d_div(T(1), t);
        m43 = m1 * inv_t;
        T rest = poly * m43;
        logs = d_fma(d_fma(r[kHmPhiL + 8], t * t, T(0)), t, d_fma(r[kHmPhiL + 6], t, r[kHmPhiL + 5]));
        rest = d_fma(lg, logs, rest);
        rest = d_fma(r[kHmC1Insp], mf, rest);
        const double full = fma(wide_slot(r, kHmT0F64), p.m53, wide_slot(r, kHmKInsF64));
        phi = lead_turned(T(0), full) + rest;
    } else if (ph_rd) {
        T rest;
        if (mixing) {
            rest = d_fma(r[kHmC1Rd], mf, d_atan2(rot_im, rot_re));
        } else {
            rest = d_fma(r[kHmC1Rd] + r[kHmAlpha0], mf, -(f_ring * f_ring * r[kHmAlpha2]) * m1);
            rest = d_fma(r[kHmAlphaL], atan_z, rest);
        }
        phi = rest + r[kHmCRd];
    } else {
        T rest = d_fma(r[kHmC0], mf, r[kHmC1] * lg);
        rest = d_fma(-r[kHmC2], m1, rest);
        const T m12 = m1 * m1;
        rest = d_fma(-(T(1) / T(3)) * r[kHmC4], m12 * m1, rest);
        if (mixing) rest = d_fma(-T(0.5) * r[kHmC3], m12, rest);
        rest = d_fma(r[kHmCL], atan_z, rest);
        phi = rest + r[kHmDeltaPhi];
    }
    phi = d_reduce(phi);
    T sn, cs;
    m_sincos(phi, &sn, &cs);
    const T amp0 = r[kHmAmp0];
    const T val = amp0 * a;
    h_re = val * cs;
    h_im = val * sn;

    // ---------------------------------------------------------- reverse
    const T g_val = d_fma(g_re, cs, g_im * sn);
    const T g_phi = val * d_fma(g_im, cs, -g_re * sn);
    acc[aAmp0] += g_val * a;
    const T g_a = clamped ? T(0) : g_val * amp0;
    MfGrad<T> g{T(0), T(0), T(0), T(0), T(0)};
    T g_rot_re = T(0), g_rot_im = T(0);
    // where both the amplitude and the phase take the rotation, h = amp0
    // rot e^{i (c1_rd M f + c_rd)} exactly (a = |rot|, phi = arg rot +
    // c1_rd M f + c_rd): rot's cotangent then needs no division by |rot|,
    // which the deep ringdown takes below FP32's smallest normal
    const bool mix_rd = mixing && amp_rd && ph_rd;
    if (mix_rd) {
        T e_s, e_c;
        m_sincos(d_reduce(d_fma(r[kHmC1Rd], mf, r[kHmCRd])), &e_s, &e_c);
        g_rot_re = amp0 * d_fma(g_re, e_c, g_im * e_s);
        g_rot_im = amp0 * d_fma(g_im, e_c, -g_re * e_s);
    }

    // the phase's coefficients
    if (ph_ins) {
        acc[aT0] += g_phi * (T)p.m53;
        acc[aKIns] += g_phi;
        acc[aC1Insp] += g_phi * mf;
        g.mf += g_phi * r[kHmC1Insp];
        g.lg += g_phi * logs;
        const T g_logs = g_phi * lg;
        const T t2 = t * t;
        acc[aPhiL5] += g_logs;
        acc[aPhiL6] += g_logs * t;
        acc[aPhiL8] += g_logs * t2 * t;
        g.t += g_logs * d_fma(T(3) * r[kHmPhiL + 8], t2, r[kHmPhiL + 6]);
        const T g_poly = g_phi * m43;
        const T g_m43 = g_phi * poly;
        g.m1 += g_m43 * inv_t;
        g.t -= g_m43 * m43 * inv_t;
        T pw = T(1), dpoly = T(0);
#pragma unroll
        for (int j = 0; j < 12; ++j) {
            acc[aPhi + j] += g_poly * pw;
            if (j < 11) dpoly = d_fma((T)(j + 1) * r[kHmPhi + j + 2], pw, dpoly);
            pw *= t;
        }
        g.t += g_poly * dpoly;
        __asm__ volatile("");
    } else if (ph_rd) {
        acc[aC1Rd] += g_phi * mf;
        acc[aCRd] += g_phi;
        if (mixing) {
            g.mf += g_phi * r[kHmC1Rd];
            if (!mix_rd) {
                // d atan2(y, x) = (x dy - y dx) / (x^2 + y^2), scaled by
                // max(|x|, |y|) first; zero below FP32's smallest normal
                const T s = fmax(fabs(rot_re), fabs(rot_im));
                const bool normal = s >= tiny_normal(T(0));
                const T inv_s = normal ? d_rcp(s) : T(0);
                const T xs = rot_re * inv_s, ys = rot_im * inv_s;
                const T rr = normal ? d_rcp(s * d_fma(xs, xs, ys * ys)) : T(0);
                g_rot_re -= g_phi * ys * rr;
                g_rot_im += g_phi * xs * rr;
            }
        } else {
            acc[aAlpha0] += g_phi * mf;
            g.mf += g_phi * (r[kHmC1Rd] + r[kHmAlpha0]);
            const T a2 = r[kHmAlpha2];
            acc[aAlpha2] -= g_phi * f_ring * f_ring * m1;
            acc[aFRing] -= g_phi * T(2) * f_ring * a2 * m1;
            g.m1 -= g_phi * f_ring * f_ring * a2;
            acc[aAlphaL] += g_phi * atan_z;
            atan_z_adjoint(g_phi * r[kHmAlphaL] * datan_z, z, inv_damp, acc, g);
        }
        __asm__ volatile("");
    } else {
        acc[aC0] += g_phi * mf;
        g.mf += g_phi * r[kHmC0];
        acc[aC1] += g_phi * lg;
        g.lg += g_phi * r[kHmC1];
        acc[aC2] -= g_phi * m1;
        g.m1 -= g_phi * r[kHmC2];
        const T m12 = m1 * m1;
        acc[aC4] -= g_phi * m12 * m1 * (T(1) / T(3));
        g.m1 -= g_phi * r[kHmC4] * m12;
        if (mixing) {
            acc[aC3] -= T(0.5) * g_phi * m12;
            g.m1 -= g_phi * r[kHmC3] * m1;
        }
        acc[aCL] += g_phi * atan_z;
        atan_z_adjoint(g_phi * r[kHmCL] * datan_z, z, inv_damp, acc, g);
        acc[aDeltaPhi] += g_phi;
    }

    // the amplitude's coefficients
    if (amp_ins) {
        const T glob = r[kHmPnGlobal] * r[kHmAmpNorm], dom = r[kHmPnDominant];
        g.m76 += g_a * d_fma(mag, glob, pseudo * dom);
        const T g_in = g_a * m76;
        acc[aPnGlobal] += g_in * mag * r[kHmAmpNorm];
        acc[aAmpNorm] += g_in * mag * r[kHmPnGlobal];
        acc[aPnDominant] += g_in * pseudo;
        const T g_mag = g_in * glob;
        const T inv_mag = mag > T(0) ? d_rcp(mag) : T(0);
        const T g_pre = g_mag * pn_re * inv_mag, g_pim = g_mag * pn_im * inv_mag;
        T pw = T(1), d_re = T(0), d_im = T(0);
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            acc[aPnRe + k] += g_pre * pw;
            acc[aPnIm + k] += g_pim * pw;
            if (k < 6) {
                d_re = d_fma((T)(k + 1) * r[kHmPnRe + k + 1], pw, d_re);
                d_im = d_fma((T)(k + 1) * r[kHmPnIm + k + 1], pw, d_im);
            }
            pw *= t;
        }
        g.t += d_fma(g_pre, d_re, g_pim * d_im);
        const T g_ps = g_in * dom;
        acc[aIns] += g_ps * t7;
        acc[aIns + 1] += g_ps * t7 * t;
        acc[aIns + 2] += g_ps * t7 * t * t;
        const T t6 = t7 * d_div(T(1), t);
        g.t += g_ps * d_fma(T(7) * t6, ins_poly, t7 * d_fma(T(2) * r[kHmIns + 2], t, r[kHmIns + 1]));
        __asm__ volatile("");
    } else if (amp_rd) {
        if (mixing) {
            if (!mix_rd) {
                const T inv_a = a >= tiny_normal(T(0)) ? d_rcp(a) : T(0);
                g_rot_re += g_a * rot_re * inv_a;
                g_rot_im += g_a * rot_im * inv_a;
            }
        } else if (falloff) {
            falloff_adjoint(r, mf, a, g_a, acc, g);
        } else {
            lorentzian_adjoint(r, mf, a, g_a, acc, g);
        }
        __asm__ volatile("");
    } else {
        g.m76 += g_a * inter;
        const T g_res = g_a * m76;
        T pw = T(1), d_u = T(0);
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if (j < n_inter) acc[aInter + j] += g_res * pw;
            if (j + 1 < n_inter) d_u = d_fma((T)(j + 1) * r[kHmInter + j + 1], pw, d_u);
            pw *= u;
        }
        const T g_u = g_res * d_u;
        acc[aInterH] += g_u * (mf - r[kHmInterC]);
        const T g_x = g_u * r[kHmInterH];
        g.mf += g_x;
        acc[aInterC] -= g_x;
    }

    // the mixing's rotation, back to w, the spheroidal phase and amplitude
    gw_re = T(0);
    gw_im = T(0);
    if (rotating) {
        const T m2r = r[kHmMix], m2i = r[kHmMix + 1], m3r = r[kHmMix + 2], m3i = r[kHmMix + 3];
        acc[aMix] += d_fma(g_rot_re, w_re, g_rot_im * w_im);
        acc[aMix + 1] += d_fma(g_rot_re, w_im, -g_rot_im * w_re);
        gw_re = d_fma(g_rot_re, m2r, -g_rot_im * m2i);
        gw_im = d_fma(g_rot_re, m2i, g_rot_im * m2r);
        const T b_re = amplm * m3r, b_im = -(m3i * amplm);
        const T g_bre = d_fma(g_rot_re, cs_s, g_rot_im * sn_s);
        const T g_bim = d_fma(-g_rot_re, sn_s, g_rot_im * cs_s);
        const T g_phl = g_rot_re * (-b_re * sn_s - b_im * cs_s) + g_rot_im * d_fma(b_re, cs_s, -b_im * sn_s);
        acc[aMix + 2] += g_bre * amplm;
        acc[aMix + 3] -= g_bim * amplm;
        const T g_aml = d_fma(g_bre, m3r, -g_bim * m3i);
        const T* sp = r + kHmSph;
        acc[aSph] += g_phl * mf;
        g.mf += g_phl * sp[0];
        acc[aSph + 4] += g_phl;
        acc[aSph + 2] -= g_phl * m1;
        g.m1 -= g_phl * sp[2];
        acc[aSph + 3] -= g_phl * m1 * m1 * m1 * (T(1) / T(3));
        g.m1 -= g_phl * sp[3] * m1 * m1;
        acc[aSph + 1] += g_phl * atan_z;
        atan_z_adjoint(g_phl * sp[1] * datan_z, z, inv_damp, acc, g);
        if (mf < r[kHmFRdAux]) {
            const T* x = r + kHmRdAux;
            acc[aRdAux] += g_aml;
            acc[aRdAux + 1] += g_aml * mf;
            acc[aRdAux + 2] += g_aml * mf * mf;
            acc[aRdAux + 3] += g_aml * mf * mf * mf;
            g.mf += g_aml * d_fma(d_fma(T(3) * x[3], mf, T(2) * x[2]), mf, x[1]);
        } else if (mf >= r[kHmFalloff]) {
            falloff_adjoint(r, mf, amplm, g_aml, acc, g);
        } else {
            lorentzian_adjoint(r, mf, amplm, g_aml, acc, g);
        }
    }

    // the powers of M f, back to the binary's M slots
    g.m1 += g.m76 * rs;
    g.t -= T(0.5) * g.m76 * m76 * d_div(T(1), t);
    acc[aMsec] += g.mf * p.f;
    acc[aCbrtM] += g.t * p.t;
    acc[aInvM] += g.m1 * p.m1;
    acc[aLogM] += g.lg;
    if (Grid) {
        // f moving: d(M f)/df = M, dt/df = t / (3 f), d(1/(M f))/df = -m1 / f,
        // dlog(M f)/df = 1 / f, and the leading term T0 f^(-5/3)'s
        // -5/3 T0 f^(-5/3) / f (its cotangent the phase's)
        T lead = T(0);
        if (ph_ins) lead = (T)(-5.0 / 3.0 * wide_slot(r, kHmT0F64) * p.m53) * g_phi;
        gf = d_fma(g.mf, r[kHmMsec], (d_fma(g.t, t * (T(1) / T(3)), g.lg - g.m1 * m1) + lead) * p.m1);
    }
}

// ----------------------------------------------------- the intermediate

// A mode's bins in the intermediate region of both its amplitude and its
// phase (most of a mode's bins before its ringdown on a linear grid): the
// amplitude's polynomial in u and the phase's closed form, through M f,
// log(M f), (M f)^-1 and (M f)^(-7/6); 24 accumulators. Such a bin is below
// both ringdown bounds, so the mixing mode does not rotate there: its
// phase only adds the c3 term.
enum InterAcc : int {
    iAmp0 = 0, iMsec = 1, iCbrtM = 2, iInvM = 3, iLogM = 4, iFRing = 5, iFDamp = 6, iInterC = 7, iInterH = 8,
    iInter = 9,  // 8
    iC0 = 17, iC1 = 18, iC2 = 19, iC4 = 20, iCL = 21, iDeltaPhi = 22, iC3 = 23,
    kInterAcc = 24
};
WFX_CONST int kInterAccSlot[kInterAcc] = {kHmAmp0, kHmMsec, kHmCbrtM, kHmInvM, kHmLogM, kHmFRing, kHmFDamp,
                                          kHmInterC, kHmInterH,
                                          kHmInter, kHmInter + 1, kHmInter + 2, kHmInter + 3, kHmInter + 4,
                                          kHmInter + 5, kHmInter + 6, kHmInter + 7,
                                          kHmC0, kHmC1, kHmC2, kHmC4, kHmCL, kHmDeltaPhi, kHmC3};
struct InterIx {
    static constexpr int FRing = iFRing, FDamp = iFDamp, Rd = -1, FalloffAt = -1;
};

// the predicate of those bins (the regions' tests of hm_adjoint_split)
template <class T>
__device__ __forceinline__ bool hm_intermediate_bin(const T* __restrict__ r, T mf) {
    return !(mf < r[kHmFAmpIn]) && !(mf >= r[kHmFAmpIm]) && !(mf < r[kHmFPhaseIn]) && !(mf >= r[kHmFPhaseIm]);
}

// hm_adjoint_split for a mode's intermediate bin (hm_intermediate_bin):
// the same operations, in the same order, on the 24 compact accumulators
template <bool Grid, class T>
__device__ __forceinline__ void hm_adjoint_intermediate(const T* __restrict__ r, const BinPowers<T>& p, T g_re,
                                                        T g_im, T* acc, T& h_re, T& h_im, T& gf) {
    const T mf = r[kHmMsec] * p.f;
    const T t = p.t * r[kHmCbrtM];
    const T m1 = p.m1 * r[kHmInvM];
    const T lg = p.lf + r[kHmLogM];
    const T rs = d_rsq(t);
    const T m76 = m1 * rs;
    const T f_ring = r[kHmFRing], f_damp = r[kHmFDamp];
    const T z = d_div(mf - f_ring, f_damp);
    const T atan_z = d_atan(z);
    const T inv_damp = d_rcp(f_damp);
    const T datan_z = d_rcp(d_fma(z, z, T(1)));
    const int n_inter = (int)r[kHmNInter];
    const T u = (mf - r[kHmInterC]) * r[kHmInterH];
    T inter = T(0);
#pragma unroll
    for (int j = 7; j >= 0; --j)
        if (j < n_inter) inter = d_fma(inter, u, r[kHmInter + j]);
    T a = inter * m76;
    const bool clamped = a < T(0);
    if (clamped) a = T(1.0e-15);
    const bool mixing = r[kHmMixing] != T(0);
    T rest = d_fma(r[kHmC0], mf, r[kHmC1] * lg);
    rest = d_fma(-r[kHmC2], m1, rest);
    const T m12 = m1 * m1;
    rest = d_fma(-(T(1) / T(3)) * r[kHmC4], m12 * m1, rest);
    if (mixing) rest = d_fma(-T(0.5) * r[kHmC3], m12, rest);
    rest = d_fma(r[kHmCL], atan_z, rest);
    const T phi = d_reduce(rest + r[kHmDeltaPhi]);
    T sn, cs;
    m_sincos(phi, &sn, &cs);
    const T amp0 = r[kHmAmp0];
    const T val = amp0 * a;
    h_re = val * cs;
    h_im = val * sn;

    const T g_val = d_fma(g_re, cs, g_im * sn);
    const T g_phi = val * d_fma(g_im, cs, -g_re * sn);
    acc[iAmp0] += g_val * a;
    const T g_a = clamped ? T(0) : g_val * amp0;
    MfGrad<T> g{T(0), T(0), T(0), T(0), T(0)};
    acc[iC0] += g_phi * mf;
    g.mf += g_phi * r[kHmC0];
    acc[iC1] += g_phi * lg;
    g.lg += g_phi * r[kHmC1];
    acc[iC2] -= g_phi * m1;
    g.m1 -= g_phi * r[kHmC2];
    acc[iC4] -= g_phi * m12 * m1 * (T(1) / T(3));
    g.m1 -= g_phi * r[kHmC4] * m12;
    if (mixing) {
        acc[iC3] -= T(0.5) * g_phi * m12;
        g.m1 -= g_phi * r[kHmC3] * m1;
    }
    acc[iCL] += g_phi * atan_z;
    atan_z_adjoint<InterIx>(g_phi * r[kHmCL] * datan_z, z, inv_damp, acc, g);
    acc[iDeltaPhi] += g_phi;
    g.m76 += g_a * inter;
    const T g_res = g_a * m76;
    T pw = T(1), d_u = T(0);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (j < n_inter) acc[iInter + j] += g_res * pw;
        if (j + 1 < n_inter) d_u = d_fma((T)(j + 1) * r[kHmInter + j + 1], pw, d_u);
        pw *= u;
    }
    const T g_u = g_res * d_u;
    acc[iInterH] += g_u * (mf - r[kHmInterC]);
    const T g_x = g_u * r[kHmInterH];
    g.mf += g_x;
    acc[iInterC] -= g_x;
    g.m1 += g.m76 * rs;
    g.t -= T(0.5) * g.m76 * m76 * d_div(T(1), t);
    acc[iMsec] += g.mf * p.f;
    acc[iCbrtM] += g.t * p.t;
    acc[iInvM] += g.m1 * p.m1;
    acc[iLogM] += g.lg;
    if (Grid) {
        const T lead = T(0);
        gf = d_fma(g.mf, r[kHmMsec], (d_fma(g.t, t * (T(1) / T(3)), g.lg - g.m1 * m1) + lead) * p.m1);
    }
}

// ---------------------------------------------------------- the inspiral

// A mode's bins in the inspiral of both its amplitude and its phase (below
// both inspiral bounds, and for the mixing mode below both ringdown bounds,
// so that it does not rotate): the PN and pseudo-PN amplitude and the phase's
// polynomial in (M f)^(1/3) with its logarithms and leading term; 43
// accumulators.
enum InsAcc : int {
    nAmp0 = 0, nMsec = 1, nCbrtM = 2, nInvM = 3, nLogM = 4,
    nT0 = 5, nKIns = 6, nC1Insp = 7, nPhiL5 = 8, nPhiL6 = 9, nPhiL8 = 10,
    nPhi = 11,         // 12: phi_1 .. phi_12
    nPnGlobal = 23, nAmpNorm = 24, nPnDominant = 25,
    nPnRe = 26,        // 7
    nPnIm = 33,        // 7
    nIns = 40,         // 3
    kInsAcc = 43
};
WFX_CONST int kInsAccSlot[kInsAcc] = {
    kHmAmp0, kHmMsec, kHmCbrtM, kHmInvM, kHmLogM,
    kHmT0F64, kHmKInsF64, kHmC1Insp, kHmPhiL + 5, kHmPhiL + 6, kHmPhiL + 8,
    kHmPhi + 1, kHmPhi + 2, kHmPhi + 3, kHmPhi + 4, kHmPhi + 5, kHmPhi + 6, kHmPhi + 7, kHmPhi + 8, kHmPhi + 9,
    kHmPhi + 10, kHmPhi + 11, kHmPhi + 12,
    kHmPnGlobal, kHmAmpNorm, kHmPnDominant,
    kHmPnRe, kHmPnRe + 1, kHmPnRe + 2, kHmPnRe + 3, kHmPnRe + 4, kHmPnRe + 5, kHmPnRe + 6,
    kHmPnIm, kHmPnIm + 1, kHmPnIm + 2, kHmPnIm + 3, kHmPnIm + 4, kHmPnIm + 5, kHmPnIm + 6,
    kHmIns, kHmIns + 1, kHmIns + 2,
};

// the predicate of those bins (the regions' tests of hm_adjoint_split)
template <class T>
__device__ __forceinline__ bool hm_inspiral_bin(const T* __restrict__ r, T mf) {
    const bool rotating = r[kHmMixing] != T(0) && (mf >= r[kHmFAmpIm] || mf >= r[kHmFPhaseIm]);
    return mf < r[kHmFAmpIn] && mf < r[kHmFPhaseIn] && !rotating;
}

// hm_adjoint_split for a mode's inspiral bin (hm_inspiral_bin): the same
// operations, in the same order, on the 43 compact accumulators
template <bool Grid, class T>
__device__ __forceinline__ void hm_adjoint_inspiral(const T* __restrict__ r, const BinPowers<T>& p, T g_re, T g_im,
                                                    T* acc, T& h_re, T& h_im, T& gf) {
    const T mf = r[kHmMsec] * p.f;
    const T t = p.t * r[kHmCbrtM];
    const T m1 = p.m1 * r[kHmInvM];
    const T lg = p.lf + r[kHmLogM];
    const T rs = d_rsq(t);
    const T m76 = m1 * rs;
    T pn_re = r[kHmPnRe + 6], pn_im = r[kHmPnIm + 6];
#pragma unroll
    for (int k = 5; k >= 0; --k) {
        pn_re = d_fma(pn_re, t, r[kHmPnRe + k]);
        pn_im = d_fma(pn_im, t, r[kHmPnIm + k]);
    }
    const T t2 = t * t;
    const T t7 = t2 * t2 * t2 * t;
    const T ins_poly = d_fma(d_fma(r[kHmIns + 2], t, r[kHmIns + 1]), t, r[kHmIns]);
    const T pseudo = t7 * ins_poly;
    const T mag = d_sqrt(d_fma(pn_re, pn_re, pn_im * pn_im));
    T a = m76 * d_fma(mag, r[kHmPnGlobal] * r[kHmAmpNorm], pseudo * r[kHmPnDominant]);
    const bool clamped = a < T(0);
    if (clamped) a = T(1.0e-15);
    T poly = r[kHmPhi + 12];
#pragma unroll
    for (int i = 11; i >= 1; --i) poly = d_fma(poly, t, r[kHmPhi + i]);
    const T inv_t = d_div(T(1), t);
    const T m43 = m1 * inv_t;
    T rest = poly * m43;
    const T logs = d_fma(d_fma(r[kHmPhiL + 8], t * t, T(0)), t, d_fma(r[kHmPhiL + 6], t, r[kHmPhiL + 5]));
    rest = d_fma(lg, logs, rest);
    rest = d_fma(r[kHmC1Insp], mf, rest);
    const double full = fma(wide_slot(r, kHmT0F64), p.m53, wide_slot(r, kHmKInsF64));
    const T phi = d_reduce(lead_turned(T(0), full) + rest);
    T sn, cs;
    m_sincos(phi, &sn, &cs);
    const T amp0 = r[kHmAmp0];
    const T val = amp0 * a;
    h_re = val * cs;
    h_im = val * sn;

    const T g_val = d_fma(g_re, cs, g_im * sn);
    const T g_phi = val * d_fma(g_im, cs, -g_re * sn);
    acc[nAmp0] += g_val * a;
    const T g_a = clamped ? T(0) : g_val * amp0;
    MfGrad<T> g{T(0), T(0), T(0), T(0), T(0)};
    acc[nT0] += g_phi * (T)p.m53;
    acc[nKIns] += g_phi;
    acc[nC1Insp] += g_phi * mf;
    g.mf += g_phi * r[kHmC1Insp];
    g.lg += g_phi * logs;
    const T g_logs = g_phi * lg;
    acc[nPhiL5] += g_logs;
    acc[nPhiL6] += g_logs * t;
    acc[nPhiL8] += g_logs * t2 * t;
    g.t += g_logs * d_fma(T(3) * r[kHmPhiL + 8], t2, r[kHmPhiL + 6]);
    const T g_poly = g_phi * m43;
    const T g_m43 = g_phi * poly;
    g.m1 += g_m43 * inv_t;
    g.t -= g_m43 * m43 * inv_t;
    T pw = T(1), dpoly = T(0);
#pragma unroll
    for (int j = 0; j < 12; ++j) {
        acc[nPhi + j] += g_poly * pw;
        if (j < 11) dpoly = d_fma((T)(j + 1) * r[kHmPhi + j + 2], pw, dpoly);
        pw *= t;
    }
    g.t += g_poly * dpoly;
    const T glob = r[kHmPnGlobal] * r[kHmAmpNorm], dom = r[kHmPnDominant];
    g.m76 += g_a * d_fma(mag, glob, pseudo * dom);
    const T g_in = g_a * m76;
    acc[nPnGlobal] += g_in * mag * r[kHmAmpNorm];
    acc[nAmpNorm] += g_in * mag * r[kHmPnGlobal];
    acc[nPnDominant] += g_in * pseudo;
    const T g_mag = g_in * glob;
    const T inv_mag = mag > T(0) ? d_rcp(mag) : T(0);
    const T g_pre = g_mag * pn_re * inv_mag, g_pim = g_mag * pn_im * inv_mag;
    T qw = T(1), d_re = T(0), d_im = T(0);
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        acc[nPnRe + k] += g_pre * qw;
        acc[nPnIm + k] += g_pim * qw;
        if (k < 6) {
            d_re = d_fma((T)(k + 1) * r[kHmPnRe + k + 1], qw, d_re);
            d_im = d_fma((T)(k + 1) * r[kHmPnIm + k + 1], qw, d_im);
        }
        qw *= t;
    }
    g.t += d_fma(g_pre, d_re, g_pim * d_im);
    const T g_ps = g_in * dom;
    acc[nIns] += g_ps * t7;
    acc[nIns + 1] += g_ps * t7 * t;
    acc[nIns + 2] += g_ps * t7 * t * t;
    const T t6 = t7 * d_div(T(1), t);
    g.t += g_ps * d_fma(T(7) * t6, ins_poly, t7 * d_fma(T(2) * r[kHmIns + 2], t, r[kHmIns + 1]));
    g.m1 += g.m76 * rs;
    g.t -= T(0.5) * g.m76 * m76 * d_div(T(1), t);
    acc[nMsec] += g.mf * p.f;
    acc[nCbrtM] += g.t * p.t;
    acc[nInvM] += g.m1 * p.m1;
    acc[nLogM] += g.lg;
    if (Grid) {
        const T lead = (T)(-5.0 / 3.0 * wide_slot(r, kHmT0F64) * p.m53) * g_phi;
        gf = d_fma(g.mf, r[kHmMsec], (d_fma(g.t, t * (T(1) / T(3)), g.lg - g.m1 * m1) + lead) * p.m1);
    }
}

// --------------------------------------------------------- the ringdown

// The ringdown adjoints' compact accumulators. A plain mode's ringdown (the
// Lorentzian or the fall-off, the phase with its arctangent) touches 16
// slots; the mixing mode's (the (3,2): its rotation of the (2,2), amplitude
// |rot| and phase arg rot) 26. M f's powers enter through M f and 1/(M f)
// only there.
enum RdAcc : int {
    rAmp0 = 0, rMsec = 1, rInvM = 2, rFRing = 3, rFDamp = 4, rFalloffAt = 5,
    rRd = 6,  // 5
    rC1Rd = 11, rCRd = 12, rAlpha0 = 13, rAlpha2 = 14, rAlphaL = 15,
    kRdAcc = 16
};
WFX_CONST int kRdAccSlot[kRdAcc] = {kHmAmp0, kHmMsec, kHmInvM, kHmFRing, kHmFDamp, kHmFalloffAt,
                                    kHmRd, kHmRd + 1, kHmRd + 2, kHmRd + 3, kHmRd + 4,
                                    kHmC1Rd, kHmCRd, kHmAlpha0, kHmAlpha2, kHmAlphaL};
struct RdIx {
    static constexpr int FRing = rFRing, FDamp = rFDamp, Rd = rRd, FalloffAt = rFalloffAt;
};
enum MixAcc : int {
    xAmp0 = 0, xMsec = 1, xInvM = 2, xFRing = 3, xFDamp = 4, xFalloffAt = 5,
    xRd = 6,      // 5
    xRdAux = 11,  // 4
    xC1Rd = 15, xCRd = 16,
    xSph = 17,    // 5
    xMix = 22,    // 4
    kMixAcc = 26
};
WFX_CONST int kMixAccSlot[kMixAcc] = {kHmAmp0, kHmMsec, kHmInvM, kHmFRing, kHmFDamp, kHmFalloffAt,
                                      kHmRd, kHmRd + 1, kHmRd + 2, kHmRd + 3, kHmRd + 4,
                                      kHmRdAux, kHmRdAux + 1, kHmRdAux + 2, kHmRdAux + 3,
                                      kHmC1Rd, kHmCRd,
                                      kHmSph, kHmSph + 1, kHmSph + 2, kHmSph + 3, kHmSph + 4,
                                      kHmMix, kHmMix + 1, kHmMix + 2, kHmMix + 3};
struct MixIx {
    static constexpr int FRing = xFRing, FDamp = xFDamp, Rd = xRd, FalloffAt = xFalloffAt;
};

// M f's cotangents back to the record's M slots and, with Grid, to f
template <bool Grid, int Msec, int InvM, class T>
__device__ __forceinline__ void ringdown_mf_adjoint(const T* __restrict__ r, const BinPowers<T>& p, T m1,
                                                    const MfGrad<T>& g, T* acc, T& gf) {
    acc[Msec] += g.mf * p.f;
    acc[InvM] += g.m1 * p.m1;
    if (Grid) gf = d_fma(g.mf, r[kHmMsec], (-(g.m1 * m1)) * p.m1);
}

// hm_adjoint_split for a plain mode's bin in the ringdown of both its
// amplitude and its phase (hm_ringdown_bin), the same operations on the
// 16 compact accumulators
template <bool Grid, class T>
__device__ __forceinline__ void hm_adjoint_ringdown(const T* __restrict__ r, const BinPowers<T>& p, T g_re, T g_im,
                                                    T* acc, T& h_re, T& h_im, T& gf) {
    const T mf = r[kHmMsec] * p.f;
    const T m1 = p.m1 * r[kHmInvM];
    const T f_ring = r[kHmFRing], f_damp = r[kHmFDamp];
    const T z = d_div(mf - f_ring, f_damp);
    const T atan_z = d_atan(z);
    const T inv_damp = d_rcp(f_damp);
    const T datan_z = d_rcp(d_fma(z, z, T(1)));
    const bool falloff = mf >= r[kHmFalloff];
    T a;
    if (falloff) {
        a = r[kHmRd + 3] * d_exp(-r[kHmRd + 4] * (mf - r[kHmFalloffAt]));
    } else {
        const T dfr = mf - f_ring;
        a = d_div(r[kHmRd], d_exp(r[kHmRd + 1] * dfr) * d_fma(dfr, dfr, r[kHmRd + 2]));
    }
    const bool clamped = a < T(0);
    if (clamped) a = T(1.0e-15);
    T rest = d_fma(r[kHmC1Rd] + r[kHmAlpha0], mf, -(f_ring * f_ring * r[kHmAlpha2]) * m1);
    rest = d_fma(r[kHmAlphaL], atan_z, rest);
    const T phi = d_reduce(rest + r[kHmCRd]);
    T sn, cs;
    m_sincos(phi, &sn, &cs);
    const T amp0 = r[kHmAmp0];
    const T val = amp0 * a;
    h_re = val * cs;
    h_im = val * sn;

    const T g_val = d_fma(g_re, cs, g_im * sn);
    const T g_phi = val * d_fma(g_im, cs, -g_re * sn);
    acc[rAmp0] += g_val * a;
    const T g_a = clamped ? T(0) : g_val * amp0;
    MfGrad<T> g{T(0), T(0), T(0), T(0), T(0)};
    acc[rC1Rd] += g_phi * mf;
    acc[rCRd] += g_phi;
    acc[rAlpha0] += g_phi * mf;
    g.mf += g_phi * (r[kHmC1Rd] + r[kHmAlpha0]);
    const T a2 = r[kHmAlpha2];
    acc[rAlpha2] -= g_phi * f_ring * f_ring * m1;
    acc[rFRing] -= g_phi * T(2) * f_ring * a2 * m1;
    g.m1 -= g_phi * f_ring * f_ring * a2;
    acc[rAlphaL] += g_phi * atan_z;
    atan_z_adjoint<RdIx>(g_phi * r[kHmAlphaL] * datan_z, z, inv_damp, acc, g);
    if (falloff)
        falloff_adjoint<RdIx>(r, mf, a, g_a, acc, g);
    else
        lorentzian_adjoint<RdIx>(r, mf, a, g_a, acc, g);
    ringdown_mf_adjoint<Grid, rMsec, rInvM>(r, p, m1, g, acc, gf);
}

// hm_adjoint_split for the mixing mode's bin in the ringdown of both its
// amplitude and its phase: h = amp0 rot e^{i (c1_rd M f + c_rd)}, rot the
// rotation of w = h_22 / |amp0|, on the 26 compact accumulators; (gw_re,
// gw_im) the cotangent of w
template <bool Grid, class T>
__device__ __forceinline__ void hm_adjoint_ringdown_mixing(const T* __restrict__ r, const BinPowers<T>& p, T w_re,
                                                           T w_im, T g_re, T g_im, T* acc, T& h_re, T& h_im,
                                                           T& gw_re, T& gw_im, T& gf) {
    const T mf = r[kHmMsec] * p.f;
    const T m1 = p.m1 * r[kHmInvM];
    const T f_ring = r[kHmFRing], f_damp = r[kHmFDamp];
    const T z = d_div(mf - f_ring, f_damp);
    const T atan_z = d_atan(z);
    const T inv_damp = d_rcp(f_damp);
    const T datan_z = d_rcp(d_fma(z, z, T(1)));
    const bool aux = mf < r[kHmFRdAux];
    const bool falloff = !aux && mf >= r[kHmFalloff];
    T amplm;
    if (aux) {
        const T* x = r + kHmRdAux;
        amplm = d_fma(d_fma(d_fma(x[3], mf, x[2]), mf, x[1]), mf, x[0]);
    } else if (falloff) {
        amplm = r[kHmRd + 3] * d_exp(-r[kHmRd + 4] * (mf - r[kHmFalloffAt]));
    } else {
        const T dfr = mf - f_ring;
        amplm = d_div(r[kHmRd], d_exp(r[kHmRd + 1] * dfr) * d_fma(dfr, dfr, r[kHmRd + 2]));
    }
    const T* sp = r + kHmSph;
    T philm = d_fma(sp[0], mf, sp[4]);
    philm = d_fma(-sp[2], m1, philm);
    philm = d_fma(-(T(1) / T(3)) * sp[3], m1 * m1 * m1, philm);
    philm = d_fma(sp[1], atan_z, philm);
    T sn_s, cs_s;
    m_sincos(d_reduce(philm), &sn_s, &cs_s);
    const T m2r = r[kHmMix], m2i = r[kHmMix + 1], m3r = r[kHmMix + 2], m3i = r[kHmMix + 3];
    const T b_re = amplm * m3r, b_im = -(m3i * amplm);
    const T rot_re = d_fma(w_re, m2r, w_im * m2i) + d_fma(b_re, cs_s, -b_im * sn_s);
    const T rot_im = d_fma(w_im, m2r, -w_re * m2i) + d_fma(b_re, sn_s, b_im * cs_s);
    T a = d_sqrt(d_fma(rot_re, rot_re, rot_im * rot_im));
    if (a < T(0)) a = T(1.0e-15);
    const T phi = d_reduce(d_fma(r[kHmC1Rd], mf, d_atan2(rot_im, rot_re)) + r[kHmCRd]);
    T sn, cs;
    m_sincos(phi, &sn, &cs);
    const T amp0 = r[kHmAmp0];
    const T val = amp0 * a;
    h_re = val * cs;
    h_im = val * sn;

    const T g_val = d_fma(g_re, cs, g_im * sn);
    const T g_phi = val * d_fma(g_im, cs, -g_re * sn);
    acc[xAmp0] += g_val * a;
    MfGrad<T> g{T(0), T(0), T(0), T(0), T(0)};
    // h = amp0 rot e^{i (c1_rd M f + c_rd)} exactly: rot's cotangent needs no
    // division by |rot|
    T e_s, e_c;
    m_sincos(d_reduce(d_fma(r[kHmC1Rd], mf, r[kHmCRd])), &e_s, &e_c);
    const T g_rot_re = amp0 * d_fma(g_re, e_c, g_im * e_s);
    const T g_rot_im = amp0 * d_fma(g_im, e_c, -g_re * e_s);
    acc[xC1Rd] += g_phi * mf;
    acc[xCRd] += g_phi;
    g.mf += g_phi * r[kHmC1Rd];
    // the rotation, back to w, the spheroidal phase and amplitude
    acc[xMix] += d_fma(g_rot_re, w_re, g_rot_im * w_im);
    acc[xMix + 1] += d_fma(g_rot_re, w_im, -g_rot_im * w_re);
    gw_re = d_fma(g_rot_re, m2r, -g_rot_im * m2i);
    gw_im = d_fma(g_rot_re, m2i, g_rot_im * m2r);
    const T g_bre = d_fma(g_rot_re, cs_s, g_rot_im * sn_s);
    const T g_bim = d_fma(-g_rot_re, sn_s, g_rot_im * cs_s);
    const T g_phl = g_rot_re * (-b_re * sn_s - b_im * cs_s) + g_rot_im * d_fma(b_re, cs_s, -b_im * sn_s);
    acc[xMix + 2] += g_bre * amplm;
    acc[xMix + 3] -= g_bim * amplm;
    const T g_aml = d_fma(g_bre, m3r, -g_bim * m3i);
    acc[xSph] += g_phl * mf;
    g.mf += g_phl * sp[0];
    acc[xSph + 4] += g_phl;
    acc[xSph + 2] -= g_phl * m1;
    g.m1 -= g_phl * sp[2];
    acc[xSph + 3] -= g_phl * m1 * m1 * m1 * (T(1) / T(3));
    g.m1 -= g_phl * sp[3] * m1 * m1;
    acc[xSph + 1] += g_phl * atan_z;
    atan_z_adjoint<MixIx>(g_phl * sp[1] * datan_z, z, inv_damp, acc, g);
    if (aux) {
        const T* x = r + kHmRdAux;
        acc[xRdAux] += g_aml;
        acc[xRdAux + 1] += g_aml * mf;
        acc[xRdAux + 2] += g_aml * mf * mf;
        acc[xRdAux + 3] += g_aml * mf * mf * mf;
        g.mf += g_aml * d_fma(d_fma(T(3) * x[3], mf, T(2) * x[2]), mf, x[1]);
    } else if (falloff) {
        falloff_adjoint<MixIx>(r, mf, amplm, g_aml, acc, g);
    } else {
        lorentzian_adjoint<MixIx>(r, mf, amplm, g_aml, acc, g);
    }
    ringdown_mf_adjoint<Grid, xMsec, xInvM>(r, p, m1, g, acc, gf);
}

}  // namespace wfx
