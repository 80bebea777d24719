// The per-binary coefficient record of IMRPhenomXAS in HBM: 64 slots
// (float for the FP32 paths, 256 B; double for FP64, 512 B), one record per
// binary, contiguous, so that a wave reads its binary's record with scalar
// loads. It holds what the reference's 86-float table holds
// (xas/core.py:68-101), without the zero columns of the 3 x 16 phase rows:
// each region keeps only the basis terms it uses (core.py:409-415), in the
// reference's summation order.
#pragma once

namespace wfx {

enum XasSlot : int {
    // inspiral phase (the leading term T0 f^(-5/3) apart): constant, then
    // f^-1, f^-2/3, f^-1/3, log f, f^1/3, f^1/3 log f, f^2/3, f, f log f,
    // f^4/3, f^5/3, f^2
    kInsConst = 0, kInsM1, kInsM23, kInsM13, kInsLog, kInsP13, kInsP13L, kInsP23, kInsP1, kInsP1L, kInsP43,
    kInsP53, kInsP2,
    // intermediate: constant, f, log f, f^-1, f^-2, f^-3
    kIntConst = 13, kIntP1, kIntLog, kIntM1, kIntM2, kIntM3,
    // ringdown: constant, f, f^2/3, f^-1, f^-3
    kRdConst = 19, kRdP1, kRdP23, kRdM1, kRdM3,
    // a atan(w (f - shift)) of the intermediate and ringdown
    kAtanA1 = 24, kAtanW1, kAtanA2, kAtanW2, kShift,
    // inspiral amplitude f^-7/6 sum_k c_k f^(k/3), k = 0..9
    kAmpIns = 29,
    // intermediate amplitude n / sum_j d_j f^j, j = 0..4
    kAmpIntN = 39, kAmpIntD = 40,
    // ringdown amplitude n f^-7/6 exp(-gr u) / (u^2 + g), u = s (f - shift)
    kAmpRdN = 45, kAmpRdGr, kAmpRdS, kAmpRdG,
    // region bounds and cutoff (Hz)
    kFPhaseIn = 49, kFPhaseIm, kFAmpIn, kFAmpRd, kFCut,
    // polarization factors (NaN where LAL fails)
    kKp = 54, kKc = 55,
    // the leading term's T0 and its value A at f_ref, each as an unevaluated
    // sum hi + lo (lo is zero in the FP64 record)
    kT0 = 56, kT0Lo = 57, kAnchor = 58, kAnchorLo = 59,
    kSlots = 64
};

// the parameters of a binary, in the packed (B, 8) double array the setup
// reads
enum XasParam : int {
    kMass1 = 0, kMass2, kChi1, kChi2, kDistance, kInclination, kPhiRef, kFRef, kParams = 8
};

// LAL's error flags of the waveform variables (xpnr/internals.py:136-149)
enum XasFlag : unsigned {
    kFlagMassRatio = 1u, kFlagSpin = 2u, kFlagFinalMass = 4u, kFlagFinalSpin = 8u, kFlagMeco = 16u
};

}  // namespace wfx
