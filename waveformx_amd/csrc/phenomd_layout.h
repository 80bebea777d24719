// The per-binary coefficient record of IMRPhenomD in HBM: 64 slots like
// the IMRPhenomXAS record (float for the FP32 paths, double for FP64), one
// record per binary. It holds what the reference's table holds
// (phenomd/core.py:68-89) without the zero columns of the 3 x 14 phase rows:
// each region keeps its constant and the basis terms it sums
// (core.py:57-65), in the reference's summation order. T0, A and their
// FP64 copies sit in the same slots as in the IMRPhenomXAS record.
#pragma once

#include "xas_layout.h"

namespace wfx {

enum PhenomDSlot : int {
    // inspiral phase (the leading term T0 f^(-5/3) apart): constant, then
    // f^-1, f^-2/3, f^-1/3, log f, f^1/3, f^1/3 log f, f^2/3, f, f^4/3,
    // f^5/3, f^2
    kPdInsConst = 0, kPdInsM1, kPdInsM23, kPdInsM13, kPdInsLog, kPdInsP13, kPdInsP13L, kPdInsP23, kPdInsP1,
    kPdInsP43, kPdInsP53, kPdInsP2,
    // intermediate: constant, f, f^-3, log f
    kPdIntConst = 12, kPdIntP1, kPdIntM3, kPdIntLog,
    // merger-ringdown: constant, f, f^-1, f^3/4, and a atan(w (f - shift))
    kPdMrdConst = 16, kPdMrdP1, kPdMrdM1, kPdMrdP34, kPdAtanA, kPdAtanW, kPdAtanShift,
    // inspiral amplitude f^-7/6 sum_k c_k f^(k/3), k = 0..9
    kPdAmpIns = 23,
    // intermediate amplitude f^-7/6 sum_j d_j y^j, y = s f, j = 0..4
    kPdAmpInt = 33, kPdAmpIntS = 38,
    // merger-ringdown amplitude n f^-7/6 exp(-gr u) / (u^2 + 1), u = s (f - shift)
    kPdAmpRdN = 39, kPdAmpRdGr, kPdAmpRdS, kPdAmpRdShift,
    // region bounds (Hz): the inspiral phase below f_phase_ins, the
    // merger-ringdown phase from f_phase_mrd on; the amplitude likewise
    kPdFPhaseIns = 43, kPdFPhaseMrd, kPdFAmpIns, kPdFAmpMrd,
    // polarization factors: h+ = kp A e^{i phi}, hx = i kc A e^{i phi}
    kPdKp = 47, kPdKc = 48,
};

static_assert(int(kPdKc) < int(kT0), "the IMRPhenomD slots end before T0");

}  // namespace wfx
