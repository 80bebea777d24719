"""IMRPhenomD in numpy FP64: the oracle of the HIP path's IMRPhenomD.

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker; the
product (``waveformx_amd``) never imports anything here.

A restatement, vectorised over binaries, of the reference's table + formula
model of IMRPhenomD (``src/waveformx/models/phenomd/core.py``), with the
fits of ``phenomd/fits.py`` and the TaylorF2 phasing of
``taylorf2/pn.py``; the fits' data (``phenomd_data.json``) are written
mechanically from the reference's text by ``gen_phenomd.py``. Each
function names the lines it follows. The order of the arithmetic is the
reference's (FP64 parity of the GPU path is checked against it to ~1e-12),
with two fixed choices where the reference leaves it to JAX:

* integer powers ``x ** n`` by binary exponentiation (``ipow``; ``1 /
  x^|n|`` for n < 0), the scheme of ``jax.lax.integer_pow``;
* the 5 x 5 amplitude system by Gaussian elimination with partial pivoting,
  unfused, row by row (``solve``; LAPACK's ``getrf``/``getrs`` choose the
  same pivots), checked against ``numpy.linalg.solve``.

The frequency powers and the leading term ``T0 f^(-5/3) - A`` (up to ~1e4
rad) take exp, log and cbrt from the shared-source functions that the FP64
kernels use too (``oracle.libm``, pinned against long double in
``tests/test_libm.py``): numpy's and the device's own functions differ by an
ulp, ~1e-12 rad there.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from .oracle import libm

DATA = json.loads((Path(__file__).resolve().parent / "phenomd_data.json").read_text())
COEFFICIENTS = {k: np.asarray(v, dtype=np.float64) for k, v in DATA["coefficients"].items()}
FINAL_SPIN = np.asarray(DATA["final_spin"], dtype=np.float64)
FRING = np.asarray(DATA["fring"], dtype=np.float64)
FDAMP = np.asarray(DATA["fdamp"], dtype=np.float64)

# xpnr/_math.py:38-50
PI = 3.141592653589793
EGAMMA = 0.577215664901532860606512090082402431
MTSUN_SI = 4.925490947641266978197229498498379006e-6
MRSUN_SI = 1.476625038050124729627979840144936351e3
MPC_SI = 1.0e6 * 3.085677581491367278913937957796471611e16
LN_4 = float(np.log(4.0))
# xas/core.py:235
K0 = -3.0 / 128.0 * float(np.pi * MTSUN_SI) ** (-5.0 / 3.0)
# phenomd/core.py:107-108
K_AMP = 2.0 * float(np.sqrt(5.0 / (64.0 * np.pi))) * MRSUN_SI * MTSUN_SI / MPC_SI \
    * float(np.sqrt(2.0 / 3.0) * np.pi ** (-1.0 / 6.0)) * MTSUN_SI ** (-7.0 / 6.0)
# phenomd/core.py:44-46
F_PHASE_INS = 0.018
F_AMP_INS = 0.014
LOG_6 = -6848.0 / 21.0

#: the 14 basis functions f^a (log f)^l of the phase and the ones each
#: region sums besides the constant (phenomd/core.py:50-65)
BASIS = ((-1, 0), (-2 / 3, 0), (-1 / 3, 0), (0, 1), (0, 0), (1 / 3, 0), (1 / 3, 1), (2 / 3, 0), (1, 0), (4 / 3, 0),
         (5 / 3, 0), (2, 0), (-3, 0), (3 / 4, 0))
INDEX = {key: i for i, key in enumerate(BASIS)}
CONST = INDEX[(0, 0)]
REGION_TERMS = (tuple(i for i, key in enumerate(BASIS) if key not in ((0, 0), (-3, 0), (3 / 4, 0))),
                (INDEX[(1, 0)], INDEX[(-3, 0)], INDEX[(0, 1)]),
                (INDEX[(1, 0)], INDEX[(-1, 0)], INDEX[(3 / 4, 0)]))

NAMES = ("mass_1", "mass_2", "chi_1", "chi_2", "luminosity_distance", "inclination", "phi_ref", "f_ref")


def ipow(x, n):
    """x ** n for an integer n by binary exponentiation (the order of
    ``jax.lax.integer_pow``)."""
    x = np.asarray(x, dtype=np.float64)
    if n == 0:
        return np.ones_like(x)
    recip, n = n < 0, abs(n)
    acc = None
    while n > 0:
        if n & 1:
            acc = x if acc is None else acc * x
        n >>= 1
        if n > 0:
            x = x * x
    return 1.0 / acc if recip else acc


def solve(a, b):
    """Solutions (B, n) of the systems a (B, n, n) x = b (B, n): Gaussian
    elimination with partial pivoting (the first largest |pivot|), then back
    substitution; a and b are not modified."""
    a = np.array(a, dtype=np.float64)
    b = np.array(b, dtype=np.float64)
    nb, n = b.shape
    rows = np.arange(nb)
    for k in range(n):
        piv = k + np.argmax(np.abs(a[:, k:, k]), axis=1)
        swap = piv != k
        if np.any(swap):
            r, p = rows[swap], piv[swap]
            a[r, k], a[r, p] = a[r, p].copy(), a[r, k].copy()
            b[r, k], b[r, p] = b[r, p].copy(), b[r, k].copy()
        for i in range(k + 1, n):
            l = a[:, i, k] / a[:, k, k]
            for j in range(k + 1, n):
                a[:, i, j] = a[:, i, j] - l * a[:, k, j]
            b[:, i] = b[:, i] - l * b[:, k]
    x = np.zeros_like(b)
    for i in range(n - 1, -1, -1):
        s = b[:, i]
        for j in range(i + 1, n):
            s = s - a[:, i, j] * x[:, j]
        x[:, i] = s / a[:, i, i]
    return x


# --- fits (phenomd/fits.py) ---------------------------------------------------


def chi_pn(eta, delta, chi1, chi2):
    """fits.py:25-29."""
    chi_s = 0.5 * (chi1 + chi2)
    chi_a = 0.5 * (chi1 - chi2)
    return chi_s * (1.0 - eta * 76.0 / 113.0) + delta * chi_a


def phenomenological(eta, xi, rows):
    """fits.py:32-37: sum_i xi^i sum_j c_ij eta^j, Horner in both."""
    total = np.zeros_like(eta)
    for row in rows[::-1]:
        total = total * xi + (row[0] + eta * (row[1] + eta * row[2]))
    return total


def final_spin(eta, delta, chi1, chi2):
    """FinalSpin0815 (fits.py:131-147)."""
    x1, x2 = 0.5 * (1.0 + delta), 0.5 * (1.0 - delta)
    s = x1 * x1 * chi1 + x2 * x2 * chi2
    total = np.zeros_like(eta)
    for row in FINAL_SPIN[::-1]:
        inner = np.zeros_like(eta)
        for c in row[::-1]:
            inner = inner * eta + c
        total = total * s + inner
    return total


def radiated_energy(eta, delta, chi1, chi2):
    """EradRational0815 (fits.py:150-161)."""
    x1, x2 = 0.5 * (1.0 + delta), 0.5 * (1.0 - delta)
    m1s, m2s = x1 * x1, x2 * x2
    s = (m1s * chi1 + m2s * chi2) / (m1s + m2s)
    eta2 = eta * eta
    eta3 = eta2 * eta
    return (eta * (0.055974469826360077 + 0.5809510763115132 * eta - 0.9606726679372312 * eta2
                   + 3.352411249771192 * eta3)
            * (1.0 + (-0.0030302335878845507 - 2.0066110851351073 * eta + 7.7050567802399215 * eta2) * s)) \
        / (1.0 + (-0.6714403054720589 - 1.4756929437702908 * eta + 7.304676214885011 * eta2) * s)


def clenshaw(coefficients, t):
    """fits.py:164-169."""
    b1 = np.zeros_like(t)
    b2 = np.zeros_like(t)
    for c in coefficients[1:][::-1]:
        b1, b2 = 2.0 * t * b1 - b2 + c, b1
    return t * b1 - b2 + coefficients[0]


def qnm(af):
    """(Mf_ring, Mf_damp) in units of the final mass (fits.py:172-178)."""
    t = 2.0 * np.cbrt(0.5 * (1.0 - af)) - 1.0
    return clenshaw(FRING, t), clenshaw(FDAMP, t)


# --- TaylorF2 phasing (taylorf2/pn.py) ---------------------------------------


def _pairs(x1, x2, chi1, chi2):
    return ((x1, chi1), (x2, chi2))


def phase_3(x1, x2, chi1, chi2):
    """pn.py:29-30."""
    total = 0.0
    for x, chi in _pairs(x1, x2, chi1, chi2):
        total = total + x * (25.0 + 38.0 / 3.0 * x) * chi
    return -16.0 * PI + total


def phase_4(eta, x1, x2, chi1, chi2):
    """pn.py:33-38."""
    phase = 15293365.0 / 508032.0 + 27145.0 / 504.0 * eta + 3085.0 / 72.0 * eta * eta
    phase = phase - 395.0 / 4.0 * eta * chi1 * chi2
    for x, chi in _pairs(x1, x2, chi1, chi2):
        phase = phase - (50.0 + 5.0 / 8.0) * x * x * chi * chi
    return phase


def phase_5(eta, x1, x2, chi1, chi2):
    """pn.py:41-46."""
    phase = 5.0 / 9.0 * (7729.0 / 84.0 - 13.0 * eta) * PI
    for x, chi in _pairs(x1, x2, chi1, chi2):
        phase = phase - chi * x * (13915.0 / 84.0 - x * (1.0 - x) * 10.0 / 3.0
                                   + x * (12760.0 / 81.0 + x * (1.0 - x) * 170.0 / 9.0))
    return phase


def spin_spin_3pn(eta, x1, x2, chi1, chi2):
    """pn.py:49-56 (Subtract3PNSS)."""
    phase = (32675.0 / 112.0 + 5575.0 / 18.0 * eta) * eta * chi1 * chi2
    for x, chi in _pairs(x1, x2, chi1, chi2):
        phase = phase + ((47035.0 / 84.0 + 2935.0 / 6.0 * x - 120.0 * x * x)
                         + (-410825.0 / 672.0 - 1085.0 / 12.0 * x + 1255.0 / 36.0 * x * x)) * x * x * chi * chi
    return phase


def phase_6(eta, x1, x2, chi1, chi2):
    """pn.py:62-70."""
    phase = 11583231236531.0 / 4694215680.0 - 640.0 / 3.0 * PI ** 2 - 6848.0 / 21.0 * EGAMMA
    phase = phase + eta * (-15737765635.0 / 3048192.0 + 2255.0 / 12.0 * PI ** 2)
    phase = phase + eta * eta * 76055.0 / 1728.0 - eta * eta * eta * 127825.0 / 1296.0
    phase = phase + LOG_6 * LN_4 + spin_spin_3pn(eta, x1, x2, chi1, chi2)
    for x, chi in _pairs(x1, x2, chi1, chi2):
        phase = phase + PI * x * (1490.0 / 3.0 + x * 260.0) * chi
    return phase


def phase_7(eta, x1, x2, chi1, chi2):
    """pn.py:73-80."""
    phase = PI * (77096675.0 / 254016.0 + 378515.0 / 1512.0 * eta - 74045.0 / 756.0 * eta * eta)
    for x, chi in _pairs(x1, x2, chi1, chi2):
        phase = phase + chi * x * (-170978035.0 / 48384.0 + eta * 2876425.0 / 672.0 + eta * eta * 4735.0 / 144.0
                                   + x * (-7189233785.0 / 1524096.0 + eta * 458555.0 / 3024.0
                                          - eta * eta * 5345.0 / 72.0))
    return phase


def taylorf2(x1, x2, chi1, chi2):
    """(v_0..v_7, {5: l_5, 6: l_6}) with the factor 3 / (128 eta)
    (pn.py:83-91)."""
    eta = x1 * x2
    scale = 3.0 / (128.0 * eta)
    v5 = phase_5(eta, x1, x2, chi1, chi2)
    v = [np.ones_like(eta), np.zeros_like(eta), 55.0 * eta / 9.0 + 3715.0 / 756.0, phase_3(x1, x2, chi1, chi2),
         phase_4(eta, x1, x2, chi1, chi2), v5, phase_6(eta, x1, x2, chi1, chi2), phase_7(eta, x1, x2, chi1, chi2)]
    return [scale * c for c in v], {5: scale * 3.0 * v5, 6: scale * LOG_6}


# --- the model in Mf (phenomd/core.py:128-268) -------------------------------


def amplitude_pn(eta, delta, chi1, chi2):
    """The coefficients of v^k, k < 7, of the inspiral amplitude
    (core.py:128-151, LAL's init_amp_ins_prefactors)."""
    eta2 = eta * eta
    eta3 = eta2 * eta
    chi12, chi22 = chi1 * chi1, chi2 * chi2
    d1 = 1.0 + delta
    return [
        np.ones_like(eta), np.zeros_like(eta), (-969.0 + 1804.0 * eta) / 672.0,
        (chi1 * (81.0 * d1 - 44.0 * eta) + chi2 * (81.0 - 81.0 * delta - 44.0 * eta)) / 48.0,
        (-27312085.0 - 10287648.0 * chi22 - 10287648.0 * chi12 * d1 + 10287648.0 * chi22 * delta
         + 24.0 * (-1975055.0 + 857304.0 * chi12 - 994896.0 * chi1 * chi2 + 857304.0 * chi22) * eta
         + 35371056.0 * eta2) / 8.128512e6,
        (chi2 * (-285197.0 * (-1.0 + delta) + 4.0 * (-91902.0 + 1579.0 * delta) * eta - 35632.0 * eta2)
         + chi1 * (285197.0 * d1 - 4.0 * (91902.0 + 1579.0 * delta) * eta - 35632.0 * eta2)
         + 42840.0 * (-1.0 + 4.0 * eta) * PI) / 32256.0,
        -(-336.0 * (-3248849057.0 + 2943675504.0 * chi12 - 3339284256.0 * chi1 * chi2 + 2943675504.0 * chi22) * eta2
          - 324322727232.0 * eta3
          - 7.0 * (-177520268561.0 + 107414046432.0 * chi22 + 107414046432.0 * chi12 * d1
                   - 107414046432.0 * chi22 * delta + 11087290368.0 * (chi1 + chi2 + chi1 * delta - chi2 * delta) * PI)
          + 12.0 * eta * (-545384828789.0 - 176491177632.0 * chi1 * chi2 + 202603761360.0 * chi22
                          + 77616.0 * chi12 * (2610335.0 + 995766.0 * delta) - 77287373856.0 * chi22 * delta
                          + 5841690624.0 * (chi1 + chi2) * PI + 21384760320.0 * PI * PI)) / 6.0085960704e10,
    ]


class Coefficients:
    """The coefficients of IMRPhenomD in Mf for masses x1, x2 relative to
    the total (core.py:154-268), with the phase and amplitude pieces as
    functions of Mf (arrays of one value per binary)."""

    def __init__(self, x1, x2, chi1, chi2):
        self.eta = eta = x1 * x2
        self.delta = delta = x1 - x2
        xi = chi_pn(eta, delta, chi1, chi2) - 1.0
        fit = {name: phenomenological(eta, xi, c) for name, c in COEFFICIENTS.items()}
        self.final_spin = af = final_spin(eta, delta, chi1, chi2)
        ring, damp = qnm(af)
        remnant = 1.0 - radiated_energy(eta, delta, chi1, chi2)
        self.f_ring, self.f_damp = ring / remnant, damp / remnant

        v, logs = taylorf2(x1, x2, chi1, chi2)
        v = list(v) + [np.zeros_like(eta)] * 4
        v[5] = v[5] - PI / 4.0
        v[6] = v[6] - spin_spin_3pn(eta, x1, x2, chi1, chi2) * v[0]
        v[8] = fit["sigma1"] / eta / PI
        v[9] = fit["sigma2"] * 0.75 / eta / PI ** (4.0 / 3.0)
        v[10] = fit["sigma3"] * 0.6 / eta / PI ** (5.0 / 3.0)
        v[11] = fit["sigma4"] * 0.5 / eta / PI ** 2
        self.v, self.logs = v, logs
        self.beta = [fit["beta1"], fit["beta2"], fit["beta3"]]
        self.alpha = [fit[f"alpha{i}"] for i in range(1, 6)]

        f1, f2 = F_PHASE_INS, 0.5 * self.f_ring
        self.c2_int = self.dins(f1) - self.dinter(f1)
        self.c1_int = self.ins(f1) - self.inter(f1) - self.c2_int * f1
        joined = self.inter(f2) + self.c1_int + self.c2_int * f2
        self.c2_mrd = self.c2_int + self.dinter(f2) - self.dmrd(f2)
        self.c1_mrd = joined - self.mrd(f2) - self.c2_mrd * f2

        self.rho = amplitude_pn(eta, delta, chi1, chi2) + [fit["rho1"] / PI ** (7.0 / 3.0),
                                                           fit["rho2"] / PI ** (8.0 / 3.0), fit["rho3"] / PI ** 3]
        self.gamma = g1, g2, g3 = fit["gamma1"], fit["gamma2"], fit["gamma3"]
        below = g2 <= 1.0
        root = np.sqrt(np.where(below, 1.0 - g2 * g2, 1.0))
        self.f_peak = np.abs(np.where(below, self.f_ring + self.f_damp * (root - 1.0) * g3 / g2,
                                      self.f_ring - self.f_damp * g3 / g2))
        self.t0 = self.dmrd(self.f_peak)
        self.sigma = self.f_damp * g3

        y1 = F_AMP_INS / self.f_peak
        y2 = 0.5 * (1.0 + y1)
        one, zero = np.ones_like(y1), np.zeros_like(y1)
        rows = np.stack([
            np.stack([one, y1, ipow(y1, 2), ipow(y1, 3), ipow(y1, 4)], -1),
            np.stack([one, y2, ipow(y2, 2), ipow(y2, 3), ipow(y2, 4)], -1),
            np.stack([one, one, one, one, one], -1),
            np.stack([zero, one, 2.0 * y1, 3.0 * ipow(y1, 2), 4.0 * ipow(y1, 3)], -1),
            np.stack([zero, one, 2.0 * one, 3.0 * one, 4.0 * one], -1)], 1)
        values = np.stack([self.amp_ins(F_AMP_INS), fit["collocation"], self.amp_mrd(self.f_peak),
                           self.damp_ins(F_AMP_INS) * self.f_peak, self.damp_mrd(self.f_peak) * self.f_peak], -1)
        self.delta_int = solve(rows, values)
        self.fit = fit

    # the phase pieces (core.py:178-211)
    def ins(self, mf):
        u = np.cbrt(PI * mf)
        log_u = np.log(u)
        total = 0.0
        for k, c in enumerate(self.v):
            if k != 1:
                total = total + c * ipow(u, k - 5)
        extra = 0.0
        for k, c in self.logs.items():
            extra = extra + c * ipow(u, k - 5) * log_u
        return total + extra

    def dins(self, mf):
        u = np.cbrt(PI * mf)
        log_u = np.log(u)
        total = 0.0
        for k, c in enumerate(self.v):
            if k != 1:
                total = total + (k - 5) / 3.0 * c * ipow(u, k - 5)
        extra = 0.0
        for k, c in self.logs.items():
            extra = extra + c * ipow(u, k - 5) * ((k - 5) / 3.0 * log_u + 1.0 / 3.0)
        return (total + extra) / mf

    def inter(self, mf):
        b = self.beta
        return (b[0] * mf - b[2] / 3.0 / ipow(mf, 3) + b[1] * np.log(mf)) / self.eta

    def dinter(self, mf):
        b = self.beta
        return (b[0] + b[2] / ipow(mf, 4) + b[1] / mf) / self.eta

    def mrd(self, mf):
        a = self.alpha
        return (a[0] * mf - a[1] / mf + 4.0 / 3.0 * a[2] * mf ** 0.75
                + a[3] * np.arctan((mf - a[4] * self.f_ring) / self.f_damp)) / self.eta

    def dmrd(self, mf):
        a = self.alpha
        y = (mf - a[4] * self.f_ring) / self.f_damp
        return (a[0] + a[1] / ipow(mf, 2) + a[2] / mf ** 0.25 + a[3] / (self.f_damp * (1.0 + y * y))) / self.eta

    # the amplitude pieces (core.py:226-245)
    def amp_ins(self, mf):
        u = np.cbrt(PI * mf)
        total = 0.0
        for k, c in enumerate(self.rho):
            if k != 1:
                total = total + c * ipow(u, k)
        return total

    def damp_ins(self, mf):
        u = np.cbrt(PI * mf)
        total = 0.0
        for k, c in enumerate(self.rho):
            if k > 1:
                total = total + k / 3.0 * c * ipow(u, k)
        return total / mf

    def amp_mrd(self, mf):
        g1, g2, _ = self.gamma
        d = mf - self.f_ring
        return g1 * self.sigma / (d * d + self.sigma * self.sigma) * np.exp(-g2 * d / self.sigma)

    def damp_mrd(self, mf):
        g1, g2, _ = self.gamma
        d = mf - self.f_ring
        s2 = d * d + self.sigma * self.sigma
        return (-2.0 * d * self.sigma * g1 / s2 - g2 * g1) / (np.exp(g2 * d / self.sigma) * s2)

    def phase(self, mf, leading=True):
        """-Psi + t0 Mf but for its constant (core.py:261-268)."""
        inspiral = self.ins(mf)
        if not leading:
            inspiral = inspiral - self.v[0] * ipow(np.cbrt(PI * mf), -5)
        f1, f2 = F_PHASE_INS, 0.5 * self.f_ring
        psi = np.where(mf < f1, inspiral,
                       np.where(mf >= f2, self.mrd(mf) + self.c1_mrd + self.c2_mrd * mf,
                                self.inter(mf) + self.c1_int + self.c2_int * mf))
        return -psi + self.t0 * mf


# --- the table (core.py:274-377) and the formula (core.py:395-485) -----------


def frequency_powers(f):
    """f, f^(1/3), 1/f, log f, f^(-7/6), f^(3/4), f^(-5/3) (core.py:395-407)."""
    f = np.asarray(f, dtype=np.float64)
    m1 = 1.0 / f
    log_f = libm("log", f)
    t = libm("exp", log_f / 3.0)
    t = t * (4.0 - t * t * t * m1) / 3.0
    root = np.sqrt(f)
    return f, t, m1, log_f, m1 / np.sqrt(t), root * np.sqrt(root), libm("exp", -5.0 / 3.0 * libm("log", f))


def phase_basis(fp):
    """core.py:410-416 (None for the constant)."""
    f, t, m1, log_f, _, f34, _ = fp
    t2 = t * t
    mt1 = t2 * m1
    return [m1, mt1 * mt1, mt1, log_f, None, t, t * log_f, t2, f, f * t, f * t2, f * f, m1 * m1 * m1, f34]


def table(params, f_min, amp_scale=1e21, reference_mf=None):
    """The tables of binaries ``params`` (B, 8) at first frequency ``f_min``
    (core.py:274-377): a dict of arrays with a leading binary axis."""
    params = np.asarray(params, dtype=np.float64)
    m1, m2, chi1, chi2, distance, inclination, phi_ref, f_ref = params.T
    f_min = np.broadcast_to(np.asarray(f_min, dtype=np.float64), m1.shape)
    total = m1 + m2
    c = Coefficients(m1 / total, m2 / total, chi1, chi2)
    eta = c.eta
    s = total * MTSUN_SI
    log_s = np.log(s)
    rows = np.zeros((m1.shape[0], 3, len(BASIS)))

    def add(r, a, l, value):
        rows[:, r, INDEX[(a, l)]] = rows[:, r, INDEX[(a, l)]] + value

    log_pis = np.log(PI * s)
    u = np.cbrt(PI * s)
    u_pow = [np.ones_like(u), u]
    for _ in range(8):
        u_pow.append(u_pow[-1] * u)
    u_inv = 1.0 / u
    u_pow_inv = {1: u_inv, 2: u_inv * u_inv, 3: u_inv * u_inv * u_inv}
    for k, coefficient in enumerate(c.v):
        if k < 2:
            continue
        a = (k - 5) / 3.0
        scale = u_pow[k - 5] if k >= 5 else u_pow_inv[5 - k]
        if k in c.logs:
            coefficient = coefficient + c.logs[k] * log_pis / 3.0
            add(0, a, 1, -c.logs[k] / 3.0 * scale)
        add(0, a, 0, -coefficient * scale)
    b1, b2, b3 = c.beta
    add(1, 1, 0, -(b1 / eta + c.c2_int) * s)
    add(1, -3, 0, b3 / (3.0 * eta) / ipow(s, 3))
    add(1, 0, 1, -b2 / eta)
    add(1, 0, 0, -b2 / eta * log_s - c.c1_int)
    a1, a2, a3, a4, a5 = c.alpha
    add(2, 1, 0, -(a1 / eta + c.c2_mrd) * s)
    add(2, -1, 0, a2 / eta / s)
    add(2, 3 / 4, 0, -4.0 / 3.0 * a3 / eta * np.sqrt(s) * np.sqrt(np.sqrt(s)))
    add(2, 0, 0, -c.c1_mrd)
    for r in range(3):
        add(r, 1, 0, c.t0 * s)
    fields = dict(atan_a=-a4 / eta, atan_w=s / c.f_damp, atan_shift=a5 * c.f_ring / s,
                  f_phase_ins=F_PHASE_INS / s, f_phase_mrd=0.5 * c.f_ring / s)

    # the reference phase: phi(f_ref) = 2 phi_ref (core.py:340-358)
    if reference_mf is None:
        invalid = np.zeros(m1.shape, dtype=bool)
        f_ref = np.where(f_ref == 0.0, f_min, f_ref)
        fp_ref = tuple(x[:, None] for x in frequency_powers(f_ref))
        rest, inspiral = phase_rest(dict(fields, phase=rows), fp_ref)
        rest, inspiral = rest[:, 0], inspiral[:, 0]
        f_ref64 = f_ref
    else:
        invalid = f_ref != 0.0
        inspiral = np.full(m1.shape, reference_mf < F_PHASE_INS)
        rest = c.phase(np.full(m1.shape, float(reference_mf)), leading=False)
        f_ref64 = reference_mf / (MTSUN_SI * (m1 + m2))
    rows[:, :, CONST] = rows[:, :, CONST] + (2.0 * phi_ref - rest)[:, None]

    t0 = K0 * libm("cbrt", m1 + m2) / (m1 * m2)
    with np.errstate(invalid="ignore"):
        anchor = np.where(inspiral, t0 * libm("exp", -5.0 / 3.0 * libm("log", f_ref64)), 0.0)

    prefactor = K_AMP * amp_scale * np.sqrt(m1 * m2) * np.exp(-np.log(total) / 6.0) / distance
    amp_ins = np.stack([prefactor * x * u_pow[k] for k, x in enumerate(c.rho)], -1)
    g1, g2, g3 = c.gamma
    width = g3 * c.f_damp
    error = (c.final_spin > 1.0) | invalid
    cos_i = np.cos(inclination)
    return dict(
        phase=rows, big=np.stack([t0, anchor], -1), atan_a=fields["atan_a"], atan_w=fields["atan_w"],
        atan_shift=fields["atan_shift"], amp_ins=amp_ins, amp_int=prefactor[:, None] * c.delta_int,
        amp_int_s=s / c.f_peak, amp_rd_n=prefactor * g1 / width, amp_rd_gr=g2, amp_rd_s=s / width,
        amp_rd_shift=c.f_ring / s, f_phase_ins=fields["f_phase_ins"], f_phase_mrd=fields["f_phase_mrd"],
        f_amp_ins=F_AMP_INS / s, f_amp_mrd=c.f_peak / s,
        kp=np.where(error, np.nan, 0.5 * (1.0 + cos_i * cos_i)), kc=np.where(error, np.nan, -cos_i),
        final_spin=c.final_spin, coefficients=c,
    )


def phase_rest(fields, fp):
    """The phase but for the leading term and -A, and whether each
    frequency is in the inspiral (core.py:425-440); fields (B,) or (B, 1)
    broadcast against the powers."""
    f = fp[0]

    def col(x):
        return np.asarray(x)[..., None] if np.ndim(x) == 1 else np.asarray(x)

    inspiral = f < col(fields["f_phase_ins"])
    mrd = f >= col(fields["f_phase_mrd"])
    basis = phase_basis(fp)
    sums = []
    for r, terms in enumerate(REGION_TERMS):
        total = col(fields["phase"][:, r, CONST])
        for k in terms:
            total = total + col(fields["phase"][:, r, k]) * basis[k]
        sums.append(total)
    sums[2] = sums[2] + col(fields["atan_a"]) * np.arctan(col(fields["atan_w"]) * (f - col(fields["atan_shift"])))
    return np.where(inspiral, sums[0], np.where(mrd, sums[2], sums[1])), inspiral


def polarizations(params, freqs, amp_scale=1e21, reference_mf=None):
    """(h+, hx) (B, N) complex128 of binaries ``params`` (B, 8) on the shared
    grid ``freqs`` (N,) (core.py:443-485, model.py:33-50: the table at the
    grid's first frequency, every frequency evaluated)."""
    freqs = np.asarray(freqs, dtype=np.float64)
    fields = table(params, freqs[0], amp_scale, reference_mf)
    return evaluate(fields, freqs)


def evaluate(fields, freqs):
    """core.py:443-485 for tables ``fields`` on the grid ``freqs``."""
    fp = frequency_powers(np.asarray(freqs, dtype=np.float64)[None, :])
    rest, inspiral = phase_rest(fields, fp)
    t0, anchor = fields["big"][:, 0, None], fields["big"][:, 1, None]
    phi = np.where(inspiral, t0 * fp[6], 0.0) + rest - anchor
    turns = np.round(phi / (2.0 * np.pi))
    phi = phi - turns * (2.0 * np.pi)

    f, t, m76 = fp[0], fp[1], fp[4]

    def col(name, *index):
        x = fields[name]
        return x[(slice(None),) + index][:, None]

    poly = col("amp_ins", 9)
    for k in range(8, -1, -1):
        poly = poly * t + col("amp_ins", k)
    y = col("amp_int_s") * f
    inter = col("amp_int", 4)
    for j in range(3, -1, -1):
        inter = inter * y + col("amp_int", j)
    u = col("amp_rd_s") * (f - col("amp_rd_shift"))
    mrd = col("amp_rd_n") * np.exp(-col("amp_rd_gr") * u) / (u * u + 1.0)
    amp = m76 * np.where(f < col("f_amp_ins"), poly, np.where(f >= col("f_amp_mrd"), mrd, inter))
    c, s = np.cos(phi), np.sin(phi)
    ap = col("kp") * amp
    ac = col("kc") * amp
    return ap * c + 1j * (ap * s), -ac * s + 1j * (ac * c)


def reference_phase(m1, m2, chi1, chi2, mf):
    """-Psi + t0 M f at M f = mf but for its constant (reference.py:18-22)."""
    m1, m2, chi1, chi2, mf = (np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in (m1, m2, chi1, chi2, mf))
    m1, m2, chi1, chi2, mf = np.broadcast_arrays(m1, m2, chi1, chi2, mf)
    total = m1 + m2
    return Coefficients(m1 / total, m2 / total, chi1, chi2).phase(mf)


def draw(n, seed=0):
    """The benchmark workload's parameter draw, as the XAS oracle's (m1 ~ U[10, 80],
    q ~ U[0.125, 1], chi ~ U[-0.9, 0.9], d_L ~ U[100, 2000], cos i ~ U[-1,
    1], phi_ref ~ U[0, 2 pi], f_ref = 20)."""
    rng = np.random.default_rng(seed)
    m1 = rng.uniform(10.0, 80.0, n)
    q = rng.uniform(0.125, 1.0, n)
    return np.stack([m1, m1 * q, rng.uniform(-0.9, 0.9, n), rng.uniform(-0.9, 0.9, n),
                     rng.uniform(100.0, 2000.0, n), np.arccos(rng.uniform(-1.0, 1.0, n)),
                     rng.uniform(0.0, 2 * np.pi, n), np.full(n, 20.0)], axis=1)


def jacobian(params, freqs, amp_scale=1e21, reference_mf=None, step=1e-5):
    """(dh+/dparams, dhx/dparams) (B, N, 8) complex128 by fourth-order
    central differences of :func:`polarizations` in each of the 8
    parameters, step ``step * max(|x|, 1)``: the region bounds move with the
    parameters, so the differences follow the formula of whichever region a
    frequency falls in (away from a bound: the value and slope of the phase,
    core.py:213-217, and of the amplitude, core.py:250-262, are continuous
    across it). At the default step, the benchmark draw's vector-Jacobian
    products agree with those at steps 3e-6 and 3e-5 to 5e-9 of their
    largest value (rounding grows below, truncation above)."""
    params = np.asarray(params, dtype=np.float64)
    freqs = np.asarray(freqs, dtype=np.float64)
    b, n = params.shape[0], freqs.shape[0]
    dhp = np.zeros((b, n, 8), dtype=np.complex128)
    dhc = np.zeros_like(dhp)
    for k in range(8):
        h = step * np.maximum(np.abs(params[:, k]), 1.0)
        taps = []
        for m in (-2, -1, 1, 2):
            p = params.copy()
            p[:, k] = p[:, k] + m * h
            taps.append(polarizations(p, freqs, amp_scale, reference_mf))
        for j, out in enumerate((dhp, dhc)):
            out[:, :, k] = (taps[0][j] - 8.0 * taps[1][j] + 8.0 * taps[2][j] - taps[3][j]) / (12.0 * h[:, None])
    return dhp, dhc
