"""Pins of the IMRPhenomD oracle (CPU). The reference's IMRPhenomD tests
compare live against LALSimulation (absent here, as is JAX) and hold no
golden vectors, so the oracle is pinned by:

* closed forms: the non-spinning TaylorF2 3.5PN coefficients (Buonanno,
  Iyer, Ochsner, Pan & Sathyaprakash 2009, Table I), the 1.5PN spin-orbit
  coefficient 4 beta with beta = sum_i chi_i (113 x_i^2 + 75 eta) / 12
  (Kidder 1995), the Schwarzschild and a = 0.9 (2,2,0) quasi-normal
  frequencies (Berti, Cardoso & Will 2006), the remnant of an equal-mass
  non-spinning merger (a_f = 0.686, E_rad = 0.048 M);
* two arrangements of one model: the table in powers of f (Hz) against the
  coefficients in M f, through the continuity of the phase (C1) and of the
  amplitude (C1) at the region boundaries, which the connection and
  collocation conditions impose, and the peak alignment (the phase is
  stationary at the amplitude's peak frequency);
* invariants of the polarizations: the reference phase at f_ref in each
  region, 1 / d_L, e^{2 i delta} under phi_ref + delta, the inclination
  factors, the strain scale, the exchange of the bodies, NaN rows when the
  final spin exceeds one, the mass-scaled reference;
* brute force: the pivoted solve against numpy.linalg.solve.
"""

import numpy as np
import pytest

from conftest import MTSUN_SI, compare
from oracle import phenomd_oracle as pd

# the reference's IMRPhenomD cases (tests/test_phenomd.py:32-44): mass_1,
# mass_2, chi_1, chi_2, distance, inclination, phi_ref, f_min, f_ref
CASES = [
    (30.0, 30.0, 0.0, 0.0, 400.0, 0.4, 0.0, 20.0, 20.0),
    (36.0, 29.0, 0.3, -0.4, 500.0, 1.0, 1.2, 20.0, 30.0),
    (60.0, 15.0, 0.8, 0.5, 1000.0, 2.5, 4.0, 15.0, 20.0),
    (80.0, 10.0, -0.7, 0.2, 2000.0, 2.0, 0.3, 10.0, 25.0),
    (1.6, 1.3, 0.05, 0.02, 40.0, 0.2, 2.0, 30.0, 40.0),
    (8.0, 1.4, 0.9, 0.0, 200.0, 0.3, 5.5, 20.0, 50.0),
    (150.0, 100.0, 0.95, 0.95, 3000.0, 0.7, 3.0, 10.0, 15.0),
    (25.0, 20.0, -0.9, -0.9, 300.0, 1.5, 0.7, 20.0, 20.0),
    (20.0, 45.0, 0.3, -0.4, 300.0, 2.9, 0.7, 20.0, 0.0),
    (50.0, 2.0, 0.995, 0.3, 300.0, 1.1, 1.7, 20.0, 25.0),
    (90.0, 5.0, 1.0, 1.0, 300.0, 1.1, 1.7, 20.0, 0.0),
]


def case_params(case):
    return np.array([case[0], case[1], case[2], case[3], case[4], case[5], case[6], case[8]])


def case_grid(case, n=1500):
    """From f_min to beyond LAL's uniform-grid cutoff M f = 0.2
    (tests/test_phenomd.py:47-51)."""
    return np.geomspace(case[7], 0.25 / ((case[0] + case[1]) * MTSUN_SI), n)


def test_solve_matches_numpy():
    rng = np.random.default_rng(3)
    a = rng.normal(size=(64, 5, 5)) + 3.0 * np.eye(5)
    a[::3] = a[::3][:, ::-1]  # pivots away from the diagonal
    b = rng.normal(size=(64, 5))
    x = pd.solve(a, b)
    np.testing.assert_allclose(x, np.linalg.solve(a, b[..., None])[..., 0], rtol=1e-12, atol=1e-12)


def test_integer_powers():
    x = np.array([0.3, 1.7, 2.9])
    for n in range(-6, 10):
        np.testing.assert_allclose(pd.ipow(x, n), x ** float(n), rtol=4e-16 * (abs(n) + 1))


@pytest.mark.parametrize("q", [1.0, 0.5, 0.1])
def test_taylorf2_nonspinning(q):
    x1, x2 = np.array([1.0 / (1.0 + q)]), np.array([q / (1.0 + q)])
    eta = x1 * x2
    v, logs = pd.taylorf2(x1, x2, np.zeros(1), np.zeros(1))
    n = 3.0 / (128.0 * eta)
    pi, gamma = np.pi, 0.5772156649015329
    expected = {
        2: 3715.0 / 756.0 + 55.0 / 9.0 * eta,
        3: -16.0 * pi,
        4: 15293365.0 / 508032.0 + 27145.0 / 504.0 * eta + 3085.0 / 72.0 * eta ** 2,
        5: pi * (38645.0 / 756.0 - 65.0 / 9.0 * eta),
        6: (11583231236531.0 / 4694215680.0 - 640.0 / 3.0 * pi ** 2 - 6848.0 / 21.0 * gamma
            - 6848.0 / 21.0 * np.log(4.0) + (-15737765635.0 / 3048192.0 + 2255.0 / 12.0 * pi ** 2) * eta
            + 76055.0 / 1728.0 * eta ** 2 - 127825.0 / 1296.0 * eta ** 3),
        7: pi * (77096675.0 / 254016.0 + 378515.0 / 1512.0 * eta - 74045.0 / 756.0 * eta ** 2),
    }
    for k, value in expected.items():
        np.testing.assert_allclose(v[k], n * value, rtol=1e-13)
    np.testing.assert_allclose(logs[5], 3.0 * n * expected[5], rtol=1e-13)
    np.testing.assert_allclose(logs[6], n * -6848.0 / 21.0, rtol=1e-13)


def test_spin_orbit_coefficient():
    rng = np.random.default_rng(5)
    x1 = rng.uniform(0.5, 0.95, 16)
    x2 = 1.0 - x1
    chi1, chi2 = rng.uniform(-1, 1, 16), rng.uniform(-1, 1, 16)
    eta = x1 * x2
    beta = (chi1 * (113.0 * x1 ** 2 + 75.0 * eta) + chi2 * (113.0 * x2 ** 2 + 75.0 * eta)) / 12.0
    np.testing.assert_allclose(pd.phase_3(x1, x2, chi1, chi2), -16.0 * np.pi + 4.0 * beta, rtol=1e-13)


def test_quasi_normal_modes_and_remnant():
    ring, damp = pd.qnm(np.array([0.0, 0.9]))
    # (2,2,0): M omega = 0.37367 - 0.08896 i at a = 0, 0.67161 - 0.06486 i at a = 0.9
    np.testing.assert_allclose(ring, np.array([0.37367, 0.67161]) / (2 * np.pi), rtol=2e-4)
    np.testing.assert_allclose(damp, np.array([0.08896, 0.06486]) / (2 * np.pi), rtol=2e-3)
    eta, delta, zero = np.array([0.25]), np.zeros(1), np.zeros(1)
    assert pd.final_spin(eta, delta, zero, zero)[0] == pytest.approx(0.686, abs=2e-3)
    assert pd.radiated_energy(eta, delta, zero, zero)[0] == pytest.approx(0.048, abs=1.5e-3)


def region_phase(fields, r, f):
    """Region r's phase sum (leading term included in the inspiral) of the
    table at the frequencies f (B, n), whatever region f falls in."""
    fp = pd.frequency_powers(f)
    basis = pd.phase_basis(fp)
    total = fields["phase"][:, r, pd.CONST, None]
    for k in pd.REGION_TERMS[r]:
        total = total + fields["phase"][:, r, k, None] * basis[k]
    if r == 0:
        total = total + fields["big"][:, 0, None] * fp[6]
    if r == 2:
        total = total + fields["atan_a"][:, None] * np.arctan(fields["atan_w"][:, None] * (f - fields["atan_shift"][:, None]))
    return total


def region_amplitude(fields, r, f):
    fp = pd.frequency_powers(f)
    t, m76 = fp[1], fp[4]
    if r == 0:
        poly = fields["amp_ins"][:, 9, None]
        for k in range(8, -1, -1):
            poly = poly * t + fields["amp_ins"][:, k, None]
        return m76 * poly
    if r == 1:
        y = fields["amp_int_s"][:, None] * f
        inter = fields["amp_int"][:, 4, None]
        for j in range(3, -1, -1):
            inter = inter * y + fields["amp_int"][:, j, None]
        return m76 * inter
    u = fields["amp_rd_s"][:, None] * (f - fields["amp_rd_shift"][:, None])
    return m76 * fields["amp_rd_n"][:, None] * np.exp(-fields["amp_rd_gr"][:, None] * u) / (u * u + 1.0)


@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_continuity_and_peak(case):
    """C1 phase at the inspiral/intermediate and intermediate/merger
    boundaries, C1 amplitude at both amplitude boundaries, and a stationary
    phase at the amplitude's peak (the time shift), from the table in f."""
    fields = pd.table(case_params(case)[None], case[7])
    if not np.isfinite(fields["kp"][0]):
        pytest.skip("LAL fails for this binary")
    for (a, b, bound) in ((0, 1, "f_phase_ins"), (1, 2, "f_phase_mrd")):
        fb = fields[bound][:, None]
        h = 1e-4 * fb
        grid = fb + np.array([[-1.0, 0.0, 1.0]]) * h
        left, right = region_phase(fields, a, grid), region_phase(fields, b, grid)
        scale = np.abs(left[:, 1]) + 1.0
        assert np.abs(left[:, 1] - right[:, 1]) / scale < 1e-12
        dl = (left[:, 2] - left[:, 0]) / (2 * h[:, 0])
        dr = (right[:, 2] - right[:, 0]) / (2 * h[:, 0])
        assert np.abs(dl - dr) / np.abs(dl) < 1e-6
    for (a, b, bound) in ((0, 1, "f_amp_ins"), (1, 2, "f_amp_mrd")):
        fb = fields[bound][:, None]
        h = 1e-4 * fb
        grid = fb + np.array([[-1.0, 0.0, 1.0]]) * h
        left, right = region_amplitude(fields, a, grid), region_amplitude(fields, b, grid)
        assert np.abs(left[:, 1] - right[:, 1]) / np.abs(right[:, 1]) < 1e-10
        dl = (left[:, 2] - left[:, 0]) / (2 * h[:, 0])
        dr = (right[:, 2] - right[:, 0]) / (2 * h[:, 0])
        assert np.abs(dl - dr) / np.abs(dl) < 1e-5
    # LAL's time shift t0 = dPhi_MRD/dMf at the amplitude's peak: there the
    # phase's slope in M f is the merger-ringdown connection's, -c2_mrd
    c = fields["coefficients"]
    s = (case[0] + case[1]) * MTSUN_SI
    fpk = fields["f_amp_mrd"][:, None]
    h = 1e-5 * fpk
    grid = fpk + np.array([[-1.0, 1.0]]) * h
    assert fields["f_amp_mrd"][0] >= fields["f_phase_mrd"][0]
    phi = region_phase(fields, 2, grid)
    slope = (phi[:, 1] - phi[:, 0]) / (2 * h[:, 0]) / s
    typical = np.abs(c.c2_mrd) + np.abs(c.t0)
    assert np.abs(slope + c.c2_mrd) / typical < 1e-6


@pytest.mark.parametrize("f_ref_mf", [0.005, 0.03, 0.2])
def test_reference_phase_in_each_region(f_ref_mf):
    """phi(f_ref) = 2 phi_ref (kp > 0: arg h+), with f_ref in the inspiral,
    the intermediate and the merger-ringdown."""
    params = pd.draw(12, seed=4)
    total = params[:, 0] + params[:, 1]
    params[:, 7] = f_ref_mf / (total * MTSUN_SI)
    for i, p in enumerate(params):
        hp, _ = pd.polarizations(p[None], np.array([10.0, p[7]]))
        assert abs(np.angle(hp[0, 1] * np.exp(-2j * p[6]))) < 1e-10


def test_scalings_and_symmetries():
    p = pd.draw(6, seed=7)
    f = np.geomspace(15.0, 900.0, 700)
    hp, hc = pd.polarizations(p, f)
    # 1 / d_L
    q = p.copy()
    q[:, 4] *= 2.5
    hp2, hc2 = pd.polarizations(q, f)
    np.testing.assert_allclose(hp2 * 2.5, hp, rtol=1e-13, atol=0)
    # e^{2 i delta} under phi_ref + delta
    q = p.copy()
    q[:, 6] += 0.3
    hp2, _ = pd.polarizations(q, f)
    np.testing.assert_allclose(hp2, hp * np.exp(0.6j), rtol=1e-11, atol=1e-11 * np.abs(hp).max())
    # inclination factors: h+ ~ (1 + cos^2 i) / 2, hx ~ -i cos i
    ratio = hc / hp
    cos_i = np.cos(p[:, 5])
    expected = (-1j * cos_i / (0.5 * (1.0 + cos_i ** 2)))[:, None]
    np.testing.assert_allclose(ratio, np.broadcast_to(expected, ratio.shape), rtol=1e-12)
    # the strain scale
    hp3, _ = pd.polarizations(p, f, amp_scale=1.0)
    np.testing.assert_allclose(hp3 * 1e21, hp, rtol=1e-14)
    # exchange of the bodies
    q = p.copy()
    q[:, [0, 1, 2, 3]] = p[:, [1, 0, 3, 2]]
    hp2, hc2 = pd.polarizations(q, f)
    assert compare(hp2[0], hp[0])[0] < 1e-11
    assert compare(hc2[3], hc[3])[0] < 1e-11


def test_final_spin_failure_and_mass_scaled_reference():
    # aligned spins of 1.3 (beyond Kerr: the setup's flag, not the API's
    # spin check, is what is tested) give a final spin above one
    p = np.array([[30.0, 15.0, 1.3, 1.3, 400.0, 0.4, 0.0, 20.0], case_params(CASES[10])])
    with np.errstate(invalid="ignore"):
        fields = pd.table(p, 20.0)
    assert fields["final_spin"][0] > 1.0 and np.isnan(fields["kp"][0])
    assert fields["final_spin"][1] < 1.0 and np.isfinite(fields["kp"][1])
    # mass-scaled reference: f_ref must be zero (NaN otherwise)
    p = pd.draw(4, seed=9)
    f = np.geomspace(20.0, 500.0, 300)
    p0 = p.copy()
    p0[:, 7] = 0.0
    hp, _ = pd.polarizations(p0, f, reference_mf=0.01)
    assert np.isfinite(hp).all()
    hp_bad, _ = pd.polarizations(p, f, reference_mf=0.01)
    assert np.isnan(hp_bad).all()
    # and it equals the physical reference at f_ref = reference_mf / M
    p1 = p0.copy()
    p1[:, 7] = 0.01 / ((p[:, 0] + p[:, 1]) * MTSUN_SI)
    hp1, _ = pd.polarizations(p1, f)
    for i in range(4):
        assert compare(hp[i], hp1[i])[0] < 1e-9


def test_finite_difference_jacobian():
    """The oracle's difference Jacobian (the checker of the GPU gradients):
    closed forms where the parameter enters the polarizations in one place
    (h ~ 1/d_L; h+ ~ e^{i 2 phi_ref}; h+ ~ (1 + cos^2 i) / 2, hx ~ -cos i),
    and, for the masses, spins and f_ref, agreement of steps 10x apart
    (their truncation and rounding errors differ)."""
    params = pd.draw(6, seed=21)
    f = np.linspace(20.0, 1024.0, 300)
    hp, hc = pd.polarizations(params, f)
    dhp, dhc = pd.jacobian(params, f)
    assert dhp.shape == (6, 300, 8) and dhc.shape == (6, 300, 8)
    d, iota = params[:, 4:5], params[:, 5:6]
    scale_p, scale_c = np.abs(hp).max(axis=1, keepdims=True), np.abs(hc).max(axis=1, keepdims=True)
    assert np.all(np.abs(dhp[:, :, 4] + hp / d) <= 1e-8 * scale_p / d)
    assert np.all(np.abs(dhc[:, :, 4] + hc / d) <= 1e-8 * scale_c / d)
    assert np.all(np.abs(dhp[:, :, 6] - 2j * hp) <= 1e-8 * scale_p)
    assert np.all(np.abs(dhc[:, :, 6] - 2j * hc) <= 1e-8 * scale_c)
    kp, kc = 0.5 * (1 + np.cos(iota) ** 2), -np.cos(iota)
    assert np.allclose(dhp[:, :, 5], hp * (-np.cos(iota) * np.sin(iota)) / kp, rtol=0, atol=1e-8 * scale_p.max())
    assert np.allclose(dhc[:, :, 5], hc * np.sin(iota) / kc, rtol=0, atol=1e-8 * scale_c.max())
    coarse = pd.jacobian(params, f, step=1e-4)
    for k in (0, 1, 2, 3, 7):
        for fine, wide in ((dhp, coarse[0]), (dhc, coarse[1])):
            col_scale = np.abs(fine[:, :, k]).max()
            assert np.abs(fine[:, :, k] - wide[:, :, k]).max() <= 1e-5 * col_scale, k
