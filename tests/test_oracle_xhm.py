"""The IMRPhenomXHM oracle (``oracle/xhm_oracle.h``) pinned without the
reference: no LALSimulation and no JAX here, and the reference's own XHM
tests compare against live LAL (``tests/test_xhm.py:91-128``), so there are
no golden vectors. What the model's mathematics fixes instead:

* the phase of every mode is C1 and its amplitude continuous (C1 where
  LAL's intermediate amplitude takes the boundary derivative) across the
  region boundaries: the connection coefficients and collocation systems
  (``xhm/reference.py:275-334,555-563``) put them there, so a wrong term in
  an ansatz or its derivative breaks it;
* the exact derivatives of the rotated (3,2) ringdown phase against
  well-conditioned finite differences, as ``tests/test_xhm.py:131-153``;
* the (2,2) mode is IMRPhenomXAS (``xhm/reference.py:656-666``);
* the odd modes vanish for equal masses and spins
  (``xpnr/hm/variables.py:144-147``);
* at low frequency each mode over the (2,2) at the same orbital frequency
  tends to its leading post-Newtonian ratio (Blanchet, Living Rev. Rel. 17
  (2014) 2, eq. 327, with the stationary-phase factor sqrt(2/m));
* the harmonics against their closed forms;
* exact invariances: swapping the bodies, 1/d_L, and h_{l,-m} ~ e^{i m phi_ref}.
"""

import numpy as np
import pytest

from oracle import oracle, xhm

MTSUN = 4.925490947641267e-06
# mass_1, mass_2, chi_1, chi_2, distance, inclination, phi_ref, f_min, f_ref
# (the cases of the reference's tests/test_xhm.py:30-40)
CASES = [
    (36.0, 29.0, 0.3, -0.4, 500.0, 1.0, 1.2, 20.0, 30.0),
    (60.0, 15.0, 0.8, 0.5, 1000.0, 2.5, 4.0, 15.0, 20.0),
    (80.0, 10.0, -0.7, 0.2, 2000.0, 2.0, 0.3, 10.0, 25.0),
    (1.6, 1.3, 0.05, 0.02, 40.0, 0.2, 2.0, 30.0, 40.0),
    (150.0, 100.0, 0.95, 0.95, 3000.0, 0.7, 3.0, 10.0, 15.0),
    (20.0, 45.0, 0.3, -0.4, 300.0, 2.9, 0.7, 20.0, 0.0),
    (50.0, 2.0, 0.995, 0.3, 300.0, 1.1, 1.7, 20.0, 25.0),
    (100.0, 1.0, 0.5, 0.0, 300.0, 1.1, 1.7, 10.0, 0.0),
    (30.0, 30.0, 0.2, 0.2, 400.0, 0.4, 0.0, 20.0, 20.0),
]


def params(case):
    return np.array([case[k] for k in (0, 1, 2, 3, 4, 5, 6, 8)])


def grid(case, n=1000):
    return np.geomspace(case[7], min(0.3 / ((case[0] + case[1]) * MTSUN), 2048.0), n)


@pytest.mark.parametrize("case", CASES)
def test_regions_join_smoothly(case):
    p = params(case)
    for k, mode in enumerate(xhm.HIGHER_MODES):
        internals = xhm.mode_internals(p, case[7], k)
        if case[0] == case[1] and case[2] == case[3] and mode in ((2, 1), (3, 3)):
            continue  # (vanishing modes: their set-up divides by zero)
        for name in ("f_phase_match_in", "f_phase_match_im", "f_amp_match_in", "f_amp_match_im"):
            x = internals[name]
            h = 1e-6 * x
            pts = np.array([x - 2 * h, x - h, np.nextafter(x, 0.0), x, x + h, x + 2 * h])
            amp, phase = xhm.mode_amplitude_phase(p, case[7], k, pts)
            q = phase if "phase" in name else amp
            jump = abs(q[3] - q[2]) / abs(q[3])
            left = (3 * q[2] - 4 * q[1] + q[0]) / (2 * h)
            right = (-3 * q[3] + 4 * q[4] - q[5]) / (2 * h)
            kink = abs(right - left) / max(abs(right), abs(left))
            assert jump < 1e-12, (mode, name, jump)
            if mode == (2, 1) and name == "f_amp_match_in":
                continue  # (LAL's (2,1) takes the inspiral value, not its derivative)
            assert kink < 1e-4, (mode, name, kink)


@pytest.mark.parametrize("case", [CASES[0], CASES[6], CASES[2]])
def test_rotated_phase_derivatives(case):
    """The first and second derivatives of the rotated (3,2) ringdown phase
    at the intermediate boundary, against fourth-order differences (the
    reference's bars: rel 1e-7 and 1e-5)."""
    p = params(case)
    x = xhm.mode_internals(p, case[7], 2)["f_phase_match_im"]
    h = 1e-5
    _, _, rot = xhm.mode_amplitude_phase(p, case[7], 2, x + h * np.arange(-2, 3), rotation=True)
    v = np.unwrap(rot[0])
    fd1 = (-v[4] + 8 * v[3] - 8 * v[1] + v[0]) / (12 * h)
    fd2 = (-v[4] + 16 * v[3] - 30 * v[2] + 16 * v[1] - v[0]) / (12 * h * h)
    assert rot[1][2] == pytest.approx(fd1, rel=1e-7)
    assert rot[2][2] == pytest.approx(fd2, rel=1e-5)


def test_harmonics_closed_forms():
    """{}_{-2}Y_{lm} against the closed forms LAL lists
    (XLALSpinWeightedSphericalHarmonic) for every mode and its mirror."""
    for theta, phi in ((0.7, 0.3), (2.4, -1.1), (0.0, 0.5)):
        c, s, ct = np.cos(theta / 2), np.sin(theta / 2), np.cos(theta)
        closed = {
            (2, 2): np.sqrt(5 / (64 * np.pi)) * (1 + ct) ** 2,
            (2, -2): np.sqrt(5 / (64 * np.pi)) * (1 - ct) ** 2,
            (2, 1): np.sqrt(5 / (16 * np.pi)) * np.sin(theta) * (1 + ct),
            (2, -1): np.sqrt(5 / (16 * np.pi)) * np.sin(theta) * (1 - ct),
            (3, 3): -np.sqrt(21 / (2 * np.pi)) * c ** 5 * s,
            (3, -3): np.sqrt(21 / (2 * np.pi)) * c * s ** 5,
            (3, 2): np.sqrt(7 / np.pi) * c ** 4 * (3 * ct - 2) / 2,
            (3, -2): np.sqrt(7 / (4 * np.pi)) * (2 + 3 * ct) * s ** 4,
            (4, 4): 3 * np.sqrt(7 / np.pi) * c ** 6 * s ** 2,
            (4, -4): 3 * np.sqrt(7 / np.pi) * c ** 2 * s ** 6,
        }
        for (ell, m), value in closed.items():
            y = xhm.swsh(theta, phi, ell, m)
            assert abs(y - value * np.exp(1j * m * phi)) <= 1e-15 * max(1.0, abs(value)), (ell, m, theta)


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[6], CASES[3]])
def test_22_mode_is_imrphenomxas(case):
    p = params(case)
    f = np.linspace(case[7], 1024.0, 1500)
    _, _, modes = xhm.polarizations(p, f, modes=True)
    hp, _ = oracle.polarizations(p[None], f, amp_scale=1.0)
    kp = -np.sqrt(5 / (16 * np.pi)) * 0.5 * (1 + np.cos(p[5]) ** 2)
    h22 = hp[0] / kp
    # (the LAL form against the oracle's table form, whose inspiral phase of
    # up to 1e4 rad is arranged differently: the bar of tests/test_oracle.py)
    assert np.max(np.abs(modes[0, 0] - h22)) <= 5e-11 * np.max(np.abs(h22))


def test_odd_modes_vanish_for_equal_masses_and_spins():
    p = params(CASES[8])
    _, _, modes = xhm.polarizations(p, grid(CASES[8]), modes=True)
    assert not np.any(modes[1]) and not np.any(modes[2])
    assert np.all(np.isfinite(modes.view(np.float64)))
    assert np.all(np.abs(modes[3]) > 0) and np.all(np.abs(modes[4]) > 0)
    q = p.copy()
    q[0] = 30.0 * (1 + 1e-6)  # unequal: the odd modes are back
    _, _, modes = xhm.polarizations(q, grid(CASES[8]), modes=True)
    assert np.all(np.abs(modes[1]) > 0)


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[7]])
def test_low_frequency_post_newtonian_ratios(case):
    """|h_lm(f)| / |h_22(2 f / m)| -> sqrt(2/m) |H_lm / H_22| at leading
    order, with v = (2 pi M f / m)^(1/3): sqrt2/3 delta v (2,1),
    3/4 sqrt(5/7) delta v (3,3), 1/3 sqrt(5/7) (1 - 3 eta) v^2 (3,2),
    4 sqrt2/9 sqrt(5/7) (1 - 3 eta) v^2 (4,4); the corrections are O(v)."""
    p = params(case)
    m_total = p[0] + p[1]
    eta = p[0] * p[1] / m_total ** 2
    delta = abs(p[0] - p[1]) / m_total
    for mf, bar in ((1e-5, 0.25), (1e-7, 0.05)):
        f = mf / (m_total * MTSUN)
        for k, (ell, m) in enumerate(xhm.HIGHER_MODES):
            _, _, at_f = xhm.polarizations(p, np.array([f, 2 * f]), modes=True)
            _, _, at_22 = xhm.polarizations(p, np.array([2 * f / m, 4 * f / m]), modes=True)
            ratio = abs(at_f[k + 1, 0, 0]) / abs(at_22[0, 0, 0])
            v = (2 * np.pi * mf / m) ** (1 / 3)
            leading = {(2, 1): np.sqrt(2) / 3 * delta * v, (3, 3): 0.75 * np.sqrt(5 / 7) * delta * v,
                       (3, 2): np.sqrt(5 / 7) / 3 * (1 - 3 * eta) * v ** 2,
                       (4, 4): 4 * np.sqrt(2) / 9 * np.sqrt(5 / 7) * (1 - 3 * eta) * v ** 2}[(ell, m)]
            assert abs(ratio / leading - 1) < bar, ((ell, m), mf, ratio / leading)


@pytest.mark.parametrize("case", [CASES[0], CASES[5], CASES[6]])
def test_exact_invariances(case):
    p = params(case)
    f = grid(case, 400)
    hp, hc, modes = xhm.polarizations(p, f, modes=True)
    # the bodies swapped: LAL orders them first (internals.py:53-58)
    s = p.copy()
    s[[0, 1, 2, 3]] = p[[1, 0, 3, 2]]
    assert np.array_equal(xhm.polarizations(s, f)[0], hp)
    # twice the distance, half of every output, bit for bit
    d = p.copy()
    d[4] *= 2.0
    hp2, hc2 = xhm.polarizations(d, f)
    assert np.array_equal(hp2 * 2, hp) and np.array_equal(hc2 * 2, hc)
    # phi_ref + delta: h_{l,-m} times e^{i m delta}
    r = p.copy()
    r[6] += 0.4
    _, _, turned = xhm.polarizations(r, f, modes=True)
    for k, (ell, m) in enumerate(xhm.MODES):
        expect = modes[k] * np.exp(1j * m * 0.4)
        scale = np.max(np.abs(modes[k]))
        if scale:
            assert np.max(np.abs(turned[k] - expect)) <= 1e-9 * scale, (ell, m)


@pytest.mark.parametrize("case", CASES)
def test_collocation_conditions(case):
    """The phase and amplitude pass through their collocation points: the
    intermediate phase derivative (fourth-order differences of the phase in
    M f) equals the collocation values at their frequencies, and the
    intermediate amplitude the fitted amplitudes
    (``xhm/reference.py:275-334,397-528``): a wrong row, term or derivative
    in a system, or a point given the wrong value, breaks it where the
    region joins do not."""
    p = params(case)
    eta = case[0] * case[1] / (case[0] + case[1]) ** 2
    for k, mode in enumerate(xhm.HIGHER_MODES):
        if case[0] == case[1] and case[2] == case[3] and mode in ((2, 1), (3, 3)):
            continue
        if mode == (3, 2) and eta < 0.05:
            continue  # (the EMR (3,2) shifts its intermediate derivative to meet the ringdown's)
        c = xhm.collocation(p, case[7], k)
        assert len(c["phase_f"]) == 5
        bounds = xhm.mode_internals(p, case[7], k)
        checked = 0
        for f, v in zip(c["phase_f"], c["phase_v"]):
            if not bounds["f_phase_match_in"] <= f <= bounds["f_phase_match_im"]:
                continue  # (a point past fRD - fdamp: there the model is the ringdown's formula)
            checked += 1
            h = 1e-6 * f
            _, ph = xhm.mode_amplitude_phase(p, case[7], k, f + h * np.array([-2.0, -1.0, 1.0, 2.0]))
            d = (ph[0] - 8 * ph[1] + 8 * ph[2] - ph[3]) / (12 * h)
            assert d == pytest.approx(v, rel=1e-6, abs=1e-6 * abs(v) + 1e-3), (mode, f)
        assert checked >= 3
        for f, v in zip(c["amp_f"], c["amp_v"]):
            a, _ = xhm.mode_amplitude_phase(p, case[7], k, np.array([f]))
            assert a[0] == pytest.approx(v, rel=1e-10), (mode, f)


def test_op_count():
    """The algorithmic work of IMRPhenomXHM per bin (oracle_xhm_count: the
    branches each bin takes, the M f powers from the frequency's): one
    sine-cosine per mode (five) below every ringdown, one more where the
    (3,2)'s ringdown rotates the (2,2); the frequency's powers counted once
    per frequency; the add/multiply count frozen within a range on the
    multimode workload."""
    params = xhm.draw(16, seed=4)
    m_total = (params[:, 0] + params[:, 1]).max()
    low = np.geomspace(20.0, 0.002 / (m_total * MTSUN), 64)  # every mode in its inspiral
    per_freq, per_bin = xhm.count_ops(params, low)
    bins = 16 * len(low)
    assert per_bin["sincos"] == 5 * bins
    # per frequency: log f, f^(1/3) = exp(log f / 3), and f^(-5/3) through
    # its own log and exp (the reference's leading-term power)
    assert per_freq["exp"] == 2 * len(low) and per_freq["log"] == 2 * len(low) and per_freq["sincos"] == 0
    f = np.linspace(20.0, 2048.0, 8193)
    per_freq, per_bin = xhm.count_ops(params, f)
    bins = 16 * len(f)
    assert 5 * bins < per_bin["sincos"] <= 6 * bins
    flops = (per_bin["add"] + per_bin["mul"]) / bins
    assert 180 < flops < 280, flops
    assert 350 < xhm.op_slots(per_bin, bins) < 550
    # bench.py prices the IMRPhenomXHM rooflines with these counts on its workload (4096 binaries,
    # XHM_SPLIT_SLOTS_PER_BIN 445.9, XHM_FP64_OPS_PER_BIN 267.4): its first 16 binaries agree within 5%
    import bench

    p16 = bench.draw_xhm(4096)[:16]
    _, pb = xhm.count_ops(p16, f)
    assert abs(xhm.op_slots(pb, bins) / bench.XHM_SPLIT_SLOTS_PER_BIN[(4096, 8193)] - 1) < 0.05
    ops = (pb["add"] + pb["mul"] + pb["div"] + pb["cmp"]) / bins
    assert abs(ops / bench.XHM_FP64_OPS_PER_BIN[(4096, 8193)] - 1) < 0.05
