"""Pins of the oracle (CPU). The reference holds no golden vectors for
IMRPhenomXAS and cannot run here (JAX absent), so the oracle is pinned by:

* a second, independent arrangement of the same model: the LAL form of
  ``xpnr/xas.py:20-61`` (per ``M f``, QNM fits in powers of ``a_f``, the
  reference phase through ``phase_22``) against the table form
  (``xas/core.py``: powers of ``f`` in Hz, the mass folded into the
  coefficients, QNM fits in powers of ``1 - a_f``);
* closed forms: the Kerr ISCO (Bardeen, Press & Teukolsky 1972), the
  Newtonian phase coefficient ``-3 / (128 eta) (pi M)^(-5/3)``, the
  non-spinning TaylorF2 3.5PN coefficients (e.g. Buonanno, Iyer, Ochsner,
  Pan & Sathyaprakash 2009, Table I);
* invariants of the polarizations: the reference phase at ``f_ref``, the
  ``1 / d_L`` and ``phi_ref`` dependence, the inclination factors, the
  strain scale, the exchange of the bodies, zeros above the cutoff, NaN
  rows where LAL fails, continuity across the region boundaries;
* brute force: the pivoted solves against ``numpy.linalg.solve``; the QNM
  re-expansion against direct evaluation (tests/test_xas.py:142-154);
  forward-mode derivatives against central differences.
"""

import numpy as np
import pytest

from conftest import CASES, MTSUN_SI, case_grid, case_params, compare
from oracle import oracle


@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_table_form_matches_lal_form(case):
    f = case_grid(case)
    p = case_params(case)
    hp, hc = oracle.polarizations(p[None], f)
    lp, lc = oracle.lal_form(p, f)
    for a, b in ((hp[0], lp), (hc[0], lc)):
        error, mismatch = compare(a, b)
        # rounding of ~1e4 rad of BNS phase in two arrangements
        assert error < 5e-11, error
        assert mismatch < 1e-14, mismatch


def test_table_form_matches_lal_form_random():
    params = oracle.draw(24, seed=11, m1=(5.0, 100.0))
    f = np.linspace(20.0, 1024.0, 2000)
    hp, hc = oracle.polarizations(params, f)
    for i, p in enumerate(params):
        lp, lc = oracle.lal_form(p, f)
        assert compare(hp[i], lp)[0] < 5e-11
        assert compare(hc[i], lc)[0] < 5e-11


def test_qnm_one_minus_a_form():
    """The QNM fits re-expanded in powers of 1 - a_f against the fits in
    powers of a_f (the reference's bound, tests/test_xas.py:142-154)."""
    for af in [-0.99, -0.5, 0.0, 0.5, 0.9, 0.99, 0.999]:
        ring, damp, ring_a, damp_a = oracle.qnm22(af)
        assert ring == pytest.approx(ring_a, rel=1e-10)
        assert damp == pytest.approx(damp_a, rel=1e-10)


def _kerr_isco(a):
    """Bardeen-Press-Teukolsky prograde ISCO radius and its orbital
    frequency M f = Omega / pi (Omega = 1 / (r^1.5 + a))."""
    z1 = 1 + (1 - a * a) ** (1 / 3) * ((1 + a) ** (1 / 3) + (1 - a) ** (1 / 3))
    z2 = np.sqrt(3 * a * a + z1 * z1)
    r = 3 + z2 - np.sign(a) * np.sqrt((3 - z1) * (3 + z1 + 2 * z2))
    return 1.0 / (r ** 1.5 + a) / np.pi


@pytest.mark.parametrize("a", [0.0, 0.3, 0.7, 0.95, -0.5, -0.9])
def test_f_isco_closed_form(a):
    assert oracle.f_isco(a) == pytest.approx(_kerr_isco(a), rel=1e-13)
    if a == 0.0:
        assert oracle.f_isco(0.0) == pytest.approx(6.0 ** -1.5 / np.pi, rel=1e-15)


@pytest.mark.parametrize("m1,m2", [(36.0, 29.0), (1.4, 1.3), (80.0, 10.0)])
def test_newtonian_coefficient(m1, m2):
    """T0 = -3 / (128 eta) (pi M_sec)^(-5/3): the coefficient of f^(-5/3)."""
    info = oracle.internals([m1, m2, 0.1, -0.2, 400.0, 0.3, 0.2, 20.0], 20.0)
    eta = m1 * m2 / (m1 + m2) ** 2
    expected = -3.0 / (128.0 * eta) * (np.pi * (m1 + m2) * MTSUN_SI) ** (-5.0 / 3.0)
    assert info["T0"] == pytest.approx(expected, rel=1e-13)
    assert info["eta"] == pytest.approx(eta, rel=1e-14)
    assert info["m_sec"] == pytest.approx((m1 + m2) * MTSUN_SI, rel=1e-15)


@pytest.mark.parametrize("q", [1.0, 0.5, 0.2])
def test_taylorf2_nonspinning_coefficients(q):
    """The non-spinning TaylorF2 coefficients in LAL's normalisation
    (coefficient of (pi M f)^(k/3), times pi^(k/3)) against the 3.5PN
    textbook values."""
    m1, m2 = 30.0, 30.0 * q
    eta = m1 * m2 / (m1 + m2) ** 2
    info = oracle.internals([m1, m2, 0.0, 0.0, 400.0, 0.3, 0.2, 20.0], 20.0)
    pi, g = np.pi, 0.5772156649015329
    assert info["phi2"] == pytest.approx((3715 / 756 + 55 * eta / 9) * pi ** (2 / 3), rel=1e-14)
    assert info["phi3"] == pytest.approx(-16 * pi * pi, rel=1e-14)
    assert info["phi4"] == pytest.approx((15293365 / 508032 + 27145 * eta / 504 + 3085 * eta ** 2 / 72)
                                         * pi ** (4 / 3), rel=1e-14)
    assert info["phi5l"] == pytest.approx(pi * (38645 / 756 - 65 * eta / 9) * pi ** (5 / 3), rel=1e-13)
    phi6 = (11583231236531 / 4694215680 - 640 * pi ** 2 / 3 - 6848 * g / 21
            - 6848 / 21 * (2 * np.log(2) + np.log(pi) / 3)
            + (-15737765635 / 3048192 + 2255 * pi ** 2 / 12) * eta + 76055 * eta ** 2 / 1728
            - 127825 * eta ** 3 / 1296) * pi ** 2
    assert info["phi6"] == pytest.approx(phi6, rel=1e-13)
    phi7 = pi * (77096675 / 254016 + 378515 * eta / 1512 - 74045 * eta ** 2 / 756) * pi ** (7 / 3)
    assert info["phi7"] == pytest.approx(phi7, rel=1e-13)


@pytest.mark.parametrize("n", [4, 5])
def test_pivoted_solve_matches_numpy(n):
    rng = np.random.default_rng(n)
    for _ in range(20):
        a = rng.normal(size=(n, n))
        b = rng.normal(size=n)
        np.testing.assert_allclose(oracle.solve(a, b), np.linalg.solve(a, b), rtol=1e-10, atol=1e-12)
    # collocation-like rows [1, x^(1/3), x^(2/3), x] on [0.0026, 0.02]
    x = np.linspace(0.0026, 0.02, n)
    a = np.stack([x ** (k / 3) for k in range(n)], axis=1)
    b = rng.normal(size=n)
    np.testing.assert_allclose(oracle.solve(a, b), np.linalg.solve(a, b), rtol=1e-8)


def _grid_with(f_ref, n=3000):
    f = np.geomspace(f_ref * 0.5, 1024.0, n)
    return np.sort(np.append(f, f_ref))


@pytest.mark.parametrize("case", CASES[:8], ids=[f"case{i}" for i in range(8)])
def test_reference_phase(case):
    """arg h+(f_ref) = 2 phi_ref + pi/4 + pi (kp < 0), arg hx(f_ref) =
    2 phi_ref + pi/4 + pi/2 sign(cos i) (xas/core.py:351-358,504-507)."""
    p = case_params(case)
    f_ref = p[7] if p[7] > 0 else case[7]
    f = _grid_with(f_ref) if p[7] > 0 else np.geomspace(f_ref, 1024.0, 3000)
    hp, hc = oracle.polarizations(p[None], f)
    i = int(np.flatnonzero(f == f_ref)[0])
    phi = 2 * p[6] + np.pi / 4
    assert np.angle(hp[0, i] * np.exp(-1j * (phi + np.pi))) == pytest.approx(0.0, abs=1e-9)
    sign = np.sign(np.cos(p[5]))
    assert np.angle(hc[0, i] * sign * np.exp(-1j * (phi + np.pi / 2))) == pytest.approx(0.0, abs=1e-9)


def test_scalings_and_symmetries():
    f = np.linspace(20.0, 1024.0, 1500)
    p = np.array([36.0, 29.0, 0.3, -0.4, 500.0, 1.0, 1.2, 30.0])
    hp, hc = oracle.polarizations(p[None], f)
    # 1 / d_L
    q = p.copy()
    q[4] = 1000.0
    hp2, _ = oracle.polarizations(q[None], f)
    np.testing.assert_allclose(hp2, hp / 2, rtol=1e-14, atol=0)
    # phi_ref -> phi_ref + d: e^{2 i d}
    q = p.copy()
    q[6] += 0.3
    hp3, hc3 = oracle.polarizations(q[None], f)
    np.testing.assert_allclose(hp3, hp * np.exp(2j * 0.3), rtol=1e-11, atol=1e-11 * abs(hp).max())
    # inclination factors: h+ ~ (1 + cos^2 i) / 2, hx ~ cos i
    q = p.copy()
    q[5] = 0.0
    hp4, hc4 = oracle.polarizations(q[None], f)
    np.testing.assert_allclose(hp, hp4 * (1 + np.cos(1.0) ** 2) / 2, rtol=1e-13, atol=1e-15 * abs(hp).max())
    np.testing.assert_allclose(hc, hc4 * np.cos(1.0), rtol=1e-13, atol=1e-15 * abs(hc).max())
    # hx = i cos(i) / (-(1 + cos^2 i) / 2) h+
    np.testing.assert_allclose(hc, hp * 1j * np.cos(1.0) / (-(1 + np.cos(1.0) ** 2) / 2), rtol=1e-12,
                               atol=1e-15 * abs(hc).max())
    # the strain scale
    hp5, _ = oracle.polarizations(p[None], f, amp_scale=1.0)
    np.testing.assert_allclose(hp5 * 1e21, hp, rtol=1e-14, atol=0)
    # exchanging the bodies (internals.py:53-58)
    q = p.copy()
    q[[0, 1, 2, 3]] = p[[1, 0, 3, 2]]
    hp6, _ = oracle.polarizations(q[None], f)
    np.testing.assert_allclose(hp6, hp, rtol=1e-13, atol=1e-14 * abs(hp).max())


def test_cutoff_and_failures():
    f = np.linspace(20.0, 1024.0, 4017)
    p = np.array([[80.0, 70.0, 0.2, 0.1, 500.0, 0.3, 0.5, 20.0],
                  [2000.0, 1.0, 0.0, 0.0, 400.0, 0.4, 0.0, 20.0],
                  [30.0, 20.0, 1.2, 0.0, 400.0, 0.4, 0.0, 20.0],
                  [30.0, 20.0, 0.999, 0.999, 400.0, 0.4, 0.0, 20.0]])
    hp, hc = oracle.polarizations(p, f)
    f_cut = 0.3 / (150.0 * MTSUN_SI)
    assert np.all(hp[0, f > f_cut] == 0) and np.all(hp[0, f <= f_cut] != 0)
    assert np.all(np.isnan(hp[1])) and np.all(np.isnan(hc[2]))
    # chi_eff > 0.99: the cutoff moves to 0.33 / M
    _, flags = oracle.tables(p, 20.0)
    assert flags[1] & 1 and flags[2] & 2 and flags[0] == 0
    info = oracle.internals(p[3], 20.0)
    assert info["f_cut"] == pytest.approx(0.33 / (50.0 * MTSUN_SI), rel=1e-14)


@pytest.mark.parametrize("case", CASES[:6], ids=[f"case{i}" for i in range(6)])
def test_continuity_at_region_boundaries(case):
    """The phase and its slope (C1, phase22.py:266-286) and the amplitude
    are continuous across f_phase_in, f_phase_im, f_amp_in, f_amp_rd."""
    p = case_params(case)
    info = oracle.internals(p, case[7])
    for key in ("f_phase_in", "f_phase_im", "f_amp_in", "f_amp_rd"):
        fb = info[key]
        if not case[7] < fb < info["f_cut"]:
            continue
        eps = fb * 1e-7
        f = np.array([fb - 2 * eps, fb - eps, fb + eps, fb + 2 * eps])
        hp, _ = oracle.polarizations(p[None], np.concatenate([[case[7]], f]))
        h = hp[0, 1:]
        amp = np.abs(h)
        phase = np.unwrap(np.angle(h))
        assert abs(amp[2] - amp[1]) / amp[1] < 1e-5, key
        slope_lo = (phase[1] - phase[0]) / eps
        slope_hi = (phase[3] - phase[2]) / eps
        jump = phase[2] - phase[1]
        assert abs(jump - 0.5 * (slope_lo + slope_hi) * 2 * eps) < 1e-6 * max(1.0, abs(phase[1])), key
        if key.startswith("f_phase"):
            assert abs(slope_hi - slope_lo) < 1e-3 * abs(slope_lo) + 1e-6, key


def test_forward_mode_matches_central_differences():
    """The oracle's forward-mode derivatives of (h+, hx) against central
    differences, step 1e-6 max(|x|, 1) (tests/test_transforms_xas.py:47-51)."""
    p = np.array([36.0, 29.0, 0.3, -0.4, 500.0, 1.0, 1.2, 30.0])
    f = np.linspace(20.0, 512.0, 400)
    dhp, dhc = oracle.jacobian(p[None], f)
    for k in range(8):
        h = 1e-6 * max(abs(p[k]), 1.0)
        a, b = p.copy(), p.copy()
        a[k] += h
        b[k] -= h
        fd = (oracle.polarizations(a[None], f)[0] - oracle.polarizations(b[None], f)[0])[0] / (2 * h)
        scale = np.max(np.abs(dhp[0, :, k]))
        if scale == 0:
            assert np.max(np.abs(fd)) < 1e-6
            continue
        assert np.max(np.abs(fd - dhp[0, :, k])) / scale < 1e-5, k


def test_op_count():
    """The algorithmic work of the formula per bin (frozen from the code)."""
    params = oracle.draw(64, seed=0)
    f = np.linspace(20.0, 1024.0, 4017)
    tables, _ = oracle.tables(params, f[0])
    per_freq, per_bin = oracle.count_ops(tables, f)
    bins = 64 * len(f)
    assert per_freq["log"] == 2 * len(f) and per_freq["exp"] == 2 * len(f)
    assert per_bin["sincos"] == bins
    flops = (per_bin["add"] + per_bin["mul"]) / bins
    assert 30 < flops < 60, flops
