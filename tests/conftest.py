"""Shared test helpers. ``gpu``-marked tests need a HIP device and the
built library; the others run on the CPU (the oracle, host logic, the ABI's
exported symbols)."""

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

MTSUN_SI = 4.925490947641267e-06

# mass_1, mass_2, chi_1, chi_2, distance, inclination, phi_ref, f_min, f_ref:
# the reference's cases (tests/test_xas.py:28-40)
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
]


def case_params(case):
    m1, m2, c1, c2, distance, inclination, phi, _, f_ref = case
    return np.array([m1, m2, c1, c2, distance, inclination, phi, f_ref])


def case_grid(case, n=1500):
    """The reference's grid of a case (tests/test_xas.py:49-51): geometric
    from f_min to 1.05 x the high-spin cutoff."""
    m1, m2, f_min = case[0], case[1], case[7]
    return np.geomspace(f_min, 0.33 / ((m1 + m2) * MTSUN_SI) * 1.05, n)


def compare(test, reference):
    """``(max |a - b| / max |b|, mismatch)`` with uniform weights
    (tests/conftest.py:65-75 of the reference)."""
    test = np.asarray(test)
    reference = np.asarray(reference)
    assert test.shape == reference.shape
    scale = np.max(np.abs(reference))
    error = np.max(np.abs(test - reference)) / scale if scale > 0 else np.max(np.abs(test))
    aa = np.vdot(test, test).real
    bb = np.vdot(reference, reference).real
    mismatch = 0.0 if aa == 0 and bb == 0 else 1.0 - np.vdot(reference, test).real / np.sqrt(aa * bb)
    return error, mismatch


def psd(f):
    """The synthetic noise curve S(f) = (f/100)^-4 + 1 + (f/100)^2."""
    x = np.asarray(f) / 100.0
    return x ** -4 + 1.0 + x ** 2


def weighted_mismatch(a, b, weights):
    """1 - <a, b> / sqrt(<a, a> <b, b>), <a, b> = sum Re(a b*) w, no time
    or phase maximisation (report/common.py:125-132 of the reference)."""
    ab = np.sum((a * np.conj(b)).real * weights, axis=-1)
    aa = np.sum(np.abs(a) ** 2 * weights, axis=-1)
    bb = np.sum(np.abs(b) ** 2 * weights, axis=-1)
    return 1.0 - ab / np.sqrt(aa * bb)


@pytest.fixture(scope="session")
def cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda", 0)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")
