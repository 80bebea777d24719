"""The shared-source FP64 exp, log and cbrt (waveformx_amd/csrc/wfx_libm.h,
read by the FP64 kernels and by the oracles for the leading phase term and
the frequency powers) against long double references, and their special
values. A wrong coefficient or reduction constant in them would move both
sides of every FP64 parity test alike, so they are pinned here on their own."""

import numpy as np
import pytest

from oracle import oracle


def ulp_error(got, x, fn):
    ref = fn(x.astype(np.longdouble))
    r = ref.astype(np.float64)
    ulp = np.spacing(np.abs(r))
    return np.abs((got.astype(np.longdouble) - ref) / ulp.astype(np.longdouble)).astype(np.float64)


@pytest.mark.skipif(np.finfo(np.longdouble).nmant < 60, reason="needs x87 long double")
def test_within_one_ulp_of_long_double():
    rng = np.random.default_rng(5)
    x = rng.uniform(-40.0, 40.0, 400000)
    assert ulp_error(oracle.libm("exp", x), x, np.exp).max() < 1.0
    # log and cbrt over 1e-12 .. 1e12 (frequencies, masses, M f)
    y = np.exp(rng.uniform(-27.6, 27.6, 400000))
    assert ulp_error(oracle.libm("log", y), y, np.log).max() < 1.0
    assert ulp_error(oracle.libm("cbrt", y), y, np.cbrt).max() < 1.0
    # the leading term's own chain, f^(-5/3) = exp(-5/3 log f) on 10 Hz .. 4 kHz
    # (the reference's formula): log f's rounding times 5/3 |log f| <= 14
    # ulp of the argument, so within 16 ulp
    f = rng.uniform(10.0, 4096.0, 100000)
    m53 = oracle.libm("exp", -5.0 / 3.0 * oracle.libm("log", f))
    exact = np.exp(np.longdouble(-5) / 3 * np.log(f.astype(np.longdouble)))
    assert np.max(np.abs((m53 - exact) / exact)) < 16 * np.finfo(float).eps


def test_special_values():
    inf, nan = np.inf, np.nan
    e = oracle.libm("exp", np.array([0.0, 1.0, -1.0, 710.0, -746.0, inf, -inf, nan, 709.7, -744.0]))
    assert e[0] == 1.0 and abs(e[1] - np.e) <= np.spacing(np.e) and abs(e[2] - np.exp(-1.0)) <= np.spacing(np.exp(-1.0))
    assert e[3] == inf and e[4] == 0.0 and e[5] == inf and e[6] == 0.0 and np.isnan(e[7])
    assert np.isfinite(e[8]) and 0.0 < e[9] < 1e-300  # subnormal results by exact scaling
    lg = oracle.libm("log", np.array([1.0, 2.0, 0.5, 0.0, -1.0, inf, nan, 1e-310, np.e]))
    assert lg[0] == 0.0 and lg[1] == np.log(2.0) and lg[2] == -np.log(2.0)
    assert lg[3] == -inf and np.isnan(lg[4]) and lg[5] == inf and np.isnan(lg[6])
    assert abs(lg[7] - np.log(1e-310)) < 1e-13 and abs(lg[8] - 1.0) <= 2 * np.finfo(float).eps
    c = oracle.libm("cbrt", np.array([27.0, -8.0, 0.0, -0.0, inf, -inf, nan, 1e-300]))
    assert c[0] == 3.0 and c[1] == -2.0 and c[2] == 0.0 and np.signbit(c[3]) and c[4] == inf and c[5] == -inf
    assert np.isnan(c[6]) and abs(c[7] / 1e-100 - 1.0) < 2 * np.finfo(float).eps
    # exact cubes
    k = np.arange(1.0, 2000.0)
    assert np.array_equal(oracle.libm("cbrt", k ** 3), k)
