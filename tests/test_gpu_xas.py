"""The HIP path against the oracle: FP64 to 1e-12 (relative to the row's
largest value), the split-precision FP32 paths to a mismatch of 1e-6 at the
FP32-rounded inputs (the reference's own float32 bar, tests/test_xas.py:76-87),
NaN rows where LAL fails, zeros above the cutoff, ragged grids."""

import numpy as np
import pytest

from conftest import CASES, case_grid, case_params, compare, psd, weighted_mismatch
from oracle import oracle

pytestmark = pytest.mark.gpu

CONFIG0_GRID = np.linspace(20.0, 1024.0, 4017)


@pytest.fixture(scope="module")
def xas(cuda):
    from waveformx_amd import xas as module

    module._check_library()
    return module


def run(xas, cuda, params, freqs, precision, amp_scale=1e21, reference_mf=None):
    import torch

    p = torch.as_tensor(np.asarray(params, dtype=np.float64), device=cuda)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    f = torch.as_tensor(freqs, dtype=dtype, device=cuda)
    hp, hc = xas.polarizations(p, f, amp_scale, reference_mf, precision)
    torch.cuda.synchronize()
    return hp.cpu().numpy().astype(np.complex128), hc.cpu().numpy().astype(np.complex128)


def rowwise_error(test, reference):
    scale = np.max(np.abs(reference), axis=1)
    return np.max(np.abs(test - reference), axis=1) / scale


def test_fp64_config0_matches_oracle(xas, cuda):
    """64 seeded binaries on linspace(20, 1024, 4017): FP64 GPU vs the
    oracle to 1e-12 relative."""
    params = oracle.draw(64, seed=0)
    hp, hc = run(xas, cuda, params, CONFIG0_GRID, "fp64")
    op, oc = oracle.polarizations(params, CONFIG0_GRID)
    assert np.all(np.isfinite(hp)) and np.all(np.isfinite(hc))
    err = max(rowwise_error(hp, op).max(), rowwise_error(hc, oc).max())
    print("fp64 config0 max rel err", err)
    assert err < 1e-12, err


@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_fp64_reference_cases(xas, cuda, case):
    """The reference's cases, 1e-12 relative: also the binaries of a few
    solar masses, whose leading phase term T0 f^(-5/3) carries ~1e4 rad
    (its exp, log and cbrt are the same source on both sides)."""
    f = case_grid(case)
    p = case_params(case)[None]
    hp, hc = run(xas, cuda, p, f, "fp64")
    op, oc = oracle.polarizations(p, f)
    for a, b in ((hp, op), (hc, oc)):
        error, mismatch = compare(a[0], b[0])
        print("fp64 case error", error)
        assert error < 1e-12, error
        assert mismatch < 1e-14, mismatch


@pytest.mark.parametrize("precision", ["split", "split-df"])
def test_split_config0_mismatch(xas, cuda, precision):
    """FP32 evaluation with the leading term split: mismatch < 1e-6 against
    the FP64 oracle at the FP32-rounded parameters and grid, uniform and
    noise-weighted."""
    params = oracle.draw(64, seed=0).astype(np.float32).astype(np.float64)
    f = CONFIG0_GRID.astype(np.float32).astype(np.float64)
    hp, hc = run(xas, cuda, params, f, precision)
    op, oc = oracle.polarizations(params, f)
    w = psd(f)
    worst = 0.0
    for a, b in ((hp, op), (hc, oc)):
        for i in range(len(params)):
            _, m = compare(a[i], b[i])
            worst = max(worst, m)
        worst = max(worst, weighted_mismatch(a, b, 1.0 / w).max())
    print(precision, "worst mismatch", worst)
    assert worst < 1e-6, worst


@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_split_reference_cases(xas, cuda, case):
    """The reference's float32 test (tests/test_xas.py:76-87) on the split
    path: mismatch < 1e-6 at the float32-rounded inputs."""
    f = case_grid(case, 1000).astype(np.float32).astype(np.float64)
    p = case_params(case).astype(np.float32).astype(np.float64)[None]
    hp, hc = run(xas, cuda, p, f, "split")
    op, oc = oracle.polarizations(p, f)
    for a, b in ((hp, op), (hc, oc)):
        _, mismatch = compare(a[0], b[0])
        assert mismatch < 1e-6, mismatch


def test_split_binary_neutron_star(xas, cuda):
    """A BNS from 20 Hz on a 1/256 Hz grid (half a million bins; 1e4 rad of
    inspiral phase): mismatch < 1e-6 (tests/test_xas.py:90-100)."""
    f = np.float32(20.0 + np.arange((2048 - 20) * 256) / 256.0).astype(np.float64)
    p = np.float32([1.42, 1.27, 0.05, -0.03, 100.0, 0.5, 1.0, 0.0]).astype(np.float64)[None]
    hp, hc = run(xas, cuda, p, f, "split")
    op, oc = oracle.polarizations(p, f)
    for a, b in ((hp, op), (hc, oc)):
        assert compare(a[0], b[0])[1] < 1e-6


def test_plain_fp32_accuracy(xas, cuda):
    """The plain FP32 arithmetic (every term, the leading inspiral phase
    too, in FP32) on the config-1 distribution: finite, and within a
    mismatch of 1e-6 of the oracle at the same float32 frequencies on every
    binary and polarization (measured ~2e-7; its inspiral phase is good to
    ~0.06 rad, so the bar is the reference's float32 mismatch bar, not a
    pointwise one)."""
    params = oracle.draw(64, seed=1)
    f32 = CONFIG0_GRID.astype(np.float32).astype(np.float64)
    hp, hc = run(xas, cuda, params, f32, "fp32")
    assert np.all(np.isfinite(hp)) and np.all(np.isfinite(hc))
    op, oc = oracle.polarizations(params, f32)
    worst = max(compare(a[b], r[b])[1] for a, r in ((hp, op), (hc, oc)) for b in range(len(params)))
    print("plain fp32, worst mismatch", worst)
    assert worst < 1e-6, worst


def test_fast_atan_bound(cuda):
    """fast_atan, the FP32 arctangent of the split and FP32 phases, through
    wfx_xas_fast_atan: within 2e-7 of atan over [0, inf) (a dense grid of
    [0, 4], a geometric one to 1e38, infinity), odd, NaN to NaN, and its
    coefficients' recipe (csrc/gen_atan.py) bounds the same error in float32
    arithmetic."""
    import ctypes

    import torch

    from waveformx_amd._lib import call

    x = np.concatenate([np.linspace(0.0, 4.0, 400001), np.geomspace(1e-38, 1e38, 200001), [np.inf]])
    x32 = torch.as_tensor(np.concatenate([x, -x, [np.nan]]), dtype=torch.float32, device=cuda)
    y = torch.empty_like(x32)
    call("wfx_xas_fast_atan", ctypes.c_void_p(x32.data_ptr()), ctypes.c_void_p(y.data_ptr()), x32.numel(),
         ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream))
    got = y.cpu().numpy().astype(np.float64)
    xs = x32.cpu().numpy().astype(np.float64)
    n = len(x)
    err = np.abs(got[:2 * n] - np.arctan(xs[:2 * n]))
    print("fast_atan, worst |error|", err.max())
    assert err.max() < 2e-7, (err.max(), xs[err.argmax()])
    np.testing.assert_array_equal(got[n:2 * n], -got[:n])
    assert np.isnan(got[-1])


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_nan_where_lal_fails(xas, cuda, precision):
    """q > 1000 gives an all-NaN row, above the cutoff too (xas/core.py:336-340)."""
    f = np.geomspace(20.0, 1024.0, 100)
    p = np.array([[2000.0, 1.0, 0.0, 0.0, 400.0, 0.4, 0.0, 20.0], [36.0, 29.0, 0.3, -0.4, 500.0, 1.0, 1.2, 30.0]])
    hp, hc = run(xas, cuda, p, f, precision)
    assert np.all(np.isnan(hp[0])) and np.all(np.isnan(hc[0]))
    assert np.all(np.isfinite(hp[1]))


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_zero_above_cutoff(xas, cuda, precision):
    """Heavy binaries: exactly zero above f_cut = 0.3 / M, non-zero below."""
    p = np.array([[80.0, 70.0, 0.2, 0.1, 500.0, 0.3, 0.5, 20.0]])
    f_cut = 0.3 / (150.0 * 4.925490947641267e-06)
    f = np.linspace(20.0, 1024.0, 4017)
    hp, _ = run(xas, cuda, p, f, precision)
    assert np.all(hp[0, f > f_cut * 1.0001] == 0)
    assert np.all(np.abs(hp[0, f < f_cut * 0.999]) > 0)


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_ragged_matches_oracle(cuda, precision):
    """Binaries on their own grids (explicit values and generated uniform
    grids) in one launch, against the oracle binary by binary."""
    import torch

    from waveformx_amd import ragged_polarizations

    rng = np.random.default_rng(7)
    params = oracle.draw(12, seed=3)
    table = np.zeros((12, 12))
    table[:, [0, 1, 4, 7, 8, 9, 10, 11]] = params
    grids = []
    for i in range(12):
        f0 = rng.uniform(10, 30)
        df = 1.0 / np.exp(rng.uniform(np.log(4), np.log(16)))
        n = int(rng.integers(50, 3000))
        grids.append((f0, df, n))
    hp, hc, off = ragged_polarizations("IMRPhenomXAS", grids, table, precision=precision)
    hp, hc, off = hp.cpu().numpy(), hc.cpu().numpy(), off.cpu().numpy()
    arrays = [f0 + np.arange(n) * df for f0, df, n in grids]
    hp2, hc2, off2 = ragged_polarizations("IMRPhenomXAS", arrays, table, precision=precision)
    assert np.array_equal(off, off2.cpu().numpy())
    for i, f in enumerate(arrays):
        fr = f if precision == "fp64" else f.astype(np.float32).astype(np.float64)
        pr = params[i:i + 1] if precision == "fp64" else params[i:i + 1].astype(np.float32).astype(np.float64)
        op, oc = oracle.polarizations(pr, fr)
        a = hp[off[i]:off[i + 1]]
        b = hp2.cpu().numpy()[off[i]:off[i + 1]]
        if precision == "fp64":
            assert compare(a, op[0])[0] < 1e-12
            assert compare(b, op[0])[0] < 1e-12
        else:
            assert compare(a, op[0])[1] < 1e-6
            assert compare(b, op[0])[1] < 1e-6
    torch.cuda.synchronize()


@pytest.mark.parametrize("precision", ["split", "split-df", "fp32"])
def test_launch_shapes_agree(xas, cuda, precision):
    """Every launch shape of the FP32 evaluation writes the same bits (the
    group and the frequencies per thread only move work between threads),
    including groups and tiles that overhang the batch and the grid."""
    import torch

    params = torch.as_tensor(oracle.draw(37, seed=9), device=cuda)
    f = torch.as_tensor(np.linspace(20.0, 1024.0, 1500), dtype=torch.float32, device=cuda)
    records, _ = xas.setup(params, f[:1], 1e21, None, precision)
    ref = [t.view(torch.int32).clone() for t in xas.evaluate(records, f, precision, 1, freqs_per_thread=1)]
    for group, fpt in ((0, 0), (1, 2), (3, 4), (8, 8), (16, 1), (64, 4)):
        hp, hc = xas.evaluate(records, f, precision, group, freqs_per_thread=fpt)
        assert torch.equal(hp.view(torch.int32), ref[0]), (group, fpt)
        assert torch.equal(hc.view(torch.int32), ref[1]), (group, fpt)


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_degenerate_sizes(xas, cuda, precision):
    """One binary at one frequency, two frequencies, a group larger than the
    batch, and empty batches and grids (nothing launched, nothing written),
    for both models; every shape against the oracle or the larger call."""
    import torch

    from oracle import phenomd_oracle as pd
    from waveformx_amd import phenomd

    dtype = torch.float64 if precision == "fp64" else torch.float32
    params = oracle.draw(3, seed=31)
    p = torch.as_tensor(params, device=cuda)
    for module, reference in ((xas, oracle), (phenomd, pd)):
        for n in (1, 2, 257):
            f = np.linspace(25.0, 400.0, n)
            ft = torch.as_tensor(f, dtype=dtype, device=cuda)
            records, _ = module.setup(p, ft[:1], 1e21, None, precision)
            full = module.evaluate(records, ft, precision)
            for group in (1, 16):
                hp, hc = module.evaluate(records, ft, precision, group)
                assert torch.equal(hp, full[0]) and torch.equal(hc, full[1])
            one, _ = module.evaluate(records[:1], ft, precision)
            assert torch.equal(one[0], full[0][0])
            op, _ = reference.polarizations(params, f if precision == "fp64" else f.astype(np.float32).astype(float))
            got = full[0].cpu().numpy().astype(np.complex128)
            if precision == "fp64":
                assert np.max(np.abs(got - op)) <= 1e-12 * np.max(np.abs(op))
            else:
                assert np.max(np.abs(got - op)) <= 1e-3 * np.max(np.abs(op))
        ft = torch.as_tensor([30.0, 40.0], dtype=dtype, device=cuda)
        records, _ = module.setup(p, ft[:1], 1e21, None, precision)
        empty_grid = module.evaluate(records, ft[:0], precision)
        assert empty_grid[0].shape == (3, 0)
        empty_batch = module.evaluate(records[:0], ft, precision)
        assert empty_batch[0].shape == (0, 2)
