"""The HIP path at the benchmark's full size, 65,536 binaries x 4017 bins
(linspace(20, 1024) Hz, the parameter draw of the oracle), both models,
through
properties that do not depend on the size: every row finite; the
split-precision outputs within the reference's float32 bar (noise-weighted
mismatch <= 1e-6, tests/test_xas.py:76-87) of the FP64 outputs on every
row; the FP64 outputs within 1e-12 of the oracle on rows sampled across the
whole batch; every launch shape writing the same bits; and h ~ 1/d_L exactly
(doubling every distance halves every output bit for bit, since the
distance enters once, as a division, and halving is exact)."""

import numpy as np
import pytest

from conftest import psd
from oracle import oracle

pytestmark = pytest.mark.gpu

NB, N = 65536, 4017
GRID = np.linspace(20.0, 1024.0, N)


@pytest.fixture(scope="module")
def batch(cuda):
    import torch

    from waveformx_amd import xas

    xas._check_library()
    params = oracle.draw(NB, seed=1)
    p = torch.as_tensor(params, device=cuda)
    f32 = torch.as_tensor(GRID, dtype=torch.float32, device=cuda)
    return xas, params, p, f32


def row_mismatch(torch, a, b, w, rows=4096):
    """Noise-weighted mismatch per row, in chunks of rows, in FP64."""
    out = []
    for r0 in range(0, a.shape[0], rows):
        x = a[r0:r0 + rows].to(torch.complex128)
        y = b[r0:r0 + rows].to(torch.complex128)
        xy = (x * y.conj()).real @ w
        xx = (x.abs() ** 2) @ w
        yy = (y.abs() ** 2) @ w
        out.append(1.0 - xy / torch.sqrt(xx * yy))
    return torch.cat(out)


def test_split_within_float32_bar_of_fp64_on_every_row(batch, cuda):
    import torch

    xas, _, p, f32 = batch
    hp32, hc32 = xas.polarizations(p, f32, precision="split")
    hp64, hc64 = xas.polarizations(p, f32.to(torch.float64), precision="fp64")
    assert bool(torch.isfinite(torch.view_as_real(hp32)).all()) and bool(torch.isfinite(torch.view_as_real(hc32)).all())
    assert bool(torch.isfinite(torch.view_as_real(hp64)).all()) and bool(torch.isfinite(torch.view_as_real(hc64)).all())
    w = torch.as_tensor(1.0 / psd(GRID.astype(np.float32).astype(np.float64)), device=cuda)
    worst = max(float(row_mismatch(torch, hp32, hp64, w).max()), float(row_mismatch(torch, hc32, hc64, w).max()))
    print("worst noise-weighted mismatch, split vs FP64, all rows:", worst)
    assert worst <= 1e-6


def test_fp64_matches_oracle_across_the_batch(batch, cuda):
    import torch

    xas, params, p, _ = batch
    f64 = torch.as_tensor(GRID, device=cuda)
    hp, hc = xas.polarizations(p, f64, precision="fp64")
    rows = np.unique(np.linspace(0, NB - 1, 48).astype(np.int64))
    op, oc = oracle.polarizations(params[rows], GRID)
    for got, ref in ((hp, op), (hc, oc)):
        g = got[torch.as_tensor(rows, device=cuda)].cpu().numpy()
        err = np.max(np.abs(g - ref), axis=1) / np.max(np.abs(ref), axis=1)
        assert np.all(err < 1e-12), err.max()


def test_launch_shapes_agree_at_full_size(batch, cuda):
    import torch

    xas, _, p, f32 = batch
    records, _ = xas.setup(p, f32[:1], 1e21, None, "split")
    a = xas.evaluate(records, f32, "split")
    for group, fpt in ((8, 8), (1, 2)):
        b = xas.evaluate(records, f32, "split", group, freqs_per_thread=fpt)
        assert torch.equal(a[0].view(torch.int32), b[0].view(torch.int32)), (group, fpt)
        assert torch.equal(a[1].view(torch.int32), b[1].view(torch.int32)), (group, fpt)
        del b


@pytest.mark.parametrize("precision", ["split", "fp64"])
def test_doubling_the_distance_halves_every_output(batch, cuda, precision):
    import torch

    xas, _, p, f32 = batch
    f = f32 if precision == "split" else f32.to(torch.float64)
    hp, hc = xas.polarizations(p, f, precision=precision)
    far = p.clone()
    far[:, 4] *= 2.0
    hp2, hc2 = xas.polarizations(far, f, precision=precision)
    assert torch.equal(hp2 * 2, hp) and torch.equal(hc2 * 2, hc)


@pytest.fixture(scope="module")
def phenomd_batch(cuda):
    import torch

    from oracle import phenomd_oracle as pd
    from waveformx_amd import phenomd

    params = pd.draw(NB, seed=2)
    return phenomd, pd, params, torch.as_tensor(params, device=cuda), \
        torch.as_tensor(GRID, dtype=torch.float32, device=cuda)


def test_phenomd_at_full_size(phenomd_batch, cuda):
    """IMRPhenomD at the same size: split within the float32 bar of FP64 on
    every row, FP64 within 1e-12 of its oracle on rows across the batch,
    launch shapes bit-identical, 1/d_L exact."""
    import torch

    phenomd, pd, params, p, f32 = phenomd_batch
    hp32, hc32 = phenomd.polarizations(p, f32, precision="split")
    f64r = f32.to(torch.float64)
    hp64, hc64 = phenomd.polarizations(p, f64r, precision="fp64")
    for t in (hp32, hc32, hp64, hc64):
        assert bool(torch.isfinite(torch.view_as_real(t)).all())
    w = torch.as_tensor(1.0 / psd(GRID.astype(np.float32).astype(np.float64)), device=cuda)
    worst = max(float(row_mismatch(torch, hp32, hp64, w).max()), float(row_mismatch(torch, hc32, hc64, w).max()))
    print("IMRPhenomD worst mismatch, split vs FP64, all rows:", worst)
    assert worst <= 1e-6
    del hp32, hc32
    rows = np.unique(np.linspace(0, NB - 1, 48).astype(np.int64))
    op, oc = pd.polarizations(params[rows], f64r.cpu().numpy())
    idx = torch.as_tensor(rows, device=cuda)
    for got, ref in ((hp64, op), (hc64, oc)):
        g = got[idx].cpu().numpy()
        err = np.max(np.abs(g - ref), axis=1) / np.max(np.abs(ref), axis=1)
        assert np.all(err < 1e-12), err.max()
    del hp64, hc64
    records, _ = phenomd.setup(p, f32[:1], 1e21, None, "split")
    a = phenomd.evaluate(records, f32, "split")
    b = phenomd.evaluate(records, f32, "split", 8, freqs_per_thread=8)
    assert torch.equal(a[0].view(torch.int32), b[0].view(torch.int32))
    del b
    far = p.clone()
    far[:, 4] *= 2.0
    hp2, _ = phenomd.polarizations(far, f32, precision="split")
    assert torch.equal(hp2 * 2, a[0])


def test_xhm_at_full_size(cuda):
    """IMRPhenomXHM at the multimode workload's size, 4096 binaries x 8193 bins
    (linspace(20, 2048) Hz): every row finite; the split arithmetic within the
    reference's float32 bar of FP64 (noise-weighted mismatch <= 1e-6) on every
    row; FP64 within 2e-10 of the oracle on rows spread over the batch; 1/d_L
    exact in both arithmetics (FP32: in its normal range)."""
    import torch

    from oracle import xhm as oxhm
    from waveformx_amd import xhm

    nb, n = 4096, 8193
    grid = np.linspace(20.0, 2048.0, n)
    params = oxhm.draw(nb, seed=4)
    p = torch.as_tensor(params, device=cuda)
    f32 = torch.as_tensor(grid, dtype=torch.float32, device=cuda)
    f64r = f32.to(torch.float64)
    hp32, hc32 = xhm.polarizations(p, f32, precision="split")
    hp64, hc64 = xhm.polarizations(p, f64r)
    for t in (hp32, hc32, hp64, hc64):
        assert bool(torch.isfinite(torch.view_as_real(t)).all())
    w = torch.as_tensor(1.0 / psd(grid.astype(np.float32).astype(np.float64)), device=cuda)
    worst = max(float(row_mismatch(torch, hp32, hp64, w).max()), float(row_mismatch(torch, hc32, hc64, w).max()))
    print("IMRPhenomXHM worst mismatch, split vs FP64, all rows:", worst)
    assert worst <= 1e-6
    rows = np.unique(np.linspace(0, nb - 1, 24).astype(np.int64))
    op, oc = oxhm.polarizations(params[rows], f64r.cpu().numpy(), amp_scale=1e21)
    idx = torch.as_tensor(rows, device=cuda)
    for got, ref in ((hp64, op), (hc64, oc)):
        g = got[idx].cpu().numpy()
        err = np.max(np.abs(g - ref), axis=1) / np.max(np.abs(ref), axis=1)
        assert np.all(err < 2e-10), err.max()
    far = p.clone()
    far[:, 4] *= 2.0
    assert torch.equal(xhm.polarizations(far, f64r)[0] * 2, hp64)
    # (split: exact where the outputs are normal FP32 numbers; the higher modes' ringdown tails, evaluated at every
    # frequency, reach FP32's subnormal range, where halving rounds)
    normal = hp32.abs() > 1e-30
    far32 = xhm.polarizations(far, f32, precision="split")[0]
    assert torch.equal((far32 * 2)[normal], hp32[normal])
