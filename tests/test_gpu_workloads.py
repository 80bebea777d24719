"""The ragged and gradient workloads at their full sizes, through the same
generators the benchmark uses (bench.ragged_grids, bench.draw):

* 16,384 binaries each on its own uniform grid (f_lower ~ U[10, 30] Hz,
  T ~ log-U[4, 64] s, up to ~1.3e5 bins, 3.5e8 bins in all): the
  graph-replayed persistent launch equals the eager launch bit for bit on
  two replays with eager work between them, every row is finite, the split
  arithmetic is within a mismatch of 1e-6 of FP64 on every row (at the same
  FP32-rounded grid), and FP64 matches the oracle to 1e-12 on 32 rows
  spanning the lengths;
* the fused loss and gradient at 8192 binaries x 4017 bins: the split
  gradient within the reference's float32 bar of the FP64 one on every
  binary, the FP64 gradient within 1e-8 of the oracle's forward-mode
  gradient on 32 rows and of central differences of the oracle's loss on
  four.
"""

import numpy as np
import pytest

import bench
from conftest import compare, psd
from oracle import oracle

pytestmark = pytest.mark.gpu


def per_row(values, lengths):
    """Each row's sum of a flat ragged array (a segmented reduction)."""
    import torch

    return torch.segment_reduce(values, "sum", lengths=lengths)


def test_ragged_workload(cuda):
    import torch

    from waveformx_amd import xas
    from waveformx_amd.ragged import RaggedPlan

    p, grids = bench.ragged_grids(16384)
    nb = len(grids)
    lengths = np.array([g[2] for g in grids])
    print(f"ragged workload: {nb} binaries, {lengths.sum()} bins, lengths {lengths.min()}..{lengths.max()}",
          flush=True)
    import time

    start = time.perf_counter()
    plan = RaggedPlan(grids, precision="split", tune=False, bins_per_thread=8, blocks=16384)
    params = torch.as_tensor(p, device=cuda)
    first = plan(params)[0].clone()
    records, _ = xas.setup(params, plan.f_min, 1e21, None, "split")
    eager_p, eager_c = xas.evaluate_ragged(records, plan.offsets, grid_f0=plan.grid_f0, grid_df=plan.grid_df,
                                           precision="split", bins_per_thread=1)
    torch.cuda.synchronize()
    hp, hc = plan(params)
    torch.cuda.synchronize()
    assert torch.equal(first.view(torch.int32), eager_p.view(torch.int32))
    assert torch.equal(hp.view(torch.int32), eager_p.view(torch.int32))
    assert torch.equal(hc.view(torch.int32), eager_c.view(torch.int32))
    assert bool(torch.isfinite(torch.view_as_real(hp)).all() and torch.isfinite(torch.view_as_real(hc)).all())

    # FP64 at the split path's own (FP32-rounded) frequencies, same setup inputs
    offsets = plan.offsets
    lengths_t = torch.as_tensor(lengths, device=cuda)
    segment = torch.repeat_interleave(torch.arange(nb, device=cuda), lengths_t)
    k = torch.arange(plan.total, device=cuda) - offsets[:-1][segment]
    f64 = plan.grid_f0[segment] + k.to(torch.float64) * plan.grid_df[segment]
    f_split = f64.to(torch.float32).to(torch.float64)
    del k
    rec64, _ = xas.setup(params, plan.f_min, 1e21, None, "fp64")
    ref_p, ref_c = xas.evaluate_ragged(rec64, offsets, freqs=f_split, precision="fp64")
    weights = 1.0 / ((f_split / 100.0) ** -4 + 1.0 + (f_split / 100.0) ** 2)
    for split, ref in ((hp, ref_p), (hc, ref_c)):
        a = split.to(torch.complex128)
        for w in (None, weights):
            scale = 1.0 if w is None else w
            ab = per_row((a * ref.conj()).real * scale, lengths_t)
            aa = per_row(a.abs() ** 2 * scale, lengths_t)
            bb = per_row(ref.abs() ** 2 * scale, lengths_t)
            mism = 1.0 - ab / torch.sqrt(aa * bb)
            worst = float(mism.max())
            print(f"split vs fp64 ragged, worst row mismatch {worst} ({time.perf_counter() - start:.1f} s)",
                  flush=True)
            assert worst < 1e-6, worst
    del ref_p, ref_c, f_split, weights, segment

    # FP64 on its generated grids against the oracle, 32 rows over the lengths
    hp64, _ = xas.evaluate_ragged(rec64, offsets, grid_f0=plan.grid_f0, grid_df=plan.grid_df, precision="fp64")
    off = offsets.cpu().numpy()
    order = np.argsort(lengths)
    rows = order[np.linspace(0, nb - 1, 32).astype(int)]
    for b in rows:
        f0, df, n = grids[b]
        op, _ = oracle.polarizations(p[b:b + 1], f0 + np.arange(n) * df)
        err = compare(hp64[off[b]:off[b + 1]].cpu().numpy(), op[0])[0]
        assert err < 1e-12, (b, n, err)
    print(f"ragged workload done in {time.perf_counter() - start:.1f} s", flush=True)


def test_gradient_workload(cuda):
    import torch

    from waveformx_amd import adjoint, xas

    nb, n = 8192, 4017
    freqs = np.linspace(20.0, 1024.0, n)
    p = bench.draw(nb, seed=3)
    rng = np.random.default_rng(4)
    perturbed = p * (1 + 1e-2 * rng.normal(size=p.shape))
    perturbed[:, 7] = p[:, 7]
    print(f"gradient workload: {nb} x {n}", flush=True)
    params = torch.as_tensor(p, device=cuda)
    f32 = torch.as_tensor(freqs, dtype=torch.float32, device=cuda)
    f_split = f32.to(torch.float64)  # the split path's grid, in FP64
    d_plus, d_cross = xas.polarizations(torch.as_tensor(perturbed, device=cuda), f_split, 1e21, None, "fp64")
    w64 = torch.as_tensor(1.0 / psd(f_split.cpu().numpy()), device=cuda)
    loss64, g64 = adjoint.loss_and_grad(params, f_split, d_plus, d_cross, w64, precision="fp64")
    loss32, g32 = adjoint.loss_and_grad(params, f32, d_plus.to(torch.complex64), d_cross.to(torch.complex64),
                                        w64.to(torch.float32), precision="split")
    torch.cuda.synchronize()
    g64n, g32n = g64.cpu().numpy()[:, :7], g32.cpu().numpy()[:, :7]
    assert np.all(np.isfinite(g64n)) and np.all(np.isfinite(g32n))
    # the reference's float32 gradient bar (tests/test_xas.py:157-171): 1e-2 |g| + 1e-3 max |g|
    bar = 1e-2 * np.abs(g64n) + 1e-3 * np.abs(g64n).max(axis=0)
    excess = np.abs(g32n - g64n) / bar
    print("split gradient, worst fraction of the float32 bar", excess.max(), flush=True)
    assert np.all(excess <= 1.0), np.unravel_index(excess.argmax(), excess.shape)
    # the loss of residuals at 1% perturbed parameters, from FP32 phases good to ~3e-4 rad
    loss_err = np.abs(loss32.cpu().numpy() / loss64.cpu().numpy() - 1.0).max()
    print("split loss, worst relative error", loss_err, flush=True)
    assert loss_err < 1e-2

    # FP64 against the oracle's forward-mode gradient on 32 rows
    rows = np.linspace(0, nb - 1, 32).astype(int)
    f = f_split.cpu().numpy()
    dp, dc = d_plus[rows].cpu().numpy(), d_cross[rows].cpu().numpy()
    wts = w64.cpu().numpy()
    hp, hc = oracle.polarizations(p[rows], f)
    jp, jc = oracle.jacobian(p[rows], f)
    analytic = 2 * (np.einsum("bf,bfk->bk", np.conj(hp - dp) * wts, jp).real
                    + np.einsum("bf,bfk->bk", np.conj(hc - dc) * wts, jc).real)[:, :7]
    got = g64n[rows]
    err = np.abs(got - analytic) / np.abs(analytic).max(axis=0)
    print("fp64 gradient vs the oracle, worst", err.max(), flush=True)
    assert err.max() < 1e-8, err.max()

    # and central differences of the oracle's loss (step 1e-6 relative) on four
    def oracle_loss(q, r):
        a, c = oracle.polarizations(q, f)
        return np.sum((np.abs(a - dp[r]) ** 2 + np.abs(c - dc[r]) ** 2) * wts, axis=1)

    few = np.arange(4)
    base = p[rows[few]]
    fd = np.zeros((4, 7))
    for j in range(7):
        h = 1e-6 * np.maximum(np.abs(base[:, j]), 1.0)
        a, b = base.copy(), base.copy()
        a[:, j] += h
        b[:, j] -= h
        fd[:, j] = (oracle_loss(a, few) - oracle_loss(b, few)) / (2 * h)
    assert np.allclose(got[few], fd, rtol=1e-4, atol=1e-6 * np.abs(fd).max()), (got[few], fd)
