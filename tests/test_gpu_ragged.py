"""Ragged batches in one persistent launch, captured in a HIP graph: every
replay (with eager work and synchronisations between replays, new
parameters each time) against the eager evaluation and the oracle."""

import numpy as np
import pytest

from conftest import compare
from oracle import oracle

pytestmark = pytest.mark.gpu


def grids_for(params, rng):
    grids = []
    for p in params:
        f_lower = rng.uniform(10.0, 30.0)
        df = 1.0 / np.exp(rng.uniform(np.log(4.0), np.log(32.0)))
        f_max = min(2048.0, 0.3 / ((p[0] + p[1]) * 4.925490947641267e-06))
        grids.append((f_lower, df, int(np.floor((f_max - f_lower) / df + 0.5)) + 1))
    return grids


@pytest.mark.parametrize("precision", ["split", "fp64"])
def test_graph_replays_match_eager_and_oracle(cuda, precision):
    import torch

    from waveformx_amd import xas
    from waveformx_amd.ragged import RaggedPlan

    rng = np.random.default_rng(21)
    params = oracle.draw(48, seed=8, m1=(5.0, 100.0))
    grids = grids_for(params, rng)
    plan = RaggedPlan(grids, precision=precision, tune=False, bins_per_thread=2, blocks=512)
    off = plan.offsets.cpu().numpy()
    for replay in range(3):
        p = params.copy()
        p[:, 6] += 0.5 * replay  # new phases each replay
        pt = torch.as_tensor(p, device=cuda)
        hp, hc = plan(pt)
        hp = hp.clone()
        # eager work between replays
        records, _ = xas.setup(pt, plan.f_min, 1e21, None, precision)
        ep, ec = xas.evaluate_ragged(records, plan.offsets, grid_f0=plan.grid_f0, grid_df=plan.grid_df,
                                     precision=precision, bins_per_thread=1)
        torch.cuda.synchronize()
        assert torch.equal(hp, ep), replay
        hp_np = hp.cpu().numpy().astype(np.complex128)
        for b in range(0, 48, 7):
            f0, df, n = grids[b]
            f = f0 + np.arange(n) * df
            pb = p[b:b + 1]
            if precision != "fp64":
                f = f.astype(np.float32).astype(np.float64)
                pb = pb.astype(np.float32).astype(np.float64)
            op, _ = oracle.polarizations(pb, f)
            err, mism = compare(hp_np[off[b]:off[b + 1]], op[0])
            if precision == "fp64":
                assert err < 1e-12, (b, err)
            else:
                assert mism < 1e-6, (b, mism)


def test_tuner_and_shapes(cuda):
    import torch

    from waveformx_amd.ragged import RaggedPlan

    rng = np.random.default_rng(3)
    params = oracle.draw(64, seed=9, m1=(5.0, 100.0))
    grids = grids_for(params, rng)
    pt = torch.as_tensor(params, device=cuda)
    reference = None
    for bpt in (1, 2, 4, 8, 16):
        for blocks in (256, 2048, 16384):
            plan = RaggedPlan(grids, precision="split", tune=False, graph=False, bins_per_thread=bpt, blocks=blocks)
            hp, _ = plan(pt)
            if reference is None:
                reference = hp.clone()
            assert torch.equal(hp, reference), (bpt, blocks)
    tuned = RaggedPlan(grids, precision="split", tune=True)
    assert tuned.bins_per_thread in (2, 4, 8, 16) and tuned.blocks > 0
    hp, _ = tuned(pt)
    torch.cuda.synchronize()
    assert torch.equal(hp, reference)


@pytest.mark.parametrize("bpt", [1, 8])
def test_many_binaries_with_empty_ranges_match_the_shared_grid(cuda, bpt):
    """5,000 binaries on one grid, every seventh with an empty range: the
    persistent launch's owner search (three rounds at this size) gives each
    work item its binary, and the ragged evaluation writes the shared-grid
    evaluation's bits."""
    import torch

    from waveformx_amd import xas

    nb, n = 5000, 700
    params = torch.as_tensor(oracle.draw(nb, seed=13), device=cuda)
    f = torch.as_tensor(np.linspace(20.0, 900.0, n), dtype=torch.float32, device=cuda)
    records, _ = xas.setup(params, f[:1], 1e21, None, "split")
    shared_p, shared_c = xas.evaluate(records, f, "split")
    lengths = np.where(np.arange(nb) % 7 == 3, 0, n)
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    keep = torch.as_tensor(lengths > 0, device=cuda)
    flat = f.repeat(int(keep.sum()))
    hp, hc = xas.evaluate_ragged(records, torch.as_tensor(offsets, device=cuda), freqs=flat, precision="split",
                                 bins_per_thread=bpt, blocks=1024)
    torch.cuda.synchronize()
    assert torch.equal(hp.view(torch.int32), shared_p[keep].reshape(-1).view(torch.int32))
    assert torch.equal(hc.view(torch.int32), shared_c[keep].reshape(-1).view(torch.int32))


@pytest.mark.parametrize("approximant", ["IMRPhenomXAS", "IMRPhenomD"])
def test_bad_descriptors_and_offsets_raise_before_any_launch(cuda, approximant):
    """A negative length, a non-positive f0 or df, or decreasing offsets
    raise ValueError on the host: no kernel ever sees them."""
    import torch

    import waveformx_amd as wx
    from waveformx_amd import phenomd, xas

    params = oracle.draw(3, seed=2)
    table = {"mass_1": params[:, 0], "mass_2": params[:, 1], "spin_1": np.zeros((3, 3)), "spin_2": np.zeros((3, 3))}
    for grids in ([(20.0, 0.25, 100), (20.0, 0.25, -50), (20.0, 0.25, 10)],
                  [(20.0, 0.0, 100), (20.0, 0.25, 50), (20.0, 0.25, 10)],
                  [(0.0, 0.25, 100), (20.0, 0.25, 50), (20.0, 0.25, 10)]):
        with pytest.raises(ValueError, match="f0 > 0, df > 0 and n >= 0"):
            wx.ragged_polarizations(approximant, grids, table)
    module = xas if approximant == "IMRPhenomXAS" else phenomd
    pt = torch.as_tensor(params, device=cuda)
    f = torch.full((1,), 20.0, dtype=torch.float64, device=cuda)
    records, _ = module.setup(pt, f, 1e21, None, "split")
    offsets = torch.as_tensor([0, 100, 50, 60], dtype=torch.int64, device=cuda)
    with pytest.raises(ValueError, match="start at 0 and never decrease"):
        module.evaluate_ragged(records, offsets, freqs=torch.full((60,), 30.0, device=cuda), precision="split")
