"""The scratch-free fused loss (wfx_xas_loss_grad_static, through the C ABI)
against the allocating path it shares its kernels with (wfx_xas_setup_jvp,
wfx_xas_power_table, wfx_xas_loss_adjoint): the same bits, on two calls in
a row, the second smaller (the static array keeps the first call's
contents), and sizes beyond the array refused before any launch."""

import numpy as np
import pytest

from conftest import psd
from oracle import oracle

pytestmark = pytest.mark.gpu


def _inputs(cuda, nb, n, precision, seed):
    import torch

    rng = np.random.default_rng(seed)
    params = oracle.draw(nb, seed=seed)
    perturbed = params * (1 + 1e-2 * rng.normal(size=params.shape))
    perturbed[:, 7] = params[:, 7]
    f = np.linspace(20.0, 1024.0, n)
    rtype = torch.float64 if precision == "fp64" else torch.float32
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    from waveformx_amd import xas

    pt = torch.as_tensor(params, device=cuda)
    ft = torch.as_tensor(f, dtype=rtype, device=cuda)
    records, _ = xas.setup(torch.as_tensor(perturbed, device=cuda), ft[:1], 1e21, None, precision)
    dp, dc = xas.evaluate(records, ft, precision)
    wts = torch.as_tensor(1.0 / psd(f), dtype=rtype, device=cuda)
    return pt, ft, dp.to(ctype), dc.to(ctype), wts


@pytest.mark.parametrize("precision", ["fp64", "split", "split-df"])
def test_static_loss_equals_the_allocating_path(cuda, precision):
    import torch

    from waveformx_amd import adjoint

    for nb, n, seed in ((8192, 4017, 1), (300, 1500, 2)):
        print(f"static loss: precision {precision}, {nb} x {n}, seed {seed}", flush=True)
        args = _inputs(cuda, nb, n, precision, seed)
        loss_s, grads_s = adjoint.loss_and_grad_static(*args, precision=precision)
        loss_a, grads_a = adjoint.loss_and_grad(*args, precision=precision)
        torch.cuda.synchronize()
        assert torch.isfinite(loss_s).all() and torch.isfinite(grads_s).all()
        assert torch.equal(loss_s, loss_a), (nb, n)
        assert torch.equal(grads_s, grads_a), (nb, n)


def test_static_loss_refuses_sizes_beyond_its_scratch(cuda):
    import torch

    from waveformx_amd import adjoint

    # 16,384 binaries need 16384 x 4608 B = 72 MiB: refused before any launch
    nb, n = 16384, 16
    params = torch.as_tensor(oracle.draw(nb, seed=3), device=cuda)
    f = torch.linspace(20.0, 40.0, n, dtype=torch.float64, device=cuda)
    data = torch.zeros((nb, n), dtype=torch.complex128, device=cuda)
    wts = torch.ones((n,), dtype=torch.float64, device=cuda)
    with pytest.raises(RuntimeError, match="exceed the 48 MiB scratch"):
        adjoint.loss_and_grad_static(params, f, data, data, wts, precision="fp64")
    with pytest.raises(ValueError, match="fp64, split or split-df"):
        adjoint.loss_and_grad_static(params, f.float(), data.to(torch.complex64), data.to(torch.complex64),
                                     wts.float(), precision="fp32")
