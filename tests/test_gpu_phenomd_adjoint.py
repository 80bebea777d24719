"""IMRPhenomD's hand-written adjoint against fourth-order central
differences of its FP64 oracle (``phenomd_oracle.jacobian``, itself pinned
by closed forms in test_oracle_phenomd.py): the FP64 path to 1e-7 of each
parameter's largest gradient, the split FP32 path to the reference's
float32 gradient bar ``|g32 - g64| <= 1e-2 |g64| + 1e-3 max|g64|``
(tests/test_transforms_xas.py:58-63), the torch backward (parameters and
grid), the grid's gradient, the fused noise-weighted loss, the mass-scaled
reference frequency, and the records the derivative setup writes."""

import numpy as np
import pytest

from conftest import psd
from oracle import phenomd_oracle as pd

pytestmark = pytest.mark.gpu

GRID = np.linspace(20.0, 1024.0, 1024)


def oracle_vjp(params, freqs, w, reference_mf=None):
    """sum_f (w0 Re dh+ + w1 Im dh+ + w2 Re dhx + w3 Im dhx) per parameter."""
    dhp, dhc = pd.jacobian(params, freqs, reference_mf=reference_mf)
    return (np.einsum("bf,bfk->bk", w[0], dhp.real) + np.einsum("bf,bfk->bk", w[1], dhp.imag)
            + np.einsum("bf,bfk->bk", w[2], dhc.real) + np.einsum("bf,bfk->bk", w[3], dhc.imag))


def device_vjp(cuda, params, freqs, w, precision, reference_mf=None):
    import torch

    from waveformx_amd import adjoint

    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    rtype = torch.float64 if precision == "fp64" else torch.float32
    p = torch.as_tensor(params, device=cuda)
    f = torch.as_tensor(freqs, dtype=rtype, device=cuda)
    gp = torch.as_tensor(w[0] + 1j * w[1], dtype=ctype, device=cuda)
    gc = torch.as_tensor(w[2] + 1j * w[3], dtype=ctype, device=cuda)
    grads, _ = adjoint.vjp(p, f, gp, gc, 1e21, reference_mf, precision, approximant="IMRPhenomD")
    return grads.cpu().numpy()


def relative_to_column(got, expected):
    scale = np.max(np.abs(expected), axis=0)
    return np.abs(got - expected) / np.where(scale > 0, scale, 1.0)


@pytest.mark.parametrize("reference_mf", [None, 0.002, 0.05])
def test_fp64_adjoint_matches_differences(cuda, reference_mf):
    """FP64 gradients of a random linear functional, for f_ref in Hz and for
    a mass-scaled reference frequency in the inspiral (0.002) and in the
    merger-ringdown (0.05)."""
    params = pd.draw(16, seed=4)
    if reference_mf is not None:
        params[:, 7] = 0.0
    w = np.random.default_rng(5).normal(size=(4, 16, len(GRID)))
    g = device_vjp(cuda, params, GRID, w, "fp64", reference_mf)
    expected = oracle_vjp(params, GRID, w, reference_mf)
    cols = slice(0, 8) if reference_mf is None else slice(0, 7)
    err = relative_to_column(g[:, cols], expected[:, cols])
    print("fp64 adjoint max rel err per parameter", err.max(axis=0))
    assert np.all(err < 1e-7), err.max(axis=0)
    if reference_mf is not None:
        assert np.all(g[:, 7] == 0.0)  # f_ref = 0 selects the mass-scaled reference
        # given with f_ref (NaN values), still no derivative with respect to
        # it (tests/test_transforms_phenomd.py:91-109 of the reference)
        bad = params.copy()
        bad[:, 7] = 20.0
        assert np.all(device_vjp(cuda, bad, GRID, w, "fp64", reference_mf)[:, 7] == 0.0)


def test_split_adjoint_within_float32_bar(cuda):
    params = pd.draw(16, seed=4)
    rounded = params.astype(np.float32).astype(np.float64)
    f = GRID.astype(np.float32).astype(np.float64)
    w = np.random.default_rng(5).normal(size=(4, 16, len(GRID)))
    g = device_vjp(cuda, params, f, w, "split")
    expected = oracle_vjp(rounded, f, w)
    bar = 1e-2 * np.abs(expected) + 1e-3 * np.max(np.abs(expected), axis=0)
    assert np.all(np.abs(g - expected) <= bar), np.max(np.abs(g - expected) / bar, axis=0)


def test_empty_grid_gives_zero_gradients(cuda):
    import torch

    from waveformx_amd import adjoint

    p = torch.as_tensor(pd.draw(3, seed=1), device=cuda)
    f = torch.tensor([20.0], dtype=torch.float64, device=cuda)
    records, jac = adjoint._records(p, f, 1e21, None, "fp64", "IMRPhenomD")
    empty = torch.empty((3, 0), dtype=torch.complex128, device=cuda)
    grads = adjoint.adjoint(records, jac, f[:0], empty, empty, "fp64", approximant="IMRPhenomD")
    assert torch.equal(grads, torch.zeros_like(grads))


def test_torch_backward(cuda):
    """loss.backward() through the torch function reaches every field and
    the grid, as the reference's tests/test_torch.py:58-96."""
    import torch

    import waveformx_amd.torch as wt
    from waveformx_amd import Binary

    rows = np.array([[36.0, 29.0, 0.0, 0.0, 0.3, 0.0, 0.0, -0.4, 500.0, 0.7, 1.2, 25.0],
                     [50.0, 8.0, 0.0, 0.0, 0.9, 0.0, 0.0, 0.3, 300.0, 2.1, 0.3, 30.0],
                     [150.0, 100.0, 0.0, 0.0, -0.5, 0.0, 0.0, 0.85, 3000.0, 0.4, 2.0, 0.0]])
    freqs = np.geomspace(20.0, 900.0, 300)
    fields = [torch.tensor(rows[:, 0]), torch.tensor(rows[:, 1]), torch.tensor(rows[:, 2:5]),
              torch.tensor(rows[:, 5:8]), torch.tensor(rows[:, 8]), torch.tensor(rows[:, 9]),
              torch.tensor(rows[:, 10]), torch.tensor(rows[:, 11])]
    fields = [x.to(cuda).requires_grad_(True) for x in fields]
    f = wt.polarizations_function("IMRPhenomD")
    hp, hc = f(Binary(*fields), torch.tensor(freqs, device=cuda))
    w = np.random.default_rng(5).normal(size=(4, 3, len(freqs)))
    wt_ = torch.tensor(w, device=cuda)
    loss = (wt_[0] * hp.real + wt_[1] * hp.imag + wt_[2] * hc.real + wt_[3] * hc.imag).sum()
    grads = torch.autograd.grad(loss, fields)
    params = rows[:, [0, 1, 4, 7, 8, 9, 10, 11]]
    expected = oracle_vjp(params, freqs, w)
    got = np.stack([grads[0].cpu().numpy(), grads[1].cpu().numpy(), grads[2][:, 2].cpu().numpy(),
                    grads[3][:, 2].cpu().numpy(), grads[4].cpu().numpy(), grads[5].cpu().numpy(),
                    grads[6].cpu().numpy(), grads[7].cpu().numpy()], axis=1)
    assert np.all(grads[2][:, :2].cpu().numpy() == 0)
    # the third binary's f_ref = 0 takes the grid's first frequency: no
    # gradient (and no finite difference: f_ref < 0 is outside the model)
    assert got[2, 7] == 0.0
    expected[2, 7] = 0.0
    err = relative_to_column(got, expected)
    assert np.all(err < 1e-7), err.max(axis=0)

    # the grid's gradient through the torch function: that of the linear
    # functional, against central differences in each frequency
    grid = torch.tensor(freqs, device=cuda, requires_grad=True)
    hp, hc = f(Binary(*[x.detach() for x in fields]), grid)
    loss = (wt_[0] * hp.real + wt_[1] * hp.imag + wt_[2] * hc.real + wt_[3] * hc.imag).sum()
    (dldf,) = torch.autograd.grad(loss, grid)
    expected = grid_differences(params, freqs, w)
    scale = np.max(np.abs(expected))
    assert np.max(np.abs(dldf.cpu().numpy() - expected)) / scale < 1e-6


def grid_differences(params, f, w):
    """sum over binaries of w . dh/df by central differences of the oracle's
    formula in each frequency, the tables held at the first frequency (the
    reference's stop_gradient on f_min)."""
    fields = pd.table(params, f[0])
    eps = 1e-6 * f
    hp_a, hc_a = pd.evaluate(fields, f + eps)
    hp_b, hc_b = pd.evaluate(fields, f - eps)
    dhp, dhc = (hp_a - hp_b) / (2 * eps), (hc_a - hc_b) / (2 * eps)
    return np.sum(w[0] * dhp.real + w[1] * dhp.imag + w[2] * dhc.real + w[3] * dhc.imag, axis=0)


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_grid_gradient(cuda, precision):
    """dL/df of the shared grid (IMRPhenomD's grid adjoint) against central
    differences of the oracle in each frequency."""
    import torch

    from waveformx_amd import adjoint

    params = pd.draw(6, seed=12)
    f = np.geomspace(20.0, 1500.0, 700)
    if precision != "fp64":
        params = params.astype(np.float32).astype(np.float64)
        f = f.astype(np.float32).astype(np.float64)
    w = np.random.default_rng(2).normal(size=(4, 6, len(f)))
    expected = grid_differences(params, f, w)
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    rtype = torch.float64 if precision == "fp64" else torch.float32
    gp = torch.as_tensor(w[0] + 1j * w[1], dtype=ctype, device=cuda)
    gc = torch.as_tensor(w[2] + 1j * w[3], dtype=ctype, device=cuda)
    _, grid = adjoint.vjp(torch.as_tensor(params, device=cuda), torch.as_tensor(f, dtype=rtype, device=cuda), gp, gc,
                          1e21, None, precision, wrt_grid=True, approximant="IMRPhenomD")
    grid = grid.cpu().numpy()
    scale = np.max(np.abs(expected))
    tol = 1e-6 if precision == "fp64" else 2e-3
    assert np.max(np.abs(grid - expected)) / scale < tol, np.max(np.abs(grid - expected)) / scale


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_fused_loss_gradient(cuda, precision):
    """L = sum_f (|h+ - d+|^2 + |hx - dx|^2) / S(f), d at perturbed
    parameters: the fused kernel's loss against the oracle's, its gradient
    against central differences of the oracle's loss."""
    import torch

    from waveformx_amd import adjoint

    rng = np.random.default_rng(9)
    params = pd.draw(8, seed=6)
    perturbed = params * (1 + 1e-2 * rng.normal(size=params.shape))
    perturbed[:, 7] = params[:, 7]
    if precision != "fp64":
        params = params.astype(np.float32).astype(np.float64)
    f = GRID if precision == "fp64" else GRID.astype(np.float32).astype(np.float64)
    dp, dc = pd.polarizations(perturbed, f)
    wts = 1.0 / psd(f)

    def oracle_loss(p):
        hp, hc = pd.polarizations(p, f)
        return np.sum((np.abs(hp - dp) ** 2 + np.abs(hc - dc) ** 2) * wts, axis=1)

    hp, hc = pd.polarizations(params, f)
    jp, jc = pd.jacobian(params, f)
    analytic = 2 * (np.einsum("bf,bfk->bk", (np.conj(hp - dp) * wts), jp).real
                    + np.einsum("bf,bfk->bk", (np.conj(hc - dc) * wts), jc).real)
    rtype = torch.float64 if precision == "fp64" else torch.float32
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    loss, grads = adjoint.loss_and_grad(torch.as_tensor(params, device=cuda),
                                        torch.as_tensor(f, dtype=rtype, device=cuda),
                                        torch.as_tensor(dp, dtype=ctype, device=cuda),
                                        torch.as_tensor(dc, dtype=ctype, device=cuda),
                                        torch.as_tensor(wts, dtype=rtype, device=cuda), precision=precision,
                                        approximant="IMRPhenomD")
    loss, grads = loss.cpu().numpy(), grads.cpu().numpy()
    expected_loss = oracle_loss(params)
    if precision == "fp64":
        assert np.allclose(loss, expected_loss, rtol=1e-10)
        assert np.all(relative_to_column(grads, analytic) < 1e-7)
    else:
        assert np.allclose(loss, expected_loss, rtol=1e-4)
        bar = 1e-2 * np.abs(analytic) + 1e-3 * np.max(np.abs(analytic), axis=0)
        assert np.all(np.abs(grads - analytic) <= bar)


@pytest.mark.parametrize("precision", ["fp64", "split", "split-df"])
def test_jacobian_setup_writes_the_forward_records(cuda, precision):
    """The record the derivative setup writes beside the Jacobian is the
    forward setup's record, bit for bit."""
    import torch

    from waveformx_amd import phenomd

    params = torch.as_tensor(pd.draw(200, seed=8), device=cuda)
    f_min = torch.tensor([20.0], dtype=torch.float64, device=cuda)
    expected, flags = phenomd.setup(params, f_min, 1e21, None, precision)
    records, jac, jflags = phenomd.setup_jvp(params, f_min, 1e21, None, precision)
    view = torch.int64 if precision == "fp64" else torch.int32
    assert records.dtype == expected.dtype
    assert torch.equal(records.view(view), expected.view(view))
    assert torch.equal(flags, jflags)
    assert jac.shape == (200, 64, 8) and bool(torch.isfinite(jac[:, :60]).all())


def test_fp32_arithmetic_is_refused(cuda):
    import torch

    from waveformx_amd import adjoint

    p = torch.as_tensor(pd.draw(2, seed=1), device=cuda)
    f = torch.as_tensor(GRID, dtype=torch.float32, device=cuda)
    cot = torch.zeros((2, len(GRID)), dtype=torch.complex64, device=cuda)
    with pytest.raises(ValueError):
        adjoint.vjp(p, f, cot, cot, 1e21, None, "fp32", approximant="IMRPhenomD")
