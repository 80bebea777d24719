"""The hand-written adjoint against the oracle's forward-mode FP64
derivatives and against central differences of the oracle: the FP64 path
to 1e-8, the split FP32 path to the reference's float32 gradient bar
``|g32 - g64| <= 1e-2 |g64| + 1e-3 max|g64|`` (tests/test_transforms_xas.py:58-63),
the torch backward, and the fused noise-weighted loss."""

import numpy as np
import pytest

from conftest import psd
from oracle import oracle

pytestmark = pytest.mark.gpu

GRID = np.linspace(20.0, 1024.0, 1024)


def oracle_vjp(params, freqs, w):
    """sum_f (w0 Re dh+ + w1 Im dh+ + w2 Re dhx + w3 Im dhx) per parameter."""
    dhp, dhc = oracle.jacobian(params, freqs)
    return (np.einsum("bf,bfk->bk", w[0], dhp.real) + np.einsum("bf,bfk->bk", w[1], dhp.imag)
            + np.einsum("bf,bfk->bk", w[2], dhc.real) + np.einsum("bf,bfk->bk", w[3], dhc.imag))


def device_vjp(cuda, params, freqs, w, precision):
    import torch

    from waveformx_amd import adjoint

    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    rtype = torch.float64 if precision == "fp64" else torch.float32
    p = torch.as_tensor(params, device=cuda)
    f = torch.as_tensor(freqs, dtype=rtype, device=cuda)
    gp = torch.as_tensor(w[0] + 1j * w[1], dtype=ctype, device=cuda)
    gc = torch.as_tensor(w[2] + 1j * w[3], dtype=ctype, device=cuda)
    grads, _ = adjoint.vjp(p, f, gp, gc, 1e21, None, precision)
    return grads.cpu().numpy()


def test_fp64_adjoint_matches_oracle(cuda):
    params = oracle.draw(16, seed=4)
    w = np.random.default_rng(5).normal(size=(4, 16, len(GRID)))
    g = device_vjp(cuda, params, GRID, w, "fp64")
    expected = oracle_vjp(params, GRID, w)
    scale = np.max(np.abs(expected), axis=0)
    err = np.abs(g - expected) / np.where(scale > 0, scale, 1.0)
    print("fp64 adjoint max rel err per parameter", err.max(axis=0))
    assert np.all(err < 1e-8), err.max(axis=0)


def test_split_adjoint_within_float32_bar(cuda):
    params = oracle.draw(16, seed=4)
    rounded = params.astype(np.float32).astype(np.float64)
    f = GRID.astype(np.float32).astype(np.float64)
    w = np.random.default_rng(5).normal(size=(4, 16, len(GRID)))
    g = device_vjp(cuda, params, f, w, "split")
    expected = oracle_vjp(rounded, f, w)
    bar = 1e-2 * np.abs(expected) + 1e-3 * np.max(np.abs(expected), axis=0)
    assert np.all(np.abs(g - expected) <= bar), np.max(np.abs(g - expected) / bar, axis=0)


def test_torch_backward(cuda):
    """loss.backward() through the torch function reaches every field, as
    the reference's tests/test_torch.py:58-96 (gradients vs the FP64
    oracle, rtol 1e-10 here)."""
    import torch

    import waveformx_amd.torch as wt
    from waveformx_amd import Binary

    rows = np.array([[36.0, 29.0, 0.0, 0.0, 0.3, 0.0, 0.0, -0.4, 500.0, 0.7, 1.2, 25.0],
                     [50.0, 8.0, 0.0, 0.0, 0.9, 0.0, 0.0, 0.3, 300.0, 2.1, 0.3, 30.0],
                     [150.0, 100.0, 0.0, 0.0, -0.5, 0.0, 0.0, 0.95, 3000.0, 0.4, 2.0, 0.0]])
    freqs = np.geomspace(20.0, 900.0, 300)
    fields = [torch.tensor(rows[:, 0]), torch.tensor(rows[:, 1]), torch.tensor(rows[:, 2:5]),
              torch.tensor(rows[:, 5:8]), torch.tensor(rows[:, 8]), torch.tensor(rows[:, 9]),
              torch.tensor(rows[:, 10]), torch.tensor(rows[:, 11])]
    fields = [x.to(cuda).requires_grad_(True) for x in fields]
    f = wt.polarizations_function("IMRPhenomXAS")
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
    scale = np.maximum(np.max(np.abs(expected), axis=0), 1e-30)
    assert np.all(np.abs(got - expected) / scale < 1e-8)


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_grid_gradient(cuda, precision):
    """dL/df against central differences of the oracle's formula in each
    frequency, the tables held at the first frequency (the reference's
    stop_gradient on f_min, xas/model.py:59)."""
    import torch

    from waveformx_amd import adjoint

    params = oracle.draw(6, seed=12)
    f = np.geomspace(20.0, 1500.0, 700)
    if precision != "fp64":
        params = params.astype(np.float32).astype(np.float64)
        f = f.astype(np.float32).astype(np.float64)
    w = np.random.default_rng(2).normal(size=(4, 6, len(f)))
    tables, _ = oracle.tables(params, f[0])
    eps = 1e-6 * f
    hp_a, hc_a = oracle.evaluate(tables, f + eps)
    hp_b, hc_b = oracle.evaluate(tables, f - eps)
    dhp, dhc = (hp_a - hp_b) / (2 * eps), (hc_a - hc_b) / (2 * eps)
    expected = np.sum(w[0] * dhp.real + w[1] * dhp.imag + w[2] * dhc.real + w[3] * dhc.imag, axis=0)
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    rtype = torch.float64 if precision == "fp64" else torch.float32
    p = torch.as_tensor(params, device=cuda)
    ft = torch.as_tensor(f, dtype=rtype, device=cuda)
    gp = torch.as_tensor(w[0] + 1j * w[1], dtype=ctype, device=cuda)
    gc = torch.as_tensor(w[2] + 1j * w[3], dtype=ctype, device=cuda)
    _, grid = adjoint.vjp(p, ft, gp, gc, 1e21, None, precision, wrt_grid=True)
    grid = grid.cpu().numpy()
    scale = np.max(np.abs(expected))
    tol = 1e-6 if precision == "fp64" else 2e-3
    assert np.max(np.abs(grid - expected)) / scale < tol, np.max(np.abs(grid - expected)) / scale
    # the sums over the binaries run in a fixed order: every call gives the same bits, also for a batch of
    # 200 binaries (four groups of the kernel's partial sums)
    many = torch.as_tensor(np.tile(params, (34, 1))[:200], device=cuda)
    gpm, gcm = (torch.as_tensor(np.tile(x.cpu().numpy(), (34, 1))[:200], device=cuda) for x in (gp, gc))
    runs = [adjoint.vjp(many, ft, gpm, gcm, 1e21, None, precision, wrt_grid=True)[1].cpu().numpy() for _ in range(3)]
    for run in runs[1:]:
        np.testing.assert_array_equal(run, runs[0])


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_fused_loss_gradient(cuda, precision):
    """L = sum_f (|h+ - d+|^2 + |hx - dx|^2) / S(f), d at perturbed
    parameters: the fused kernel's loss and gradient against the FP64
    oracle (analytic, and central differences step 1e-6 relative)."""
    import torch

    from waveformx_amd import adjoint

    rng = np.random.default_rng(9)
    params = oracle.draw(8, seed=6)
    perturbed = params * (1 + 1e-2 * rng.normal(size=params.shape))
    perturbed[:, 7] = params[:, 7]
    if precision != "fp64":
        params = params.astype(np.float32).astype(np.float64)
    f = GRID if precision == "fp64" else GRID.astype(np.float32).astype(np.float64)
    dp, dc = oracle.polarizations(perturbed, f)
    wts = 1.0 / psd(f)

    def oracle_loss(p):
        hp, hc = oracle.polarizations(p, f)
        return np.sum((np.abs(hp - dp) ** 2 + np.abs(hc - dc) ** 2) * wts, axis=1)

    hp, hc = oracle.polarizations(params, f)
    jp, jc = oracle.jacobian(params, f)
    analytic = 2 * (np.einsum("bf,bfk->bk", (np.conj(hp - dp) * wts), jp).real
                    + np.einsum("bf,bfk->bk", (np.conj(hc - dc) * wts), jc).real)
    rtype = torch.float64 if precision == "fp64" else torch.float32
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    loss, grads = adjoint.loss_and_grad(torch.as_tensor(params, device=cuda),
                                        torch.as_tensor(f, dtype=rtype, device=cuda),
                                        torch.as_tensor(dp, dtype=ctype, device=cuda),
                                        torch.as_tensor(dc, dtype=ctype, device=cuda),
                                        torch.as_tensor(wts, dtype=rtype, device=cuda), precision=precision)
    loss, grads = loss.cpu().numpy(), grads.cpu().numpy()
    expected_loss = oracle_loss(params)
    # central differences of the oracle, step 1e-6 max(|x|, 1)
    fd = np.zeros_like(analytic)
    for k in range(7):
        h = 1e-6 * np.maximum(np.abs(params[:, k]), 1.0)
        a, b = params.copy(), params.copy()
        a[:, k] += h
        b[:, k] -= h
        fd[:, k] = (oracle_loss(a) - oracle_loss(b)) / (2 * h)
    assert np.allclose(fd[:, :7], analytic[:, :7], rtol=1e-4, atol=1e-6 * np.abs(analytic[:, :7]).max())
    if precision == "fp64":
        assert np.allclose(loss, expected_loss, rtol=1e-10)
        assert np.allclose(grads[:, :7], analytic[:, :7], rtol=1e-7, atol=1e-9 * np.abs(analytic).max())
    else:
        assert np.allclose(loss, expected_loss, rtol=1e-4)
        bar = 1e-2 * np.abs(analytic[:, :7]) + 1e-3 * np.max(np.abs(analytic[:, :7]), axis=0)
        assert np.all(np.abs(grads[:, :7] - analytic[:, :7]) <= bar)


@pytest.mark.parametrize("precision", ["fp64", "split", "split-df"])
def test_jacobian_setup_writes_the_forward_records(cuda, precision):
    """The record that the derivative setup writes alongside the Jacobian is
    the forward setup's record, bit for bit (FP64, and the FP32 split record
    with the anchor folded into the post-inspiral constants)."""
    import torch

    from waveformx_amd import xas

    params = torch.as_tensor(oracle.draw(200, seed=8), device=cuda)
    f_min = torch.tensor([20.0], dtype=torch.float64, device=cuda)
    expected, flags = xas.setup(params, f_min, 1e21, None, precision)
    records, jac, jflags = xas.setup_jvp(params, f_min, 1e21, None, precision)
    assert records.dtype == expected.dtype
    assert torch.equal(records.view(torch.int32 if precision != "fp64" else torch.int64),
                       expected.view(torch.int32 if precision != "fp64" else torch.int64))
    assert torch.equal(flags, jflags)
    assert jac.shape == (200, 64, 8) and bool(torch.isfinite(jac[:, :60]).all())


def oracle_vjp_differences(params, freqs, w, reference_mf, columns=range(7), step=1e-5):
    """sum_f w . dh/dparam by fourth-order central differences of the FP64
    oracle (step * max(|x|, 1))."""
    out = np.zeros((len(params), 8))
    for k in columns:
        h = step * np.maximum(np.abs(params[:, k]), 1.0)
        taps = []
        for m in (-2, -1, 1, 2):
            p = params.copy()
            p[:, k] += m * h
            taps.append(oracle.polarizations(p, freqs, reference_mf=reference_mf))
        for j, (a, b) in enumerate(((0, 1), (2, 3))):
            d = (taps[0][j] - 8 * taps[1][j] + 8 * taps[2][j] - taps[3][j]) / (12 * h[:, None])
            out[:, k] += np.sum(w[a] * d.real + w[b] * d.imag, axis=1)
    return out


def test_reference_mf_gradient(cuda):
    """The gradient with the mass-scaled reference frequency (f_ref zero:
    f_ref = reference_mf / M, whose mass dependence the gradient carries),
    against central differences of the oracle; a zero derivative with
    respect to f_ref, also for a binary given with f_ref (NaN values), as the
    reference's tests/test_transforms_xas.py:89-111."""
    import torch

    from waveformx_amd import adjoint

    params = oracle.draw(6, seed=19)
    params[:, 7] = 0.0
    w = np.random.default_rng(7).normal(size=(4, 6, len(GRID)))
    gp = torch.as_tensor(w[0] + 1j * w[1], device=cuda)
    gc = torch.as_tensor(w[2] + 1j * w[3], device=cuda)
    f = torch.as_tensor(GRID, device=cuda)
    g, _ = adjoint.vjp(torch.as_tensor(params, device=cuda), f, gp, gc, 1e21, 0.01, "fp64")
    g = g.cpu().numpy()
    expected = oracle_vjp_differences(params, GRID, w, 0.01)
    scale = np.max(np.abs(expected[:, :7]), axis=0)
    err = np.abs(g[:, :7] - expected[:, :7]) / scale
    assert np.all(err < 1e-6), err.max(axis=0)
    assert np.all(g[:, 7] == 0.0)
    bad = params.copy()
    bad[:, 7] = 20.0
    gb, _ = adjoint.vjp(torch.as_tensor(bad, device=cuda), f, gp, gc, 1e21, 0.01, "fp64")
    assert np.all(gb.cpu().numpy()[:, 7] == 0.0)
    hp, _ = oracle.polarizations(bad, GRID, reference_mf=0.01)
    assert np.isnan(hp).all()
