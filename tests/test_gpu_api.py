"""The drop-in API on the GPU: the torch function's input forms, dtypes,
shapes and errors (the reference's tests/test_torch.py:119-145), and
batch_polarizations' ok/errors (tests/test_xas.py:103-139)."""

import numpy as np
import pytest

from conftest import compare
from oracle import oracle

pytestmark = pytest.mark.gpu

ALIGNED = [36.0, 29.0, 0.0, 0.0, 0.3, 0.0, 0.0, -0.4, 500.0, 0.7, 1.2, 25.0]
FREQUENCIES = np.geomspace(20.0, 900.0, 300)


def row_params(row):
    row = np.asarray(row)
    return row[..., [0, 1, 4, 7, 8, 9, 10, 11]]


def test_single_and_batch_forms(cuda):
    import torch

    import waveformx_amd as wx
    import waveformx_amd.torch as wt

    f = wt.polarizations_function("IMRPhenomXAS")
    x = torch.tensor(ALIGNED, dtype=torch.float64)
    binary = wx.Binary(x[0], x[1], x[2:5], x[5:8], x[8], x[9], x[10], x[11])
    hp, hc = f(binary, torch.tensor(FREQUENCIES))
    assert hp.shape == (len(FREQUENCIES),) and hp.dtype == torch.complex128
    assert hp.device.type == "cpu"  # results come back to the grid's device
    op, oc = oracle.polarizations(row_params(ALIGNED)[None], FREQUENCIES)
    assert compare(hp.numpy(), op[0])[0] < 1e-12
    assert compare(hc.numpy(), oc[0])[0] < 1e-12

    mapping = dict(mass_1=36.0, mass_2=29, spin_1=(0.0, 0.0, 0.3), spin_2=[0.0, 0.0, -0.4],
                   luminosity_distance=500.0, inclination=0.7, phi_ref=1.2, f_ref=25.0)
    hp2, _ = f(mapping, FREQUENCIES)
    assert torch.equal(hp2, hp)

    masses = torch.tensor([36.0, 40.0], dtype=torch.float64, device=cuda)
    hp3, hc3 = f({**mapping, "mass_1": masses}, torch.tensor(FREQUENCIES, device=cuda))
    assert hp3.shape == (2, len(FREQUENCIES)) and hp3.device.type == "cuda"
    assert compare(hp3[0].cpu().numpy(), op[0])[0] < 1e-12

    with pytest.raises(TypeError, match="unknown"):
        f({**mapping, "mass": 1.0}, FREQUENCIES)
    with pytest.raises(ValueError, match="batch"):
        f({**mapping, "mass_1": masses, "mass_2": torch.ones(3, dtype=torch.float64)}, FREQUENCIES)
    with pytest.raises(wx.UnsupportedError):
        wt.polarizations_function("NRSur7dq4")
    with pytest.raises(wx.UnknownApproximantError):
        wt.polarizations_function("IMRPhenomZZZ")


def test_float32_inputs_give_complex64(cuda):
    import torch

    import waveformx_amd.torch as wt

    f = wt.polarizations_function("IMRPhenomXAS")
    x = torch.tensor(ALIGNED, dtype=torch.float32, device=cuda)
    grid = torch.tensor(FREQUENCIES, dtype=torch.float32, device=cuda)
    hp, hc = f((x[0], x[1], x[2:5], x[5:8], x[8], x[9], x[10], x[11]), grid)
    assert hp.dtype == torch.complex64
    rounded = row_params(np.float32(ALIGNED).astype(np.float64))[None]
    op, _ = oracle.polarizations(rounded, np.float32(FREQUENCIES).astype(np.float64))
    assert compare(hp.cpu().numpy().astype(np.complex128), op[0])[1] < 1e-6


def test_scaled_strain_and_reference_mf(cuda):
    import torch

    import waveformx_amd.torch as wt

    grid = torch.tensor(FREQUENCIES, device=cuda)
    x = torch.tensor(ALIGNED, dtype=torch.float64, device=cuda)
    binary = (x[0], x[1], x[2:5], x[5:8], x[8], x[9], x[10], x[11])
    scaled, _ = wt.polarizations_function("IMRPhenomXAS")(binary, grid)
    raw, _ = wt.polarizations_function("IMRPhenomXAS", scaled_strain=False)(binary, grid)
    assert torch.allclose(raw * 1e21, scaled, rtol=1e-15, atol=0)
    # reference_mf: f_ref must be zero (NaN otherwise), then it is reference_mf / M
    nan_hp, _ = wt.polarizations_function("IMRPhenomXAS", reference_mf=0.01)(binary, grid)
    assert torch.isnan(nan_hp).all()
    zero_ref = (x[0], x[1], x[2:5], x[5:8], x[8], x[9], x[10], torch.zeros((), dtype=torch.float64, device=cuda))
    hp, _ = wt.polarizations_function("IMRPhenomXAS", reference_mf=0.01)(zero_ref, grid)
    f_ref = 0.01 / (4.925490947641266978197229498498379006e-6 * 65.0)
    p = row_params(ALIGNED).copy()
    p[7] = f_ref
    op, _ = oracle.polarizations(p[None], FREQUENCIES)
    assert compare(hp.cpu().numpy(), op[0])[0] < 1e-11


def test_batch_polarizations_errors(cuda):
    import waveformx_amd as wx

    f = np.geomspace(20.0, 1024.0, 700)
    table = np.array([ALIGNED,
                      [2000.0, 1.0] + ALIGNED[2:],                       # q > 1000: LAL refuses
                      [36.0, 29.0, 0.1, 0.0, 0.3] + ALIGNED[5:],         # in-plane spin
                      [-1.0, 29.0] + ALIGNED[2:],                        # negative mass
                      [50.0, 8.0, 0.0, 0.0, 0.9, 0.0, 0.0, 0.3, 300.0, 2.1, 0.3, 30.0]])
    batch = wx.batch_polarizations("IMRPhenomXAS", f, table)
    assert list(batch.ok) == [True, False, False, False, True]
    assert "1000" in str(batch.errors[1]) and "in-plane" in str(batch.errors[2])
    assert "positive" in str(batch.errors[3])
    hp = batch.hplus.cpu().numpy()
    assert np.all(hp[1:4] == 0)
    op, _ = oracle.polarizations(row_params(table[[0, 4]]), f)
    assert compare(hp[0], op[0])[0] < 1e-12 and compare(hp[4], op[1])[0] < 1e-12
    with pytest.raises(wx.InvalidParameterError):
        wx.batch_polarizations("IMRPhenomXAS", f, table, strict=True)
    chunked = wx.batch_polarizations("IMRPhenomXAS", f, table, chunk_size=2)
    assert np.array_equal(chunked.hplus.cpu().numpy(), hp)


def test_single_binary_polarizations(cuda):
    """polarizations() and polarizations_on_grid() (src/waveformx/_api.py:
    156-177): one binary from keyword arguments, FP64, against the oracle;
    options as keywords; LAL's checks raised with the reference's messages."""
    import torch

    import waveformx_amd as wx

    f = np.linspace(20.0, 1024.0, 2000)
    kw = dict(mass_1=36.0, mass_2=29.0, spin_1=(0.0, 0.0, 0.3), spin_2=(0.0, 0.0, -0.4), luminosity_distance=500.0,
              inclination=0.7, phi_ref=1.2, f_ref=25.0)
    hp, hc = wx.polarizations("IMRPhenomXAS", f, **kw)
    assert hp.dtype == torch.complex128 and hp.shape == (2000,) and hp.device.type == "cuda"
    op, oc = oracle.polarizations(np.array([[36.0, 29.0, 0.3, -0.4, 500.0, 0.7, 1.2, 25.0]]), f)
    assert compare(hp.cpu().numpy(), op[0])[0] < 1e-12 and compare(hc.cpu().numpy(), oc[0])[0] < 1e-12
    gp, gc = wx.polarizations_on_grid("IMRPhenomXAS", wx.FrequencyGrid(f), **kw)
    assert torch.equal(gp, hp) and torch.equal(gc, hc)
    raw, _ = wx.polarizations("IMRPhenomXAS", f, scaled_strain=False, **kw)
    assert torch.allclose(raw * 1e21, hp, rtol=1e-14, atol=0)
    with pytest.raises(wx.InvalidParameterError, match="in-plane"):
        wx.polarizations("IMRPhenomXAS", f, **dict(kw, spin_1=(0.1, 0.0, 0.3)))
    with pytest.raises(wx.InvalidParameterError, match="positive"):
        wx.polarizations("IMRPhenomXAS", f, **dict(kw, luminosity_distance=-1.0))
    dp, _ = wx.polarizations("IMRPhenomD", f, **kw)
    assert dp.dtype == torch.complex128 and bool(torch.isfinite(torch.view_as_real(dp)).all())


@pytest.mark.parametrize("approximant", ["IMRPhenomXAS", "IMRPhenomD"])
def test_gradcheck(cuda, approximant):
    """torch.autograd.gradcheck of the torch function with respect to the
    grid and every field (the reference's tests/test_torch.py:108-116): the
    hand-written backward, conjugation convention included, against torch's
    own finite differences. The grid's first frequency also sets the tables
    (f_min), whose effect the gradient leaves out as the reference's
    stop_gradient does: f_ref is given, so f_min enters nothing but the
    grid."""
    import torch

    import waveformx_amd.torch as wt

    f = wt.polarizations_function(approximant)
    frequencies = torch.linspace(20.0, 200.0, 20, dtype=torch.float64, device=cuda, requires_grad=True)
    row = np.asarray(ALIGNED)
    values = [row[0], row[1], row[2:5], row[5:8], row[8], row[9], row[10], row[11]]
    fields = [torch.tensor(v, dtype=torch.float64, device=cuda, requires_grad=True) for v in values]

    def function(frequencies, *fields):
        return f(tuple(fields), frequencies)

    assert torch.autograd.gradcheck(function, (frequencies, *fields))


def test_gradient_reaches_a_tensor_inside_a_spin(cuda):
    """A tensor inside a spin sequence gets its gradient (the reference's
    tests/test_torch.py:119-131)."""
    import torch

    import waveformx_amd.torch as wt

    f = wt.polarizations_function("IMRPhenomXAS")
    chi = torch.tensor(0.3, dtype=torch.float64, requires_grad=True)
    binary = dict(mass_1=36.0, mass_2=29, spin_1=(0.0, 0.0, chi), spin_2=[0.0, 0.0, -0.4],
                  luminosity_distance=500.0, inclination=0.7, phi_ref=1.2, f_ref=25.0)
    hp, _ = f(binary, FREQUENCIES)
    (grad,) = torch.autograd.grad(hp.abs().sum(), chi)
    assert torch.isfinite(grad) and grad != 0
    # against a central difference of the same function
    h = 1e-6
    up = f({**binary, "spin_1": (0.0, 0.0, 0.3 + h)}, FREQUENCIES)[0].abs().sum()
    down = f({**binary, "spin_1": (0.0, 0.0, 0.3 - h)}, FREQUENCIES)[0].abs().sum()
    assert abs(float(grad) - float(up - down) / (2 * h)) <= 1e-5 * abs(float(grad))


def test_model_objects(cuda):
    """models.get(...).validate raises LAL's errors from the setup kernel's
    flags (xas/api.py:52-60, phenomd/api.py:33-41 of the reference), and
    .polarizations is the pure function."""
    import torch

    import waveformx_amd as wx
    from waveformx_amd import models

    options = wx.WaveformOptions()
    f = np.linspace(20.0, 512.0, 300)
    xas_model, pd_model = models.get("IMRPhenomXAS"), models.get("IMRPhenomD")
    good = wx.Binary(36.0, 29.0, (0.0, 0.0, 0.3), (0.0, 0.0, -0.4), 500.0, 0.7, 1.2, 25.0)
    xas_model.validate(good, f, options)
    pd_model.validate(good, f, options)
    with pytest.raises(wx.InvalidParameterError, match="mass ratios beyond 1000"):
        xas_model.validate(good._replace(mass_1=2000.0, mass_2=1.0), f, options)
    with pytest.raises(wx.InvalidParameterError, match="final spin"):
        pd_model.validate(good._replace(spin_1=(0.0, 0.0, 1.3), spin_2=(0.0, 0.0, 1.3)), [], options)
    hp, hc = xas_model.polarizations(good, torch.as_tensor(f, device=cuda), options)
    op, oc = oracle.polarizations(np.array([[36.0, 29.0, 0.3, -0.4, 500.0, 0.7, 1.2, 25.0]]), f)
    assert compare(hp.cpu().numpy(), op[0])[0] < 1e-12 and compare(hc.cpu().numpy(), oc[0])[0] < 1e-12
