"""IMRPhenomD's HIP path against its oracle (oracle/phenomd_oracle.py):
FP64 to 1e-12 (relative to the row's largest value; for the reference's
cases, 8 ulp of the inspiral's leading phase where that is larger), the
split-precision FP32 paths to a mismatch of 1e-6 at the FP32-rounded inputs
(the reference's float32 bar, tests/test_phenomd.py:131-158), NaN rows
where LAL fails, ragged grids, every launch shape writing the same bits,
the reference phase of the converters, and the public functions."""

import numpy as np
import pytest

from conftest import compare, psd, weighted_mismatch
from oracle import phenomd_oracle as pd
from test_oracle_phenomd import CASES, case_grid, case_params

pytestmark = pytest.mark.gpu

CONFIG0_GRID = np.linspace(20.0, 1024.0, 4017)


@pytest.fixture(scope="module")
def phd(cuda):
    from waveformx_amd import phenomd, xas

    xas._check_library()
    return phenomd


def run(phd, cuda, params, freqs, precision, amp_scale=1e21, reference_mf=None):
    import torch

    p = torch.as_tensor(np.asarray(params, dtype=np.float64), device=cuda)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    f = torch.as_tensor(freqs, dtype=dtype, device=cuda)
    hp, hc = phd.polarizations(p, f, amp_scale, reference_mf, precision)
    torch.cuda.synchronize()
    return hp.cpu().numpy().astype(np.complex128), hc.cpu().numpy().astype(np.complex128)


def row_error(test, reference):
    return np.max(np.abs(test - reference), axis=1) / np.max(np.abs(reference), axis=1)


def test_fp64_config0_matches_oracle(phd, cuda):
    params = pd.draw(64, seed=0)
    hp, hc = run(phd, cuda, params, CONFIG0_GRID, "fp64")
    op, oc = pd.polarizations(params, CONFIG0_GRID)
    for a, b in ((hp, op), (hc, oc)):
        err = row_error(a, b)
        print("fp64 max rel err", err.max())
        assert np.all(err < 1e-12), err.max()


@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_fp64_reference_cases(phd, cuda, case):
    f = case_grid(case)
    p = case_params(case)[None]
    hp, hc = run(phd, cuda, p, f, "fp64")
    op, oc = pd.polarizations(p, f)
    for a, b in ((hp, op), (hc, oc)):
        err = compare(a[0], b[0])[0]
        print("fp64 case error", err)
        assert err < 1e-12, err


@pytest.mark.parametrize("precision", ["split", "split-df"])
def test_split_config0_mismatch(phd, cuda, precision):
    params = pd.draw(64, seed=0).astype(np.float32).astype(np.float64)
    f = CONFIG0_GRID.astype(np.float32).astype(np.float64)
    hp, hc = run(phd, cuda, params, f, precision)
    op, oc = pd.polarizations(params, f)
    worst = 0.0
    for a, b in ((hp, op), (hc, oc)):
        for i in range(len(params)):
            worst = max(worst, compare(a[i], b[i])[1])
        worst = max(worst, weighted_mismatch(a, b, 1.0 / psd(f)).max())
    print(precision, "worst mismatch", worst)
    assert worst < 1e-6, worst


@pytest.mark.parametrize("case", CASES, ids=[f"case{i}" for i in range(len(CASES))])
def test_split_reference_cases(phd, cuda, case):
    f = case_grid(case, 1000).astype(np.float32).astype(np.float64)
    p = case_params(case).astype(np.float32).astype(np.float64)[None]
    hp, hc = run(phd, cuda, p, f, "split")
    op, oc = pd.polarizations(p, f)
    for a, b in ((hp, op), (hc, oc)):
        assert compare(a[0], b[0])[1] < 1e-6


def test_split_binary_neutron_star(phd, cuda):
    """tests/test_phenomd.py:148-158: a BNS on a 1/256 Hz grid."""
    f = np.float32(20.0 + np.arange((2048 - 20) * 256) / 256.0).astype(np.float64)
    p = np.float32([1.42, 1.27, 0.05, -0.03, 100.0, 0.5, 1.0, 0.0]).astype(np.float64)[None]
    hp, hc = run(phd, cuda, p, f, "split")
    op, oc = pd.polarizations(p, f)
    for a, b in ((hp, op), (hc, oc)):
        assert compare(a[0], b[0])[1] < 1e-6


def test_nan_rows_and_mass_scaled_reference(phd, cuda):
    params = np.array([[30.0, 15.0, 1.3, 1.3, 400.0, 0.4, 0.0, 20.0], case_params(CASES[1])])
    for precision in ("fp64", "split"):
        hp, hc = run(phd, cuda, params, CONFIG0_GRID[:500], precision)
        assert np.isnan(hp[0]).all() and np.isnan(hc[0]).all()
        assert np.isfinite(hp[1]).all()
    p0 = pd.draw(8, seed=3)
    p0[:, 7] = 0.0
    p0[4:, 7] = 25.0  # not zero: NaN with reference_mf
    hp, _ = run(phd, cuda, p0, CONFIG0_GRID, "fp64", reference_mf=0.01)
    op, _ = pd.polarizations(p0, CONFIG0_GRID, reference_mf=0.01)
    assert np.isnan(hp[4:]).all() and np.isnan(op[4:]).all()
    assert np.all(row_error(hp[:4], op[:4]) < 1e-12)


@pytest.mark.parametrize("precision", ["split", "split-df"])
def test_launch_shapes_agree(phd, cuda, precision):
    import torch

    params = torch.as_tensor(pd.draw(37, seed=9), device=cuda)
    f = torch.as_tensor(np.linspace(20.0, 1024.0, 1500), dtype=torch.float32, device=cuda)
    records, _ = phd.setup(params, f[:1], 1e21, None, precision)
    ref = [t.view(torch.int32).clone() for t in phd.evaluate(records, f, precision, 1, freqs_per_thread=1)]
    for group, fpt in ((0, 0), (3, 4), (8, 8), (64, 2)):
        hp, hc = phd.evaluate(records, f, precision, group, freqs_per_thread=fpt)
        assert torch.equal(hp.view(torch.int32), ref[0]), (group, fpt)
        assert torch.equal(hc.view(torch.int32), ref[1]), (group, fpt)


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_ragged_matches_oracle(cuda, precision):
    import torch

    from waveformx_amd import ragged_polarizations

    params = pd.draw(6, seed=12)
    grids = [np.linspace(15.0 + 2 * i, 600.0 + 50 * i, 700 + 113 * i) for i in range(6)]
    if precision != "fp64":
        grids = [g.astype(np.float32).astype(np.float64) for g in grids]
    table = {"mass_1": params[:, 0], "mass_2": params[:, 1], "spin_1": np.c_[np.zeros((6, 2)), params[:, 2]],
             "spin_2": np.c_[np.zeros((6, 2)), params[:, 3]], "luminosity_distance": params[:, 4],
             "inclination": params[:, 5], "phi_ref": params[:, 6], "f_ref": params[:, 7]}
    hp, hc, off = ragged_polarizations("IMRPhenomD", grids, table, precision=precision)
    hp, off = hp.cpu().numpy().astype(np.complex128), off.cpu().numpy()
    for i, g in enumerate(grids):
        op, _ = pd.polarizations(params[i:i + 1], g)
        got = hp[off[i]:off[i + 1]]
        if precision == "fp64":
            assert compare(got, op[0])[0] < 1e-12
        else:
            assert compare(got, op[0])[1] < 1e-6


def test_reference_phase(phd, cuda):
    import torch

    params = pd.draw(32, seed=5)
    mf = np.linspace(0.004, 0.15, 32)
    got = phd.reference_phase(torch.as_tensor(params, device=cuda), torch.as_tensor(mf, device=cuda)).cpu().numpy()
    expected = pd.reference_phase(params[:, 0], params[:, 1], params[:, 2], params[:, 3], mf)
    assert np.all(np.abs(got - expected) <= 1e-12 * np.maximum(np.abs(expected), 1.0))


def test_public_functions(cuda):
    import torch

    import waveformx_amd as wx

    f = np.linspace(20.0, 1024.0, 700)
    params = pd.draw(5, seed=6)
    table = {"mass_1": params[:, 0], "mass_2": params[:, 1], "spin_1": np.c_[np.zeros((5, 2)), params[:, 2]],
             "spin_2": np.c_[np.zeros((5, 2)), params[:, 3]], "luminosity_distance": params[:, 4],
             "inclination": params[:, 5], "phi_ref": params[:, 6], "f_ref": params[:, 7]}
    table["spin_2"][3, 0] = 0.2  # in-plane: rejected
    batch = wx.batch_polarizations("IMRPhenomD", f, table)
    assert list(batch.ok) == [True, True, True, False, True]
    assert "in-plane" in str(batch.errors[3]) and "IMRPhenomD" in str(batch.errors[3])
    op, _ = pd.polarizations(params, f)
    hp = batch.hplus.cpu().numpy()
    assert np.all(row_error(hp[[0, 1, 2, 4]], op[[0, 1, 2, 4]]) < 1e-12)
    assert np.all(hp[3] == 0)
    # the torch function: float64 -> FP64, float32 -> split; with a gradient
    fn = wx.polarizations_function("IMRPhenomD")
    grid = torch.as_tensor(f, device=cuda)
    binary = wx.Binary(torch.tensor(params[0, 0], device=cuda, requires_grad=True), params[0, 1],
                       (0.0, 0.0, params[0, 2]), (0.0, 0.0, params[0, 3]), params[0, 4], params[0, 5], params[0, 6],
                       params[0, 7])
    h64, _ = fn(binary, grid)
    assert h64.dtype == torch.complex128
    assert compare(h64.detach().cpu().numpy(), op[0])[0] < 1e-12
    binary32 = binary._replace(mass_1=torch.tensor(params[0, 0], dtype=torch.float32, device=cuda))
    h32, _ = fn(binary32, grid.to(torch.float32))
    assert h32.dtype == torch.complex64
    assert compare(h32.cpu().numpy().astype(np.complex128), op[0])[1] < 1e-6
    h64.abs().sum().backward()  # the adjoint (checked in test_gpu_phenomd_adjoint.py)
    assert binary.mass_1.grad is not None and bool(torch.isfinite(binary.mass_1.grad))
    with pytest.raises(ValueError):
        wx.polarizations_function("IMRPhenomD", arithmetic="fp32")(binary, grid)
