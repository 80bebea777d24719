"""The reference-convention converters (the reference's
tests/test_xas.py:157-178): a binary given at f_ref in LAL's convention and
the converted binary under ``reference_mf`` give the same waveform; the two
converters are inverse; f_ref = 0 takes f_min (NaN without)."""

import numpy as np
import pytest

from conftest import CASES, compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [c for c in CASES if c[8] > 0], ids=lambda c: f"{c[0]}+{c[1]}")
def test_reference_invariance(cuda, case):
    import torch

    import waveformx_amd.torch as wt
    from waveformx_amd import Binary, reference

    m1, m2, c1, c2, distance, inclination, phi, f_min, f_ref = case
    f = torch.tensor(np.geomspace(f_min, 512.0, 256), device=cuda)
    binary = Binary(torch.tensor(m1, dtype=torch.float64, device=cuda), m2, (0.0, 0.0, c1), (0.0, 0.0, c2),
                    distance, inclination, phi, f_ref)
    hp, hc = wt.polarizations_function("IMRPhenomXAS")(binary, f)
    scaled = reference.to_mass_scaled_reference(binary, 0.01)
    assert float(scaled.f_ref) == 0.0
    hp2, hc2 = wt.polarizations_function("IMRPhenomXAS", reference_mf=0.01)(scaled, f)
    for a, b in ((hp2, hp), (hc2, hc)):
        err, _ = compare(a.cpu().numpy(), b.cpu().numpy())
        assert err < 1e-10, err
    back = reference.to_physical_reference(scaled, f_ref, 0.01)
    assert np.cos(float(back.phi_ref) - phi) == pytest.approx(1.0, abs=1e-12)
    assert float(back.f_ref) == f_ref


def test_zero_f_ref_and_batches(cuda):
    import torch

    from waveformx_amd import Binary, reference

    b = Binary(torch.tensor([36.0, 20.0], dtype=torch.float64), torch.tensor([29.0, 10.0], dtype=torch.float64),
               torch.tensor([[0, 0, 0.3], [0, 0, 0.1]], dtype=torch.float64),
               torch.tensor([[0, 0, -0.4], [0, 0, 0.2]], dtype=torch.float64), 400.0, 0.3,
               torch.tensor([1.2, 0.4], dtype=torch.float64), torch.tensor([0.0, 25.0], dtype=torch.float64))
    nan = reference.to_mass_scaled_reference(b, 0.01)
    assert np.isnan(float(nan.phi_ref[0])) and np.isfinite(float(nan.phi_ref[1]))
    out = reference.to_mass_scaled_reference(b, 0.01, f_min=20.0)
    assert np.all(np.isfinite(out.phi_ref.numpy()))
    # each row as its own binary
    for i in range(2):
        single = Binary(float(b.mass_1[i]), float(b.mass_2[i]), b.spin_1[i].numpy(), b.spin_2[i].numpy(), 400.0, 0.3,
                        float(b.phi_ref[i]), float(b.f_ref[i]))
        one = reference.to_mass_scaled_reference(single, 0.01, f_min=20.0)
        assert float(one.phi_ref) == pytest.approx(float(out.phi_ref[i]), abs=1e-12)
    single32 = reference.to_mass_scaled_reference(
        Binary(*(torch.as_tensor(np.asarray(v), dtype=torch.float32) for v in
                 Binary(36.0, 29.0, (0, 0, 0.3), (0, 0, -0.4), 400.0, 0.3, 1.2, 25.0))), 0.01)
    assert single32.phi_ref.dtype == torch.float32


def test_phenomd_reference_invariance(cuda):
    """IMRPhenomD (tests/test_phenomd.py:190-206): the converted binary under
    reference_mf gives the same waveform; f_ref = 0 takes f_min; the phase
    change is kept modulo 2 pi (phi_ref modulo pi)."""
    import torch

    import waveformx_amd.torch as wt
    from test_oracle_phenomd import CASES as PD_CASES
    from waveformx_amd import Binary, reference

    to_scaled, to_physical = reference.converters("IMRPhenomD")
    for case in PD_CASES[:7]:
        m1, m2, c1, c2, distance, inclination, phi, f_min, f_ref = case
        f = torch.tensor(np.geomspace(f_min, 512.0, 256), device=cuda)
        binary = Binary(torch.tensor(m1, dtype=torch.float64, device=cuda), m2, (0.0, 0.0, c1), (0.0, 0.0, c2),
                        distance, inclination, phi, f_ref)
        hp, hc = wt.polarizations_function("IMRPhenomD")(binary, f)
        scaled = to_scaled(binary, 0.01, f_min=f_min)
        assert float(scaled.f_ref) == 0.0
        hp2, hc2 = wt.polarizations_function("IMRPhenomD", reference_mf=0.01)(scaled, f)
        for a, b in ((hp2, hp), (hc2, hc)):
            err, _ = compare(a.cpu().numpy(), b.cpu().numpy())
            assert err < 1e-10, (case, err)
        back = to_physical(scaled, f_ref if f_ref > 0 else f_min, 0.01)
        # phi_ref modulo pi
        assert np.cos(2.0 * (float(back.phi_ref) - phi)) == pytest.approx(1.0, abs=1e-12)
    nan = to_scaled(Binary(20.0, 45.0, (0, 0, 0.3), (0, 0, -0.4), 300.0, 2.9, 0.7, 0.0), 0.01)
    assert np.isnan(float(nan.phi_ref))
