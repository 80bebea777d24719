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


@pytest.mark.parametrize("approximant", ["IMRPhenomXAS", "IMRPhenomD", "IMRPhenomXHM"])
def test_converters_are_differentiable(cuda, approximant):
    """The gradient of the converted phi_ref with respect to each input
    (the reference takes jax.grad through its converters,
    tests/test_xas.py:174-175) against central differences of the
    converter itself: masses and spins step 1e-6 relative, f_ref and f_min
    1e-5 Hz; d phi_ref' / d phi_ref is 1."""
    import torch

    import waveformx_amd as wx

    def convert(m1, m2, c1, c2, phi, f_ref, f_min, back=False):
        binary = wx.Binary(m1, m2, torch.stack([torch.zeros_like(c1), torch.zeros_like(c1), c1]),
                           torch.stack([torch.zeros_like(c2), torch.zeros_like(c2), c2]), 400.0, 0.3, phi, f_ref)
        if back:
            return wx.to_physical_reference(binary, f_ref, 0.012, approximant=approximant).phi_ref
        return wx.to_mass_scaled_reference(binary, 0.012, f_min=f_min, approximant=approximant).phi_ref

    for values, back in (((36.0, 29.0, 0.3, -0.4, 1.2, 25.0, 20.0), False),
                         ((60.0, 15.0, 0.8, 0.5, 4.0, 0.0, 15.0), False),
                         ((1.6, 1.3, 0.05, 0.02, 2.0, 40.0, 30.0), False),
                         ((36.0, 29.0, 0.3, -0.4, 1.2, 25.0, 20.0), True)):
        x = [torch.tensor(v, dtype=torch.float64, device=cuda, requires_grad=True) for v in values]
        out = convert(*x, back=back)
        grads = torch.autograd.grad(out, x, allow_unused=True)
        for k, v in enumerate(values):
            h = 1e-5 if k >= 5 else 1e-6 * max(abs(v), 1.0)
            up = [torch.tensor(w + (h if j == k else 0.0), dtype=torch.float64, device=cuda)
                  for j, w in enumerate(values)]
            down = [torch.tensor(w - (h if j == k else 0.0), dtype=torch.float64, device=cuda)
                    for j, w in enumerate(values)]
            # the same branch of the modulo on both sides
            wrap = np.pi if approximant == "IMRPhenomD" else 2 * np.pi  # phi_ref modulo pi or 2 pi
            fd = float(torch.remainder(convert(*up, back=back) - convert(*down, back=back) + wrap / 2, wrap)
                       - wrap / 2) / (2 * h)
            got = 0.0 if grads[k] is None else float(grads[k])
            if k == 6 and (values[5] != 0.0 or back):  # f_min only where f_ref = 0
                assert got == 0.0
                continue
            if k == 5 and values[5] == 0.0:  # f_ref = 0 takes f_min
                continue
            scale = max(abs(fd), 1e-3)
            assert abs(got - fd) <= 1e-5 * scale + 1e-7, (approximant, values, k, got, fd)
        assert float(grads[4]) == 1.0


def test_package_level_converters(cuda):
    """waveformx's package-level converters (src/waveformx/_reference.py:
    36-79): approximant selects the model, IMRPhenomXHM takes IMRPhenomXAS's,
    IMRPhenomXPNR (the reference's default) is outside this build, an
    unknown name raises UnknownApproximantError."""
    import waveformx_amd as wx
    from waveformx_amd import reference

    binary = wx.Binary(36.0, 29.0, (0.0, 0.0, 0.3), (0.0, 0.0, -0.4), 500.0, 1.0, 1.2, 30.0)
    xas = reference.to_mass_scaled_reference(binary, 0.01)
    for name in ("IMRPhenomXAS", "IMRPhenomXHM"):
        got = wx.to_mass_scaled_reference(binary, 0.01, approximant=name)
        assert float(got.phi_ref) == float(xas.phi_ref) and float(got.f_ref) == 0.0
        back = wx.to_physical_reference(got, 30.0, 0.01, approximant=name)
        assert np.cos(float(back.phi_ref) - 1.2) == pytest.approx(1.0, abs=1e-12)
    phd = wx.to_mass_scaled_reference(binary, 0.01, approximant="IMRPhenomD")
    assert float(phd.phi_ref) == float(reference.converters("IMRPhenomD")[0](binary, 0.01).phi_ref)
    with pytest.raises(wx.UnsupportedError, match="IMRPhenomXPNR"):
        wx.to_mass_scaled_reference(binary, 0.01)
    with pytest.raises(wx.UnknownApproximantError):
        wx.to_physical_reference(binary, 30.0, 0.01, approximant="IMRPhenomQ")
    assert wx.REFERENCE_MODULES["IMRPhenomXHM"] == "IMRPhenomXAS"
