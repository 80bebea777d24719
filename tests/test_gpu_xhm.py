"""IMRPhenomXHM on the GPU (FP64) against the oracle (``oracle/xhm.py``):
the polarizations and each mode h_{l,-m}, on the reference's test cases
(``tests/test_xhm.py:30-40``, their grids ``:48-50``) and on a random batch
of the multimode workload's distribution.

The bar: ``max|a - b| <= 2e-10 max|b|`` per binary, mode and polarization
(the ``conftest.compare`` error of the reference, whose own float64 bar
against LALSimulation is 1e-8 per mode, ``tests/test_xhm.py:112-123``).
The differences are rounding of large phases: the inspiral phase of a
binary neutron star reaches ~1e4 rad (m/2 times that for mode m), summed
from terms of that size that cancel; the device's cbrt, log, pow and
sin/cos differ from the host's by an ulp, which those sums carry to
~1e-10 of the mode at most (~1e-12 for the heavier binaries). The (2,2)
mode is moreover the IMRPhenomXAS record's table form and the oracle's
the LAL form (up to ~5e-11 apart, the IMRPhenomXAS oracle's own
table-vs-LAL bar), and the (3,2) ringdown recycles it. The higher modes
otherwise follow the same LAL form on both sides."""

import numpy as np
import pytest

from oracle import oracle
from oracle import xhm as oxhm

pytestmark = pytest.mark.gpu

MTSUN = 4.925490947641267e-06
CASES = [
    (36.0, 29.0, 0.3, -0.4, 500.0, 1.0, 1.2, 20.0, 30.0),
    (60.0, 15.0, 0.8, 0.5, 1000.0, 2.5, 4.0, 15.0, 20.0),
    (80.0, 10.0, -0.7, 0.2, 2000.0, 2.0, 0.3, 10.0, 25.0),
    (1.6, 1.3, 0.05, 0.02, 40.0, 0.2, 2.0, 30.0, 40.0),
    (150.0, 100.0, 0.95, 0.95, 3000.0, 0.7, 3.0, 10.0, 15.0),
    (20.0, 45.0, 0.3, -0.4, 300.0, 2.9, 0.7, 20.0, 0.0),
    (50.0, 2.0, 0.995, 0.3, 300.0, 1.1, 1.7, 20.0, 25.0),
    (100.0, 1.0, 0.5, 0.0, 300.0, 1.1, 1.7, 10.0, 0.0),
    (30.0, 30.0, 0.2, 0.2, 400.0, 0.4, 0.0, 20.0, 20.0),
]
ULP = np.finfo(float).eps
# the (2,2) mode is IMRPhenomXAS's table form on the device and LAL's form in
# the XHM oracle: the two arrangements agree to 5e-11 (tests/test_oracle.py)
TABLE_VS_LAL = 5e-11


def mode_bars(params, f_min):
    """Each mode's FP64 bar against the oracle: 1e-12, or 128 ulp of its
    leading inspiral phase |T0| f_min^(-5/3) (m/2)^(8/3) where that is
    larger. The LAL form sums ~13 inspiral terms of up to that size, which
    the device takes from the frequency's powers times the binary's powers
    of M and the oracle computes in M f directly: measured 9-40 ulp, up to
    183 ulp where higher-order terms that cancel outgrow the leading one
    (q = 100's (2,1), (250 Msun)'s (3,2))."""
    t0 = abs(oracle.internals(np.asarray(params, dtype=float), f_min)["T0"])
    return [max(1e-12, 128 * ULP * t0 * f_min ** (-5.0 / 3.0) * (m / 2.0) ** (8.0 / 3.0)) for _, m in oxhm.MODES]


def polarization_bar(params, f_min):
    return max(TABLE_VS_LAL, max(mode_bars(params, f_min)))


def relative_error(got, ref):
    scale = np.max(np.abs(ref))
    return 0.0 if scale == 0 and not np.any(got) else np.max(np.abs(got - ref)) / scale


@pytest.mark.parametrize("case", CASES)
def test_reference_cases(cuda, case):
    import torch

    from waveformx_amd import xhm

    params = np.array([[case[k] for k in (0, 1, 2, 3, 4, 5, 6, 8)]])
    f = np.geomspace(case[7], min(0.3 / ((case[0] + case[1]) * MTSUN), 2048.0), 1000)
    print("case", case, flush=True)
    hp, hc, modes = xhm.polarizations(torch.as_tensor(params, device=cuda), torch.as_tensor(f, device=cuda)) + \
        (xhm.modes(torch.as_tensor(params, device=cuda), torch.as_tensor(f, device=cuda)),)
    op, oc, om = oxhm.polarizations(params, f, amp_scale=1e21, modes=True)
    bar = polarization_bar(params[0], f[0])
    for got, ref in ((hp, op), (hc, oc)):
        error = relative_error(got.cpu().numpy()[0], ref[0])
        print("polarization", error, bar)
        assert error <= bar, (error, bar)
    got_modes = modes.cpu().numpy()
    for k, (mode, bar) in enumerate(zip(oxhm.MODES, mode_bars(params[0], f[0]))):
        error = relative_error(got_modes[k, 0], om[k, 0])
        print(mode, error, bar)
        assert error <= bar, (mode, error, bar)
    # the (2,2) mode against IMRPhenomXAS's oracle, the same table form: its
    # leading term is the same source on both sides, so 1e-12 but where the
    # other inspiral terms reach ~1e3 rad (q = 100), the (2,2)'s mode bar
    xp, _ = oracle.polarizations(params, f)
    kp = -0.5 * (1.0 + np.cos(params[0, 5]) ** 2) * 2.0 * np.sqrt(5.0 / (64.0 * np.pi))
    error = relative_error(got_modes[0, 0], xp[0] / kp)
    print("(2,2) vs IMRPhenomXAS's oracle", error)
    assert error <= max(1e-12, mode_bars(params[0], f[0])[0] if case[0] / case[1] > 20 else 1e-12)


def test_random_batch(cuda):
    import torch

    from waveformx_amd import xhm

    params = oxhm.draw(48, seed=4)
    f = np.linspace(20.0, 2048.0, 8193)
    p = torch.as_tensor(params, device=cuda)
    records, head, modes_, flags = xhm.setup(p, torch.as_tensor(f[:1], device=cuda))
    hp, hc, modes = xhm.evaluate(records, head, modes_, torch.as_tensor(f, device=cuda), with_modes=True)
    assert not bool(flags.any())
    op, oc, om = oxhm.polarizations(params, f, amp_scale=1e21, modes=True)
    hp, hc, modes = hp.cpu().numpy(), hc.cpu().numpy(), modes.cpu().numpy()
    worst = 0.0
    for b in range(params.shape[0]):
        bars = [polarization_bar(params[b], f[0])] * 2 + mode_bars(params[b], f[0])
        refs = [(hp[b], op[b]), (hc[b], oc[b])] + [(modes[k, b], om[k, b]) for k in range(5)]
        for (got, ref), bar in zip(refs, bars):
            error = relative_error(got, ref)
            worst = max(worst, error / bar)
            assert error <= bar, (b, error, bar)
    print("worst error over the batch, in units of its bar:", worst)
    # a float32 grid is the same grid, read as float32
    hp32, _ = xhm.evaluate(records, head, modes_, torch.as_tensor(f, dtype=torch.float32, device=cuda))
    f32 = f.astype(np.float32).astype(np.float64)
    hp64, _ = xhm.evaluate(records, head, modes_, torch.as_tensor(f32, device=cuda))
    assert torch.equal(hp32, hp64)


def test_public_functions(cuda):
    """IMRPhenomXHM through the reference's interface: the torch function
    (float64 and float32 inputs, a batch axis, a backward pass),
    batch_polarizations with failing rows, the mass-scaled reference
    (``torch.py:175-183``, ``_api.py:307-360``, ``xhm/model.py:40-42``)."""
    import torch

    import waveformx_amd as wx
    import waveformx_amd.torch as wt

    case = CASES[0]
    params = np.array([[case[k] for k in (0, 1, 2, 3, 4, 5, 6, 8)]])
    f = np.linspace(20.0, 1024.0, 2001)
    op, oc = oxhm.polarizations(params, f, amp_scale=1e21)
    binary = wx.Binary(case[0], case[1], (0.0, 0.0, case[2]), (0.0, 0.0, case[3]), case[4], case[5], case[6], case[8])
    function = wt.polarizations_function("IMRPhenomXHM")
    hp, hc = function(binary, torch.as_tensor(f, device=cuda))
    assert hp.dtype == torch.complex128 and hp.shape == (f.size,)
    bar = polarization_bar(params[0], f[0])
    assert relative_error(hp.cpu().numpy(), op[0]) <= bar and relative_error(hc.cpu().numpy(), oc[0]) <= bar
    b32 = wx.Binary(*(torch.tensor(v, dtype=torch.float32, device=cuda) for v in binary))
    hp32, _ = function(b32, torch.as_tensor(f, dtype=torch.float32, device=cuda))
    assert hp32.dtype == torch.complex64
    # a batch of two, the second binary heavier
    m1 = torch.tensor([case[0], 60.0], dtype=torch.float64, device=cuda, requires_grad=True)
    batch = wx.Binary(m1, case[1], (0.0, 0.0, case[2]), (0.0, 0.0, case[3]), case[4], case[5], case[6], case[8])
    bp, _ = function(batch, torch.as_tensor(f, device=cuda))
    assert bp.shape == (2, f.size) and torch.equal(bp[0], hp)
    bp.abs().sum().backward()
    assert m1.grad.shape == (2,) and bool(torch.isfinite(m1.grad).all())
    # batch_polarizations: a failing row (q > 1000) is zero and reported
    table = wx.parameter_table({"mass_1": np.array([case[0], 2000.0]), "mass_2": np.array([case[1], 1.0]),
                                "spin_1": np.array([[0, 0, case[2]], [0, 0, 0]]),
                                "spin_2": np.array([[0, 0, case[3]], [0, 0, 0]]),
                                "luminosity_distance": np.array([case[4], 100.0]),
                                "inclination": np.array([case[5], 0.3]), "phi_ref": np.array([case[6], 0.0]),
                                "f_ref": np.array([case[8], 20.0])})
    result = wx.batch_polarizations("IMRPhenomXHM", f, table)
    assert list(result.ok) == [True, False] and "1000" in str(result.errors[1])
    assert torch.equal(result.hplus[0], hp) and not bool(result.hplus[1].any())
    with pytest.raises(wx.InvalidParameterError, match="IMRPhenomXHM is not valid at mass ratios beyond 1000"):
        wx.batch_polarizations("IMRPhenomXHM", f, table, strict=True)
    # the mass-scaled reference: f_ref = reference_mf / M, NaN where f_ref != 0
    scaled = params.copy()
    scaled[0, 7] = 0.012 / (MTSUN * (case[0] + case[1]))
    sp, _ = oxhm.polarizations(scaled, f, amp_scale=1e21)
    zero_ref = wx.Binary(*binary[:7], 0.0)
    rp, _ = wt.polarizations_function("IMRPhenomXHM", reference_mf=0.012)(zero_ref, torch.as_tensor(f, device=cuda))
    assert relative_error(rp.cpu().numpy(), sp[0]) <= polarization_bar(scaled[0], f[0])
    nan_p, _ = wt.polarizations_function("IMRPhenomXHM", reference_mf=0.012)(binary, torch.as_tensor(f, device=cuda))
    assert bool(torch.isnan(nan_p).all())


def test_modes_and_their_sum(cuda):
    """``modes`` in ``SimInspiralChooseFDModes``'s two-sided conventions:
    their sum with the harmonics at (inclination, pi/2) is the
    polarizations (``tests/test_xhm.py:126-128``: to 1e-12), the negative
    frequencies are (-1)^l conj(h_{l,-m}), and ``choose_fd_modes`` holds the
    same modes on LAL's uniform grid."""
    import torch

    import waveformx_amd as wx

    case = CASES[2]
    keywords = dict(mass_1=case[0], mass_2=case[1], spin_1=(0.0, 0.0, case[2]), spin_2=(0.0, 0.0, case[3]),
                    luminosity_distance=case[4], inclination=case[5], phi_ref=case[6], f_ref=case[8])
    f = np.geomspace(case[7], 0.3 / ((case[0] + case[1]) * MTSUN), 700)
    hp, hc = wx.polarizations("IMRPhenomXHM", f, **keywords)
    fm = wx.modes("IMRPhenomXHM", f, **keywords)
    assert fm.frequencies.shape == (2 * f.size,) and len(fm.modes) == 10
    p2, c2 = fm.polarizations()
    assert float((p2 - hp).abs().max()) <= 1e-12 * float(hp.abs().max())
    assert float((c2 - hc).abs().max()) <= 1e-12 * float(hc.abs().max())
    h21, h33 = fm[(2, -1)][f.size:], fm[(3, -3)][f.size:]
    assert torch.equal(fm[(2, 1)][:f.size], torch.flip(torch.conj(h21), dims=(0,)))
    assert torch.equal(fm[(3, 3)][:f.size], torch.flip(-torch.conj(h33), dims=(0,)))
    params = np.array([[case[k] for k in (0, 1, 2, 3, 4, 5, 6, 8)]])
    _, _, om = oxhm.polarizations(params, f, amp_scale=1e21, modes=True)
    for k, ((ell, m), bar) in enumerate(zip(oxhm.MODES, mode_bars(params[0], f[0]))):
        assert relative_error(fm[(ell, -m)][f.size:].cpu().numpy(), om[k, 0]) <= bar
    # choose_fd_modes: the band [f_min, f_cut] of k delta_f, zero elsewhere
    delta_f, f_min = 0.25, 20.0
    chosen = wx.choose_fd_modes("IMRPhenomXHM", delta_f, f_min, 1024.0, **{**keywords, "f_ref": 25.0})
    freqs = chosen.frequencies.cpu().numpy()
    assert freqs[0] < 0 and np.count_nonzero(freqs == 0.0) == 1
    h22 = chosen[(2, -2)].cpu().numpy()
    positive = freqs >= 0
    band = positive & (freqs >= int(f_min / delta_f) * delta_f) & (freqs <= 1024.0)
    assert not np.any(h22[positive & ~band & (freqs < f_min)])
    direct = wx.modes("IMRPhenomXHM", freqs[band], **{**keywords, "f_ref": 25.0})
    assert relative_error(h22[band], direct[(2, -2)][np.count_nonzero(band):].cpu().numpy()) <= 1e-15


def mismatch(a, b, w=None):
    """1 - Re<a, b> / sqrt(<a, a><b, b>), uniform or with weights w."""
    w = np.ones(a.shape[-1]) if w is None else w
    ab = np.sum((a * np.conj(b)).real * w)
    return 1.0 - ab / np.sqrt(np.sum(np.abs(a) ** 2 * w) * np.sum(np.abs(b) ** 2 * w))


def test_split_against_fp64(cuda):
    """The split arithmetic (FP32, each mode's leading inspiral phase in
    FP64) against the FP64 path on the same float32 grid, at the reference's
    float32 bars (``tests/test_xhm.py:156-178``): mismatch <= 1e-6 for the
    polarizations (uniform, and noise-weighted) and <= 2e-4 per mode, on the
    reference's cases, a random batch and a binary neutron star on a
    1/256 Hz grid."""
    import torch

    from conftest import psd
    from waveformx_amd import xhm

    batches = []
    for case in CASES:
        f = np.geomspace(case[7], min(0.3 / ((case[0] + case[1]) * MTSUN), 2048.0), 1000)
        batches.append((np.array([[case[k] for k in (0, 1, 2, 3, 4, 5, 6, 8)]]), f))
    batches.append((oxhm.draw(64, seed=7), np.linspace(20.0, 2048.0, 8193)))
    bns = np.array([[1.42, 1.27, 0.05, -0.03, 100.0, 0.5, 1.0, 0.0]])
    batches.append((bns, 20.0 + np.arange(2 ** 18) / 256.0))
    worst = {"polarization": 0.0, "weighted": 0.0, "mode": 0.0}
    for params, f in batches:
        f32 = torch.as_tensor(f, dtype=torch.float32, device=cuda)
        p = torch.as_tensor(params, device=cuda)
        m32 = xhm.modes(p, f32, precision="split")
        m64 = xhm.modes(p, f32.to(torch.float64), precision="fp64")
        hp32, hc32 = xhm.polarizations(p, f32, precision="split")
        hp64, hc64 = xhm.polarizations(p, f32.to(torch.float64), precision="fp64")
        assert m32.dtype == torch.complex64 and hp32.dtype == torch.complex64
        w = 1.0 / psd(f32.cpu().numpy().astype(np.float64))
        for b in range(params.shape[0]):
            for a32, a64 in ((hp32[b], hp64[b]), (hc32[b], hc64[b])):
                a32 = a32.cpu().numpy().astype(np.complex128)
                a64 = a64.cpu().numpy()
                worst["polarization"] = max(worst["polarization"], mismatch(a32, a64))
                worst["weighted"] = max(worst["weighted"], mismatch(a32, a64, w))
            for k in range(5):
                a64 = m64[k, b].cpu().numpy()
                if np.any(a64):
                    worst["mode"] = max(worst["mode"], mismatch(m32[k, b].cpu().numpy().astype(np.complex128), a64))
    print("split vs FP64, worst mismatches:", worst)
    assert worst["polarization"] <= 1e-6 and worst["weighted"] <= 1e-6 and worst["mode"] <= 2e-4


@pytest.mark.parametrize("case", [c for c in CASES if c[8] > 0][:5], ids=lambda c: f"{c[0]}+{c[1]}")
def test_reference_invariance_of_the_modes(cuda, case):
    """The reference's ``tests/test_xhm.py:181-197``: a binary given at f_ref and
    the same binary converted by IMRPhenomXAS's converters to the mass-scaled
    reference give the same modes under ``reference_mf``, the odd modes up
    to a half-turn (the converters fix phi_ref modulo pi, the odd modes turn
    by e^{i m delta})."""
    import torch

    import waveformx_amd as wx
    from waveformx_amd import reference

    m1, m2, c1, c2, distance, inclination, phi, f_min, f_ref = case
    f = np.geomspace(f_min, min(0.3 / ((m1 + m2) * MTSUN), 1024.0), 400)
    keywords = dict(mass_1=m1, mass_2=m2, spin_1=(0.0, 0.0, c1), spin_2=(0.0, 0.0, c2), luminosity_distance=distance,
                    inclination=inclination, phi_ref=phi, f_ref=f_ref)
    physical = wx.modes("IMRPhenomXHM", f, **keywords)
    binary = wx.Binary(torch.tensor(m1, dtype=torch.float64), m2, (0.0, 0.0, c1), (0.0, 0.0, c2), distance,
                       inclination, phi, f_ref)
    scaled = reference.to_mass_scaled_reference(binary, 0.01)
    assert float(scaled.f_ref) == 0.0
    converted = wx.modes("IMRPhenomXHM", f, reference_mf=0.01,
                         **{**keywords, "phi_ref": float(scaled.phi_ref), "f_ref": 0.0})
    for (ell, m), h in physical.modes.items():
        a, b = converted[(ell, m)].cpu().numpy(), h.cpu().numpy()
        if not np.any(b):
            continue
        err = min(relative_error(a, b), relative_error(-a, b)) if m % 2 else relative_error(a, b)
        assert err < 1e-9, ((ell, m), err)


def test_degenerate_sizes_and_errors(cuda):
    """Empty batches and grids, one frequency, a failing row (NaN in every
    output and mode), and the argument checks of the device functions."""
    import torch

    from waveformx_amd import xhm

    f = torch.linspace(20.0, 512.0, 300, dtype=torch.float64, device=cuda)
    none = torch.zeros((0, 8), dtype=torch.float64, device=cuda)
    for precision in ("fp64", "split"):
        hp, hc = xhm.polarizations(none, f, precision=precision)
        assert hp.shape == (0, 300) and hc.shape == (0, 300)
    params = torch.as_tensor(oxhm.draw(3, seed=11), device=cuda)
    records, head, modes_, flags = xhm.setup(params, f[:1])
    hp, hc = xhm.evaluate(records, head, modes_, f[:0])
    assert hp.shape == (3, 0)
    one, _ = xhm.evaluate(records, head, modes_, f[:1])
    full, _ = xhm.evaluate(records, head, modes_, f)
    assert torch.equal(one[:, 0], full[:, 0])
    bad = params.clone()
    bad[1, 0], bad[1, 1] = 2000.0, 1.0  # q > 1000: LAL's error, NaN row
    records, head, modes_, flags = xhm.setup(bad, f[:1])
    assert int(flags[1].item()) & 1 and not bool(flags[[0, 2]].any())
    hp, hc, m = xhm.evaluate(records, head, modes_, f, with_modes=True)
    assert bool(torch.isnan(hp[1]).all()) and bool(torch.isnan(m[:, 1]).all())
    assert bool(torch.isfinite(hp[[0, 2]]).all()) and bool(torch.isfinite(torch.view_as_real(m[:, [0, 2]])).all())
    with pytest.raises(ValueError, match="float32 grid"):
        s = xhm.setup(params, f[:1], precision="split")
        xhm.evaluate(*s[:3], f)
    with pytest.raises(ValueError, match="out"):
        xhm.evaluate(records, head, modes_, f, out=(torch.empty((3, 300), dtype=torch.complex64, device=cuda),) * 2)
    with pytest.raises(ValueError, match="arithmetic"):
        xhm.setup(params, f[:1], precision="fp32")
    with pytest.raises(ValueError, match="params"):
        xhm.setup(params[:, :7].contiguous(), f[:1])


@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[6]], ids=lambda c: f"{c[0]}+{c[1]}")
def test_vjp_against_oracle_differences(cuda, case):
    """The vector-Jacobian product of the FP64 path (xhm.vjp: the exact
    reverse-mode adjoint in FP64) and the device's own differences of the
    FP64 evaluation (xhm.vjp_differences, its region selections held)
    against the oracle's fourth-order differences, for a random cotangent,
    on a grid below the (2,2)'s cutoff: within 1e-6 of each parameter's
    largest derivative; and the torch backward equals xhm.vjp."""
    import torch

    import waveformx_amd as wx
    import waveformx_amd.torch as wt
    from waveformx_amd import xhm

    params = np.array([[case[k] for k in (0, 1, 2, 3, 4, 5, 6, 8)]])
    f = np.geomspace(case[7], 0.2 / ((case[0] + case[1]) * MTSUN), 500)
    rng = np.random.default_rng(3)
    op, oc = oxhm.polarizations(params, f, amp_scale=1e21)
    scale = np.abs(op).max()
    cot_p = (rng.normal(size=f.size) + 1j * rng.normal(size=f.size))[None] * scale
    cot_c = (rng.normal(size=f.size) + 1j * rng.normal(size=f.size))[None] * scale
    jp, jc = oxhm.jacobian(params, f, amp_scale=1e21)
    expected = (np.einsum("bf,bfk->bk", np.conj(cot_p), jp) + np.einsum("bf,bfk->bk", np.conj(cot_c), jc)).real
    args = (torch.as_tensor(params, device=cuda), torch.as_tensor(f, device=cuda),
            torch.as_tensor(cot_p, device=cuda), torch.as_tensor(cot_c, device=cuda), 1e21)
    got = xhm.vjp(*args).cpu().numpy()
    differences = xhm.vjp_differences(*args).cpu().numpy()
    scale_g = np.max(np.abs(expected), axis=0)
    for name, value in (("exact", got), ("differences", differences)):
        err = np.abs(value - expected) / np.where(scale_g > 0, scale_g, 1.0)
        print(f"vjp ({name}) vs oracle differences:", err.max(axis=0))
        assert np.all(err <= 1e-6), err
    # the torch backward: L = Re sum conj(cot) h
    fields = [torch.tensor(case[k], dtype=torch.float64, device=cuda, requires_grad=True) for k in range(9)]
    binary = wx.Binary(fields[0], fields[1], torch.stack([fields[2] * 0, fields[2] * 0, fields[2]]),
                       torch.stack([fields[3] * 0, fields[3] * 0, fields[3]]), fields[4], fields[5], fields[6],
                       fields[8])
    hp, hc = wt.polarizations_function("IMRPhenomXHM")(binary, torch.as_tensor(f, device=cuda))
    loss = (torch.conj(torch.as_tensor(cot_p[0], device=cuda)) * hp).real.sum() + \
        (torch.conj(torch.as_tensor(cot_c[0], device=cuda)) * hc).real.sum()
    loss.backward()
    torch_grads = [float(fields[k].grad) for k in (0, 1, 2, 3, 4, 5, 6, 8)]
    assert np.allclose(torch_grads, got[0], rtol=1e-12, atol=0.0)


@pytest.mark.parametrize("case", [CASES[1], CASES[5], CASES[6]], ids=lambda c: f"{c[0]}+{c[1]}")
def test_grid_gradient(cuda, case):
    """dL/df of each frequency (xhm.adjoint with wrt_grid, L = Re sum_f
    conj(cot) h summed over a batch) against the oracle's fourth-order
    differences in each frequency (f_ref given, so the first frequency's
    set-up role enters nothing), on a grid below the (2,2)'s cutoff: FP64
    within 1e-6 of each frequency's own value plus 1e-4 of the largest; the
    split gradient on the float32 grid within the float32 bar of FP64's at
    the same frequencies."""
    import torch

    from waveformx_amd import xhm

    base = np.array([case[k] for k in (0, 1, 2, 3, 4, 5, 6, 8)])
    if base[7] == 0.0:
        base[7] = case[7]  # the reference frequency given (f_ref = 0 would move with the grid's first bin)
    params = np.stack([base, base * np.array([1.02, 0.97, 0.9, 1.1, 1.0, 1.0, 1.0, 1.0])])
    m_total = params[:, :2].sum(1).max()
    f = np.geomspace(case[7], 0.2 / (m_total * MTSUN), 400)
    rng = np.random.default_rng(31)
    op, _ = oxhm.polarizations(params, f, amp_scale=1e21)
    scale = np.abs(op).max()
    cot_p = (rng.normal(size=(2, f.size)) + 1j * rng.normal(size=(2, f.size))) * scale
    cot_c = (rng.normal(size=(2, f.size)) + 1j * rng.normal(size=(2, f.size))) * scale
    # the oracle's derivative of each bin's polarizations in its own frequency
    step = 1e-5 * f
    dh_p = np.zeros_like(op)
    dh_c = np.zeros_like(op)
    for offset, weight in ((-2, 1.0 / 12.0), (-1, -8.0 / 12.0), (1, 8.0 / 12.0), (2, -1.0 / 12.0)):
        hp, hc = oxhm.polarizations(params, f + offset * step, amp_scale=1e21)
        dh_p += weight * hp / step
        dh_c += weight * hc / step
    expected = (np.conj(cot_p) * dh_p + np.conj(cot_c) * dh_c).real.sum(0)
    p = torch.as_tensor(params, device=cuda)
    _, _, got = xhm.adjoint(p, torch.as_tensor(f, device=cuda), torch.as_tensor(cot_p, device=cuda),
                            torch.as_tensor(cot_c, device=cuda), precision="fp64", wrt_grid=True)
    got = got.cpu().numpy()
    err = np.abs(got - expected) / (np.abs(expected) + 1e-4 * np.abs(expected).max())
    print("FP64 dL/df vs oracle differences:", err.max())
    assert np.all(err <= 1e-6), (err.argmax(), err.max())
    # split on the float32 grid against FP64 at the same frequencies
    f32 = torch.as_tensor(f, dtype=torch.float32, device=cuda)
    cp32 = torch.as_tensor(cot_p, dtype=torch.complex64, device=cuda)
    cc32 = torch.as_tensor(cot_c, dtype=torch.complex64, device=cuda)
    _, _, split = xhm.adjoint(p, f32, cp32, cc32, precision="split", wrt_grid=True)
    _, _, ref = xhm.adjoint(p, f32.double(), cp32.to(torch.complex128), cc32.to(torch.complex128),
                            precision="fp64", wrt_grid=True)
    split, ref = split.cpu().numpy(), ref.cpu().numpy()
    excess = np.abs(split - ref) / (1e-2 * np.abs(ref) + 1e-3 * np.abs(ref).max())
    print("split dL/df, worst fraction of the float32 bar:", excess.max())
    assert np.all(excess <= 1.0)


def test_grid_gradient_torch(cuda):
    """torch.autograd.gradcheck of the IMRPhenomXHM torch function with
    respect to the grid and every field (the reference's
    tests/test_torch.py:108-116): the backward (the exact FP64 adjoint)
    against torch's own finite differences of the FP64 evaluation; f_ref is
    given, so the grid's first frequency enters nothing but the grid."""
    import torch

    import waveformx_amd.torch as wt

    f = wt.polarizations_function("IMRPhenomXHM")
    frequencies = torch.linspace(20.0, 200.0, 20, dtype=torch.float64, device=cuda, requires_grad=True)
    values = [36.0, 29.0, (0.0, 0.0, 0.3), (0.0, 0.0, -0.4), 500.0, 1.0, 1.2, 30.0]
    fields = [torch.tensor(v, dtype=torch.float64, device=cuda, requires_grad=True) for v in values]

    def function(frequencies, *fields):
        return f(tuple(fields), frequencies)

    assert torch.autograd.gradcheck(function, (frequencies, *fields))


def test_fused_projection(cuda):
    """xhm.project (the inner products of the gradient, taken in the
    evaluation's launch) against the same inner products of the written
    polarizations, with cotangent rows shared by binaries (row b % R) and a
    grid that does not fill its last block; a failing binary gives NaN."""
    import torch

    from waveformx_amd import xhm

    params = oxhm.draw(6, seed=11)
    params[5, 2] = 1.5  # fails LAL's checks (the device flags it)
    p = torch.as_tensor(params, device=cuda)
    f = torch.linspace(20.0, 700.0, 300, dtype=torch.float64, device=cuda)
    rec, head, modes_, _ = xhm.setup(p, f[:1])
    hp, hc = xhm.evaluate(rec, head, modes_, f)
    rng = np.random.default_rng(5)
    for rows in (6, 2, 1):
        cp = torch.as_tensor(rng.normal(size=(rows, 300)) + 1j * rng.normal(size=(rows, 300)), device=cuda) * 1e-21
        cc = torch.as_tensor(rng.normal(size=(rows, 300)) + 1j * rng.normal(size=(rows, 300)), device=cuda) * 1e-21
        index = torch.arange(6, device=cuda) % rows
        expected = ((torch.conj(cp[index]) * hp).real.sum(1) + (torch.conj(cc[index]) * hc).real.sum(1)).cpu().numpy()
        got = xhm.project(rec, head, modes_, f, cp, cc).cpu().numpy()
        assert np.isnan(got[5]) and np.isnan(expected[5])
        np.testing.assert_allclose(got[:5], expected[:5], rtol=1e-12, atol=1e-12 * np.abs(expected[:5]).max())
    # several heads (other factors) over the same records and modes, in one launch: each as alone
    other = head.clone()
    other[:, :20] = torch.as_tensor(rng.normal(size=(6, 20)), device=cuda)
    heads = torch.stack([head, other, head * 2.0])
    many = xhm.project(rec, heads, modes_, f, cp, cc).cpu().numpy()
    for j in range(3):
        np.testing.assert_array_equal(many[j], xhm.project(rec, heads[j], modes_, f, cp, cc).cpu().numpy())
    with pytest.raises(ValueError):
        xhm.project(rec, torch.stack([head] * 5), modes_, f, cp, cc)
    with pytest.raises(ValueError):
        xhm.project(rec, head, modes_, f.float(), cp, cc)
    assert xhm.project(rec[:0], head[:0], modes_[:0], f, cp, cc).shape == (0,)


def test_vjp_of_a_batch_is_that_of_each_binary(cuda):
    """The FP64 gradient of a batch (xhm.vjp, the exact adjoint) equals each
    binary's own, bit for bit, and so does that of the differences
    (xhm.vjp_differences: the stencil points set up and projected as one
    batch); f_ref = 0 has no f_ref derivative, a failing binary a NaN row."""
    import torch

    from waveformx_amd import xhm

    params = oxhm.draw(5, seed=12)
    params[1, 7] = 0.0
    params[3, 2] = 1.5  # fails LAL's checks (the device flags it)
    p = torch.as_tensor(params, device=cuda)
    f = torch.as_tensor(np.geomspace(20.0, 300.0, 400), device=cuda)
    rng = np.random.default_rng(6)
    cp = torch.as_tensor(rng.normal(size=(5, 400)) + 1j * rng.normal(size=(5, 400)), device=cuda) * 1e-21
    cc = torch.as_tensor(rng.normal(size=(5, 400)) + 1j * rng.normal(size=(5, 400)), device=cuda) * 1e-21
    for vjp in (xhm.vjp, xhm.vjp_differences):
        batch = vjp(p, f, cp, cc).cpu().numpy()
        for b in range(5):
            one = vjp(p[b:b + 1], f, cp[b:b + 1], cc[b:b + 1]).cpu().numpy()[0]
            np.testing.assert_array_equal(batch[b], one)
        assert batch[1, 7] == 0.0 and np.all(np.isfinite(batch[[0, 1, 2, 4]]))
        assert np.all(np.isnan(batch[3, :7]))


def float32_bar(ref):
    """The reference's float32 gradient bar (tests/test_xas.py:157-171):
    1e-2 |g| + 1e-3 max |g| per parameter."""
    return 1e-2 * np.abs(ref) + 1e-3 * np.abs(ref).max(axis=0)


def test_loss_and_grad(cuda):
    """adjoint.loss_and_grad for IMRPhenomXHM (the loss fused into the
    exact adjoint, in each arithmetic): the loss of the forward pass (FP64:
    the adjoint's residual is in the split layout's formula, 1e-9 of the
    FP64 evaluation's loss); FP64's gradient is xhm.vjp's for the
    cotangents 2 w (h - d) of the FP64 forward pass to 1e-9 of each
    parameter's largest, the split one within the reference's float32 bar
    of it; adjoint.vjp routes each arithmetic, with the grid's gradient;
    the argument checks."""
    import torch

    from waveformx_amd import adjoint, xhm

    params = oxhm.draw(4, seed=13)
    p = torch.as_tensor(params, device=cuda)
    d_params = torch.as_tensor(params * (1.0 + 1e-2 * np.random.default_rng(8).normal(size=params.shape)),
                               device=cuda)
    d_params[:, 7] = p[:, 7]
    freqs = np.linspace(20.0, 600.0, 700)
    psd = (freqs / 100.0) ** -4 + 1.0 + (freqs / 100.0) ** 2
    for precision, rdtype in (("fp64", torch.float64), ("split", torch.float32)):
        f = torch.as_tensor(freqs, dtype=rdtype, device=cuda)
        w = torch.as_tensor(1.0 / psd, dtype=rdtype, device=cuda)
        dp, dc = xhm.polarizations(d_params, f, precision=precision)
        print("loss_and_grad", precision, flush=True)
        loss, grads = adjoint.loss_and_grad(p, f, dp, dc, w, precision=precision, approximant="IMRPhenomXHM")
        hp, hc = xhm.polarizations(p, f, precision=precision)
        expected = ((hp - dp).abs() ** 2 + (hc - dc).abs() ** 2) @ w.to(hp.real.dtype)
        np.testing.assert_allclose(loss.cpu().numpy(), expected.double().cpu().numpy(),
                                   rtol=1e-5 if precision == "split" else 1e-9)
        w64 = w.double()
        ref = xhm.vjp(p, f.double(), 2.0 * w64 * (hp - dp).to(torch.complex128),
                      2.0 * w64 * (hc - dc).to(torch.complex128)).cpu().numpy()
        if precision == "fp64":
            err = np.abs(grads.cpu().numpy() - ref) / np.abs(ref).max(axis=0)
            print("FP64 loss gradient vs xhm.vjp of the residual", err.max(axis=0), flush=True)
            assert np.all(err <= 1e-9), err
        else:
            excess = np.abs(grads.cpu().numpy() - ref) / float32_bar(ref)
            print("split loss gradient, worst fraction of the float32 bar", excess.max(), flush=True)
            assert np.all(excess <= 1.0), excess
        assert bool(torch.isfinite(grads[:, :7]).all()) and bool((loss > 0).all())
    # adjoint.vjp routes each arithmetic to the exact adjoint, with the grid's gradient on request
    cot = (2.0 * w * (hp - dp)).to(torch.complex64)
    got, grid = adjoint.vjp(p, f, cot, None, 1e21, None, "split", approximant="IMRPhenomXHM")
    assert grid is None
    np.testing.assert_array_equal(got.cpu().numpy(),
                                  xhm.adjoint(p, f, cot, torch.zeros_like(cot))[0].cpu().numpy())
    f64 = f.double()
    cot64 = cot.to(torch.complex128)
    got, _ = adjoint.vjp(p, f64, cot64, None, 1e21, None, "fp64", approximant="IMRPhenomXHM")
    np.testing.assert_array_equal(got.cpu().numpy(), xhm.vjp(p, f64, cot64, torch.zeros_like(cot64)).cpu().numpy())
    got, grid = adjoint.vjp(p, f64, cot64, None, 1e21, None, "fp64", wrt_grid=True, approximant="IMRPhenomXHM")
    np.testing.assert_array_equal(grid.cpu().numpy(), xhm.adjoint(p, f64, cot64, torch.zeros_like(cot64),
                                                                  precision="fp64", wrt_grid=True)[2].cpu().numpy())
    assert grid.shape == (f64.shape[0],) and bool(torch.isfinite(grid).all())
    with pytest.raises(ValueError):
        adjoint.vjp(p, f64, cot64, None, 1e21, None, "fp32", approximant="IMRPhenomXHM")
    with pytest.raises(ValueError):
        adjoint.loss_and_grad(p, f, dp, dc, w, precision="fp32", approximant="IMRPhenomXHM")
    with pytest.raises(ValueError):
        adjoint.loss_and_grad(p, f, dp, dc, w, precision="split", powers=w, approximant="IMRPhenomXHM")


def test_single_mode_gradient(cuda):
    """xhm.adjoint_with: from an unchanged set-up, adjoint's bits; with all
    but one mode's factors zeroed in the set-up's head (and in the data),
    the exact gradient of that mode's own loss: against central differences
    of that loss (the FP64 evaluation with the same head, steps of 1e-6
    times max(|x|, 0.1) in the masses and spins) to 1e-5 of each
    parameter's largest, and
    the split gradient within the reference's float32 bar of the FP64 one
    (every mode; the (3,2)'s ringdown still recycles the (2,2))."""
    import torch

    from waveformx_amd import xhm

    params = oxhm.draw(3, seed=21)
    p = torch.as_tensor(params, device=cuda)
    perturbed = params * (1.0 + 1e-2 * np.random.default_rng(5).normal(size=params.shape))
    perturbed[:, 7] = params[:, 7]
    freqs = np.linspace(20.0, 600.0, 700)
    psd = (freqs / 100.0) ** -4 + 1.0 + (freqs / 100.0) ** 2
    f64 = torch.as_tensor(freqs, device=cuda)
    f32 = f64.float()
    w64 = torch.as_tensor(1.0 / psd, device=cuda)

    def only(head, k):
        out = torch.zeros_like(head)
        out[:, 20:] = head[:, 20:]
        out[:, 4 * k:4 * k + 4] = head[:, 4 * k:4 * k + 4]
        return out

    def forward(theta, k):
        records, head, modes_, _ = xhm.setup(torch.as_tensor(theta, device=cuda), f64[:1])
        return xhm.evaluate(records, only(head, k), modes_, f64)

    cot = torch.as_tensor(np.random.default_rng(6).normal(size=(3, 700)) * (1 + 1j), device=cuda)
    j = xhm.setup_jvp(p, f64[:1], precision="fp64")
    np.testing.assert_array_equal(xhm.adjoint_with(j, p, f64, cot, cot, precision="fp64")[0].cpu().numpy(),
                                  xhm.adjoint(p, f64, cot, cot, precision="fp64")[0].cpu().numpy())
    for k in range(5):
        dp, dc = forward(perturbed, k)
        j = xhm.setup_jvp(p, f64[:1], precision="fp64")
        j["head"], j["dhead"] = only(j["head"], k), only(j["dhead"], k)
        grads, loss, _ = xhm.adjoint_with(j, p, f64, dp, dc, w64, precision="fp64")
        grads = grads.cpu().numpy()
        hp, hc = forward(params, k)
        np.testing.assert_allclose(loss.cpu().numpy(),
                                   (((hp - dp).abs() ** 2 + (hc - dc).abs() ** 2) @ w64).cpu().numpy(), rtol=1e-9)
        diff = np.zeros((3, 4))
        for d in range(4):
            step = 1e-6 * np.maximum(np.abs(params[:, d]), 0.1)
            losses = []
            for sign in (1.0, -1.0):
                theta = params.copy()
                theta[:, d] += sign * step
                hp, hc = forward(theta, k)
                losses.append((((hp - dp).abs() ** 2 + (hc - dc).abs() ** 2) @ w64).cpu().numpy())
            diff[:, d] = (losses[0] - losses[1]) / (2.0 * step)
        err = np.abs(grads[:, :4] - diff) / np.abs(diff).max(axis=0)
        print("mode", oxhm.MODES[k], "gradient vs central differences", err.max(axis=0), flush=True)
        assert np.all(err <= 1e-5), (k, err)
        j32 = xhm.setup_jvp(p, f32[:1].double(), precision="split")
        j32["head"], j32["dhead"] = only(j32["head"], k), only(j32["dhead"], k)
        dp64, dc64 = forward(perturbed, k)
        g32 = xhm.adjoint_with(j32, p, f32, dp64.to(torch.complex64), dc64.to(torch.complex64), w64.float(),
                               precision="split")[0].cpu().numpy()
        j64 = xhm.setup_jvp(p, f32[:1].double(), precision="fp64")
        j64["head"], j64["dhead"] = only(j64["head"], k), only(j64["dhead"], k)
        g64 = xhm.adjoint_with(j64, p, f32.double(), dp64, dc64, w64, precision="fp64")[0].cpu().numpy()
        excess = np.abs(g32 - g64)[:, :7] / float32_bar(g64[:, :7])
        print("mode", oxhm.MODES[k], "split gradient, worst fraction of the float32 bar", excess.max(), flush=True)
        assert np.all(excess <= 1.0), (k, excess)


def test_gradient_on_a_permuted_grid(cuda):
    """The adjoint's region waves on an ascending grid (each kind of wave
    takes only its own blocks, between the cuts found by search) against
    the same bins in a shuffled order (no cuts: every wave tests every bin):
    the same loss and gradient up to the sums' grouping, in both arithmetics,
    and dL/df permuted with the bins; a grid that ascends but for one pair of
    bins takes the shuffled path too."""
    import torch

    from waveformx_amd import xhm

    params = oxhm.draw(6, seed=31)
    p = torch.as_tensor(params, device=cuda)
    perturbed = params * (1.0 + 1e-2 * np.random.default_rng(9).normal(size=params.shape))
    perturbed[:, 7] = params[:, 7]
    freqs = np.linspace(20.0, 1024.0, 1500)
    w = 1.0 / ((freqs / 100.0) ** -4 + 1.0 + (freqs / 100.0) ** 2)
    order = np.random.default_rng(10).permutation(freqs.size)
    swapped = np.arange(freqs.size)
    swapped[[700, 701]] = swapped[[701, 700]]
    for precision, rdtype, ctype, tol in (("fp64", torch.float64, torch.complex128, 1e-11),
                                          ("split", torch.float32, torch.complex64, 1e-5)):
        f = torch.as_tensor(freqs, dtype=rdtype, device=cuda)
        wt = torch.as_tensor(w, dtype=rdtype, device=cuda)
        dp, dc = xhm.polarizations(torch.as_tensor(perturbed, device=cuda), f.double())
        dp, dc = dp.to(ctype), dc.to(ctype)
        grads, loss, dldf = xhm.adjoint(p, f, dp, dc, wt, precision=precision, wrt_grid=True)
        for perm in (order, swapped):
            idx = torch.as_tensor(perm, device=cuda)
            # the shuffled grid's first frequency is not f_min: the set-up takes the sorted grid's
            j = xhm.setup_jvp(p, f[:1].double(), precision=precision)
            g2, l2, d2 = xhm.adjoint_with(j, p, f[idx], dp[:, idx], dc[:, idx], wt[idx], precision=precision,
                                          wrt_grid=True)
            err = (g2 - grads).abs().amax(0) / grads.abs().amax(0)
            print(precision, "permuted grid: gradient", err.max().item(), flush=True)
            assert bool((err[:7] <= tol).all()), err
            np.testing.assert_allclose(l2.cpu().numpy(), loss.cpu().numpy(), rtol=tol)
            np.testing.assert_allclose(d2.cpu().numpy(), dldf[idx].cpu().numpy(), rtol=0,
                                       atol=tol * dldf.abs().max().item())


def test_split_gradient(cuda):
    """The exact split gradient (xhm.adjoint: the dual-number set-up and
    the reverse-mode adjoint of the split evaluation) for random cotangents
    on a float32 grid over every region of the modes: within the
    reference's float32 bar of the FP64 path's (xhm.vjp, the exact FP64
    adjoint at the same float32 frequencies) on a random batch and edge rows (equal masses
    and spins, m1 < m2, a spin of 0.99, q = 18, f_ref = 0 and the
    mass-scaled reference), and at the boundaries where the differences are
    one-sided (masses 1e-5 apart with unequal spins: the labels' swap; a
    spin of 1 - 1e-7; q = 1000); the swapped binary's gradient is the permuted one
    bit for bit; a binary that fails LAL's checks has a NaN row; a batch
    equals each binary alone; the torch backward of the float32 function is
    this gradient."""
    import torch

    import waveformx_amd as wx
    import waveformx_amd.torch as wt
    from waveformx_amd import xhm

    params = oxhm.draw(12, seed=21)
    edge = np.array([[30.0, 30.0, 0.2, 0.2, 400.0, 0.4, 0.0, 20.0],
                     [20.0, 45.0, 0.3, -0.4, 300.0, 2.9, 0.7, 20.0],   # swapped, then
                     [45.0, 20.0, -0.4, 0.3, 300.0, 2.9, 0.7, 20.0],   # in order
                     [50.0, 12.0, 0.99, 0.3, 300.0, 1.1, 1.7, 20.0],
                     [90.0, 5.0, 0.5, 0.0, 300.0, 1.1, 1.7, 0.0],      # f_ref = 0
                     # at the boundaries that the differences take one-sided: masses 1e-5 apart with
                     # unequal spins (the labels' swap within the central stencil's reach; at exactly
                     # equal masses the exact gradient takes |m1 - m2|'s derivative as zero, jax's
                     # convention, which no difference sees), a spin at its bound, q = 1000
                     [30.0, 30.0 - 1e-5, 0.5, -0.3, 400.0, 0.9, 0.3, 20.0],
                     [40.0, 20.0, 1.0 - 1e-7, 0.2, 400.0, 0.9, 0.3, 20.0],
                     [200.0, 0.2, 0.3, 0.0, 100.0, 0.9, 0.3, 20.0],
                     [30.0, 20.0, 1.5, 0.0, 300.0, 1.1, 1.7, 20.0]])   # fails LAL's checks
    params = np.concatenate([params, edge])
    nb = params.shape[0]
    swapped, ordered, no_f_ref = 13, 14, 16
    p = torch.as_tensor(params, device=cuda)
    f = torch.linspace(20.0, 1024.0, 2049, dtype=torch.float32, device=cuda)
    rng = np.random.default_rng(22)
    cp = torch.as_tensor(rng.normal(size=(nb, 2049)) + 1j * rng.normal(size=(nb, 2049)), dtype=torch.complex64,
                         device=cuda)
    cc = torch.as_tensor(rng.normal(size=(nb, 2049)) + 1j * rng.normal(size=(nb, 2049)), dtype=torch.complex64,
                         device=cuda)
    cp[swapped], cc[swapped] = cp[ordered], cc[ordered]  # the swapped pair sees the same cotangents
    for reference_mf in (None, 0.01):
        print("split gradient, reference_mf", reference_mf, flush=True)
        pr = p.clone()
        if reference_mf is not None:
            pr[:, 7] = 0.0  # given at the mass-scaled reference
        got = xhm.adjoint(pr, f, cp, cc, reference_mf=reference_mf)[0].cpu().numpy()
        ref = xhm.vjp(pr, f.double(), cp.to(torch.complex128), cc.to(torch.complex128),
                      reference_mf=reference_mf).cpu().numpy()
        ok = np.arange(nb) != nb - 1
        assert np.all(np.isnan(got[~ok, :7])) and np.all(np.isfinite(got[ok]))
        diff, bar = np.abs(got[ok] - ref[ok]), float32_bar(ref[ok])
        excess = np.where(bar > 0, diff / np.where(bar > 0, bar, 1.0), np.where(diff == 0, 0.0, np.inf))
        print("worst fraction of the float32 bar per parameter", excess.max(axis=0), flush=True)
        assert np.all(excess <= 1.0), np.unravel_index(excess.argmax(), excess.shape)
        np.testing.assert_array_equal(got[swapped][[1, 0, 3, 2, 4, 5, 6, 7]], got[ordered])
        assert got[no_f_ref, 7] == 0.0 and (reference_mf is None or np.all(got[:, 7][ok] == 0.0))
    for b in (0, ordered):
        one = xhm.adjoint(p[b:b + 1], f, cp[b:b + 1], cc[b:b + 1])[0].cpu().numpy()[0]
        np.testing.assert_array_equal(one, xhm.adjoint(p, f, cp, cc)[0].cpu().numpy()[b])
    # the torch backward of the float32 function (split arithmetic), at the float32-rounded parameters
    b = 3
    fields = [torch.tensor(params[b, k], dtype=torch.float32, device=cuda, requires_grad=True) for k in range(8)]
    zero = fields[2] * 0
    binary = wx.Binary(fields[0], fields[1], torch.stack([zero, zero, fields[2]]),
                       torch.stack([zero, zero, fields[3]]), fields[4], fields[5], fields[6], fields[7])
    hp, hc = wt.polarizations_function("IMRPhenomXHM")(binary, f)
    loss = (torch.conj(cp[b]) * hp).real.sum() + (torch.conj(cc[b]) * hc).real.sum()
    loss.backward()
    torch_grads = np.array([float(x.grad) for x in fields])
    p32 = p[b:b + 1].float().double()
    expected = xhm.adjoint(p32, f, cp[b:b + 1], cc[b:b + 1])[0].cpu().numpy()[0]
    np.testing.assert_allclose(torch_grads, expected.astype(np.float32), rtol=1e-6)


def test_against_lalsimulation_aligned_xpnr(cuda):
    """The device's FP64 IMRPhenomXHM against LALSimulation 6.2.0's own
    polarizations of the two aligned-spin binaries of the reference's
    IMRPhenomXPNR dump (tests/golden/lal_xpnr_aligned.json): within a
    mismatch of 5e-6 and 5e-3 of the largest value, the residual of
    IMRPhenomXPNR's choices (the higher modes' release 122019, its final
    spin), as the oracle (tests/test_oracle_pins.py)."""
    import torch

    from test_oracle_pins import lal_cases, mismatch_and_error
    from waveformx_amd import xhm

    for index, params, f, hp, hc in lal_cases():
        got_p, got_c = xhm.polarizations(torch.as_tensor(params, device=cuda), torch.as_tensor(f, device=cuda),
                                         amp_scale=1.0)
        for got, ref in ((got_p, hp), (got_c, hc)):
            mism, err = mismatch_and_error(got.cpu().numpy()[0], ref)
            assert mism < 5e-6 and err < 5e-3, (index, mism, err)


def own_grids(params, rng):
    """Per-binary uniform grids as the ragged workload draws them: f_lower
    ~ U[10, 30] Hz, df = 1/T with T ~ log-U[4, 16] s, up to min(2048 Hz,
    0.3/M)."""
    grids = []
    for p in params:
        f_lower = rng.uniform(10.0, 30.0)
        df = 1.0 / np.exp(rng.uniform(np.log(4.0), np.log(16.0)))
        f_max = min(2048.0, 0.3 / ((p[0] + p[1]) * MTSUN))
        grids.append((f_lower, df, int(np.floor((f_max - f_lower) / df + 0.5)) + 1))
    return grids


def same_bits(a, b):
    """Equal values and equal NaN positions (complex tensors)."""
    import torch

    ra, rb = torch.view_as_real(a), torch.view_as_real(b)
    na, nb = torch.isnan(ra), torch.isnan(rb)
    return bool(torch.equal(na, nb)) and bool(torch.equal(torch.where(na, 0.0, ra), torch.where(nb, 0.0, rb)))


def test_own_grids(cuda):
    """Binaries each on their own grid (xhm.evaluate_ragged: one launch over
    256-bin tiles; the reference's per-binary (B, N) frequencies,
    xhm/mlx/_kernels.py:565-576, with a length per binary): each binary's
    slice, and its modes, equal the shared-grid evaluation of that binary
    alone on its grid bit for bit, in FP64 and split; the uniform grids
    generated on the device from (f0, df) equal the same grids given
    explicitly; FP64 against the oracle within the polarization bar on
    every binary; split against FP64 within the reference's float32
    mismatch bar; ragged_polarizations routes IMRPhenomXHM here; a failing
    binary's slice is NaN; empty grids are skipped."""
    import torch

    import waveformx_amd as wx
    from waveformx_amd import xhm

    rng = np.random.default_rng(41)
    params = oxhm.draw(10, seed=42)
    params[3, 2] = 1.5  # fails LAL's checks
    grids = own_grids(params, rng)
    grids[6] = (grids[6][0], grids[6][1], 0)  # an empty grid
    lengths = np.array([g[2] for g in grids])
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    f0 = np.array([g[0] for g in grids])
    df = np.array([g[1] for g in grids])
    flat = np.concatenate([g[0] + np.arange(g[2]) * g[1] for g in grids])
    p = torch.as_tensor(params, device=cuda)
    off = torch.as_tensor(offsets, device=cuda)
    results = {}
    for precision, rdtype in (("fp64", torch.float64), ("split", torch.float32)):
        print("own grids", precision, flush=True)
        records, head, modes_, _ = xhm.setup(p, torch.as_tensor(f0, device=cuda), precision=precision)
        fl = torch.as_tensor(flat, dtype=rdtype, device=cuda)
        hp, hc, hm = xhm.evaluate_ragged(records, head, modes_, off, freqs=fl, with_modes=True)
        gp, gc = xhm.evaluate_ragged(records, head, modes_, off, grid_f0=torch.as_tensor(f0, device=cuda),
                                     grid_df=torch.as_tensor(df, device=cuda))
        assert same_bits(hp, gp) and same_bits(hc, gc)
        for b in range(len(grids)):
            lo, hi = offsets[b], offsets[b + 1]
            if hi == lo:
                continue
            sp, sc, sm = xhm.evaluate(records[b:b + 1], head[b:b + 1], modes_[b:b + 1], fl[lo:hi], with_modes=True)
            assert same_bits(hp[lo:hi], sp[0]) and same_bits(hc[lo:hi], sc[0]), (precision, b)
            assert same_bits(hm[:, lo:hi].contiguous(), sm[:, 0].contiguous()), (precision, b)
        assert bool(torch.isnan(torch.view_as_real(hp[offsets[3]:offsets[4]])).all())
        results[precision] = (hp.cpu().numpy(), hc.cpu().numpy())
    # FP64 against the oracle, every binary that LAL accepts
    for b in range(len(grids)):
        lo, hi = offsets[b], offsets[b + 1]
        if hi == lo or b == 3:
            continue
        op, oc = oxhm.polarizations(params[b:b + 1], flat[lo:hi], amp_scale=1e21)
        bar = polarization_bar(params[b], flat[lo])
        for got, ref in ((results["fp64"][0][lo:hi], op[0]), (results["fp64"][1][lo:hi], oc[0])):
            assert relative_error(got, ref) <= bar, b
        # split against FP64 (at the same float32-rounded frequencies: the FP64 path on the float32 grid)
    records, head, modes_, _ = xhm.setup(p, torch.as_tensor(f0, device=cuda), precision="fp64")
    ref_p, _ = xhm.evaluate_ragged(records, head, modes_, off,
                                   freqs=torch.as_tensor(flat, dtype=torch.float32, device=cuda))
    ref_p = ref_p.cpu().numpy()
    for b in range(len(grids)):
        lo, hi = offsets[b], offsets[b + 1]
        if hi == lo or b == 3:
            continue
        a, r = results["split"][0][lo:hi].astype(np.complex128), ref_p[lo:hi]
        mism = 1.0 - np.vdot(r, a).real / np.sqrt(np.vdot(a, a).real * np.vdot(r, r).real)
        assert mism <= 1e-6, (b, mism)
    # the public route: ragged_polarizations with descriptors and with explicit grids
    spins = [np.zeros((len(params), 3)), np.zeros((len(params), 3))]
    spins[0][:, 2], spins[1][:, 2] = params[:, 2], params[:, 3]
    binaries = {"mass_1": params[:, 0], "mass_2": params[:, 1], "spin_1": spins[0], "spin_2": spins[1],
                "luminosity_distance": params[:, 4], "inclination": params[:, 5], "phi_ref": params[:, 6],
                "f_ref": params[:, 7]}
    hp_d, hc_d, off_d = wx.ragged_polarizations("IMRPhenomXHM", grids, binaries, device=cuda)
    np.testing.assert_array_equal(off_d.cpu().numpy(), offsets)
    np.testing.assert_array_equal(np.nan_to_num(hp_d.cpu().numpy()), np.nan_to_num(results["fp64"][0]))
    explicit = [g[0] + np.arange(g[2]) * g[1] for g in grids]
    hp_e, _, _ = wx.ragged_polarizations("IMRPhenomXHM", explicit, binaries, device=cuda)
    np.testing.assert_array_equal(np.nan_to_num(hp_e.cpu().numpy()), np.nan_to_num(results["fp64"][0]))


def test_per_binary_grid_rows(cuda):
    """Per-binary grids (B, N), the reference's batched frequencies
    (xhm/mlx/_api.py:79-93): xhm.polarizations and xhm.modes equal, row by
    row, the shared-grid evaluation of each binary on its own row (each set
    up at its own first frequency), in FP64 and split."""
    import torch

    from waveformx_amd import xhm

    params = oxhm.draw(5, seed=43)
    rows = np.stack([np.linspace(f_lo, 900.0, 700) for f_lo in (12.0, 20.0, 25.5, 30.0, 17.25)])
    p = torch.as_tensor(params, device=cuda)
    for precision in ("fp64", "split"):
        f = torch.as_tensor(rows, device=cuda)
        hp, hc = xhm.polarizations(p, f, precision=precision)
        modes = xhm.modes(p, f, precision=precision)
        assert hp.shape == (5, 700) and modes.shape == (5, 5, 700)
        for b in range(5):
            one_p, one_c = xhm.polarizations(p[b:b + 1], f[b], precision=precision)
            assert same_bits(hp[b], one_p[0]) and same_bits(hc[b], one_c[0]), (precision, b)
            assert same_bits(modes[:, b].contiguous(), xhm.modes(p[b:b + 1], f[b], precision=precision)[:, 0].contiguous())
    with pytest.raises(ValueError):
        xhm.polarizations(p, torch.as_tensor(rows[:3], device=cuda))
