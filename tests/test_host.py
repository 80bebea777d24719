"""Host logic on the CPU: the value types and checks of the drop-in
boundary (the reference's _binary.py, _options.py, _results.py semantics),
the approximant registry's errors, and that the C-ABI library loads and
exports every entry point include/waveformx_amd.h declares (no kernel is
launched)."""

import re
from pathlib import Path

import numpy as np
import pytest

import waveformx_amd as wx
from waveformx_amd import _api, _types

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_header_symbols():
    header = (ROOT / "include" / "waveformx_amd.h").read_text()
    declared = set(re.findall(r"\b(wfx_\w+)\s*\(", header))
    assert {"wfx_xas_setup", "wfx_xas_eval", "wfx_xas_adjoint", "wfx_last_error"} <= declared
    from waveformx_amd import _lib

    lib = _lib.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert set(_lib.SIGNATURES) == declared
    assert lib.wfx_xas_record_slots() == 64
    assert lib.wfx_xas_record_bytes(0) == 512 and lib.wfx_xas_record_bytes(2) == 256
    # errors are reported, not raised, across the ABI
    assert lib.wfx_xas_record_bytes(7) == -1
    assert b"precision" in lib.wfx_last_error()
    assert lib.wfx_xas_setup(None, 4, None, 0, 1.0, float("nan"), 0, None, None, None) == -1
    assert b"null" in lib.wfx_last_error()
    assert lib.wfx_xas_eval(None, -1, None, 0, 10, 0, 0, 0, None, None, None) == -1
    assert lib.wfx_xas_eval(None, 4, None, 0, 10, 2, 0, 3, None, None, None) == -1
    assert lib.wfx_xas_setup_jvp(None, 4, None, 0, 1.0, float("nan"), None, None, None, None, None) == -1
    assert b"null" in lib.wfx_last_error()
    # empty batches are a no-op
    assert lib.wfx_xas_setup(None, 0, None, 0, 1.0, float("nan"), 0, None, None, None) == 0


def header_arities():
    header = (ROOT / "include" / "waveformx_amd.h").read_text()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    arities = {}
    for name, args in re.findall(r"\b(wfx_\w+)\s*\(([^)]*)\)\s*;", header):
        args = args.strip()
        arities[name] = 0 if args in ("", "void") else args.count(",") + 1
    return arities


def test_binding_and_stub_match_the_header():
    """The ctypes binding and the maintainer's stub in INTEGRATION.md declare
    as many arguments as the header does, and the library reports the
    binding's ABI version."""
    from waveformx_amd import _lib

    arities = header_arities()
    assert set(arities) == set(_lib.SIGNATURES)
    for name, (args, _) in _lib.SIGNATURES.items():
        assert len(args) == arities[name], name
    assert _lib.lib().wfx_abi_version() == _lib.ABI_VERSION
    stub = (ROOT / "INTEGRATION.md").read_text()
    found = re.findall(r"_lib\.(wfx_\w+)\.argtypes = \[([^\]]*)\]", stub)
    assert found
    for name, args in found:
        assert len([a for a in args.split(",") if a.strip()]) == arities[name], name
    assert f"wfx_abi_version() == {_lib.ABI_VERSION}" in stub


def test_parameter_table_forms():
    n = 3
    mapping = {"mass_1": np.array([30.0, 20.0, 10.0]), "mass_2": np.array([20.0, 10.0, 5.0]),
               "spin_1": np.zeros((n, 3)), "spin_2z": np.array([0.1, 0.2, 0.3])}
    t = wx.parameter_table(mapping)
    assert t.shape == (3, 12)
    assert np.all(t[:, 8] == 1.0) and np.all(t[:, 11] == 0.0)
    assert np.all(t[:, 7] == [0.1, 0.2, 0.3])
    assert np.array_equal(wx.parameter_table(t), t)
    with pytest.raises(TypeError, match="mass_2"):
        wx.parameter_table({"mass_1": np.ones(3)})
    with pytest.raises(TypeError, match="unknown"):
        wx.parameter_table({**mapping, "spin": np.ones(3)})
    with pytest.raises(ValueError, match="12 columns"):
        wx.parameter_table(np.ones((3, 5)))


def test_check_table_messages():
    t = np.zeros((6, 12))
    t[:, 0], t[:, 1], t[:, 8] = 30.0, 20.0, 400.0
    t[1, 0] = -1.0
    t[2, 2:5] = [0.1, 0.0, 0.2]
    t[3, 5:8] = [0.0, 0.0, 1.5]
    t[4, 8] = 0.0
    t[5, 11] = -3.0
    errors = _types.check_table(t)
    assert sorted(errors) == [1, 2, 3, 4, 5]
    assert "positive" in str(errors[1])
    assert "in-plane" in str(errors[2])
    assert "spin_2" in str(errors[3])
    assert "luminosity_distance" in str(errors[4])
    assert "f_ref" in str(errors[5])
    assert all(isinstance(e, wx.InvalidParameterError) and isinstance(e, ValueError) for e in errors.values())


def test_options():
    o = wx.WaveformOptions()
    assert o.scaled_strain and o.reference_mf is None
    assert _types.strain_scale(o) == 1e21
    assert _types.strain_scale(o.replace(scaled_strain=False)) == 1.0
    with pytest.raises(wx.InvalidParameterError):
        wx.WaveformOptions(reference_mf=0.02)
    with pytest.raises(TypeError):
        wx.WaveformOptions(scaled_strain=1)
    with pytest.raises(TypeError, match="unexpected"):
        _types.parse_options(None, {"bogus": 1})
    assert _types.parse_options(None, {"reference_mf": 0.01}).reference_mf == 0.01
    assert hash(o) == hash(wx.WaveformOptions())
    assert wx.WaveformOptions(precession_version="MSA").precession_version == 223


def test_frequency_grid():
    g = wx.FrequencyGrid.uniform(20.0, 30.0, 0.25)
    assert len(g) == 41 and g.frequencies[-1] == pytest.approx(30.0)
    with pytest.raises(wx.InvalidParameterError):
        wx.FrequencyGrid([20.0, 10.0])
    with pytest.raises(wx.InvalidParameterError):
        wx.FrequencyGrid([0.0, 10.0])
    with pytest.raises(wx.InvalidParameterError):
        wx.FrequencyGrid(np.ones((2, 2)))


def test_registry():
    assert wx.approximants() == ["IMRPhenomD", "IMRPhenomXAS", "IMRPhenomXHM"]
    assert _api.model("IMRPhenomXAS") == "IMRPhenomXAS"
    assert _api.model("IMRPhenomD") == "IMRPhenomD"
    assert _api.model("IMRPhenomXHM") == "IMRPhenomXHM"
    with pytest.raises(wx.UnknownApproximantError):
        _api.model("IMRPhenomQQ")
    with pytest.raises(wx.UnsupportedError):
        _api.model("IMRPhenomXPNR")
    with pytest.raises(wx.UnsupportedError):
        _api.model("NRSur7dq4")
    with pytest.raises(ValueError):
        wx.approximants("both")


def test_torch_function_input_errors():
    torch = pytest.importorskip("torch")
    import waveformx_amd.torch as wt

    f = wt.polarizations_function("IMRPhenomXAS")
    with pytest.raises(TypeError, match="unknown"):
        f({"mass_1": 1.0, "mass_2": 1.0, "spin_1": (0, 0, 0), "spin_2": (0, 0, 0), "mass": 2.0}, [20.0, 30.0])
    with pytest.raises(TypeError, match="missing"):
        f({"mass_1": 1.0}, [20.0, 30.0])
    with pytest.raises(TypeError):
        f(3.0, [20.0])
    with pytest.raises(wx.UnsupportedError):
        wt.polarizations_function("NRSur7dq4")
    if not torch.cuda.is_available():
        # there is no CPU path: a call without a HIP device raises
        with pytest.raises(RuntimeError, match="HIP device"):
            f(wx.Binary(30.0, 20.0, (0, 0, 0), (0, 0, 0)), torch.tensor([20.0, 30.0]))


def test_registry_queries_and_single_binary_argument_errors():
    """is_precessing / is_frequency_domain over the reference's registry
    (src/waveformx/_api.py:79-84), and polarizations()'s keyword errors,
    raised on the host before any device work (_binary.py:82-106)."""
    import waveformx_amd as wx

    assert wx.is_precessing("IMRPhenomXPNR") and not wx.is_precessing("IMRPhenomXAS")
    assert wx.is_frequency_domain("IMRPhenomD") and not wx.is_frequency_domain("NRSur7dq4")
    with pytest.raises(wx.UnknownApproximantError):
        wx.is_precessing("IMRPhenomXYZ")
    with pytest.raises(TypeError, match="mass_2"):
        wx.polarizations("IMRPhenomXAS", [20.0, 30.0], mass_1=30.0)
    with pytest.raises(TypeError, match="spin_1"):
        wx.polarizations("IMRPhenomXAS", [20.0, 30.0], mass_1=30.0, mass_2=20.0, spin_1=(0.0, 0.1))
    with pytest.raises(TypeError, match="unexpected keyword"):
        wx.polarizations("IMRPhenomXAS", [20.0, 30.0], mass_1=30.0, mass_2=20.0, colour="blue")
    with pytest.raises(wx.UnsupportedError):
        wx.polarizations("IMRPhenomXPNR", [20.0, 30.0], mass_1=30.0, mass_2=20.0)
    # parameter and grid errors come from the host checks, GPU or not
    # (the reference's tests/test_api.py:64-80)
    with pytest.raises(wx.InvalidParameterError, match="masses"):
        wx.polarizations("IMRPhenomXAS", [20.0, 30.0], mass_1=-30.0, mass_2=20.0)
    with pytest.raises(wx.InvalidParameterError, match="spin_1"):
        wx.polarizations("IMRPhenomXAS", [20.0, 30.0], mass_1=30.0, mass_2=20.0, spin_1=(0.9, 0.9, 0.0))
    with pytest.raises(wx.InvalidParameterError):
        wx.polarizations("IMRPhenomXAS", [30.0, 20.0], mass_1=30.0, mass_2=20.0)
    with pytest.raises(wx.InvalidParameterError, match="in-plane"):
        wx.polarizations("IMRPhenomD", [20.0, 30.0], mass_1=30.0, mass_2=20.0, spin_2=(0.0, 0.2, 0.3))
    with pytest.raises(ValueError):
        wx.polarizations("SEOBNRv4_ROM", [20.0, 30.0], mass_1=1.4, mass_2=1.4)


def test_model_contract_on_the_host():
    """models.get and the MODEL objects' host-side checks
    (src/waveformx/models/__init__.py:62-100, xas/api.py:42-48)."""
    from waveformx_amd import models

    xas_model = models.get("IMRPhenomXAS")
    assert xas_model.domain == "frequency" and not xas_model.precessing
    assert models.get("IMRPhenomD").approximant == "IMRPhenomD"
    with pytest.raises(wx.UnknownApproximantError):
        models.get("IMRPhenomXYZ")
    with pytest.raises(wx.UnsupportedError):
        models.get("NRSur7dq4")
    options = wx.WaveformOptions()
    with pytest.raises(wx.InvalidParameterError, match="in-plane"):
        xas_model.validate(wx.Binary(30.0, 20.0, (0.1, 0.0, 0.0), (0.0, 0.0, 0.0)), [20.0, 30.0], options)
    with pytest.raises(wx.InvalidParameterError, match="f_ref must be zero"):
        xas_model.validate(wx.Binary(30.0, 20.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), f_ref=20.0), [20.0, 30.0],
                           wx.WaveformOptions(reference_mf=0.01))
    xas_model.validate(wx.Binary(30.0, 20.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)), [], options)  # no grid: no flags
    with pytest.raises(wx.UnsupportedError):
        xas_model.modes(None, None, options)
    with pytest.raises(wx.InvalidParameterError, match="does not take"):
        models.check_reference_mf(wx.Binary(30.0, 20.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)),
                                  wx.WaveformOptions(reference_mf=0.01), "TaylorF2", supported=False)


def test_xhm_host_surface():
    """IMRPhenomXHM's model object, modes and harmonics on the host
    (``models/xhm/api.py:54-126``, ``_harmonics.py:33-38``)."""
    from waveformx_amd import _modes, models

    model = models.get("IMRPhenomXHM")
    assert model.domain == "frequency" and not model.precessing
    options = wx.WaveformOptions()
    with pytest.raises(wx.InvalidParameterError, match="IMRPhenomXHM is a non-precessing"):
        model.validate(wx.Binary(30.0, 20.0, (0.1, 0.0, 0.0), (0.0, 0.0, 0.0)), [20.0, 30.0], options)
    with pytest.raises(wx.UnsupportedError, match="every mode"):
        model.validate(wx.Binary(30.0, 20.0, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0)), [20.0, 30.0],
                       wx.WaveformOptions(mode_array=((2, 2),)))
    with pytest.raises(wx.UnsupportedError):
        wx.modes("IMRPhenomD", [20.0, 30.0], mass_1=30.0, mass_2=20.0)
    with pytest.raises(wx.InvalidParameterError, match="masses"):
        wx.modes("IMRPhenomXHM", [20.0, 30.0], mass_1=-30.0, mass_2=20.0)
    with pytest.raises(wx.InvalidParameterError, match="delta_f"):
        wx.choose_fd_modes("IMRPhenomXHM", 0.0, 20.0, mass_1=30.0, mass_2=20.0)
    with pytest.raises(wx.UnsupportedError):
        wx.ragged_polarizations("IMRPhenomXPNR", [np.array([20.0, 30.0])], np.ones((1, 12)))
    theta, phi = 0.8, 0.4
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    y44 = 3 * np.sqrt(7 / np.pi) * c ** 6 * s ** 2 * np.exp(4j * phi)
    assert abs(_modes.spin_weighted_spherical_harmonic(theta, phi, 4, 4) - y44) < 1e-15


def test_xhm_selection_slots_match_the_headers():
    """The record slots xhm.vjp holds fixed are the headers' region bounds
    and branch flags (csrc/xhm_setup.h HmSlot, csrc/xas_layout.h XasSlot)."""
    from waveformx_amd import xhm

    csrc = ROOT / "waveformx_amd" / "csrc"

    def slots(text):
        out, last = {}, -1
        body = re.sub(r"//.*", "", text)
        for name, value in re.findall(r"\b(k[A-Z]\w*)\s*(?:=\s*(\d+))?\s*[,}]", body):
            last = int(value) if value else last + 1
            out[name] = last
        return out

    hm = slots((csrc / "xhm_setup.h").read_text().split("enum HmSlot : int {")[1].split("};")[0] + "}")
    xas = slots((csrc / "xas_layout.h").read_text().split("enum XasSlot : int {")[1].split("};")[0] + "}")
    assert xhm.MODE_SELECTION_SLOTS == tuple(hm[k] for k in ("kHmTag", "kHmOn", "kHmFalloff", "kHmMixing",
                                                             "kHmFAmpIn", "kHmFAmpIm", "kHmFPhaseIn",
                                                             "kHmFPhaseIm", "kHmFRdAux", "kHmNInter"))
    assert xhm.XAS_SELECTION_SLOTS == tuple(xas[k] for k in ("kFPhaseIn", "kFPhaseIm", "kFAmpIn", "kFAmpRd", "kFCut"))
    assert (xhm._XAS_INS_CONST, xhm._XAS_ANCHOR) == (xas["kInsConst"], xas["kAnchor"])



def test_ragged_descriptor_and_offset_checks():
    """Uniform-grid descriptors need f0 > 0, df > 0 and n >= 0, and offsets
    read on the host must start at 0 and never decrease: lengths
    [100, -50, 10] would give offsets [0, 100, 50, 60] and hand binary 0
    bins past the end of a 60-bin output."""
    import torch

    from waveformx_amd import ragged, xas

    ok = np.array([20.0, 10.0]), np.array([0.25, 0.5]), np.array([100, 0])
    ragged.check_descriptors(*ok)
    for f0, df, n in ((ok[0], ok[1], np.array([100, -50])), (np.array([0.0, 10.0]), ok[1], ok[2]),
                      (ok[0], np.array([0.25, 0.0]), ok[2]), (ok[0], np.array([np.nan, 0.5]), ok[2])):
        with pytest.raises(ValueError, match="f0 > 0, df > 0 and n >= 0"):
            ragged.check_descriptors(f0, df, n)
    assert xas.checked_total(torch.tensor([0, 100, 100, 160])) == 160
    for bad in ([0, 100, 50, 60], [5, 10, 20]):
        with pytest.raises(ValueError, match="start at 0 and never decrease"):
            xas.checked_total(torch.tensor(bad))


def test_xhm_difference_stencils_stay_on_their_side():
    """The FP64 XHM gradient's differences: central away from the
    boundaries, one-sided towards the parameters' side where the central
    stencil would cross the set-up's swap of the labels (m1 = m2), the spin
    bound |chi| = 1 or the mass-ratio bound q = 1000 (xhm._directions); the
    one-sided stencil differentiates a quartic exactly."""
    import torch

    from waveformx_amd import xhm

    p = torch.tensor([[30.0, 20.0, 0.1, 0.2, 400.0, 0.4, 0.0, 20.0],
                      [30.0, 30.0, 0.5, -0.3, 400.0, 0.4, 0.0, 20.0],
                      [30.0, 30.0 + 1e-6, 0.5, -0.3, 400.0, 0.4, 0.0, 20.0],
                      [40.0, 20.0, 1.0 - 1e-7, -1.0 + 1e-7, 400.0, 0.4, 0.0, 20.0],
                      [1000.0, 1.0, 0.0, 0.0, 400.0, 0.4, 0.0, 20.0],
                      [1.0, 1000.0, 0.0, 0.0, 400.0, 0.4, 0.0, 20.0]], dtype=torch.float64)
    d = xhm._directions(p, xhm._steps(p))
    expected = {0: [0, 1, -1, 0, -1, 1], 1: [0, -1, 1, 0, 1, -1], 2: [0, 0, 0, -1, 0, 0], 3: [0, 0, 0, 1, 0, 0]}
    for j, values in expected.items():
        assert d[j].tolist() == values, (j, d[j].tolist())
    # the stencils: exact on polynomials of degree 4
    x0, h = 0.7, 1e-2
    poly = lambda x: 3 * x ** 4 - 2 * x ** 3 + x - 5  # noqa: E731
    exact = 12 * x0 ** 3 - 6 * x0 ** 2 + 1
    for side in (1, -1):
        est = xhm._FORWARD_BASE * side * poly(x0) + sum(side * w * poly(x0 + side * o * h) for o, w in xhm._FORWARD)
        assert abs(est / h - exact) < 1e-9
    assert abs(sum(w * poly(x0 + o * h) for o, w in xhm._STENCIL) / h - exact) < 1e-9


def test_fast_atan_coefficients_come_from_their_recipe():
    """The kernel's arctangent coefficients (csrc/xas_point.h) are the
    float32 output of csrc/gen_atan.py, literally; the recipe's fit bounds
    |error| by 7e-8 in exact arithmetic on [0, 1] and by 2e-7 in float32
    arithmetic over [0, 1e30]."""
    import importlib.util
    import re
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    spec = importlib.util.spec_from_file_location("gen_atan", root / "waveformx_amd" / "csrc" / "gen_atan.py")
    gen = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(gen)
    c = gen.coefficients()
    text = (root / "waveformx_amd" / "csrc" / "xas_point.h").read_text()
    body = text[text.index("__device__ inline float fast_atan"):]
    body = body[body.index("float p = "):body.index("p *= r;")]
    found = re.findall(r"(-?\d\.\d+(?:e-?\d+)?f)", body)
    assert found == gen.literals(c), found
    x = np.linspace(0.0, 1.0, 100001)
    exact = np.max(np.abs(x * np.polyval(c[::-1].astype(np.float64), x * x) - np.arctan(x)))
    assert exact < 7e-8, exact
    z = np.concatenate([x, np.geomspace(1.0, 1e30, 100001)])
    assert np.max(np.abs(gen.fast_atan_f32(z, c).astype(np.float64) - np.arctan(z))) < 2e-7
