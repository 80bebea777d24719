"""The models and their contract with the public functions
(``src/waveformx/models/__init__.py:1-100``): :func:`get` gives the
``MODEL`` of an approximant built here, an object with

* ``validate(binary, frequencies, options)``: LALSimulation's checks of the
  model for concrete values (raises :class:`InvalidParameterError`): the
  mass-scaled reference's rule, no in-plane spins, and LAL's flags, which
  come from the setup kernel on the device (``xas/api.py:31-60``,
  ``phenomd/api.py:27-43``);
* ``polarizations(binary, frequencies, options)``: the pure function (no
  checks; NaN rows where LAL fails), through the torch function;
* ``modes`` / ``choose_fd_modes``: IMRPhenomXHM's in the conventions of
  ``SimInspiralChooseFDModes``; not implemented for IMRPhenomXAS and
  IMRPhenomD, as in the reference (:class:`UnsupportedError`).
"""

from __future__ import annotations

from typing import Any

import numpy as np
import torch

from ._api import APPROXIMANTS, compute_device, model as _check_built
from ._types import WaveformOptions, strain_scale
from .exceptions import InvalidParameterError, UnsupportedError

__all__ = ["APPROXIMANTS", "get", "check_reference_mf", "AlignedSpinModel", "XHMModel"]


def check_reference_mf(binary: Any, options: WaveformOptions, approximant: str, supported: bool = True) -> None:
    """The checks of the mass-scaled reference (``options.reference_mf``)
    for concrete values: refused where not ``supported``, and
    ``binary.f_ref`` must be zero with it (``models/__init__.py:87-100``)."""
    if options.reference_mf is None:
        return
    if not supported:
        raise InvalidParameterError(f"{approximant} does not take reference_mf")
    if float(binary.f_ref) != 0.0:
        raise InvalidParameterError(
            "with reference_mf, the reference frequency is reference_mf / M: f_ref must be zero "
            "(to_mass_scaled_reference converts a binary given at f_ref)")


def _spin(value: Any) -> np.ndarray:
    if isinstance(value, torch.Tensor):
        value = value.detach().cpu().numpy()
    return np.asarray([float(x.detach().cpu()) if isinstance(x, torch.Tensor) else float(x) for x in value]
                      if isinstance(value, (list, tuple)) else value, dtype=float).reshape(3)


def _scalar(value: Any) -> float:
    return float(value.detach().cpu()) if isinstance(value, torch.Tensor) else float(value)


class AlignedSpinModel:
    """The ``MODEL`` of IMRPhenomXAS or IMRPhenomD
    (``SimInspiralChooseFDWaveformSequence``)."""

    domain = "frequency"
    precessing = False

    def __init__(self, approximant: str):
        self.approximant = approximant

    def _module(self):
        from . import phenomd, xas

        return (phenomd, phenomd.LAL_FLAGS) if self.approximant == "IMRPhenomD" else (xas, xas.LAL_FLAGS)

    def validate(self, binary: Any, frequencies: Any, options: WaveformOptions) -> None:
        """LALSimulation's checks of the model for one concrete binary."""
        check_reference_mf(binary, options, self.approximant)
        spins = [_spin(binary.spin_1), _spin(binary.spin_2)]
        if any(np.any(s[:2] != 0.0) for s in spins):
            raise InvalidParameterError(f"non-zero in-plane spins were given, but {self.approximant} is a "
                                        "non-precessing approximant")
        frequencies = np.asarray(frequencies, dtype=float)
        if frequencies.size == 0 and self.approximant != "IMRPhenomD":
            return  # (IMRPhenomD's one flag, the final spin, needs no grid)
        f_min = float(frequencies[0]) if frequencies.size else 20.0
        row = [_scalar(binary.mass_1), _scalar(binary.mass_2), spins[0][2], spins[1][2],
               _scalar(binary.luminosity_distance), _scalar(binary.inclination), _scalar(binary.phi_ref),
               _scalar(binary.f_ref)]
        dev = compute_device()
        module, flags = self._module()
        params = torch.as_tensor([row], dtype=torch.float64, device=dev)
        _, bits = module.setup(params, torch.tensor([f_min], dtype=torch.float64, device=dev),
                               strain_scale(options), options.reference_mf, "fp64")
        value = int(bits[0].item())
        for bit, message in flags:
            if value & bit:
                raise InvalidParameterError(message)

    def polarizations(self, binary: Any, frequencies: Any, options: WaveformOptions):
        """``SimInspiralChooseFDWaveformSequence`` at ``frequencies`` (the
        torch function: FP64 for float64 inputs, split FP32 for float32)."""
        from .torch import polarizations_function

        return polarizations_function(self.approximant, options)(binary, frequencies)

    def modes(self, binary: Any, frequencies: Any, options: WaveformOptions):
        raise UnsupportedError(f"the modes of {self.approximant} are not implemented in waveformx")

    def choose_fd_modes(self, binary: Any, delta_f: float, f_min: float, f_max: float, options: WaveformOptions):
        raise UnsupportedError(f"the modes of {self.approximant} are not implemented in waveformx")


class XHMModel(AlignedSpinModel):
    """The ``MODEL`` of IMRPhenomXHM (``models/xhm/api.py:54-126``):
    ``validate`` as IMRPhenomXAS's with the name IMRPhenomXHM (and no mode
    array), ``polarizations`` through the torch function, ``modes`` and
    ``choose_fd_modes`` in ``SimInspiralChooseFDModes``'s conventions."""

    def __init__(self):
        super().__init__("IMRPhenomXHM")

    def _module(self):
        from . import xhm

        return xhm, xhm.LAL_FLAGS

    def validate(self, binary: Any, frequencies: Any, options: WaveformOptions) -> None:
        if options.mode_array is not None:
            raise UnsupportedError("IMRPhenomXHM computes every mode: select them from modes()")
        check_reference_mf(binary, options, self.approximant)
        spins = [_spin(binary.spin_1), _spin(binary.spin_2)]
        if any(np.any(s[:2] != 0.0) for s in spins):
            raise InvalidParameterError("non-zero in-plane spins were given, but IMRPhenomXHM is a non-precessing "
                                        "approximant")
        frequencies = np.asarray(frequencies, dtype=float)
        if frequencies.size == 0:
            return
        from . import xhm

        row = [_scalar(binary.mass_1), _scalar(binary.mass_2), spins[0][2], spins[1][2],
               _scalar(binary.luminosity_distance), _scalar(binary.inclination), _scalar(binary.phi_ref),
               _scalar(binary.f_ref)]
        dev = compute_device()
        params = torch.as_tensor([row], dtype=torch.float64, device=dev)
        *_, flags = xhm.setup(params, torch.tensor([float(frequencies[0])], dtype=torch.float64, device=dev),
                              strain_scale(options), options.reference_mf)
        value = int(flags[0].item())
        for bit, message in xhm.LAL_FLAGS:
            if value & bit:
                raise InvalidParameterError(message)

    def modes(self, binary: Any, frequencies: Any, options: WaveformOptions):
        from ._modes import modes

        return modes(self.approximant, frequencies, options=options, **_keywords(binary))

    def choose_fd_modes(self, binary: Any, delta_f: float, f_min: float, f_max: float, options: WaveformOptions):
        from ._modes import choose_fd_modes

        return choose_fd_modes(self.approximant, delta_f, f_min, f_max, options=options, **_keywords(binary))


def _keywords(binary: Any) -> dict:
    return dict(mass_1=_scalar(binary.mass_1), mass_2=_scalar(binary.mass_2), spin_1=tuple(_spin(binary.spin_1)),
                spin_2=tuple(_spin(binary.spin_2)), luminosity_distance=_scalar(binary.luminosity_distance),
                inclination=_scalar(binary.inclination), phi_ref=_scalar(binary.phi_ref), f_ref=_scalar(binary.f_ref))


_MODELS = {name: AlignedSpinModel(name) for name in ("IMRPhenomXAS", "IMRPhenomD")}
_MODELS["IMRPhenomXHM"] = XHMModel()


def get(approximant: str) -> AlignedSpinModel:
    """The ``MODEL`` of the approximant named ``approximant`` (LAL's name):
    :class:`UnknownApproximantError` for names the reference does not know,
    :class:`UnsupportedError` for those not built here."""
    _check_built(approximant)
    return _MODELS[approximant]
