"""The modes of IMRPhenomXHM in the conventions of
``SimInspiralChooseFDModes`` (``src/waveformx/_api.py:180-204``,
``models/xhm/api.py:40-126``, ``_results.py:72-97``):

* :class:`FrequencyModes` -- two-sided frequencies and ``{(l, m): h_lm}``
  with the line of sight ``(theta, phi)``; :meth:`FrequencyModes.polarizations`
  sums them with the spin-weighted spherical harmonics
  (``_harmonics.py:41-73``);
* :func:`modes` -- the modes at ``-f[::-1], f``: ``h_{l,-m}`` at the
  positive frequencies, ``h_{l,m}(-f) = (-1)^l conj(h_{l,-m}(f))`` at the
  negative ones;
* :func:`choose_fd_modes` -- the same on LAL's uniform two-sided grid
  ``k delta_f`` (zero below ``f_min`` and above ``min(f_max, f_cut)``).

The modes come from the FP64 IMRPhenomXHM kernels (:mod:`.xhm`).
"""

from __future__ import annotations

import dataclasses
from math import factorial, sqrt
from typing import Any, Dict, Optional, Tuple

import numpy as np
import torch

from ._types import WaveformOptions, check_frequencies, parse_options, strain_scale
from .exceptions import InvalidParameterError, UnsupportedError

__all__ = ["FrequencyModes", "spin_weighted_spherical_harmonic", "modes", "choose_fd_modes"]

MTSUN_SI = 4.925490947641267e-06


def _wigner_small_d(ell: int, mp: int, m: int, beta: float) -> float:
    """Wigner's d^l_{m' m}(beta) in LAL's convention (XLALWignerdMatrix)."""
    c, s = np.cos(beta / 2), np.sin(beta / 2)
    norm = sqrt(factorial(ell + mp) * factorial(ell - mp) * factorial(ell + m) * factorial(ell - m))
    total = 0.0
    for k in range(max(0, m - mp), min(ell + m, ell - mp) + 1):
        coefficient = (-1) ** (k - m + mp) * norm / (
            factorial(ell + m - k) * factorial(k) * factorial(ell - k - mp) * factorial(k - m + mp))
        total += coefficient * c ** (2 * ell - 2 * k + m - mp) * s ** (2 * k - m + mp)
    return total


def spin_weighted_spherical_harmonic(theta: float, phi: float, ell: int, m: int, s: int = -2) -> complex:
    """``{}_sY_{lm}(theta, phi)`` as ``XLALSpinWeightedSphericalHarmonic``."""
    return (-1) ** s * sqrt((2 * ell + 1) / (4 * np.pi)) * _wigner_small_d(ell, m, -s, theta) * \
        np.exp(1j * m * phi)


@dataclasses.dataclass
class FrequencyModes:
    """Modes ``h_lm`` on ``frequencies`` (two-sided), with the angles of the
    line of sight in the frame of the modes and the polarization angle."""

    frequencies: Any
    modes: Dict[Tuple[int, int], Any]
    theta: float = 0.0
    phi: float = 0.0
    polarization_angle: float = 0.0

    def __getitem__(self, mode):
        return self.modes[tuple(mode)]

    def __iter__(self):
        return iter((self.frequencies, self.modes))

    def polarizations(self) -> Tuple[torch.Tensor, torch.Tensor]:
        """``(h_+, h_x)`` on the non-negative frequencies:
        ``h_+ = 1/2 sum (h_lm(f) Y + conj(h_lm(-f) Y))``, ``h_x = i/2 sum
        (h_lm(f) Y - conj(h_lm(-f) Y))``, rotated by twice the polarization
        angle."""
        f = torch.as_tensor(self.frequencies)
        n2 = f.shape[0]
        n = int((f >= 0).sum())
        hp = hc = None
        for (ell, m), h in self.modes.items():
            h = torch.as_tensor(h)
            positive = h[n2 - n:]
            negative = torch.flip(h[:n], dims=(0,))
            y = spin_weighted_spherical_harmonic(self.theta, self.phi, ell, m)
            a = positive * y
            b = torch.conj(negative) * np.conj(y)
            hp = 0.5 * (a + b) if hp is None else hp + 0.5 * (a + b)
            hc = 0.5j * (a - b) if hc is None else hc + 0.5j * (a - b)
        cos, sin = np.cos(2 * self.polarization_angle), np.sin(2 * self.polarization_angle)
        return hp * cos + hc * sin, hc * cos - hp * sin


def _two_sided(modes5: torch.Tensor, frequencies: torch.Tensor, inclination: float) -> FrequencyModes:
    from .xhm import MODES

    zero = torch.zeros_like(modes5[0])
    out = {}
    for (ell, emm), h in zip(MODES, modes5):
        sign = -1.0 if ell % 2 else 1.0
        out[(ell, -emm)] = torch.cat([zero, h])
        out[(ell, emm)] = torch.cat([torch.flip(sign * torch.conj(h), dims=(0,)), zero])
    two = torch.cat([-torch.flip(frequencies, dims=(0,)), frequencies])
    return FrequencyModes(two, dict(sorted(out.items())), float(inclination), float(np.pi / 2), 0.0)


def _binary_modes(approximant: str, frequencies: np.ndarray, options: WaveformOptions, keywords: dict,
                  device: Optional[torch.device]) -> Tuple[torch.Tensor, np.ndarray]:
    """The five modes (5, N) of one binary at ``frequencies``, after the
    host checks and LAL's flags (raised as the reference raises them)."""
    from . import _api, xhm
    from ._types import BINARY_NAMES, check_table

    if approximant != "IMRPhenomXHM":
        _api.model(approximant)
        raise UnsupportedError(f"the modes of {approximant} are not implemented in waveformx")
    if options.mode_array is not None:
        raise UnsupportedError("IMRPhenomXHM computes every mode: select them from modes()")
    unknown = set(keywords) - set(BINARY_NAMES)
    if unknown:
        raise TypeError(f"unexpected keyword arguments {sorted(unknown)}")
    row = _api._binary_row(keywords)
    errors = check_table(row, approximant)
    _api._reference_mf_errors(row, options, errors)
    if errors:
        raise errors[0]
    dev = _api.compute_device(device)
    params = _api._table_params(row, dev)
    f = torch.as_tensor(frequencies, dtype=torch.float64, device=dev)
    records, head, modes_, flags = xhm.setup(params, f[:1], strain_scale(options), options.reference_mf)
    value = int(flags[0].item())
    for bit, message in xhm.LAL_FLAGS:
        if value & bit:
            raise InvalidParameterError(message)
    out = xhm.evaluate(records, head, modes_, f, with_modes=True)[2]
    return out[:, 0], row[0]


def modes(approximant: str, frequencies: Any, *, options: Optional[WaveformOptions] = None,
          device: Optional[torch.device] = None, **keywords) -> FrequencyModes:
    """The modes ``h_lm`` at ``+-frequencies`` (``SimInspiralChooseFDModes``
    conventions; ``src/waveformx/_api.py:180-187``) of one binary, FP64
    complex128 device tensors."""
    from ._types import BINARY_NAMES

    binary = {k: v for k, v in keywords.items() if k in BINARY_NAMES}
    options = parse_options(options, {k: v for k, v in keywords.items() if k not in BINARY_NAMES})
    frequencies = check_frequencies(frequencies)
    out, row = _binary_modes(approximant, frequencies, options, binary, device)
    f = torch.as_tensor(frequencies, dtype=torch.float64, device=out.device)
    return _two_sided(out, f, row[9])


def choose_fd_modes(approximant: str, delta_f: float, f_min: float, f_max: float = 0.0, *,
                    options: Optional[WaveformOptions] = None, device: Optional[torch.device] = None,
                    **keywords) -> FrequencyModes:
    """``SimInspiralChooseFDModes`` (``xhm/api.py:84-123``): the modes on
    LAL's two-sided uniform grid ``k delta_f``, zero below ``f_min`` and
    above ``min(f_max, f_cut)``, with ``f_ref = 0`` meaning ``f_min``."""
    from ._types import BINARY_NAMES

    if not delta_f > 0 or not f_min > 0 or f_max < 0:
        raise InvalidParameterError("delta_f and f_min must be positive and f_max non-negative")
    binary = {k: v for k, v in keywords.items() if k in BINARY_NAMES}
    options = parse_options(options, {k: v for k, v in keywords.items() if k not in BINARY_NAMES})
    m1, m2 = float(binary.get("mass_1", 0.0)), float(binary.get("mass_2", 0.0))
    chi1 = float(np.asarray(binary.get("spin_1", (0.0, 0.0, 0.0)), dtype=float)[2])
    chi2 = float(np.asarray(binary.get("spin_2", (0.0, 0.0, 0.0)), dtype=float)[2])
    if m1 < m2:
        m1, m2, chi1, chi2 = m2, m1, chi2, chi1
    eta = m1 * m2 / (m1 + m2) ** 2 if m1 > 0 and m2 > 0 else 0.25
    delta = np.sqrt(max(1.0 - 4.0 * eta, 0.0))
    chi_eff = 0.5 * (1.0 + delta) * chi1 + 0.5 * (1.0 - delta) * chi2
    f_cut = (0.33 if chi_eff > 0.99 else 0.3) / ((m1 + m2) * MTSUN_SI) if m1 + m2 > 0 else np.inf
    f_max_prime = min(f_max if f_max != 0.0 else f_cut, f_cut)
    i_start = int(f_min / delta_f)
    i_stop = int(f_max_prime / delta_f) + 1
    last = f_max if f_max_prime < f_max else f_max_prime
    n_full = int(2.0 ** np.ceil(np.log2(int(last / delta_f)))) + 1
    n = max(n_full, i_stop)
    if float(binary.get("f_ref", 0.0)) == 0.0 and options.reference_mf is None:
        binary = {**binary, "f_ref": f_min}
    band = np.arange(i_start, i_stop) * delta_f
    out, row = _binary_modes(approximant, band, options, binary, device)
    full = torch.zeros((5, n), dtype=torch.complex128, device=out.device)
    full[:, i_start:i_stop] = out
    full = full[:, :n_full]
    frequencies = torch.arange(n_full, dtype=torch.float64, device=out.device) * delta_f
    two = _two_sided(full, frequencies, row[9])
    # (a zero frequency appears once)
    keep = torch.as_tensor(np.r_[np.arange(n_full - 1), np.arange(n_full, 2 * n_full)], device=out.device)
    return FrequencyModes(two.frequencies[keep], {k: v[keep] for k, v in two.modes.items()}, two.theta, two.phi,
                          0.0)
