"""Converters between the reference conventions of IMRPhenomXAS and
IMRPhenomD (``src/waveformx/models/xas/reference.py:29-38``,
``src/waveformx/models/phenomd/reference.py:18-26``, both through
``src/waveformx/models/taylorf2/reference.py:46-127``).

In LAL's convention ``phi_ref`` is given at ``Binary.f_ref`` (Hz; zero: the
first frequency ``f_min``); in the mass-scaled one
(``WaveformOptions.reference_mf``) at ``M f_ref = reference_mf``. Only
``phi_ref`` changes: by half the change of the model's phase ``Phi``
between the two references, reduced modulo the model's period. For
IMRPhenomXAS ``Phi`` is LAL's (2,2) phase ``phase_22 / eta + linb M f`` and
the period ``4 pi`` (``phi_ref`` modulo ``2 pi``, which leaves every mode
unchanged); for IMRPhenomD it is ``-Psi + t0 M f`` and the period ``2 pi``
(``phi_ref`` modulo ``pi``). ``Phi`` is evaluated in FP64 on the GPU
(``wfx_xas_reference_phase``, ``wfx_phd_reference_phase``).

:func:`to_mass_scaled_reference` and :func:`to_physical_reference` are
IMRPhenomXAS's; :func:`converters` gives either model's pair.
"""

from __future__ import annotations

import ctypes
from typing import Any, Callable, Optional, Tuple

import numpy as np
import torch

from ._lib import call
from ._types import REFERENCE_MF, Binary

__all__ = ["phase", "to_mass_scaled_reference", "to_physical_reference", "converters"]

MTSUN_SI = 4.925490947641266978197229498498379006e-6
#: approximant -> (entry point of Phi, period of the phase change)
MODELS = {"IMRPhenomXAS": ("wfx_xas_reference_phase", 4.0 * np.pi),
          "IMRPhenomD": ("wfx_phd_reference_phase", 2.0 * np.pi)}
PERIOD = MODELS["IMRPhenomXAS"][1]


def phase(mass_1, mass_2, chi_1, chi_2, mf, device: Optional[torch.device] = None,
          approximant: str = "IMRPhenomXAS") -> torch.Tensor:
    """The model's phase ``Phi`` at ``M f = mf`` but for its constant
    (float64, one value per binary, on the device)."""
    from ._api import compute_device

    entry = MODELS[approximant][0]
    dev = compute_device(device)
    values = [torch.as_tensor(x, dtype=torch.float64).to(dev) for x in (mass_1, mass_2, chi_1, chi_2, mf)]
    values = torch.broadcast_tensors(*values)
    shape = values[0].shape
    m1, m2, c1, c2, mf_t = (v.reshape(-1) for v in values)
    ones, zeros = torch.ones_like(m1), torch.zeros_like(m1)
    params = torch.stack([m1, m2, c1, c2, ones, zeros, zeros, ones], dim=1).contiguous()
    mf_t = mf_t.contiguous()
    out = torch.empty_like(m1)
    call(entry, ctypes.c_void_p(params.data_ptr()), m1.shape[0], ctypes.c_void_p(mf_t.data_ptr()),
         ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    return out.reshape(shape)


def _fields(binary: Binary):
    """The fields as float64 tensors on the compute device, and the device
    of the input tensors (None: host values)."""
    from ._api import compute_device

    dev = None
    for value in binary:
        if isinstance(value, torch.Tensor):
            dev = value.device
            break
    run = compute_device(dev)
    as64 = [torch.as_tensor(np.asarray(v) if not isinstance(v, torch.Tensor) else v, dtype=torch.float64).to(run)
            for v in binary]
    return as64, dev


def _chi(spin: torch.Tensor) -> torch.Tensor:
    return spin[..., 2]


def _dtype(binary: Binary) -> torch.dtype:
    floats = [v for v in binary if isinstance(v, torch.Tensor) and v.is_floating_point()]
    return torch.float32 if floats and all(v.dtype == torch.float32 for v in floats) else torch.float64


def _shifted(binary: Binary, phi_from: torch.Tensor, phi_to: torch.Tensor, f_ref, device, period) -> Binary:
    change = phi_to - phi_from
    turns = torch.round(change / period)
    phi = torch.as_tensor(binary.phi_ref, dtype=torch.float64).to(change.device) + 0.5 * (change - turns * period)
    dtype = _dtype(binary)
    phi = phi.to(dtype)
    f_ref = torch.full_like(phi, float(f_ref)) if not isinstance(f_ref, torch.Tensor) else f_ref.to(phi)
    if device is not None:
        phi, f_ref = phi.to(device), f_ref.to(device)
    else:
        phi, f_ref = phi.cpu(), f_ref.cpu()
    return binary._replace(phi_ref=phi, f_ref=f_ref)


def converters(approximant: str = "IMRPhenomXAS") -> Tuple[Callable, Callable]:
    """``(to_mass_scaled_reference, to_physical_reference)`` of an
    approximant (taylorf2/reference.py:46-127 with the model's phase and
    period, f_ref = 0 meaning the first frequency)."""
    if approximant not in MODELS:
        raise ValueError(f"no reference converters for {approximant!r} (have {sorted(MODELS)})")
    period = MODELS[approximant][1]

    def to_mass_scaled(binary: Binary, reference_mf: float = REFERENCE_MF, f_min: Any = None) -> Binary:
        """``binary`` given at ``binary.f_ref`` (Hz; zero: ``f_min``, NaN
        without it) as a binary given at ``reference_mf / M`` (``f_ref``
        zero), for ``WaveformOptions(reference_mf=reference_mf)``."""
        (m1, m2, s1, s2, _, _, _, f_ref), device = _fields(binary)
        f = f_ref
        if f_min is None:
            f = torch.where(f == 0.0, torch.full_like(f, float("nan")), f)
        else:
            f = torch.where(f == 0.0, torch.as_tensor(f_min, dtype=torch.float64, device=f.device).expand_as(f), f)
        f = torch.where(f > 0.0, f, torch.full_like(f, float("nan")))
        phi_from = phase(m1, m2, _chi(s1), _chi(s2), f * (m1 + m2) * MTSUN_SI, device, approximant)
        phi_to = phase(m1, m2, _chi(s1), _chi(s2), torch.full_like(m1, float(reference_mf)), device, approximant)
        return _shifted(binary, phi_from, phi_to, 0.0, device, period)

    def to_physical(binary: Binary, f_ref: Any, reference_mf: float = REFERENCE_MF) -> Binary:
        """The inverse of the mass-scaled conversion: ``binary`` given at
        ``reference_mf / M`` (its ``f_ref`` ignored) as a binary given at
        ``f_ref`` (Hz)."""
        (m1, m2, s1, s2, _, _, _, _), device = _fields(binary)
        f = torch.as_tensor(f_ref, dtype=torch.float64, device=m1.device).expand_as(m1)
        f = torch.where(f > 0.0, f, torch.full_like(f, float("nan")))
        phi_from = phase(m1, m2, _chi(s1), _chi(s2), torch.full_like(m1, float(reference_mf)), device, approximant)
        phi_to = phase(m1, m2, _chi(s1), _chi(s2), f * (m1 + m2) * MTSUN_SI, device, approximant)
        return _shifted(binary, phi_from, phi_to, f_ref, device, period)

    to_mass_scaled.__name__ = "to_mass_scaled_reference"
    to_physical.__name__ = "to_physical_reference"
    return to_mass_scaled, to_physical


to_mass_scaled_reference, to_physical_reference = converters("IMRPhenomXAS")
