"""Converters between IMRPhenomXAS's reference conventions
(``src/waveformx/models/xas/reference.py:29-38`` through
``src/waveformx/models/taylorf2/reference.py:46-127``).

In LAL's convention ``phi_ref`` is given at ``Binary.f_ref`` (Hz; zero: the
first frequency ``f_min``); in the mass-scaled one
(``WaveformOptions.reference_mf``) at ``M f_ref = reference_mf``. Only
``phi_ref`` changes: by half the change of LAL's (2,2) phase ``Phi`` (its
``phase_22 / eta + linb M f``) between the two references, reduced modulo
``4 pi`` (``phi_ref`` modulo ``2 pi``, which leaves every mode unchanged).
``Phi`` is evaluated in FP64 on the GPU (``wfx_xas_reference_phase``).
"""

from __future__ import annotations

import ctypes
from typing import Any, Optional

import numpy as np
import torch

from ._lib import call
from ._types import REFERENCE_MF, Binary

__all__ = ["phase", "to_mass_scaled_reference", "to_physical_reference"]

MTSUN_SI = 4.925490947641266978197229498498379006e-6
PERIOD = 4.0 * np.pi


def phase(mass_1, mass_2, chi_1, chi_2, mf, device: Optional[torch.device] = None) -> torch.Tensor:
    """LAL's (2,2) phase at ``M f = mf`` but for its constant (float64, one
    value per binary, on the device)."""
    from ._api import compute_device

    dev = compute_device(device)
    values = [torch.as_tensor(x, dtype=torch.float64).to(dev) for x in (mass_1, mass_2, chi_1, chi_2, mf)]
    values = torch.broadcast_tensors(*values)
    shape = values[0].shape
    m1, m2, c1, c2, mf_t = (v.reshape(-1) for v in values)
    ones, zeros = torch.ones_like(m1), torch.zeros_like(m1)
    params = torch.stack([m1, m2, c1, c2, ones, zeros, zeros, ones], dim=1).contiguous()
    mf_t = mf_t.contiguous()
    out = torch.empty_like(m1)
    call("wfx_xas_reference_phase", ctypes.c_void_p(params.data_ptr()), m1.shape[0],
         ctypes.c_void_p(mf_t.data_ptr()), ctypes.c_void_p(out.data_ptr()),
         ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
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


def _shifted(binary: Binary, phi_from: torch.Tensor, phi_to: torch.Tensor, f_ref, device) -> Binary:
    change = phi_to - phi_from
    turns = torch.round(change / PERIOD)
    phi = torch.as_tensor(binary.phi_ref, dtype=torch.float64).to(change.device) + 0.5 * (change - turns * PERIOD)
    dtype = _dtype(binary)
    phi = phi.to(dtype)
    f_ref = torch.full_like(phi, float(f_ref)) if not isinstance(f_ref, torch.Tensor) else f_ref.to(phi)
    if device is not None:
        phi, f_ref = phi.to(device), f_ref.to(device)
    else:
        phi, f_ref = phi.cpu(), f_ref.cpu()
    return binary._replace(phi_ref=phi, f_ref=f_ref)


def to_mass_scaled_reference(binary: Binary, reference_mf: float = REFERENCE_MF, f_min: Any = None) -> Binary:
    """``binary`` given at ``binary.f_ref`` (Hz; zero: ``f_min``, NaN
    without it) as a binary given at ``reference_mf / M`` (``f_ref`` zero),
    for ``WaveformOptions(reference_mf=reference_mf)``."""
    (m1, m2, s1, s2, _, _, _, f_ref), device = _fields(binary)
    f = f_ref
    if f_min is None:
        f = torch.where(f == 0.0, torch.full_like(f, float("nan")), f)
    else:
        f = torch.where(f == 0.0, torch.as_tensor(f_min, dtype=torch.float64, device=f.device).expand_as(f), f)
    f = torch.where(f > 0.0, f, torch.full_like(f, float("nan")))
    phi_from = phase(m1, m2, _chi(s1), _chi(s2), f * (m1 + m2) * MTSUN_SI, device)
    phi_to = phase(m1, m2, _chi(s1), _chi(s2), torch.full_like(m1, float(reference_mf)), device)
    return _shifted(binary, phi_from, phi_to, 0.0, device)


def to_physical_reference(binary: Binary, f_ref: Any, reference_mf: float = REFERENCE_MF) -> Binary:
    """The inverse of :func:`to_mass_scaled_reference`: ``binary`` given at
    ``reference_mf / M`` (its ``f_ref`` ignored) as a binary given at
    ``f_ref`` (Hz)."""
    (m1, m2, s1, s2, _, _, _, _), device = _fields(binary)
    f = torch.as_tensor(f_ref, dtype=torch.float64, device=m1.device).expand_as(m1)
    f = torch.where(f > 0.0, f, torch.full_like(f, float("nan")))
    phi_from = phase(m1, m2, _chi(s1), _chi(s2), torch.full_like(m1, float(reference_mf)), device)
    phi_to = phase(m1, m2, _chi(s1), _chi(s2), f * (m1 + m2) * MTSUN_SI, device)
    return _shifted(binary, phi_from, phi_to, f_ref, device)
