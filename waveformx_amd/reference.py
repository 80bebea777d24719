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
IMRPhenomXAS's; :func:`converters` gives either model's pair, and
IMRPhenomXHM's, which are IMRPhenomXAS's (its higher modes are aligned with
the (2,2) at the reference, so the period 4 pi keeps every mode). The
package-level ``waveformx_amd.to_mass_scaled_reference`` and
``to_physical_reference`` dispatch on ``approximant=`` as the reference's
(``src/waveformx/_reference.py:36-79``). The converters are differentiable:
``Phi``'s derivatives come from the same FP64 kernel in forward mode
(``wfx_xas_reference_phase_grad``, ``wfx_phd_reference_phase_grad``), so a
gradient of the converted ``phi_ref`` reaches the masses, spins, ``phi_ref``,
``f_ref`` and ``f_min`` given as tensors, as ``jax.grad`` does through the
reference's (``tests/test_xas.py:174-175``).
"""

from __future__ import annotations

import ctypes
from typing import Any, Callable, Optional, Tuple

import numpy as np
import torch

from ._lib import call
from ._types import REFERENCE_MF, Binary

__all__ = ["phase", "to_mass_scaled_reference", "to_physical_reference", "converters", "REFERENCE_MODULES",
           "reference_module", "package_to_mass_scaled_reference", "package_to_physical_reference"]

MTSUN_SI = 4.925490947641266978197229498498379006e-6
#: approximant -> (entry point of Phi and its derivatives, period of the
#: phase change); IMRPhenomXHM's reference is its (2,2)'s, IMRPhenomXAS's
#: (src/waveformx/_reference.py:40)
MODELS = {"IMRPhenomXAS": ("wfx_xas_reference_phase_grad", 4.0 * np.pi),
          "IMRPhenomXHM": ("wfx_xas_reference_phase_grad", 4.0 * np.pi),
          "IMRPhenomD": ("wfx_phd_reference_phase_grad", 2.0 * np.pi)}
PERIOD = MODELS["IMRPhenomXAS"][1]


class _Phase(torch.autograd.Function):
    """``Phi`` of flat float64 device tensors (m1, m2, chi1, chi2, mf); the
    backward multiplies by its forward-mode derivatives from the same
    kernel, as the reference's jax.grad through the converters."""

    @staticmethod
    def forward(ctx, entry, m1, m2, c1, c2, mf):
        ones, zeros = torch.ones_like(m1), torch.zeros_like(m1)
        params = torch.stack([m1, m2, c1, c2, ones, zeros, zeros, ones], dim=1).contiguous()
        mf = mf.contiguous()
        out = torch.empty_like(m1)
        grad = torch.empty((m1.shape[0], 5), dtype=torch.float64, device=m1.device)
        call(entry, ctypes.c_void_p(params.data_ptr()), m1.shape[0], ctypes.c_void_p(mf.data_ptr()),
             ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(grad.data_ptr()),
             ctypes.c_void_p(torch.cuda.current_stream(m1.device).cuda_stream))
        ctx.save_for_backward(grad)
        return out

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return (None,) + tuple(g * grad[:, k] for k in range(5))


def phase(mass_1, mass_2, chi_1, chi_2, mf, device: Optional[torch.device] = None,
          approximant: str = "IMRPhenomXAS") -> torch.Tensor:
    """The model's phase ``Phi`` at ``M f = mf`` but for its constant
    (float64, one value per binary, on the device), differentiable with
    respect to every argument given as a tensor."""
    from ._api import compute_device

    entry = MODELS[approximant][0]
    dev = compute_device(device)
    values = [torch.as_tensor(x, dtype=torch.float64).to(dev) for x in (mass_1, mass_2, chi_1, chi_2, mf)]
    values = torch.broadcast_tensors(*values)
    shape = values[0].shape
    return _Phase.apply(entry, *(v.reshape(-1) for v in values)).reshape(shape)


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
    period, f_ref = 0 meaning the first frequency); IMRPhenomXHM's are
    IMRPhenomXAS's (src/waveformx/_reference.py:40). Differentiable: the
    results' ``phi_ref`` carries gradients to the input tensors (masses,
    spins, ``phi_ref``, ``f_ref``, ``f_min``)."""
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


#: approximant -> the approximant whose converters it uses (the reference's
#: REFERENCE_MODULES, src/waveformx/_reference.py:36-47, for the models built here)
REFERENCE_MODULES = {"IMRPhenomXAS": "IMRPhenomXAS", "IMRPhenomXHM": "IMRPhenomXAS", "IMRPhenomD": "IMRPhenomD"}
_PRECESSING = ("IMRPhenomXPNR", "IMRPhenomXPNR_NNLO", "IMRPhenomXPNR_MSA", "NRSur7dq4", "TaylorF2")


def reference_module(approximant: str) -> Tuple[Callable, Callable]:
    """The converters of ``approximant`` (LAL's name): the pair of
    :func:`converters` (src/waveformx/_reference.py:50-58)."""
    from .exceptions import UnknownApproximantError, UnsupportedError

    if approximant in _PRECESSING:
        raise UnsupportedError(f"{approximant}'s converters are outside this build, which has those of "
                               f"{', '.join(REFERENCE_MODULES)}")
    if approximant not in REFERENCE_MODULES:
        raise UnknownApproximantError(f"unknown approximant {approximant!r}: waveformx implements "
                                      f"{', '.join(REFERENCE_MODULES)}")
    return converters(REFERENCE_MODULES[approximant])


def package_to_mass_scaled_reference(binary: Binary, reference_mf: float = REFERENCE_MF, *args,
                                     approximant: str = "IMRPhenomXPNR", **keywords) -> Binary:
    """``waveformx.to_mass_scaled_reference`` (src/waveformx/_reference.py:
    61-71): ``binary`` at ``binary.f_ref`` (Hz) as a binary at the
    mass-scaled reference ``reference_mf / M``, for ``approximant`` (whose
    further arguments follow: ``f_min`` where ``f_ref = 0`` is the first
    frequency). ``approximant`` defaults to IMRPhenomXPNR, as the
    reference's, whose converters are not built here."""
    return reference_module(approximant)[0](binary, reference_mf, *args, **keywords)


def package_to_physical_reference(binary: Binary, f_ref: Any, reference_mf: float = REFERENCE_MF, *args,
                                  approximant: str = "IMRPhenomXPNR", **keywords) -> Binary:
    """``waveformx.to_physical_reference`` (src/waveformx/_reference.py:
    74-79): the inverse of :func:`package_to_mass_scaled_reference`."""
    return reference_module(approximant)[1](binary, f_ref, reference_mf, *args, **keywords)
