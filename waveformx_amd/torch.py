"""The PyTorch interface of the reference (``src/waveformx/torch.py``):
``polarizations_function(approximant, options=None, **option_keywords)``
returns ``f(binary, frequencies) -> (hplus, hcross)``.

Same inputs as the reference (``torch.py:122-172``): a :class:`Binary` or a
mapping of its fields (unknown field: ``TypeError``; missing masses or
spins: ``TypeError``), as tensors, numbers or sequences; float32 if every
floating tensor is float32, else float64; an optional leading batch axis
(``spin_*`` as ``(n, 3)``, the other fields broadcast; inconsistent batch
sizes: ``ValueError``). Same outputs: complex tensors ``(N,)`` or ``(n,
N)`` on the grid's device, the strain times 1e21 unless
``scaled_strain=False``, NaN rows where LAL fails, no parameter checks.

What runs is the HIP path: float64 inputs run the FP64 kernels (complex128
out), float32 inputs the split-precision FP32 kernels (complex64 out); the
keyword ``arithmetic`` ("fp64", "split", "split-df", "fp32"; IMRPhenomD has
no "fp32": its table is FP64) overrides the choice. Tensors on the CPU are
computed on the current HIP device and the results copied back.
Approximants: IMRPhenomXAS, IMRPhenomD and IMRPhenomXHM (arithmetics
"fp64" and "split"; its backward is the reverse-mode adjoint of
``xhm.adjoint``, to the binary's fields and the grid).

The backward pass (``torch.autograd.Function``, once differentiable, the
cotangents conjugated as in ``torch.py:108-119``) runs the model's
hand-written adjoint kernel: it recomputes the formula per bin, reduces the
cotangents into derivatives of each binary's record, and contracts them
with the record's forward-mode derivatives from the FP64 setup. Gradients
reach every field (the in-plane spins get zero) and the grid. As in the
reference, the function is opaque to ``torch.compile`` and ``torch.func``.
"""

from __future__ import annotations

from typing import Any, Callable, Mapping, Optional

import numpy as np
import torch

from . import _api, phenomd, xas, xhm
from ._types import BINARY_NAMES, Binary, WaveformOptions, parse_options, strain_scale

__all__ = ["polarizations_function"]


def _fields(binary: Any) -> list:
    if isinstance(binary, Binary):
        return list(binary)
    if isinstance(binary, Mapping):
        unknown = set(binary) - set(BINARY_NAMES)
        if unknown:
            raise TypeError(f"unknown fields of a Binary: {sorted(unknown)}")
        missing = [name for name in BINARY_NAMES[:4] if name not in binary]
        if missing:
            raise TypeError(f"missing fields of a Binary: {missing}")
        return [binary.get(name, Binary._field_defaults.get(name)) for name in BINARY_NAMES]
    if isinstance(binary, tuple) and len(binary) == len(BINARY_NAMES):
        return list(binary)
    raise TypeError(f"binary must be a Binary or a mapping of its fields, got {type(binary).__name__}")


def _tensors(value: Any) -> list:
    if isinstance(value, torch.Tensor):
        return [value]
    if isinstance(value, (list, tuple)):
        return [x for item in value for x in _tensors(item)]
    return []


def _as_tensor(value: Any, dtype: torch.dtype, device: torch.device) -> torch.Tensor:
    if isinstance(value, torch.Tensor):
        return value.to(device=device, dtype=dtype)
    if isinstance(value, (list, tuple)) and _tensors(value):
        return torch.stack([_as_tensor(x, dtype, device) for x in value])
    return torch.as_tensor(np.asarray(value, dtype=np.float64), dtype=dtype, device=device)


def _arithmetic(approximant: str, dtype: torch.dtype, arithmetic: Optional[str]) -> str:
    allowed = phenomd.PRECISIONS if approximant == "IMRPhenomD" else xhm.PRECISIONS \
        if approximant == "IMRPhenomXHM" else tuple(xas.PRECISIONS)
    if arithmetic is None:
        return "fp64" if dtype == torch.float64 else "split"
    if arithmetic not in allowed:
        raise ValueError(f"arithmetic of {approximant} must be one of {sorted(allowed)}")
    return arithmetic


class _ModelFunction(torch.autograd.Function):
    """(hplus, hcross) of the grid and the binary's fields (B,)-shaped, with
    the adjoint kernel as the backward pass (IMRPhenomXAS)."""

    @staticmethod
    def forward(ctx, key, grid, mass_1, mass_2, spin_1, spin_2, distance, inclination, phi_ref, f_ref):
        approximant, amp_scale, reference_mf, arithmetic = key
        params = xas.pack_params(mass_1, mass_2, spin_1[:, 2], spin_2[:, 2], distance, inclination, phi_ref, f_ref)
        freqs = grid.contiguous()
        if approximant == "IMRPhenomXHM":
            hp, hc = xhm.polarizations(params, freqs, amp_scale, reference_mf, arithmetic)
            if grid.dtype == torch.float32 and hp.dtype == torch.complex128:
                hp, hc = hp.to(torch.complex64), hc.to(torch.complex64)
        else:
            module = phenomd if approximant == "IMRPhenomD" else xas
            hp, hc = module.polarizations(params, freqs, amp_scale, reference_mf, arithmetic)
        ctx.key = key
        ctx.save_for_backward(grid, params)
        return hp, hc

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, grad_hplus, grad_hcross):
        from . import adjoint

        approximant, amp_scale, reference_mf, arithmetic = ctx.key
        grid, params = ctx.saved_tensors
        wrt_grid = ctx.needs_input_grad[1]
        grads, grad_grid = adjoint.vjp(params, grid.contiguous(), grad_hplus, grad_hcross, amp_scale,
                                       reference_mf, arithmetic, wrt_grid, approximant)
        dtype = grid.dtype
        g = grads.to(dtype)
        spin_1 = torch.zeros((g.shape[0], 3), dtype=dtype, device=g.device)
        spin_2 = torch.zeros_like(spin_1)
        spin_1[:, 2] = g[:, 2]
        spin_2[:, 2] = g[:, 3]
        out = [grad_grid.to(dtype) if grad_grid is not None else None, g[:, 0], g[:, 1], spin_1, spin_2, g[:, 4],
               g[:, 5], g[:, 6], g[:, 7]]
        return (None, *(o if need else None for o, need in zip(out, ctx.needs_input_grad[1:])))


def _wrap(approximant: str, options: WaveformOptions, arithmetic: Optional[str]) -> Callable:
    amp_scale = strain_scale(options)

    def function(binary: Any, grid: Any):
        values = [*_fields(binary), grid]
        tensors = [t for v in values for t in _tensors(v)]
        floats = [t for t in tensors if t.is_floating_point()]
        dtype = torch.float32 if floats and all(t.dtype == torch.float32 for t in floats) else torch.float64
        device = grid.device if isinstance(grid, torch.Tensor) else tensors[0].device if tensors else \
            torch.device("cpu")
        run = _api.compute_device(device)
        *fields, grid_t = (_as_tensor(v, dtype, run) for v in values)
        if grid_t.ndim != 1:
            raise ValueError("frequencies must be a 1-d array")
        sizes = {f.shape[0] for i, f in enumerate(fields) if f.ndim > (1 if i in (2, 3) else 0)}
        if len(sizes) > 1:
            raise ValueError(f"the batch axes of the binary's fields differ: {sorted(sizes)}")
        batched = bool(sizes)
        n = sizes.pop() if batched else 1
        fields = [torch.broadcast_to(f, (n, 3) if i in (2, 3) else (n,)) for i, f in enumerate(fields)]
        key = (approximant, amp_scale, options.reference_mf, _arithmetic(approximant, dtype, arithmetic))
        hp, hc = _ModelFunction.apply(key, grid_t, *fields)
        if not batched:
            hp, hc = hp[0], hc[0]
        if device != run:
            hp, hc = hp.to(device), hc.to(device)
        return hp, hc

    function.__name__ = "polarizations"
    function.__qualname__ = f"{approximant}.polarizations"
    return function


def polarizations_function(approximant: str, options: Optional[WaveformOptions] = None, *,
                           arithmetic: Optional[str] = None, **option_keywords) -> Callable:
    """The PyTorch function ``f(binary, frequencies) -> (hplus, hcross)`` of
    a frequency-domain approximant (``torch.py:175-183``)."""
    _api.model(approximant, "frequency")
    return _wrap(approximant, parse_options(options, option_keywords), arithmetic)
