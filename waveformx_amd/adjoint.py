"""Gradients of IMRPhenomXAS and IMRPhenomD on the GPU through the
hand-written adjoints (IMRPhenomXHM's, :mod:`waveformx_amd.xhm`'s, are
routed here too).

:func:`vjp` is the backward pass of the torch function (the reference's
``jax.vjp``, ``src/waveformx/torch.py:87-119``): the FP64 setup with
forward-mode derivatives gives each binary's record and d record /
d(m1, m2, chi1, chi2, d_L, inclination, phi_ref, f_ref); the adjoint kernel
recomputes the formula per bin, reduces the cotangents into d L / d record
and contracts with those derivatives. Nothing of the forward pass is stored.

:func:`loss_and_grad` fuses the forward pass with a noise-weighted
least-squares loss ``L = sum_f w(f) (|h+ - d+|^2 + |hx - dx|^2)``: it reads
the data and writes only each binary's loss and gradient. IMRPhenomXHM's
gradient, in either arithmetic, is :func:`waveformx_amd.xhm.adjoint`'s (the
dual-number set-up and a reverse-mode pass per mode over the bins, the
loss fused, the grid's gradient on request).

Both models share the kernels' shape (one wave per binary, the same
reduce-scatter and contraction; one thread per frequency for the grid's
gradient); each has its own per-bin adjoint and its own entry points.
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch

from . import phenomd, xas
from ._lib import call, lib
from .xas import _code, _ptr, _require_device, _stream

__all__ = ["vjp", "adjoint", "loss_and_grad", "PARAMETERS", "MODELS"]

PARAMETERS = ("mass_1", "mass_2", "chi_1", "chi_2", "luminosity_distance", "inclination", "phi_ref", "f_ref")

#: approximant -> (derivative setup, adjoint, fused loss adjoint, grid
#: adjoint) entry points
MODELS = {
    "IMRPhenomXAS": ("wfx_xas_setup_jvp", "wfx_xas_adjoint", "wfx_xas_loss_adjoint", "wfx_xas_grid_adjoint"),
    "IMRPhenomD": ("wfx_phd_setup_jvp", "wfx_phd_adjoint", "wfx_phd_loss_adjoint", "wfx_phd_grid_adjoint"),
}


def _entries(approximant: str, precision: str) -> Tuple[str, str, str, str]:
    if approximant not in MODELS:
        raise ValueError(f"no adjoint for {approximant!r}; gradients exist for {sorted(MODELS)}")
    if approximant == "IMRPhenomD" and precision not in phenomd.PRECISIONS:
        raise ValueError(f"IMRPhenomD's arithmetic is one of {phenomd.PRECISIONS}, got {precision!r}")
    return MODELS[approximant]


def _records(params, freqs, amp_scale, reference_mf, precision, approximant="IMRPhenomXAS"):
    """The forward precision's records and the FP64 records' derivatives
    (one launch; plain FP32 records come from their own FP32 setup)."""
    setup_entry = _entries(approximant, precision)[0]
    if precision in ("fp64", "split", "split-df"):
        records, jac, _ = xas.run_setup_jvp(setup_entry, params, freqs[:1], amp_scale, reference_mf, precision)
        return records, jac
    _, jac, _ = xas.run_setup_jvp(setup_entry, params, freqs[:1], amp_scale, reference_mf, "split")
    records, _ = xas.setup(params, freqs[:1], amp_scale, reference_mf, precision)
    return records, jac


power_table = xas.power_table


def adjoint(records: torch.Tensor, jac: torch.Tensor, freqs: torch.Tensor, cot_plus: torch.Tensor,
            cot_cross: torch.Tensor, precision: str, powers: Optional[torch.Tensor] = None,
            approximant: str = "IMRPhenomXAS") -> torch.Tensor:
    """(B, 8) float64 gradients from the cotangents (B, N) (complex, the
    torch convention dL/dRe + i dL/dIm) on the shared grid ``freqs``
    (``powers``: its :func:`power_table`, made here if not given)."""
    entry = _entries(approximant, precision)[1]
    for name, t in (("records", records), ("jac", jac), ("freqs", freqs), ("cot_plus", cot_plus),
                    ("cot_cross", cot_cross)):
        _require_device(t, name)
    nb, n = records.shape[0], freqs.shape[0]
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    if cot_plus.shape != (nb, n) or cot_cross.shape != (nb, n) or cot_plus.dtype != ctype or \
            cot_cross.dtype != ctype:
        raise ValueError(f"cotangents must be ({nb}, {n}) {ctype}")
    grads = torch.empty((nb, 8), dtype=torch.float64, device=records.device)
    if powers is None:
        powers = power_table(freqs, precision)
    call(entry, _ptr(records), _ptr(jac), nb, _ptr(freqs), int(freqs.dtype == torch.float64), n,
         _code(precision), _ptr(cot_plus), _ptr(cot_cross), _ptr(powers), _ptr(grads), _stream(records.device))
    return grads


def grid_adjoint(records: torch.Tensor, freqs: torch.Tensor, cot_plus: torch.Tensor,
                 cot_cross: torch.Tensor, precision: str, approximant: str = "IMRPhenomXAS") -> torch.Tensor:
    """dL/df (N,) float64 of the shared grid from the cotangents (B, N), the
    tables held (the first frequency's effect through the setup excluded, as
    the reference's stop_gradient on f_min)."""
    entry = _entries(approximant, precision)[3]
    nb, n = records.shape[0], freqs.shape[0]
    dldf = torch.empty((n,), dtype=torch.float64, device=records.device)
    size = lib().wfx_xas_grid_adjoint_workspace(nb, n)
    if size < 0:
        raise ValueError("the grid's gradient is too large")
    workspace = torch.empty((size,), dtype=torch.uint8, device=records.device)
    call(entry, _ptr(records), nb, _ptr(freqs), int(freqs.dtype == torch.float64), n,
         _code(precision), _ptr(cot_plus), _ptr(cot_cross), _ptr(dldf), _ptr(workspace), workspace.numel(),
         _stream(records.device))
    return dldf


def vjp(params: torch.Tensor, freqs: torch.Tensor, grad_plus: Optional[torch.Tensor],
        grad_cross: Optional[torch.Tensor], amp_scale: float, reference_mf: Optional[float], precision: str,
        wrt_grid: bool = False, approximant: str = "IMRPhenomXAS") -> Tuple[torch.Tensor, Optional[torch.Tensor]]:
    """Gradients (B, 8) with respect to the packed parameters, and (N,)
    with respect to the grid if ``wrt_grid`` (else None). IMRPhenomXHM's is
    :func:`waveformx_amd.xhm.adjoint` (exact, FP64 on the float64 grid or
    FP32 on the float32 grid)."""
    if approximant == "IMRPhenomXHM":
        from . import xhm

        if precision not in xhm.PRECISIONS:
            raise ValueError(f"IMRPhenomXHM's arithmetic is one of {xhm.PRECISIONS}, got {precision!r}")
        nb, n = params.shape[0], freqs.shape[0]
        ctype = torch.complex128 if precision == "fp64" else torch.complex64
        zeros = torch.zeros((nb, n), dtype=ctype, device=params.device)
        gp = zeros if grad_plus is None else grad_plus.reshape(nb, n)
        gc = zeros if grad_cross is None else grad_cross.reshape(nb, n)
        grads, _, grid = xhm.adjoint(params, freqs, gp, gc, None, amp_scale, reference_mf, precision, wrt_grid)
        return grads, grid
    _entries(approximant, precision)
    nb, n = params.shape[0], freqs.shape[0]
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    zeros = None
    if grad_plus is None or grad_cross is None:
        zeros = torch.zeros((nb, n), dtype=ctype, device=params.device)
    gp = zeros if grad_plus is None else grad_plus.to(ctype).reshape(nb, n).contiguous()
    gc = zeros if grad_cross is None else grad_cross.to(ctype).reshape(nb, n).contiguous()
    records, jac = _records(params, freqs, amp_scale, reference_mf, precision, approximant)
    grid = grid_adjoint(records, freqs, gp, gc, precision, approximant) if wrt_grid else None
    return adjoint(records, jac, freqs, gp, gc, precision, approximant=approximant), grid


def loss_and_grad(params: torch.Tensor, freqs: torch.Tensor, data_plus: torch.Tensor, data_cross: torch.Tensor,
                  weights: torch.Tensor, amp_scale: float = 1e21, reference_mf: Optional[float] = None,
                  precision: str = "split", records: Optional[Tuple[torch.Tensor, torch.Tensor]] = None,
                  powers: Optional[torch.Tensor] = None,
                  approximant: str = "IMRPhenomXAS") -> Tuple[torch.Tensor, torch.Tensor]:
    """Per-binary loss (B,) and gradient (B, 8), float64, of
    ``sum_f w(f) (|h+ - d+|^2 + |hx - dx|^2)`` in one fused pass
    (``records``: (records, jacobian) if already set up; ``powers``: the
    grid's :func:`power_table`, made here if not given; neither for
    IMRPhenomXHM)."""
    xhm_model = approximant == "IMRPhenomXHM"
    if xhm_model:
        from . import xhm

        if precision not in xhm.PRECISIONS:
            raise ValueError(f"IMRPhenomXHM's arithmetic is one of {xhm.PRECISIONS}, got {precision!r}")
        if records is not None or powers is not None:
            raise ValueError("IMRPhenomXHM's loss takes neither records nor powers")
    else:
        entry = _entries(approximant, precision)[2]
    for name, t in (("params", params), ("freqs", freqs), ("data_plus", data_plus), ("data_cross", data_cross),
                    ("weights", weights)):
        _require_device(t, name)
    nb, n = params.shape[0], freqs.shape[0]
    rdtype = torch.float64 if precision == "fp64" else torch.float32
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    if data_plus.shape != (nb, n) or data_cross.shape != (nb, n) or data_plus.dtype != ctype or \
            data_cross.dtype != ctype:
        raise ValueError(f"data must be ({nb}, {n}) {ctype}")
    if weights.shape != (n,) or weights.dtype != rdtype:
        raise ValueError(f"weights must be ({n},) {rdtype}")
    if xhm_model:
        grads, loss, _ = xhm.adjoint(params, freqs, data_plus, data_cross, weights, amp_scale, reference_mf,
                                     precision)
        return loss, grads
    rec, jac = records if records is not None else _records(params, freqs, amp_scale, reference_mf, precision,
                                                            approximant)
    loss = torch.empty((nb,), dtype=torch.float64, device=params.device)
    grads = torch.empty((nb, 8), dtype=torch.float64, device=params.device)
    if powers is None:
        powers = power_table(freqs, precision)
    call(entry, _ptr(rec), _ptr(jac), nb, _ptr(freqs), int(freqs.dtype == torch.float64), n,
         _code(precision), _ptr(data_plus), _ptr(data_cross), _ptr(weights), _ptr(powers), _ptr(loss), _ptr(grads),
         _stream(params.device))
    return loss, grads


#: the bytes of the static scratch behind :func:`loss_and_grad_static`
STATIC_SCRATCH_BYTES = 48 << 20


def loss_and_grad_static(params: torch.Tensor, freqs: torch.Tensor, data_plus: torch.Tensor,
                         data_cross: torch.Tensor, weights: torch.Tensor, amp_scale: float = 1e21,
                         reference_mf: Optional[float] = None, precision: str = "split",
                         out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """IMRPhenomXAS's :func:`loss_and_grad` through
    ``wfx_xas_loss_grad_static``: the library allocates nothing (its records,
    derivatives and power table live in a 48 MiB static device array per
    GPU, shared by every call on it), for callers that cannot allocate.
    ``out``: (loss (B,), grads (B, 8)) float64 to write into. Calls that
    overlap on one GPU race."""
    for name, t in (("params", params), ("freqs", freqs), ("data_plus", data_plus), ("data_cross", data_cross),
                    ("weights", weights)):
        _require_device(t, name)
    if precision not in ("fp64", "split", "split-df"):
        raise ValueError(f"the static loss runs fp64, split or split-df, not {precision!r}")
    nb, n = params.shape[0], freqs.shape[0]
    rdtype = torch.float64 if precision == "fp64" else torch.float32
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    if data_plus.shape != (nb, n) or data_cross.shape != (nb, n) or data_plus.dtype != ctype or \
            data_cross.dtype != ctype:
        raise ValueError(f"data must be ({nb}, {n}) {ctype}")
    if weights.shape != (n,) or weights.dtype != rdtype:
        raise ValueError(f"weights must be ({n},) {rdtype}")
    if out is None:
        loss = torch.empty((nb,), dtype=torch.float64, device=params.device)
        grads = torch.empty((nb, 8), dtype=torch.float64, device=params.device)
    else:
        loss, grads = out
    f_min = freqs[:1].to(torch.float64)
    call("wfx_xas_loss_grad_static", _ptr(params), nb, _ptr(f_min), 0, float(amp_scale),
         float("nan") if reference_mf is None else float(reference_mf), _ptr(freqs),
         int(freqs.dtype == torch.float64), n, _code(precision), _ptr(data_plus), _ptr(data_cross), _ptr(weights),
         _ptr(loss), _ptr(grads), _stream(params.device))
    return loss, grads
