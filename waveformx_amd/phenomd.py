"""IMRPhenomD on the GPU: device-tensor wrappers of its HIP kernels, which
share the evaluation kernels of IMRPhenomXAS (the same launch shapes, a
record of 64 slots per binary, the same split-precision arithmetic).

* :func:`setup` -- the per-binary records (the reference's ``core.table``,
  ``src/waveformx/models/phenomd/core.py:274-377``), always computed in FP64
  (``phenomd/model.py:33-50``) and written as FP64 records (``"fp64"``) or as
  FP32 records (``"split"``, ``"split-df"``);
* :func:`evaluate`, :func:`evaluate_ragged` -- ``(h+, hx)`` on one shared
  grid or on per-binary grids (``core.evaluate``, ``core.py:443-485``):
  every frequency is evaluated (LAL's sequence has no cutoff);
* :func:`polarizations` -- ``model.polarizations``.

There is no plain-FP32 arithmetic: the reference computes the table in FP64
whatever the binary's precision.
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch

from ._lib import call
from .xas import (RECORD_SLOTS, _code, _f_min, _ptr, _require_device, _stream, checked_total, run_setup_jvp,
                  tile_prefix)

__all__ = ["PRECISIONS", "LAL_FLAGS", "setup", "setup_jvp", "evaluate", "evaluate_ragged", "polarizations",
           "reference_phase"]

PRECISIONS = ("fp64", "split", "split-df")
#: the setup's flag bits -> the reference's message (phenomd/api.py:39-45)
LAL_FLAGS = ((8, "the final spin of IMRPhenomD exceeds one"),)


def _check_precision(precision: str) -> None:
    if precision not in PRECISIONS:
        raise ValueError(f"IMRPhenomD's arithmetic is one of {PRECISIONS} (its table is FP64), got {precision!r}")


def setup(params: torch.Tensor, f_min: torch.Tensor, amp_scale: float = 1e21, reference_mf: Optional[float] = None,
          precision: str = "split") -> Tuple[torch.Tensor, torch.Tensor]:
    """Records (B, 64) (float64 for ``"fp64"``, else float32) and the int32
    flags (B,) of binaries ``params`` (B, 8) at first frequency ``f_min``
    (one value or (B,))."""
    _check_precision(precision)
    _require_device(params, "params")
    nb = params.shape[0]
    if params.ndim != 2 or params.shape[1] != 8 or params.dtype != torch.float64:
        raise ValueError("params must be a (B, 8) float64 tensor")
    f_min, stride = _f_min(f_min, nb)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    records = torch.empty((nb, RECORD_SLOTS), dtype=dtype, device=params.device)
    flags = torch.empty((nb,), dtype=torch.int32, device=params.device)
    call("wfx_phd_setup", _ptr(params), nb, _ptr(f_min), stride, float(amp_scale),
         float("nan") if reference_mf is None else float(reference_mf), _code(precision), _ptr(records),
         _ptr(flags), _stream(params.device))
    return records, flags


def setup_jvp(params: torch.Tensor, f_min: torch.Tensor, amp_scale: float = 1e21,
              reference_mf: Optional[float] = None,
              precision: str = "fp64") -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Records (B, 64) of ``precision`` (as :func:`setup` writes them), the
    derivatives (B, 64, 8) of the FP64 record with respect to (m1, m2, chi1,
    chi2, distance, inclination, phi_ref, f_ref) (forward mode through the
    FP64 table; the region bounds carry none, as the reference's
    ``jnp.where`` selections), and the flags."""
    _check_precision(precision)
    return run_setup_jvp("wfx_phd_setup_jvp", params, f_min, amp_scale, reference_mf, precision)


def evaluate(records: torch.Tensor, freqs: torch.Tensor, precision: str = "split", group: int = 0,
             out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None,
             freqs_per_thread: int = 0) -> Tuple[torch.Tensor, torch.Tensor]:
    """(h+, hx) (B, N) on one shared grid ``freqs`` (N,) (float32 or
    float64); complex128 for ``"fp64"``, else complex64."""
    _check_precision(precision)
    _require_device(records, "records")
    _require_device(freqs, "freqs")
    if freqs.dtype not in (torch.float32, torch.float64):
        raise ValueError("freqs must be float32 or float64")
    nb, n = records.shape[0], freqs.shape[0]
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    if out is None:
        hp = torch.empty((nb, n), dtype=ctype, device=records.device)
        hc = torch.empty((nb, n), dtype=ctype, device=records.device)
    else:
        hp, hc = out
        if (hp.shape != (nb, n) or hc.shape != (nb, n) or hp.dtype != ctype or hc.dtype != ctype
                or not hp.is_contiguous() or not hc.is_contiguous()):
            raise ValueError(f"outputs must be contiguous ({nb}, {n}) {ctype}")
    call("wfx_phd_eval", _ptr(records), nb, _ptr(freqs), int(freqs.dtype == torch.float64), n, _code(precision),
         int(group), int(freqs_per_thread), _ptr(hp), _ptr(hc), _stream(records.device))
    return hp, hc


def evaluate_ragged(records: torch.Tensor, offsets: torch.Tensor, freqs: Optional[torch.Tensor] = None,
                    grid_f0: Optional[torch.Tensor] = None, grid_df: Optional[torch.Tensor] = None,
                    precision: str = "split", tiles: Optional[torch.Tensor] = None, blocks: int = 0,
                    total: Optional[int] = None, bins_per_thread: int = 1,
                    out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """(h+, hx) flat (total,) of binaries on their own grids, as
    :func:`waveformx_amd.xas.evaluate_ragged`."""
    _check_precision(precision)
    _require_device(records, "records")
    _require_device(offsets, "offsets")
    nb = records.shape[0]
    if offsets.dtype != torch.int64 or offsets.shape[0] != nb + 1:
        raise ValueError("offsets must be (B + 1,) int64")
    if freqs is None and (grid_f0 is None or grid_df is None):
        raise ValueError("give freqs or (grid_f0, grid_df)")
    if total is None:
        total = checked_total(offsets)
    if freqs is not None and freqs.shape[0] < total:
        raise ValueError("fewer frequencies than offsets[-1]")
    if tiles is None:
        tiles = tile_prefix(offsets, bins_per_thread)
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    if out is None:
        hp = torch.empty((total,), dtype=ctype, device=records.device)
        hc = torch.empty((total,), dtype=ctype, device=records.device)
    else:
        hp, hc = out
        if hp.shape[0] < total or hc.shape[0] < total or hp.dtype != ctype or hc.dtype != ctype:
            raise ValueError(f"outputs must hold {total} {ctype}")
    if grid_f0 is not None:
        grid_f0 = grid_f0.to(torch.float64).contiguous()
        grid_df = grid_df.to(torch.float64).contiguous()
    f64 = int(freqs is not None and freqs.dtype == torch.float64)
    call("wfx_phd_eval_ragged", _ptr(records), nb, _ptr(freqs), f64, _ptr(grid_f0), _ptr(grid_df), _ptr(offsets),
         _ptr(tiles), _code(precision), int(bins_per_thread), int(blocks), _ptr(hp), _ptr(hc),
         _stream(records.device))
    return hp, hc


def polarizations(params: torch.Tensor, freqs: torch.Tensor, amp_scale: float = 1e21,
                  reference_mf: Optional[float] = None, precision: str = "split",
                  group: int = 0) -> Tuple[torch.Tensor, torch.Tensor]:
    """``model.polarizations`` of a batch on a shared grid: the FP64 setup
    at the grid's first frequency, then the evaluation."""
    records, _ = setup(params, freqs[:1], amp_scale, reference_mf, precision)
    return evaluate(records, freqs, precision, group)


def reference_phase(params: torch.Tensor, mf: torch.Tensor) -> torch.Tensor:
    """IMRPhenomD's phase -Psi + t0 M f at ``M f = mf`` (B,) but for its
    constant, float64 (``phenomd/reference.py:18-22``), from params (B, 8)."""
    _require_device(params, "params")
    mf = mf.to(torch.float64).contiguous()
    out = torch.empty((params.shape[0],), dtype=torch.float64, device=params.device)
    call("wfx_phd_reference_phase", _ptr(params), params.shape[0], _ptr(mf), _ptr(out), _stream(params.device))
    return out
