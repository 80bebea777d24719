"""IMRPhenomXAS on the GPU: device-tensor wrappers of the HIP kernels.

* :func:`setup` -- the per-binary records (the reference's ``core.table``,
  ``src/waveformx/models/xas/core.py:257-379``) and LAL's error flags;
* :func:`evaluate` -- ``(h+, hx)`` on one shared grid (``core.evaluate``,
  ``core.py:496-507``, under ``jax.vmap``);
* :func:`evaluate_ragged` -- each binary on its own grid, in one
  persistent launch (no counterpart in the reference, whose batches share
  one grid);
* :func:`polarizations` -- the reference's ``model.polarizations``
  (``xas/model.py:48-60``): setup at the grid's first frequency, then the
  evaluation.

Arithmetics (``precision``): ``"fp64"`` (FP64 throughout, complex128),
``"split"`` (FP64 setup, FP32 evaluation with the inspiral's leading term
and the phase reduction in FP64, complex64), ``"split-df"`` (the same with
the leading term in double-float FP32 arithmetic) and ``"fp32"`` (FP32
setup and evaluation, no split: for comparison only).
"""

from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch

from ._lib import call, lib

__all__ = ["PRECISIONS", "RECORD_SLOTS", "setup", "setup_jvp", "evaluate", "evaluate_ragged", "tile_prefix",
           "power_table", "polarizations", "pack_params", "LAL_FLAGS"]

PRECISIONS = {"fp64": 0, "fp32": 1, "split": 2, "split-df": 3}
RECORD_SLOTS = 64
#: LAL's error flags of the setup (bit -> the reference's message,
#: ``xas/api.py:113-120``), in the order the reference checks them
LAL_FLAGS = ((1, "IMRPhenomXAS is not valid at mass ratios beyond 1000"),
             (4, "the fitted final mass is not physical"),
             (8, "the fitted final spin exceeds one"),
             (16, "f_MECO cannot be greater than f_ISCO"))


def _ptr(t: Optional[torch.Tensor]):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(device: torch.device):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _code(precision: str) -> int:
    try:
        return PRECISIONS[precision]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}, got {precision!r}") from None


def _require_device(t: torch.Tensor, name: str) -> None:
    if t.device.type != "cuda":
        raise RuntimeError(f"{name} must be on a HIP device (waveformx_amd has no CPU path)")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def pack_params(mass_1, mass_2, chi_1, chi_2, distance, inclination, phi_ref, f_ref) -> torch.Tensor:
    """The (B, 8) float64 parameter array the setup reads."""
    return torch.stack([torch.as_tensor(x).to(torch.float64) for x in
                        (mass_1, mass_2, chi_1, chi_2, distance, inclination, phi_ref, f_ref)], dim=-1).contiguous()


def _f_min(f_min: torch.Tensor, nb: int) -> Tuple[torch.Tensor, int]:
    f_min = f_min.to(torch.float64).reshape(-1).contiguous()
    if f_min.numel() == 1:
        return f_min, 0
    if f_min.numel() != nb:
        raise ValueError("f_min must hold one value or one per binary")
    return f_min, 1


def setup(params: torch.Tensor, f_min: torch.Tensor, amp_scale: float = 1e21, reference_mf: Optional[float] = None,
          precision: str = "split") -> Tuple[torch.Tensor, torch.Tensor]:
    """Records (B, 64) (float64 for ``"fp64"``, else float32) and the int32
    LAL error flags (B,) of binaries ``params`` (B, 8) at first frequency
    ``f_min`` (one value or (B,))."""
    _require_device(params, "params")
    nb = params.shape[0]
    if params.ndim != 2 or params.shape[1] != 8 or params.dtype != torch.float64:
        raise ValueError("params must be a (B, 8) float64 tensor")
    f_min, stride = _f_min(f_min, nb)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    records = torch.empty((nb, RECORD_SLOTS), dtype=dtype, device=params.device)
    flags = torch.empty((nb,), dtype=torch.int32, device=params.device)
    call("wfx_xas_setup", _ptr(params), nb, _ptr(f_min), stride, float(amp_scale),
         float("nan") if reference_mf is None else float(reference_mf), _code(precision), _ptr(records),
         _ptr(flags), _stream(params.device))
    return records, flags


def setup_jvp(params: torch.Tensor, f_min: torch.Tensor, amp_scale: float = 1e21,
              reference_mf: Optional[float] = None,
              precision: str = "fp64") -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Records (B, 64) of ``precision`` ("fp64": float64; "split" or
    "split-df": the float32 records :func:`setup` writes for them), the
    derivatives (B, 64, 8) of the FP64 record with respect to (m1, m2, chi1,
    chi2, distance, inclination, phi_ref, f_ref), and flags."""
    return run_setup_jvp("wfx_xas_setup_jvp", params, f_min, amp_scale, reference_mf, precision)


def run_setup_jvp(entry: str, params: torch.Tensor, f_min: torch.Tensor, amp_scale: float,
                  reference_mf: Optional[float], precision: str) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """:func:`setup_jvp` through the C entry point ``entry`` (a model's
    derivative setup: they share one signature)."""
    _require_device(params, "params")
    if precision not in ("fp64", "split", "split-df"):
        raise ValueError(f"setup_jvp writes fp64 or split records, not {precision!r}")
    nb = params.shape[0]
    f_min, stride = _f_min(f_min, nb)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    records = torch.empty((nb, RECORD_SLOTS), dtype=dtype, device=params.device)
    jac = torch.empty((nb, RECORD_SLOTS, 8), dtype=torch.float64, device=params.device)
    flags = torch.empty((nb,), dtype=torch.int32, device=params.device)
    wide, split = (_ptr(records), None) if precision == "fp64" else (None, _ptr(records))
    call(entry, _ptr(params), nb, _ptr(f_min), stride, float(amp_scale),
         float("nan") if reference_mf is None else float(reference_mf), wide, split, _ptr(jac), _ptr(flags),
         _stream(params.device))
    return records, jac, flags


def evaluate(records: torch.Tensor, freqs: torch.Tensor, precision: str = "split", group: int = 0,
             out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None,
             freqs_per_thread: int = 0) -> Tuple[torch.Tensor, torch.Tensor]:
    """(h+, hx) (B, N) on one shared grid ``freqs`` (N,) (float32 or
    float64); complex128 for ``"fp64"``, else complex64. ``group``
    (binaries per block) and ``freqs_per_thread`` (1, 2, 4, 8) shape the
    launch (0: defaults); every shape writes the same bits."""
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
    if out is not None and (hp.shape != (nb, n) or hc.shape != (nb, n) or hp.dtype != ctype or hc.dtype != ctype
                            or not hp.is_contiguous() or not hc.is_contiguous()):
        raise ValueError(f"outputs must be contiguous ({nb}, {n}) {ctype}")
    call("wfx_xas_eval", _ptr(records), nb, _ptr(freqs), int(freqs.dtype == torch.float64), n, _code(precision),
         int(group), int(freqs_per_thread), _ptr(hp), _ptr(hc), _stream(records.device))
    return hp, hc


def checked_total(offsets: torch.Tensor) -> int:
    """``offsets[-1]``, read on the host after checking that the offsets
    start at 0 and never decrease (a decreasing pair would hand a binary
    bins past the end of the outputs)."""
    host = offsets.cpu()
    if int(host[0]) != 0 or bool((host[1:] < host[:-1]).any()):
        raise ValueError("offsets must start at 0 and never decrease")
    return int(host[-1])


def tile_prefix(offsets: torch.Tensor, bins_per_thread: int = 1) -> torch.Tensor:
    """The prefix sum (B+1,) of the (256 x bins_per_thread)-bin tiles per
    binary, for :func:`evaluate_ragged`."""
    _require_device(offsets, "offsets")
    nb = offsets.shape[0] - 1
    tiles = torch.empty((nb + 1,), dtype=torch.int64, device=offsets.device)
    call("wfx_xas_tile_prefix", _ptr(offsets), nb, int(bins_per_thread), _ptr(tiles), _stream(offsets.device))
    return tiles


def evaluate_ragged(records: torch.Tensor, offsets: torch.Tensor, freqs: Optional[torch.Tensor] = None,
                    grid_f0: Optional[torch.Tensor] = None, grid_df: Optional[torch.Tensor] = None,
                    precision: str = "split", tiles: Optional[torch.Tensor] = None, blocks: int = 0,
                    total: Optional[int] = None, bins_per_thread: int = 1,
                    out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """(h+, hx) flat (total,) of binaries on their own grids: binary b owns
    [offsets[b], offsets[b+1]) of the concatenated ``freqs``, or the uniform
    grid ``grid_f0[b] + k grid_df[b]`` of that length. ``tiles`` (from
    :func:`tile_prefix` with the same ``bins_per_thread``) and ``total``
    given, nothing synchronises with the host (graph capture)."""
    _require_device(records, "records")
    _require_device(offsets, "offsets")
    if offsets.dtype != torch.int64:
        raise ValueError("offsets must be int64")
    nb = records.shape[0]
    if offsets.shape[0] != nb + 1:
        raise ValueError("offsets must have B + 1 entries")
    if freqs is None and (grid_f0 is None or grid_df is None):
        raise ValueError("give freqs or (grid_f0, grid_df)")
    if total is None:
        total = checked_total(offsets)
        if freqs is not None and freqs.shape[0] != total:
            raise ValueError("offsets[-1] must equal the number of frequencies")
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
    call("wfx_xas_eval_ragged", _ptr(records), nb, _ptr(freqs), f64, _ptr(grid_f0), _ptr(grid_df), _ptr(offsets),
         _ptr(tiles), _code(precision), int(bins_per_thread), int(blocks), _ptr(hp), _ptr(hc),
         _stream(records.device))
    return hp, hc


def power_table(freqs: torch.Tensor, precision: str) -> torch.Tensor:
    """The grid's table of frequency powers (bytes, on the device) that the
    adjoint kernels, which walk every frequency once per binary, read
    instead of computing the powers per bin (the same values)."""
    _require_device(freqs, "freqs")
    row = lib().wfx_xas_power_table_bytes(_code(precision))
    table = torch.empty((freqs.shape[0], row), dtype=torch.uint8, device=freqs.device)
    call("wfx_xas_power_table", _ptr(freqs), int(freqs.dtype == torch.float64), freqs.shape[0], _code(precision),
         _ptr(table), _stream(freqs.device))
    return table


def polarizations(params: torch.Tensor, freqs: torch.Tensor, amp_scale: float = 1e21,
                  reference_mf: Optional[float] = None, precision: str = "split",
                  group: int = 0) -> Tuple[torch.Tensor, torch.Tensor]:
    """``model.polarizations`` of a batch on a shared grid: the setup at the
    grid's first frequency (in the working precision, as the reference),
    then the evaluation."""
    records, _ = setup(params, freqs[:1], amp_scale, reference_mf, precision)
    return evaluate(records, freqs, precision, group)


def _check_library() -> None:
    lib()
