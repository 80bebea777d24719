"""IMRPhenomXHM (release 122022) on the GPU: device-tensor wrappers of its
HIP kernels.

* :func:`setup` -- per binary, the FP64 record of its (2,2) mode (that of
  IMRPhenomXAS), a head (the factors of the five modes in the
  polarizations, ``xhm/reference.py:646-653``) and one record per higher
  mode, (2,1), (3,3), (3,2), (4,4) (``xhm/reference.py:580-608``: LAL's
  set-up with exact derivatives where LAL differentiates numerically); one
  FP64 thread per binary and mode;
* :func:`evaluate` -- ``(h+, hx)`` (B, N) complex128 on a shared grid, and
  optionally the five modes ``h_{l,-m}`` (5, B, N): each higher mode by
  region in Mf, the (3,2) ringdown recycling the (2,2) mode
  (``xhm/reference.py:620-677``);
* :func:`polarizations` / :func:`modes` -- set-up at the grid's first
  frequency, then the evaluation.

Two arithmetics, both with an FP64 set-up: ``"fp64"`` (complex128 out; the
reference's float64 path) and ``"split"`` (FP32 with each mode's leading
inspiral phase term and constant in FP64 and the later regions' constants
reduced modulo 2 pi: the reference's float32 design, ``xhm/core.py:13-21``;
complex64 out, float32 grid).
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

from ._lib import call, lib
from .xas import RECORD_SLOTS, _f_min, _ptr, _require_device, _stream

__all__ = ["MODES", "LAL_FLAGS", "PRECISIONS", "setup", "setup_jvp", "evaluate", "evaluate_ragged", "polarizations",
           "modes", "adjoint", "vjp", "vjp_differences"]

#: the modes (l, m) of the outputs, the (2,2) first (``xhm/reference.py:55-57``)
MODES = ((2, 2), (2, 1), (3, 3), (3, 2), (4, 4))
PRECISIONS = ("fp64", "split")
_CODES = {"fp64": 0, "split": 2}
#: the setup's flag bits -> LAL's messages: those of the waveform variables
#: (``xas/api.py:40-60``, checked for IMRPhenomXHM by ``xhm/api.py:67-72``)
LAL_FLAGS = ((1, "IMRPhenomXHM is not valid at mass ratios beyond 1000"),
             (4, "the fitted final mass is not physical"),
             (8, "the fitted final spin exceeds one"),
             (16, "f_MECO cannot be greater than f_ISCO"))


def setup(params: torch.Tensor, f_min: torch.Tensor, amp_scale: float = 1e21, reference_mf: Optional[float] = None,
          precision: str = "fp64") -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor, torch.Tensor]:
    """``(records22 (B, 64), head (B, H), modes (B, 4, S), flags (B,))``,
    float64 for ``"fp64"``, float32 for ``"split"`` (int32 flags), for
    binaries ``params`` (B, 8) at first frequency ``f_min`` (one value or
    (B,)); with ``reference_mf`` the binaries are given at ``M f_ref =
    reference_mf`` (their f_ref zero, NaN rows otherwise:
    ``xhm/model.py:40-42``)."""
    if precision not in PRECISIONS:
        raise ValueError(f"IMRPhenomXHM's arithmetic is one of {PRECISIONS}, got {precision!r}")
    _require_device(params, "params")
    if params.ndim != 2 or params.shape[1] != 8 or params.dtype != torch.float64:
        raise ValueError("params must be a (B, 8) float64 tensor")
    nb = params.shape[0]
    params = params.contiguous()
    f_min, stride = _f_min(f_min, nb)
    handle = lib()
    dev = params.device
    dtype = torch.float64 if precision == "fp64" else torch.float32
    records = torch.empty((nb, RECORD_SLOTS), dtype=dtype, device=dev)
    head = torch.empty((nb, handle.wfx_xhm_head_slots()), dtype=dtype, device=dev)
    modes_ = torch.empty((nb, 4, handle.wfx_xhm_record_slots()), dtype=dtype, device=dev)
    flags = torch.empty((nb,), dtype=torch.int32, device=dev)
    call("wfx_xhm_setup", _ptr(params), nb, _ptr(f_min), stride, float(amp_scale),
         float("nan") if reference_mf is None else float(reference_mf), _CODES[precision], _ptr(records),
         _ptr(head), _ptr(modes_), _ptr(flags), _stream(dev))
    return records, head, modes_, flags


def setup_jvp(params: torch.Tensor, f_min: torch.Tensor, amp_scale: float = 1e21,
              reference_mf: Optional[float] = None, precision: str = "fp64") -> dict:
    """The gradient's set-up with its derivatives, in dual numbers through
    the same FP64 code as :func:`setup`: ``records22``, ``head``, ``flags``
    as :func:`setup`'s; ``modes`` the higher modes' records in the split
    evaluation's layout (float32: :func:`setup`'s ``"split"`` records;
    float64 for ``"fp64"``: the same slots in FP64, not the FP64
    evaluation's LAL form); ``jac22`` (B, 64, 8) float64, the (2,2) record's
    derivatives in the 8 parameters (IMRPhenomXAS's derivative setup: the
    same record); ``dhead`` (B, H), the head's derivative in the
    inclination; ``dmodes`` (B, 4, S, 4), the higher modes' records'
    derivatives in (m1, m2, chi1z, chi2z) (the last axis). One call
    (``wfx_xhm_setup_jvp``): the (2,2)'s derivative set-up hands its
    coefficients to the higher modes' through a workspace."""
    if precision not in PRECISIONS:
        raise ValueError(f"IMRPhenomXHM's arithmetic is one of {PRECISIONS}, got {precision!r}")
    _require_device(params, "params")
    if params.ndim != 2 or params.shape[1] != 8 or params.dtype != torch.float64:
        raise ValueError("params must be a (B, 8) float64 tensor")
    params = params.contiguous()
    nb = params.shape[0]
    f_min_t, stride = _f_min(f_min, nb)
    handle = lib()
    dev = params.device
    dtype = torch.float64 if precision == "fp64" else torch.float32
    slots = handle.wfx_xhm_record_slots()
    records22 = torch.empty((nb, RECORD_SLOTS), dtype=dtype, device=dev)
    jac22 = torch.empty((nb, RECORD_SLOTS, 8), dtype=torch.float64, device=dev)
    flags = torch.empty((nb,), dtype=torch.int32, device=dev)
    head = torch.empty((nb, handle.wfx_xhm_head_slots()), dtype=dtype, device=dev)
    dhead = torch.empty_like(head)
    modes_ = torch.empty((nb, 4, slots), dtype=dtype, device=dev)
    dmodes = torch.empty((nb, 4, slots, 4), dtype=dtype, device=dev)
    workspace = torch.empty((max(handle.wfx_xhm_setup_jvp_workspace(nb), 0),), dtype=torch.uint8, device=dev)
    call("wfx_xhm_setup_jvp", _ptr(params), nb, _ptr(f_min_t), stride, float(amp_scale),
         float("nan") if reference_mf is None else float(reference_mf), _CODES[precision], _ptr(records22),
         _ptr(jac22), _ptr(flags), _ptr(head), _ptr(dhead), _ptr(modes_), _ptr(dmodes), _ptr(workspace),
         workspace.numel(), _stream(dev))
    return {"records22": records22, "jac22": jac22, "head": head, "dhead": dhead, "modes": modes_,
            "dmodes": dmodes, "flags": flags}


def evaluate(records: torch.Tensor, head: torch.Tensor, modes_: torch.Tensor, freqs: torch.Tensor,
             with_modes: bool = False, out: Optional[Tuple[torch.Tensor, torch.Tensor]] = None):
    """``(h+, hx)`` (B, N) on the shared grid ``freqs`` (N,): complex128 from
    FP64 records (float32 or float64 grid), complex64 from split records
    (float32 grid); with ``with_modes`` also the modes (5, B, N); ``out``
    two contiguous (B, N) tensors of that type to write into."""
    for t, name in ((records, "records"), (head, "head"), (modes_, "modes"), (freqs, "freqs")):
        _require_device(t, name)
    if freqs.dtype not in (torch.float32, torch.float64):
        raise ValueError("freqs must be float32 or float64")
    split = records.dtype == torch.float32
    if split and freqs.dtype != torch.float32:
        raise ValueError("the split evaluation takes a float32 grid")
    if head.dtype != records.dtype or modes_.dtype != records.dtype:
        raise ValueError("records, head and modes must come from one setup")
    nb, n = records.shape[0], freqs.shape[0]
    if head.shape[0] != nb or modes_.shape[0] != nb:
        raise ValueError("records, head and modes must describe the same binaries")
    freqs = freqs.contiguous()
    dev = records.device
    ctype = torch.complex64 if split else torch.complex128
    if out is None:
        hp = torch.empty((nb, n), dtype=ctype, device=dev)
        hc = torch.empty((nb, n), dtype=ctype, device=dev)
    else:
        hp, hc = out
        for t in out:
            if t.shape != (nb, n) or t.dtype != ctype or not t.is_contiguous() or t.device != dev:
                raise ValueError(f"out must be two contiguous (B, N) {ctype} tensors on the records' device")
    mode_out = torch.empty((5, nb, n), dtype=ctype, device=dev) if with_modes else None
    call("wfx_xhm_eval", _ptr(records), _ptr(head), _ptr(modes_), nb, _ptr(freqs),
         int(freqs.dtype == torch.float64), n, 2 if split else 0, _ptr(hp), _ptr(hc), _ptr(mode_out), _stream(dev))
    return (hp, hc, mode_out) if with_modes else (hp, hc)


def evaluate_ragged(records: torch.Tensor, head: torch.Tensor, modes_: torch.Tensor, offsets: torch.Tensor,
                    freqs: Optional[torch.Tensor] = None, grid_f0: Optional[torch.Tensor] = None,
                    grid_df: Optional[torch.Tensor] = None, with_modes: bool = False, tiles: Optional[torch.Tensor] = None,
                    total: Optional[int] = None, blocks: int = 0):
    """``(h+, hx)`` flat (total,) of binaries on their own grids: binary b
    owns ``[offsets[b], offsets[b+1])`` of the concatenated ``freqs``
    (float32 for split records; float32 or float64 for FP64 records), or the
    uniform grid ``grid_f0[b] + k grid_df[b]`` of that length; with
    ``with_modes`` also the modes (5, total). One persistent launch over
    256-bin tiles; each bin has the bits :func:`evaluate` gives it on a
    shared grid. ``tiles`` (:func:`waveformx_amd.xas.tile_prefix` of the
    offsets) and ``total`` given, nothing synchronises with the host."""
    from .xas import checked_total, tile_prefix

    for t, name in ((records, "records"), (head, "head"), (modes_, "modes"), (offsets, "offsets")):
        _require_device(t, name)
    if offsets.dtype != torch.int64:
        raise ValueError("offsets must be int64")
    nb = records.shape[0]
    if offsets.shape[0] != nb + 1 or head.shape[0] != nb or modes_.shape[0] != nb:
        raise ValueError("records, head, modes and offsets (B + 1) must describe the same binaries")
    if head.dtype != records.dtype or modes_.dtype != records.dtype:
        raise ValueError("records, head and modes must come from one setup")
    split = records.dtype == torch.float32
    if freqs is None and (grid_f0 is None or grid_df is None):
        raise ValueError("give freqs or (grid_f0, grid_df)")
    if freqs is not None:
        _require_device(freqs, "freqs")
        if freqs.dtype not in (torch.float32, torch.float64) or (split and freqs.dtype != torch.float32):
            raise ValueError("freqs must be float32 (split) or float32/float64 (FP64)")
        freqs = freqs.contiguous()
    if total is None:
        total = checked_total(offsets)
        if freqs is not None and freqs.shape[0] != total:
            raise ValueError("offsets[-1] must equal the number of frequencies")
    if freqs is not None and freqs.shape[0] < total:
        raise ValueError("fewer frequencies than offsets[-1]")
    if grid_f0 is not None:
        grid_f0 = grid_f0.to(torch.float64).contiguous()
        grid_df = grid_df.to(torch.float64).contiguous()
    if tiles is None:
        tiles = tile_prefix(offsets, 1)
    dev = records.device
    ctype = torch.complex64 if split else torch.complex128
    hp = torch.empty((total,), dtype=ctype, device=dev)
    hc = torch.empty((total,), dtype=ctype, device=dev)
    mode_out = torch.empty((5, total), dtype=ctype, device=dev) if with_modes else None
    if nb and total:
        call("wfx_xhm_eval_ragged", _ptr(records), _ptr(head), _ptr(modes_), nb, _ptr(freqs),
             int(freqs is not None and freqs.dtype == torch.float64), _ptr(grid_f0), _ptr(grid_df), _ptr(offsets),
             _ptr(tiles), 2 if split else 0, int(blocks), _ptr(hp), _ptr(hc), _ptr(mode_out), _stream(dev))
    return (hp, hc, mode_out) if with_modes else (hp, hc)


def _per_binary(params: torch.Tensor, freqs: torch.Tensor, amp_scale, reference_mf, precision, with_modes):
    """Per-binary grids (B, N), as the reference's (B, N) frequencies
    (``xhm/mlx/_api.py:79-93``): each binary set up at its own first
    frequency, the grid rows evaluated as one ragged batch."""
    nb, n = freqs.shape
    if nb != params.shape[0]:
        raise ValueError(f"per-binary grids must be ({params.shape[0]}, N), got {tuple(freqs.shape)}")
    freqs = freqs.contiguous()
    records, head, modes_, _ = setup(params, freqs[:, 0].to(torch.float64), amp_scale, reference_mf, precision)
    offsets = torch.arange(nb + 1, dtype=torch.int64, device=params.device) * n
    out = evaluate_ragged(records, head, modes_, offsets, freqs=freqs.reshape(-1), with_modes=with_modes,
                          total=nb * n)
    if with_modes:
        return out[2].reshape(5, nb, n)
    return out[0].reshape(nb, n), out[1].reshape(nb, n)


def project(records: torch.Tensor, head: torch.Tensor, modes_: torch.Tensor, freqs: torch.Tensor,
            cot_plus: torch.Tensor, cot_cross: torch.Tensor) -> torch.Tensor:
    """``Re sum_f (conj(cot+) h+ + conj(cotx) hx)`` (B,) float64 of the FP64
    polarizations of ``records, head, modes`` (B binaries, FP64 set-up) on
    the float64 grid ``freqs`` (N,), computed without writing them (one
    fused launch); binary b takes row ``b % R`` of the complex128 cotangents
    (R, N). ``head`` (H, B, S), H = 1..4 heads (the modes' factors in the
    polarizations) over the same records and modes, gives (H, B) in the
    same launch, the modes computed once."""
    for t, name in ((records, "records"), (head, "head"), (modes_, "modes"), (freqs, "freqs"),
                    (cot_plus, "cot_plus"), (cot_cross, "cot_cross")):
        _require_device(t, name)
    if records.dtype != torch.float64 or head.dtype != torch.float64 or modes_.dtype != torch.float64:
        raise ValueError("project takes the FP64 set-up")
    if freqs.dtype != torch.float64:
        raise ValueError("project takes a float64 grid")
    stacked = head.dim() == 3
    heads = head if stacked else head[None]
    n_heads = heads.shape[0]
    if not 1 <= n_heads <= 4:
        raise ValueError("project takes 1 to 4 heads")
    nb, n = records.shape[0], freqs.shape[0]
    if heads.shape[1] != nb or modes_.shape[0] != nb:
        raise ValueError("records, head and modes must describe the same binaries")
    heads = heads.contiguous()
    cot_plus = cot_plus.to(torch.complex128).contiguous()
    cot_cross = cot_cross.to(torch.complex128).contiguous()
    rows = cot_plus.shape[0]
    if cot_plus.shape != (rows, n) or cot_cross.shape != (rows, n) or (rows == 0 and nb > 0):
        raise ValueError("the cotangents must be two (R, N) arrays on the grid, R > 0")
    chunks = (n + 255) // 256
    partial = torch.zeros((n_heads, nb, max(chunks, 1)), dtype=torch.float64, device=records.device)
    call("wfx_xhm_project", _ptr(records), _ptr(heads), n_heads, _ptr(modes_), nb, _ptr(freqs.contiguous()), n,
         _ptr(cot_plus), _ptr(cot_cross), rows, _ptr(partial), _stream(records.device))
    out = partial.sum(2)
    return out if stacked else out[0]


def polarizations(params: torch.Tensor, freqs: torch.Tensor, amp_scale: float = 1e21,
                  reference_mf: Optional[float] = None, precision: str = "fp64") -> Tuple[torch.Tensor, torch.Tensor]:
    """``(h+, hx)`` of a batch on a shared grid (N,) or on per-binary grids
    (B, N): the set-up at each grid's first frequency, then the evaluation
    (``"split"`` on the float32 grid)."""
    if precision == "split":
        freqs = freqs.to(torch.float32)
    if freqs.ndim == 2:
        return _per_binary(params, freqs, amp_scale, reference_mf, precision, False)
    records, head, modes_, _ = setup(params, freqs[:1], amp_scale, reference_mf, precision)
    return evaluate(records, head, modes_, freqs)


def modes(params: torch.Tensor, freqs: torch.Tensor, amp_scale: float = 1e21,
          reference_mf: Optional[float] = None, precision: str = "fp64") -> torch.Tensor:
    """The five modes ``h_{l,-m}`` of :data:`MODES`, (5, B, N), on a shared
    grid (N,) or per-binary grids (B, N)."""
    if precision == "split":
        freqs = freqs.to(torch.float32)
    if freqs.ndim == 2:
        return _per_binary(params, freqs, amp_scale, reference_mf, precision, True)
    records, head, modes_, _ = setup(params, freqs[:1], amp_scale, reference_mf, precision)
    return evaluate(records, head, modes_, freqs, with_modes=True)[2]


# -------------------------------------------------------------- gradient

#: the record slots that select a region or a branch (xhm_setup.h HmSlot:
#: kHmTag, kHmOn, kHmFalloff, kHmMixing, kHmFAmpIn, kHmFAmpIm, kHmFPhaseIn,
#: kHmFPhaseIm, kHmFRdAux, kHmNInter; xas_layout.h: kFPhaseIn, kFPhaseIm,
#: kFAmpIn, kFAmpRd, kFCut): held at their values at the parameters, as the
#: reference's jnp.where selections carry no derivative
MODE_SELECTION_SLOTS = (0, 1, 6, 7, 8, 9, 10, 11, 44, 55)
#: the IMRPhenomXAS record's inspiral phase constant and anchor (csrc/xas_layout.h)
_XAS_INS_CONST, _XAS_ANCHOR = 0, 58
XAS_SELECTION_SLOTS = (49, 50, 51, 52, 53)
#: the fourth-order central-difference stencil (offset, weight)
_STENCIL = ((-2, 1.0 / 12.0), (-1, -8.0 / 12.0), (1, 8.0 / 12.0), (2, -1.0 / 12.0))


#: the fourth-order one-sided stencils (offsets 1..4 and the base point's
#: weight), for a parameter whose central stencil would cross a boundary
_FORWARD = ((1, 48.0 / 12.0), (2, -36.0 / 12.0), (3, 16.0 / 12.0), (4, -3.0 / 12.0))
_FORWARD_BASE = -25.0 / 12.0
#: the mass ratio beyond which LAL's checks fail the binary (xas/api.py:40-60)
_MAX_MASS_RATIO = 1000.0


def _steps(params: torch.Tensor) -> list:
    """Per-parameter steps (B,) of the differences, or None for a parameter
    taken in closed form (d_L, the inclination, phi_ref and f_ref)."""
    m = params[:, :2].abs()
    ones = torch.ones_like(params[:, 0])
    return [3e-7 * m[:, 0], 3e-7 * m[:, 1], 1e-6 * ones, 1e-6 * ones, None, None, None, None]


def _directions(params: torch.Tensor, steps: list) -> dict:
    """Per parameter j with a step, the stencil's direction per binary (B,):
    0 central, +1 forward, -1 backward. A one-sided stencil keeps every point
    on the parameters' side of a boundary the central one would cross: the
    set-up's swap of the labels where m1 < m2, LAL's spin bound |chi| <= 1
    and its mass-ratio bound q <= 1000 (each a branch or a NaN there, not
    one smooth formula)."""
    m1, m2, c1, c2 = params[:, 0], params[:, 1], params[:, 2], params[:, 3]
    ordered = m1 >= m2
    reach = [None if h is None else 4.0 * h for h in steps]
    zero = torch.zeros_like(m1)
    out = {}
    # m1: the swap at m1 = m2; q = m1 / m2 grows with m1 when ordered
    crosses = torch.where(ordered, m1 - reach[0] < m2, m1 + reach[0] >= m2)
    d = torch.where(crosses, torch.where(ordered, zero + 1, zero - 1), zero)
    q_bound = ordered & ((m1 + reach[0]) / m2 > _MAX_MASS_RATIO)
    d = torch.where(q_bound, zero - 1, d)
    q_bound = ~ordered & (m2 / (m1 - reach[0]) > _MAX_MASS_RATIO)
    out[0] = torch.where(q_bound, zero + 1, d)
    # m2 likewise
    crosses = torch.where(ordered, m2 + reach[1] > m1, m2 - reach[1] <= m1)
    d = torch.where(crosses, torch.where(ordered, zero - 1, zero + 1), zero)
    q_bound = ordered & (m1 / (m2 - reach[1]) > _MAX_MASS_RATIO)
    d = torch.where(q_bound, zero + 1, d)
    q_bound = ~ordered & ((m2 + reach[1]) / m1 > _MAX_MASS_RATIO)
    out[1] = torch.where(q_bound, zero - 1, d)
    # the spins: |chi| <= 1
    for j, chi in ((2, c1), (3, c2)):
        d = torch.where(chi + reach[j] > 1.0, zero - 1, zero)
        out[j] = torch.where(chi - reach[j] < -1.0, zero + 1, d)
    return out


def _factor_derivatives(inclination) -> Tuple:
    """d(F+, Fx)_lm / d(inclination) (B, 5) complex128 of the five modes, by
    fourth-order differences of the harmonics (xhm/reference.py:646-653),
    vectorised over the binaries."""
    from ._modes import spin_weighted_spherical_harmonic as swsh

    def factors(theta):
        out_p, out_c = [], []
        for ell, emm in MODES:
            sign = -1.0 if ell % 2 else 1.0
            ym = swsh(theta, np.pi / 2.0, ell, -emm)
            ystar = np.conj(swsh(theta, np.pi / 2.0, ell, emm))
            out_p.append(0.5 * (ym + sign * ystar))
            out_c.append(0.5j * (ym - sign * ystar))
        return np.stack(out_p, axis=-1), np.stack(out_c, axis=-1)

    theta = np.asarray(inclination, dtype=float)
    step = 1e-3
    dp = np.zeros((theta.size, 5), np.complex128)
    dc = np.zeros((theta.size, 5), np.complex128)
    for offset, weight in _STENCIL:
        fp, fc = factors(theta + offset * step)
        dp += weight * fp / step
        dc += weight * fc / step
    return dp, dc


def vjp_differences(params: torch.Tensor, freqs: torch.Tensor, cot_plus: torch.Tensor, cot_cross: torch.Tensor,
                    amp_scale: float = 1e21, reference_mf: Optional[float] = None) -> torch.Tensor:
    """The FP64 gradient by differences, an independent check of
    :func:`vjp`'s exact one (another arithmetic: the FP64 evaluation's LAL
    form, differenced): dL/d(m1, m2, chi1z, chi2z, d_L, inclination, phi_ref,
    f_ref) (B, 8) float64 of the FP64 polarizations, for the cotangents
    dL/dRe h + i dL/dIm h (B, N) of h+ and hx: Re sum_f conj(cot) dh/dtheta.
    The masses' and
    spins' derivatives are fourth-order central differences of the FP64
    set-up and evaluation with every region selection (bounds, branches)
    held at the parameters, the reference's jnp.where semantics, so that
    each difference is of one smooth formula; relative accuracy ~1e-8.
    Where the central stencil would cross the set-up's swap of the labels
    (m1 = m2), the spin bound |chi| = 1 or the mass-ratio bound q = 1000,
    the stencil is one-sided (fourth order, on the parameters' side).
    d_L's, phi_ref's, the inclination's and f_ref's are closed forms (the
    projection with the modes' factors in the polarizations replaced: f_ref
    turns each mode h_{l,-m} by m/2 times the (2,2)'s constant phase); f_ref's
    is zero where f_ref is zero or with reference_mf. Every inner product is
    taken in the evaluation's launch (:func:`project`): no waveform is
    written."""
    _require_device(params, "params")
    params = params.to(torch.float64).contiguous()
    freqs = freqs.to(torch.float64).contiguous()
    cot_plus = cot_plus.to(torch.complex128)
    cot_cross = cot_cross.to(torch.complex128)
    nb = params.shape[0]
    grads = torch.zeros((nb, 8), dtype=torch.float64, device=params.device)
    if nb == 0 or freqs.shape[0] == 0:
        return grads
    rec0, head0, modes0, _ = setup(params, freqs[:1], amp_scale, reference_mf)
    dev = params.device

    # closed forms, each the projection of the set-up at the parameters with other factors (F+, Fx)_lm in
    # the head: h ~ 1/d_L (the factors as they are); h_{l,-m} ~ e^{i m phi_ref} (each mode's phase holds
    # m phi_ref: the (2,2)'s 2 phi_ref and its alignment m/2 phi_22), so i m (F+, Fx)_lm; the inclination
    # enters the factors only (their derivatives differenced on the host: smooth functions of order one),
    # NaN where LAL fails as the factors are
    factors = head0[:, :20].reshape(nb, 5, 4)
    fp = torch.complex(factors[..., 0], factors[..., 1])
    fc = torch.complex(factors[..., 2], factors[..., 3])
    emm = torch.tensor([float(m) for _, m in MODES], dtype=torch.float64, device=dev)
    dfp, dfc = (torch.as_tensor(a, device=dev) for a in _factor_derivatives(params[:, 5].cpu().numpy()))
    failed = torch.isnan(fp)
    dfp = torch.where(failed, fp, dfp)
    dfc = torch.where(failed, fc, dfc)

    def with_factors(f_plus, f_cross):
        head = head0.clone()
        head[:, :20] = torch.stack([f_plus.real, f_plus.imag, f_cross.real, f_cross.imag], -1).reshape(nb, 20)
        return head

    # f_ref turns every mode by a constant: h_{l,-m} by m/2 times the (2,2)'s phase constant, whose derivative
    # is that of the (2,2) record's phase constants less its anchor's (the XAS set-up's Jacobian); none where
    # f_ref is zero (it takes f_min) or with the mass-scaled reference
    if reference_mf is None:
        from . import xas

        jac = xas.setup_jvp(params, freqs[:1], amp_scale, None, "fp64")[1]
        turn = (jac[:, _XAS_INS_CONST, 7] - jac[:, _XAS_ANCHOR, 7])[:, None]
    else:
        turn = torch.zeros((nb, 1), dtype=torch.float64, device=dev)
    heads = torch.stack([head0, with_factors(1j * emm * fp, 1j * emm * fc), with_factors(dfp, dfc),
                         with_factors(0.5j * emm * turn * fp, 0.5j * emm * turn * fc)])
    closed = project(rec0, heads, modes0, freqs, cot_plus, cot_cross)
    grads[:, 4] = -closed[0] / params[:, 4]
    grads[:, 6] = closed[1]
    grads[:, 5] = closed[2]
    grads[:, 7] = closed[3]

    # the differences: every stencil point of every parameter as one batch of binaries (in slices of at
    # most 65,536), set up at once with the selections held at the parameters, projected in one launch
    # each parameter's stencil per binary: central, or one-sided where the central one would cross a
    # boundary (offsets and weights (B, 4), the base point's weight (B,))
    steps = _steps(params)
    stencils = {}
    as_t = lambda values: torch.tensor(values, dtype=torch.float64, device=dev)  # noqa: E731
    c_off, c_w = as_t([o for o, _ in _STENCIL]), as_t([w for _, w in _STENCIL])
    f_off, f_w = as_t([o for o, _ in _FORWARD]), as_t([w for _, w in _FORWARD])
    for j, d in _directions(params, steps).items():
        side = (d != 0)[:, None]
        stencils[j] = (torch.where(side, d[:, None] * f_off, c_off), torch.where(side, d[:, None] * f_w, c_w),
                       d * _FORWARD_BASE)
    points = [(j, s) for j in stencils for s in range(len(_STENCIL))]
    per_slice = max(1, 65536 // nb)
    acc = torch.zeros((8, nb), dtype=torch.float64, device=dev)
    for first in range(0, len(points), per_slice):
        chunk = points[first:first + per_slice]
        p = params.repeat(len(chunk), 1)
        for s, (j, k) in enumerate(chunk):
            p[s * nb:(s + 1) * nb, j] += stencils[j][0][:, k] * steps[j]
        rec, head, modes_, _ = setup(p, freqs[:1], amp_scale, reference_mf)
        rec.view(len(chunk), nb, -1)[:, :, list(XAS_SELECTION_SLOTS)] = rec0[:, list(XAS_SELECTION_SLOTS)]
        modes_.view(len(chunk), nb, 4, -1)[:, :, :, list(MODE_SELECTION_SLOTS)] = \
            modes0[:, :, list(MODE_SELECTION_SLOTS)]
        sums = project(rec, head, modes_, freqs, cot_plus, cot_cross).view(len(chunk), nb)
        for s, (j, k) in enumerate(chunk):
            acc[j] += stencils[j][1][:, k] * sums[s]
        del rec, head, modes_
    for j, (_, _, base) in stencils.items():
        # the base point's own projection is the closed forms' first (the head unchanged)
        grads[:, j] = (acc[j] + torch.where(base != 0, base * closed[0], torch.zeros_like(base))) / steps[j]
    return grads


# ------------------------------------------------ the exact gradient

def adjoint(params: torch.Tensor, freqs: torch.Tensor, in_plus: torch.Tensor, in_cross: torch.Tensor,
            weights: Optional[torch.Tensor] = None, amp_scale: float = 1e21, reference_mf: Optional[float] = None,
            precision: str = "split", wrt_grid: bool = False
            ) -> Tuple[torch.Tensor, Optional[torch.Tensor], Optional[torch.Tensor]]:
    """The exact gradient ``(grads (B, 8), loss (B,) or None, dldf (N,) or
    None)``, float64, of the polarizations on the grid ``freqs``: FP32 on
    the float32 grid (``"split"``) or FP64 on the float64 grid (``"fp64"``),
    in the split evaluation's layout (the same formula in either scalar).
    With ``weights`` None, ``in_plus``/``in_cross`` are the cotangents
    (B, N) (dL/dRe + i dL/dIm) and the loss is None; else they are the data
    and the loss is ``sum_f w (|h+ - d+|^2 + |hx - dx|^2)``. ``wrt_grid``:
    also dL/df of each frequency (the grid's gradient, summed over the
    binaries; the set-up's f_min held, as the reference's stop_gradient).
    One dual-number set-up (the records and their derivatives) and
    ``wfx_xhm_adjoint``: reverse mode per mode and span of frequencies, each
    record's slots' totals contracted once with their derivatives."""
    if precision not in PRECISIONS:
        raise ValueError(f"IMRPhenomXHM's arithmetic is one of {PRECISIONS}, got {precision!r}")
    _require_device(params, "params")
    params = params.to(torch.float64).contiguous()
    rdtype = torch.float64 if precision == "fp64" else torch.float32
    nb, n = params.shape[0], freqs.shape[0]
    j = None
    if nb > 0 and n > 0:
        j = setup_jvp(params, freqs[:1].to(rdtype).to(torch.float64), amp_scale, reference_mf, precision)
    return adjoint_with(j, params, freqs, in_plus, in_cross, weights, precision, wrt_grid)


def adjoint_with(setup: Optional[dict], params: torch.Tensor, freqs: torch.Tensor, in_plus: torch.Tensor,
                 in_cross: torch.Tensor, weights: Optional[torch.Tensor] = None, precision: str = "split",
                 wrt_grid: bool = False
                 ) -> Tuple[torch.Tensor, Optional[torch.Tensor], Optional[torch.Tensor]]:
    """:func:`adjoint` from a derivative set-up already made
    (:func:`setup_jvp`'s dict, in ``precision``, of these ``params`` and
    the grid's first frequency; None only for an empty batch or grid): the
    same launch of ``wfx_xhm_adjoint``. The set-up's head holds each mode's
    factors in the polarizations; a caller may change them (zeroing all but
    one mode's factors and their inclination derivatives gives that mode's
    loss and its exact gradient)."""
    if precision not in PRECISIONS:
        raise ValueError(f"IMRPhenomXHM's arithmetic is one of {PRECISIONS}, got {precision!r}")
    _require_device(params, "params")
    params = params.to(torch.float64).contiguous()
    rdtype = torch.float64 if precision == "fp64" else torch.float32
    ctype = torch.complex128 if precision == "fp64" else torch.complex64
    grid = freqs.to(rdtype).contiguous()
    nb, n = params.shape[0], grid.shape[0]
    in_plus = in_plus.to(ctype).reshape(nb, n).contiguous()
    in_cross = in_cross.to(ctype).reshape(nb, n).contiguous()
    dev = params.device
    if weights is not None:
        weights = weights.to(rdtype).contiguous()
        if weights.shape != (n,):
            raise ValueError(f"weights must be ({n},)")
    dldf = torch.zeros((n,), dtype=torch.float64, device=dev) if wrt_grid else None
    if nb == 0 or n == 0:
        zeros = torch.zeros((nb, 8), dtype=torch.float64, device=dev)
        return zeros, (None if weights is None else torch.zeros((nb,), dtype=torch.float64, device=dev)), dldf
    j = setup
    if j is None or j["records22"].shape[0] != nb or j["head"].dtype != rdtype:
        raise ValueError(f"the set-up must be setup_jvp's of these {nb} binaries in {precision!r}")
    code = _CODES[precision]
    size = lib().wfx_xhm_adjoint_workspace(nb, n, code, int(weights is not None), int(wrt_grid))
    workspace = torch.empty((size,), dtype=torch.uint8, device=dev)
    grads = torch.empty((nb, 8), dtype=torch.float64, device=dev)
    loss = None if weights is None else torch.empty((nb,), dtype=torch.float64, device=dev)
    call("wfx_xhm_adjoint", _ptr(params), _ptr(j["records22"]), _ptr(j["jac22"]), _ptr(j["head"]), _ptr(j["dhead"]),
         _ptr(j["modes"]), _ptr(j["dmodes"]), nb, _ptr(grid), n, code, _ptr(in_plus), _ptr(in_cross), _ptr(weights),
         _ptr(workspace), workspace.numel(), _ptr(loss), _ptr(grads), _ptr(dldf), _stream(dev))
    return grads, loss, dldf


def vjp(params: torch.Tensor, freqs: torch.Tensor, cot_plus: torch.Tensor, cot_cross: torch.Tensor,
        amp_scale: float = 1e21, reference_mf: Optional[float] = None) -> torch.Tensor:
    """dL/d(m1, m2, chi1z, chi2z, d_L, inclination, phi_ref, f_ref) (B, 8)
    float64 of the FP64 polarizations on the float64 grid ``freqs``, for
    the cotangents dL/dRe h + i dL/dIm h (B, N) of h+ and hx: the exact FP64
    gradient (:func:`adjoint` in FP64). f_ref's derivative is zero where
    f_ref is zero or with ``reference_mf``."""
    return adjoint(params, freqs.to(torch.float64), cot_plus, cot_cross, None, amp_scale, reference_mf, "fp64")[0]
