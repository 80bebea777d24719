"""The public functions of the drop-in boundary: the approximant registry
(``src/waveformx/models/__init__.py:62-84``, ``_api.py:65-84``),
:func:`polarizations` and :func:`polarizations_on_grid` (one binary,
``_api.py:156-177``), :func:`batch_polarizations` (``_api.py:307-360``)
and :func:`ragged_polarizations` (each binary on its own grid, which the
reference does not have).

IMRPhenomXAS, IMRPhenomD and IMRPhenomXHM are built here; the
reference's other names raise :class:`UnsupportedError`, unknown names
:class:`UnknownApproximantError` (as ``models.get`` does).
"""

from __future__ import annotations

from typing import Any, Mapping, Optional, Sequence, Tuple

import numpy as np
import torch

from . import phenomd, xas, xhm
from ._types import (BINARY_NAMES, BatchPolarizations, Binary, FrequencyGrid, WaveformOptions, check_frequencies,
                     check_table, parameter_table, parse_options, strain_scale)
from .ragged import check_descriptors
from .exceptions import InvalidParameterError, ModelError, UnknownApproximantError, UnsupportedError

__all__ = ["APPROXIMANTS", "approximants", "is_precessing", "is_frequency_domain", "polarizations",
           "polarizations_on_grid", "batch_polarizations", "ragged_polarizations", "compute_device"]

#: approximant -> (domain, precessing, implemented here)
APPROXIMANTS = {
    "TaylorF2": ("frequency", False, False),
    "IMRPhenomD": ("frequency", False, True),
    "IMRPhenomXAS": ("frequency", False, True),
    "IMRPhenomXHM": ("frequency", False, True),
    "IMRPhenomXPNR": ("frequency", True, False),
    "IMRPhenomXPNR_NNLO": ("frequency", True, False),
    "IMRPhenomXPNR_MSA": ("frequency", True, False),
    "NRSur7dq4": ("time", True, False),
}


def approximants(domain: Optional[str] = None) -> list:
    """The approximants built here (of one domain)."""
    if domain not in (None, "frequency", "time"):
        raise ValueError('domain must be None, "frequency" or "time"')
    return [n for n, (d, _, built) in APPROXIMANTS.items() if built and domain in (None, d)]


def _entry(approximant: str) -> Tuple[str, bool, bool]:
    if not isinstance(approximant, str) or approximant not in APPROXIMANTS:
        raise UnknownApproximantError(f"unknown approximant {approximant!r}: waveformx implements "
                                      f"{', '.join(APPROXIMANTS)}")
    return APPROXIMANTS[approximant]


def is_precessing(approximant: str) -> bool:
    """Whether the approximant is precessing (``_api.py:79-80``; any name of
    the registry, built here or not)."""
    return _entry(approximant)[1]


def is_frequency_domain(approximant: str) -> bool:
    """Whether the approximant is a frequency-domain one (``_api.py:83-84``)."""
    return _entry(approximant)[0] == "frequency"


def model(approximant: str, domain: str = "frequency") -> str:
    """Check that ``approximant`` is built here for ``domain``."""
    d, _, built = _entry(approximant)
    if d != domain:
        what = "frequency-domain" if domain == "frequency" else "time-domain"
        raise UnsupportedError(f"{approximant} has no {what} waveforms")
    if not built:
        raise UnsupportedError(f"{approximant} is not implemented in waveformx_amd (IMRPhenomXAS, IMRPhenomD and "
                               "IMRPhenomXHM are)")
    return approximant


def _module(approximant: str):
    """The device module of a built approximant and its flag messages."""
    if approximant == "IMRPhenomD":
        return phenomd, phenomd.LAL_FLAGS
    if approximant == "IMRPhenomXHM":
        return xhm, xhm.LAL_FLAGS
    return xas, xas.LAL_FLAGS


def compute_device(device: Optional[torch.device] = None) -> torch.device:
    """The HIP device to run on: ``device`` if it is one, else the current
    one; raises if there is none (there is no CPU path)."""
    if device is not None and torch.device(device).type == "cuda":
        return torch.device(device)
    if not torch.cuda.is_available():
        raise RuntimeError("waveformx_amd needs a HIP device (MI355X); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def _reference_mf_errors(table: np.ndarray, options: WaveformOptions, errors: dict) -> None:
    # models/__init__.py:87-100
    if options.reference_mf is None:
        return
    for index in np.flatnonzero(table[:, 11] != 0.0):
        errors.setdefault(int(index), InvalidParameterError(
            "with reference_mf, the reference frequency is reference_mf / M: f_ref must be zero "
            "(to_mass_scaled_reference converts a binary given at f_ref)"))


def _table_params(table: np.ndarray, device: torch.device) -> torch.Tensor:
    cols = table[:, [0, 1, 4, 7, 8, 9, 10, 11]]
    return torch.as_tensor(np.ascontiguousarray(cols), dtype=torch.float64, device=device)


def _binary_row(keywords: Mapping[str, Any]) -> np.ndarray:
    """A one-row table of the binary keyword arguments, with the errors of
    ``binary_from_keywords`` (``src/waveformx/_binary.py:82-106``)."""
    for name in ("mass_1", "mass_2"):
        if name not in keywords:
            raise TypeError(f"missing required keyword argument {name!r}")

    def real(name: str) -> float:
        value = keywords.get(name, Binary._field_defaults.get(name))
        try:
            return float(value)
        except (TypeError, ValueError):
            raise TypeError(f"{name} must be a real number, got {type(value).__name__}") from None

    def spin(name: str) -> list:
        try:
            value = np.asarray(keywords.get(name, (0.0, 0.0, 0.0)), dtype=float)
        except (TypeError, ValueError):
            raise TypeError(f"{name} must be a sequence of three real numbers (x, y, z)") from None
        if value.shape != (3,):
            raise TypeError(f"{name} must be a sequence of three real numbers (x, y, z)")
        return list(value)

    row = [real("mass_1"), real("mass_2"), *spin("spin_1"), *spin("spin_2"), real("luminosity_distance"),
           real("inclination"), real("phi_ref"), real("f_ref")]
    return np.asarray(row, dtype=float)[None, :]


def polarizations(approximant: str, frequencies: Any, *, options: Optional[WaveformOptions] = None,
                  device: Optional[torch.device] = None, **keywords) -> Tuple[torch.Tensor, torch.Tensor]:
    """``(h_+, h_x)`` of one binary (keyword arguments ``mass_1``, ``mass_2``,
    ``spin_1``, ... and options) at ``frequencies`` (Hz), FP64 on the GPU,
    as complex128 device tensors (``src/waveformx/_api.py:156-170``). A
    binary that fails LALSimulation's checks raises
    :class:`InvalidParameterError` with the reference's message."""
    model(approximant)
    binary = {k: v for k, v in keywords.items() if k in BINARY_NAMES}
    option_keywords = {k: v for k, v in keywords.items() if k not in BINARY_NAMES}
    row = _binary_row(binary)
    options = parse_options(options, option_keywords)
    # the parameter checks on the host, before any device work (check_binary,
    # _binary.py:108-126, then the grid and the model's checks)
    errors = check_table(row, approximant)
    frequencies = check_frequencies(frequencies)
    _reference_mf_errors(row, options, errors)
    if errors:
        raise errors[0]
    batch = batch_polarizations(approximant, frequencies, row, strict=True, options=options, device=device)
    return batch.hplus[0], batch.hcross[0]


def polarizations_on_grid(approximant: str, grid: Any, *, options: Optional[WaveformOptions] = None,
                          device: Optional[torch.device] = None, **keywords) -> Tuple[torch.Tensor, torch.Tensor]:
    """:func:`polarizations` on a :class:`FrequencyGrid` (or frequencies)
    (``src/waveformx/_api.py:173-177``)."""
    if not isinstance(grid, FrequencyGrid):
        grid = FrequencyGrid(grid)
    return polarizations(approximant, grid.frequencies, options=options, device=device, **keywords)


def batch_polarizations(approximant: str, grid: Any, parameters: Any, *, threads: int = 1, strict: bool = False,
                        exact: bool = False, chunk_size: Optional[int] = None,
                        options: Optional[WaveformOptions] = None, device: Optional[torch.device] = None,
                        **option_keywords) -> BatchPolarizations:
    """Polarizations of many binaries on one grid (complex128, FP64 on the
    GPU), with ``ok`` and ``errors``: rows that fail LALSimulation's checks,
    or whose waveform is not finite, are zero and reported, unless
    ``strict`` (the first error is raised). ``threads`` and ``exact`` are
    accepted and ignored. ``hplus`` and ``hcross`` are device tensors.

    The checks run on the whole table at once (host arrays for the
    parameter checks, the setup kernel's flags for LAL's), where the
    reference loops over rows."""
    model(approximant)
    options = parse_options(options, option_keywords)
    if not isinstance(grid, FrequencyGrid):
        grid = FrequencyGrid(grid)
    table = parameter_table(parameters)
    n, nf = len(table), len(grid)
    dev = compute_device(device)
    errors = check_table(table, approximant)
    _reference_mf_errors(table, options, errors)
    hplus = torch.zeros((n, nf), dtype=torch.complex128, device=dev)
    hcross = torch.zeros((n, nf), dtype=torch.complex128, device=dev)
    if n == 0:
        return BatchPolarizations(hplus, hcross, np.ones(0, dtype=bool), {})
    freqs = torch.as_tensor(grid.frequencies, dtype=torch.float64, device=dev)
    params = _table_params(table, dev)
    step = n if chunk_size is None else max(1, int(chunk_size))
    flags = torch.zeros((n,), dtype=torch.int32, device=dev)
    module, lal_flags = _module(approximant)
    if nf:
        for start in range(0, n, step):
            stop = min(n, start + step)
            if approximant == "IMRPhenomXHM":
                records, head, modes_, flags[start:stop] = xhm.setup(params[start:stop], freqs[:1],
                                                                     strain_scale(options), options.reference_mf)
                hplus[start:stop], hcross[start:stop] = xhm.evaluate(records, head, modes_, freqs)
                continue
            records, flags[start:stop] = module.setup(params[start:stop], freqs[:1], strain_scale(options),
                                                      options.reference_mf, "fp64")
            module.evaluate(records, freqs, "fp64", out=(hplus[start:stop], hcross[start:stop]))
    flags_host = flags.cpu().numpy()
    for bit, message in lal_flags:
        for index in np.flatnonzero(flags_host & bit):
            errors.setdefault(int(index), InvalidParameterError(message))
    ok = np.ones(n, dtype=bool)
    ok[list(errors)] = False
    if nf:
        finite = (torch.isfinite(hplus).all(dim=1) & torch.isfinite(hcross).all(dim=1)).cpu().numpy()
        for index in np.flatnonzero(ok & ~finite):
            ok[index] = False
            errors[int(index)] = ModelError(f"{approximant} could not be evaluated for binary {index}")
    if strict and errors:
        raise errors[min(errors)]
    bad = torch.as_tensor(~ok, device=dev)
    hplus[bad] = 0
    hcross[bad] = 0
    return BatchPolarizations(hplus, hcross, ok, errors)


def ragged_polarizations(approximant: str, grids: Sequence[Any], parameters: Any, *,
                         precision: str = "fp64", options: Optional[WaveformOptions] = None,
                         device: Optional[torch.device] = None,
                         **option_keywords) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Polarizations of binaries each on its own grid, in one launch.
    ``grids``: one frequency array per binary, or ``(f0, df, n)`` triples
    (uniform grids generated on the device). Returns the flat (total,)
    ``hplus``, ``hcross`` and the (B+1,) int64 ``offsets`` of each binary's
    bins. No parameter checks (as the reference's pure functions)."""
    model(approximant)
    options = parse_options(options, option_keywords)
    table = parameter_table(parameters)
    dev = compute_device(device)
    if len(grids) != len(table):
        raise ValueError("give one grid per binary")
    params = _table_params(table, dev)
    module, _ = _module(approximant)
    descriptors = all(isinstance(g, tuple) and len(g) == 3 for g in grids)
    if descriptors:
        f0 = np.array([float(g[0]) for g in grids])
        df = np.array([float(g[1]) for g in grids])
        lengths = np.array([int(g[2]) for g in grids], dtype=np.int64)
        check_descriptors(f0, df, lengths)
        offsets = np.concatenate([[0], np.cumsum(lengths)])
        off = torch.as_tensor(offsets, device=dev)
        f0_t = torch.as_tensor(f0, device=dev)
        df_t = torch.as_tensor(df, device=dev)
        f_min = f0_t if precision in ("fp64", "split", "split-df") else f0_t.to(torch.float32).to(torch.float64)
        if approximant == "IMRPhenomXHM":
            records, head, modes_, _ = module.setup(params, f_min, strain_scale(options), options.reference_mf,
                                                    precision)
            hp, hc = module.evaluate_ragged(records, head, modes_, off, grid_f0=f0_t, grid_df=df_t,
                                            total=int(offsets[-1]))
            return hp, hc, off
        records, _ = module.setup(params, f_min, strain_scale(options), options.reference_mf, precision)
        hp, hc = module.evaluate_ragged(records, off, grid_f0=f0_t, grid_df=df_t, precision=precision,
                                        total=int(offsets[-1]))
        return hp, hc, off
    arrays = [np.asarray(g, dtype=float) for g in grids]
    offsets = np.concatenate([[0], np.cumsum([len(a) for a in arrays])]).astype(np.int64)
    dtype = torch.float64 if precision == "fp64" else torch.float32
    freqs = torch.as_tensor(np.concatenate(arrays) if arrays else np.zeros(0), dtype=dtype, device=dev)
    off = torch.as_tensor(offsets, device=dev)
    starts = torch.as_tensor(offsets[:-1], device=dev).clamp(max=max(freqs.shape[0] - 1, 0))
    f_min = freqs[starts].to(torch.float64) if freqs.shape[0] else torch.zeros(len(table), device=dev,
                                                                                 dtype=torch.float64)
    if approximant == "IMRPhenomXHM":
        records, head, modes_, _ = module.setup(params, f_min, strain_scale(options), options.reference_mf,
                                                precision)
        hp, hc = module.evaluate_ragged(records, head, modes_, off, freqs=freqs)
        return hp, hc, off
    records, _ = module.setup(params, f_min, strain_scale(options), options.reference_mf, precision)
    hp, hc = module.evaluate_ragged(records, off, freqs=freqs, precision=precision)
    return hp, hc, off
