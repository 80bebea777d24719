"""The value types of the drop-in boundary, with the reference's fields,
defaults and errors:

* :class:`Binary` (``src/waveformx/_binary.py:18-44``) and the 12-column
  table of binaries (``_binary.py:50-63,129-167``), with LALSimulation's
  common checks applied to a whole table at once (``check_binary``,
  ``_binary.py:108-126``, row by row in the reference);
* :class:`WaveformOptions` (``src/waveformx/_options.py:74-185``);
* :class:`FrequencyGrid`, :func:`check_frequencies` and
  :class:`BatchPolarizations` (``src/waveformx/_results.py:31-68,157-169``).
"""

from __future__ import annotations

import dataclasses
from typing import Any, Dict, Mapping, NamedTuple, Optional, Tuple

import numpy as np

from .exceptions import InvalidParameterError

__all__ = [
    "Binary", "BINARY_NAMES", "COLUMNS", "parameter_table", "check_table", "WaveformOptions", "OPTION_NAMES",
    "REFERENCE_MF", "STRAIN_SCALE", "strain_scale", "parse_options", "FrequencyGrid", "check_frequencies",
    "BatchPolarizations",
]


class Binary(NamedTuple):
    """The parameters of ``SimInspiralChooseFDWaveform``: masses in solar
    masses, dimensionless spin vectors ``(x, y, z)``, distance in Mpc,
    inclination and ``phi_ref`` in radians, ``f_ref`` in Hz (zero: the
    first frequency). Fields may carry a leading batch axis (spins then
    ``(n, 3)``)."""

    mass_1: Any
    mass_2: Any
    spin_1: Any
    spin_2: Any
    luminosity_distance: Any = 1.0
    inclination: Any = 0.0
    phi_ref: Any = 0.0
    f_ref: Any = 0.0


BINARY_NAMES = Binary._fields

COLUMNS = ("mass_1", "mass_2", "spin_1x", "spin_1y", "spin_1z", "spin_2x", "spin_2y", "spin_2z",
           "luminosity_distance", "inclination", "phi_ref", "f_ref")


def parameter_table(parameters: Any) -> np.ndarray:
    """An ``(n, 12)`` float64 table (columns :data:`COLUMNS`) from a mapping
    of arrays (``spin_1``/``spin_2`` as ``(n, 3)`` or the per-component
    columns; absent optional fields take their defaults), a structured
    array or a DataFrame, or an ``(n, 12)`` array."""
    if isinstance(parameters, Mapping) or getattr(getattr(parameters, "dtype", None), "names", None) \
            or hasattr(parameters, "columns"):
        names = set(parameters.keys() if hasattr(parameters, "keys") else parameters.dtype.names)
        if not names:
            raise TypeError("parameters is empty")
        n = len(np.atleast_1d(np.asarray(parameters[next(iter(names))])))
        unknown = names - set(COLUMNS) - set(BINARY_NAMES)
        if unknown:
            raise TypeError(f"unknown parameters {sorted(unknown)}: the columns are {', '.join(COLUMNS)}")
        columns = []
        for column in COLUMNS:
            if column in names:
                values = np.asarray(parameters[column], dtype=float)
            elif column[:-1] in ("spin_1", "spin_2") and column[:-1] in names:
                values = np.asarray(parameters[column[:-1]], dtype=float).reshape(n, 3)[:, "xyz".index(column[-1])]
            elif column in ("mass_1", "mass_2"):
                raise TypeError(f"parameters must contain {column!r}")
            else:
                values = np.full(n, Binary._field_defaults.get(column, 0.0), dtype=float)
            columns.append(np.broadcast_to(values, (n,)))
        return np.stack(columns, axis=1)
    table = np.asarray(parameters, dtype=float)
    if table.ndim != 2 or table.shape[1] != len(COLUMNS):
        raise ValueError(f"parameters must have {len(COLUMNS)} columns ({', '.join(COLUMNS)})")
    return table


def check_table(table: np.ndarray, approximant: str = "IMRPhenomXAS") -> Dict[int, Exception]:
    """LALSimulation's common checks (``_binary.py:108-126``) and the
    aligned-spin check of IMRPhenomX (``xas/api.py:103-108``) on every row
    of a table at once: ``{row: InvalidParameterError}`` for the rows that
    fail, with the reference's first failing check and message."""
    t = np.asarray(table, dtype=float)
    errors: Dict[int, Exception] = {}
    m1, m2 = t[:, 0], t[:, 1]
    s1, s2 = t[:, 2:5], t[:, 5:8]
    checks = [
        (~(np.isfinite(m1) & np.isfinite(m2)) | (m1 <= 0) | (m2 <= 0), "masses must be positive"),
        (~np.all(np.isfinite(s1), axis=1) | (np.linalg.norm(s1, axis=1) > 1.0), "spin_1 must have magnitude <= 1"),
        (~np.all(np.isfinite(s2), axis=1) | (np.linalg.norm(s2, axis=1) > 1.0), "spin_2 must have magnitude <= 1"),
        (~np.isfinite(t[:, 8]) | (t[:, 8] <= 0), "luminosity_distance must be positive"),
        (~np.isfinite(t[:, 11]) | (t[:, 11] < 0), "f_ref must be positive, or zero for the model's default"),
        (~np.isfinite(t[:, 9]), "inclination must be finite"),
        (~np.isfinite(t[:, 10]), "phi_ref must be finite"),
        (np.any(s1[:, :2] != 0.0, axis=1) | np.any(s2[:, :2] != 0.0, axis=1),
         f"non-zero in-plane spins were given, but {approximant} is a non-precessing approximant"),
    ]
    for failed, message in checks:
        for index in np.flatnonzero(failed):
            errors.setdefault(int(index), InvalidParameterError(message))
    return errors


# ------------------------------------------------------------------ options

PRECESSION_VERSIONS = {"nnlo": 102, "msa": 223, "spintaylor": 320, "spin_taylor": 320}
REFERENCE_MF = 0.01
STRAIN_SCALE = 1e21


def _precession_version(value: Any) -> int:
    if isinstance(value, str):
        try:
            return PRECESSION_VERSIONS[value.lower()]
        except KeyError:
            raise InvalidParameterError(f"unknown precession version {value!r}: use 102 (\"NNLO\"), 223 (\"MSA\") "
                                        "or 320 (\"SpinTaylor\")") from None
    if isinstance(value, bool) or not isinstance(value, int):
        raise TypeError("precession_version must be an int (LAL's PhenomXPrecVersion) or a str")
    if value not in (102, 223, 320):
        raise InvalidParameterError(f"precession version {value} is not implemented: use 102 (NNLO), 223 (MSA) or "
                                    "320 (SpinTaylor)")
    return value


def _mode_array(value: Any) -> Optional[Tuple[Tuple[int, int], ...]]:
    if value is None:
        return None
    if isinstance(value, str):
        raise TypeError("mode_array must be a sequence of (l, m) pairs or None")
    modes = []
    for item in value:
        try:
            ell, emm = (int(x) for x in item)
        except (TypeError, ValueError):
            raise TypeError("mode_array must be a sequence of (l, m) pairs or None") from None
        modes.append((ell, emm))
    return tuple(modes)


def _reference_mf(value: Any) -> Optional[float]:
    if value is None:
        return None
    if isinstance(value, bool):
        raise TypeError("reference_mf must be a float or None")
    try:
        value = float(value)
    except (TypeError, ValueError):
        raise TypeError("reference_mf must be a float or None") from None
    if not 0.0 < value < 0.015:
        raise InvalidParameterError(
            f"reference_mf = {value} must lie in (0, 0.015): the inspiral of some binaries ends at M f = 0.015")
    return value


@dataclasses.dataclass(frozen=True)
class WaveformOptions:
    """LALSimulation's ``LALDict`` settings (immutable, hashable). For
    IMRPhenomXAS only ``scaled_strain`` (strain x 1e21, default True) and
    ``reference_mf`` (the reference frequency as ``M f_ref``) act; the
    others are validated and ignored, as in the reference."""

    precession_version: int = 320
    mode_array: Optional[Tuple[Tuple[int, int], ...]] = None
    extended_phase_grid: bool = False
    scaled_strain: bool = True
    reference_mf: Optional[float] = None

    def __post_init__(self):
        object.__setattr__(self, "precession_version", _precession_version(self.precession_version))
        object.__setattr__(self, "mode_array", _mode_array(self.mode_array))
        for name in ("extended_phase_grid", "scaled_strain"):
            if not isinstance(getattr(self, name), bool):
                raise TypeError(f"{name} must be a bool")
        object.__setattr__(self, "reference_mf", _reference_mf(self.reference_mf))

    @staticmethod
    def lalsimulation_defaults() -> "WaveformOptions":
        return WaveformOptions(precession_version=223, scaled_strain=False)

    def replace(self, **options: Any) -> "WaveformOptions":
        _check_names(options)
        return dataclasses.replace(self, **options)

    def as_dict(self) -> dict:
        return {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}


OPTION_NAMES = tuple(f.name for f in dataclasses.fields(WaveformOptions))


def strain_scale(options: WaveformOptions) -> float:
    return STRAIN_SCALE if options.scaled_strain else 1.0


def _check_names(options) -> None:
    for name in options:
        if name not in OPTION_NAMES:
            raise TypeError(f"unexpected keyword argument {name!r}: the options are {', '.join(OPTION_NAMES)}")


def parse_options(options: Optional[WaveformOptions], keywords: Mapping[str, Any]) -> WaveformOptions:
    """``options`` (default ``WaveformOptions()``) changed by ``keywords``
    (``_options.py:177-185``)."""
    if options is None:
        options = WaveformOptions()
    elif not isinstance(options, WaveformOptions):
        raise TypeError("options must be a WaveformOptions or None")
    return options.replace(**keywords) if keywords else options


# -------------------------------------------------------------------- grids

def check_frequencies(frequencies: Any) -> np.ndarray:
    """``frequencies`` as a 1-d float64 array, positive, finite and strictly
    increasing (``_results.py:31-41``)."""
    values = np.asarray(frequencies, dtype=float)
    if values.ndim != 1:
        raise InvalidParameterError("frequencies must be a 1-d array")
    if values.size and (not np.all(np.isfinite(values)) or values[0] <= 0):
        raise InvalidParameterError("frequencies must be positive and finite")
    if np.any(np.diff(values) <= 0):
        raise InvalidParameterError("frequencies must be strictly increasing")
    return values


class FrequencyGrid:
    """Frequencies (Hz) shared by many waveforms (``_results.py:44-68``)."""

    def __init__(self, frequencies: Any):
        self._frequencies = check_frequencies(frequencies)

    @staticmethod
    def uniform(f_min: float, f_max: float, delta_f: float) -> "FrequencyGrid":
        """``f_min, f_min + delta_f, ...`` up to and including ``f_max`` (to
        within half a step)."""
        n = int(np.floor((f_max - f_min) / delta_f + 0.5)) + 1
        return FrequencyGrid(f_min + np.arange(n) * delta_f)

    @property
    def frequencies(self) -> np.ndarray:
        return self._frequencies

    def __len__(self) -> int:
        return int(self._frequencies.shape[0])

    def __repr__(self) -> str:
        f = self._frequencies
        span = f"{f[0]:g}-{f[-1]:g}" if len(self) else "0-0"
        return f"FrequencyGrid({len(self)} frequencies, {span} Hz)"


@dataclasses.dataclass
class BatchPolarizations:
    """``(n_binaries, n_frequencies)`` ``hplus`` and ``hcross``, with ``ok``
    (whether each binary succeeded) and ``errors`` (``{index: exception}``);
    failed rows are zero (``_results.py:157-169``)."""

    hplus: Any
    hcross: Any
    ok: np.ndarray
    errors: Dict[int, Exception]

    def __iter__(self):
        return iter((self.hplus, self.hcross))
