"""The error types of the reference's public API (``src/waveformx/exceptions.py:41-67``),
with the same bases, so that ``except ValueError`` and the like keep working."""

from __future__ import annotations

__all__ = ["WaveformError", "InvalidParameterError", "UnsupportedError", "UnknownApproximantError", "ModelError"]


class WaveformError(Exception):
    """Base class of the errors of waveformx_amd."""


class InvalidParameterError(WaveformError, ValueError):
    """A parameter that LALSimulation rejects for this approximant."""


class UnsupportedError(WaveformError, ValueError):
    """The approximant does not support the request."""


class UnknownApproximantError(WaveformError, ValueError):
    """The name is not one of the implemented approximants."""


class ModelError(WaveformError, RuntimeError):
    """The model could not be evaluated for this binary."""
