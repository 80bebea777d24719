"""waveformx_amd: IMRPhenomXAS (and IMRPhenomD, IMRPhenomXHM) frequency-domain
polarizations on MI355X.

A drop-in for the reference's PyTorch path (``waveformx.torch.
polarizations_function``, ``waveformx.batch_polarizations``): the same
:class:`Binary` and options in, the same ``(hplus, hcross)`` out, computed
by hand-written HIP kernels for gfx950 (``libwfx_xas.so``, C ABI in
``include/waveformx_amd.h``). There is no CPU path.
"""

from ._api import (APPROXIMANTS, approximants, batch_polarizations, compute_device, is_frequency_domain,
                   is_precessing, polarizations, polarizations_on_grid, ragged_polarizations)
from ._modes import FrequencyModes, choose_fd_modes, modes
from ._types import (BINARY_NAMES, COLUMNS, REFERENCE_MF, BatchPolarizations, Binary, FrequencyGrid,
                     WaveformOptions, check_frequencies, parameter_table)
from .exceptions import (InvalidParameterError, ModelError, UnknownApproximantError, UnsupportedError,
                         WaveformError)
from .reference import REFERENCE_MODULES, reference_module
from .reference import package_to_mass_scaled_reference as to_mass_scaled_reference
from .reference import package_to_physical_reference as to_physical_reference

__all__ = [
    "APPROXIMANTS", "approximants", "is_precessing", "is_frequency_domain", "polarizations", "polarizations_on_grid",
    "batch_polarizations", "ragged_polarizations", "compute_device", "Binary",
    "BINARY_NAMES", "COLUMNS", "REFERENCE_MF", "BatchPolarizations", "FrequencyGrid", "WaveformOptions",
    "check_frequencies", "parameter_table", "WaveformError", "InvalidParameterError", "UnsupportedError",
    "UnknownApproximantError", "ModelError", "polarizations_function", "modes", "choose_fd_modes",
    "FrequencyModes", "REFERENCE_MODULES", "reference_module", "to_mass_scaled_reference", "to_physical_reference",
]


def polarizations_function(approximant, options=None, **option_keywords):
    """The PyTorch function of an approximant (see
    :func:`waveformx_amd.torch.polarizations_function`)."""
    from .torch import polarizations_function as f

    return f(approximant, options, **option_keywords)
