"""ctypes binding of ``libwfx_xas.so`` (the C ABI of include/waveformx_amd.h).

There is no fallback: without the built library, or without a HIP device,
every entry point raises.
"""

from __future__ import annotations

import ctypes
from pathlib import Path

LIBRARY = Path(__file__).resolve().parent / "libwfx_xas.so"

_lib = None

_P, _I64, _I, _D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_double
ABI_VERSION = 25
SIGNATURES = {
    "wfx_last_error": ([], ctypes.c_char_p),
    "wfx_abi_version": ([], _I),
    "wfx_xas_record_slots": ([], _I),
    "wfx_xas_record_bytes": ([_I], _I),
    "wfx_xas_setup": ([_P, _I64, _P, _I64, _D, _D, _I, _P, _P, _P], _I),
    "wfx_xas_setup_jvp": ([_P, _I64, _P, _I64, _D, _D, _P, _P, _P, _P, _P], _I),
    "wfx_xas_eval": ([_P, _I64, _P, _I, _I64, _I, _I, _I, _P, _P, _P], _I),
    "wfx_xas_reference_phase": ([_P, _I64, _P, _P, _P], _I),
    "wfx_xas_reference_phase_grad": ([_P, _I64, _P, _P, _P, _P], _I),
    "wfx_xas_fast_atan": ([_P, _P, _I64, _P], _I),
    "wfx_xas_tile_prefix": ([_P, _I64, _I, _P, _P], _I),
    "wfx_xas_eval_ragged": ([_P, _I64, _P, _I, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P], _I),
    "wfx_xas_power_table_bytes": ([_I], _I),
    "wfx_xas_power_table": ([_P, _I, _I64, _I, _P, _P], _I),
    "wfx_xas_adjoint": ([_P, _P, _I64, _P, _I, _I64, _I, _P, _P, _P, _P, _P], _I),
    "wfx_xas_grid_adjoint_workspace": ([_I64, _I64], _I64),
    "wfx_xas_grid_adjoint": ([_P, _I64, _P, _I, _I64, _I, _P, _P, _P, _P, _I64, _P], _I),
    "wfx_xas_loss_adjoint": ([_P, _P, _I64, _P, _I, _I64, _I, _P, _P, _P, _P, _P, _P, _P], _I),
    "wfx_xas_loss_grad_static": ([_P, _I64, _P, _I64, _D, _D, _P, _I, _I64, _I, _P, _P, _P, _P, _P, _P], _I),
    "wfx_phd_setup": ([_P, _I64, _P, _I64, _D, _D, _I, _P, _P, _P], _I),
    "wfx_phd_reference_phase": ([_P, _I64, _P, _P, _P], _I),
    "wfx_phd_reference_phase_grad": ([_P, _I64, _P, _P, _P, _P], _I),
    "wfx_phd_eval": ([_P, _I64, _P, _I, _I64, _I, _I, _I, _P, _P, _P], _I),
    "wfx_phd_eval_ragged": ([_P, _I64, _P, _I, _P, _P, _P, _P, _I, _I, _I, _P, _P, _P], _I),
    "wfx_phd_setup_jvp": ([_P, _I64, _P, _I64, _D, _D, _P, _P, _P, _P, _P], _I),
    "wfx_phd_adjoint": ([_P, _P, _I64, _P, _I, _I64, _I, _P, _P, _P, _P, _P], _I),
    "wfx_phd_loss_adjoint": ([_P, _P, _I64, _P, _I, _I64, _I, _P, _P, _P, _P, _P, _P, _P], _I),
    "wfx_phd_grid_adjoint": ([_P, _I64, _P, _I, _I64, _I, _P, _P, _P, _P, _I64, _P], _I),
    "wfx_xhm_record_slots": ([], _I),
    "wfx_xhm_head_slots": ([], _I),
    "wfx_xhm_setup": ([_P, _I64, _P, _I64, _D, _D, _I, _P, _P, _P, _P, _P], _I),
    "wfx_xhm_eval_ragged": ([_P, _P, _P, _I64, _P, _I, _P, _P, _P, _P, _I, _I, _P, _P, _P, _P], _I),
    "wfx_xhm_setup_jvp_workspace": ([_I64], _I64),
    "wfx_xhm_setup_jvp": ([_P, _I64, _P, _I64, _D, _D, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _P], _I),
    "wfx_xhm_adjoint_workspace": ([_I64, _I64, _I, _I, _I], _I64),
    "wfx_xhm_adjoint": ([_P, _P, _P, _P, _P, _P, _P, _I64, _P, _I64, _I, _P, _P, _P, _P, _I64, _P, _P, _P, _P], _I),
    "wfx_xhm_eval": ([_P, _P, _P, _I64, _P, _I, _I64, _I, _P, _P, _P, _P], _I),
    "wfx_xhm_project": ([_P, _P, _I, _P, _I64, _P, _I64, _P, _P, _I64, _P, _P], _I),
}


def lib() -> ctypes.CDLL:
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise RuntimeError(f"{LIBRARY.name} is not built: run `python -m waveformx_amd.build` "
                               "(waveformx_amd has no CPU fallback)")
        handle = ctypes.CDLL(str(LIBRARY))
        version = getattr(handle, "wfx_abi_version", None)
        found = version() if version is not None else None
        if found != ABI_VERSION:
            raise RuntimeError(f"{LIBRARY.name} has ABI version {found}, this binding needs {ABI_VERSION}: rebuild "
                               "it with `python -m waveformx_amd.build`")
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = res
        _lib = handle
    return _lib


def call(name: str, *args) -> None:
    """Call an entry point and raise RuntimeError with its message on a
    non-zero status."""
    status = getattr(lib(), name)(*args)
    if status != 0:
        message = lib().wfx_last_error().decode()
        raise RuntimeError(message if message.startswith(name) else f"{name}: {message}")
