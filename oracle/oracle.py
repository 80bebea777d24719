"""numpy front end of the IMRPhenomXAS oracle (``liboracle_xas.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke``
and ``bench.py``'s CPU-baseline leg, as the checker. The product
(``waveformx_amd``) never imports or links anything here.

The C++ restatement (``xas_oracle.cpp``) follows the reference's FP64 JAX
path (``src/waveformx/models/xas/core.py``, ``xpnr/{internals,amp22,
phase22}.py``); see its header for the file:line list.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_xas.so"
TABLE_SIZE = 86
NAMES = ("mass_1", "mass_2", "chi_1", "chi_2", "luminosity_distance", "inclination", "phi_ref", "f_ref")
INTERNALS = ("eta", "delta", "chi_eff", "m_final", "a_final", "f_ring", "f_damp", "f_meco", "f_isco", "m_sec", "T0",
             "f_phase_in", "f_phase_im", "f_amp_in", "f_amp_rd", "f_cut", "phi2", "phi3", "phi4", "phi5l", "phi6",
             "phi7", "linb", "flags")

_lib = None


def build():
    """Compile the library (``make`` in this folder)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle_xas.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = ctypes.CDLL(str(LIB))
        d, i64, u32, ll = (np.ctypeslib.ndpointer(np.float64, flags="C"), np.ctypeslib.ndpointer(np.int64, flags="C"),
                           np.ctypeslib.ndpointer(np.uint32, flags="C"), np.ctypeslib.ndpointer(np.int64, flags="C"))
        L, D, I = ctypes.c_long, ctypes.c_double, ctypes.c_int
        _lib.oracle_xas_tables.argtypes = [d, d, L, D, D, d, u32]
        _lib.oracle_xas_eval.argtypes = [d, L, d, L, d, d]
        _lib.oracle_xas_polarizations.argtypes = [d, L, d, ctypes.c_void_p, L, D, D, d, d, I]
        _lib.oracle_xas_jacobian.argtypes = [d, L, d, L, D, d, d]
        _lib.oracle_xas_table_jacobian.argtypes = [d, d, L, D, d, d]
        _lib.oracle_xas_lal_form.argtypes = [d, d, L, D, d, d]
        _lib.oracle_qnm22.argtypes = [D, d]
        _lib.oracle_f_isco.argtypes = [D]
        _lib.oracle_f_isco.restype = D
        _lib.oracle_solve.argtypes = [I, d, d, d]
        _lib.oracle_xas_internals.argtypes = [d, D, D, d]
        _lib.oracle_xas_count.argtypes = [d, L, d, L, ll]
        _lib.oracle_libm.argtypes = [I, d, d, L]
    return _lib


def libm(name, x):
    """The product's shared-source FP64 ``exp``, ``log`` or ``cbrt``
    (``waveformx_amd/csrc/wfx_libm.h``, compiled into this library) on an
    array: the functions the leading phase term reads on both sides."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    y = np.empty_like(x)
    lib().oracle_libm(("exp", "log", "cbrt").index(name), x.reshape(-1), y.reshape(-1), x.size)
    return y


def _params(params):
    p = np.ascontiguousarray(np.atleast_2d(np.asarray(params, dtype=np.float64)))
    if p.shape[1] != 8:
        raise ValueError("params must be (B, 8): m1, m2, chi1, chi2, distance, inclination, phi_ref, f_ref")
    return p


def _complex(raw):
    return raw[..., 0] + 1j * raw[..., 1]


def tables(params, f_min, amp_scale=1e21, reference_mf=None):
    """The reference's 86-entry tables (B, 86) and LAL error flags (B,)."""
    p = _params(params)
    b = len(p)
    f_min = np.ascontiguousarray(np.broadcast_to(np.asarray(f_min, dtype=np.float64), (b,)))
    out = np.zeros((b, TABLE_SIZE))
    flags = np.zeros(b, dtype=np.uint32)
    lib().oracle_xas_tables(p, f_min, b, amp_scale, np.nan if reference_mf is None else reference_mf, out, flags)
    return out, flags


def evaluate(table, freqs):
    """(h+, hx) (B, N) complex128 of tables on one shared grid."""
    t = np.ascontiguousarray(np.atleast_2d(table), dtype=np.float64)
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    b, n = len(t), len(f)
    hp = np.zeros((b, n, 2))
    hc = np.zeros((b, n, 2))
    lib().oracle_xas_eval(t, b, f, n, hp, hc)
    return _complex(hp), _complex(hc)


def polarizations(params, freqs, amp_scale=1e21, reference_mf=None, offsets=None, threads=0):
    """(h+, hx) of binaries ``params`` (B, 8) as the reference's
    ``model.polarizations``: f_min is the (each binary's) first frequency.
    Shared grid: (B, N) complex128. Ragged (``offsets`` (B+1,) into the
    concatenated ``freqs``): flat (total,) complex128 each."""
    p = _params(params)
    b = len(p)
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    if offsets is None:
        n = len(f)
        hp = np.zeros((b, n, 2))
        hc = np.zeros((b, n, 2))
        lib().oracle_xas_polarizations(p, b, f, None, n, amp_scale,
                                       np.nan if reference_mf is None else reference_mf, hp, hc, threads)
    else:
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        total = int(off[-1])
        hp = np.zeros((total, 2))
        hc = np.zeros((total, 2))
        lib().oracle_xas_polarizations(p, b, f, off.ctypes.data_as(ctypes.c_void_p), 0, amp_scale,
                                       np.nan if reference_mf is None else reference_mf, hp, hc, threads)
    return _complex(hp), _complex(hc)


def jacobian(params, freqs, amp_scale=1e21):
    """d(h+, hx)/d(params) (B, N, 8) complex128 by forward mode, shared grid
    (the first frequency, f_min, carries no derivative)."""
    p = _params(params)
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    b, n = len(p), len(f)
    dhp = np.zeros((b, n, 8, 2))
    dhc = np.zeros((b, n, 8, 2))
    lib().oracle_xas_jacobian(p, b, f, n, amp_scale, dhp, dhc)
    return _complex(dhp), _complex(dhc)


def table_jacobian(params, f_min, amp_scale=1e21):
    """tables (B, 86) and d tables / d params (B, 86, 8)."""
    p = _params(params)
    b = len(p)
    f_min = np.ascontiguousarray(np.broadcast_to(np.asarray(f_min, dtype=np.float64), (b,)))
    t = np.zeros((b, TABLE_SIZE))
    dt = np.zeros((b, TABLE_SIZE, 8))
    lib().oracle_xas_table_jacobian(p, f_min, b, amp_scale, t, dt)
    return t, dt


def lal_form(params, freqs, amp_scale=1e21):
    """(h+, hx) of one binary by the LAL-form arrangement (xpnr/xas.py)."""
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(8))
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    hp = np.zeros((len(f), 2))
    hc = np.zeros((len(f), 2))
    lib().oracle_xas_lal_form(p, f, len(f), amp_scale, hp, hc)
    return _complex(hp), _complex(hc)


def qnm22(a):
    """(f_ring, f_damp) in powers of 1 - a, and in powers of a."""
    out = np.zeros(4)
    lib().oracle_qnm22(float(a), out)
    return out


def f_isco(a):
    return lib().oracle_f_isco(float(a))


def solve(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(len(b))
    if lib().oracle_solve(len(b), a.reshape(-1), b, x) != 0:
        raise ValueError("n must be 4 or 5")
    return x


def internals(params, f_min, amp_scale=1e21):
    """Named intermediate values of one binary (``INTERNALS``)."""
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(8))
    out = np.zeros(len(INTERNALS))
    lib().oracle_xas_internals(p, float(f_min), amp_scale, out)
    return dict(zip(INTERNALS, out))


def count_ops(table, freqs):
    """Algorithmic op counts (per-frequency part, per-bin part) of the
    formula on these tables and grid: dicts of add, mul, div, sqrt, exp,
    log, sincos, atan, cmp, cbrt, pow."""
    t = np.ascontiguousarray(np.atleast_2d(table), dtype=np.float64)
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    out = np.zeros(22, dtype=np.int64)
    lib().oracle_xas_count(t, len(t), f, len(f), out)
    keys = ("add", "mul", "div", "sqrt", "exp", "log", "sincos", "atan", "cmp", "cbrt", "pow")
    return dict(zip(keys, out[:11].tolist())), dict(zip(keys, out[11:].tolist()))


def draw(n, seed, m1=(10.0, 80.0), q=(0.125, 1.0), chi=(-0.9, 0.9), distance=(100.0, 2000.0), f_ref=20.0):
    """Seeded binaries (B, 8) in the order of the benchmark configurations:
    m1 ~ U, q = m2/m1 ~ U, chi1z, chi2z ~ U, d_L ~ U, inclination =
    arccos(U[-1, 1]), phi_ref ~ U[0, 2 pi], f_ref fixed."""
    rng = np.random.default_rng(seed)
    m1v = rng.uniform(*m1, n)
    qv = rng.uniform(*q, n)
    c1 = rng.uniform(*chi, n)
    c2 = rng.uniform(*chi, n)
    dl = rng.uniform(*distance, n)
    inc = np.arccos(rng.uniform(-1.0, 1.0, n))
    phi = rng.uniform(0.0, 2 * np.pi, n)
    return np.stack([m1v, m1v * qv, c1, c2, dl, inc, phi, np.full(n, f_ref)], axis=1)


if os.environ.get("WFX_ORACLE_BUILD") == "1":
    build()
