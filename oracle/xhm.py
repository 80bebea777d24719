"""numpy front end of the IMRPhenomXHM oracle (``xhm_oracle.h``, compiled
into ``liboracle_xas.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke``
and ``bench.py``'s CPU-baseline leg, as the checker. The product
(``waveformx_amd``) never imports or links anything here.

The C++ restatement follows the reference's LAL form of IMRPhenomXHM
(release 122022): ``src/waveformx/models/xhm/reference.py`` and
``xpnr/hm/{variables,inspiral,phase}.py``; see the header of
``xhm_oracle.h`` for the file:line list.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import oracle

#: The modes (l, m) in the order of the outputs: the (2,2), then LAL's
#: higher modes (xhm/reference.py:55-57).
MODES = ((2, 2), (2, 1), (3, 3), (3, 2), (4, 4))
HIGHER_MODES = MODES[1:]
INTERNALS = ("f_ring", "f_damp", "f_amp_match_in", "f_amp_match_im", "f_phase_match_in", "f_phase_match_im",
             "c1_insp", "c_insp", "c1_rd", "c_rd", "deltaphi", "c0", "c1", "c2", "c3", "c4", "c_l", "amp0")

_ready = False


def lib():
    global _ready
    handle = oracle.lib()
    if not _ready:
        d = np.ctypeslib.ndpointer(np.float64, flags="C")
        L, D, I, P = ctypes.c_long, ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        handle.oracle_xhm_polarizations.argtypes = [d, L, d, L, D, d, d, P, I]
        handle.oracle_xhm_mode_internals.argtypes = [d, D, I, d]
        handle.oracle_xhm_mode_amplitude_phase.argtypes = [d, D, I, d, L, d, d, P]
        handle.oracle_swsh.argtypes = [D, D, I, I, d]
        handle.oracle_xhm_collocation.argtypes = [d, D, I, d]
        handle.oracle_xhm_count.argtypes = [d, L, d, L, np.ctypeslib.ndpointer(np.int64, flags="C")]
        _ready = True
    return handle


def draw(nb: int, seed: int = 4) -> np.ndarray:
    """The parameters of the multimode workload, (nb, 8) = (m1, m2, chi1z,
    chi2z, d_L, inclination, phi_ref, f_ref): m1 ~ U[10, 80], q ~ U[0.1, 1],
    spins ~ U[-0.9, 0.9], d_L ~ U[100, 2000] Mpc, cos(inclination) ~
    U[-1, 1], phi_ref ~ U[0, 2 pi], f_ref = 20 Hz."""
    rng = np.random.default_rng(seed)
    m1 = rng.uniform(10.0, 80.0, nb)
    q = rng.uniform(0.1, 1.0, nb)
    chi1 = rng.uniform(-0.9, 0.9, nb)
    chi2 = rng.uniform(-0.9, 0.9, nb)
    distance = rng.uniform(100.0, 2000.0, nb)
    inclination = np.arccos(rng.uniform(-1.0, 1.0, nb))
    phi = rng.uniform(0.0, 2.0 * np.pi, nb)
    return np.stack([m1, q * m1, chi1, chi2, distance, inclination, phi, np.full(nb, 20.0)], axis=1)


def polarizations(params, freqs, amp_scale: float = 1.0, modes: bool = False, threads: int = 0):
    """``(h_+, h_x)`` (complex128, (nb, n)) of IMRPhenomXHM on the grid
    ``freqs`` (f_min = freqs[0]) and, with ``modes``, the five modes
    ``h_{l,-m}`` of :data:`MODES` as (5, nb, n); ``threads`` OpenMP threads
    (0: the default)."""
    params = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
    freqs = np.ascontiguousarray(freqs, dtype=np.float64)
    nb, n = params.shape[0], freqs.shape[0]
    hp = np.empty((nb, n), np.complex128)
    hc = np.empty((nb, n), np.complex128)
    out = np.empty((5, nb, n), np.complex128) if modes else None
    lib().oracle_xhm_polarizations(params, nb, freqs, n, amp_scale, hp.view(np.float64), hc.view(np.float64),
                                   out.ctypes.data if modes else None, threads)
    return (hp, hc, out) if modes else (hp, hc)


def mode_internals(params, f_min: float, k: int) -> dict:
    """The set-up of higher mode ``k`` (0..3 for (2,1), (3,3), (3,2), (4,4))
    of one binary: :data:`INTERNALS`."""
    out = np.empty(len(INTERNALS))
    lib().oracle_xhm_mode_internals(np.ascontiguousarray(params, dtype=np.float64), f_min, k, out)
    return dict(zip(INTERNALS, out))


def mode_amplitude_phase(params, f_min: float, k: int, mf, rotation: bool = False):
    """Amplitude (without ``(-1)^l amp0``) and phase of higher mode ``k`` at
    the mass-scaled frequencies ``mf``, and for the (3,2) with
    ``rotation`` the rotated ringdown phase and its first two derivatives
    in ``Mf`` (3, n)."""
    mf = np.ascontiguousarray(mf, dtype=np.float64)
    amp = np.empty_like(mf)
    phase = np.empty_like(mf)
    rot = np.empty((mf.size, 3)) if rotation else None
    lib().oracle_xhm_mode_amplitude_phase(np.ascontiguousarray(params, dtype=np.float64), f_min, k, mf, mf.size,
                                          amp, phase, rot.ctypes.data if rotation else None)
    return (amp, phase, rot.T.copy()) if rotation else (amp, phase)


def collocation(params, f_min: float, k: int) -> dict:
    """The collocation conditions of higher mode ``k``'s set-up: the phase
    derivative's points (M f) and values, and the intermediate amplitude's
    fitted points and amplitudes (without amp0)."""
    out = np.empty(26)
    lib().oracle_xhm_collocation(np.ascontiguousarray(params, dtype=np.float64), f_min, k, out)
    n_p, n_a = int(out[0]), int(out[13])
    return {"phase_f": out[1:1 + n_p], "phase_v": out[7:7 + n_p], "amp_f": out[14:14 + n_a],
            "amp_v": out[20:20 + n_a]}


def swsh(theta: float, phi: float, ell: int, m: int) -> complex:
    """The spin-weighted spherical harmonic of spin weight -2."""
    out = np.empty(2)
    lib().oracle_swsh(theta, phi, ell, m, out)
    return complex(out[0], out[1])


def jacobian(params, freqs, amp_scale: float = 1.0, threads: int = 0):
    """d(h_+, h_x)/d(m1, m2, chi1z, chi2z, d_L, inclination, phi_ref, f_ref),
    each (nb, n, 8) complex128, by fourth-order central differences of
    :func:`polarizations` (steps: masses 1e-6 relative, spins 2e-6,
    d_L 1e-5 relative, angles 2e-3, f_ref 2e-6 relative; zero for f_ref = 0,
    which takes f_min). The gradient checker of the GPU path's differences,
    on grids away from the (2,2)'s cutoff and region bounds, where the
    difference of the whole model is that of one smooth formula."""
    params = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
    nb, n = params.shape[0], len(freqs)
    jp = np.zeros((nb, n, 8), np.complex128)
    jc = np.zeros((nb, n, 8), np.complex128)
    scale = np.abs(params)
    steps = [1e-6 * scale[:, 0], 1e-6 * scale[:, 1], np.full(nb, 2e-6), np.full(nb, 2e-6), 1e-5 * scale[:, 4],
             np.full(nb, 2e-3), np.full(nb, 2e-3), 2e-6 * scale[:, 7]]
    for j, h in enumerate(steps):
        if j == 7 and not np.any(params[:, 7] > 0):
            continue
        acc_p = np.zeros((nb, n), np.complex128)
        acc_c = np.zeros((nb, n), np.complex128)
        for offset, weight in ((-2, 1.0 / 12.0), (-1, -8.0 / 12.0), (1, 8.0 / 12.0), (2, -1.0 / 12.0)):
            p = params.copy()
            p[:, j] += offset * h
            hp, hc = polarizations(p, freqs, amp_scale, threads=threads)
            acc_p += weight * hp
            acc_c += weight * hc
        safe = np.where(h > 0, h, 1.0)[:, None]
        jp[:, :, j] = np.where(h[:, None] > 0, acc_p / safe, 0.0)
        jc[:, :, j] = np.where(h[:, None] > 0, acc_c / safe, 0.0)
    return jp, jc


#: the op tally's kinds, in oracle_xhm_count's order
OP_KINDS = ("add", "mul", "div", "sqrt", "exp", "log", "sincos", "atan", "cmp", "cbrt", "pow")


def count_ops(params, freqs):
    """Algorithmic op counts (per-frequency part, per-bin part) of the
    polarizations of ``params`` (B, 8) on the shared grid ``freqs``: the
    branches each bin takes, the powers of M f derived from the frequency's
    (``oracle_xhm_count``): dicts of :data:`OP_KINDS`."""
    p = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    out = np.zeros(22, dtype=np.int64)
    lib().oracle_xhm_count(p, len(p), f, len(f), out)
    return dict(zip(OP_KINDS, out[:11].tolist())), dict(zip(OP_KINDS, out[11:].tolist()))


def op_slots(per_bin: dict, bins: int) -> float:
    """Issue slots per bin of a tally in FP32 hardware terms: one per add,
    multiplication, comparison; a division as a reciprocal and a
    multiplication; the transcendentals (exp, log, sqrt, cbrt, a sine with
    its cosine counted twice, an arctangent as its fitted polynomial's eight
    operations and a reciprocal) at the quarter rate of the hardware's
    transcendental unit, four slots each."""
    quarter = 4.0
    slots = per_bin["add"] + per_bin["mul"] + per_bin["cmp"] + per_bin["div"] * (1 + quarter)
    slots += quarter * (per_bin["exp"] + per_bin["log"] + per_bin["sqrt"] + per_bin["cbrt"] + 2 * per_bin["sincos"])
    slots += per_bin["atan"] * (8 + quarter)
    return slots / bins

