"""Emit ``phenomd_fits.inc``: the calibration data of IMRPhenomD as
``constexpr double`` arrays for the per-binary setup kernels.

The data are LALSimulation's (via waveformrs), which the reference holds as
Python literals: the phenomenological fits ``COEFFICIENTS`` and the
final-spin fit ``_FINAL_SPIN`` (``src/waveformx/models/phenomd/fits.py:
40-128``) and the Chebyshev series of the QNM frequencies ``FRING``,
``FDAMP`` (``src/waveformx/models/phenomd/_qnm_fit.py:43-89``). They are
read from that text as an AST (never imported) and written with
``repr`` (the shortest string that round-trips the double). The constants
that the reference computes in Python floats from pi (``pi ** (4/3)`` and
the like, ``fits`` scalings in ``core.py:169-172,222``) and the amplitude
prefactor ``K_AMP`` (``core.py:107-108``) are evaluated here in the same
Python float arithmetic, so that device and host read identical doubles.

Usage: ``python waveformx_amd/csrc/gen_phenomd.py [REFERENCE_ROOT]``.
"""

from __future__ import annotations

import ast
import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent

#: the fits in the order of the setup's enum (phenomd_setup.h)
FIT_NAMES = ("rho1", "rho2", "rho3", "gamma1", "gamma2", "gamma3", "collocation", "alpha1", "alpha2", "alpha3",
             "alpha4", "alpha5", "beta1", "beta2", "beta3", "sigma1", "sigma2", "sigma3", "sigma4")


def literals(path: Path, names):
    tree = ast.parse(path.read_text())
    found = {}
    for node in tree.body:
        if isinstance(node, (ast.Assign, ast.AnnAssign)):
            targets = node.targets if isinstance(node, ast.Assign) else [node.target]
            for target in targets:
                if isinstance(target, ast.Name) and target.id in names:
                    value = node.value
                    if isinstance(value, ast.Call) and getattr(value.func, "id", "") == "dict":
                        found[target.id] = {kw.arg: ast.literal_eval(kw.value) for kw in value.keywords}
                    else:
                        found[target.id] = ast.literal_eval(value)
    missing = set(names) - set(found)
    if missing:
        raise SystemExit(f"{path}: not found: {sorted(missing)}")
    return found


def array(name, values, shape):
    flat = np.asarray(values, dtype=np.float64).reshape(-1)
    body = ", ".join(repr(float(x)) for x in flat)
    dims = "".join(f"[{n}]" for n in shape)
    return f"constexpr double {name}{dims} = {{{body}}};"


def generate(root: Path) -> str:
    base = root / "src/waveformx/models/phenomd"
    fits = literals(base / "fits.py", ["COEFFICIENTS", "_FINAL_SPIN"])
    qnm = literals(base / "_qnm_fit.py", ["FRING", "FDAMP"])
    coefficients = fits["COEFFICIENTS"]
    if set(coefficients) != set(FIT_NAMES):
        raise SystemExit(f"unexpected fits: {sorted(coefficients)}")
    pi = math.pi
    mtsun, mrsun = 4.925490947641266978197229498498379006e-6, 1.476625038050124729627979840144936351e3
    mpc = 1.0e6 * 3.085677581491367278913937957796471611e16
    k_amp = 2.0 * float(np.sqrt(5.0 / (64.0 * np.pi))) * mrsun * mtsun / mpc \
        * float(np.sqrt(2.0 / 3.0) * np.pi ** (-1.0 / 6.0)) * mtsun ** (-7.0 / 6.0)
    scalars = {
        "kPdPi2": pi ** 2, "kPdPi3": pi ** 3, "kPdPi43": pi ** (4.0 / 3.0), "kPdPi53": pi ** (5.0 / 3.0),
        "kPdPi73": pi ** (7.0 / 3.0), "kPdPi83": pi ** (8.0 / 3.0), "kPdLn4": float(np.log(4.0)), "kPdKAmp": k_amp,
        "kPdK0": -3.0 / 128.0 * float(np.pi * mtsun) ** (-5.0 / 3.0),
        "kPdEulerGamma": 0.577215664901532860606512090082402431,
    }
    lines = [
        "// GENERATED by waveformx_amd/csrc/gen_phenomd.py from the reference's",
        "// src/waveformx/models/phenomd/fits.py and _qnm_fit.py; do not edit.",
        "#pragma once",
        "",
        "namespace wfx {",
        "",
        "// the phenomenological fits: [fit][power of xi][power of eta] (fits.py:40-128)",
        array("kPdFits", [coefficients[k] for k in FIT_NAMES], (len(FIT_NAMES), 4, 3)),
        "// FinalSpin0815: [power of s][power of eta] (fits.py:118-128)",
        array("kPdFinalSpin", fits["_FINAL_SPIN"], (5, 5)),
        "// Chebyshev series of Mf_ring and Mf_damp, lowest degree first (_qnm_fit.py:43-89)",
        array("kPdFring", qnm["FRING"], (len(qnm["FRING"]),)),
        array("kPdFdamp", qnm["FDAMP"], (len(qnm["FDAMP"]),)),
        "",
    ]
    for name, value in scalars.items():
        lines.append(f"constexpr double {name} = {value!r};")
    lines += ["", "}  // namespace wfx", ""]
    return "\n".join(lines)


def main():
    root = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
    (HERE / "phenomd_fits.inc").write_text(generate(root))


if __name__ == "__main__":
    main()
