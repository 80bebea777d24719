"""Emit ``phenomd_data.json``: the calibration data of IMRPhenomD that the
oracle's restatement (``phenomd_oracle.py``) evaluates.

TEST INFRASTRUCTURE ONLY (the oracle's input data). The data are
LALSimulation's (via waveformrs), which the reference holds as Python
literals: the phenomenological fits ``COEFFICIENTS`` and the final-spin fit
``_FINAL_SPIN`` (``src/waveformx/models/phenomd/fits.py:40-128``) and the
Chebyshev series of the QNM frequencies ``FRING``, ``FDAMP``
(``src/waveformx/models/phenomd/_qnm_fit.py:43-89``). They are read from
that text as an AST (never imported) and written out unchanged.

Usage: ``python oracle/gen_phenomd.py [REFERENCE_ROOT]`` (default
``/root/reference``); it rewrites ``oracle/phenomd_data.json``.
"""

from __future__ import annotations

import ast
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def literals(path: Path, names):
    """The module-level literal assignments ``names`` of ``path``."""
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


def generate(root: Path) -> dict:
    base = root / "src/waveformx/models/phenomd"
    fits = literals(base / "fits.py", ["COEFFICIENTS", "_FINAL_SPIN"])
    qnm = literals(base / "_qnm_fit.py", ["FRING", "FDAMP"])
    return {
        "source": "src/waveformx/models/phenomd/fits.py, _qnm_fit.py",
        "coefficients": {k: [list(row) for row in v] for k, v in fits["COEFFICIENTS"].items()},
        "final_spin": [list(row) for row in fits["_FINAL_SPIN"]],
        "fring": list(qnm["FRING"]),
        "fdamp": list(qnm["FDAMP"]),
    }


def main():
    root = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
    data = generate(root)
    (HERE / "phenomd_data.json").write_text(json.dumps(data, indent=1) + "\n")


if __name__ == "__main__":
    main()
