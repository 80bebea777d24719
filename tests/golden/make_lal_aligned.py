"""Write ``lal_xpnr_aligned.json``: the two exactly aligned-spin binaries of
the reference's LALSimulation 6.2.0 dump of IMRPhenomXPNR (cases 7 and 8 of
``tests/data/imrphenomxpnr_parityfix.json``, in-plane spins zero), their
parameters, grids and LAL's polarizations, unchanged (data, not code).

With no in-plane spin, IMRPhenomXPNR's angles and antisymmetric mode
vanish and it is IMRPhenomXHM but for its choices of release (122019 for
the higher modes, against IMRPhenomXHM's default 122022) and of the final
spin: the only LAL-computed values this image holds for the aligned-spin
models. ``tests/test_oracle_pins.py`` compares the oracles with them.

Usage: ``python tests/golden/make_lal_aligned.py [REFERENCE_ROOT]``
(default ``/root/reference``).
"""

import json
import sys
from pathlib import Path

CASES = (7, 8)
KEYS = ("mass_1", "mass_2", "spin_1", "spin_2", "luminosity_distance", "inclination", "phi_ref", "f_min", "f_ref",
        "f_max", "frequencies", "hplus_real", "hplus_imag", "hcross_real", "hcross_imag")


def main():
    root = Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference")
    dump = json.loads((root / "tests" / "data" / "imrphenomxpnr_parityfix.json").read_text())
    cases = []
    for index in CASES:
        case = dump["cases"][index]
        assert case["spin_1"][:2] == [0.0, 0.0] and case["spin_2"][:2] == [0.0, 0.0], index
        assert "mode_array" not in case, index
        cases.append({"source_index": index, **{k: case[k] for k in KEYS}})
    out = {"source": "tests/data/imrphenomxpnr_parityfix.json of the reference",
           "approximant": "IMRPhenomXPNR", "lalsimulation_version": dump["lalsimulation_version"],
           "lal_version": dump["lal_version"], "cases": cases}
    (Path(__file__).resolve().parent / "lal_xpnr_aligned.json").write_text(json.dumps(out) + "\n")


if __name__ == "__main__":
    main()
