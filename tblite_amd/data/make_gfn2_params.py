"""Regenerate ``gfn2.json`` from tblite's exported GFN2-xTB parameter file.

tblite writes its built-in GFN2 tables (``src/tblite/xtb/gfn2.f90:196-468``)
in the TOML schema of ``doc/spec/parameter.rst``; the reference keeps such an
export at ``test/cli/test/gfn2-molmom.toml``.  Only the entries the
zeroth-order Hamiltonian needs are carried over, unit conversions are left
to the loader (``tblite_amd/params.py``), exactly as the reference's loader
applies them (``src/tblite/param/element.f90:394-412, 501-519``):

- ``levels`` and ``kcn`` stay in eV;
- ``shpoly`` is stored already scaled by 0.01;
- ``en`` is the Pauling electronegativity (mctc-lib ``get_pauling_en``).

Usage:  python make_gfn2_params.py /path/to/gfn2-molmom.toml
"""
import json
import sys
from pathlib import Path

import tomli

LETTER = {"s": 0, "p": 1, "d": 2, "f": 3, "g": 4}


def main(path):
    data = tomli.loads(Path(path).read_text())
    ham = data["hamiltonian"]["xtb"]
    shell = dict(ham["shell"])
    elements = []
    for z, (sym, rec) in enumerate(data["element"].items(), start=1):
        shells = [(int(s[:-1]), LETTER[s[-1]]) for s in rec["shells"]]
        elements.append({
            "z": z, "sym": sym,
            "pqn": [n for n, _ in shells], "ang": [l for _, l in shells],
            "ngauss": rec["ngauss"], "slater": rec["slater"],
            "levels_ev": rec["levels"], "kcn_ev": rec.get("kcn", [0.0] * len(shells)),
            "shpoly": rec.get("shpoly", [0.0] * len(shells)),
            "refocc": rec["refocc"], "en": rec["en"],
        })
    out = {
        "name": data["meta"]["name"],
        "reference": data["meta"]["reference"],
        "wexp": ham["wexp"], "enscale": ham["enscale"], "cn": ham["cn"],
        "shell": shell,
        "elements": elements,
    }
    dst = Path(__file__).with_name("gfn2.json")
    dst.write_text(json.dumps(out, indent=1))
    print("wrote", dst, len(elements), "elements")


if __name__ == "__main__":
    main(sys.argv[1])
