"""Regenerate ``sto_ng.json``: the STO-nG expansion constants of R. F. Stewart,
J. Chem. Phys. 52, 431-438 (1970), Tables I-V.

The constants are published data.  The copy that tblite ships
(``src/tblite/basis/slater.f90:54-380``) is read here once so the numbers are
identical to the digit; the JSON layout is this package's own:

    {"<ng>": {"<n><l>": {"alpha": [...], "coeff": [...]}}}

with ``l`` given as a letter (s, p, d, f, g).  Exponents are for zeta = 1; the
expansion multiplies them by zeta**2 (``slater.f90:455-487``).

Usage:  python make_sto_ng.py /path/to/reference/src/tblite/basis/slater.f90
"""
import json
import re
import sys
from pathlib import Path

# row order of the 15 (n, l) function types in the published tables
TYPES = ["1s", "2s", "3s", "4s", "5s", "2p", "3p", "4p", "5p",
         "3d", "4d", "5d", "4f", "5f", "5g"]


def _arrays(text):
    out = {}
    pat = re.compile(r"real\(wp\), parameter :: (p\w+)\(?[^=]*=\s*(?:reshape\()?\[(.*?)\]",
                     re.S)
    for name, body in pat.findall(text):
        body = re.sub(r"!.*", "", body)
        nums = re.findall(r"[-+]?\d+\.\d+e[-+]?\d+", body)
        out[name] = [float(x) for x in nums]
    return out


def main(path):
    arrays = _arrays(Path(path).read_text())
    table = {}
    table["1"] = {t: {"alpha": [arrays["pAlpha1"][k]], "coeff": [1.0]}
                  for k, t in enumerate(TYPES)}
    for ng in range(2, 7):
        a = arrays[f"pAlpha{ng}"]
        c = arrays[f"pCoeff{ng}"]
        assert len(a) == ng * len(TYPES) and len(c) == len(a), (ng, len(a), len(c))
        table[str(ng)] = {t: {"alpha": a[k * ng:(k + 1) * ng], "coeff": c[k * ng:(k + 1) * ng]}
                          for k, t in enumerate(TYPES)}
    table["6"]["6s"] = {"alpha": arrays["pAlpha6s"], "coeff": arrays["pCoeff6s"]}
    table["6"]["6p"] = {"alpha": arrays["pAlpha6p"], "coeff": arrays["pCoeff6p"]}
    dst = Path(__file__).with_name("sto_ng.json")
    dst.write_text(json.dumps({"source": "R. F. Stewart, J. Chem. Phys. 52, 431 (1970), Tables I-V",
                               "table": table}, indent=1))
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1])
