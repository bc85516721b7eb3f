"""Write ``reference_literals.json``: golden data held by the reference's own
unit tests for this path, as plain numbers.

Sources (tblite 0.7.0, read as text from the reference tree):

- ``test/unit/test_integral_overlap.f90``: the test basis tables of
  ``make_basis`` (:90-210), the literal STO-6G overlap matrices of
  ``test_overlap_{alh3,bh3,beh2,ch4,cl2,f2,h2,lih}`` (:273-823) and the
  diatomic-frame scaled overlaps of ``test_overlap_diat_{alh3,cl2}``
  (:337-399, :622-744; ksig = kpi = kdel = 1.2, :231-233);
- ``test/unit/test_hamiltonian.f90``: the test basis tables (:60-110) and the
  literal GFN2 H0 matrices of ``test_hamiltonian_{h2,lih,s2,sih4}`` (:510-745).

The reference obtains the geometries from mstore, which is not part of the
reference tree, so only the element sequence is recorded here (heavy atom
first, as the block structure of every matrix shows); the pin tests recover
the geometry by fitting it.

Usage:  python extract_reference_fixtures.py /path/to/reference
"""
import json
import re
import sys
from pathlib import Path

OVERLAP_CASES = {"alh3": [13, 1, 1, 1], "bh3": [5, 1, 1, 1], "beh2": [4, 1, 1],
                 "ch4": [6, 1, 1, 1, 1], "cl2": [17, 17], "f2": [9, 9], "h2": [1, 1],
                 "lih": [3, 1]}
HAMILTONIAN_CASES = {"h2": [1, 1], "lih": [3, 1], "s2": [16, 16], "sih4": [14, 1, 1, 1, 1]}
# test_overlap_diat_<name>: literal variable name (CeCl3 has f functions: l > 2, left out)
OVERLAP_DIAT_CASES = {"alh3": "overlap_diat", "cl2": "overlap"}

NUM = r"[-+]?\d+\.\d+E[-+]?\d+"


def _numbers(body):
    body = re.sub(r"!.*", "", body)
    return [float(x) for x in re.findall(NUM, body.replace("_wp", ""), flags=re.I)]


def _ints(body):
    body = re.sub(r"!.*", "", body)
    return [int(x) for x in re.findall(r"(?<![\d.])\d+(?![\d.])", body)]


def _param_block(text, name):
    m = re.search(rf"parameter :: {name}\([^)]*\) = (?:reshape\()?\[(.*?)\]", text, re.S)
    return m.group(1)


def _basis_tables(text):
    nsh = _ints(_param_block(text, "nsh"))
    lsh = _ints(_param_block(text, "lsh"))
    pqn = _ints(_param_block(text, "pqn"))
    zeta = [float(x) for x in re.findall(r"\d+\.\d+", re.sub(r"!.*", "",
                                                               _param_block(text, "zeta")))]
    nz = len(nsh)
    msh = len(lsh) // nz
    return {"nsh": nsh, "msh": msh, "lsh": lsh, "pqn": pqn, "zeta": zeta}


def _literal(text, sub, var):
    m = re.search(rf"subroutine {sub}\(error\)(.*?)end subroutine {sub}", text, re.S)
    body = m.group(1)
    nao = int(re.search(r"nao = (\d+)", body).group(1))
    vals = _numbers(re.search(rf"{var}\(nao, nao\) = reshape\(\[(.*?)\]", body, re.S).group(1))
    assert len(vals) == nao * nao, (sub, len(vals), nao)
    return nao, vals


def main(root):
    root = Path(root)
    ov = (root / "test/unit/test_integral_overlap.f90").read_text()
    ham = (root / "test/unit/test_hamiltonian.f90").read_text()
    out = {"source": "tblite 0.7.0 test/unit literals", "overlap": {}, "hamiltonian": {},
           "basis_overlap": _basis_tables(ov), "basis_hamiltonian": _basis_tables(ham)}
    for name, nums in OVERLAP_CASES.items():
        nao, vals = _literal(ov, f"test_overlap_{name}", "overlap")
        out["overlap"][name] = {"numbers": nums, "nao": nao, "fortran_order": vals,
                                "thr": "1e5*epsilon (test_integral_overlap.f90:37)"}
    out["overlap_diat"] = {}
    for name, var in OVERLAP_DIAT_CASES.items():
        nao, vals = _literal(ov, f"test_overlap_diat_{name}", var)
        out["overlap_diat"][name] = {"numbers": OVERLAP_CASES[name], "nao": nao,
                                     "fortran_order": vals, "ksig": 1.2, "kpi": 1.2, "kdel": 1.2,
                                     "thr": "1e6*epsilon (thr1*10, test_integral_overlap.f90:251)"}
    for name, nums in HAMILTONIAN_CASES.items():
        nao, vals = _literal(ham, f"test_hamiltonian_{name}", "hamiltonian")
        out["hamiltonian"][name] = {"numbers": nums, "nao": nao, "fortran_order": vals,
                                    "thr": "1e2*sqrt(epsilon) (test_hamiltonian.f90:40)"}
    dst = Path(__file__).with_name("reference_literals.json")
    dst.write_text(json.dumps(out, indent=1))
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1])
