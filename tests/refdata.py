"""Helpers that rebuild the reference unit tests' own setups (basis tables,
literal matrices) from ``tests/golden/reference_literals.json``."""
from __future__ import annotations

import json
from functools import lru_cache
from pathlib import Path

import numpy as np

from tblite_amd import params, structure

GOLDEN = Path(__file__).with_name("golden")


@lru_cache(maxsize=None)
def literals():
    return json.loads((GOLDEN / "reference_literals.json").read_text())


def literal_matrix(kind: str, name: str) -> np.ndarray:
    rec = literals()[kind][name]
    n = rec["nao"]
    return np.array(rec["fortran_order"]).reshape((n, n), order="F")


def test_basis_cgto(table: str, nums, ng: int = 6):
    """``make_basis`` of the reference unit tests (e.g.
    ``test/unit/test_integral_overlap.f90:90-210``): STO-``ng`` for every shell
    with the test's own (nsh, lsh, pqn, zeta) tables."""
    t = literals()[table]
    msh = t["msh"]
    out, nsh_id = [], []
    for z in nums:
        shells = []
        for ish in range(t["nsh"][z - 1]):
            k = ish + msh * (z - 1)
            a, c, info = params.slater_to_gauss(ng, t["pqn"][k], t["lsh"][k], t["zeta"][k], True)
            assert info == 0
            shells.append(params.CGTO(t["lsh"][k], a, c))
        out.append(shells)
        nsh_id.append(len(shells))
    return out, np.array(nsh_id, dtype=np.int32)


def custom_setup(numbers, xyz, cgto=None, nsh_id=None, rad=None):
    """``Setup`` for ``numbers``/``xyz`` with GFN2 H0 parameters and optionally
    a replacement basis (same shell count/angular momenta as GFN2) and radii."""
    mol = structure.Structure(numbers, xyz)
    spec = params.gfn2_species(mol.nums)
    if cgto is not None:
        assert list(nsh_id) == list(spec.nsh_id)
        spec.basis_cgto = cgto
    if rad is not None:
        spec.rad = np.asarray(rad, dtype=np.float64)
    return structure.setup(mol, spec=spec)
