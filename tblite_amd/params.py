"""GFN2-xTB zeroth-order-Hamiltonian parameters and the STO-nG basis.

Host-side setup that precedes the integral path; none of it is timed with the
kernels.  Each function names the tblite routine whose behaviour it keeps.

Data sources (both generated once by the scripts beside them in ``data/``):

- ``data/gfn2.json``: tblite's exported GFN2 parameter file; units converted
  here as ``src/tblite/param/element.f90:394-412, 501-519`` does (eV -> Eh with
  the older 27.21138505 factor, ``xtb/gfn2.f90:51``).
- ``data/sto_ng.json``: the Stewart (1970) STO-nG constants
  (``src/tblite/basis/slater.f90:54-380``).

Atomic radii for the shell polynomial (``h0%rad``, ``xtb/spec.f90:176-187``)
come from mctc-lib's ``get_atomic_rad``, which is not part of the reference
tree; the values below are the published table that function returns
(Pyykko-style atomic radii in Angstrom, converted to bohr).  Their values are
pinned only where the reference's literal H0 matrices fix them (H, Li, S, Si
through the fits in ``tests/test_oracle_pins.py``); the kernels take ``rad``
as an input either way.
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from functools import lru_cache
from pathlib import Path

import numpy as np

DATA = Path(__file__).with_name("data")

EV_TO_AU = 1.0 / 27.21138505          # xtb/gfn2.f90:51, param/element.f90:410,517
AUTOAA = 0.52917721067                # mctc-lib bohr radius (CODATA 2014)
AATOAU = 1.0 / AUTOAA
MAXG = 12                             # basis/type.f90:32
TOP = 2.0 / math.pi
ELEMENT_SYMBOLS = [
    "H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al", "Si", "P", "S",
    "Cl", "Ar", "K", "Ca", "Sc", "Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga",
    "Ge", "As", "Se", "Br", "Kr", "Rb", "Sr", "Y", "Zr", "Nb", "Mo", "Tc", "Ru", "Rh", "Pd",
    "Ag", "Cd", "In", "Sn", "Sb", "Te", "I", "Xe", "Cs", "Ba", "La", "Ce", "Pr", "Nd", "Pm",
    "Sm", "Eu", "Gd", "Tb", "Dy", "Ho", "Er", "Tm", "Yb", "Lu", "Hf", "Ta", "W", "Re", "Os",
    "Ir", "Pt", "Au", "Hg", "Tl", "Pb", "Bi", "Po", "At", "Rn",
]

# mctc-lib atomic radii (Angstrom), Z = 1..86
_ATOMIC_RAD_AA = [
    0.32, 0.37, 1.30, 0.99, 0.84, 0.75, 0.71, 0.64, 0.60, 0.62, 1.60, 1.40, 1.24, 1.14, 1.09,
    1.04, 1.00, 1.01, 2.00, 1.74, 1.59, 1.48, 1.44, 1.30, 1.29, 1.24, 1.18, 1.17, 1.22, 1.20,
    1.23, 1.20, 1.20, 1.18, 1.17, 1.16, 2.15, 1.90, 1.76, 1.64, 1.56, 1.46, 1.38, 1.36, 1.34,
    1.30, 1.36, 1.40, 1.42, 1.40, 1.40, 1.37, 1.36, 1.36, 2.38, 2.06, 1.94, 1.84, 1.90, 1.88,
    1.86, 1.85, 1.83, 1.82, 1.81, 1.80, 1.79, 1.77, 1.77, 1.78, 1.74, 1.64, 1.58, 1.50, 1.41,
    1.36, 1.32, 1.30, 1.30, 1.32, 1.44, 1.45, 1.50, 1.42, 1.48, 1.46,
]


def atomic_rad(z: int) -> float:
    """Atomic radius in bohr (mctc-lib ``get_atomic_rad``)."""
    return _ATOMIC_RAD_AA[z - 1] * AATOAU


# Covalent radii of the coordination number (Angstrom), Z = 1..86: the
# DFT-D3 table, i.e. the single-bond radii of P. Pyykko, M. Atsumi, Chem. Eur.
# J. 15, 186 (2009) with the values of metals reduced.  mctc-lib's
# get_covalent_rad (the radii of the GFN2 coordination number,
# xtb/gfn2.f90:640) returns 4/3 of these in bohr.  mctc-lib is not part of the
# reference tree; the H, Li, S and Si entries are fixed by the reference's
# literal H0 matrices (test/unit/test_hamiltonian.f90), whose CN comes from it.
_COVALENT_RAD_D3_AA = [
    0.32, 0.46, 1.20, 0.94, 0.77, 0.75, 0.71, 0.63, 0.64, 0.67, 1.40, 1.25, 1.13, 1.04, 1.10,
    1.02, 0.99, 0.96, 1.76, 1.54, 1.33, 1.22, 1.21, 1.10, 1.07, 1.04, 1.00, 0.99, 1.01, 1.09,
    1.12, 1.09, 1.15, 1.10, 1.14, 1.17, 1.89, 1.67, 1.47, 1.39, 1.32, 1.24, 1.15, 1.13, 1.13,
    1.08, 1.15, 1.23, 1.28, 1.26, 1.26, 1.23, 1.32, 1.31, 2.09, 1.76, 1.62, 1.47, 1.58, 1.57,
    1.56, 1.55, 1.51, 1.52, 1.51, 1.50, 1.49, 1.49, 1.48, 1.53, 1.46, 1.37, 1.31, 1.23, 1.18,
    1.16, 1.11, 1.12, 1.13, 1.32, 1.30, 1.30, 1.36, 1.31, 1.38, 1.42,
]

# GFN2 double-exponential coordination number (mctc-lib ncoord "dexp"):
# steepness of the two counting functions, offset of the second (bohr) and
# the real-space cutoff (bohr, xtb/singlepoint.f90:55)
CN_KA, CN_KB, CN_RSHIFT, CN_CUTOFF = 10.0, 20.0, 2.0, 25.0


def covalent_rad(z: int) -> float:
    """Covalent radius of the GFN2 coordination number in bohr (4/3 of the
    D3 covalent radius, mctc-lib ``get_covalent_rad``)."""
    return 4.0 / 3.0 * _COVALENT_RAD_D3_AA[z - 1] * AATOAU


@lru_cache(maxsize=None)
def _sto_table():
    return json.loads((DATA / "sto_ng.json").read_text())["table"]


@lru_cache(maxsize=None)
def gfn2_table():
    return json.loads((DATA / "gfn2.json").read_text())


_LETTER = "spdfg"
# (2l-1)!! for l = 0..7 (the reference indexes its 1-based copy with l+1)
_DFACT = [1.0, 1.0, 3.0, 15.0, 105.0, 945.0, 10395.0, 135135.0]


def slater_to_gauss(ng: int, n: int, l: int, zeta: float, norm: bool = True):
    """STO-nG expansion of a Slater function (``basis/slater.f90:409-511``).

    Returns ``(alpha, coeff, info)``; ``info`` is 0 on success or the position
    of the offending argument, as in the reference: 1 = ng, 2 = n/l, 3 = l,
    4 = zeta.
    """
    if n > 5 or n <= l:
        if not (n == 6 and ng == 6):
            return None, None, 2
    if zeta <= 0.0:
        return None, None, 4
    if l > 4:
        return None, None, 3
    if ng < 1 or ng > 6:
        return None, None, 1
    key = f"{n}{_LETTER[l]}"
    tab = _sto_table()[str(ng)]
    if key not in tab:
        return None, None, 2
    alpha = np.array(tab[key]["alpha"], dtype=np.float64) * zeta ** 2
    coeff = np.array(tab[key]["coeff"], dtype=np.float64)
    if norm:
        coeff = coeff * (TOP * alpha) ** 0.75 * np.sqrt(4 * alpha) ** l / math.sqrt(_DFACT[l])
    return alpha, coeff, 0


@dataclass
class CGTO:
    """``cgto_type`` (``basis/type.f90:43-55``)."""
    ang: int
    alpha: np.ndarray
    coeff: np.ndarray

    @property
    def nprim(self) -> int:
        return len(self.alpha)


def integral_cutoff(acc: float = 1.0) -> float:
    """``basis/type.f90:250-260``.  The reference clips ``acc`` with its bounds
    swapped, so the result is 25 for every accuracy; kept as is."""
    min_intcut, max_intcut, max_acc, min_acc = 5.0, 25.0, 1.0e-4, 1.0e3
    clip = lambda v, lo, hi: min(max(v, lo), hi)  # noqa: E731
    return clip(max_intcut - 10 * math.log10(clip(acc, min_acc, max_acc)), min_intcut, max_intcut)


@dataclass
class Basis:
    """``basis_type`` + ``new_basis`` (``basis/type.f90:58-192``) for one molecule.

    ``cgto[isp][ish]``; all index maps 0-based.
    """
    nsh_id: np.ndarray
    cgto: list
    id: np.ndarray
    nsh: int = 0
    nao: int = 0
    maxl: int = 0
    intcut: float = 25.0
    min_alpha: float = math.inf
    ish_at: np.ndarray = field(default=None)
    iao_sh: np.ndarray = field(default=None)
    nao_sh: np.ndarray = field(default=None)
    ao2at: np.ndarray = field(default=None)
    sh2at: np.ndarray = field(default=None)

    @classmethod
    def new(cls, id_, nsh_id, cgto, acc: float = 1.0):
        id_ = np.asarray(id_, dtype=np.int32)
        nsh_id = np.asarray(nsh_id, dtype=np.int32)
        self = cls(nsh_id=nsh_id, cgto=cgto, id=id_)
        self.intcut = integral_cutoff(acc)
        nsh_at = nsh_id[id_]
        self.nsh = int(nsh_at.sum())
        self.ish_at = np.concatenate([[0], np.cumsum(nsh_at)[:-1]]).astype(np.int32)
        self.sh2at = np.repeat(np.arange(len(id_), dtype=np.int32), nsh_at)
        nao_sh = []
        for iat, isp in enumerate(id_):
            for ish in range(nsh_id[isp]):
                nao_sh.append(2 * cgto[isp][ish].ang + 1)
        self.nao_sh = np.asarray(nao_sh, dtype=np.int32)
        self.nao = int(self.nao_sh.sum())
        self.iao_sh = np.concatenate([[0], np.cumsum(self.nao_sh)[:-1]]).astype(np.int32)
        self.ao2at = np.repeat(self.sh2at, self.nao_sh).astype(np.int32)
        self.maxl = max(c.ang for sp in cgto for c in sp)
        self.min_alpha = min(float(np.min(c.alpha)) for sp in cgto for c in sp)
        return self

    @property
    def nid(self) -> int:
        return len(self.nsh_id)

    @property
    def msh(self) -> int:
        return int(np.max(self.nsh_id))

    def get_cutoff(self, acc: float | None = None) -> float:
        """``basis/type.f90:227-246``."""
        intcut = self.intcut if acc is None else integral_cutoff(acc)
        return min(math.sqrt(2.0 * intcut / self.min_alpha), 40.0)


def _kshell(k: int, l: int, shell: dict) -> float:
    """``kshell`` (``xtb/gfn2.f90:1024-1028``) from the exported shell table;
    a missing off-diagonal key is the average of the diagonal ones
    (``param/hamiltonian.f90:198-199``)."""
    key = "spdfg"[min(k, l)] + "spdfg"[max(k, l)]
    if key in shell:
        return shell[key]
    return 0.5 * (shell["spdfg"[k] * 2] + shell["spdfg"[l] * 2])


@dataclass
class GFN2Species:
    """Per-species slice of the GFN2 tables, indexed by the molecule's
    species list (``mol%num`` -> ``id``)."""
    nums: list
    basis_cgto: list
    nsh_id: np.ndarray
    selfenergy: np.ndarray   # (msh, nid), Eh
    kcn: np.ndarray          # (msh, nid), Eh
    shpoly: np.ndarray       # (msh, nid)
    refocc: np.ndarray       # (msh, nid)
    hscale: np.ndarray       # (msh, msh, nid, nid) Fortran order (jsh, ish, jsp, isp)
    rad: np.ndarray          # (nid,), bohr
    slater: np.ndarray       # (msh, nid)
    # diatomic-frame scaling of the overlap in H0 (h0%ksig/kpi/kdel(izp, jzp),
    # xtb/h0.f90:64-65, 107-108): (nid, nid) each, or None (GFN2: off)
    ksig: np.ndarray | None = None
    kpi: np.ndarray | None = None
    kdel: np.ndarray | None = None


def gfn2_species(nums) -> GFN2Species:
    """GFN2 basis and H0 tables for the species ``nums`` (atomic numbers).

    Mirrors ``add_basis`` (``xtb/gfn2.f90:607-630``) and the ``gfn2_h0spec``
    generators (``xtb/gfn2.f90:739-889``)."""
    tab = gfn2_table()
    els = tab["elements"]
    nid = len(nums)
    recs = [els[z - 1] for z in nums]
    if any(z < 1 or z > 86 for z in nums):
        raise ValueError("No support for elements with Z > 86.")
    nsh_id = np.array([len(r["ang"]) for r in recs], dtype=np.int32)
    msh = int(nsh_id.max())
    cg = []
    for r in recs:
        shells = []
        for ish in range(len(r["ang"])):
            a, c, info = slater_to_gauss(r["ngauss"][ish], r["pqn"][ish], r["ang"][ish],
                                         r["slater"][ish], True)
            if info != 0:
                raise ValueError(f"slater_to_gauss failed for {r['sym']} shell {ish}: {info}")
            shells.append(CGTO(r["ang"][ish], a, c))
        cg.append(shells)
    se = np.zeros((msh, nid))
    kcn = np.zeros((msh, nid))
    shp = np.zeros((msh, nid))
    ref = np.zeros((msh, nid))
    zet = np.zeros((msh, nid))
    for isp, r in enumerate(recs):
        for ish in range(len(r["ang"])):
            se[ish, isp] = r["levels_ev"][ish] * EV_TO_AU
            kcn[ish, isp] = r["kcn_ev"][ish] * EV_TO_AU
            shp[ish, isp] = r["shpoly"][ish]
            ref[ish, isp] = r["refocc"][ish]
            zet[ish, isp] = r["slater"][ish]
    hs = np.zeros((msh, msh, nid, nid))
    for isp, ri in enumerate(recs):
        for jsp, rj in enumerate(recs):
            den = (ri["en"] - rj["en"]) ** 2
            for ish, il in enumerate(ri["ang"]):
                for jsh, jl in enumerate(rj["ang"]):
                    zi, zj = ri["slater"][ish], rj["slater"][jsh]
                    zij = (2 * math.sqrt(zi * zj) / (zi + zj)) ** tab["wexp"]
                    enp = 1.0 + tab["enscale"] * den
                    km = 1.0 * _kshell(jl, il, tab["shell"]) * enp
                    hs[jsh, ish, jsp, isp] = zij * km
    rad = np.array([atomic_rad(z) for z in nums])
    return GFN2Species(list(nums), cg, nsh_id, se, kcn, shp, ref, hs, rad, zet)


def species_of(numbers):
    """Map atomic numbers to (``nums``, ``id``) like mctc-lib's structure_type:
    species in order of first appearance, ``id`` 0-based."""
    nums, ids = [], []
    for z in numbers:
        z = int(z)
        if z not in nums:
            nums.append(z)
        ids.append(nums.index(z))
    return nums, np.asarray(ids, dtype=np.int32)
