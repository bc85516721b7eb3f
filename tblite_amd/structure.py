"""Molecular structure, neighbour list and the flattened tables the integral
kernels consume.

``Structure`` carries what the path reads from mctc-lib's ``structure_type``
(``nat``, ``nid``, ``id``, ``num``, ``xyz`` in bohr).  ``Setup`` gathers the
GFN2 basis, H0 parameters and half neighbour list for one structure, in the
0-based flat layout the C-ABI (``include/tbx.h``) documents.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import params

EPS = float(np.finfo(np.float64).eps)


@dataclass
class Structure:
    numbers: np.ndarray   # atomic numbers, (nat,)
    xyz: np.ndarray       # Cartesian coordinates in bohr, (nat, 3) row-major = Fortran xyz(3, nat)
    charge: float = 0.0

    def __post_init__(self):
        self.numbers = np.asarray(self.numbers, dtype=np.int32)
        self.xyz = np.ascontiguousarray(self.xyz, dtype=np.float64).reshape(-1, 3)
        self.nums, self.id = params.species_of(self.numbers)

    @property
    def nat(self) -> int:
        return len(self.numbers)

    @property
    def nid(self) -> int:
        return len(self.nums)


def get_lattice_points(periodic, lattice, rthr: float) -> np.ndarray:
    """``get_lattice_points_cutoff`` (``cutoff.f90:97-121``) through the C-ABI:
    the lattice points within ``rthr`` (bohr) of the cell whose lattice
    vectors are the rows of ``lattice`` (3, 3), origin first, shape (ntr, 3).
    No periodic direction: the single zero translation."""
    import ctypes as ct

    from . import lib as L
    per = np.ascontiguousarray([1 if p else 0 for p in periodic], dtype=np.int32)
    lat = np.ascontiguousarray(np.asarray(lattice, dtype=np.float64).reshape(3, 3))
    p = lambda a: a.ctypes.data_as(ct.c_void_p)  # noqa: E731
    n = ct.c_int(0)  # no room: the call reports the number of points
    st = L.lib().tbx_get_lattice_points(p(per), p(lat), float(rthr), ct.byref(n), None)
    if n.value <= 0:
        L.check(st)
    out = np.zeros((n.value, 3))
    L.check(L.lib().tbx_get_lattice_points(p(per), p(lat), float(rthr), ct.byref(n), p(out)))
    return out


def lattice_points_molecule() -> np.ndarray:
    """``get_lattice_points_cutoff`` for a non-periodic system (``cutoff.f90:97-121``):
    a single zero translation, shape (ntr, 3)."""
    return np.zeros((1, 3))


def new_adjacency_list(xyz: np.ndarray, trans: np.ndarray, cutoff: float):
    """Half neighbour list in CSR form (``adjlist.f90:96-143``, complete=.false.).

    For every atom ``iat`` the entries ``(jat, itr)`` with ``jat <= iat`` and
    ``eps <= r2 <= cutoff**2``, ordered by ``jat`` then ``itr``.
    Returns ``inl, nnl, nlat, nltr`` (0-based int32).
    """
    xyz = np.asarray(xyz, dtype=np.float64)
    trans = np.asarray(trans, dtype=np.float64).reshape(-1, 3)
    nat = len(xyz)
    cutoff2 = cutoff * cutoff
    inl = np.zeros(nat, dtype=np.int32)
    nnl = np.zeros(nat, dtype=np.int32)
    nlat, nltr = [], []
    img = 0
    for iat in range(nat):
        inl[iat] = img
        v = xyz[iat][None, None, :] - xyz[: iat + 1][:, None, :] - trans[None, :, :]
        r2 = (v[..., 0] ** 2 + v[..., 1] ** 2) + v[..., 2] ** 2
        keep = (r2 >= EPS) & (r2 <= cutoff2)
        jj, tt = np.nonzero(keep)
        nlat.append(jj.astype(np.int32))
        nltr.append(tt.astype(np.int32))
        img += len(jj)
        nnl[iat] = img - inl[iat]
    nlat = np.concatenate(nlat) if nlat else np.zeros(0, np.int32)
    nltr = np.concatenate(nltr) if nltr else np.zeros(0, np.int32)
    return inl, nnl, nlat.astype(np.int32), nltr.astype(np.int32)


@dataclass
class Setup:
    """Everything ``get_hamiltonian`` / ``get_hamiltonian_gradient`` read,
    flattened.  Index conventions (0-based):

    - cgto of shell ``ish`` of species ``isp``: slot ``ish + msh*isp`` in
      ``ang``, ``nprim``, ``alpha[12*slot:]``, ``coeff[12*slot:]``;
    - ``hscale`` in Fortran order ``(jsh, ish, jsp, isp)``;
    - ``shpoly`` ``(ish, isp)``, ``rad`` ``(isp)``, ``selfenergy`` per shell.
    """
    mol: Structure
    spec: params.GFN2Species
    bas: params.Basis
    trans: np.ndarray
    cutoff: float
    inl: np.ndarray
    nnl: np.ndarray
    nlat: np.ndarray
    nltr: np.ndarray
    ang: np.ndarray
    nprim: np.ndarray
    alpha: np.ndarray
    coeff: np.ndarray

    @property
    def msh(self) -> int:
        return int(self.spec.nsh_id.max())

    def selfenergy(self, cn=None):
        """``get_selfenergy`` (``xtb/h0.f90:113-217``), CN branch with derivative."""
        return get_selfenergy(self.spec, self.mol.id, self.bas.ish_at, self.spec.nsh_id, cn)


def get_selfenergy(spec: params.GFN2Species, id_, ish_at, nsh_id, cn=None):
    nsh = int(sum(nsh_id[i] for i in id_))
    se = np.zeros(nsh)
    dse = np.zeros(nsh)
    for iat, izp in enumerate(id_):
        ii = ish_at[iat]
        for ish in range(nsh_id[izp]):
            se[ii + ish] = spec.selfenergy[ish, izp]
            if cn is not None:
                se[ii + ish] -= spec.kcn[ish, izp] * cn[iat]
                dse[ii + ish] = -spec.kcn[ish, izp]
    return se, dse


def flat_cgto(spec: params.GFN2Species, msh: int):
    nid = len(spec.nums)
    ang = np.full(nid * msh, -1, dtype=np.int32)
    nprim = np.zeros(nid * msh, dtype=np.int32)
    alpha = np.zeros(nid * msh * params.MAXG)
    coeff = np.zeros(nid * msh * params.MAXG)
    for isp, shells in enumerate(spec.basis_cgto):
        for ish, c in enumerate(shells):
            s = ish + msh * isp
            ang[s] = c.ang
            nprim[s] = c.nprim
            alpha[params.MAXG * s: params.MAXG * s + c.nprim] = c.alpha
            coeff[params.MAXG * s: params.MAXG * s + c.nprim] = c.coeff
    return ang, nprim, alpha, coeff


def setup(mol: Structure, trans: np.ndarray | None = None, spec=None) -> Setup:
    """GFN2 setup for one structure: basis (``xtb/gfn2.f90:607-630``), H0
    spec, cutoff and half neighbour list (``xtb/singlepoint.f90:231-233``)."""
    if spec is None:
        spec = params.gfn2_species(mol.nums)
    bas = params.Basis.new(mol.id, spec.nsh_id, spec.basis_cgto, acc=1.0)
    if trans is None:
        trans = lattice_points_molecule()
    cutoff = bas.get_cutoff()
    inl, nnl, nlat, nltr = new_adjacency_list(mol.xyz, trans, cutoff)
    msh = int(spec.nsh_id.max())
    ang, nprim, alpha, coeff = flat_cgto(spec, msh)
    return Setup(mol, spec, bas, np.ascontiguousarray(trans, dtype=np.float64).reshape(-1, 3),
                 cutoff, inl, nnl, nlat, nltr, ang, nprim, alpha, coeff)
