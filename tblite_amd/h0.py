"""Host-side mirror of tblite's ``tblite_xtb_h0`` procedures for this path,
running on the gfx950 kernels through the C-ABI.

Same names, argument meaning and layout as the Fortran
(``src/tblite/xtb/h0.f90``): Fortran-ordered numpy arrays, ``intent(out)``
integrals, ``intent(inout)`` (accumulated) gradient outputs.  The structure,
basis, H0 parameters and neighbour list are carried by a ``Setup``
(``tblite_amd.structure.setup``).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import lib as L
from .structure import Setup


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


class _Args:
    """Keeps the numpy arrays alive while ctypes structures point into them."""

    def __init__(self, setup: Setup, with_list: bool = True):
        mol, bas, spec = setup.mol, setup.bas, setup.spec
        self.keep = []

        def i32(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            self.keep.append(a)
            return a

        def f64(a):
            a = np.ascontiguousarray(a, dtype=np.float64)
            self.keep.append(a)
            return a

        msh = setup.msh
        self.basis = L.Basis(mol.nid, msh, _p(i32(spec.nsh_id)), _p(i32(setup.ang)),
                             _p(i32(setup.nprim)), _p(f64(setup.alpha)), _p(f64(setup.coeff)),
                             bas.intcut)
        kd = [None] * 3
        if spec.ksig is not None:  # diatomic-frame scaling tables (nid, nid)
            kd = [_p(f64(np.asarray(k, dtype=np.float64).reshape(-1, order="F")))
                  for k in (spec.ksig, spec.kpi, spec.kdel)]
        self.h0 = L.H0Par(_p(f64(spec.hscale.reshape(-1, order="F"))),
                          _p(f64(spec.shpoly.reshape(-1, order="F"))), _p(f64(spec.rad)), *kd)
        trans = f64(setup.trans)
        self.geom = L.Geom(mol.nat, _p(i32(mol.id)), _p(f64(mol.xyz)), len(setup.trans),
                           _p(trans), setup.cutoff)
        self.list = None
        if with_list:
            self.list = L.List(_p(i32(setup.inl)), _p(i32(setup.nnl)), len(setup.nlat),
                               _p(i32(setup.nlat)), _p(i32(setup.nltr)))
        self.ish_at = i32(bas.ish_at)
        self.iao_sh = i32(bas.iao_sh)
        self.nao = bas.nao


def get_selfenergy(setup: Setup, cn=None):
    """``get_selfenergy`` (``xtb/h0.f90:113-217``): returns ``(selfenergy, dsedcn)``."""
    spec, bas, mol = setup.spec, setup.bas, setup.mol
    se = np.zeros(bas.nsh)
    dse = np.zeros(bas.nsh)
    id_ = np.ascontiguousarray(mol.id, dtype=np.int32)
    ish = np.ascontiguousarray(bas.ish_at, dtype=np.int32)
    nsh = np.ascontiguousarray(spec.nsh_id, dtype=np.int32)
    sesp = np.ascontiguousarray(spec.selfenergy.reshape(-1, order="F"))
    kcn = np.ascontiguousarray(spec.kcn.reshape(-1, order="F"))
    cnp = None if cn is None else np.ascontiguousarray(cn, dtype=np.float64)
    L.check(L.lib().tbx_get_selfenergy(mol.nat, _p(id_), _p(ish), _p(nsh), setup.msh, _p(sesp),
                                       _p(kcn), None if cnp is None else _p(cnp), _p(se), _p(dse)))
    return se, dse


def get_hamiltonian(setup: Setup, selfenergy, device: int = 0):
    """``get_hamiltonian`` (``xtb/h0.f90:220-429``).

    Returns Fortran-ordered ``overlap(nao,nao)``, ``dpint(3,nao,nao)``,
    ``qpint(6,nao,nao)`` and ``hamiltonian(nao,nao)``."""
    a = _Args(setup)
    nao = a.nao
    S = np.zeros((nao, nao), order="F")
    D = np.zeros((3, nao, nao), order="F")
    Q = np.zeros((6, nao, nao), order="F")
    H = np.zeros((nao, nao), order="F")
    se = np.ascontiguousarray(selfenergy, dtype=np.float64)
    L.check(L.lib().tbx_get_hamiltonian(L.context(device), ct.byref(a.geom), ct.byref(a.list),
                                        ct.byref(a.basis), ct.byref(a.h0), _p(a.ish_at),
                                        _p(a.iao_sh), nao, _p(se), _p(S), _p(D), _p(Q), _p(H)))
    return S, D, Q, H


def get_multipole_integrals(setup: Setup, device: int = 0):
    """``get_multipole_integrals`` (``integral/native/integrals.f90:1394-1465``,
    the full double loop over the pairs within the cutoff).

    Returns Fortran-ordered ``overlap(nao,nao)``, ``dpint(3,nao,nao)`` and
    ``qpint(6,nao,nao)``."""
    a = _Args(setup)
    nao = a.nao
    S = np.zeros((nao, nao), order="F")
    D = np.zeros((3, nao, nao), order="F")
    Q = np.zeros((6, nao, nao), order="F")
    L.check(L.lib().tbx_get_multipole_integrals(L.context(device), ct.byref(a.geom),
                                                ct.byref(a.basis), _p(a.ish_at), _p(a.iao_sh),
                                                nao, _p(S), _p(D), _p(Q)))
    return S, D, Q


def get_dipole_integrals(setup: Setup, device: int = 0):
    """``get_dipole_integrals`` (``integral/native/integrals.f90:994-1059``):
    ``overlap(nao,nao)`` and ``dpint(3,nao,nao)``."""
    a = _Args(setup)
    nao = a.nao
    S = np.zeros((nao, nao), order="F")
    D = np.zeros((3, nao, nao), order="F")
    L.check(L.lib().tbx_get_dipole_integrals(L.context(device), ct.byref(a.geom),
                                             ct.byref(a.basis), _p(a.ish_at), _p(a.iao_sh), nao,
                                             _p(S), _p(D)))
    return S, D


def get_overlap(setup: Setup, device: int = 0):
    """``get_overlap`` (``integral/native/integrals.f90:411-470``):
    ``overlap(nao,nao)``."""
    a = _Args(setup)
    nao = a.nao
    S = np.zeros((nao, nao), order="F")
    L.check(L.lib().tbx_get_overlap(L.context(device), ct.byref(a.geom), ct.byref(a.basis),
                                    _p(a.ish_at), _p(a.iao_sh), nao, _p(S)))
    return S


def get_hamiltonian_gradient(setup: Setup, selfenergy, dsedcn, vao, vdp, vqp, pmat, xmat,
                             dEdcn, gradient, sigma, device: int = 0):
    """``get_hamiltonian_gradient`` (``xtb/h0.f90:432-666``).

    ``pmat``/``xmat`` ``(nao,nao,nspin)``, ``vao`` ``(nao,nspin)``, ``vdp``
    ``(3,nat,nspin)``, ``vqp`` ``(6,nat,nspin)``; ``dEdcn(nat)``,
    ``gradient(3,nat)`` and ``sigma(3,3)`` are accumulated in place."""
    a = _Args(setup)
    nao, nat = a.nao, setup.mol.nat
    pm = np.asfortranarray(np.asarray(pmat, dtype=np.float64).reshape(nao, nao, -1, order="F"))
    nspin = pm.shape[2]
    xm = np.asfortranarray(np.asarray(xmat, dtype=np.float64).reshape(nao, nao, -1, order="F"))
    va = np.asfortranarray(np.asarray(vao, dtype=np.float64).reshape(nao, -1, order="F"))
    vd = np.asfortranarray(np.asarray(vdp, dtype=np.float64).reshape(3, nat, -1, order="F"))
    vq = np.asfortranarray(np.asarray(vqp, dtype=np.float64).reshape(6, nat, -1, order="F"))
    for arr, shape in ((dEdcn, (nat,)), (gradient, (3, nat)), (sigma, (3, 3))):
        if arr.shape != shape or arr.dtype != np.float64 or not arr.flags.f_contiguous:
            raise ValueError(f"output must be a Fortran-ordered float64 array of shape {shape}")
    se = np.ascontiguousarray(selfenergy, dtype=np.float64)
    dse = np.ascontiguousarray(dsedcn, dtype=np.float64)
    L.check(L.lib().tbx_get_hamiltonian_gradient(
        L.context(device), ct.byref(a.geom), ct.byref(a.list), ct.byref(a.basis), ct.byref(a.h0),
        _p(a.ish_at), _p(a.iao_sh), nao, _p(se), _p(dse), nspin, _p(va), _p(vd), _p(vq), _p(pm),
        _p(xm), _p(dEdcn), _p(gradient), _p(sigma)))
    return dEdcn, gradient, sigma


def get_hamiltonian_adjoint(setup: Setup, selfenergy, sbar, hbar, dbar, qbar, gradient, dEdse,
                            sigma=None, device: int = 0):
    """Adjoint (backward) of ``get_hamiltonian`` for one structure: with
    ``E = <sbar,S> + <hbar,H> + <dbar,D> + <qbar,Q>`` (adjoints Fortran-ordered
    ``(nao,nao)``, ``(nao,nao)``, ``(3,nao,nao)``, ``(6,nao,nao)``),
    accumulates dE/dxyz into ``gradient(3,nat)``, dE/dselfenergy into
    ``dEdse(nsh)`` and, if given, the strain derivative into ``sigma(3,3)``.
    The integrals are recomputed on the device (no derivative tensors)."""
    a = _Args(setup)
    nao, nat, nsh = a.nao, setup.mol.nat, setup.bas.nsh

    def fa(x, shape):
        return np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(shape, order="F"))

    sb, hb = fa(sbar, (nao, nao)), fa(hbar, (nao, nao))
    db, qb = fa(dbar, (3, nao, nao)), fa(qbar, (6, nao, nao))
    outs = [(gradient, (3, nat)), (dEdse, (nsh,))]
    if sigma is not None:
        outs.append((sigma, (3, 3)))
    for arr, shape in outs:
        if arr.shape != shape or arr.dtype != np.float64 or not arr.flags.f_contiguous:
            raise ValueError(f"output must be a Fortran-ordered float64 array of shape {shape}")
    se = np.ascontiguousarray(selfenergy, dtype=np.float64)
    L.check(L.lib().tbx_get_hamiltonian_adjoint(
        L.context(device), ct.byref(a.geom), ct.byref(a.list), ct.byref(a.basis), ct.byref(a.h0),
        _p(a.ish_at), _p(a.iao_sh), nao, _p(se), _p(sb), _p(hb), _p(db), _p(qb), _p(gradient),
        _p(dEdse), None if sigma is None else _p(sigma)))
    return gradient, dEdse, sigma
