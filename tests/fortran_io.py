"""Stream-file exchange with the Fortran host driver of the tbx_h0 binding
(tblite_amd/fortran/tbx_h0_driver.f90): the input carries one structure with
its flattened GFN2 tables, density, potential and output adjoints in the
driver's read order; the output the dense integrals, the accumulated gradient
terms and the adjoint's dE/dxyz, dE/dse and sigma."""
from __future__ import annotations

import numpy as np

from tblite_amd import lib as L

DRIVER = L.LIB_PATH.parent / "tbx_h0_driver"


def write_input(path, st, se, dse, vao, vdp, vqp, P, W, bars):
    mol, bas, spec = st.mol, st.bas, st.spec
    nat, nid, msh, nao, nsh = mol.nat, mol.nid, st.msh, bas.nao, bas.nsh
    nspin = P.shape[2]
    i32 = lambda a: np.ascontiguousarray(a, dtype="<i4").ravel(order="C")  # noqa: E731
    f64F = lambda a: np.asarray(a, dtype="<f8").ravel(order="F")  # noqa: E731
    parts = [
        i32([nat, nid, msh, nao, nsh, len(st.trans), nspin, len(st.nlat)]),
        i32(np.asarray(mol.id) + 1),
        np.ascontiguousarray(mol.xyz, dtype="<f8").ravel(),       # (nat,3) C = (3,nat) F
        np.ascontiguousarray(st.trans, dtype="<f8").ravel(),      # (ntrans,3) C = (3,ntrans) F
        np.array([st.cutoff], dtype="<f8"),
        i32(bas.ish_at), i32(bas.iao_sh),
        i32(spec.nsh_id), i32(st.ang), i32(st.nprim),
        np.ascontiguousarray(st.alpha, dtype="<f8"), np.ascontiguousarray(st.coeff, dtype="<f8"),
        f64F(spec.hscale), f64F(spec.shpoly), np.ascontiguousarray(spec.rad, dtype="<f8"),
        np.array([bas.intcut], dtype="<f8"),
        np.ascontiguousarray(se, dtype="<f8"), np.ascontiguousarray(dse, dtype="<f8"),
        f64F(vao), f64F(vdp), f64F(vqp), f64F(P), f64F(W),
        i32(st.inl), i32(st.nnl), i32(np.asarray(st.nlat) + 1), i32(np.asarray(st.nltr) + 1),
    ] + [f64F(b) for b in bars]
    with open(path, "wb") as f:
        for p in parts:
            f.write(p.tobytes())


def read_output(path, st):
    nat, nao = st.mol.nat, st.bas.nao
    raw = open(path, "rb").read()
    nsh = st.bas.nsh
    stat = np.frombuffer(raw[:16], dtype="<i4")
    v = np.frombuffer(raw[16:], dtype="<f8")
    shapes = [(nao, nao), (3, nao, nao), (6, nao, nao), (nao, nao), (nat,), (3, nat), (3, 3),
              (nao, nao), (3, nao, nao), (6, nao, nao), (3, nat), (nsh,), (3, 3)]
    sizes = [int(np.prod(sh)) for sh in shapes]
    out, o = [], 0
    for n, sh in zip(sizes, shapes):
        out.append(v[o:o + n].reshape(sh, order="F"))
        o += n
    assert o == v.size
    return stat, out
