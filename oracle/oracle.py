"""ctypes front end of the oracle (``tbx_oracle.c``).

TEST INFRASTRUCTURE: imported only by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``.  The product never imports it.

Build with ``make -C oracle`` (``__graft_entry__.build()`` does this).

Parity status of each oracle function (what pins it; tests/test_oracle_pins.py):

- ``overlap_cgto`` / STO-nG tables: the reference's 93 self-overlap KATs
  (``test/unit/test_slater_expansion.f90``); literal overlap matrices of
  ``test/unit/test_integral_overlap.f90`` reproduced after fitting the
  unknown geometry; Gaussian product theorem closed form for s-s.
- ``multipole_cgto``: dipole translation rule of
  ``test/unit/test_integral_multipole.f90:108-211``; brute-force 3D
  quadrature of the moment integrals on tiny inputs.
- ``multipole_grad_cgto``: central finite differences and centre-exchange
  antisymmetry (``test_integral_overlap.f90:1458-1691``,
  ``test_integral_multipole.f90:214-264``).
- ``get_hamiltonian``: literal GFN2 H0 matrices of
  ``test/unit/test_hamiltonian.f90:510-745`` reproduced after fitting the
  unknown geometry, coordination number and radius.
- ``get_hamiltonian_gradient``: central finite differences of the energy
  functional it differentiates (method of ``test_hamiltonian.f90:179-262``).
"""
from __future__ import annotations

import ctypes as ct
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libtbx_oracle.so"

MAXG = 12
CGTO_DTYPE = np.dtype([("ang", np.int32), ("nprim", np.int32),
                       ("alpha", np.float64, (MAXG,)), ("coeff", np.float64, (MAXG,))],
                      align=True)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: run `make -C {HERE}`")
        _lib = ct.CDLL(str(LIB_PATH))
        _lib.orc_overlap_1d.restype = ct.c_double
        _lib.orc_overlap_1d.argtypes = [ct.c_int, ct.c_double]
    return _lib


def _p(a):
    return a.ctypes.data_as(ct.c_void_p)


def cgto_record(ang, alpha, coeff):
    rec = np.zeros(1, dtype=CGTO_DTYPE)
    rec["ang"] = ang
    rec["nprim"] = len(alpha)
    rec["alpha"][0, : len(alpha)] = alpha
    rec["coeff"][0, : len(coeff)] = coeff
    return rec


def _vec(vec):
    v = np.ascontiguousarray(vec, dtype=np.float64).reshape(3)
    return v, float(v @ v)


def overlap_cgto(cj, ci, vec, intcut=25.0, r2=None):
    """Spherical overlap block, Fortran layout (nsj, nsi)."""
    v, rr = _vec(vec)
    r2 = rr if r2 is None else r2
    nj, ni = 2 * int(cj["ang"][0]) + 1, 2 * int(ci["ang"][0]) + 1
    out = np.zeros((nj, ni), order="F")
    st = lib().orc_overlap_cgto(_p(cj), _p(ci), ct.c_double(r2), _p(v), ct.c_double(intcut), _p(out))
    if st != 0:
        raise ValueError("oracle: l > 2 not supported for spherical blocks")
    return out


def cart_overlap(cj, ci, vec, intcut=25.0):
    v, r2 = _vec(vec)
    nj = (int(cj["ang"][0]) + 1) * (int(cj["ang"][0]) + 2) // 2
    ni = (int(ci["ang"][0]) + 1) * (int(ci["ang"][0]) + 2) // 2
    out = np.zeros((nj, ni), order="F")
    st = lib().orc_cart_overlap(_p(cj), _p(ci), ct.c_double(r2), _p(v), ct.c_double(intcut), _p(out))
    if st != 0:
        raise ValueError("oracle: l > 4")
    return out


def multipole_cgto(cj, ci, vec, intcut=25.0, r2=None):
    """(S (nj,ni), D (3,nj,ni), Q (6,nj,ni)), operator on centre i."""
    v, rr = _vec(vec)
    r2 = rr if r2 is None else r2
    nj, ni = 2 * int(cj["ang"][0]) + 1, 2 * int(ci["ang"][0]) + 1
    s = np.zeros((nj, ni), order="F")
    d = np.zeros((3, nj, ni), order="F")
    q = np.zeros((6, nj, ni), order="F")
    st = lib().orc_multipole_cgto(_p(cj), _p(ci), ct.c_double(r2), _p(v), ct.c_double(intcut),
                                  _p(s), _p(d), _p(q))
    if st != 0:
        raise ValueError("oracle: l > 2")
    return s, d, q


def multipole_grad_cgto(cj, ci, vec, intcut=25.0):
    """Returns dict with s, d (op j), q (op j), ds (3,nj,ni), ddj, dqj, ddi, dqi
    ((3,3,nj,ni) / (3,6,nj,ni), first index derivative direction)."""
    v, r2 = _vec(vec)
    nj, ni = 2 * int(cj["ang"][0]) + 1, 2 * int(ci["ang"][0]) + 1
    out = dict(s=np.zeros((nj, ni), order="F"), d=np.zeros((3, nj, ni), order="F"),
               q=np.zeros((6, nj, ni), order="F"), ds=np.zeros((3, nj, ni), order="F"),
               ddj=np.zeros((3, 3, nj, ni), order="F"), dqj=np.zeros((3, 6, nj, ni), order="F"),
               ddi=np.zeros((3, 3, nj, ni), order="F"), dqi=np.zeros((3, 6, nj, ni), order="F"))
    st = lib().orc_multipole_grad_cgto(
        _p(cj), _p(ci), ct.c_double(r2), _p(v), ct.c_double(intcut), _p(out["s"]), _p(out["d"]),
        _p(out["q"]), _p(out["ds"]), _p(out["ddj"]), _p(out["dqj"]), _p(out["ddi"]),
        _p(out["dqi"]))
    if st != 0:
        raise ValueError("oracle: l > 2")
    return out


def shift_operator(vec, s, di, qi):
    """Forward shift of the multipole operator from centre i to centre j (h0.f90:704-747)."""
    v = np.ascontiguousarray(vec, dtype=np.float64)
    di = np.ascontiguousarray(di, dtype=np.float64)
    qi = np.ascontiguousarray(qi, dtype=np.float64)
    dj, qj = np.zeros(3), np.zeros(6)
    lib().orc_shift_operator(_p(v), ct.c_double(s), _p(di), _p(qi), _p(dj), _p(qj))
    return dj, qj


def overlap_1d(moment, alpha):
    return lib().orc_overlap_1d(int(moment), float(alpha))


def _cgto_array(setup):
    nid, msh = len(setup.spec.nums), setup.msh
    arr = np.zeros(nid * msh, dtype=CGTO_DTYPE)
    arr["ang"] = setup.ang
    arr["nprim"] = setup.nprim
    arr["alpha"] = setup.alpha.reshape(nid * msh, MAXG)
    arr["coeff"] = setup.coeff.reshape(nid * msh, MAXG)
    return arr


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _common_args(setup):
    mol, bas, spec = setup.mol, setup.bas, setup.spec
    keep = [_i32(mol.id), _f64(mol.xyz), _f64(setup.trans), _i32(setup.inl), _i32(setup.nnl),
            _i32(setup.nlat), _i32(setup.nltr), _i32(spec.nsh_id), _i32(bas.ish_at),
            _i32(bas.iao_sh), _cgto_array(setup),
            _f64(spec.hscale.reshape(-1, order="F")), _f64(spec.shpoly.reshape(-1, order="F")),
            _f64(spec.rad)]
    args = [ct.c_int(mol.nat), ct.c_int(mol.nid), ct.c_int(setup.msh), ct.c_int(bas.nsh),
            ct.c_int(bas.nao), _p(keep[0]), _p(keep[1]), ct.c_int(len(setup.trans)), _p(keep[2]),
            _p(keep[3]), _p(keep[4]), _p(keep[5]), _p(keep[6]), _p(keep[7]), _p(keep[8]),
            _p(keep[9]), _p(keep[10]), ct.c_double(bas.intcut), _p(keep[11]), _p(keep[12]),
            _p(keep[13])]
    return args, keep


def get_hamiltonian(setup, selfenergy):
    """Dense Fortran-order S (nao,nao), D (3,nao,nao), Q (6,nao,nao), H (nao,nao).
    With ``setup.spec.ksig`` set, H0 is built from the diatomic-frame scaled
    overlap (``do_diat_scale``, xtb/h0.f90:345-390)."""
    nao = setup.bas.nao
    S = np.zeros((nao, nao), order="F")
    D = np.zeros((3, nao, nao), order="F")
    Q = np.zeros((6, nao, nao), order="F")
    H = np.zeros((nao, nao), order="F")
    se = _f64(selfenergy)
    args, keep = _common_args(setup)
    spec = setup.spec
    if spec.ksig is None:
        st = lib().orc_get_hamiltonian(*args, _p(se), _p(S), _p(D), _p(Q), _p(H))
    else:
        ks = [_f64(np.asarray(k, dtype=np.float64).reshape(-1, order="F"))
              for k in (spec.ksig, spec.kpi, spec.kdel)]
        st = lib().orc_get_hamiltonian_diat(*args, *[_p(k) for k in ks], _p(se), _p(S), _p(D),
                                            _p(Q), _p(H))
    if st == -2:
        raise ValueError("oracle: diatomic-frame scaling needs shells s, p, d in this order")
    if st != 0:
        raise ValueError("oracle: unsupported basis")
    return S, D, Q, H


def harmtr(maxl, vec):
    """Diatomic-frame rotation (integral/diat_trafo.f90:221-335) as a (9, 9)
    array: rows the molecular-frame functions (s; p y, z, x; d m = -2..2),
    columns those of the frame whose z axis is ``vec``."""
    R = np.zeros((9, 9), order="F")
    lib().orc_harmtr(ct.c_int(maxl), _p(_f64(vec)), _p(R))
    return R


def diat_trafo(vec, maxlj, maxli, ksig, kpi, kdel, block):
    """diat_trafo (integral/diat_trafo.f90:99-156, no derivative) of a (9, 9)
    atom-pair block (rows atom j, columns atom i); returns the scaled block."""
    B = np.array(block, dtype=np.float64, order="F").reshape(9, 9, order="F").copy(order="F")
    lib().orc_diat_trafo(_p(_f64(vec)), ct.c_int(maxlj), ct.c_int(maxli), ct.c_double(ksig),
                         ct.c_double(kpi), ct.c_double(kdel), _p(B))
    return B


def get_overlap_diat_lat(setup, ksig, kpi, kdel, cutoff=None):
    """get_overlap_diat_lat (integral/native/integrals.f90:473-582): S and
    the diatomic-frame scaled S, dense, Fortran order; ksig/kpi/kdel are
    (nid, nid) tables indexed (izp, jzp)."""
    mol, bas, nao = setup.mol, setup.bas, setup.bas.nao
    cut = bas.get_cutoff() if cutoff is None else cutoff
    S = np.zeros((nao, nao), order="F")
    Sd = np.zeros((nao, nao), order="F")
    keep = [_i32(mol.id), _f64(mol.xyz), _f64(setup.trans), _i32(setup.spec.nsh_id),
            _i32(bas.ish_at), _i32(bas.iao_sh), _cgto_array(setup)]
    ks = [_f64(np.broadcast_to(np.asarray(k, dtype=np.float64), (mol.nid, mol.nid))
               .reshape(-1, order="F")) for k in (ksig, kpi, kdel)]
    st = lib().orc_get_overlap_diat_lat(
        ct.c_int(mol.nat), ct.c_int(mol.nid), ct.c_int(setup.msh), ct.c_int(nao), _p(keep[0]),
        _p(keep[1]), ct.c_int(len(setup.trans)), _p(keep[2]), ct.c_double(cut), _p(keep[3]),
        _p(keep[4]), _p(keep[5]), _p(keep[6]), ct.c_double(bas.intcut), *[_p(k) for k in ks],
        _p(S), _p(Sd))
    if st == -2:
        raise ValueError("oracle: diatomic-frame scaling needs shells s, p, d in this order")
    if st != 0:
        raise ValueError("oracle: unsupported basis")
    return S, Sd


def get_multipole_integrals_lat(setup, cutoff=None, which="multipole"):
    """Full double-loop builders (integral/native/integrals.f90:411-470,
    994-1059, 1394-1465): which = "overlap" -> S; "dipole" -> S, D;
    "multipole" -> S, D, Q.  Dense, Fortran order."""
    mol, bas, nao = setup.mol, setup.bas, setup.bas.nao
    cut = bas.get_cutoff() if cutoff is None else cutoff
    S = np.zeros((nao, nao), order="F")
    D = np.zeros((3, nao, nao), order="F") if which in ("dipole", "multipole") else None
    Q = np.zeros((6, nao, nao), order="F") if which == "multipole" else None
    keep = [_i32(mol.id), _f64(mol.xyz), _f64(setup.trans), _i32(setup.spec.nsh_id),
            _i32(bas.ish_at), _i32(bas.iao_sh), _cgto_array(setup)]
    st = lib().orc_get_multipole_integrals_lat(
        ct.c_int(mol.nat), ct.c_int(setup.msh), ct.c_int(nao), _p(keep[0]), _p(keep[1]),
        ct.c_int(len(setup.trans)), _p(keep[2]), ct.c_double(cut), _p(keep[3]), _p(keep[4]),
        _p(keep[5]), _p(keep[6]), ct.c_double(bas.intcut), _p(S),
        _p(D) if D is not None else None, _p(Q) if Q is not None else None)
    if st != 0:
        raise ValueError("oracle: unsupported basis")
    return tuple(x for x in (S, D, Q) if x is not None)


def get_hamiltonian_gradient(setup, selfenergy, dsedcn, vao, vdp, vqp, pmat, xmat,
                             dEdcn=None, gradient=None, sigma=None):
    """Accumulates into dEdcn (nat), gradient (3,nat) and sigma (3,3) (Fortran order)."""
    mol, nao = setup.mol, setup.bas.nao
    nspin = pmat.shape[2] if pmat.ndim == 3 else 1
    pm = np.asfortranarray(pmat.reshape(nao, nao, nspin))
    xm = np.asfortranarray(xmat.reshape(nao, nao, -1))
    va = np.asfortranarray(np.asarray(vao, dtype=np.float64).reshape(nao, -1))
    vd = np.asfortranarray(np.asarray(vdp, dtype=np.float64).reshape(3, mol.nat, -1))
    vq = np.asfortranarray(np.asarray(vqp, dtype=np.float64).reshape(6, mol.nat, -1))
    dEdcn = np.zeros(mol.nat) if dEdcn is None else dEdcn
    gradient = np.zeros((3, mol.nat), order="F") if gradient is None else gradient
    sigma = np.zeros((3, 3), order="F") if sigma is None else sigma
    se, dse = _f64(selfenergy), _f64(dsedcn)
    args, keep = _common_args(setup)
    st = lib().orc_get_hamiltonian_gradient(*args, _p(se), _p(dse), ct.c_int(nspin), _p(va),
                                            _p(vd), _p(vq), _p(pm), _p(xm), _p(dEdcn),
                                            _p(gradient), _p(sigma))
    if st != 0:
        raise ValueError("oracle: unsupported basis")
    return dEdcn, gradient, sigma


def get_hamiltonian_adjoint(setup, selfenergy, sbar, hbar, dbar, qbar, gradient=None, dEdse=None,
                            sigma=None):
    """Adjoint of ``get_hamiltonian``: for E = <sbar,S> + <hbar,H> + <dbar,D> +
    <qbar,Q> (Fortran-ordered (nao,nao), (nao,nao), (3,nao,nao), (6,nao,nao)
    adjoints) accumulates dE/dxyz into ``gradient`` (3,nat), dE/dselfenergy
    into ``dEdse`` (nsh) and the strain derivative into ``sigma`` (3,3)."""
    mol, nao = setup.mol, setup.bas.nao
    sb = np.asfortranarray(np.asarray(sbar, dtype=np.float64).reshape(nao, nao, order="F"))
    hb = np.asfortranarray(np.asarray(hbar, dtype=np.float64).reshape(nao, nao, order="F"))
    db = np.asfortranarray(np.asarray(dbar, dtype=np.float64).reshape(3, nao, nao, order="F"))
    qb = np.asfortranarray(np.asarray(qbar, dtype=np.float64).reshape(6, nao, nao, order="F"))
    gradient = np.zeros((3, mol.nat), order="F") if gradient is None else gradient
    dEdse = np.zeros(setup.bas.nsh) if dEdse is None else dEdse
    sigma = np.zeros((3, 3), order="F") if sigma is None else sigma
    se = _f64(selfenergy)
    args, keep = _common_args(setup)
    st = lib().orc_get_hamiltonian_adjoint(*args, _p(se), _p(sb), _p(hb), _p(db), _p(qb),
                                           _p(gradient), _p(dEdse), _p(sigma))
    if st != 0:
        raise ValueError("oracle: unsupported basis")
    return gradient, dEdse, sigma


def ncoord_dexp(ids, xyz, trans, rcov, ka=10.0, kb=20.0, rshift=2.0, cutoff=25.0,
                derivatives=True):
    """Double-exponential coordination number (mctc-lib ncoord "dexp", GFN2):
    cn (nat), and with ``derivatives`` dcndr (3, nat, nat) = d cn(i)/d xyz(:, j)
    as [c, j, i] in Fortran order, and dcndL (3, 3, nat)."""
    ids = _i32(ids)
    x = _f64(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
    tr = _f64(np.asarray(trans, dtype=np.float64).reshape(-1, 3))
    rc = _f64(rcov)
    nat = len(ids)
    cn = np.zeros(nat)
    dcndr = np.zeros((3, nat, nat), order="F") if derivatives else None
    dcndL = np.zeros((3, 3, nat), order="F") if derivatives else None
    lib().orc_ncoord_dexp(ct.c_int(nat), _p(ids), _p(x), ct.c_int(len(tr)), _p(tr), _p(rc),
                          ct.c_double(ka), ct.c_double(kb), ct.c_double(rshift),
                          ct.c_double(cutoff), _p(cn),
                          _p(dcndr) if derivatives else None,
                          _p(dcndL) if derivatives else None)
    return (cn, dcndr, dcndL) if derivatives else cn


def get_lattice_points(periodic, lattice, rthr):
    """get_lattice_points_cutoff (cutoff.f90:97-121) in plain Python: lattice
    vectors are the rows of ``lattice``; returns (ntr, 3), origin first."""
    lat = np.asarray(lattice, dtype=np.float64).reshape(3, 3)
    if not any(periodic):
        return np.zeros((1, 3))
    rep = []
    for k in range(3):  # get_translations, cutoff.f90:127-160
        n = np.cross(lat[(k + 1) % 3], lat[(k + 2) % 3])
        n = n / np.sqrt(np.sum(n * n))
        rep.append(int(np.ceil(abs(rthr / np.sum(n * lat[k])))))
    out = []
    for ix in range(rep[0] + 1):  # get_lattice_points_rep_3d, cutoff.f90:39-94
        for iy in range(rep[1] + 1):
            for iz in range(rep[2] + 1):
                for jx in ((1, -1) if ix > 0 else (1,)):
                    for jy in ((1, -1) if iy > 0 else (1,)):
                        for jz in ((1, -1) if iz > 0 else (1,)):
                            out.append(lat[0] * ix * jx + lat[1] * iy * jy + lat[2] * iz * jz)
    return np.array(out)
