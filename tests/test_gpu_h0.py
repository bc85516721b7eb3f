"""Parity of the gfx950 path (through the C-ABI) with the oracle.

Tolerances (the accuracy bar of the path): integrals and H0 to 1e-10 absolute, gradient
contractions to 1e-8 (Eh/bohr); the kernels use the same FP64 formulas in a
different order, so the observed differences are ~1e-14.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tblite_amd import h0, lib as L, structure, synthetic

pytestmark = pytest.mark.gpu

TOL_INT = 1e-10
TOL_GRAD = 1e-8


def _case(name):
    if name == "octane":
        z, x = synthetic.octane()
    elif name == "water32":
        z, x = synthetic.water_cluster(32, seed=7)
    elif name == "mixed80":
        z, x = synthetic.mixed_cluster(80, seed=11, density=1.2)
    elif name == "sulfur_dimer":
        z, x = np.array([16, 16, 17, 1]), np.array([[0, 0, 0], [0.3, 0.2, 3.7], [3.1, -1.0, 1.2],
                                                    [-1.9, 1.1, -0.4]])
    else:
        raise KeyError(name)
    return structure.setup(structure.Structure(z, x))


def _flush(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["sulfur_dimer", "octane", "water32", "mixed80"])
def test_get_hamiltonian_parity(name):
    st = _case(name)
    _flush(f"[gpu] get_hamiltonian {name} nat={st.mol.nat} nao={st.bas.nao}")
    rng = np.random.default_rng(5)
    se, _ = st.selfenergy(rng.uniform(0.5, 3.5, st.mol.nat))
    ref = O.get_hamiltonian(st, se)
    got = h0.get_hamiltonian(st, se)
    for r, g, lab in zip(ref, got, "SDQH"):
        err = np.abs(r - g).max()
        assert err < TOL_INT, (lab, err)
    # symmetry of S and H0 and the dipole shift rule between the two blocks
    S, D, Q, H = got
    assert np.abs(S - S.T).max() < TOL_INT and np.abs(H - H.T).max() < TOL_INT


def _random_density(st, nspin, seed):
    rng = np.random.default_rng(seed)
    nao, nat = st.bas.nao, st.mol.nat

    def sym():
        g = rng.normal(size=(nao, nao))
        return 0.5 * (g + g.T)
    P = np.asfortranarray(np.stack([sym() for _ in range(nspin)], axis=2))
    W = np.asfortranarray(np.stack([sym() for _ in range(nspin)], axis=2))
    vao = np.asfortranarray(rng.normal(scale=1e-2, size=(nao, nspin)))
    vdp = np.asfortranarray(rng.normal(scale=1e-3, size=(3, nat, nspin)))
    vqp = np.asfortranarray(rng.normal(scale=1e-3, size=(6, nat, nspin)))
    return P, W, vao, vdp, vqp


@pytest.mark.parametrize("name,nspin", [("sulfur_dimer", 1), ("octane", 1), ("water32", 2),
                                        ("mixed80", 1)])
def test_get_hamiltonian_gradient_parity(name, nspin):
    st = _case(name)
    _flush(f"[gpu] get_hamiltonian_gradient {name} nspin={nspin}")
    rng = np.random.default_rng(9)
    se, dse = st.selfenergy(rng.uniform(0.5, 3.5, st.mol.nat))
    P, W, vao, vdp, vqp = _random_density(st, nspin, 13)
    nat = st.mol.nat
    # accumulate semantics: start from non-zero values
    c0 = rng.normal(size=nat)
    g0 = np.asfortranarray(rng.normal(size=(3, nat)))
    s0 = np.asfortranarray(rng.normal(size=(3, 3)))
    cr, gr, sr = O.get_hamiltonian_gradient(st, se, dse, vao, vdp, vqp, P, W,
                                            c0.copy(), g0.copy(order="F"), s0.copy(order="F"))
    cg, gg, sg = h0.get_hamiltonian_gradient(st, se, dse, vao, vdp, vqp, P, W, c0.copy(),
                                             g0.copy(order="F"), s0.copy(order="F"))
    assert np.abs(gr - gg).max() < TOL_GRAD
    assert np.abs(cr - cg).max() < TOL_GRAD
    assert np.abs(sr - sg).max() < TOL_GRAD


def test_periodic_images_parity():
    """Lattice images (ntrans > 1) in both forward and gradient."""
    z = np.array([8, 1, 1, 6, 1, 16])
    x = np.array([[0.0, 0.0, 0.0], [1.8, 0.0, 0.0], [-0.5, 1.7, 0.0], [4.0, 3.0, 3.0],
                  [5.9, 3.2, 3.4], [1.0, 5.0, 6.5]])
    cell = 14.0
    trans = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)],
                     dtype=float) * cell
    st = structure.setup(structure.Structure(z, x), trans=trans)
    _flush(f"[gpu] periodic nat={st.mol.nat} ntrans={len(trans)} nnbr={len(st.nlat)}")
    rng = np.random.default_rng(3)
    se, dse = st.selfenergy(rng.uniform(0.5, 3.5, st.mol.nat))
    ref = O.get_hamiltonian(st, se)
    got = h0.get_hamiltonian(st, se)
    for r, g in zip(ref, got):
        assert np.abs(r - g).max() < TOL_INT
    P, W, vao, vdp, vqp = _random_density(st, 1, 4)
    cr, gr, sr = O.get_hamiltonian_gradient(st, se, dse, vao, vdp, vqp, P, W)
    cg, gg, sg = h0.get_hamiltonian_gradient(st, se, dse, vao, vdp, vqp, P, W,
                                             np.zeros(st.mol.nat),
                                             np.zeros((3, st.mol.nat), order="F"),
                                             np.zeros((3, 3), order="F"))
    assert np.abs(gr - gg).max() < TOL_GRAD
    assert np.abs(cr - cg).max() < TOL_GRAD
    assert np.abs(sr - sg).max() < TOL_GRAD


def _periodic_case():
    z = np.array([8, 1, 1, 6, 1, 16])
    x = np.array([[0.0, 0.0, 0.0], [1.8, 0.0, 0.0], [-0.5, 1.7, 0.0], [4.0, 3.0, 3.0],
                  [5.9, 3.2, 3.4], [1.0, 5.0, 6.5]])
    trans = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)],
                     dtype=float) * 14.0
    return structure.setup(structure.Structure(z, x), trans=trans)


@pytest.mark.parametrize("name", ["sulfur_dimer", "octane", "mixed80", "periodic"])
def test_get_hamiltonian_adjoint_parity(name):
    """Host-pointer adjoint of get_hamiltonian against the oracle's adjoint
    with non-symmetric random adjoints (dE/dxyz, dE/dse, sigma; accumulated)."""
    st = _periodic_case() if name == "periodic" else _case(name)
    nao, nat, nsh = st.bas.nao, st.mol.nat, st.bas.nsh
    _flush(f"[gpu] get_hamiltonian_adjoint {name} nat={nat} nao={nao}")
    rng = np.random.default_rng(13)
    se, _ = st.selfenergy(rng.uniform(0.5, 3.5, nat))
    bars = [np.asfortranarray(rng.normal(size=sh)) for sh in
            ((nao, nao), (nao, nao), (3, nao, nao), (6, nao, nao))]
    rg, rdse, rsig = O.get_hamiltonian_adjoint(st, se, *bars)
    g0 = rng.normal(size=(3, nat))
    e0 = rng.normal(size=nsh)
    s0 = rng.normal(size=(3, 3))
    gg, gdse, gsig = (np.asfortranarray(g0.copy()), e0.copy(), np.asfortranarray(s0.copy()))
    h0.get_hamiltonian_adjoint(st, se, *bars, gg, gdse, gsig)
    assert np.abs(rg - (gg - g0)).max() < TOL_GRAD
    assert np.abs(rdse - (gdse - e0)).max() < TOL_GRAD
    assert np.abs(rsig - (gsig - s0)).max() < TOL_GRAD
    # sigma is optional
    g2, d2 = np.zeros((3, nat), order="F"), np.zeros(nsh)
    h0.get_hamiltonian_adjoint(st, se, *bars, g2, d2)
    assert np.abs(g2 - rg).max() < TOL_GRAD and np.abs(d2 - rdse).max() < TOL_GRAD


@pytest.mark.parametrize("name", ["sulfur_dimer", "octane", "periodic"])
def test_full_builders_parity(name):
    """get_multipole_integrals / get_dipole_integrals / get_overlap against the
    oracle's full double loop (integrals.f90:411-470, 994-1059, 1394-1465)."""
    if name == "periodic":
        z = np.array([8, 1, 1, 6, 1, 16])
        x = np.array([[0.0, 0.0, 0.0], [1.8, 0.0, 0.0], [-0.5, 1.7, 0.0], [4.0, 3.0, 3.0],
                      [5.9, 3.2, 3.4], [1.0, 5.0, 6.5]])
        trans = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)],
                         dtype=float) * 14.0
        st = structure.setup(structure.Structure(z, x), trans=trans)
    else:
        st = _case(name)
    _flush(f"[gpu] full builders {name} nat={st.mol.nat} nao={st.bas.nao}")
    Sr, Dr, Qr = O.get_multipole_integrals_lat(st)
    S, D, Q = h0.get_multipole_integrals(st)
    for r, g in ((Sr, S), (Dr, D), (Qr, Q)):
        assert np.abs(r - g).max() < TOL_INT
    # same kernel; the LDS accumulation order may differ between launches
    S2, D2 = h0.get_dipole_integrals(st)
    S3 = h0.get_overlap(st)
    for r, g in ((S, S2), (D, D2), (S, S3)):
        assert np.abs(r - g).max() < 1e-13


def test_list_mismatch_is_reported():
    """A neighbour list built with another cutoff is refused, not silently used."""
    st = _case("octane")
    short = structure.new_adjacency_list(st.mol.xyz, st.trans, 6.0)
    st.inl, st.nnl, st.nlat, st.nltr = short
    se, _ = st.selfenergy(None)
    with pytest.raises(L.TbxError) as e:
        h0.get_hamiltonian(st, se)
    assert e.value.code == 1


def test_unsupported_shell_is_reported():
    st = _case("octane")
    st.ang = st.ang.copy()
    st.ang[0] = 3
    se, _ = st.selfenergy(None)
    with pytest.raises(L.TbxError) as e:
        h0.get_hamiltonian(st, se)
    assert e.value.code == 2


def test_nan_coordinates_are_reported():
    st = _case("octane")
    st.mol.xyz = st.mol.xyz.copy()
    st.mol.xyz[3, 1] = np.nan
    se, _ = st.selfenergy(None)
    with pytest.raises(L.TbxError) as e:
        h0.get_hamiltonian(st, se)
    assert e.value.code == 1
