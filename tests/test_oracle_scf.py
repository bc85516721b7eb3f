"""Pin the oracle of the SCF-loop consumers (oracle/scf.py) before trusting it.

The checks are identities the physics fixes, computed by a different route
(traces and plain matrix algebra) from the loops of the restatement:

- a constant potential: add_vao_to_h1 adds -c S;
- the variational pairs: tr(P dh1) of add_vao_to_h1 / add_vmp_to_h1 equals the
  potential contracted with get_mulliken_shell_charges / _atomic_multipoles
  of the same density (the Mulliken charges and multipoles are the
  derivatives of the energy with respect to those potentials);
- the charge sum sum(qsh) = sum(n0sh) - tr(P S);
- the spin representation: equal up and down densities give the closed-shell
  charges with zero magnetisation (updown_to_magnet, wavefunction/spin.f90);
- the molecular dipole: sum_a (mu_a + q_a R_a) of the Mulliken atomic dipoles
  and charges equals -tr(P r) with the origin dipole integrals of the full
  double-loop builder (an independent route to the integrals and the sum).
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import scf as OS
from tblite_amd import structure, synthetic


def _setup(name):
    if name == "octane":
        z, x = synthetic.octane()
    else:  # d shells
        z, x = np.array([16, 16, 17, 1, 8, 1, 1]), np.array(
            [[0, 0, 0], [0.3, 0.2, 3.7], [3.1, -1.0, 1.2], [-1.9, 1.1, -0.4], [5.0, 2.0, 1.0],
             [6.7, 2.1, 1.5], [4.6, 3.6, 0.5]])
    st = structure.setup(structure.Structure(z, x))
    S, D, Q, _ = O.get_hamiltonian(st, np.zeros(st.bas.nsh))
    return st, S, D, Q


def _density(nao, rng):
    c = rng.normal(size=(nao, nao // 2))
    return c @ c.T / nao


def _atom_ranges(bas):
    nat = len(bas.ish_at)
    first = bas.iao_sh[bas.ish_at]
    nao_at = np.bincount(bas.ao2at, minlength=nat)
    return first, nao_at


@pytest.mark.parametrize("name", ["octane", "sulfur"])
def test_constant_potential(name):
    st, S, _, _ = _setup(name)
    nao = st.bas.nao
    h1 = np.zeros((nao, nao, 2))
    vao = np.array([[0.3, -1.1]] * nao)
    OS.add_vao_to_h1(S, vao, h1)
    assert np.abs(h1[:, :, 0] + 0.3 * S).max() < 1e-14
    assert np.abs(h1[:, :, 1] - 1.1 * S).max() < 1e-14


@pytest.mark.parametrize("name", ["octane", "sulfur"])
def test_vao_against_shell_charges(name):
    st, S, _, _ = _setup(name)
    bas, nao = st.bas, st.bas.nao
    rng = np.random.default_rng(1)
    P = _density(nao, rng)
    vsh = rng.normal(size=bas.nsh)
    vao = np.repeat(vsh, bas.nao_sh)[:, None]
    h1 = np.zeros((nao, nao, 1))
    OS.add_vao_to_h1(S, vao, h1)
    n0 = rng.uniform(0, 2, bas.nsh)
    qsh = OS.get_mulliken_shell_charges(S, P[:, :, None], n0, bas.iao_sh, bas.nao_sh)
    assert abs(np.sum(P * h1[:, :, 0]) - vsh @ (qsh[:, 0] - n0)) < 1e-12
    assert abs(qsh[:, 0].sum() - (n0.sum() - np.sum(P * S))) < 1e-12


@pytest.mark.parametrize("name", ["octane", "sulfur"])
@pytest.mark.parametrize("which", ["dipole", "quadrupole"])
def test_vmp_against_atomic_multipoles(name, which):
    st, S, D, Q = _setup(name)
    bas, nao, nat = st.bas, st.bas.nao, st.mol.nat
    mp = D if which == "dipole" else Q
    nmp = mp.shape[0]
    rng = np.random.default_rng(2)
    P = _density(nao, rng)
    vmp = rng.normal(size=(nmp, nat, 1))
    h1 = np.zeros((nao, nao, 1))
    OS.add_vmp_to_h1(mp, vmp, bas.ao2at, h1)
    first, nao_at = _atom_ranges(bas)
    mpat = OS.get_mulliken_atomic_multipoles(mp, P[:, :, None], first, nao_at)
    assert abs(np.sum(P * h1[:, :, 0]) - np.sum(vmp[:, :, 0] * mpat[:, :, 0])) < 1e-12
    # the update is symmetric and only the leading min(nmp, nvmp) components enter
    assert np.abs(h1[:, :, 0] - h1[:, :, 0].T).max() < 1e-14
    h2 = np.zeros((nao, nao, 1))
    OS.add_vmp_to_h1(mp, np.concatenate([vmp, rng.normal(size=(2, nat, 1))]), bas.ao2at, h2)
    assert np.array_equal(h1, h2)


def test_spin_representation():
    st, S, D, _ = _setup("sulfur")
    bas, nao = st.bas, st.bas.nao
    rng = np.random.default_rng(3)
    P = _density(nao, rng)
    n0 = rng.uniform(0, 2, bas.nsh)
    q1 = OS.get_mulliken_shell_charges(S, P[:, :, None], n0, bas.iao_sh, bas.nao_sh)
    q2 = OS.get_mulliken_shell_charges(S, np.stack([P / 2, P / 2], axis=2), n0, bas.iao_sh,
                                       bas.nao_sh)
    assert np.abs(q2[:, 0] - q1[:, 0]).max() < 1e-13 and np.abs(q2[:, 1]).max() < 1e-13
    Pu, Pd = _density(nao, rng), _density(nao, rng)
    q3 = OS.get_mulliken_shell_charges(S, np.stack([Pu, Pd], axis=2), n0, bas.iao_sh, bas.nao_sh)
    qu = OS.get_mulliken_shell_charges(S, Pu[:, :, None], 0 * n0, bas.iao_sh, bas.nao_sh)[:, 0]
    qd = OS.get_mulliken_shell_charges(S, Pd[:, :, None], 0 * n0, bas.iao_sh, bas.nao_sh)[:, 0]
    assert np.abs(q3[:, 0] - (qu + qd + n0)).max() < 1e-13
    assert np.abs(q3[:, 1] - (qu - qd)).max() < 1e-13
    first, nao_at = _atom_ranges(bas)
    m2 = OS.get_mulliken_atomic_multipoles(D, np.stack([Pu, Pd], axis=2), first, nao_at)
    mu = OS.get_mulliken_atomic_multipoles(D, Pu[:, :, None], first, nao_at)[:, :, 0]
    md = OS.get_mulliken_atomic_multipoles(D, Pd[:, :, None], first, nao_at)[:, :, 0]
    assert np.abs(m2[:, :, 0] - (mu + md)).max() < 1e-13
    assert np.abs(m2[:, :, 1] - (mu - md)).max() < 1e-13


def test_identity_detects_planted_mistake():
    """A transposed operand in add_vmp_to_h1 (the atom of the row instead of
    the column) breaks the variational identity."""
    st, S, D, _ = _setup("sulfur")
    bas, nao, nat = st.bas, st.bas.nao, st.mol.nat
    rng = np.random.default_rng(4)
    P = _density(nao, rng)
    vmp = rng.normal(size=(3, nat, 1))
    h1 = np.zeros((nao, nao, 1))
    wrong = np.transpose(D, (0, 2, 1)).copy()
    OS.add_vmp_to_h1(wrong, vmp, bas.ao2at, h1)
    first, nao_at = _atom_ranges(bas)
    mpat = OS.get_mulliken_atomic_multipoles(D, P[:, :, None], first, nao_at)
    assert abs(np.sum(P * h1[:, :, 0]) - np.sum(vmp[:, :, 0] * mpat[:, :, 0])) > 1e-6


# -- density matrices ----------------------------------------------------------

def _eigen_case(n=12, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.normal(size=(n, n))
    s = a @ a.T / n + np.eye(n)          # positive definite overlap
    h = rng.normal(size=(n, n))
    return 0.5 * (h + h.T), s


def test_oracle_density_matrix_identities():
    """For S-orthonormal C: tr(P S) = sum focc, P S C = C diag(focc), P
    symmetric; W is the density of focc * eval; a transposed coefficient
    matrix breaks them."""
    h, s = _eigen_case()
    e, c = OS.sygvd(h, s)
    assert np.abs(c.T @ s @ c - np.eye(len(e))).max() < 1e-12
    assert np.abs(h @ c - s @ c * e[None, :]).max() < 1e-11
    focc = np.random.default_rng(1).uniform(0.0, 2.0, len(e))
    p = OS.get_density_matrix(focc, c)
    assert abs(np.trace(p @ s) - focc.sum()) < 1e-12
    assert np.abs(p @ s @ c - c * focc[None, :]).max() < 1e-12
    assert np.abs(p - p.T).max() < 1e-14
    w = OS.get_wdensity(focc, e, c)
    assert np.abs(w @ s @ c - c * (focc * e)[None, :]).max() < 1e-11
    assert abs(np.trace(p @ h) - np.trace(w @ s)) < 1e-11  # both are sum f e
    bad = OS.get_density_matrix(focc, c.T)
    assert abs(np.trace(bad @ s) - focc.sum()) > 1e-3


def test_oracle_fermi_filling():
    """The occupations sum to the electron count; they are the Fermi function
    of the returned level (2 e_fermi); small kt gives the aufbau filling."""
    emo = np.sort(np.random.default_rng(4).normal(size=20))
    kt = 300.0 * 3.166808578545117e-6
    for nel in (7.0, 7.3, 7.6, 0.4):
        homo, occ, ef = OS.get_fermi_filling(nel, kt, emo)
        if homo > 0:
            occt = float(homo)
            assert abs(occ.sum() - occt) < np.sqrt(np.finfo(float).eps)
            # occupations of the level before the last Newton step: within one step
            with np.errstate(over="ignore"):  # exp overflows to inf far above the level: f = 0
                f = 1.0 / (np.exp((emo - 2 * ef) / kt) + 1.0)
            assert abs(f.sum() - occt) < 1e-6
    homo, occ, _ = OS.get_fermi_filling(8.0, 1e-6, emo)
    assert homo == 8 and np.allclose(occ[:8], 1.0) and np.allclose(occ[8:], 0.0)
    homo, occ, ef = OS.get_fermi_filling(0.4, kt, emo)  # homo = 0: aufbau only
    assert homo == 0 and ef == 0.0 and occ[0] == 0.4 and occ[1:].sum() == 0.0


@pytest.mark.parametrize("name", ["octane", "sulfur"])
def test_molecular_dipole_closed_form(name):
    """sum_a (mu_a + q_a R_a) = -tr(P r) about the origin: the atomic dipoles
    (about each atom) and the electronic Mulliken charges of the restatement
    against the origin dipole integrals of the full double-loop builder,
    <j|r|i> = <j|r - R_i|i> + R_i S_ji (no half list, no shift_operator).
    A dipole of the wrong sign, or charges placed on the wrong atoms, misses it."""
    st, S, D, _ = _setup(name)
    bas, nao, nat = st.bas, st.bas.nao, st.mol.nat
    rng = np.random.default_rng(8)
    P = _density(nao, rng)
    S_lat, D_lat = O.get_multipole_integrals_lat(st, which="dipole")
    assert np.abs(S_lat - S).max() < 1e-13
    R_ao = st.mol.xyz[bas.ao2at]  # centre of the column function
    r_origin = D_lat + R_ao.T[:, None, :] * S_lat[None, :, :]
    mu_ref = -np.einsum("ji,kji->k", P, r_origin)
    first, nao_at = _atom_ranges(bas)
    mpat = OS.get_mulliken_atomic_multipoles(D, P[:, :, None], first, nao_at)[:, :, 0]
    n0 = np.zeros(bas.nsh)
    qsh = OS.get_mulliken_shell_charges(S, P[:, :, None], n0, bas.iao_sh, bas.nao_sh)[:, 0]
    q_at = np.bincount(np.repeat(np.arange(nat), st.spec.nsh_id[st.mol.id]), weights=qsh,
                       minlength=nat)
    mu = mpat.sum(axis=1) + st.mol.xyz.T @ q_at
    assert np.abs(mu - mu_ref).max() < 1e-11
    assert np.abs(mu_ref).max() > 1e-2
    # planted mistakes
    assert np.abs(-mpat.sum(axis=1) + st.mol.xyz.T @ q_at - mu_ref).max() > 1e-3
    assert np.abs(mpat.sum(axis=1) + st.mol.xyz.T @ np.roll(q_at, 1) - mu_ref).max() > 1e-3
