"""Pin the oracle's diatomic-frame scaling (integral/diat_trafo.f90, used by
get_hamiltonian with do_diat_scale, xtb/h0.f90:345-390) before trusting it.

- the rotation (harmtr): orthogonal, and its bond-axis columns are the
  closed forms of the p_z' and d_z'2 functions along the bond;
- the literal scaled overlaps of the reference's unit tests
  (``test_overlap_diat_{alh3,cl2}``, ``test/unit/test_integral_overlap.f90``,
  ksig = kpi = kdel = 1.2) at the fitted geometry.  With the three factors
  equal, a correct frame gives 1.2 S off-site: an element the rotation fails
  to put on the sigma/pi/delta slots stays unscaled and misses the literal
  (a wrong sigma/pi/delta assignment is planted below and must miss it);
- an independent route with unequal factors: projectors onto the |m| = 0, 1,
  2 subspaces about the bond, from a rotation of the real harmonic
  polynomials fitted numerically (no use of harmtr's formulas):
  S' = S + sum_m (k_m - 1) P_m(j) S P_m(i);
- H0 with the scaling: S, D, Q unchanged, on-site H unchanged, off-site
  shell blocks hij * S', and all factors 1 give the GFN2 H bit for bit.
"""
from functools import lru_cache

import numpy as np
import pytest
from scipy.optimize import least_squares

import refdata
from oracle import oracle as O
from tblite_amd import params, structure

EPS = np.finfo(np.float64).eps
S3 = np.sqrt(3.0)


def _harmonics(l, r):
    """Real solid harmonics of degree l in tblite's order (p: y, z, x; d: m = -2..2)."""
    x, y, z = r[:, 0], r[:, 1], r[:, 2]
    if l == 0:
        return np.ones((len(r), 1))
    if l == 1:
        return np.stack([y, z, x], axis=1)
    return np.stack([S3 * x * y, S3 * y * z, z * z - 0.5 * (x * x + y * y), S3 * x * z,
                     0.5 * S3 * (x * x - y * y)], axis=1)


def _bond_frame(n):
    """Orthonormal axes (e1, e2, n) as the columns of Q (any e1 perpendicular to n)."""
    n = n / np.linalg.norm(n)
    a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
    e1 = a - n * (a @ n)
    e1 /= np.linalg.norm(e1)
    return np.stack([e1, np.cross(n, e1), n], axis=1)


def _projectors(l, n):
    """P_|m| of shell l about the bond n: the bond-frame functions g_k(r) =
    f_k(Q^T r) expanded in the molecular-frame ones, g_k = sum_c D[c, k] f_c."""
    if l == 0:
        return [np.ones((1, 1))]
    Q = _bond_frame(n)
    r = np.random.default_rng(7).normal(size=(64, 3))
    F, G = _harmonics(l, r), _harmonics(l, r @ Q)
    D = np.linalg.lstsq(F, G, rcond=None)[0]
    assert np.abs(D.T @ D - np.eye(2 * l + 1)).max() < 1e-12
    # bond-frame index of |m|: p (y', z', x') -> 1, 0, 1; d -> 2, 1, 0, 1, 2
    am = [1, 0, 1] if l == 1 else [2, 1, 0, 1, 2]
    return [D[:, [k for k in range(2 * l + 1) if am[k] == m]] @
            D[:, [k for k in range(2 * l + 1) if am[k] == m]].T for m in range(l + 1)]


def _scaled_block(S, lj, li, vec, k):
    """S' of a shell-pair block (rows shell lj of atom j, columns li of atom i)."""
    Pj, Pi = _projectors(lj, vec), _projectors(li, vec)
    out = S.copy()
    for m in range(min(lj, li) + 1):
        out += (k[m] - 1.0) * Pj[m] @ S @ Pi[m]
    return out


def _random_vectors():
    rng = np.random.default_rng(21)
    vs = [rng.normal(size=3) * 2.0 for _ in range(4)]
    return vs + [np.array([0.0, 0.0, 2.5]), np.array([0.0, 0.0, -3.0]),
                 np.array([1.5, -2.0, 0.0]), np.array([1e-7, 1e-7, 2.0])]


@pytest.mark.parametrize("iv", range(8))
def test_harmtr_orthogonal_with_bond_axis_columns(iv):
    v = _random_vectors()[iv]
    n = v / np.linalg.norm(v)
    R = O.harmtr(2, v)
    assert np.abs(R.T @ R - np.eye(9)).max() < 1e-14
    assert R[0, 0] == 1.0 and np.all(R[0, 1:] == 0.0) and np.all(R[1:, 0] == 0.0)
    # p_z' along the bond: (n_y, n_z, n_x) in the (y, z, x) order
    assert np.abs(R[1:4, 2] - n[[1, 2, 0]]).max() < 1e-14
    # d_z'2 along the bond: the d functions at the unit vector n
    assert np.abs(R[4:9, 6] - _harmonics(2, n[None, :])[0]).max() < 1e-14
    # no mixing of p and d
    assert np.all(R[1:4, 4:9] == 0.0) and np.all(R[4:9, 1:4] == 0.0)


def _two_atom_block(z_i, z_j, vec):
    """Oracle overlap block of atom pair (i, j) at R_i - R_j = vec: rows the AOs
    of atom j, columns those of atom i, and the setup."""
    xyz = np.array([np.asarray(vec, dtype=np.float64), np.zeros(3)])
    st = structure.setup(structure.Structure([z_i, z_j], xyz))
    (S,) = O.get_multipole_integrals_lat(st, which="overlap")
    ni = st.bas.iao_sh[st.bas.ish_at[1]] if st.mol.nat > 1 else st.bas.nao
    return S[ni:, :ni], st


@pytest.mark.parametrize("iv", range(8))
@pytest.mark.parametrize("pair", [(17, 16), (6, 17), (1, 14), (7, 1)])
def test_diat_trafo_against_projectors(iv, pair):
    """orc_diat_trafo on real overlap blocks (s, p, d shells) with unequal
    factors against the projector form, shell pair by shell pair."""
    v = _random_vectors()[iv]
    v = v / np.linalg.norm(v) * 3.1
    blk, st = _two_atom_block(pair[0], pair[1], v)
    li_max = len(st.spec.basis_cgto[0]) - 1
    lj_max = len(st.spec.basis_cgto[st.mol.id[1]]) - 1
    k = (0.7, 1.35, 0.45)
    B = np.zeros((9, 9), order="F")
    B[:blk.shape[0], :blk.shape[1]] = blk
    got = O.diat_trafo(v, lj_max, li_max, *k, B)
    for lj in range(lj_max + 1):
        for li in range(li_max + 1):
            rj, ri = slice(lj * lj, (lj + 1) ** 2), slice(li * li, (li + 1) ** 2)
            ref = _scaled_block(blk[rj, ri], lj, li, v, k)
            assert np.abs(got[rj, ri] - ref).max() < 1e-14, (lj, li)
    # equal factors of 1 leave the block alone
    assert np.array_equal(O.diat_trafo(v, lj_max, li_max, 1.0, 1.0, 1.0, B), B)


@lru_cache(maxsize=None)
def _fitted_overlap_setup(name):
    """The mstore geometry (absent) fitted to the literal unscaled overlap of
    the same molecule, as in the overlap pins."""
    nums = refdata.literals()["overlap"][name]["numbers"]
    ref = refdata.literal_matrix("overlap", name)
    nat = len(nums)
    mnums, _ = params.species_of(nums)
    cg, nsh = refdata.test_basis_cgto("basis_overlap", mnums)

    def setup_of(x):
        xyz = np.vstack([np.zeros(3), x.reshape(nat - 1, 3)])
        return refdata.custom_setup(nums, xyz, cg, nsh)

    def resid(x):
        return (O.get_multipole_integrals_lat(setup_of(x), which="overlap")[0] - ref).ravel()

    best = None
    for seed in range(10):
        g = np.random.default_rng(seed).normal(size=(nat - 1, 3))
        g = (g / np.linalg.norm(g, axis=1)[:, None] * 2.5).ravel()
        r = least_squares(resid, g, xtol=1e-15, ftol=1e-15, gtol=1e-15)
        best = r if best is None or r.cost < best.cost else best
        if np.abs(r.fun).max() < 1e5 * EPS:
            break
    assert np.abs(best.fun).max() < 1e5 * EPS
    return setup_of(best.x)


@pytest.mark.parametrize("name", ["alh3", "cl2"])
def test_literal_diat_overlap(name):
    """The reference checks these at thr1*10 = 1e6*eps (test_integral_overlap.f90:251)."""
    rec = refdata.literals()["overlap_diat"][name]
    ref = refdata.literal_matrix("overlap_diat", name)
    st = _fitted_overlap_setup(name)
    S, Sd = O.get_overlap_diat_lat(st, rec["ksig"], rec["kpi"], rec["kdel"])
    assert np.abs(Sd - ref).max() < 1e6 * EPS
    # on-site blocks are left unscaled, off-site ones are 1.2 S
    off = ~np.equal.outer(_ao_atom(st), _ao_atom(st))
    assert np.abs(Sd[~off] - S[~off]).max() < 1e-15
    assert np.abs(Sd[off] - 1.2 * S[off]).max() < 1e-13

    # planted mistake: sigma and delta swapped in the d shell (z'2 taken for
    # x'y') -- the block is then not 1.2 S and the literal is missed
    def wrong(l, n):
        P = _projectors(l, n)
        return [P[2], P[1], P[0]] if l == 2 else P

    Sw = S.copy()
    for i in range(st.mol.nat):
        for j in range(st.mol.nat):
            if i == j:
                continue
            v = st.mol.xyz[i] - st.mol.xyz[j]
            for li, ri in _shell_slices(st, i):
                for lj, rj in _shell_slices(st, j):
                    Pj, Pi = wrong(lj, v), wrong(li, v)
                    blk = S[rj, ri]
                    Sw[rj, ri] = blk + sum(0.2 * Pj[m] @ blk @ Pi[m]
                                           for m in range(min(lj, li) + 1))
    assert np.abs(Sw - ref).max() > 1e-3


def _ao_atom(st):
    out = np.zeros(st.bas.nao, dtype=int)
    for iat in range(st.mol.nat):
        a0 = st.bas.iao_sh[st.bas.ish_at[iat]]
        a1 = st.bas.iao_sh[st.bas.ish_at[iat + 1]] if iat + 1 < st.mol.nat else st.bas.nao
        out[a0:a1] = iat
    return out


def _shell_slices(st, iat):
    izp = st.mol.id[iat]
    out = []
    for ish, c in enumerate(st.spec.basis_cgto[izp]):
        a0 = st.bas.iao_sh[st.bas.ish_at[iat] + ish]
        out.append((c.ang, slice(a0, a0 + 2 * c.ang + 1)))
    return out


def _diat_setup(nums, xyz, k=None, seed=0):
    mol = structure.Structure(nums, xyz)
    spec = params.gfn2_species(mol.nums)
    nid = len(mol.nums)
    rng = np.random.default_rng(seed)
    if k is None:
        ks = [rng.uniform(0.5, 1.5, size=(nid, nid)) for _ in range(3)]
    else:
        ks = [np.full((nid, nid), kk) for kk in k]
    spec.ksig, spec.kpi, spec.kdel = [np.asfortranarray(x) for x in ks]
    return structure.setup(mol, spec=spec)


def test_hamiltonian_diat_assembly():
    rng = np.random.default_rng(5)
    nums = [14, 17, 6, 1, 8, 16, 1, 7]
    xyz = rng.uniform(-4.0, 4.0, size=(len(nums), 3))
    st = _diat_setup(nums, xyz, seed=3)
    st0 = structure.setup(structure.Structure(nums, xyz))
    se = rng.uniform(-0.7, -0.1, size=st.bas.nsh)
    S, D, Q, H = O.get_hamiltonian(st, se)
    S0, D0, Q0, H0 = O.get_hamiltonian(st0, se)
    assert np.array_equal(S, S0) and np.array_equal(D, D0) and np.array_equal(Q, Q0)
    at = _ao_atom(st)
    on = np.equal.outer(at, at)
    assert np.array_equal(H[on], H0[on])
    spec, id_ = st.spec, st.mol.id
    checked = 0
    for i in range(st.mol.nat):
        for j in range(st.mol.nat):
            if i == j:
                continue
            v = st.mol.xyz[i] - st.mol.xyz[j]
            # the half list holds (iat, jat <= iat): a block with j > i is the
            # mirror of the pair (j, i), scaled with the factors of (id_j, id_i)
            a, b = max(i, j), min(i, j)
            kk = [x[id_[a], id_[b]] for x in (spec.ksig, spec.kpi, spec.kdel)]
            for li, ri in _shell_slices(st, i):
                for lj, rj in _shell_slices(st, j):
                    blk = S[rj, ri]
                    if np.abs(blk).max() < 1e-8:
                        continue
                    e = np.unravel_index(np.abs(blk).argmax(), blk.shape)
                    hij = H0[rj, ri][e] / blk[e]
                    assert np.abs(H[rj, ri] - hij * _scaled_block(blk, lj, li, v, kk)).max() \
                        < 1e-13
                    checked += 1
    assert checked > 40
    # factors of 1: the GFN2 Hamiltonian, bit for bit
    S1, D1, Q1, H1 = O.get_hamiltonian(_diat_setup(nums, xyz, k=(1.0, 1.0, 1.0)), se)
    assert np.array_equal(H1, H0)


def test_diat_needs_s_p_d_shell_order():
    """Transition metals of GFN2 have their shells in the order d, s, p."""
    xyz = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 3.0]])
    st = _diat_setup([26, 1], xyz)
    with pytest.raises(ValueError, match="s, p, d"):
        O.get_hamiltonian(st, np.zeros(st.bas.nsh))
