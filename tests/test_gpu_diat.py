"""H0 from the diatomic-frame scaled overlap (do_diat_scale, xtb/h0.f90:345-390;
integral/diat_trafo.f90) on the device against the oracle: the drop-in
get_hamiltonian for molecules and lattice images, the batched forward with
the near-first pair order, and the gradient entry points refusing the
scaling.  Scale factors are random and not symmetric in the species pair, so
that an (izp, jzp) / (jzp, izp) slip fails.  Tolerance 1e-10 (integrals)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tblite_amd import h0, lib as L, params, structure, synthetic

pytestmark = pytest.mark.gpu

TOL_INT = 1e-10

# per-element factors: k(izp, jzp) = fa[Z_i] * fb[Z_j] for sigma, pi, delta
_FA = np.random.default_rng(41).uniform(0.6, 1.3, size=(3, 87))
_FB = np.random.default_rng(42).uniform(0.8, 1.2, size=(3, 87))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _tables(nums):
    z = np.asarray(nums)
    return [np.asfortranarray(np.outer(_FA[c, z], _FB[c, z])) for c in range(3)]


def _diat_setup(z, x, trans=None):
    mol = structure.Structure(z, x)
    spec = params.gfn2_species(mol.nums)
    spec.ksig, spec.kpi, spec.kdel = _tables(mol.nums)
    return structure.setup(mol, trans=trans, spec=spec)


def _case(name):
    if name == "sulfur_chloride":
        z = np.array([16, 16, 17, 1, 14, 15, 8, 1])
        x = np.array([[0, 0, 0], [0.3, 0.2, 3.7], [3.1, -1.0, 1.2], [-1.9, 1.1, -0.4],
                      [1.0, 4.0, -2.0], [-3.0, -2.5, 2.0], [0.0, 0.0, 6.5], [2.0, 2.0, 7.0]])
        return _diat_setup(z, x.astype(float))
    if name == "mixed40":
        return _diat_setup(*synthetic.mixed_cluster(40, seed=6, density=1.0))
    if name == "random60":
        return _diat_setup(*synthetic.random_molecule(60, np.random.default_rng(9)))
    if name == "periodic":
        z = np.array([8, 1, 1, 6, 1, 16, 17])
        x = np.array([[0.0, 0.0, 0.0], [1.8, 0.0, 0.0], [-0.5, 1.7, 0.0], [4.0, 3.0, 3.0],
                      [5.9, 3.2, 3.4], [1.0, 5.0, 6.5], [-3.0, 2.0, 4.0]])
        trans = np.array([[a, b, c] for a in (-1, 0, 1) for b in (-1, 0, 1)
                          for c in (-1, 0, 1)], dtype=float) * 14.0
        return _diat_setup(z, x, trans)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["sulfur_chloride", "mixed40", "random60", "periodic"])
def test_get_hamiltonian_diat_parity(name):
    st = _case(name)
    _say(f"[gpu] diat get_hamiltonian {name} nat={st.mol.nat} nao={st.bas.nao} "
         f"ntrans={len(st.trans)}")
    rng = np.random.default_rng(5)
    se, _ = st.selfenergy(rng.uniform(0.5, 3.5, st.mol.nat))
    ref = O.get_hamiltonian(st, se)
    got = h0.get_hamiltonian(st, se)
    for r, g, lab in zip(ref, got, "SDQH"):
        assert np.abs(r - g).max() < TOL_INT, (lab, np.abs(r - g).max())
    # the scaling took effect
    plain = structure.setup(st.mol, trans=st.trans)
    H0 = O.get_hamiltonian(plain, se)[3]
    assert np.abs(H0 - got[3]).max() > 1e-4


def test_gradient_refuses_diat():
    st = _case("sulfur_chloride")
    _say("[gpu] diat gradient refusal")
    nao, nat = st.bas.nao, st.mol.nat
    se, dse = st.selfenergy(np.full(nat, 2.0))
    P = np.asfortranarray(np.eye(nao)[:, :, None])
    outs = (np.zeros(nat), np.zeros((3, nat), order="F"), np.zeros((3, 3), order="F"))
    with pytest.raises(L.TbxError) as e:
        h0.get_hamiltonian_gradient(st, se, dse, np.zeros((nao, 1)), np.zeros((3, nat, 1)),
                                    np.zeros((6, nat, 1)), P, P, *outs)
    assert e.value.code == 1 and "diatomic" in str(e.value)
    assert all(np.all(o == 0.0) for o in outs)


def test_batch_forward_diat_parity():
    torch = pytest.importorskip("torch")
    from tblite_amd.batch import Batch
    rng = np.random.default_rng(77)
    mols = [synthetic.random_molecule(int(k), rng) for k in rng.integers(4, 120, size=9)]
    mols.append(synthetic.mixed_cluster(50, seed=8, density=1.1))
    mols.append((np.array([17]), np.zeros((1, 3))))
    b = Batch(mols, diat_scale=_tables)
    _say(f"[gpu] diat batch forward nmol={b.nmol} nat={b.nat} ndense={b.ndense}")
    cn = torch.tensor(rng.uniform(0.5, 3.5, b.nat), dtype=torch.float64, device="cuda")
    se, _ = b.selfenergy(cn)
    se = se.contiguous()
    S, D, Q, H = b.forward(b.tensor_xyz(), se)
    torch.cuda.synchronize()
    views = [b.molecule_views(t, c) for t, c in ((S, 1), (D, 3), (Q, 6), (H, 1))]
    se_h = se.cpu().numpy()
    worst = 0.0
    for m, (z, x) in enumerate(mols):
        st = _diat_setup(z, x)
        ref = O.get_hamiltonian(st, se_h[int(b.sh_off[m]):int(b.sh_off[m + 1])])
        for r, v in zip(ref, (vv[m] for vv in views)):
            g = np.transpose(v.cpu().numpy(), (2, 1, 0))
            err = np.abs(r.reshape(g.shape, order="F") - g).max()
            worst = max(worst, err)
            assert err < TOL_INT, (m, err)
    _say(f"    max |GPU - oracle| = {worst:.1e}")
    with pytest.raises(L.TbxError):
        b.adjoint(b.tensor_xyz(), se, S, D, Q, H)
