"""Edge cases of the batched and drop-in paths against the oracle: single-atom
molecules (an s-only and a d-shell atom), atoms beyond the neighbour cutoff
(no off-site pairs at all), a diatomic, and a mixed molecule with one distant
atom, batched together with an ordinary molecule.  Tolerances as in
test_gpu_batch.py."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tblite_amd import h0, structure, synthetic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL_INT = 1e-10
TOL_GRAD = 1e-8


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _edge_molecules():
    rng = np.random.default_rng(77)
    return [
        (np.array([1]), np.zeros((1, 3))),                                  # lone H
        (np.array([16]), np.array([[0.5, -0.2, 0.1]])),                     # lone S (d shell)
        (np.array([1, 1]), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 50.0]])),   # beyond the cutoff
        (np.array([6, 1]), np.array([[0.0, 0.0, 0.0], [0.3, 0.4, 2.0]])),    # diatomic
        (np.array([8, 1, 17]), np.array([[0.0, 0.0, 0.0], [1.8, 0.0, 0.0],
                                         [0.0, 45.0, 10.0]])),              # one distant atom
        synthetic.random_molecule(20, rng),
    ]


@pytest.fixture(scope="module")
def edge_case():
    from tblite_amd.batch import Batch
    mols = _edge_molecules()
    b = Batch(mols)
    rng = np.random.default_rng(5)
    cn = torch.tensor(rng.uniform(0.5, 3.5, b.nat), dtype=torch.float64, device="cuda")
    se, dse = b.selfenergy(cn)
    return mols, b, cn, se.contiguous(), dse.contiguous()


def test_edge_forward_parity(edge_case):
    mols, b, cn, se, dse = edge_case
    _say(f"[gpu] edge forward nmol={b.nmol} nat={b.nat}")
    S, D, Q, H = b.forward(b.tensor_xyz(), se)
    torch.cuda.synchronize()
    views = [b.molecule_views(t, c) for t, c in ((S, 1), (D, 3), (Q, 6), (H, 1))]
    se_h = se.cpu().numpy()
    for m, mol in enumerate(mols):
        st = structure.setup(structure.Structure(*mol))
        shs = slice(int(b.sh_off[m]), int(b.sh_off[m + 1]))
        ref = O.get_hamiltonian(st, se_h[shs])
        for r, v in zip(ref, (vw[m] for vw in views)):
            g = np.transpose(v.cpu().numpy(), (2, 1, 0))
            assert np.abs(r.reshape(g.shape, order="F") - g).max() < TOL_INT, m
    # atoms beyond the cutoff: no off-site block at all
    Sf = views[0][2].cpu().numpy()[:, :, 0]
    assert np.all(Sf[0, 1:] == 0.0) and np.all(Sf[1:, 0] == 0.0)


def test_edge_gradient_and_autograd(edge_case):
    from test_gpu_batch import _plain_adjoints, _random_inputs
    from tblite_amd.batch import h0_integrals
    mols, b, cn, se, dse = edge_case
    _say("[gpu] edge plain gradient and autograd")
    P, W, vao, vdp, vqp = _random_inputs(b, 41)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    dEdcn, grad, sigma = b.gradient(b.tensor_xyz(), se, dse, t(P), t(W), t(vao), t(vdp), t(vqp))
    torch.cuda.synchronize()
    grad_h, dEdcn_h, sigma_h = grad.cpu().numpy(), dEdcn.cpu().numpy(), sigma.cpu().numpy()
    se_h, dse_h = se.cpu().numpy(), dse.cpu().numpy()
    for m, mol in enumerate(mols):
        st = structure.setup(structure.Structure(*mol))
        ats = slice(int(b.atom_off[m]), int(b.atom_off[m + 1]))
        shs = slice(int(b.sh_off[m]), int(b.sh_off[m + 1]))
        n = st.bas.nao
        a = int(b.dense_off[m])
        ao = slice(int(b.ao_off[m]), int(b.ao_off[m + 1]))
        cr, gr, sr = O.get_hamiltonian_gradient(
            st, se_h[shs], dse_h[shs], vao[ao].reshape(n, 1, order="F"),
            vdp[3 * ats.start: 3 * ats.stop].reshape(3, -1, 1, order="F"),
            vqp[6 * ats.start: 6 * ats.stop].reshape(6, -1, 1, order="F"),
            P[a: a + n * n].reshape(n, n, 1, order="F"), W[a: a + n * n].reshape(n, n, 1, order="F"))
        assert np.abs(grad_h[3 * ats.start: 3 * ats.stop].reshape(-1, 3).T - gr).max() < TOL_GRAD
        assert np.abs(dEdcn_h[ats] - cr).max() < TOL_GRAD
        assert np.abs(sigma_h[9 * m: 9 * m + 9].reshape(3, 3, order="F") - sr).max() < TOL_GRAD
    # the lone atoms and the pair beyond the cutoff feel no force
    assert np.all(grad_h[: 3 * 4] == 0.0)
    # autograd through the adjoint kernel equals the plain gradient
    Sb, Hb, Db, Qb = (t(x) for x in _plain_adjoints(b, P, W, vao, vdp, vqp))
    xyz = b.tensor_xyz().requires_grad_(True)
    out = h0_integrals(b, xyz, se.detach())
    torch.autograd.backward(out, (Sb, Db, Qb, Hb))
    assert float((xyz.grad - grad).abs().max()) < TOL_GRAD


@pytest.mark.parametrize("k", [0, 2])
def test_edge_dropin(k):
    """Drop-in path for a lone atom and for a pair beyond the cutoff."""
    z, x = _edge_molecules()[k]
    st = structure.setup(structure.Structure(z, x))
    _say(f"[gpu] edge drop-in nat={st.mol.nat}")
    se, _ = st.selfenergy(np.full(st.mol.nat, 1.5))
    for r, g in zip(O.get_hamiltonian(st, se), h0.get_hamiltonian(st, se)):
        assert np.abs(r - g).max() < TOL_INT
    for r, g in zip(O.get_multipole_integrals_lat(st), h0.get_multipole_integrals(st)):
        assert np.abs(r - g).max() < TOL_INT


def test_misaligned_multipole_output_is_rejected(edge_case):
    """The forward stores dipoles and quadrupoles with 16-byte accesses: an
    output that is not 16-byte aligned is an argument error, and nothing is
    launched."""
    from tblite_amd import lib as L
    mols, b, cn, se, dse = edge_case
    dev = b.torch_device
    S = torch.empty(b.ndense, dtype=torch.float64, device=dev)
    H = torch.empty_like(S)
    Dbuf = torch.empty(3 * b.ndense + 1, dtype=torch.float64, device=dev)
    Q = torch.empty(6 * b.ndense, dtype=torch.float64, device=dev)
    D = Dbuf[1:]
    assert D.data_ptr() % 16 == 8
    with pytest.raises(L.TbxError):
        b.forward(b.tensor_xyz(), se, out=(S, D, Q, H))


def test_batch_of_lone_atoms_only():
    """A batch of lone atoms has no gradient work units at all: the adjoint is
    its on-site part alone.  Forward against the oracle; autograd gives zero
    dE/dR, and dE/dse_s = 1/2 sum over the on-site elements (mu, nu) of
    Hbar * S * ([shell(mu) == s] + [shell(nu) == s]), with the oracle's S."""
    from tblite_amd.batch import Batch, h0_integrals
    mols = [(np.array([1]), np.zeros((1, 3))), (np.array([16]), np.array([[1.0, 2.0, 3.0]])),
            (np.array([8]), np.array([[-4.0, 0.5, 0.0]])), (np.array([17]), np.zeros((1, 3)))]
    b = Batch(mols)
    _say(f"[gpu] lone atoms only nmol={b.nmol}")
    rng = np.random.default_rng(9)
    cn = torch.tensor(rng.uniform(0.5, 3.5, b.nat), dtype=torch.float64, device="cuda")
    se = b.selfenergy(cn)[0].contiguous()
    t = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    Sb, Hb = t(rng.standard_normal(b.ndense)), t(rng.standard_normal(b.ndense))
    Db, Qb = t(rng.standard_normal(3 * b.ndense)), t(rng.standard_normal(6 * b.ndense))
    xyz = b.tensor_xyz().requires_grad_(True)
    s = se.detach().requires_grad_(True)
    out = h0_integrals(b, xyz, s)
    torch.autograd.backward(out, (Sb, Db, Qb, Hb))
    torch.cuda.synchronize()
    assert float(xyz.grad.abs().max()) == 0.0
    se_h, Hb_h, dse = se.cpu().numpy(), Hb.cpu().numpy(), s.grad.cpu().numpy()
    S = out[0].detach().cpu().numpy()
    for m, mol in enumerate(mols):
        st = structure.setup(structure.Structure(*mol))
        shs = slice(int(b.sh_off[m]), int(b.sh_off[m + 1]))
        ref = O.get_hamiltonian(st, se_h[shs])
        n = st.bas.nao
        a = int(b.dense_off[m])
        assert np.abs(ref[0].reshape(-1, order="F") - S[a: a + n * n]).max() < TOL_INT
        # shell of every AO of the atom, in the batch's shell numbering
        ao_sh = np.concatenate([np.full(2 * c.ang + 1, k) for k, c in
                                enumerate(b.spec.basis_cgto[b.ids[m]])])
        w = 0.5 * Hb_h[a: a + n * n].reshape(n, n, order="F") * ref[0]
        want = np.array([w[ao_sh == k, :].sum() + w[:, ao_sh == k].sum()
                         for k in range(shs.stop - shs.start)])
        assert np.abs(dse[shs] - want).max() < TOL_GRAD
