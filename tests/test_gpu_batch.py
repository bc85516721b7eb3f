"""Batched device path: forward / plain gradient / adjoint (autograd) parity
with the oracle, finite differences of the adjoint, and HIP-graph replay."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tblite_amd import lib as L, structure, synthetic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL_INT = 1e-10
TOL_GRAD = 1e-8


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _molecules(n=10, seed=1005, lo=8, hi=60):
    rng = np.random.default_rng(seed)
    mols = [synthetic.random_molecule(int(k), rng) for k in rng.integers(lo, hi, size=n)]
    mols.append(synthetic.water_cluster(6, seed=3))
    mols.append(synthetic.mixed_cluster(30, seed=4, density=1.1))
    return mols


@pytest.fixture(scope="module")
def batch_case():
    from tblite_amd.batch import Batch
    mols = _molecules()
    b = Batch(mols)
    rng = np.random.default_rng(17)
    cn = torch.tensor(rng.uniform(0.5, 3.5, b.nat), dtype=torch.float64, device="cuda")
    se, dse = b.selfenergy(cn)
    return mols, b, cn, se.contiguous(), dse.contiguous()


def _setup(mol):
    return structure.setup(structure.Structure(*mol))


def _mol_slice(b, m):
    return slice(int(b.atom_off[m]), int(b.atom_off[m + 1])), \
        slice(int(b.sh_off[m]), int(b.sh_off[m + 1]))


def test_batch_forward_parity(batch_case):
    mols, b, cn, se, dse = batch_case
    _say(f"[gpu] batch forward nmol={b.nmol} nat={b.nat} ndense={b.ndense}")
    xyz = b.tensor_xyz()
    S, D, Q, H = b.forward(xyz, se)
    torch.cuda.synchronize()
    Sv, Dv, Qv, Hv = (b.molecule_views(t, c) for t, c in ((S, 1), (D, 3), (Q, 6), (H, 1)))
    se_h = se.cpu().numpy()
    for m, mol in enumerate(mols):
        st = _setup(mol)
        _, shs = _mol_slice(b, m)
        ref = O.get_hamiltonian(st, se_h[shs])
        for r, v in zip(ref, (Sv[m], Dv[m], Qv[m], Hv[m])):
            g = v.cpu().numpy()  # (nu, mu, k) = Fortran (k, mu, nu)
            g = np.transpose(g, (2, 1, 0))
            assert np.abs(r.reshape(g.shape, order="F") - g).max() < TOL_INT


def _random_inputs(b, seed):
    rng = np.random.default_rng(seed)
    P = np.zeros(b.ndense)
    W = np.zeros(b.ndense)
    for m in range(b.nmol):
        n = int(b.ao_off[m + 1] - b.ao_off[m])
        g = rng.normal(size=(n, n))
        h = rng.normal(size=(n, n))
        a = int(b.dense_off[m])
        P[a: a + n * n] = (0.5 * (g + g.T)).ravel(order="F")
        W[a: a + n * n] = (0.5 * (h + h.T)).ravel(order="F")
    vao = rng.normal(scale=1e-2, size=b.nao)
    vdp = rng.normal(scale=1e-3, size=3 * b.nat)
    vqp = rng.normal(scale=1e-3, size=6 * b.nat)
    return P, W, vao, vdp, vqp


def test_batch_gradient_parity(batch_case):
    mols, b, cn, se, dse = batch_case
    _say("[gpu] batch plain gradient")
    P, W, vao, vdp, vqp = _random_inputs(b, 23)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    dEdcn, grad, sigma = b.gradient(b.tensor_xyz(), se, dse, t(P), t(W), t(vao), t(vdp), t(vqp))
    torch.cuda.synchronize()
    dEdcn, grad, sigma = dEdcn.cpu().numpy(), grad.cpu().numpy(), sigma.cpu().numpy()
    se_h, dse_h = se.cpu().numpy(), dse.cpu().numpy()
    for m, mol in enumerate(mols):
        st = _setup(mol)
        ats, shs = _mol_slice(b, m)
        n = st.bas.nao
        a = int(b.dense_off[m])
        Pm = P[a: a + n * n].reshape(n, n, 1, order="F")
        Wm = W[a: a + n * n].reshape(n, n, 1, order="F")
        ao = slice(int(b.ao_off[m]), int(b.ao_off[m + 1]))
        cr, gr, sr = O.get_hamiltonian_gradient(
            st, se_h[shs], dse_h[shs], vao[ao].reshape(n, 1, order="F"),
            vdp[3 * ats.start: 3 * ats.stop].reshape(3, -1, 1, order="F"),
            vqp[6 * ats.start: 6 * ats.stop].reshape(6, -1, 1, order="F"), Pm, Wm)
        assert np.abs(grad[3 * ats.start: 3 * ats.stop].reshape(-1, 3).T - gr).max() < TOL_GRAD
        assert np.abs(dEdcn[ats] - cr).max() < TOL_GRAD
        assert np.abs(sigma[9 * m: 9 * m + 9].reshape(3, 3, order="F") - sr).max() < TOL_GRAD


def _plain_adjoints(b, P, W, vao, vdp, vqp):
    """Adjoints of S, H, D, Q for which d/dR of <adjoints, outputs> is the
    functional get_hamiltonian_gradient differentiates (xtb/h0.f90:552-567):
    H -> P, S -> -W - P (v_mu + v_nu)/2, D(k,mu,nu) -> -P vdp(k, atom(nu)),
    Q(k,mu,nu) -> -P vqp(k, atom(nu))."""
    Sb = np.zeros(b.ndense)
    Db = np.zeros(3 * b.ndense)
    Qb = np.zeros(6 * b.ndense)
    for m in range(b.nmol):
        n = int(b.ao_off[m + 1] - b.ao_off[m])
        a = int(b.dense_off[m])
        ao = slice(int(b.ao_off[m]), int(b.ao_off[m + 1]))
        Pm = P[a: a + n * n].reshape(n, n, order="F")
        Wm = W[a: a + n * n].reshape(n, n, order="F")
        v = vao[ao]
        Sb[a: a + n * n] = (-Wm - 0.5 * Pm * (v[:, None] + v[None, :])).ravel(order="F")
        ao2at = b.ao_atom[ao]
        vd = vdp.reshape(-1, 3)[ao2at]  # (n, 3), atom of each AO
        vq = vqp.reshape(-1, 6)[ao2at]
        Dm = -(Pm[None, :, :] * vd.T[:, None, :])  # (3, mu, nu), atom of the column nu
        Qm = -(Pm[None, :, :] * vq.T[:, None, :])
        Db[3 * a: 3 * (a + n * n)] = Dm.ravel(order="F")
        Qb[6 * a: 6 * (a + n * n)] = Qm.ravel(order="F")
    return Sb, P.copy(), Db, Qb


def test_autograd_reproduces_plain_gradient(batch_case):
    """torch.autograd through the adjoint kernel with the adjoints of the
    reference's energy functional gives get_hamiltonian_gradient's dE/dR."""
    from tblite_amd.batch import h0_integrals
    mols, b, cn, se, dse = batch_case
    _say("[gpu] autograd vs plain gradient")
    P, W, vao, vdp, vqp = _random_inputs(b, 31)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    Sb, Hb, Db, Qb = (t(x) for x in _plain_adjoints(b, P, W, vao, vdp, vqp))
    xyz = b.tensor_xyz().requires_grad_(True)
    cnv = cn.clone().requires_grad_(True)
    se_v, _ = b.selfenergy(cnv)
    outs = h0_integrals(b, xyz, se_v)
    torch.autograd.backward(outs, (Sb, Db, Qb, Hb))
    _, gref, _ = b.gradient(b.tensor_xyz(), se, dse, t(P), t(W), t(vao), t(vdp), t(vqp))
    torch.cuda.synchronize()
    assert torch.abs(xyz.grad - gref).max().item() < TOL_GRAD
    # dE/dcn: the reference keeps only diag(P) on site (h0.f90:654-663); the
    # exact derivative carries S_mumu (1 - S_mumu < 1e-9) and the off-diagonal
    # on-site overlaps (< 1e-15)
    dref, _, _ = b.gradient(b.tensor_xyz(), se, dse, t(P), t(W), t(vao), t(vdp), t(vqp))
    torch.cuda.synchronize()
    assert torch.abs(cnv.grad - dref).max().item() < 1e-7 * max(1.0, dref.abs().max().item())


def test_adjoint_finite_differences(batch_case):
    """dE/dR and dE/dse of E = <Sb,S> + <Db,D> + <Qb,Q> + <Hb,H> with random
    symmetric adjoints vs central differences (Richardson), a few atoms."""
    mols, b, cn, se, dse = batch_case
    _say("[gpu] adjoint finite differences")
    rng = np.random.default_rng(41)
    Sb = torch.tensor(rng.normal(size=b.ndense), dtype=torch.float64, device="cuda")
    Hb = torch.tensor(rng.normal(size=b.ndense), dtype=torch.float64, device="cuda")
    Db = torch.tensor(rng.normal(size=3 * b.ndense), dtype=torch.float64, device="cuda")
    Qb = torch.tensor(rng.normal(size=6 * b.ndense), dtype=torch.float64, device="cuda")
    xyz0 = b.tensor_xyz()
    dxyz, dse_ = b.adjoint(xyz0, se, Sb, Db, Qb, Hb)

    def energy(x, s):
        S, D, Q, H = b.forward(x, s)
        return float((Sb * S).sum() + (Db * D).sum() + (Qb * Q).sum() + (Hb * H).sum())

    def rich(f, h):
        d1 = (f(h) - f(-h)) / (2 * h)
        d2 = (f(h / 2) - f(-h / 2)) / h
        return (4 * d2 - d1) / 3

    for idx in [0, 5, 3 * int(b.atom_off[3]) + 1, 3 * b.nat - 2]:
        def f(h):
            x = xyz0.clone()
            x[idx] += h
            return energy(x, se)
        num = rich(f, 1e-3)
        assert abs(num - dxyz[idx].item()) < 1e-7 * max(1.0, abs(num)), (idx, num, dxyz[idx].item())
    for idx in [0, 7, b.nsh - 1]:
        def f(h):
            s = se.clone()
            s[idx] += h
            return energy(xyz0, s)
        num = rich(f, 1e-3)
        assert abs(num - dse_[idx].item()) < 1e-7 * max(1.0, abs(num))


def test_graph_replay_matches_eager(batch_case):
    """Forward + gradient + adjoint (given the forward's S and D) captured in
    a HIP graph, replayed several times with eager work in between, every
    replay equal to the eager result."""
    mols, b, cn, se, dse = batch_case
    _say("[gpu] graph replay")
    P, W, vao, vdp, vqp = _random_inputs(b, 5)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    xyz = b.tensor_xyz()
    P_, W_, va, vd, vq = t(P), t(W), t(vao), t(vdp), t(vqp)
    rng = np.random.default_rng(6)
    adj = [t(rng.normal(size=k * b.ndense)) for k in (1, 3, 6, 1)]
    outs_f = b.forward(xyz, se)
    outs_g = b.gradient(xyz, se, dse, P_, W_, va, vd, vq)
    outs_a = b.adjoint(xyz, se, *adj, S=outs_f[0], D=outs_f[1])
    torch.cuda.synchronize()
    outs = outs_f + outs_g + outs_a
    ref = [o.clone() for o in outs]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        b.forward(xyz, se, out=outs_f)
        b.gradient(xyz, se, dse, P_, W_, va, vd, vq, out=outs_g)
        b.adjoint(xyz, se, *adj, out=outs_a, S=outs_f[0], D=outs_f[1])
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        b.forward(xyz, se, out=outs_f)
        b.gradient(xyz, se, dse, P_, W_, va, vd, vq, out=outs_g)
        b.adjoint(xyz, se, *adj, out=outs_a, S=outs_f[0], D=outs_f[1])
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        junk = torch.randn(1000, device="cuda").sum()  # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        del junk
        for o, r in zip(outs, ref):
            assert torch.abs(o - r).max().item() < 1e-12, rep


def test_tile_shapes_keep_results(batch_case):
    """Other work-unit caps, chunk lengths and bands (the autotuner's search
    space) give the same integrals and adjoint gradient; invalid shapes are
    refused."""
    mols, b, cn, se, dse = batch_case
    _say("[gpu] tile shapes")
    rng = np.random.default_rng(5)
    adj = [torch.tensor(rng.normal(size=k * b.ndense), dtype=torch.float64, device="cuda")
           for k in (1, 3, 6, 1)]
    xyz = b.tensor_xyz()
    ref_f = [t.clone() for t in b.forward(xyz, se)]
    ref_a = [t.clone() for t in b.adjoint(xyz, se, *adj)]
    for r, g in zip(ref_a, b.adjoint(xyz, se, *adj, S=ref_f[0], D=ref_f[1])):
        assert torch.abs(r - g).max().item() < 1e-10
    default = b.tiles()
    try:
        for tasks, gtasks, chunk, fb, gb in [(128, 128, 4, 1, 4), (192, 112, 12, 2, 1),
                                             (256, 128, 1, 6, 16), (256, 104, 16, 1, 3)]:
            b.set_tiles(fwd_unit_tasks=tasks, fwd_chunk=chunk, grad_unit_tasks=gtasks,
                        grad_chunk=chunk, fwd_band=fb, grad_band=gb)
            for r, g in zip(ref_f, b.forward(xyz, se)):
                assert torch.abs(r - g).max().item() < 1e-12, (tasks, chunk)
            for r, g in zip(ref_a, b.adjoint(xyz, se, *adj)):
                assert torch.abs(r - g).max().item() < 1e-10, (tasks, chunk)
            for r, g in zip(ref_a, b.adjoint(xyz, se, *adj, S=ref_f[0], D=ref_f[1])):
                assert torch.abs(r - g).max().item() < 1e-10, (tasks, chunk)
        with pytest.raises(L.TbxError):
            b.set_tiles(fwd_unit_tasks=512)
        with pytest.raises(L.TbxError):
            b.set_tiles(grad_unit_tasks=257)
        with pytest.raises(L.TbxError):
            b.set_tiles(grad_chunk=0)
        with pytest.raises(L.TbxError):
            b.set_tiles(grad_band=0)
    finally:
        b.set_tiles(**default)


def test_pair_order_hint_keeps_results(batch_case):
    """The pair-order hint (near pairs first, distant ones in units the
    gradient kernels skip) changes the speed, not the results: the forward,
    the adjoint and the plain gradient agree without a hint, with the true
    geometry and with a misleading one (the molecules stretched threefold, so
    that the units the hint marks distant hold pairs within the cutoff)."""
    mols, b, cn, se, dse = batch_case
    _say("[gpu] pair-order hint")
    rng = np.random.default_rng(23)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    adj = [t(rng.normal(size=k * b.ndense)) for k in (1, 3, 6, 1)]
    P, W, vao, vdp, vqp = (t(a) for a in _random_inputs(b, 29))
    xyz = b.tensor_xyz()

    def run():
        f = [x.clone() for x in b.forward(xyz, se)]
        a = [x.clone() for x in b.adjoint(xyz, se, *adj)]
        # the per-element explicit terms (far units walked by the
        # persistent kernel when the hint is stale)
        a2 = [x.clone() for x in b.adjoint(xyz, se, *adj, S=f[0], D=f[1])]
        g = [x.clone() for x in b.gradient(xyz, se, dse, P, W, vao, vdp, vqp)]
        g2 = [x.clone() for x in b.gradient(xyz, se, dse, P, W, vao, vdp, vqp, S=f[0], D=f[1])]
        torch.cuda.synchronize()
        return f + a + a2 + g + g2

    # hints displaced by less and by more than half the skin (0.25 bohr):
    # the far units are skipped in the first case and walked in the second
    d = rng.normal(size=b.xyz0.shape)
    d /= np.linalg.norm(d, axis=1)[:, None]
    try:
        b.order_pairs(None)
        ref = run()
        for hint in (b.xyz0, 3.0 * b.xyz0, b.xyz0 + 0.2 * d, b.xyz0 + 0.3 * d):
            b.order_pairs(hint, skin=0.5)
            for k, (r, g) in enumerate(zip(ref, run())):
                assert torch.abs(r - g).max().item() < 1e-10, k
    finally:
        b.order_pairs(b.xyz0)


def test_autotuner_choice_and_cache(batch_case, tmp_path):
    """tblite_amd.tune over a reduced search space (caps, chunks, then bands):
    the chosen tiles come from that space, every candidate is in the table,
    the chosen tiles leave the forward and the adjoint unchanged, and a
    second call with the same cache file applies the stored tiles without
    timing anything."""
    from tblite_amd import tune
    mols, b, cn, se, dse = batch_case
    _say("[gpu] autotuner")
    rng = np.random.default_rng(8)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    adj = [t(rng.normal(size=k * b.ndense)) for k in (1, 3, 6, 1)]
    P, W, vao, vdp, vqp = (t(a) for a in _random_inputs(b, 9))
    xyz = b.tensor_xyz()
    default = b.tiles()
    ref_f = [x.clone() for x in b.forward(xyz, se)]
    ref_a = [x.clone() for x in b.adjoint(xyz, se, *adj, S=ref_f[0], D=ref_f[1])]
    cache = tmp_path / "tiles.json"
    space = dict(fwd_unit_tasks=(192, 256), grad_unit_tasks=(128, 256), chunks=(1, 3),
                 fwd_bands=(1, 2), grad_bands=(2, 4))
    try:
        res = tune.tune(b, xyz, se, adj, (dse, P, W, vao, vdp, vqp), reps=1, cache=cache,
                        **space)
        tl = res["tiles"]
        assert tl["fwd_unit_tasks"] in (192, 256) and tl["grad_unit_tasks"] in (128, 256)
        assert tl["fwd_chunk"] in (1, 3) and tl["grad_chunk"] in (1, 3)
        assert tl["fwd_band"] in (1, 2) and tl["grad_band"] in (2, 4)
        assert len(res["table"]) == 2 * 2 + 2  # caps x chunks, then the bands
        assert b.tiles() == tl
        S, D, Q, H = b.forward(xyz, se)
        for r, g in zip(ref_f, (S, D, Q, H)):
            assert torch.abs(r - g).max().item() < 1e-12
        for r, g in zip(ref_a, b.adjoint(xyz, se, *adj, S=S, D=D)):
            assert torch.abs(r - g).max().item() < 1e-10
        # the cached choice comes back without a sweep
        b.set_tiles(**default)
        again = tune.tune(b, xyz, se, adj, None, reps=1, cache=cache, **space)
        assert again["tiles"] == tl and b.tiles() == tl
    finally:
        b.set_tiles(**default)
