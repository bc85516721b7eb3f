"""Full-size configurations beyond the 2048-molecule benchmark batch:

- (H2O)512, 1536 atoms (plain GPU form): drop-in forward and gradient against
  the oracle at full size;
- 4000-atom molecular cluster with S/Cl d shells (plain GPU form): size-
  independent identities at full size (symmetry of S and H0, the dipole shift
  rule between the two triangles, traceless Q, translation invariance of the
  gradient, symmetric sigma) and oracle parity on a 500-atom sub-cluster;
- the differentiation set: 512 molecules of 20-200 atoms plus a 999-atom water
  cluster in one batch; autograd against central differences and against the
  plain get_hamiltonian_gradient kernel with the functional's adjoints.

Tolerances: integrals and H0 1e-10, gradients 1e-8 (absolute)."""
import sys
import time

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tblite_amd import h0, structure, synthetic
from tblite_amd.batch import Batch, h0_integrals

pytestmark = pytest.mark.gpu

TOL_INT = 1e-10
TOL_GRAD = 1e-8


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _density(st, seed):
    rng = np.random.default_rng(seed)
    nao, nat = st.bas.nao, st.mol.nat
    g = rng.normal(size=(nao, nao))
    P = np.asfortranarray((0.5 * (g + g.T))[:, :, None])
    W = np.asfortranarray(0.05 * P)
    vao = np.asfortranarray(rng.normal(scale=1e-2, size=(nao, 1)))
    vdp = np.asfortranarray(rng.normal(scale=1e-3, size=(3, nat, 1)))
    vqp = np.asfortranarray(rng.normal(scale=1e-3, size=(6, nat, 1)))
    return P, W, vao, vdp, vqp


def test_water512_full_size_parity():
    z, x = synthetic.water_cluster(512, seed=1002)
    st = structure.setup(structure.Structure(z, x))
    _say(f"[gpu] (H2O)512 nat={st.mol.nat} nao={st.bas.nao}")
    rng = np.random.default_rng(3)
    se, dse = st.selfenergy(rng.uniform(0.5, 3.5, st.mol.nat))
    t = time.perf_counter()
    got = h0.get_hamiltonian(st, se)
    _say(f"    drop-in get_hamiltonian {time.perf_counter() - t:.3f} s (PCIe included)")
    ref = O.get_hamiltonian(st, se)
    for r, g, lab in zip(ref, got, "SDQH"):
        assert np.abs(r - g).max() < TOL_INT, lab
    del ref, got
    P, W, vao, vdp, vqp = _density(st, 5)
    t = time.perf_counter()
    cg, gg, sg = h0.get_hamiltonian_gradient(st, se, dse, vao, vdp, vqp, P, W, *_zeros(st))
    _say(f"    drop-in get_hamiltonian_gradient {time.perf_counter() - t:.3f} s")
    cr, gr, sr = O.get_hamiltonian_gradient(st, se, dse, vao, vdp, vqp, P, W)
    assert np.abs(gr - gg).max() < TOL_GRAD
    assert np.abs(cr - cg).max() < TOL_GRAD
    assert np.abs(sr - sg).max() < TOL_GRAD


def _zeros(st):
    """Accumulated outputs dEdcn(nat), gradient(3,nat), sigma(3,3), starting at zero."""
    nat = st.mol.nat
    return np.zeros(nat), np.zeros((3, nat), order="F"), np.zeros((3, 3), order="F")


def _ao_atoms(st):
    nao_sp = np.array([sum(2 * c.ang + 1 for c in sh) for sh in st.spec.basis_cgto])
    return np.repeat(np.arange(st.mol.nat), nao_sp[st.mol.id])


def test_cluster4000_full_size_identities_and_subcluster_parity():
    z, x = synthetic.mixed_cluster(4000, seed=1003)
    st = structure.setup(structure.Structure(z, x))
    nat, nao = st.mol.nat, st.bas.nao
    _say(f"[gpu] 4000-atom cluster nat={nat} nao={nao} d-shell atoms="
         f"{int(np.isin(z, [16, 17]).sum())}")
    rng = np.random.default_rng(7)
    se, dse = st.selfenergy(rng.uniform(0.5, 3.5, nat))
    t = time.perf_counter()
    S, D, Q, H = h0.get_hamiltonian(st, se)
    _say(f"    drop-in get_hamiltonian {time.perf_counter() - t:.3f} s (PCIe included)")
    assert np.abs(S - S.T).max() < TOL_INT and np.abs(H - H.T).max() < TOL_INT
    # operator on the column atom: D(k,mu,nu) - D(k,nu,mu) = (R_at(mu) - R_at(nu))_k S(mu,nu)
    R = np.asarray(st.mol.xyz).reshape(-1, 3)[_ao_atoms(st)]
    for k in range(3):
        lhs = D[k] - D[k].T
        rhs = (R[:, k][:, None] - R[:, k][None, :]) * S
        assert np.abs(lhs - rhs).max() < TOL_INT, k
    assert np.abs(Q[0] + Q[2] + Q[5]).max() < TOL_INT
    del S, D, Q, H, R
    P, W, vao, vdp, vqp = _density(st, 9)
    t = time.perf_counter()
    cg, gg, sg = h0.get_hamiltonian_gradient(st, se, dse, vao, vdp, vqp, P, W, *_zeros(st))
    _say(f"    drop-in get_hamiltonian_gradient {time.perf_counter() - t:.3f} s")
    assert np.abs(gg.sum(axis=1)).max() < 1e-8 * max(1.0, np.abs(gg).max())
    assert np.abs(sg - sg.T).max() < TOL_GRAD
    del P, W
    # oracle parity on a 500-atom sub-cluster around the centre
    xyz = np.asarray(x).reshape(-1, 3)
    keep = np.argsort(((xyz - xyz.mean(0)) ** 2).sum(1))[:500]
    keep.sort()
    sub = structure.setup(structure.Structure(np.asarray(z)[keep], xyz[keep]))
    se_s, dse_s = sub.selfenergy(rng.uniform(0.5, 3.5, sub.mol.nat))
    for r, g, lab in zip(O.get_hamiltonian(sub, se_s), h0.get_hamiltonian(sub, se_s), "SDQH"):
        assert np.abs(r - g).max() < TOL_INT, lab
    P, W, vao, vdp, vqp = _density(sub, 11)
    ref = O.get_hamiltonian_gradient(sub, se_s, dse_s, vao, vdp, vqp, P, W)
    got = h0.get_hamiltonian_gradient(sub, se_s, dse_s, vao, vdp, vqp, P, W, *_zeros(sub))
    for r, g in zip(ref, got):
        assert np.abs(r - g).max() < TOL_GRAD


@pytest.fixture(scope="module")
def diff_batch():
    mols = synthetic.molecule_batch(512, 20, 200, seed=1005, workers=8)
    zw, xw = synthetic.water_cluster(333, seed=1005)
    mols.append((zw, xw))
    b = Batch(mols)
    rng = np.random.default_rng(1005)
    cn = torch.tensor(rng.uniform(0.5, 3.5, b.nat), dtype=torch.float64, device="cuda")
    return mols, b, cn


def test_diffset_autograd_finite_differences(diff_batch):
    """E = <Sb,S> + <Db,D> + <Qb,Q> + <Hb,H> with symmetrised standard-normal
    adjoints; autograd (adjoint kernel) vs central differences, h = 1e-4 bohr,
    on coordinates of small molecules and of the 999-atom water cluster."""
    mols, b, cn = diff_batch
    _say(f"[gpu] differentiation set: {b.nmol} molecules, nat={b.nat}, ndense={b.ndense}")
    rng = np.random.default_rng(77)

    def sym_adjoint(comps):
        out = torch.empty(comps * b.ndense, dtype=torch.float64, device="cuda")
        for v in b.molecule_views(out, comps):
            g = torch.tensor(rng.normal(size=tuple(v.shape)), dtype=torch.float64, device="cuda")
            v.copy_(0.5 * (g + g.transpose(0, 1)))
        return out
    Sb, Db, Qb, Hb = sym_adjoint(1), sym_adjoint(3), sym_adjoint(6), sym_adjoint(1)
    xyz0 = b.tensor_xyz()
    se0, _ = b.selfenergy(cn)
    x = xyz0.clone().requires_grad_(True)
    cnv = cn.clone().requires_grad_(True)
    se, _ = b.selfenergy(cnv)
    S, D, Q, H = h0_integrals(b, x, se)
    E = (Sb * S).sum() + (Db * D).sum() + (Qb * Q).sum() + (Hb * H).sum()
    E.backward()
    gx, gcn = x.grad, cnv.grad

    def energy(xx, cc):
        s, _ = b.selfenergy(cc)
        S, D, Q, H = b.forward(xx, s.contiguous())
        return float((Sb * S).sum() + (Db * D).sum() + (Qb * Q).sum() + (Hb * H).sum())

    h = 1e-4
    w0 = int(b.atom_off[b.nmol - 1])  # first atom of the water cluster
    for idx in [1, 3 * int(b.atom_off[100]) + 2, 3 * (w0 + 10), 3 * (w0 + 500) + 1, 3 * b.nat - 1]:
        xp, xm = xyz0.clone(), xyz0.clone()
        xp[idx] += h
        xm[idx] -= h
        num = (energy(xp, cn) - energy(xm, cn)) / (2 * h)
        ana = gx[idx].item()
        assert abs(num - ana) < 1e-6 * max(1.0, abs(ana)), (idx, num, ana)
    for a in [0, w0 + 3]:
        cp, cm = cn.clone(), cn.clone()
        cp[a] += h
        cm[a] -= h
        num = (energy(xyz0, cp) - energy(xyz0, cm)) / (2 * h)
        assert abs(num - gcn[a].item()) < 1e-6 * max(1.0, abs(num)), (a, num, gcn[a].item())


def test_diffset_autograd_equals_plain_gradient(diff_batch):
    """With the adjoints of the functional get_hamiltonian_gradient
    differentiates, the autograd gradient equals the plain gradient kernel."""
    from test_gpu_batch import _plain_adjoints

    mols, b, cn = diff_batch
    rng = np.random.default_rng(8)
    P = np.zeros(b.ndense)
    for m in range(b.nmol):
        n = int(b.ao_off[m + 1] - b.ao_off[m])
        g = rng.normal(size=(n, n))
        P[int(b.dense_off[m]): int(b.dense_off[m]) + n * n] = (0.5 * (g + g.T)).ravel()
    W = 0.05 * P
    vao = rng.normal(scale=1e-2, size=b.nao)
    vdp = rng.normal(scale=1e-3, size=3 * b.nat)
    vqp = rng.normal(scale=1e-3, size=6 * b.nat)
    Sb, Hb, Db, Qb = _plain_adjoints(b, P, W, vao, vdp, vqp)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    se, dse = b.selfenergy(cn)
    se, dse = se.contiguous(), dse.contiguous()
    xyz = b.tensor_xyz()
    _, grad, _ = b.gradient(xyz, se, dse, dev(P), dev(W), dev(vao), dev(vdp), dev(vqp))
    x = xyz.clone().requires_grad_(True)
    o = h0_integrals(b, x, se)
    torch.autograd.backward(o, (dev(Sb), dev(Db), dev(Qb), dev(Hb)))
    torch.cuda.synchronize()
    err = torch.abs(x.grad - grad).max().item()
    _say(f"    autograd vs plain gradient: max |diff| = {err:.2e}")
    assert err < TOL_GRAD


def test_diffset_autograd_vs_oracle_gradient(diff_batch):
    """Autograd dE/dR with the adjoints of the functional
    get_hamiltonian_gradient differentiates, against the oracle's
    get_hamiltonian_gradient (xtb/h0.f90:432-666): the 999-atom water cluster
    and every one of the 512 molecules."""
    from test_gpu_batch import _plain_adjoints

    mols, b, cn = diff_batch
    rng = np.random.default_rng(9)
    P = np.zeros(b.ndense)
    for m in range(b.nmol):
        n = int(b.ao_off[m + 1] - b.ao_off[m])
        g = rng.normal(size=(n, n))
        P[int(b.dense_off[m]): int(b.dense_off[m]) + n * n] = (0.5 * (g + g.T)).ravel()
    W = 0.05 * P
    vao = rng.normal(scale=1e-2, size=b.nao)
    vdp = rng.normal(scale=1e-3, size=3 * b.nat)
    vqp = rng.normal(scale=1e-3, size=6 * b.nat)
    Sb, Hb, Db, Qb = _plain_adjoints(b, P, W, vao, vdp, vqp)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    se, dse = b.selfenergy(cn)
    x = b.tensor_xyz().requires_grad_(True)
    o = h0_integrals(b, x, se.detach().contiguous())
    torch.autograd.backward(o, (dev(Sb), dev(Db), dev(Qb), dev(Hb)))
    torch.cuda.synchronize()
    gx = x.grad.cpu().numpy()
    se_h, dse_h = se.detach().cpu().numpy(), dse.cpu().numpy()
    worst = 0.0
    for m in range(b.nmol):
        st = structure.setup(structure.Structure(*mols[m]))
        a0, a1 = int(b.atom_off[m]), int(b.atom_off[m + 1])
        s0, s1 = int(b.sh_off[m]), int(b.sh_off[m + 1])
        o0, o1 = int(b.ao_off[m]), int(b.ao_off[m + 1])
        d0, n = int(b.dense_off[m]), o1 - o0
        _, gr, _ = O.get_hamiltonian_gradient(
            st, se_h[s0:s1], dse_h[s0:s1], vao[o0:o1].reshape(n, 1),
            vdp[3 * a0: 3 * a1].reshape(3, -1, 1, order="F"),
            vqp[6 * a0: 6 * a1].reshape(6, -1, 1, order="F"),
            P[d0: d0 + n * n].reshape(n, n, 1, order="F"),
            W[d0: d0 + n * n].reshape(n, n, 1, order="F"))
        err = float(np.abs(gx[3 * a0: 3 * a1].reshape(-1, 3).T - gr).max())
        worst = max(worst, err)
        assert err < TOL_GRAD, (m, err)
    _say(f"    autograd vs oracle gradient: {b.nmol} molecules, max |diff| = {worst:.2e}")


def test_diffset_autograd_vs_oracle_adjoint(diff_batch):
    """Autograd with symmetrised standard-normal adjoints (the FD test's
    functional) against the oracle's adjoint of get_hamiltonian: dE/dR and
    dE/dCN (through get_selfenergy) of every molecule and the water cluster."""
    mols, b, cn = diff_batch
    g = torch.Generator(device="cuda")
    g.manual_seed(1005)

    def sym_adjoint(comps):
        out = torch.empty(comps * b.ndense, dtype=torch.float64, device="cuda")
        for v in b.molecule_views(out, comps):
            r = torch.randn(tuple(v.shape), generator=g, dtype=torch.float64, device="cuda")
            v.copy_(0.5 * (r + r.transpose(0, 1)))
        return out
    Sb, Db, Qb, Hb = sym_adjoint(1), sym_adjoint(3), sym_adjoint(6), sym_adjoint(1)
    x = b.tensor_xyz().requires_grad_(True)
    cnv = cn.clone().requires_grad_(True)
    se, _ = b.selfenergy(cnv)
    S, D, Q, H = h0_integrals(b, x, se)
    E = (Sb * S).sum() + (Db * D).sum() + (Qb * Q).sum() + (Hb * H).sum()
    E.backward()
    torch.cuda.synchronize()
    gx, gcn = x.grad.cpu().numpy(), cnv.grad.cpu().numpy()
    se_h = se.detach().cpu().numpy()
    worst = 0.0
    for m in range(b.nmol):
        st = structure.setup(structure.Structure(*mols[m]))
        a0, a1 = int(b.atom_off[m]), int(b.atom_off[m + 1])
        s0, s1 = int(b.sh_off[m]), int(b.sh_off[m + 1])
        o0, o1 = int(b.ao_off[m]), int(b.ao_off[m + 1])
        d0, d1, n = int(b.dense_off[m]), int(b.dense_off[m + 1]), o1 - o0
        bars = [t[c * d0: c * d1].cpu().numpy().reshape((c, n, n) if c > 1 else (n, n),
                                                         order="F")
                for t, c in ((Sb, 1), (Hb, 1), (Db, 3), (Qb, 6))]
        gr, er, _ = O.get_hamiltonian_adjoint(st, se_h[s0:s1], *bars)
        _, dsedcn = st.selfenergy(np.ones(st.mol.nat))  # -kcn, independent of cn
        sh_at = np.repeat(np.arange(st.mol.nat), st.spec.nsh_id[st.mol.id])
        cr = np.bincount(sh_at, weights=dsedcn * er, minlength=st.mol.nat)
        errs = (np.abs(gx[3 * a0: 3 * a1].reshape(-1, 3).T - gr).max(),
                np.abs(gcn[a0:a1] - cr).max())
        worst = max(worst, *errs)
        assert max(errs) < TOL_GRAD, (m, errs)
    _say(f"    autograd vs oracle adjoint: {b.nmol} molecules, max |diff| = {worst:.2e}")
