"""SCF-loop consumers of the resident integrals on the batch (add_vao_to_h1,
add_vmp_to_h1, Mulliken shell charges and atomic multipoles): parity of the
gfx950 kernels, through the C-ABI, with the oracle (oracle/scf.py) on every
molecule of a mixed batch, one and two spin channels.  Tolerance 1e-12
(FP64 sums of the same terms in a different order)."""
import sys

import numpy as np
import pytest

from oracle import scf as OS
from tblite_amd import structure, synthetic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scf_case():
    from tblite_amd.batch import Batch
    rng = np.random.default_rng(1006)
    mols = [synthetic.random_molecule(int(k), rng) for k in rng.integers(4, 50, size=7)]
    mols.append(synthetic.water_cluster(14, seed=5))  # nao > 64: several 32 x 32 tiles
    mols.append((np.array([16, 16, 17, 1]), np.array([[0, 0, 0], [0.3, 0.2, 3.7],
                                                       [3.1, -1.0, 1.2], [-1.9, 1.1, -0.4]])))
    b = Batch(mols)
    xyz = b.tensor_xyz()
    se = torch.zeros(b.nsh, dtype=torch.float64, device="cuda")
    S, D, Q, _ = b.forward(xyz, se)
    torch.cuda.synchronize()
    return mols, b, S, D, Q


def _fortran(flat, n, comps):
    """Packed block (element (r, c) at c*n + r, components innermost) ->
    numpy (comps, n, n) / (n, n) in the reference's index order."""
    a = flat.reshape(n, n, comps).transpose(2, 1, 0)
    return a[0] if comps == 1 else a


def _dense_spin(b, rng, nspin, sym=True):
    out = np.zeros(nspin * b.ndense)
    for s in range(nspin):
        for m in range(b.nmol):
            n = int(b.ao_off[m + 1] - b.ao_off[m])
            a = rng.normal(size=(n, n))
            if sym:
                a = 0.5 * (a + a.T)
            o = s * b.ndense + int(b.dense_off[m])
            out[o: o + n * n] = a.T.reshape(-1)  # column-major block
    return out


def _block(arr, b, m, s=0, comps=1):
    n = int(b.ao_off[m + 1] - b.ao_off[m])
    a, e = int(b.dense_off[m]), int(b.dense_off[m + 1])
    o = s * comps * b.ndense
    return _fortran(arr[o + comps * a: o + comps * e], n, comps)


@pytest.mark.parametrize("nspin", [1, 2])
def test_add_vao_to_h1_parity(scf_case, nspin):
    mols, b, S, D, Q = scf_case
    _say(f"[gpu] add_vao_to_h1 nspin={nspin} nmol={b.nmol} ndense={b.ndense}")
    rng = np.random.default_rng(10 + nspin)
    h0 = _dense_spin(b, rng, nspin, sym=False)
    vao = rng.normal(size=nspin * b.nao)
    h1 = torch.tensor(h0, device="cuda")
    b.add_vao_to_h1(S, torch.tensor(vao, device="cuda"), h1, nspin)
    got = h1.cpu().numpy()
    Sh = S.cpu().numpy()
    for m in range(b.nmol):
        n = int(b.ao_off[m + 1] - b.ao_off[m])
        a0 = int(b.ao_off[m])
        ref = np.stack([_block(h0, b, m, s) for s in range(nspin)], axis=2)
        v = np.stack([vao[s * b.nao + a0: s * b.nao + a0 + n] for s in range(nspin)], axis=1)
        OS.add_vao_to_h1(_block(Sh, b, m), v, ref)
        for s in range(nspin):
            assert np.abs(_block(got, b, m, s) - ref[:, :, s]).max() < TOL


@pytest.mark.parametrize("nspin,which,vdim", [(1, "D", 3), (2, "Q", 6), (1, "Q", 3)])
def test_add_vmp_to_h1_parity(scf_case, nspin, which, vdim):
    mols, b, S, D, Q = scf_case
    mp = D if which == "D" else Q
    mpdim = 3 if which == "D" else 6
    _say(f"[gpu] add_vmp_to_h1 nspin={nspin} mp={which} vdim={vdim}")
    rng = np.random.default_rng(20 + nspin + vdim)
    h0 = _dense_spin(b, rng, nspin, sym=False)
    vmp = rng.normal(size=nspin * b.nat * vdim)
    h1 = torch.tensor(h0, device="cuda")
    b.add_vmp_to_h1(mp, mpdim, torch.tensor(vmp, device="cuda"), vdim, h1, nspin)
    got = h1.cpu().numpy()
    mph = mp.cpu().numpy()
    for m, mol in enumerate(mols):
        st = structure.setup(structure.Structure(*mol))
        at0, at1 = int(b.atom_off[m]), int(b.atom_off[m + 1])
        ref = np.stack([_block(h0, b, m, s) for s in range(nspin)], axis=2)
        v = np.stack([vmp[s * b.nat * vdim:(s + 1) * b.nat * vdim].reshape(b.nat, vdim).T[:, at0:at1]
                      for s in range(nspin)], axis=2)
        OS.add_vmp_to_h1(_block(mph, b, m, 0, mpdim), v, st.bas.ao2at, ref)
        for s in range(nspin):
            assert np.abs(_block(got, b, m, s) - ref[:, :, s]).max() < TOL


@pytest.mark.parametrize("nspin", [1, 2])
def test_mulliken_parity(scf_case, nspin):
    mols, b, S, D, Q = scf_case
    _say(f"[gpu] mulliken nspin={nspin}")
    rng = np.random.default_rng(30 + nspin)
    P = _dense_spin(b, rng, nspin)
    n0 = rng.uniform(0, 2, b.nsh)
    Pd = torch.tensor(P, device="cuda")
    qsh = b.mulliken_shell_charges(S, Pd, torch.tensor(n0, device="cuda"), nspin).cpu().numpy()
    mps = {3: b.mulliken_atomic_multipoles(D, 3, Pd, nspin).cpu().numpy(),
           6: b.mulliken_atomic_multipoles(Q, 6, Pd, nspin).cpu().numpy()}
    Sh, Dh, Qh = S.cpu().numpy(), D.cpu().numpy(), Q.cpu().numpy()
    for m, mol in enumerate(mols):
        st = structure.setup(structure.Structure(*mol))
        bas = st.bas
        at0, at1 = int(b.atom_off[m]), int(b.atom_off[m + 1])
        sh0, sh1 = int(b.sh_off[m]), int(b.sh_off[m + 1])
        Pm = np.stack([_block(P, b, m, s) for s in range(nspin)], axis=2)
        ref = OS.get_mulliken_shell_charges(_block(Sh, b, m), Pm, n0[sh0:sh1], bas.iao_sh,
                                            bas.nao_sh)
        for s in range(nspin):
            assert np.abs(qsh[s * b.nsh + sh0: s * b.nsh + sh1] - ref[:, s]).max() < TOL
        first = bas.iao_sh[bas.ish_at]
        nao_at = np.bincount(bas.ao2at, minlength=len(bas.ish_at))
        for mpdim, arr in ((3, Dh), (6, Qh)):
            ref = OS.get_mulliken_atomic_multipoles(_block(arr, b, m, 0, mpdim), Pm, first, nao_at)
            for s in range(nspin):
                g = mps[mpdim][s * b.nat * mpdim:(s + 1) * b.nat * mpdim].reshape(b.nat, mpdim)
                assert np.abs(g[at0:at1].T - ref[:, :, s]).max() < TOL
