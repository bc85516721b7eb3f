"""Density matrices of the SCF on the device, molecule by molecule against the
oracle (scf/diag.f90:62-166, wavefunction/fermi.f90:31-117): the generalized
eigenproblem of the batch's own H0 and S (rocSOLVER dsygvd), the Fermi
filling, and the FP64-MFMA products P = C diag(f) C^T and W = C diag(f e) C^T
(tiles of 32 x 32 with ragged edges: molecules of 1 to ~300 orbitals).

Tolerances: eigenvalues 1e-10, occupations 1e-10, densities 1e-10."""
import sys

import numpy as np
import pytest

from oracle import scf as OS
from tblite_amd import synthetic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

KT = 300.0 * 3.166808578545117e-6  # test/unit/test_gfn2_xtb.f90:37


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dens_case():
    from tblite_amd.batch import Batch
    rng = np.random.default_rng(1008)
    mols = [synthetic.random_molecule(int(k), rng) for k in rng.integers(3, 110, size=10)]
    mols.append((np.array([1]), np.zeros((1, 3))))
    mols.append(synthetic.mixed_cluster(40, seed=6, density=1.0))
    b = Batch(mols)
    cn = torch.tensor(rng.uniform(0.5, 3.5, b.nat), dtype=torch.float64, device="cuda")
    se, _ = b.selfenergy(cn)
    S, _, _, H = b.forward(b.tensor_xyz(), se.contiguous())
    torch.cuda.synchronize()
    nao_m = np.diff(b.ao_off)
    _say(f"[gpu] density: nmol={b.nmol} nao per molecule {nao_m.min()}..{nao_m.max()}")
    return b, S, H, rng


def _block(t, b, m):
    n = int(b.ao_off[m + 1] - b.ao_off[m])
    d0, d1 = int(b.dense_off[m]), int(b.dense_off[m + 1])
    return t[d0:d1].cpu().numpy().reshape(n, n, order="F")


def test_sygvd_fermi_and_densities(dens_case):
    b, S, H, rng = dens_case
    C = H.clone()
    e = b.sygvd(C, S)
    torch.cuda.synchronize()
    e_h = e.cpu().numpy()
    nel = np.array([rng.uniform(0.3, 0.9) * int(b.ao_off[m + 1] - b.ao_off[m])
                    for m in range(b.nmol)])
    focc, ef = b.fermi_filling(torch.tensor(nel, dtype=torch.float64, device="cuda"), KT, e)
    P = b.density_matrix(C, focc)
    W = b.wdensity(C, focc, e)
    torch.cuda.synchronize()
    focc_h, ef_h = focc.cpu().numpy(), ef.cpu().numpy()
    worst = {"eval": 0.0, "focc": 0.0, "P": 0.0, "W": 0.0}
    for m in range(b.nmol):
        a0, a1 = int(b.ao_off[m]), int(b.ao_off[m + 1])
        Sm, Hm, Cm = _block(S, b, m), _block(H, b, m), _block(C, b, m)
        er, cr = OS.sygvd(Hm, Sm)
        err_e = np.abs(e_h[a0:a1] - er).max() / max(1.0, np.abs(er).max())
        # device eigenvectors are S-orthonormal eigenvectors
        n = a1 - a0
        assert np.abs(Cm.T @ Sm @ Cm - np.eye(n)).max() < 1e-9, m
        assert np.abs(Hm @ Cm - Sm @ Cm * e_h[a0:a1][None, :]).max() < 1e-9, m
        homo, fr, efr = OS.get_fermi_filling(nel[m], KT, e_h[a0:a1])
        err_f = max(np.abs(focc_h[a0:a1] - fr).max(), abs(ef_h[m] - efr))
        # the MFMA products against the oracle's on the same C, occupations, energies
        Pr = OS.get_density_matrix(focc_h[a0:a1], Cm)
        Wr = OS.get_wdensity(focc_h[a0:a1], e_h[a0:a1], Cm)
        err_p = np.abs(_block(P, b, m) - Pr).max()
        err_w = np.abs(_block(W, b, m) - Wr).max()
        # and the density is the oracle's own from its own eigenvectors
        err_p2 = np.abs(_block(P, b, m) - OS.get_density_matrix(fr, cr)).max()
        for k, v in (("eval", err_e), ("focc", err_f), ("P", max(err_p, err_p2)), ("W", err_w)):
            worst[k] = max(worst[k], v)
        assert err_e < 1e-10 and err_f < 1e-10, (m, err_e, err_f)
        assert err_p < 1e-10 and err_w < 1e-10 and err_p2 < 1e-8, (m, err_p, err_w, err_p2)
    _say("    max |GPU - oracle|: " + ", ".join(f"{k} {v:.1e}" for k, v in worst.items()))
