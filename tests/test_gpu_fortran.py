"""The Fortran iso_c_binding module (tblite_amd/fortran/tbx_h0.f90), driven by
a Fortran host program, against the oracle: the drop-in boundary exercised
from Fortran exactly as tblite's xtb_singlepoint would call it (tblite's
1-based adjacency list, Fortran-ordered dense outputs, accumulated gradient),
with the tables packed by tbx_h0_pack_tables on the default context, and the
adjoint of get_hamiltonian against the oracle's."""
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tblite_amd import structure, synthetic

from fortran_io import DRIVER, read_output, write_input

pytestmark = pytest.mark.gpu


def _setup(name):
    if name == "octane":
        z, x = synthetic.octane()
    else:  # S/Cl d shells
        z, x = np.array([16, 16, 17, 1]), np.array([[0, 0, 0], [0.3, 0.2, 3.7],
                                                    [3.1, -1.0, 1.2], [-1.9, 1.1, -0.4]])
    return structure.setup(structure.Structure(z, x))


@pytest.mark.parametrize("name", ["octane", "sulfur_dimer"])
def test_fortran_driver_parity(name, tmp_path):
    st = _setup(name)
    print(f"[gpu] fortran driver {name} nat={st.mol.nat} nao={st.bas.nao}", flush=True)
    rng = np.random.default_rng(21)
    se, dse = st.selfenergy(rng.uniform(0.5, 3.5, st.mol.nat))
    nao, nat = st.bas.nao, st.mol.nat
    g = rng.normal(size=(nao, nao))
    P = np.asfortranarray((0.5 * (g + g.T))[:, :, None])
    W = np.asfortranarray(0.1 * P)
    vao = np.asfortranarray(rng.normal(scale=1e-2, size=(nao, 1)))
    vdp = np.asfortranarray(rng.normal(scale=1e-3, size=(3, nat, 1)))
    vqp = np.asfortranarray(rng.normal(scale=1e-3, size=(6, nat, 1)))
    bars = (rng.normal(size=(nao, nao)), rng.normal(size=(nao, nao)),
            rng.normal(size=(3, nao, nao)), rng.normal(size=(6, nao, nao)))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, st, se, dse, vao, vdp, vqp, P, W, bars)
    r = subprocess.run([str(DRIVER), str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    stat, (S, D, Q, H, dEdcn, grad, sigma, S2, D2, Q2, ag, adse, asig) = read_output(fout, st)
    assert list(stat) == [0, 0, 0, 0]
    for ref, got, lab in zip(O.get_hamiltonian(st, se), (S, D, Q, H), "SDQH"):
        assert np.abs(ref - got).max() < 1e-10, lab
    cr, gr, sr = O.get_hamiltonian_gradient(st, se, dse, vao, vdp, vqp, P, W)
    assert np.abs(cr - dEdcn).max() < 1e-8
    assert np.abs(gr - grad).max() < 1e-8
    assert np.abs(sr - sigma).max() < 1e-8
    # tbx_h0_get_multipole_integrals against the oracle's full double loop
    for ref, got in zip(O.get_multipole_integrals_lat(st), (S2, D2, Q2)):
        assert np.abs(ref - got).max() < 1e-10
    # tbx_h0_get_hamiltonian_adjoint against the oracle's adjoint
    rg, rdse, rsig = O.get_hamiltonian_adjoint(st, se, *[np.asfortranarray(b) for b in bars])
    assert np.abs(rg - ag).max() < 1e-8
    assert np.abs(rdse - adse).max() < 1e-8
    assert np.abs(rsig - asig).max() < 1e-8
