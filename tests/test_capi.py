"""C-ABI checks that need no GPU: the library loads, exports every function
``include/tbx.h`` declares, host-only entry points match the oracle, and
argument errors come back as status codes (no aborts)."""
import ctypes as ct

import numpy as np
import pytest

from oracle import oracle as O
from tblite_amd import h0, lib as L, structure, synthetic

TBX_ERR_ARG, TBX_ERR_BASIS = 1, 2  # include/tbx.h


def test_library_exports_header_symbols():
    names = L.header_functions()
    assert "tbx_get_hamiltonian" in names and "tbx_batch_adjoint" in names
    so = L.lib()
    for n in names:
        assert hasattr(so, n), n


def test_selfenergy_matches_oracle():
    z, x = synthetic.octane()
    st = structure.setup(structure.Structure(z, x))
    cn = np.linspace(0.5, 3.5, st.mol.nat)
    se, dse = h0.get_selfenergy(st, cn)
    se_ref, dse_ref = st.selfenergy(cn)
    assert np.array_equal(se, se_ref) and np.array_equal(dse, dse_ref)
    # oracle restatement of get_selfenergy
    nsh = st.bas.nsh
    out, dout = np.zeros(nsh), np.zeros(nsh)
    p = lambda a: a.ctypes.data_as(ct.c_void_p)  # noqa: E731
    ids = np.ascontiguousarray(st.mol.id, dtype=np.int32)
    ish = np.ascontiguousarray(st.bas.ish_at, dtype=np.int32)
    nid = np.ascontiguousarray(st.spec.nsh_id, dtype=np.int32)
    sesp = np.ascontiguousarray(st.spec.selfenergy.reshape(-1, order="F"))
    kcn = np.ascontiguousarray(st.spec.kcn.reshape(-1, order="F"))
    O.lib().orc_get_selfenergy(st.mol.nat, p(ids), p(ish), p(nid), st.msh, p(sesp), p(kcn),
                               p(np.ascontiguousarray(cn)), p(out), p(dout))
    assert np.array_equal(se, out) and np.array_equal(dse, dout)


def test_null_context_is_an_error_not_a_crash():
    z, x = synthetic.octane()
    st = structure.setup(structure.Structure(z, x))
    a = h0._Args(st)
    nao = st.bas.nao
    S = np.zeros((nao, nao), order="F")
    se = np.zeros(st.bas.nsh)
    st_code = L.lib().tbx_get_hamiltonian(None, ct.byref(a.geom), ct.byref(a.list),
                                          ct.byref(a.basis), ct.byref(a.h0), h0._p(a.ish_at),
                                          h0._p(a.iao_sh), nao, h0._p(se), h0._p(S), h0._p(S),
                                          h0._p(S), h0._p(S))
    assert st_code == 1
    assert b"context" in L.lib().tbx_last_error()


def test_selfenergy_rejects_null():
    assert L.lib().tbx_get_selfenergy(1, None, None, None, 1, None, None, None, None, None) == 1


def test_batch_create_rejects_l3_shell_without_gpu():
    """A basis with an f shell is refused with TBX_ERR_BASIS before any device
    work (the host validation runs first)."""
    ang = np.array([3], dtype=np.int32)
    nprim = np.array([1], dtype=np.int32)
    nsh = np.array([1], dtype=np.int32)
    alpha = np.zeros(12)
    alpha[0] = 1.0
    coeff = np.zeros(12)
    coeff[0] = 1.0
    hs = np.ones(1)
    p = h0._p
    bas = L.Basis(1, 1, p(nsh), p(ang), p(nprim), p(alpha), p(coeff), 25.0)
    par = L.H0Par(p(hs), p(hs), p(hs))
    mol_nat = np.array([1], dtype=np.int32)
    ids = np.array([0], dtype=np.int32)
    desc = L.BatchDesc(1, p(mol_nat), p(ids), ct.pointer(bas), ct.pointer(par), 10.0)
    out = ct.c_void_p()
    # a dummy non-null context pointer is never dereferenced before validation
    code = L.lib().tbx_batch_create(ct.c_void_p(1), ct.byref(desc), ct.byref(out))
    assert code == 2
    assert b"l=3" in L.lib().tbx_last_error()
    assert not out.value


def test_packed_maps_match_tables():
    """The kernels' register-only spherical/Cartesian maps reproduce the
    transform tables (host self-test exported by the library)."""
    assert L.lib() is not None
    import ctypes as ct
    f = ct.CDLL(str(L.LIB_PATH)).tbx_selftest_maps
    f.restype = ct.c_int
    assert f() == 0


def test_fortran_driver_reads_input_and_reports_missing_device(tmp_path):
    """The Fortran binding is built, its driver parses the stream input written
    by tests/fortran_io.py, and without a GPU it stops with the C-ABI's error
    message marshalled through tbx_h0_last_error (exit status 2)."""
    import subprocess

    torch = pytest.importorskip("torch")
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present: the parity test covers the driver")
    from fortran_io import DRIVER, write_input

    st = structure.setup(structure.Structure(*synthetic.octane()))
    se, dse = st.selfenergy(np.full(st.mol.nat, 2.0))
    nao, nat = st.bas.nao, st.mol.nat
    P = np.asfortranarray(np.eye(nao)[:, :, None])
    bars = (np.zeros((nao, nao)), np.zeros((nao, nao)), np.zeros((3, nao, nao)),
            np.zeros((6, nao, nao)))
    write_input(tmp_path / "in.bin", st, se, dse, np.zeros((nao, 1)), np.zeros((3, nat, 1)),
                np.zeros((6, nat, 1)), P, P, bars)
    r = subprocess.run([str(DRIVER), str(tmp_path / "in.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 2, r.stdout + r.stderr
    assert "no context" in r.stdout and len(r.stdout.strip()) > len("tbx_h0_driver: no context:")
    # tbx_h0_pack_tables rejects a shell with more primitives than a cgto holds
    # (basis/type.f90:32), before any device call (exit status 3)
    st.nprim = st.nprim.copy()
    st.nprim[0] = 13
    write_input(tmp_path / "bad.bin", st, se, dse, np.zeros((nao, 1)), np.zeros((3, nat, 1)),
                np.zeros((6, nat, 1)), P, P, bars)
    r = subprocess.run([str(DRIVER), str(tmp_path / "bad.bin"), str(tmp_path / "out.bin")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 3 and "rejected the tables" in r.stdout, r.stdout + r.stderr


def test_full_builders_validate_arguments():
    """get_multipole_integrals / get_dipole_integrals / get_overlap reject a
    missing basis or output before any device work."""
    z, x = synthetic.octane()
    st = structure.setup(structure.Structure(z, x))
    a = h0._Args(st)
    nao = st.bas.nao
    S = np.zeros((nao, nao), order="F")
    so = L.lib()
    args = (ct.byref(a.geom), None, h0._p(a.ish_at), h0._p(a.iao_sh), nao)
    assert so.tbx_get_multipole_integrals(None, *args, h0._p(S), None, None) == TBX_ERR_BASIS
    assert so.tbx_get_overlap(None, ct.byref(a.geom), ct.byref(a.basis), h0._p(a.ish_at),
                              h0._p(a.iao_sh), nao, None) == TBX_ERR_ARG
    assert so.tbx_get_dipole_integrals(None, ct.byref(a.geom), ct.byref(a.basis),
                                       h0._p(a.ish_at), h0._p(a.iao_sh), nao, h0._p(S),
                                       None) == TBX_ERR_ARG
    # a null context is reported, not dereferenced
    assert so.tbx_get_overlap(None, ct.byref(a.geom), ct.byref(a.basis), h0._p(a.ish_at),
                              h0._p(a.iao_sh), nao, h0._p(S)) == TBX_ERR_ARG
    assert b"context" in so.tbx_last_error()


def test_scf_consumers_reject_null_batch():
    so = L.lib()
    buf = (ct.c_double * 8)()
    assert so.tbx_batch_add_vao_to_h1(None, 1, buf, buf, buf, None) == TBX_ERR_ARG
    assert so.tbx_batch_add_vmp_to_h1(None, 1, 3, buf, 3, buf, buf, None) == TBX_ERR_ARG
    assert so.tbx_batch_mulliken_shell_charges(None, 1, buf, buf, buf, buf, None) == TBX_ERR_ARG
    assert so.tbx_batch_mulliken_atomic_multipoles(None, 1, 6, buf, buf, buf,
                                                   None) == TBX_ERR_ARG
    assert so.tbx_batch_last_scf_timing(None, buf) == TBX_ERR_ARG


def test_batch_entry_points_reject_null():
    """The batched gradient and adjoint entry points, with and without the
    forward's S and D, refuse a null batch or a null S / D before any device
    work."""
    so = L.lib()
    buf = (ct.c_double * 8)()
    assert so.tbx_batch_adjoint(None, *([buf] * 8), None) == TBX_ERR_ARG
    assert so.tbx_batch_adjoint_with_forward(None, *([buf] * 10), None) == TBX_ERR_ARG
    assert so.tbx_batch_adjoint_with_forward(None, buf, buf, None, buf, *([buf] * 6),
                                             None) == TBX_ERR_ARG
    assert b"null" in so.tbx_last_error()
    assert so.tbx_batch_gradient(None, *([buf] * 11), None) == TBX_ERR_ARG
    assert so.tbx_batch_gradient_with_forward(None, *([buf] * 13), None) == TBX_ERR_ARG
    assert so.tbx_batch_gradient_with_forward(None, buf, buf, buf, None, *([buf] * 9),
                                              None) == TBX_ERR_ARG


_SHELLS ={1: (0,), 6: (0, 1), 7: (0, 1), 8: (0, 1), 9: (0, 1), 16: (0, 1, 2), 17: (0, 1, 2)}


def _expected_plan_totals(mols):
    """Totals the unit plans must have, from the GFN2 shells of each element
    alone: dense elements sum(nao^2); forward pairs (i, j <= i), gradient
    pairs (i, j < i); Cartesian-pair tasks sum over those pairs of
    ncart_i ncart_j = ((sum ncart)^2 +- sum ncart^2) / 2."""
    tot = dict(ndense=0, npair_fwd=0, npair_grad=0, ntask_fwd=0, ntask_grad=0)
    for z, _ in mols:
        nao = np.array([sum(2 * l + 1 for l in _SHELLS[int(a)]) for a in z])
        nc = np.array([sum((l + 1) * (l + 2) // 2 for l in _SHELLS[int(a)]) for a in z])
        n = len(z)
        tot["ndense"] += int(nao.sum()) ** 2
        tot["npair_fwd"] += n * (n + 1) // 2
        tot["npair_grad"] += n * (n - 1) // 2
        tot["ntask_fwd"] += (int(nc.sum()) ** 2 + int((nc ** 2).sum())) // 2
        tot["ntask_grad"] += (int(nc.sum()) ** 2 - int((nc ** 2).sum())) // 2
    return tot


@pytest.mark.parametrize("hint", [False, True])
@pytest.mark.parametrize("tiles", [
    None,                                                    # library defaults
    dict(fwd_unit_tasks=100, fwd_unit_pairs=1, fwd_chunk=1,  # one atom pair per unit
         grad_unit_tasks=100, grad_unit_pairs=1, grad_chunk=1),
    dict(fwd_unit_tasks=192, fwd_unit_pairs=7, fwd_chunk=2,
         grad_unit_tasks=112, grad_unit_pairs=13, grad_chunk=4),
    dict(fwd_unit_tasks=256, fwd_unit_pairs=128, fwd_chunk=3,  # bands of rows
         grad_unit_tasks=256, grad_unit_pairs=128, grad_chunk=1, fwd_band=5, grad_band=16),
    dict(fwd_unit_tasks=100, fwd_unit_pairs=1, fwd_chunk=1,
         grad_unit_tasks=100, grad_unit_pairs=1, grad_chunk=1, fwd_band=3, grad_band=2),
])
def test_unit_plans_cover_every_pair_and_element_once(tiles, hint):
    """Host work-unit planner, no GPU: every atom pair (i, j <= i) of every
    molecule is in exactly one forward unit and every (i, j < i) in exactly one
    gradient unit; every dense element is written by exactly one forward
    element, as the kernel locates it from the segment words, and reads only
    its own shell pair's tasks; no unit exceeds a cap.  Lone atoms (H, S), a diatomic, S/Cl d shells and molecules
    that span many units; totals checked against the closed forms above.  With
    the geometry hint (near pairs first, distant ones in units of their own)
    the same invariants hold."""
    from tblite_amd.batch import plan_stats
    rng = np.random.default_rng(11)
    mols = [(np.array([1]), np.zeros((1, 3))), (np.array([16]), np.zeros((1, 3))),
            (np.array([6, 17]), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 3.3]]))]
    mols += [synthetic.random_molecule(n, rng) for n in (9, 23, 57, 140)]
    st = plan_stats(mols, tiles, hint=hint)
    assert st["bad_pairs"] == 0 and st["bad_elems"] == 0 and st["bad_units"] == 0, st
    for k, v in _expected_plan_totals(mols).items():
        assert st[k] == v, (k, st[k], v)
    assert st["nelem"] == st["ndense"]
    assert st["nunit_fwd"] >= 4 and st["nunit_grad"] >= 3
    if tiles is not None and tiles["fwd_unit_pairs"] == 1:
        assert st["nunit_fwd"] == st["npair_fwd"] and st["nunit_grad"] == st["npair_grad"]
    # forward plans hold per-pair and per-shell-pair words only, no per-element data
    assert 44 * st["npair_fwd"] <= st["plan_bytes_fwd"] <= 256 * st["npair_fwd"]
    assert st["plan_bytes_grad"] > 0


def test_unit_plans_reject_tiles_beyond_the_kernels():
    """Gradient units beyond 256 tasks, or below the largest atom pair (an S-S
    pair has (1 + 3 + 6)^2 = 100 Cartesian-pair tasks), are argument errors."""
    from tblite_amd.batch import plan_stats
    mols = [(np.array([16, 16]), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 3.9]]))]
    base = dict(fwd_unit_tasks=256, fwd_unit_pairs=128, fwd_chunk=3,
                grad_unit_tasks=256, grad_unit_pairs=128, grad_chunk=1)
    for bad in (dict(grad_unit_tasks=257), dict(grad_unit_tasks=99), dict(fwd_unit_tasks=99),
                dict(fwd_unit_pairs=129), dict(grad_chunk=0), dict(fwd_band=0),
                dict(grad_band=17)):
        with pytest.raises(L.TbxError) as e:
            plan_stats(mols, {**base, **bad})
        assert e.value.code == TBX_ERR_ARG
    assert plan_stats(mols, base)["bad_elems"] == 0


def test_unit_plans_of_the_2048_molecule_workload():
    """The plans the bench runs (2048 molecules of 8-256 atoms, default
    tiles), in the geometry-independent order and near-first by the
    molecules' geometry: no violation, and the totals of the closed forms.
    With the hint, a large share of the gradient units holds no pair within
    the cutoff (the kernel skips them).  About 20 s."""
    from tblite_amd.batch import plan_stats
    mols = synthetic.molecule_batch(2048, 8, 256, seed=1004, workers=8)
    want = _expected_plan_totals(mols)
    for hint in (False, True):
        st = plan_stats(mols, hint=hint)
        assert st["bad_pairs"] == 0 and st["bad_elems"] == 0 and st["bad_units"] == 0, st
        for k, v in want.items():
            assert st[k] == v, (hint, k, st[k], v)
        assert st["ndense"] == 144333216 and st["nelem"] == st["ndense"]
        if hint:
            print("gradient units", st["nunit_grad"], "with a pair within the cutoff",
                  st["nunit_grad_live"])
            assert st["nunit_grad_live"] < 0.8 * st["nunit_grad"]


def test_pair_order_hint_validates_arguments():
    """The pair-order hint refuses a negative or non-finite skin and
    non-finite coordinates (status TBX_ERR_ARG, nothing built); a null batch
    is an argument error, not a crash."""
    from tblite_amd.batch import plan_stats
    mols = [(np.array([6, 1]), np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 2.0]]))]
    assert plan_stats(mols, hint=True, skin=0.0)["bad_elems"] == 0
    for skin in (-1.0, float("nan"), float("inf")):
        with pytest.raises(L.TbxError) as e:
            plan_stats(mols, hint=True, skin=skin)
        assert e.value.code == TBX_ERR_ARG
    with pytest.raises(L.TbxError) as e:
        plan_stats([(np.array([6, 1]), np.array([[0.0, 0.0, 0.0], [0.0, np.nan, 2.0]]))],
                   hint=True)
    assert e.value.code == TBX_ERR_ARG
    assert L.lib().tbx_batch_order_pairs(None, None, 1.0) == TBX_ERR_ARG


def test_diat_scale_tables_are_checked_on_the_host():
    """The diatomic-frame scaling needs all three tables and species whose
    shells are s, p, d in this order (GFN2's transition metals are d, s, p)."""
    import ctypes as ct
    from tblite_amd.batch import _Description
    mols = [(np.array([8, 1, 1]), np.array([[0.0, 0, 0], [1.8, 0, 0], [0, 1.8, 0]])),
            (np.array([17, 6]), np.array([[0.0, 0, 0], [0, 0, 3.3]]))]

    def ones(nums):
        n = len(nums)
        return [np.ones((n, n)), np.full((n, n), 1.1), np.full((n, n), 0.9)]

    def stats(desc):
        out = L.PlanStats()
        return L.lib().tbx_batch_plan_stats(ct.byref(desc), None, None, ct.c_double(1.0),
                                            ct.byref(out))

    d = _Description(mols, ones)
    assert stats(d.desc) == 0
    d.h0.kdel = None  # one table missing
    assert stats(d.desc) == 1
    assert b"together" in L.lib().tbx_last_error()
    tm = mols + [(np.array([26, 1]), np.array([[0.0, 0, 0], [0, 0, 3.0]]))]
    assert stats(_Description(tm, ones).desc) == 2
    assert b"s, p, d" in L.lib().tbx_last_error()
    assert stats(_Description(tm).desc) == 0  # without the scaling the order is free
