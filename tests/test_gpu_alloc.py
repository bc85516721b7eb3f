"""Forced device-allocation failure: under a context memory limit smaller than
a call needs, the drop-in get_hamiltonian and the batch set-up return
TBX_ERR_ALLOC with their outputs untouched and no batch created; with the
limit lifted, the same context gives the oracle's results."""
import ctypes as ct

import numpy as np
import pytest

from oracle import oracle as O
from tblite_amd import h0, lib as L, structure, synthetic

pytestmark = pytest.mark.gpu

TBX_ERR_ALLOC = 4  # include/tbx.h


def test_allocation_failure_is_reported_without_side_effects():
    from tblite_amd.batch import _Description
    so = L.lib()
    ctx = so.tbx_context_new(0)
    assert ctx
    try:
        z, x = synthetic.mixed_cluster(80, seed=11, density=1.2)
        st = structure.setup(structure.Structure(z, x))
        nao = st.bas.nao
        print(f"[gpu] allocation failure: nat={st.mol.nat} nao={nao}", flush=True)
        a = h0._Args(st)
        se, _ = st.selfenergy(np.full(st.mol.nat, 2.0))
        se = np.ascontiguousarray(se)
        limit = 64 * 1024  # far below the 88 nao^2 bytes of the dense outputs
        assert 88 * nao * nao > 4 * limit
        assert so.tbx_context_set_memory_limit(ctx, limit) == 0
        outs = [np.full(shape, 7.0, order="F") for shape in
                ((nao, nao), (3, nao, nao), (6, nao, nao), (nao, nao))]
        p = h0._p
        args = (ctx, ct.byref(a.geom), ct.byref(a.list), ct.byref(a.basis), ct.byref(a.h0),
                p(a.ish_at), p(a.iao_sh), nao, p(se), *[p(o) for o in outs])
        assert so.tbx_get_hamiltonian(*args) == TBX_ERR_ALLOC
        assert b"memory" in so.tbx_last_error().lower()
        assert all(np.all(o == 7.0) for o in outs)
        assert 0 <= so.tbx_context_memory_used(ctx) <= limit
        # a batch whose tables exceed the limit is not created
        d = _Description([(z, x), synthetic.random_molecule(30, np.random.default_rng(1))])
        handle = ct.c_void_p()
        assert so.tbx_batch_create(ctx, ct.byref(d.desc), ct.byref(handle)) == TBX_ERR_ALLOC
        assert not handle.value
        assert so.tbx_context_memory_used(ctx) <= limit
        # without the limit the same context works
        assert so.tbx_context_set_memory_limit(ctx, -1) == 0
        assert so.tbx_get_hamiltonian(*args) == 0
        for ref, got in zip(O.get_hamiltonian(st, se), outs):
            assert np.abs(ref - got).max() < 1e-10
        assert so.tbx_context_memory_used(ctx) > limit
    finally:
        so.tbx_context_delete(ctx)
