"""tbx_batch_gradient_noalloc (static device scratch, no allocation) against
tbx_batch_gradient through the C-ABI: two calls in a row on two batches, the
second smaller, so that partial sums the first call left in the shared
scratch would show in the second if they were not cleared."""
import ctypes as ct
import sys

import numpy as np
import pytest

from tblite_amd import lib as L, synthetic

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

TOL_GRAD = 1e-8


def _inputs(b, seed):
    rng = np.random.default_rng(seed)
    P = np.zeros(b.ndense)
    for m in range(b.nmol):
        n = int(b.ao_off[m + 1] - b.ao_off[m])
        g = rng.normal(size=(n, n))
        P[int(b.dense_off[m]): int(b.dense_off[m]) + n * n] = (0.5 * (g + g.T)).ravel()
    W = 0.05 * P
    vao = rng.normal(scale=1e-2, size=b.nao)
    vdp = rng.normal(scale=1e-3, size=3 * b.nat)
    vqp = rng.normal(scale=1e-3, size=6 * b.nat)
    cn = rng.uniform(0.5, 3.5, b.nat)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")  # noqa: E731
    se, dse = b.selfenergy(dev(cn))
    return [b.tensor_xyz(), se.contiguous(), dse.contiguous(), dev(P), dev(W), dev(vao),
            dev(vdp), dev(vqp)]


def _run(fn, b, args):
    outs = (torch.full((b.nat,), np.nan, dtype=torch.float64, device="cuda"),
            torch.full((3 * b.nat,), np.nan, dtype=torch.float64, device="cuda"),
            torch.full((9 * b.nmol,), np.nan, dtype=torch.float64, device="cuda"))
    p = [ct.c_void_p(t.data_ptr()) for t in list(args) + list(outs)]
    stream = ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(fn(b.handle, *p, stream))
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in outs]


def test_noalloc_gradient_matches_gradient_twice():
    from tblite_amd.batch import Batch
    so = L.lib()
    rng = np.random.default_rng(2)
    big = [synthetic.random_molecule(int(k), rng) for k in rng.integers(20, 120, size=24)]
    small = big[:5]
    for name, mols, seed in (("first", big, 3), ("second, smaller", small, 4)):
        b = Batch(mols)
        print(f"[gpu] no-alloc gradient, {name} call: nmol={b.nmol} nsh={b.nsh}", flush=True)
        sys.stdout.flush()
        args = _inputs(b, seed)
        ref = _run(so.tbx_batch_gradient, b, args)
        got = _run(so.tbx_batch_gradient_noalloc, b, args)
        for r, g, lab in zip(ref, got, ("dEdcn", "gradient", "sigma")):
            assert np.all(np.isfinite(g)), lab
            assert np.abs(r - g).max() < TOL_GRAD, (name, lab, np.abs(r - g).max())
