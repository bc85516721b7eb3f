"""Run the batched kernels alone on the 2048-molecule workload, for profilers.

The workload and inputs are bench.py's (2048 molecules of 8-256 atoms,
seed 1004, near-first pair order, default tiles); each kernel is launched
``--reps`` times after one warm-up, with nothing else on the GPU in between,
so that ``rocprofv3 --kernel-include-regex`` can pick them out:

    python -m tblite_amd.kernel_driver [--reps 3] [--nmol 2048]
        [--kernels forward,adjoint,adjoint_fwd,gradient,gradient_fwd,density]
"""
from __future__ import annotations

import argparse
import sys
import time

import torch

from . import synthetic
from .batch import Batch


def workload(nmol: int, seed: int = 1004):
    mols = synthetic.molecule_batch(nmol, seed=seed, workers=8)
    b = Batch(mols)
    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    xyz = b.tensor_xyz()
    cn = torch.rand(b.nat, generator=g, dtype=torch.float64, device=dev) * 3.0 + 0.5
    se, dsedcn = b.selfenergy(cn)
    se, dsedcn = se.contiguous(), dsedcn.contiguous()

    def rnd(n, scale=1.0):
        return torch.randn(n, generator=g, dtype=torch.float64, device=dev) * scale

    P, W = rnd(b.ndense), rnd(b.ndense, 0.05)
    plain = (dsedcn, P, W, rnd(b.nao, 1e-2), rnd(3 * b.nat, 1e-3), rnd(6 * b.nat, 1e-3))
    adjoints = (rnd(b.ndense), rnd(3 * b.ndense), rnd(6 * b.ndense), rnd(b.ndense))
    return b, xyz, se, plain, adjoints


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--nmol", type=int, default=2048)
    ap.add_argument("--kernels", default="forward,adjoint,adjoint_fwd,gradient,gradient_fwd")
    a = ap.parse_args(argv)
    b, xyz, se, plain, adjoints = workload(a.nmol)
    outs = (torch.empty(b.ndense, dtype=torch.float64, device="cuda"),
            torch.empty(3 * b.ndense, dtype=torch.float64, device="cuda"),
            torch.empty(6 * b.ndense, dtype=torch.float64, device="cuda"),
            torch.empty(b.ndense, dtype=torch.float64, device="cuda"))
    run = {"forward": lambda: b.forward(xyz, se, out=outs),
           "adjoint": lambda: b.adjoint(xyz, se, *adjoints),
           # with the forward's S and D (the autograd backward); outs hold
           # them once the forward has run
           "adjoint_fwd": lambda: b.adjoint(xyz, se, *adjoints, S=outs[0], D=outs[1]),
           "gradient": lambda: b.gradient(xyz, se, *plain),
           "gradient_fwd": lambda: b.gradient(xyz, se, *plain, S=outs[0], D=outs[1])}
    names = [k for k in a.kernels.split(",") if k]
    if "density" in names:
        # P = C diag(f) C^T on random coefficients (packed per molecule)
        g = torch.Generator(device="cuda")
        g.manual_seed(12)
        coeff = torch.randn(b.ndense, generator=g, dtype=torch.float64, device="cuda")
        focc = torch.rand(b.nao, generator=g, dtype=torch.float64, device="cuda") * 2.0
        pd = torch.empty(b.ndense, dtype=torch.float64, device="cuda")
        run["density"] = lambda: b.density_matrix(coeff, focc, out=pd)
    for k in names:
        run[k]()  # warm-up
    torch.cuda.synchronize()
    for k in names:
        for _ in range(a.reps):
            if k == "density":  # one launch on torch's current stream
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                run[k]()
                e1.record()
                torch.cuda.synchronize()
                print(f"{k} {e0.elapsed_time(e1):.3f} ms", flush=True)
                continue
            run[k]()
            torch.cuda.synchronize()
            t = b.last_timing()
            ms = {"forward": t["forward_ms"], "adjoint": t["adjoint_ms"],
                  "adjoint_fwd": t["adjoint_ms"], "gradient": t["gradient_ms"],
                  "gradient_fwd": t["gradient_ms"]}[k]
            print(f"{k} {ms:.3f} ms", flush=True)
    time.sleep(0.1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
