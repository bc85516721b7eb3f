"""Autotuner of the batched kernels' tile shapes.

A tile shape is the work-unit cap (Cartesian-pair tasks per workgroup), the
chunk length (kept primitive pairs per work item) and the band (row atoms a
unit spans, walked column by column) of ``k_forward`` and of ``k_gradient``;
the two kernels use separate work units, so each is tuned on its own.  Caps
and chunks are searched first, then the bands at the chosen caps and chunks.
Every candidate is applied with ``Batch.set_tiles`` and timed with the
library's HIP events on the launch stream; the fastest shape per kernel is
kept (and optionally cached on disk, keyed by the batch's shape)."""
from __future__ import annotations

import itertools
import json
import os
from pathlib import Path

import torch

FWD_UNIT_TASKS = (192, 256)   # forward workgroups hold up to 256 tasks
GRAD_UNIT_TASKS = (192, 256)  # gradient workgroups hold up to 256 tasks
CHUNKS = (1, 2, 3, 4, 6)
FWD_BANDS = (1, 2)
GRAD_BANDS = (2, 4, 8)


def batch_key(b) -> str:
    return f"nmol{b.nmol}-nat{b.nat}-nao{b.nao}-nd{b.ndense}-sp{'.'.join(map(str, b.nums))}"


def _time(b, xyz, se, adjoints, plain, reps):
    """Mean forward, adjoint (given the forward's S, D) and plain-gradient ms."""
    f = a = g = 0.0
    for _ in range(reps):
        S, D, _, _ = b.forward(xyz, se)
        b.adjoint(xyz, se, *adjoints, S=S, D=D)  # as the autograd backward runs it
        torch.cuda.synchronize()
        t = b.last_timing()
        f += t["forward_ms"] / reps
        a += t["adjoint_ms"] / reps
        if plain is not None:
            b.gradient(xyz, se, *plain, S=S, D=D)
            torch.cuda.synchronize()
            g += b.last_timing()["gradient_ms"] / reps
    return f, a, g


def tune(b, xyz, se, adjoints, plain=None, reps: int = 2, fwd_unit_tasks=FWD_UNIT_TASKS,
         grad_unit_tasks=GRAD_UNIT_TASKS, chunks=CHUNKS, fwd_bands=FWD_BANDS,
         grad_bands=GRAD_BANDS, cache: str | os.PathLike | None = None, log=None):
    """Choose tile shapes for batch ``b`` on the given inputs.

    ``adjoints`` = (Sb, Db, Qb, Hb) times the adjoint kernel; ``plain`` =
    (dsedcn, P, W, vao, vdp, vqp) additionally times the plain gradient
    kernel (the gradient tile is then chosen on the sum of both).  Returns
    ``{"tiles": {...}, "table": [...]}`` and leaves the best tiles applied."""
    key = batch_key(b)
    if cache is not None and Path(cache).exists():
        known = json.loads(Path(cache).read_text())
        if key in known:
            b.set_tiles(**known[key]["tiles"])
            return known[key]
    min_tasks = int(max(sum((c.ang + 1) * (c.ang + 2) // 2 for c in sh)
                        for sh in b.spec.basis_cgto)) ** 2
    fwd = [t for t in fwd_unit_tasks if t >= min_tasks]
    grad = [t for t in grad_unit_tasks if t >= min_tasks]
    table = []
    # both kernels are timed in every step; the forward walks its task caps,
    # the gradient its own (cycled), with the same chunk length
    steps = max(len(fwd), len(grad))
    for k, chunk in itertools.product(range(steps), chunks):
        tf, tg = fwd[k % len(fwd)], grad[k % len(grad)]
        b.set_tiles(fwd_unit_tasks=tf, fwd_chunk=chunk, grad_unit_tasks=tg, grad_chunk=chunk)
        f, a, g = _time(b, xyz, se, adjoints, plain, reps)
        row = {"fwd_unit_tasks": tf, "grad_unit_tasks": tg, "chunk": chunk,
               "fwd_band": b.tiles()["fwd_band"], "grad_band": b.tiles()["grad_band"],
               "forward_ms": f, "adjoint_ms": a, "gradient_ms": g}
        table.append(row)
        if log:
            log(f"tune: tasks={tf}/{tg} chunk={chunk}: forward {f:.3f} ms, adjoint {a:.3f} ms"
                + (f", gradient {g:.3f} ms" if plain is not None else ""))
    bf = min(table, key=lambda r: r["forward_ms"])
    bg = min(table, key=lambda r: r["adjoint_ms"] + r["gradient_ms"])
    tiles = {"fwd_unit_tasks": bf["fwd_unit_tasks"], "fwd_chunk": bf["chunk"],
             "grad_unit_tasks": bg["grad_unit_tasks"], "grad_chunk": bg["chunk"]}
    # then the bands at those caps and chunks (forward and gradient bands cycled)
    start = len(table)
    for k in range(max(len(fwd_bands), len(grad_bands))):
        fb, gb = fwd_bands[k % len(fwd_bands)], grad_bands[k % len(grad_bands)]
        b.set_tiles(**tiles, fwd_band=fb, grad_band=gb)
        f, a, g = _time(b, xyz, se, adjoints, plain, reps)
        table.append({"fwd_unit_tasks": tiles["fwd_unit_tasks"],
                      "grad_unit_tasks": tiles["grad_unit_tasks"], "chunk": None,
                      "fwd_band": fb, "grad_band": gb, "forward_ms": f, "adjoint_ms": a,
                      "gradient_ms": g})
        if log:
            log(f"tune: bands={fb}/{gb}: forward {f:.3f} ms, adjoint {a:.3f} ms"
                + (f", gradient {g:.3f} ms" if plain is not None else ""))
    if len(table) > start:
        bands = table[start:]
        tiles["fwd_band"] = min(bands, key=lambda r: r["forward_ms"])["fwd_band"]
        tiles["grad_band"] = min(bands, key=lambda r: r["adjoint_ms"] + r["gradient_ms"])["grad_band"]
    b.set_tiles(**tiles)
    result = {"tiles": b.tiles(), "table": table}
    if cache is not None:
        p = Path(cache)
        known = json.loads(p.read_text()) if p.exists() else {}
        known[key] = result
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(known, indent=1))
    return result
