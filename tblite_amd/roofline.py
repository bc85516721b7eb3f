"""Algorithmic work of the integral path, for roofline reporting.

K is the set of kept primitive pairs -- every shell pair of every half-list
atom pair (``adjlist.f90:125-129``) plus every on-site shell pair, screened
with ``ai*aj*r2/(ai+aj) <= intcut`` (``integral/native/integrals.f90:1211-1215``)
-- and c = ncart(l_i)*ncart(l_j).  Per kept primitive pair and Cartesian pair:

- forward  (S, D, Q):                40 FLOP   (10 moments x (2 mul for the
  x*y*z product + 2 for the weighted accumulate))
- gradient / adjoint (S, D, Q and their R derivatives, contracted): 160 FLOP
  (40 quantities x 4)

Bytes: the forward writes S, H0, D, Q densely (88 bytes per nao^2 element);
the plain gradient reads P and W (16 B), the adjoint reads the four dense
adjoints (88 B).  Setup, spherical transform and scatter are not counted, so
a fraction of peak derived from these figures is a lower bound.
"""
from __future__ import annotations

import numpy as np

FLOP_FWD_PER_KC = 40
FLOP_BWD_PER_KC = 160
BYTES_FWD_PER_ELEM = 88
BYTES_GRAD_PER_ELEM = 16
BYTES_ADJ_PER_ELEM = 88

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0     # MI355X HBM3E


def _ncart(l):
    return (l + 1) * (l + 2) // 2


def count_kc(spec, ids_per_mol, xyz_per_mol, cutoff, intcut=25.0):
    """Return (Kc_offsite, Kc_onsite): sums over kept primitive pairs of
    ncart_i*ncart_j for the half-list pairs and for the on-site blocks."""
    nsp = len(spec.nums)
    # thresholds: kept iff r2 <= intcut*(ai+aj)/(ai*aj)
    shells = [[(c.ang, np.asarray(c.alpha)) for c in spec.basis_cgto[s]] for s in range(nsp)]
    thr = {}
    onsite = np.zeros(nsp)
    for si in range(nsp):
        for sj in range(nsp):
            lst = []
            for li, ai in shells[si]:
                for lj, aj in shells[sj]:
                    t = (intcut * (ai[:, None] + aj[None, :]) / (ai[:, None] * aj[None, :])).ravel()
                    lst.append((np.sort(t), _ncart(li) * _ncart(lj)))
            thr[(si, sj)] = lst
        onsite[si] = sum(len(ai) * len(aj) * _ncart(li) * _ncart(lj)
                         for li, ai in shells[si] for lj, aj in shells[si])
    cut2 = cutoff * cutoff
    r2_by = {k: [] for k in thr}
    kc_on = 0.0
    for ids, xyz in zip(ids_per_mol, xyz_per_mol):
        ids = np.asarray(ids)
        x = np.asarray(xyz).reshape(-1, 3)
        n = len(ids)
        kc_on += onsite[ids].sum()
        if n < 2:
            continue
        iu, ju = np.triu_indices(n, 1)
        d = x[iu] - x[ju]
        r2 = np.einsum("ij,ij->i", d, d)
        keep = r2 <= cut2
        iu, ju, r2 = iu[keep], ju[keep], r2[keep]
        code = ids[iu] * nsp + ids[ju]
        for c in np.unique(code):
            sel = code == c
            r2_by[(int(c) // nsp, int(c) % nsp)].append(r2[sel])
    kc_off = 0.0
    for key, chunks in r2_by.items():
        if not chunks:
            continue
        r2 = np.concatenate(chunks)
        for t, nc in thr[key]:
            kept = len(t) * len(r2) - np.searchsorted(t, r2, side="left").sum()
            kc_off += nc * float(kept)
    return kc_off, kc_on


def roofline_entry(bound, achieved, peak, unit, traffic=None):
    return {"bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
            "frac": achieved / peak, "traffic": traffic}


def copy_probe(gib: float = 1.0, reps: int = 3, device: int = 0):
    """Run the library's streaming copy kernel (k_copy) ``reps`` times over
    ``gib`` GiB and return (bytes read per launch, bytes written per launch,
    measured GB/s).  Profiled with FETCH_SIZE / WRITE_SIZE, its per-dispatch
    counter readings divided by these byte counts give the factors that turn
    the integral kernels' readings into HBM bytes."""
    import ctypes as ct

    import torch

    from . import lib as L

    n = int(gib * (1 << 30)) // 8
    dev = torch.device("cuda", device)
    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    stream = torch.cuda.current_stream(dev)
    h = ct.c_void_p(stream.cuda_stream)
    L.check(L.lib().tbx_probe_copy(ct.c_void_p(src.data_ptr()), ct.c_void_p(dst.data_ptr()), n, h))
    torch.cuda.synchronize(dev)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record(stream)
    for _ in range(reps):
        L.check(L.lib().tbx_probe_copy(ct.c_void_p(src.data_ptr()), ct.c_void_p(dst.data_ptr()),
                                       n, h))
    t1.record(stream)
    torch.cuda.synchronize(dev)
    ms = t0.elapsed_time(t1) / reps
    assert bool((dst == 1).all())
    return 8 * n, 8 * n, 16 * n / (ms * 1e6)


def peaks_probe(gib: float = 8.0, n_gemm: int = 8192, reps: int = 5, device: int = 0) -> dict:
    """What this GPU reaches on plain work, beside its published peaks: the
    library's streaming copy (read + write GB/s), a write-only fill of the
    same size (torch's zero_, GB/s) and an FP64 DGEMM of n_gemm^3 through
    torch.matmul (rocBLAS / hipBLASLt on FP64 MFMA, TFLOP/s)."""
    import torch

    dev = torch.device("cuda", device)
    _, _, copy_gbs = copy_probe(gib, reps, device)
    buf = torch.empty(int(gib * (1 << 30)) // 8, dtype=torch.float64, device=dev)
    buf.zero_()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize(dev)
    t0.record()
    for _ in range(reps):
        buf.zero_()
    t1.record()
    torch.cuda.synchronize(dev)
    fill_gbs = buf.numel() * 8 / (t0.elapsed_time(t1) / reps * 1e6)
    del buf
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    a = torch.rand(n_gemm, n_gemm, dtype=torch.float64, device=dev, generator=g)
    b = torch.rand(n_gemm, n_gemm, dtype=torch.float64, device=dev, generator=g)
    c = torch.matmul(a, b)
    torch.cuda.synchronize(dev)
    t0.record()
    for _ in range(reps):
        c = torch.matmul(a, b)
    t1.record()
    torch.cuda.synchronize(dev)
    tflops = 2.0 * n_gemm ** 3 / (t0.elapsed_time(t1) / reps * 1e-3) / 1e12
    return {"copy_GBs": copy_gbs, "fill_GBs": fill_gbs, "dgemm_fp64_TFLOPs": tflops,
            "gib": gib, "n_gemm": n_gemm, "reps": reps,
            "published": {"hbm_GBs": PEAK_HBM_GBS, "fp64_TFLOPs": PEAK_FP64_TFLOPS}}


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="HBM traffic calibration probe; --peaks: the "
                                 "copy, fill and FP64 DGEMM rates this GPU reaches")
    ap.add_argument("--gib", type=float, default=1.0)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--peaks", action="store_true")
    a = ap.parse_args()
    if a.peaks:
        print(json.dumps(peaks_probe(max(a.gib, 8.0))))
    else:
        rd, wr, gbs = copy_probe(a.gib, a.reps)
        print(json.dumps({"probe": "k_copy", "bytes_read": rd, "bytes_written": wr,
                          "launches": a.reps + 1, "GBs": gbs}))
