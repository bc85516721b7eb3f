"""Throughput of the GFN2 S/D/Q + H0 + gradient path on one MI355X.

Workload (one "step"): the batched GFN2 integral path over 2048 synthetic
molecules (atom counts log-uniform in [8, 256], H/C/N/O/F/S/Cl), FP64:

  forward   dense S, D, Q (traceless), H0 for every molecule  (one launch)
  backward  torch.autograd through the adjoint kernel with the adjoints of
            get_hamiltonian_gradient's energy functional, i.e. dE/dR and dE/dse
            equal to the plain gradient (one launch; the backward holds the
            forward's S and D and passes them on, so the terms linear in them
            are contracted per element)

Inputs are resident in HBM when the timed region starts.  ``value`` is
molecules per second of that batched-autograd step (eager launches through
torch.autograd).  Beside it the line reports the same step's two launches
replayed as a captured HIP graph, the batched plain-gradient step, the plain
per-molecule GPU path (one molecule per launch, inputs resident in HBM, on a
sample of the molecules) and the drop-in host-pointer entry points (PCIe
copies included), the roofline of the dominant kernel and of the step's
other kernel (HIP events on the launch stream), the CPU baseline (the oracle, a C restatement of the
reference loop nest with OpenMP on the physical cores, timed in a child
process on a bounded sample), the SCF-loop consumers of the resident S / Q,
and the path's neighbours on the same batch ("next_rows": coordination
number and its derivative, Fermi filling, the density products, and the
generalized eigensolver on a sample of molecules).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

NMOL = 2048
SEED = 1004


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(kernel: str, nmol: int):
    """HBM bytes per launch of ``kernel`` on this workload, from the committed
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy kernel); None
    when no profile of this workload exists."""
    path = ROOT / "profiles" / f"batch{nmol}_hbm_traffic.md"
    if not path.exists():
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def _cpu_info():
    """Host CPUs this process may run on: logical CPUs of the affinity mask,
    the physical cores among them (/proc/cpuinfo core ids) and the cgroup CPU
    quota, if one is set.  The CPU baseline runs one OpenMP thread per
    physical core, no more than the quota allows."""
    try:
        cpus = sorted(os.sched_getaffinity(0))
    except AttributeError:
        cpus = list(range(os.cpu_count() or 1))
    core_of = {}
    try:
        cur = {}
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                cur[k] = v
            elif cur:
                if "processor" in cur:
                    core_of[int(cur["processor"])] = (cur.get("physical id", "0"),
                                                      cur.get("core id", cur["processor"]))
                cur = {}
        if "processor" in cur:
            core_of[int(cur["processor"])] = (cur.get("physical id", "0"),
                                              cur.get("core id", cur["processor"]))
    except OSError:
        pass
    physical = len({core_of.get(c, ("0", str(c))) for c in cpus})
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    threads = physical
    if quota is not None:
        threads = max(1, min(threads, int(quota)))
    return {"logical": len(cpus), "physical": physical, "cgroup_quota_cpus": quota,
            "threads": threads}


def _cpu_threads():
    return _cpu_info()["threads"]


# --------------------------------------------------------------------------
# CPU baseline (child process: no torch, its own OpenMP runtime)

def cpu_baseline_child(budget_s: float):
    import numpy as np

    from oracle import oracle as O
    from tblite_amd import structure, synthetic

    mols = synthetic.molecule_batch(NMOL, seed=SEED, workers=min(8, _cpu_threads()))
    rng = np.random.default_rng(7)
    done, t_work, atoms = 0, 0.0, 0
    for z, x in mols:
        st = structure.setup(structure.Structure(z, x))
        nao, nat = st.bas.nao, st.mol.nat
        cn = rng.uniform(0.5, 3.5, nat)
        g = rng.normal(size=(nao, nao))
        P = np.asfortranarray((0.5 * (g + g.T))[:, :, None])
        W = np.asfortranarray(0.1 * P)
        vao = np.zeros((nao, 1), order="F")
        vdp = np.zeros((3, nat, 1), order="F")
        vqp = np.zeros((6, nat, 1), order="F")
        t0 = time.perf_counter()
        # timed region of the reference path: selfenergy, neighbour list,
        # get_hamiltonian, get_hamiltonian_gradient
        se, dse = st.selfenergy(cn)
        structure.new_adjacency_list(st.mol.xyz, st.trans, st.cutoff)
        O.get_hamiltonian(st, se)
        O.get_hamiltonian_gradient(st, se, dse, vao, vdp, vqp, P, W)
        t_work += time.perf_counter() - t0
        done += 1
        atoms += nat
        if t_work > budget_s:
            break
    print(json.dumps({"molecules": done, "atoms": atoms, "seconds": t_work}))


def run_cpu_baseline(budget_s=15.0):
    info = _cpu_info()
    threads = info["threads"]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_SCHEDULE="dynamic",
               OMP_PROC_BIND="spread", OMP_STACKSIZE="4G")
    out = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-baseline-child",
                          str(budget_s)], env=env, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        log("cpu baseline failed: " + out.stderr[-2000:])
        return None
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": r["molecules"] / r["seconds"], "unit": "molecules/s", "cores": threads,
            "kind": "port", "logical_cpus": info["logical"],
            "physical_cores": info["physical"], "cgroup_quota_cpus": info["cgroup_quota_cpus"],
            "sample": f"first {r['molecules']} of the 2048 molecules ({r['atoms']} atoms), "
                      f"oracle get_selfenergy+adjacency list+get_hamiltonian+"
                      f"get_hamiltonian_gradient, OpenMP {threads} threads (one per physical "
                      f"core of {info['physical']}, {info['logical']} logical CPUs), "
                      f"{r['seconds']:.1f} s"}


# --------------------------------------------------------------------------

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nmol", type=int, default=NMOL)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--cpu-baseline-child", type=float, default=None)
    args = ap.parse_args()

    if args.cpu_baseline_child is not None:
        cpu_baseline_child(args.cpu_baseline_child)
        return
    if args.gpus != 1:
        print("bench.py: the batched path runs on one GPU; several GPUs would each take a "
              "share of the molecules with no exchange between them, and that is not measured "
              "here", file=sys.stderr)
        sys.exit(2)

    cpu = None if args.no_cpu else run_cpu_baseline(args.cpu_budget)

    import numpy as np
    import torch

    from tblite_amd import h0 as h0mod
    from tblite_amd import params, roofline, structure, synthetic, tune
    from tblite_amd.batch import Batch, h0_integrals

    t0 = time.time()
    mols = synthetic.molecule_batch(args.nmol, seed=SEED, workers=min(8, _cpu_threads()))
    log(f"generated {len(mols)} molecules in {time.time() - t0:.1f} s")
    b = Batch(mols)
    dev = torch.device("cuda", 0)
    log(f"batch: nat={b.nat} nsh={b.nsh} nao={b.nao} ndense={b.ndense}")
    ids_m = [b.ids[b.atom_off[m]: b.atom_off[m + 1]] for m in range(b.nmol)]
    kc_off, kc_on = roofline.count_kc(b.spec, ids_m, [x for _, x in mols], b.cutoff, b.intcut)
    kc = kc_off + kc_on

    # inputs (seeded); the values do not change the work of any kernel
    g = torch.Generator(device=dev)
    g.manual_seed(11)
    xyz = b.tensor_xyz()
    cn = torch.rand(b.nat, generator=g, dtype=torch.float64, device=dev) * 3.0 + 0.5
    se, dsedcn = b.selfenergy(cn)
    se, dsedcn = se.contiguous(), dsedcn.contiguous()
    P = torch.empty(b.ndense, dtype=torch.float64, device=dev)
    W = torch.empty(b.ndense, dtype=torch.float64, device=dev)
    for m, (pv, wv) in enumerate(zip(b.molecule_views(P), b.molecule_views(W))):
        n = pv.shape[0]
        a = torch.randn(n, n, generator=g, dtype=torch.float64, device=dev)
        pv[:, :, 0] = 0.5 * (a + a.T)
        wv[:, :, 0] = 0.05 * pv[:, :, 0]
    vao = torch.randn(b.nao, generator=g, dtype=torch.float64, device=dev) * 1e-2
    vdp = torch.randn(3 * b.nat, generator=g, dtype=torch.float64, device=dev) * 1e-3
    vqp = torch.randn(6 * b.nat, generator=g, dtype=torch.float64, device=dev) * 1e-3
    # adjoints of the functional get_hamiltonian_gradient differentiates
    ao_atom = torch.as_tensor(b.ao_atom, device=dev)
    Hb = P
    Sb = torch.empty_like(P)
    Db = torch.empty(3 * b.ndense, dtype=torch.float64, device=dev)
    Qb = torch.empty(6 * b.ndense, dtype=torch.float64, device=dev)
    views = zip(b.molecule_views(P), b.molecule_views(W), b.molecule_views(Sb),
                b.molecule_views(Db, 3), b.molecule_views(Qb, 6))
    for m, (pv, wv, sv, dv, qv) in enumerate(views):
        ao = slice(int(b.ao_off[m]), int(b.ao_off[m + 1]))
        v = vao[ao]
        p = pv[:, :, 0]  # [nu, mu] = P(mu, nu)
        sv[:, :, 0] = -wv[:, :, 0] - 0.5 * p * (v[None, :] + v[:, None])
        at = ao_atom[ao]
        dv.copy_(-p[:, :, None] * vdp.view(-1, 3)[at][:, None, :])  # atom of column nu
        qv.copy_(-p[:, :, None] * vqp.view(-1, 6)[at][:, None, :])
    torch.cuda.synchronize()
    outs = (torch.empty(b.ndense, dtype=torch.float64, device=dev),
            torch.empty(3 * b.ndense, dtype=torch.float64, device=dev),
            torch.empty(6 * b.ndense, dtype=torch.float64, device=dev),
            torch.empty(b.ndense, dtype=torch.float64, device=dev))
    gouts = (torch.empty(b.nat, dtype=torch.float64, device=dev),
             torch.empty(3 * b.nat, dtype=torch.float64, device=dev),
             torch.empty(9 * b.nmol, dtype=torch.float64, device=dev))
    log(f"inputs ready; Kc={kc:.3e} (off-site {kc_off:.3e})")
    # tile shapes: timed choice before the warm-up (set-up, not part of the timed steps)
    tuned = None
    if not args.no_tune:
        tuned = tune.tune(b, xyz, se, (Sb, Db, Qb, Hb), (dsedcn, P, W, vao, vdp, vqp), log=log)
        log(f"tiles: {tuned['tiles']}")

    def step_autograd():
        x = xyz.detach().requires_grad_(True)
        s = se.detach().requires_grad_(True)
        o = h0_integrals(b, x, s)
        torch.autograd.backward(o, (Sb, Db, Qb, Hb))
        return x.grad, s.grad

    def step_plain():  # an SCF caller's step: the gradient given the forward's S and D
        b.forward(xyz, se, out=outs)
        return b.gradient(xyz, se, dsedcn, P, W, vao, vdp, vqp, out=gouts, S=outs[0], D=outs[1])

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        fwd, bwd = [], []
        t = time.perf_counter()
        for _ in range(steps):
            fn()
            torch.cuda.synchronize()
            tm = b.last_timing()
            fwd.append(tm["forward_ms"])
            bwd.append(tm["adjoint_ms"] if fn is step_autograd else tm["gradient_ms"])
        torch.cuda.synchronize()
        el = time.perf_counter() - t
        return el, float(np.mean(fwd)), float(np.mean(bwd))

    el_a, fwd_a, adj_a = timed(step_autograd, args.steps, args.warmup)
    log(f"autograd: {el_a / args.steps * 1e3:.2f} ms/step (forward {fwd_a:.2f} ms, "
        f"adjoint {adj_a:.2f} ms)")
    el_p, fwd_p, grad_p = timed(step_plain, args.steps, args.warmup)
    log(f"plain batched: {el_p / args.steps * 1e3:.2f} ms/step (forward {fwd_p:.2f} ms, "
        f"gradient {grad_p:.2f} ms)")

    # the same step as one captured HIP graph (forward + adjoint launches,
    # outputs preallocated), replayed; captured on a side stream
    aouts = (torch.empty(3 * b.nat, dtype=torch.float64, device=dev),
             torch.empty(b.nsh, dtype=torch.float64, device=dev))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        b.forward(xyz, se, out=outs)
        b.adjoint(xyz, se, Sb, Db, Qb, Hb, out=aouts, S=outs[0], D=outs[1])
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        b.forward(xyz, se, out=outs)
        b.adjoint(xyz, se, Sb, Db, Qb, Hb, out=aouts, S=outs[0], D=outs[1])
    for _ in range(args.warmup):
        graph.replay()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    el_g = time.perf_counter() - t
    log(f"graph-replayed step: {el_g / args.steps * 1e3:.2f} ms/step")

    # agreement of the two batched gradients (autograd vs plain kernel), and
    # of the graph's adjoint with both
    gx, _ = step_autograd()
    _, gp, _ = step_plain()
    graph.replay()
    torch.cuda.synchronize()
    agree = float((gx - gp).abs().max())
    agree_graph = float((aouts[0] - gx).abs().max())

    # plain per-molecule GPU path: one molecule per launch, inputs resident in
    # HBM (views into the batch's packed tensors) and outputs preallocated per
    # molecule, on every 8th molecule of the workload (all sizes); each
    # molecule is its own one-molecule batch, created before the timed loop
    sample = list(range(0, b.nmol, 8))
    singles = []
    for m in sample:
        a0, a1 = int(b.atom_off[m]), int(b.atom_off[m + 1])
        s0, s1 = int(b.sh_off[m]), int(b.sh_off[m + 1])
        o0, o1 = int(b.ao_off[m]), int(b.ao_off[m + 1])
        d0, d1 = int(b.dense_off[m]), int(b.dense_off[m + 1])
        bm = Batch([mols[m]])
        ins = (xyz[3 * a0: 3 * a1], se[s0:s1], dsedcn[s0:s1], P[d0:d1], W[d0:d1],
               vao[o0:o1], vdp[3 * a0: 3 * a1], vqp[6 * a0: 6 * a1])
        nd = d1 - d0
        fo = tuple(torch.empty(c * nd, dtype=torch.float64, device=dev) for c in (1, 3, 6, 1))
        go = tuple(torch.empty(n, dtype=torch.float64, device=dev)
                   for n in (a1 - a0, 3 * (a1 - a0), 9))
        singles.append((bm, ins, fo, go))

    def step_single():
        for bm, (x_, s_, ds_, p_, w_, va_, vd_, vq_), fo, go in singles:
            bm.forward(x_, s_, out=fo)
            bm.gradient(x_, s_, ds_, p_, w_, va_, vd_, vq_, out=go)

    step_single()
    torch.cuda.synchronize()
    t = time.perf_counter()
    reps = 3
    for _ in range(reps):
        step_single()
    torch.cuda.synchronize()
    plain_rate = reps * len(singles) / (time.perf_counter() - t)
    log(f"plain per-molecule GPU (device-resident): {plain_rate:.0f} molecules/s")
    del singles

    # drop-in host-pointer entry points (PCIe copies, host work-unit plans
    # cached across calls of the same molecule) on the first 64 molecules
    nsample = min(64, len(mols))
    rng = np.random.default_rng(3)
    dropin = []
    for z, x in mols[:nsample]:
        st = structure.setup(structure.Structure(z, x))
        nao_m, nat_m = st.bas.nao, st.mol.nat
        sem, dsem = st.selfenergy(rng.uniform(0.5, 3.5, nat_m))
        a = rng.normal(size=(nao_m, nao_m))
        Pm = np.asfortranarray((0.5 * (a + a.T))[:, :, None])
        dropin.append((st, sem, dsem, Pm))

    def run_dropin():
        for st, sem, dsem, Pm in dropin:
            nao_m, nat_m = st.bas.nao, st.mol.nat
            h0mod.get_hamiltonian(st, sem)
            h0mod.get_hamiltonian_gradient(st, sem, dsem, np.zeros((nao_m, 1), order="F"),
                                           np.zeros((3, nat_m, 1), order="F"),
                                           np.zeros((6, nat_m, 1), order="F"), Pm, 0.05 * Pm,
                                           np.zeros(nat_m), np.zeros((3, nat_m), order="F"),
                                           np.zeros((3, 3), order="F"))
    t = time.perf_counter()
    run_dropin()
    dropin_rate = nsample / (time.perf_counter() - t)

    # SCF-loop consumers of the resident S / Q (one HBM pass each, nspin = 1):
    # kernel time from the batch's events, algorithmic bytes per dense element
    S_r, _, Q_r, _ = outs
    h1 = torch.zeros(b.ndense, dtype=torch.float64, device=dev)
    n0 = torch.zeros(b.nsh, dtype=torch.float64, device=dev)
    vmp = torch.randn(6 * b.nat, generator=g, dtype=torch.float64, device=dev) * 1e-3
    scf = {}
    for name, key, fn, per_elem in (
            ("add_vao_to_h1", "add_vao_ms", lambda: b.add_vao_to_h1(S_r, vao, h1), 24),
            ("add_vmp_to_h1 (quadrupole)", "add_vmp_ms",
             lambda: b.add_vmp_to_h1(Q_r, 6, vmp, 6, h1), 64),
            ("get_mulliken_shell_charges", "shell_charges_ms",
             lambda: b.mulliken_shell_charges(S_r, P, n0), 16),
            ("get_mulliken_atomic_multipoles (quadrupole)", "atomic_multipoles_ms",
             lambda: b.mulliken_atomic_multipoles(Q_r, 6, P), 56)):
        fn()
        torch.cuda.synchronize()
        ms = []
        for _ in range(5):
            fn()
            torch.cuda.synchronize()
            ms.append(b.last_scf_timing()[key])
        t_ms = float(np.median(ms))
        gbs = per_elem * b.ndense / (t_ms * 1e-3) / 1e9
        scf[name] = {"ms": t_ms, "bytes_per_dense_element": per_elem, "GB/s": gbs,
                     "frac_hbm_peak": gbs / roofline.PEAK_HBM_GBS}
    log("scf consumers: " + ", ".join(f"{k} {v['ms']:.3f} ms ({v['GB/s']:.0f} GB/s)"
                                      for k, v in scf.items()))
    del h1

    # the path's neighbours on the same batch, timed with events on the
    # launch stream (median of 5): coordination number and its derivative,
    # Fermi filling, the density products (FP64 MFMA; P is symmetric, so the
    # algorithmic FLOP is that of a symmetric rank-nao product, nao^2 (nao + 1)
    # per molecule and matrix), and the generalized eigensolver on a sample
    # (one rocSOLVER dsygvd call per molecule)
    def ev_ms(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        t = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            t.append(e0.elapsed_time(e1))
        return float(np.median(t))

    nao_m = np.diff(b.ao_off).astype(np.float64)
    nat_m = np.diff(b.atom_off).astype(np.float64)
    rows = {}
    cn_out = torch.empty(b.nat, dtype=torch.float64, device=dev)
    t_cn = ev_ms(lambda: b.ncoord(xyz, cn_out))
    dEdcn = torch.randn(b.nat, generator=g, dtype=torch.float64, device=dev)
    gtmp = torch.zeros(3 * b.nat, dtype=torch.float64, device=dev)
    stmp = torch.zeros(9 * b.nmol, dtype=torch.float64, device=dev)
    t_dcn = ev_ms(lambda: b.ncoord_gradient(xyz, dEdcn, gtmp, stmp))
    rows["ncoord"] = {"ms": t_cn, "atom_pairs": float((nat_m ** 2).sum()),
                      "atom_pairs_per_s": float((nat_m ** 2).sum()) / (t_cn * 1e-3)}
    rows["ncoord_gradient"] = {"ms": t_dcn}
    emo, _ = torch.sort(torch.randn(b.nao, generator=g, dtype=torch.float64, device=dev))
    for m in range(b.nmol):  # ascending within each molecule
        a0, a1 = int(b.ao_off[m]), int(b.ao_off[m + 1])
        emo[a0:a1] = torch.sort(emo[a0:a1])[0]
    nel = torch.as_tensor(0.5 * nao_m, dtype=torch.float64, device=dev)
    focc, _ = b.fermi_filling(nel, 300.0 * 3.166808578545117e-6, emo)
    rows["fermi_filling"] = {"ms": ev_ms(lambda: b.fermi_filling(nel, 9.5e-4, emo))}
    coeff = outs[3]  # any dense matrix serves as coefficients for the product's timing
    Pd = torch.empty(b.ndense, dtype=torch.float64, device=dev)
    t_p = ev_ms(lambda: b.density_matrix(coeff, focc, Pd))
    t_w = ev_ms(lambda: b.wdensity(coeff, focc, emo, Pd))
    syrk_flop = float((nao_m ** 2 * (nao_m + 1.0)).sum())
    rows["density_matrix"] = {"ms": t_p, "algorithmic_flop": syrk_flop,
                              "TFLOP/s": syrk_flop / (t_p * 1e-3) / 1e12,
                              "frac_fp64_peak": syrk_flop / (t_p * 1e-3) / 1e12
                              / roofline.PEAK_FP64_TFLOPS}
    rows["wdensity"] = {"ms": t_w, "TFLOP/s": syrk_flop / (t_w * 1e-3) / 1e12}
    del Pd
    nsy = min(32, b.nmol)
    bs = Batch(mols[:nsy], order_pairs=False)
    Ss, _, _, Hs = bs.forward(bs.tensor_xyz(), se[: int(b.sh_off[nsy])].contiguous())
    Cs = Hs.clone()
    t_sy = ev_ms(lambda: (Cs.copy_(Hs), bs.sygvd(Cs, Ss)), reps=3)
    rows["sygvd_sample"] = {"ms": t_sy, "molecules": nsy,
                            "ms_per_molecule": t_sy / nsy,
                            "nao_range": [int(np.diff(bs.ao_off).min()),
                                          int(np.diff(bs.ao_off).max())]}
    del bs, Ss, Hs, Cs
    log("rows: " + ", ".join(f"{k} {v['ms']:.3f} ms" for k, v in rows.items()))

    ms_step = el_a / args.steps * 1e3
    value = b.nmol / (el_a / args.steps)
    # dominant kernel of the autograd step.  The adjoint's derivative work is
    # over the off-site primitive pairs only (its on-site part reads a
    # per-species table), so its algorithmic FLOP count uses K.c off-site.
    fwd_bytes = roofline.BYTES_FWD_PER_ELEM * b.ndense
    adj_flop = roofline.FLOP_BWD_PER_KC * kc_off
    ach = adj_flop / (adj_a * 1e-3) / 1e12
    roof_adj = roofline.roofline_entry("fp64", ach, roofline.PEAK_FP64_TFLOPS, "TFLOP/s",
                                       profiled_traffic("k_gradient<adjoint>", b.nmol))
    roof_adj["bound_note"] = ("FP64 vector ALU (the kernel issues no MFMA); peak = MI355X FP64 "
                              "78.6 TFLOP/s")
    roof_adj["kernel"] = "k_gradient<adjoint> (FP64 vector, given the forward's S and D)"
    roof_adj["kernel_ms"] = adj_a
    roof_adj["algorithmic_flop"] = adj_flop
    roof_adj["achieved_incl_onsite"] = roofline.FLOP_BWD_PER_KC * kc / (adj_a * 1e-3) / 1e12
    ach = fwd_bytes / (fwd_a * 1e-3) / 1e9
    roof_fwd = roofline.roofline_entry("hbm", ach, roofline.PEAK_HBM_GBS, "GB/s",
                                       profiled_traffic("k_forward", b.nmol))
    roof_fwd["kernel"] = "k_forward"
    roof_fwd["kernel_ms"] = fwd_a
    roof_fwd["algorithmic_bytes"] = fwd_bytes
    # the dominant kernel of the autograd step is `roofline`; the other kernel
    # of the step is reported beside it
    roof, roof_other = (roof_adj, roof_fwd) if adj_a >= fwd_a else (roof_fwd, roof_adj)
    line = {
        "metric": "molecules/s for GFN2 S/D/Q+H0+gradient, batched-autograd vs plain GPU; "
                  "% FP64 peak",
        "value": value,
        "unit": "molecules/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded random molecules; random symmetric P, W and potentials)",
        "config": {"workload": f"batched GFN2 S/D/Q+H0 forward + autograd gradient, "
                               f"{b.nmol} molecules, 8-256 atoms log-uniform",
                   "molecules": b.nmol, "atoms": b.nat, "nao": b.nao,
                   "dense_elements": b.ndense, "kept_prim_x_cart_pairs": kc,
                   "parallelism": "single GPU"},
        "graph_ms_per_step": el_g / args.steps * 1e3,
        "graph_molecules_per_s": b.nmol / (el_g / args.steps),
        "batched_plain_molecules_per_s": b.nmol / (el_p / args.steps),
        "plain_gpu_molecules_per_s": plain_rate,
        "plain_gpu_sample": f"every 8th molecule ({len(sample)}), one molecule per launch: "
                            f"forward + plain gradient kernels on inputs resident in HBM",
        "batched_autograd_over_plain_gpu": value / plain_rate,
        "dropin_pcie_molecules_per_s": dropin_rate,
        "dropin_pcie_sample": f"first {nsample} molecules, host-pointer get_hamiltonian + "
                              f"get_hamiltonian_gradient (PCIe copies included, work-unit "
                              f"plans built once per molecule)",
        "kernel_ms": {"forward": fwd_a, "adjoint": adj_a, "gradient": grad_p},
        "forward_hbm_GBs": fwd_bytes / (fwd_a * 1e-3) / 1e9,
        "adjoint_fp64_TFLOPs": adj_flop / (adj_a * 1e-3) / 1e12,
        "gradient_fp64_TFLOPs": adj_flop / (grad_p * 1e-3) / 1e12,
        "kept_prim_x_cart_pairs_offsite": kc_off,
        "autograd_vs_plain_max_abs_diff": agree,
        "graph_vs_autograd_max_abs_diff": agree_graph,
        "tiles": b.tiles(),
        "scf_consumers": scf,
        "next_rows": rows,
        "roofline": roof,
        "roofline_other_kernel": roof_other,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
