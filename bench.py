#!/usr/bin/env python3
"""Benchmark: R^k integrals/s and Fock-gradient time on one MI355X.

Workload (the headline configuration): E = sum_n c_n R^k(a,b;c,d) over 30
orbitals on an ns=200, ks=9 (ksp=8, ksq=9) exponential grid (Z=26, r_max=30),
k = 0..8, 5e4 seeded rows (c ~ U[-1,1], k, a, b, c, d uniform).  One step =
the batched forward (every R^k value + E) and adjoint (dE/d(P,Q) of all
orbitals, the direct+exchange potentials) replayed as one captured hipGraph
with the orbitals and coefficients resident in HBM.  value = rows / step.

Also reported: the plain GPU path (materialised R^k arrays + convol + SUM_AmB,
forward only) on the same rows, the per-phase roofline of the dominant
kernel (device time from HIP events on the engine's stream), and the CPU
restatement of the reference algorithm timed on a bounded sample on this
host (one multipole's rows).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFS = 78.6
SEED = 20261203 + 4


def make_workload(ns=200, ks=9, ksp=8, ksq=9, norb=30, kmax=8, nrows=50000, seed=SEED):
    from dbsr3_amd.grid import mkgrid_02
    t = mkgrid_02(ns, ks, 26.0, 0.25, 30.0)
    rng = np.random.default_rng(seed)
    nv = ns - ks + 1
    nsp, nsq = nv + ksp - 1, nv + ksq - 1
    pq = np.zeros((ns, 2, norb), order="F")
    for o in range(norb):
        pq[1:nsp - 1, 0, o] = rng.standard_normal(nsp - 2)
        pq[1:nsq - 1, 1, o] = rng.standard_normal(nsq - 2)
    rows = np.zeros((nrows, 5), dtype=np.int32)
    rows[:, 0] = rng.integers(0, kmax + 1, nrows)
    rows[:, 1:] = rng.integers(1, norb + 1, (nrows, 4))
    coef = rng.uniform(-1.0, 1.0, nrows)
    return t, pq, rows, coef


def normalise_on_gpu(d, pq):
    from dbsr3_amd.dbs import fortran_to_device
    out = pq.copy(order="F")
    for o in range(pq.shape[2]):
        f = fortran_to_device(out[:, :, o])
        n = d.quadr(f, f, 0)
        out[:, :, o] /= np.sqrt(n)
    return out


def cpu_baseline(t, ns, ks, ksp, ksq, pq, rows, budget_s):
    """The oracle (C restatement of rk()/Rk_evaluate) on the rows of one
    multipole, single thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    g = O.Grid(t, ns, ks, ksp, ksq)
    fs = [np.asfortranarray(pq[:, :, o]) for o in range(pq.shape[2])]
    k0 = int(rows[0, 0])
    sample = rows[rows[:, 0] == k0]
    t0 = time.perf_counter()
    vals = O.rk_rows(g, fs, sample, deadline=t0 + budget_s)
    t1 = time.perf_counter()
    n = len(vals)
    return n, t1 - t0, sample[:n], vals, k0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--rows", type=int, default=50000)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine shards by instance/row blocks "
              "but no multi-GPU figure is produced in this build (use --gpus 1)", file=sys.stderr)
        return 2

    import torch
    from dbsr3_amd import DBS, Engine

    ns, ks, ksp, ksq, norb = 200, 9, 8, 9, 30
    t, pq0, rows, coef = make_workload(nrows=args.rows)
    d = DBS(t, ns, ks, ksp, ksq)
    pq = normalise_on_gpu(d, pq0)
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq)
    eng.set_coef(0, coef)
    tuning = eng.autotune(with_adjoint=True, reps=3) if not args.no_autotune else eng.get_tuning()
    eng.graph_capture(with_adjoint=True)
    for _ in range(args.warmup):
        eng.graph_launch()
    eng.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eng.graph_launch()
    eng.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    value = len(rows) / (ms * 1e-3)
    vals = eng.values()
    energy = eng.energy()

    # per-phase device time and the dominant kernel's roofline
    prof = eng.profile(with_adjoint=True, reps=5)
    fwd_only = eng.profile(with_adjoint=False, reps=5)
    fock_ms = sum(v[0] for v in prof.values())
    dom = max((k for k in prof if k != "start"), key=lambda k: prof[k][0])
    dms, dfl, dby = prof[dom]
    intensity = dfl / dby if dby > 0 else float("inf")
    if intensity > PEAK_FP64_TFS * 1e12 / (PEAK_HBM_GBS * 1e9):
        roof = {"bound": "mfma", "achieved": dfl / (dms * 1e-3) / 1e12, "peak": PEAK_FP64_TFS,
                "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": dby / (dms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS,
                "unit": "GB/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None
    roof["kernel"] = dom
    roof["kernel_ms"] = dms

    out_extra = {}
    if not args.no_plain:
        eng.forward_plain()  # builds and caches the arrays
        torch.cuda.synchronize()
        p0 = time.perf_counter()
        eng.forward_plain()
        p1 = time.perf_counter()
        plain_ms = (p1 - p0) * 1e3
        pvals = eng.values()
        rel = np.abs(vals - pvals) / np.maximum(np.abs(pvals), 1e-3 * np.max(np.abs(pvals)))
        out_extra.update({"plain_gpu_integrals_per_s": len(rows) / (plain_ms * 1e-3),
                          "plain_gpu_ms": plain_ms,
                          "batched_vs_plain": value / (len(rows) / (plain_ms * 1e-3)),
                          "max_rel_err_vs_plain": float(rel.max())})
        eng.forward_adjoint()
        eng.sync()

    cpu = None
    if not args.no_cpu:
        n, secs, sample, cvals, k0 = cpu_baseline(t, ns, ks, ksp, ksq, pq, rows, args.cpu_seconds)
        sel = np.nonzero(rows[:, 0] == k0)[0][:n]
        rel = np.abs(vals[sel] - cvals) / np.maximum(np.abs(cvals), 1e-3 * np.max(np.abs(cvals)))
        cpu = {"value": n / secs, "unit": "R^k integrals/s", "cores": 1, "kind": "port",
               "sample": f"{n} rows with k={k0} (arrays built inside the timed region), "
                         f"{secs:.1f} s single-thread"}
        out_extra["max_rel_err_vs_cpu_oracle"] = float(rel.max())
        out_extra["batched_vs_cpu"] = value / cpu["value"]

    line = {
        "metric": "R^k integrals/s + Fock-gradient time, batched vs plain FP64 GPU, <=1e-12 rel err",
        "value": value,
        "unit": "R^k integrals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded orbitals and coefficient rows)",
        "config": {"workload": "E=sum c*R^k forward+adjoint: 30 orbitals, ns=200, ks=9/8/9, "
                               "k=0..8, 5e4 rows, one hipGraph per step",
                   "rows": int(len(rows)), "ns": ns, "ks": ks, "norb": norb, "kmax": 8,
                   "tuning": tuning},
        "fock_gradient_ms": fock_ms,
        "forward_ms": sum(v[0] for v in fwd_only.values()),
        "energy": energy,
        "phases_ms": {k: round(v[0], 5) for k, v in prof.items()},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    line.update(out_extra)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
