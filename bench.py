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


# engine phase -> the (main) kernel that runs it for this workload ({w}:
# pair-panel width)
PHASE_KERNEL = {
    "orb_points": "k_orb_values", "coef_gather": "k_adj_cgather", "features": "k_features_mfma",
    "prefix_scan": "k_scan_cols", "pair_panel": "{pair}<{w}, {v}, true>",
    "contract_rows": "{pair}<{w}, {v}, false>", "rows_reduce": "k_rows_reduce",
    "contract_dense": "k_contract_dense", "energy": "k_energy",
    "adj_suffix_scan": "k_scan_cols", "adj_orb_points": "k_adj_orb_acc<2>", "adj_coef": "k_adj_proj",
}
# rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload (HBM bytes per
# dispatch, copy-kernel scaled); rows "| kernel | tuning | ... | bytes |"
# (the kernel name carries the panel width, the tuning column is informative)
TRAFFIC_FILE = "engine_rk_fwd_adj_ns200_norb30_rows5e4_pmc.md"


def pair_kernel(rows, tuning):
    """the pair-panel kernel the engine launches for this row set (total
    densities of each multipole; one density per thread when they fit a
    workgroup of up to 512)"""
    nd = 0
    for k in np.unique(rows[:, 0]):
        r = rows[rows[:, 0] == k]
        pairs = {(min(a, c), max(a, c)) for a, c in r[:, [1, 3]]} | {(min(b, d), max(b, d)) for b, d in r[:, [2, 4]]}
        nd = max(nd, len(pairs))
    w = tuning[2]
    if w >= 8 and nd <= 512:
        return "k_pair_rot", "false"
    return "k_pair_panel", "true"


def measured_traffic(path, kernel, tuning):
    """HBM bytes per launch of `kernel` (name with its template arguments)
    from the committed counter summary of this workload, else None."""
    try:
        with open(path) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) >= 6 and cells[0] == kernel:
                    return float(cells[-1])
    except (OSError, ValueError):
        pass
    return None


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
    """The oracle (C restatement of mrk_*/convol/SUM_AmB as rk() and
    Rk_evaluate use them) on the rows taken in multipole order, single
    thread, until the time budget is spent: the R^k arrays of each k are
    built inside the timed region as the reference builds them on first use."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    g = O.Grid(t, ns, ks, ksp, ksq)
    fs = [np.asfortranarray(pq[:, :, o]) for o in range(pq.shape[2])]
    order = np.argsort(rows[:, 0], kind="stable")
    t0 = time.perf_counter()
    vals = O.rk_rows(g, fs, rows[order], deadline=t0 + budget_s)
    t1 = time.perf_counter()
    n = len(vals)
    return n, t1 - t0, order[:n], vals


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
    ap.add_argument("--tuning", type=str, default="",
                    help="comma-separated engine tuning (skips the autotuner)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the engine, copy MIB MiB device-to-device 3 times "
                         "(a plain copy kernel of known size for hardware-counter scaling)")
    ap.add_argument("--traffic-file", type=str,
                    default=os.path.join(ROOT, "profiles", TRAFFIC_FILE))
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine shards by instance/row blocks "
              "but no multi-GPU figure is produced in this build (use --gpus 1)", file=sys.stderr)
        return 2

    import torch
    from dbsr3_amd import DBS, Engine

    if args.copy_calibration > 0:
        n = args.copy_calibration * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ns, ks, ksp, ksq, norb = 200, 9, 8, 9, 30
    t, pq0, rows, coef = make_workload(nrows=args.rows)
    d = DBS(t, ns, ks, ksp, ksq)
    pq = normalise_on_gpu(d, pq0)
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq)
    eng.set_coef(0, coef)
    if args.tuning:
        eng.set_tuning([int(x) for x in args.tuning.split(",")])
        tuning = eng.get_tuning()
    elif args.no_autotune:
        tuning = eng.get_tuning()
    else:
        tuning = eng.autotune(with_adjoint=True, reps=3)
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
    pk, pv = pair_kernel(rows, tuning)
    kname = PHASE_KERNEL.get(dom, dom).format(w=tuning[2], pair=pk, v=pv)
    roof["traffic"] = measured_traffic(args.traffic_file, kname, tuning)
    roof["kernel"] = kname
    roof["phase"] = dom
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
        n, secs, sel, cvals = cpu_baseline(t, ns, ks, ksp, ksq, pq, rows, args.cpu_seconds)
        rel = np.abs(vals[sel] - cvals) / np.maximum(np.abs(cvals), 1e-3 * np.max(np.abs(cvals)))
        kk = sorted(set(int(x) for x in rows[sel, 0]))
        cpu = {"value": n / secs, "unit": "R^k integrals/s", "cores": 1, "kind": "port",
               "sample": f"first {n} of the rows in multipole order (k={kk[0]}..{kk[-1]}; "
                         f"R^k arrays of each k built inside the timed region), "
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
