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

`--config N` measures one of the other integral-set configurations instead
(same JSON line format, not the headline):
  1  all R^k(a,b;c,d), a<=c, b<=d, k=0..4 of 12 orbitals (30 420 rows) and the
     S^k ppqq/pqqp sets, on the Z=26 recipe grid (ns=208) and the Fe knots
     (ns=113), forward only;
  2  MAT3 channel blocks: 2e5 seeded entries (40 channels, 20 bound orbitals,
     ns=150, k=0..6, itype 3 direct / 4 exchange, atype 0..3) accumulated by
     R_data into 820 symmetric 300x300 blocks;
  3  256 uneven instances (ns 60..400, ks 7..9, kmax 2..10, 4..8 orbitals,
     all pairs) forward+adjoint in one hipGraph, autotuned;
  5  SCF-style: ns=600, 60 orbitals, k=0..12, 1e6 rows, 20 iterations of
     new orbitals + one forward+adjoint hipGraph replay;
  6  the CI3 accumulation path: Add_integral (with Breit) of 2e5 terms,
     c_data, every list by Rk_evaluate / Sk_evaluate into HM/DM.
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
    "prefix_scan": "k_scan_cols", "pair_panel": "{pair}<{w}, {v}, true{d}>",
    "contract_rows": "{pair}<{w}, {v}, false{d}>", "rows_reduce": "k_rows_reduce",
    "contract_dense": "k_contract_dense", "energy": "k_energy",
    "adj_suffix_scan": "k_scan_cols", "adj_orb_points": "k_adj_orb_acc<2, 256>", "adj_coef": "k_adj_proj",
}
# rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this workload (HBM bytes per
# dispatch, copy-kernel scaled); rows "| kernel | tuning | ... | bytes |"
# (the kernel name carries the panel width, the tuning column is informative)
TRAFFIC_FILE = "engine_rk_fwd_adj_ns200_norb30_rows5e4_pmc.md"


def pair_kernel(rows, tuning):
    """the pair-panel kernel the engine launches for this row set (total
    densities of each multipole): the rotated kernel when the largest group
    fits its workgroup (512 threads, 1024 at W = 4) at one density per
    thread, or at two for W < 16; returns (name, ONE argument, trailing
    template arguments)"""
    nd = 0
    for k in np.unique(rows[:, 0]):
        r = rows[rows[:, 0] == k]
        pairs = {(min(a, c), max(a, c)) for a, c in r[:, [1, 3]]} | {(min(b, d), max(b, d)) for b, d in r[:, [2, 4]]}
        nd = max(nd, len(pairs))
    w = tuning[2]
    tmax = 1024 if w == 4 else 512
    if nd <= (tmax if w >= 16 else 2 * tmax):
        return "k_pair_rot", "false", ", 1" if nd <= tmax else ", 2"
    return "k_pair_panel", "true", ""


def measured_traffic(path, kernel, tuning):
    """(HBM bytes per launch of `kernel`, the engine tuning they were counted
    with) from the committed counter summary of this workload: the row of
    the same tuning, else a row of the same kernel (the name carries the
    panel width) under another tuning; (None, None) when there is none."""
    tune = ",".join(str(int(x)) for x in tuning)
    other = (None, None)
    try:
        with open(path) as f:
            for line in f:
                cells = [c.strip() for c in line.strip().strip("|").split("|")]
                if len(cells) >= 6 and cells[0] == kernel:
                    if cells[1] == tune:
                        return float(cells[-1]), tune
                    if other[0] is None:
                        other = (float(cells[-1]), cells[1])
    except (OSError, ValueError):
        pass
    return other


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


def host_cpu():
    """(threads the all-cores baseline uses, logical CPUs available, CPU model):
    OMP_NUM_THREADS when set (the GPU box sets it to its CPU share), else all
    CPUs this process may run on"""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    threads = int(os.environ.get("OMP_NUM_THREADS", n))
    return max(1, min(threads, n)), n, model


def cpu_rows(kind, t, ns, ks, ksp, ksq, pq, rows, budget_s, nthreads=1):
    """CPU evaluation of `rows` (rk() type sum) in Rk_evaluate's organisation
    (one convolution per left pair, DBSR_CI3/rk_evaluate.f90:31-56), lowest
    multipoles first, until the time budget is spent, with the arrays of
    each multipole built inside the timed region: kind 'reference' -- the
    reference's own density/convol/SUM_AmB (oracle/_ref) on one thread;
    'port' -- the oracle's OpenMP port of the same on nthreads threads.
    Returns (rows done, seconds, their indices in `rows`, values) or None."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    rs, perm = O.sort_rows_by_left_pair(rows)
    if kind == "reference":
        if not O.ref_available():
            return None
        R = O.ref()
        R.ref_setup(ns, ks, ksp, ksq, O._d(O.f64(t)))
        t0 = time.perf_counter()
        v = O.ref_rk_rows(pq, rs, budget_s)
    else:
        g = O.Grid(t, ns, ks, ksp, ksq)
        t0 = time.perf_counter()
        v = O.rk_rows_par(g, pq, rs, nthreads, budget_s)
    secs = time.perf_counter() - t0
    return len(v), secs, perm[:len(v)], v


def cpu_lines(t, ns, ks, ksp, ksq, pq, rows, budget_s, gpu_vals=None, note=""):
    """(cpu_baseline, all-cores baseline, extras): the reference's own
    routines on one thread and the port on all host cores, each on a bounded
    sample (whole multipoles, lowest first), with their agreement with the
    GPU values of the same rows"""
    threads, ncpu, model = host_cpu()
    extra = {}
    base = None
    ref = cpu_rows("reference", t, ns, ks, ksp, ksq, pq, rows, budget_s)
    par = cpu_rows("port", t, ns, ks, ksp, ksq, pq, rows, budget_s, threads)
    for name, res, cores, kind in (("reference", ref, 1, "reference"), ("port", par, threads, "port")):
        if res is None or res[0] == 0:
            continue
        n, secs, sel, v = res
        ks_done = sorted(set(int(x) for x in rows[sel, 0]))
        line = {"value": n / secs, "unit": "R^k integrals/s", "cores": cores, "kind": kind,
                "sample": (f"{'the reference' + chr(39) + 's own density/convol/SUM_AmB (oracle/_ref)' if kind == 'reference' else 'the oracle port (OpenMP over (k, type) units)'}"
                           f" in Rk_evaluate's organisation (one convolution per left pair), arrays built inside "
                           f"the timed region, on {n} of {len(rows)} rows (k={ks_done[0]}..{ks_done[-1]}), "
                           f"{secs:.1f} s, {cores} thread(s){note}")}
        if kind == "port":
            line.update({"nproc": ncpu, "cpu_model": model})
        if gpu_vals is not None:
            scale = np.max(np.abs(v)) if len(v) else 1.0
            extra[f"max_rel_err_vs_cpu_{name}"] = float(np.max(np.abs(gpu_vals[sel] - v) /
                                                             np.maximum(np.abs(v), 1e-3 * scale)))
        if kind == "reference":
            base = line
        else:
            extra["cpu_all_cores"] = line
    if base is None:
        base = extra.pop("cpu_all_cores", None)
    return base, extra


def step_work(eng, prof, ms, with_adjoint=True):
    """whole-step algorithmic work from the engine's per-phase counts (each
    phase: flops of the factored form; bytes read and written once) at the
    measured step time, with the minimum bytes of the problem (orbitals,
    coefficients, outputs)"""
    fl = sum(v[1] for v in prof.values())
    by = sum(v[2] for v in prof.values())
    _, bmin = eng.work(with_adjoint=with_adjoint)
    sec = ms * 1e-3
    return {"flops": fl, "tflops": fl / sec / 1e12, "fp64_frac": fl / sec / 1e12 / PEAK_FP64_TFS,
            "bytes_phases": by, "gbs_phases": by / sec / 1e9, "hbm_frac_phases": by / sec / 1e9 / PEAK_HBM_GBS,
            "bytes_min": bmin}


def roofline_of(prof):
    """the dominant phase of an engine profile and its roofline (bytes and
    flops from the engine's algorithmic counts, time from HIP events)"""
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
    roof["phase"] = dom
    roof["kernel_ms"] = dms
    return roof, dom


def time_graph(eng, with_adjoint, steps, warmup, before=None):
    """ms per replay of the engine's captured step (before(i) runs on the
    host ahead of replay i, e.g. new orbitals)"""
    import torch
    eng.graph_capture(with_adjoint=with_adjoint)
    for i in range(warmup):
        if before:
            before(i)
        eng.graph_launch()
    eng.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        if before:
            before(i)
        eng.graph_launch()
    eng.sync()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def seeded_orbitals(t, ns, ks, ksp, ksq, norb, seed):
    """P,Q ~ N(0,1) on 2..nsp-1 / 2..nsq-1 (first and last coefficients zero)"""
    rng = np.random.default_rng(seed)
    nv = ns - ks + 1
    nsp, nsq = nv + ksp - 1, nv + ksq - 1
    pq = np.zeros((ns, 2, norb), order="F")
    for o in range(norb):
        pq[1:nsp - 1, 0, o] = rng.standard_normal(nsp - 2)
        pq[1:nsq - 1, 1, o] = rng.standard_normal(nsq - 2)
    return pq


def all_pair_rows(norb, kmax):
    """every (k, a, b, c, d) with a <= c, b <= d, k = 0..kmax, in (k, (a,c), (b,d)) order"""
    a, c = np.triu_indices(norb)
    npair = len(a)
    L = np.repeat(np.arange(npair), npair)
    R = np.tile(np.arange(npair), npair)
    one = np.stack([a[L] + 1, a[R] + 1, c[L] + 1, c[R] + 1], axis=1)
    rows = np.zeros(((kmax + 1) * len(one), 5), dtype=np.int32)
    for k in range(kmax + 1):
        rows[k * len(one):(k + 1) * len(one), 0] = k
        rows[k * len(one):(k + 1) * len(one), 1:] = one
    return rows


CONFIG3_ROWS_MAX = 16384


def config3_instances(seed=20261203 + 3, ninst=256):
    """config 3: per instance ns ~ U{60..400}, ks ~ {7,8,9} (ksp = ks-1, ksq =
    ks), kmax ~ U{2..10}, 4..24 orbitals (seeded, P,Q ~ N(0,1), zero ends, not
    normalised) on an mkgrid_02 grid (Z=26, r_max=30); the integral set is
    every R^k(a,b;c,d) with a <= c, b <= d, k <= kmax, or a seeded subset of
    CONFIG3_ROWS_MAX of them (kept in order) where there are more; c ~ U[-1,1].
    Returns [(t, ns, ks, kmax, pq, rows, coef)]."""
    from dbsr3_amd.grid import mkgrid_02
    rng = np.random.default_rng(seed)
    out = []
    for i in range(ninst):
        ns = int(rng.integers(60, 401))
        ks = int(rng.choice([7, 8, 9]))
        kmax = int(rng.integers(2, 11))
        norb = int(rng.integers(4, 25))
        t = mkgrid_02(ns, ks, 26.0, 0.25, 30.0)
        pq = seeded_orbitals(t, ns, ks, ks - 1, ks, norb, seed + 1000 * (i + 1))
        rows = all_pair_rows(norb, kmax)
        if len(rows) > CONFIG3_ROWS_MAX:
            rows = rows[np.sort(rng.choice(len(rows), CONFIG3_ROWS_MAX, replace=False))]
        out.append((t, ns, ks, kmax, pq, rows, rng.uniform(-1.0, 1.0, len(rows))))
    return out


def config_line(cfg, value, unit, ms, steps, warmup, workload, extra):
    line = {"metric": f"config {cfg}: {unit}", "value": value, "unit": unit, "n_gpus": 1,
            "steps": steps, "warmup": warmup, "ms_per_step": ms, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic (seeded orbitals and coefficients)", "config": workload}
    line.update(extra)
    return line


def run_config1(args):
    import math
    import torch
    from dbsr3_amd import DBS, Engine
    from dbsr3_amd.grid import mkgrid_01
    norb, kmax = 12, 4
    rows = all_pair_rows(norb, kmax)
    grids = {"recipe": None, "fe": None}
    t, ns = mkgrid_01(9, 26.0, hi=1e-3, he=math.exp(0.125) - 1.0, hmax=0.25, rmax=30.0)
    grids["recipe"] = (t, ns)
    grids["fe"] = (np.loadtxt(os.path.join(ROOT, "tests", "golden", "fe_knot_t.txt")), 113)
    rates, main = {}, None
    for gname, (t, ns) in grids.items():
        d = DBS(t, ns, 9, 8, 9)
        pq = normalise_on_gpu(d, seeded_orbitals(t, ns, 9, 8, 9, norb, 20261203 + 1))
        for kind, kname in ((0, "R"), (5, "S_ppqq"), (6, "S_pqqp")):
            eng = Engine([(d, norb, rows)], kind=kind)
            eng.set_orbitals(0, pq)
            eng.set_coef(0, np.ones(len(rows)))
            ms = time_graph(eng, False, args.steps, args.warmup)
            r = {"ms": ms, "integrals_per_s": len(rows) / (ms * 1e-3)}
            if kind == 0 and gname == "recipe":
                prof = eng.profile(with_adjoint=False, reps=3)
                roof, _ = roofline_of(prof)
                work = step_work(eng, prof, ms, with_adjoint=False)
                vals = eng.values()
                eng.forward_plain()
                torch.cuda.synchronize()
                p0 = time.perf_counter()
                eng.forward_plain()
                plain_ms = (time.perf_counter() - p0) * 1e3
                r["plain_gpu_ms"] = plain_ms
                main = (ms, roof, t, ns, pq, vals, work)
            rates[f"{gname}_ns{ns}_{kname}"] = r
            eng.close()
    ms, roof, t, ns, pq, vals, work = main
    value = len(rows) / (ms * 1e-3)
    workload = {"workload": "all R^k/S^k of 12 orbitals, a<=c, b<=d, k=0..4 (30420 rows per set), "
                            "forward only, one hipGraph per step; value: R^k on the recipe grid",
                "rows": int(len(rows)), "norb": norb, "kmax": kmax, "grids": {"recipe": grids["recipe"][1],
                                                                             "fe": 113}}
    cpu, extra = (None, {}) if args.no_cpu else cpu_lines(t, ns, 9, 8, 9, pq, rows, args.cpu_seconds, vals)
    if cpu:
        extra["batched_vs_cpu"] = value / cpu["value"]
    if "cpu_all_cores" in extra:
        extra["batched_vs_cpu_all_cores"] = value / extra["cpu_all_cores"]["value"]
    return config_line(1, value, "R^k integrals/s", ms, args.steps, args.warmup, workload,
                       {"sets": rates, "roofline": roof, "step_work": work, "cpu_baseline": cpu, **extra})


CONFIG2 = dict(ns=150, ks=9, ksp=8, ksq=9, nch=40, nbound=20, kmax=6, nent=200000, ibo=2 ** 15)
# 0-based positions of the two channel orbitals in R^k(j1,j2;j3,j4): (1,3) and
# (2,4) give the direct structure Rk(i . j .), (1,4) and (2,3) the exchange Rk(i . . j)
CH_POS = ((0, 2), (1, 3), (0, 3), (1, 2))


def config2_lists(seed=20261203 + 2):
    """Config 2's MAT3 channel-channel R-integrals, built the reference's way:
    2e5 seeded terms (c ~ U[-1,1], k ~ U{0..6}, two of 20 bound orbitals, a
    channel block ich >= jch uniform over the 820 blocks of 40 channels, the
    two channel orbitals in a direct or exchange position; the second bound
    orbital and the parity of k are chosen so that every term obeys the
    parity rule on l, as an angular program only emits such terms) pass through the
    library's MAT3 Add_integral (parity rule on l, Idef_Rtype) into c_data
    (DBSR_MAT3/MOD_c_data.f90), whose Add_matrix hands every merged list
    (itype 3/4, kpol) to the Gen_matrix hook.  Returns (t, bound-orbital
    pq (not normalised), [(itype, kpol, (c, K1, K2, K3, K4))] in hook order)."""
    from dbsr3_amd import CData
    from dbsr3_amd.grid import mkgrid_02
    P = CONFIG2
    ns, ks, ksp, ksq, nch, nb, kmax = P["ns"], P["ks"], P["ksp"], P["ksq"], P["nch"], P["nbound"], P["kmax"]
    t = mkgrid_02(ns, ks, 26.0, 0.25, 30.0)
    pq = seeded_orbitals(t, ns, ks, ksp, ksq, nb, seed)
    rng = np.random.default_rng(seed)
    norb = nb + nch
    ipbs = np.zeros(norb + 1, dtype=np.int32)
    ipbs[nb + 1:] = np.arange(1, nch + 1)
    lorb = np.zeros(norb + 1, dtype=np.int32)
    lorb[1:] = rng.integers(0, 4, norb)
    n = P["nent"]
    c = rng.uniform(-1.0, 1.0, n)
    k = rng.integers(0, kmax + 1, n)
    blk = rng.integers(0, nch * (nch + 1) // 2, n)
    ich_of = np.concatenate([np.arange(j, nch + 1) for j in range(1, nch + 1)])
    jch_of = np.concatenate([np.full(nch + 1 - j, j) for j in range(1, nch + 1)])
    bound = rng.integers(1, nb + 1, (n, 2))
    form = rng.integers(0, 4, n)
    by_parity = [np.nonzero(lorb[1:nb + 1] % 2 == p)[0] + 1 for p in (0, 1)]
    alt = rng.integers(0, 1 << 30, n)
    lists = []
    cd = CData(4, 0, kmax, 2000, 1000, 100, 1e-12, lambda it, kk, lst: lists.append((it, kk, lst)))
    for e in range(n):
        j = [0, 0, 0, 0]
        p, q = CH_POS[form[e]]
        j[p], j[q] = nb + int(ich_of[blk[e]]), nb + int(jch_of[blk[e]])
        r = [x for x in range(4) if x not in (p, q)]
        j[r[0]], j[r[1]] = int(bound[e, 0]), int(bound[e, 1])
        p13, p24 = (lorb[j[0]] + lorb[j[2]]) % 2, (lorb[j[1]] + lorb[j[3]]) % 2
        if p13 != p24:  # second bound orbital of the other l parity
            cand = by_parity[(lorb[j[r[1]]] + 1) % 2]
            j[r[1]] = int(cand[alt[e] % len(cand)])
            p13 = (lorb[j[0]] + lorb[j[2]]) % 2
        kk = int(k[e])
        if kk % 2 != p13:
            kk = kk + 1 if kk < kmax else kk - 1
        cd.add_integral_mat(kk, *j, float(c[e]), lorb, ipbs, ibo=P["ibo"], mk=kmax)
    for it in (3, 4):
        for kk in range(kmax + 1):
            cd.add_matrix(it, kk)
    cd.close()
    return t, pq, lists


def config2_runs(lists):
    """runs of equal K1 per list (one convolution each), by itype"""
    out = {3: 0, 4: 0}
    for it, _, lst in lists:
        K1 = np.asarray(lst[1])
        out[it] += int(np.count_nonzero(K1[1:] != K1[:-1]) + 1) if len(K1) else 0
    return out


def config2_sample(lists, blocks, icc, ibo):
    """the lists cut down to the entries whose channel block is in `blocks`
    (order kept; blocks are independent, so their values are unchanged)"""
    keep = set(int(b) for b in blocks)
    out = []
    for it, kk, lst in lists:
        K1 = np.asarray(lst[1])
        b = icc[K1 // ibo - 1, K1 % ibo - 1]
        sel = np.array([int(x) in keep for x in b], dtype=bool)
        if sel.any():
            out.append((it, kk, tuple(np.asarray(x)[sel] for x in lst)))
    return out


def run_config2(args):
    import torch
    from dbsr3_amd import DBS, MatBlocks
    from dbsr3_amd.dbs import fortran_to_device
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    P = CONFIG2
    ns, ks, ksp, ksq, nch, nb, kmax, ibo = (P[x] for x in ("ns", "ks", "ksp", "ksq", "nch", "nbound", "kmax",
                                                            "ibo"))
    t, pq0, lists = config2_lists()
    d = DBS(t, ns, ks, ksp, ksq)
    pq = normalise_on_gpu(d, pq0)
    pq_dev = fortran_to_device(pq)
    fd = [fortran_to_device(pq[:, :, o]) for o in range(nb)]
    # QUADR_PQ(i,j,k) of the bound orbitals for the asymptotic coefficients
    qpq = np.zeros((nb, nb, kmax + 1), order="F")
    for i in range(nb):
        for j in range(nb):
            for m in range(kmax + 1):
                qpq[i, j, m] = d.quadr(fd[i], fd[j], m)
    m = MatBlocks(d, nch, 0, mk=kmax)
    nl = sum(len(l[2][0]) for l in lists)

    def sweep():
        # calls queued without waiting where the path allows it (itype 4,
        # itype 3 without the asymptotic coefficients); one wait at the end
        for it, kk, lst in lists:
            for atype in range(4):
                m.R_data(it, kk, atype, lst, pq_dev, nb, ibo=ibo, quadr_pq=qpq, sync=False)
        d.sync()

    sweep()  # warm: builds and caches the R^k arrays; blocks checked below
    torch.cuda.synchronize()
    rng = np.random.default_rng(20261203 + 2)
    blocks = sorted(set(rng.choice(m.nblocks, 7, replace=False) + 1) | {int(m.icc[5, 5])})
    got = {b: m.hch[b - 1].cpu().numpy().T.copy() for b in blocks}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sweep()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    value = nl / (ms * 1e-3)
    # algorithmic work of one sweep: per run one convolution (band: 2(ns ks)^2
    # flops; exchange case 3: 4 updates of 2 flops per rkb element, 8 ns^2 ks^2),
    # four P/Q types each; HBM: every rkb element read once per run
    runs = config2_runs(lists)
    flops = 4 * (runs[3] * 2.0 * (ns * ks) ** 2 + runs[4] * 8.0 * ns * ns * ks * ks)
    roof = {"bound": "mfma", "achieved": flops / (ms * 1e-3) / 1e12, "peak": PEAK_FP64_TFS, "unit": "TFLOP/s"}
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof.update({"traffic": None, "runs": runs, "flops_per_sweep": flops,
                 "note": "whole sweep (all lists, four P/Q types): densities, band and exchange convolutions, "
                         "UPDATE_HS; flops counted as the reference's convol loops do them"})
    # oracle on the sampled blocks (every entry of those blocks, every type)
    g = O.Grid(t, ns, ks, ksp, ksq)
    icc_s = np.zeros((nch, nch), dtype=np.int32, order="F")
    for i, b in enumerate(blocks):
        icc_s[m.icc == b] = i + 1
    sample = config2_sample(lists, blocks, m.icc, ibo)
    href = np.zeros((2 * ns, 2 * ns, len(blocks)), order="F")
    ns_entries = sum(len(l[2][0]) for l in sample)
    t0 = time.perf_counter()
    for it, kk, lst in sample:
        for atype in range(4):
            O.rdata_34(g, it, atype, kk, lst[:4], pq, ibo, icc_s, href)
    secs = time.perf_counter() - t0
    err = max(np.max(np.abs(got[b] - href[:, :, i])) / np.max(np.abs(href[:, :, i]))
              for i, b in enumerate(blocks))
    if not err <= 1e-12:
        raise SystemExit(f"config 2: channel blocks differ from the oracle (max rel err {err:.3e})")
    cpu = {"value": ns_entries / secs, "unit": "list entries/s", "cores": 1, "kind": "port",
           "sample": f"every entry of {len(blocks)} seeded channel blocks ({ns_entries} merged entries, four "
                     f"P/Q types each, arrays built inside the timed region), {secs:.1f} s single-thread"}
    workload = {"workload": "MAT3 channel blocks: 2e5 seeded R^k terms (40 channels, 20 bound orbitals, ns=150, "
                            "ks=9/8/9, k=0..6) through Add_integral + c_data into itype 3/4 lists; each list by "
                            "R_data for atype 0..3 into 820 blocks of 300x300; one step = every list once",
                "entries": P["nent"], "merged_entries": int(nl), "lists": len(lists), "ns": ns, "channels": nch}
    return config_line(2, value, "list entries/s", ms, args.steps, 1, workload,
                       {"roofline": roof, "cpu_baseline": cpu, "max_rel_err_sampled_blocks": err,
                        "sampled_blocks": [int(b) for b in blocks]})


def run_config3(args):
    from dbsr3_amd import DBS, Engine
    spec = config3_instances()
    insts, host = [], []
    for (t, ns, ks, kmax, pq0, rows, coef) in spec:
        d = DBS(t, ns, ks, ks - 1, ks)
        pq = normalise_on_gpu(d, pq0)
        insts.append((d, pq.shape[2], rows))
        host.append((t, ns, ks, pq, rows, coef))
    nrows = sum(len(h[4]) for h in host)
    eng = Engine(insts)
    for i, h in enumerate(host):
        eng.set_orbitals(i, h[3])
        eng.set_coef(i, h[5])
    if args.tuning:  # a fixed tuning (profiling runs without the autotuner's trials)
        eng.set_tuning([int(x) for x in args.tuning.split(",")])
        tuning = eng.get_tuning()
    else:
        tuning = eng.autotune(with_adjoint=True, reps=2)
    ms = time_graph(eng, True, args.steps, args.warmup)
    value = nrows / (ms * 1e-3)
    prof = eng.profile(with_adjoint=True, reps=3)
    roof, _ = roofline_of(prof)
    work = step_work(eng, prof, ms)
    extra = {}
    cpu = None
    if not args.no_cpu:
        # whole multipoles of the smallest instances first, until the budget is spent
        threads, ncpu, model = host_cpu()
        order = sorted(range(len(host)), key=lambda j: host[j][1])
        for kind, cores in (("reference", 1), ("port", threads)):
            n = 0
            t0 = time.perf_counter()
            for j in order:
                t, ns, ks, pq, rows, _ = host[j]
                left = args.cpu_seconds - (time.perf_counter() - t0)
                if left <= 0:
                    break
                res = cpu_rows(kind, t, ns, ks, ks - 1, ks, pq, rows, left, cores)
                if res is None:
                    break
                n += res[0]
            secs = time.perf_counter() - t0
            if n == 0:
                continue
            line = {"value": n / secs, "unit": "R^k integrals/s", "cores": cores, "kind": kind,
                    "sample": f"{n} rows of the smallest instances first, whole multipoles (arrays built inside "
                              f"the timed region), Rk_evaluate's organisation, {secs:.1f} s, {cores} thread(s); "
                              f"the GPU figure is forward+adjoint"}
            if kind == "reference":
                cpu = line
            else:
                line.update({"nproc": ncpu, "cpu_model": model})
                extra["cpu_all_cores"] = line
        if cpu is None:
            cpu = extra.pop("cpu_all_cores", None)
        if cpu:
            extra["batched_vs_cpu"] = value / cpu["value"]
        if "cpu_all_cores" in extra:
            extra["batched_vs_cpu_all_cores"] = value / extra["cpu_all_cores"]["value"]
    workload = {"workload": "256 uneven instances (ns 60..400, ks 7..9, kmax 2..10, 4..24 orbitals; all "
                            f"a<=c, b<=d pairs, or a seeded subset of {CONFIG3_ROWS_MAX} rows per instance) "
                            "forward+adjoint, one hipGraph per step, autotuned",
                "instances": len(spec), "rows": int(nrows), "tuning": tuning}
    return config_line(3, value, "R^k integrals/s", ms, args.steps, args.warmup, workload,
                       {"roofline": roof, "step_work": work, "cpu_baseline": cpu, "energy_0": eng.energy(0),
                        **extra})


def run_config5(args):
    from dbsr3_amd import DBS, Engine
    from dbsr3_amd.grid import mkgrid_02
    ns, ks, ksp, ksq, norb, kmax, nrows = 600, 9, 8, 9, 60, 12, 1_000_000
    t = mkgrid_02(ns, ks, 26.0, 0.25, 30.0)
    d = DBS(t, ns, ks, ksp, ksq)
    pq0 = normalise_on_gpu(d, seeded_orbitals(t, ns, ks, ksp, ksq, norb, 20261203 + 5))
    rng = np.random.default_rng(20261203 + 5)
    rows = np.zeros((nrows, 5), dtype=np.int32)
    rows[:, 0] = rng.integers(0, kmax + 1, nrows)
    rows[:, 1:] = rng.integers(1, norb + 1, (nrows, 4))
    coef = rng.uniform(-1, 1, nrows)
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq0)
    eng.set_coef(0, coef)
    # tuned once, as an SCF run would before its iterations
    if args.tuning:
        eng.set_tuning([int(x) for x in args.tuning.split(",")])
        tuning = eng.get_tuning()
    elif args.no_autotune:
        tuning = eng.get_tuning()
    else:
        tuning = eng.autotune(with_adjoint=True, reps=2)
    budget = {}
    if args.memory_budget_mb > 0:
        # the large-ns path: (instance, k) groups in batches whose features
        # and gradients fit the budget
        nb0, by0 = eng.memory()
        eng.set_memory_budget(int(args.memory_budget_mb * (1 << 20)))
        nb, by = eng.memory()
        budget = {"memory_budget_mb": args.memory_budget_mb, "batches": nb, "feature_bytes": by,
                  "feature_bytes_unbudgeted": by0}
    bump = np.zeros_like(pq0)
    bump[1:-2] = np.sin(np.arange(1, ns - 2))[:, None, None] * 1e-3
    # one SCF iteration = new orbitals (a 576 KB upload) + the captured step
    ms = time_graph(eng, True, 20, args.warmup, before=lambda i: eng.set_orbitals(0, pq0 + i * bump))
    value = nrows / (ms * 1e-3)
    prof = eng.profile(with_adjoint=True, reps=2)
    roof, _ = roofline_of(prof)
    work = step_work(eng, prof, ms)
    extra = {}
    cpu = None
    if not args.no_cpu:
        eng.set_orbitals(0, pq0)
        eng.forward()
        vals = eng.values()
        cpu, extra = cpu_lines(t, ns, ks, ksp, ksq, pq0, rows, args.cpu_seconds, vals,
                               "; the GPU figure is forward+adjoint")
        if cpu:
            extra["batched_vs_cpu"] = value / cpu["value"]
        if "cpu_all_cores" in extra:
            extra["batched_vs_cpu_all_cores"] = value / extra["cpu_all_cores"]["value"]
    workload = {"workload": "SCF-style: ns=600, ks=9/8/9, 60 orbitals, k=0..12, 1e6 rows; 20 iterations of "
                            "new orbitals + one forward+adjoint hipGraph replay (autotuned once)",
                "rows": nrows, "ns": ns, "norb": norb, "kmax": kmax, "iterations": 20, "tuning": tuning,
                **budget}
    return config_line(5, value, "R^k integrals/s", ms, 20, args.warmup, workload,
                       {"roofline": roof, "step_work": work, "cpu_baseline": cpu, **extra})


def run_config6(args):
    """the CI3 accumulation path: 2e5 seeded angular terms (c ~ U[-1,1],
    k ~ U{0..4}, a CSF pair ic >= jc of 400, four of 30 orbitals with
    seeded kappa) through the library's CI3 Add_integral with the Breit
    terms (DBSR_CI3/add_integral.f90) and c_data; every merged list by
    Rk_evaluate / Sk_evaluate on the device into HM(400,400) and DM; ns=200
    (ks=9/8/9).  One step = every list once."""
    import torch
    from dbsr3_amd import CData, DBS
    from dbsr3_amd.cdata import rk_evaluate
    from dbsr3_amd.dbs import fortran_to_device
    from dbsr3_amd.grid import mkgrid_02
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ns, ks, ksp, ksq, norb, ncsf, nterm, kmax = 200, 9, 8, 9, 30, 400, 200000, 4
    t = mkgrid_02(ns, ks, 26.0, 0.25, 30.0)
    d = DBS(t, ns, ks, ksp, ksq)
    pq = normalise_on_gpu(d, seeded_orbitals(t, ns, ks, ksp, ksq, norb, 20261203 + 6))
    rng = np.random.default_rng(20261203 + 6)
    kappa = np.concatenate([[0], rng.choice([-1, 1, -2, 2, -3, 3], norb)]).astype(np.int32)
    lists = []
    cd = CData(6, 0, kmax + 1, 2000, 1000, 100, 1e-12, lambda it, k, l: lists.append((it, k, l)))
    ic = rng.integers(1, ncsf + 1, nterm)
    jc = rng.integers(1, ncsf + 1, nterm)
    orbs = rng.integers(1, norb + 1, (nterm, 4))
    kk = rng.integers(0, kmax + 1, nterm)
    cc = rng.uniform(-1.0, 1.0, nterm)
    for n in range(nterm):
        i, j = max(ic[n], jc[n]), min(ic[n], jc[n])
        cd.add_integral_ci(int(kk[n]), int(i), int(j), *(int(x) for x in orbs[n]), float(cc[n]), kappa, mbreit=1)
    for it in range(1, 7):
        for k in range(kmax + 2):
            cd.add_matrix(it, k)
    cd.close()
    nl = sum(len(l[2][0]) for l in lists)
    pq_dev = fortran_to_device(pq)
    HM = torch.zeros((ncsf, ncsf), dtype=torch.float64, device="cuda")
    DM = torch.zeros(ncsf, dtype=torch.float64, device="cuda")

    def sweep():
        # lists queued without waiting: the host builds the next list's
        # tables while the device runs the current one
        for it, k, lst in lists:
            rk_evaluate(d, it, k, lst, pq_dev, norb, HM, ncsf, ncsf, DM, sync=False)
        d.sync()
    sweep()  # warm: builds the arrays
    torch.cuda.synchronize()
    hm = HM.cpu().numpy().T.copy()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sweep()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    value = nl / (ms * 1e-3)
    # the oracle on the lists of one multipole of each R type (CPU baseline and check)
    g = O.Grid(t, ns, ks, ksp, ksq)
    HMr, DMr = np.zeros((ncsf, ncsf), order="F"), np.zeros(ncsf)
    sample = [(it, k, lst) for it, k, lst in lists if k == 1 and it <= 4]
    ns_entries = sum(len(l[2][0]) for l in sample)
    t0 = time.perf_counter()
    for it, k, lst in sample:
        O.rk_evaluate(g, it, k, lst, pq, 2 ** 15, HMr, ncsf, DMr)
    secs = time.perf_counter() - t0
    HMs = torch.zeros((ncsf, ncsf), dtype=torch.float64, device="cuda")
    DMs = torch.zeros(ncsf, dtype=torch.float64, device="cuda")
    for it, k, lst in sample:
        rk_evaluate(d, it, k, lst, pq_dev, norb, HMs, ncsf, ncsf, DMs)
    err = float(np.max(np.abs(HMs.cpu().numpy().T - HMr)) / np.max(np.abs(HMr)))
    if not err <= 1e-12:
        raise SystemExit(f"config 6: HM differs from the oracle (max rel err {err:.3e})")
    cpu = {"value": ns_entries / secs, "unit": "list entries/s", "cores": 1, "kind": "port",
           "sample": f"the R^k lists of k=1 ({ns_entries} entries, arrays built inside the timed region), "
                     f"{secs:.1f} s single-thread"}
    workload = {"workload": "CI3 Add_integral (with Breit) of 2e5 seeded terms, 30 orbitals, 400 CSFs, ns=200, "
                            "k=0..4; every merged list by Rk_evaluate/Sk_evaluate into HM/DM; one step = every list",
                "terms": nterm, "merged_entries": int(nl), "lists": len(lists), "ns": ns, "csfs": ncsf}
    # algorithmic work of a sweep: per list the case-2 products of its
    # distinct left densities, 2 K^2 flops each (K = ns*kb: kb = ks for R^k,
    # 2ks-1 for S^k), and SUM_AmB, 2 K flops per entry
    flops = 0.0
    for it, k, lst in lists:
        cl = np.asarray(lst[0])
        keep = np.abs(cl) >= 1e-10
        K = ns * (ks if it <= 4 else 2 * ks - 1)
        flops += 2.0 * K * K * len(np.unique(np.asarray(lst[1])[keep])) + 2.0 * K * int(keep.sum())
    ach = flops / (ms * 1e-3) / 1e12
    roof = {"bound": "mfma", "achieved": ach, "peak": 78.6, "unit": "TFLOP/s", "frac": ach / 78.6,
            "traffic": None, "phase": "sweep", "flops_per_sweep": flops,
            "note": "whole sweep over the host-timed step (lists queued without waiting; the host's table "
                    "building is the longer side): the case-2 products (rocBLAS dgemm) and SUM_AmB"}
    return config_line(6, value, "list entries/s", ms, args.steps, 1, workload,
                       {"roofline": roof, "cpu_baseline": cpu, "max_rel_err_sampled_lists": err,
                        "hm_norm": float(np.abs(hm).sum())})


# edmax: the window's upper end (dbsr_par's Edmax; the Rb example keeps its
# channel solutions below -1 au with the core energy in the blocks): here the
# B-spline states above 1e4 au are disregarded, about 100 solutions per channel
CONFIG7 = dict(ns=150, ks=9, ksp=8, ksq=9, nch=40, z=26.0, nbound=3, kappas=(-1, 1, -2, 2, -3, 3, -4),
               cpu_channels=12, edmax=1e4)


def config7_case(d, seed=20261203 + 7):
    """Config 2's partial wave after accumulation, the input of MAT3's
    post-processing: 40 channels (kappa cycling over -1, 1, -2, 2, -3, 3, -4)
    on the ns=150 grid, the 820 blocks in alloc_dbsr_matrix's numbering
    (diagonal blocks first); diagonal block = the Dirac-Coulomb operator of
    the channel's kappa (point nucleus Z=26, MAT_dhl_pq on the device with
    the Bloch term, so symmetric) plus a seeded symmetric 1e-3 N(0,1) part,
    other blocks seeded 1e-2 N(0,1); overlaps fppqq; 3 seeded bound orbitals
    per kappa that vanish on the B-splines Zero_cond removes for that kappa
    (as bound orbitals do), B-orthonormal, as the orthogonality conditions
    (IBORT = 0 except for seeded pairs)."""
    from dbsr3_amd.dbs import device_to_fortran
    P = CONFIG7
    ns, nch = P["ns"], P["nch"]
    ms = 2 * ns
    rng = np.random.default_rng(seed)
    kap = [P["kappas"][i % len(P["kappas"])] for i in range(nch)]
    lch = [k if k > 0 else -k - 1 for k in kap]
    nuc = d.nucleus("point", P["z"])
    fppqq = np.asfortranarray(device_to_fortran(d.bspline_overlaps()[2]))
    dhl = {k: device_to_fortran(d.dhl_pq(nuc, k, bloch=True)) for k in sorted(set(kap))}
    nb = nch * (nch + 1) // 2
    hch = np.asfortranarray(rng.standard_normal((ms, ms, nb)) * 1e-2)
    for i in range(nch):
        R = hch[:, :, i] * 0.1
        hch[:, :, i] = dhl[kap[i]] + (R + R.T)
    kset = sorted(set(kap))
    nbound = P["nbound"] * len(kset)
    pv = np.zeros((ms, nbound + nch), order="F")
    kbs = []
    from dbsr3_amd.channels import Zero_cond
    for kp in kset:
        first = len(kbs)
        alive = Zero_cond(1, d, [kp if kp > 0 else -kp - 1], [kp])[0]
        for n in range(P["nbound"]):
            v = rng.standard_normal(ms) * alive
            for j in range(first, len(kbs)):
                v -= (pv[:, j] @ fppqq @ v) * pv[:, j]
            pv[:, len(kbs)] = v / np.sqrt(v @ fppqq @ v)
            kbs.append(kp)
    kbs = kbs + kap
    ipbs = [0] * nbound + list(range(1, nch + 1))
    ipch = list(range(nbound + 1, nbound + nch + 1))
    qv = np.asfortranarray(fppqq @ pv)
    skip = set((int(a), int(b)) for a, b in zip(rng.integers(nbound + 1, nbound + nch + 1, 8),
                                                rng.integers(1, nbound + 1, 8)))
    ibort = lambda i, j: 1 if (i, j) in skip else 0
    return dict(kap=kap, lch=lch, hch=hch, fppqq=fppqq, pv=pv, qv=qv, kbs=kbs, ipbs=ipbs, ipch=ipch,
                ibort=ibort)


def config7_icc(nch):
    """alloc_dbsr_matrix's numbering for idiag = 0 (DBSR_MAT3/updates.f90:20-38)"""
    icc = np.zeros((nch, nch), dtype=np.int32, order="F")
    n = 0
    for i in range(nch):
        n += 1
        icc[i, i] = n
    for i in range(1, nch):
        for j in range(i):
            n += 1
            icc[i, j] = icc[j, i] = n
    return icc


def config7_oracle(case, nch_s, ns, nsp, nsq):
    """the oracle's post-processing of channels 1..nch_s (their blocks are
    independent of the other channels): (projected hch, solutions, ipsol,
    records), the reference's sequence DBS_ORTH(+1), Diag_channels,
    DBS_ORTH(-1), Record_matrix"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    icc_f = config7_icc(CONFIG7["nch"])
    icc = config7_icc(nch_s)
    ms = 2 * ns
    nb = nch_s * (nch_s + 1) // 2
    h = np.zeros((ms, ms, nb), order="F")
    for i in range(nch_s):
        for j in range(i + 1):
            h[:, :, icc[i, j] - 1] = case["hch"][:, :, icc_f[i, j] - 1]
    dg = np.asfortranarray(np.repeat(case["fppqq"][:, :, None], nch_s, axis=2))
    orth = O.orth_conditions(nch_s, case["ipch"][:nch_s], case["ipbs"], case["kbs"], case["ibort"])
    O.dbs_orth(h, icc, 1, orth, case["pv"], case["qv"])
    iprm, _ = O.zero_cond(nch_s, ns, nsp, nsq, CONFIG7["ksp"], CONFIG7["ksq"], case["lch"], case["kap"])
    ipsol = O.diag_channels(h, dg, icc, iprm, ns, edmax=CONFIG7["edmax"])
    O.dbs_orth(h, icc, -1, orth, case["pv"], case["qv"])
    recs, ov = O.record_matrix(h, icc, dg, ipsol, 1)
    return h, dg, ipsol, recs, icc, iprm


def run_config7(args):
    """MAT3's channel-block post-processing of config 2's partial wave: one
    step = DBS_ORTH on the diagonal blocks, Diag_channels, DBS_ORTH on the
    other blocks and Record_matrix with the records left in HBM
    (DBSR_MAT3/sub1.f90:133-170, 258-265)"""
    import torch
    from dbsr3_amd import DBS, MatBlocks
    from dbsr3_amd.channels import (DBS_ORTH, Diag_channels, Zero_cond, orth_conditions, record_matrix_device,
                                    record_sizes)
    from dbsr3_amd.dbs import fortran_to_device
    from dbsr3_amd.grid import mkgrid_02
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    P = CONFIG7
    ns, ks, ksp, ksq, nch = P["ns"], P["ks"], P["ksp"], P["ksq"], P["nch"]
    ms = 2 * ns
    t = mkgrid_02(ns, ks, P["z"], 0.25, 30.0)
    d = DBS(t, ns, ks, ksp, ksq)
    case = config7_case(d)
    m = MatBlocks(d, nch, 0)
    assert np.array_equal(m.icc, config7_icc(nch))
    h0 = torch.from_numpy(np.ascontiguousarray(case["hch"].T)).cuda()
    d0 = torch.from_numpy(np.ascontiguousarray(np.repeat(case["fppqq"].T[None], nch, axis=0))).cuda()
    dg = torch.empty_like(d0)
    iprm, nhm = Zero_cond(nch, d, case["lch"], case["kap"])
    nort, iort = orth_conditions(case["ipch"], case["ipbs"], case["kbs"], case["ibort"])
    pv, qv = fortran_to_device(case["pv"]), fortran_to_device(case["qv"])
    bufs = {}

    def step():
        m.hch.copy_(h0)
        dg.copy_(d0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        DBS_ORTH(m, 1, nort, iort, pv, qv)
        t1 = time.perf_counter()
        ipsol = Diag_channels(m, dg, iprm, edmax=P["edmax"])
        t2 = time.perf_counter()
        DBS_ORTH(m, -1, nort, iort, pv, qv)
        t3 = time.perf_counter()
        if "c" not in bufs:
            nc, nw = record_sizes(m, ipsol)
            bufs["c"] = torch.zeros(max(nc, 1), dtype=torch.float64, device="cuda")
            bufs["w"] = torch.zeros(max(nw, 1), dtype=torch.float64, device="cuda")
        coff, woff, ov = record_matrix_device(m, dg, ipsol, bufs["c"], bufs["w"])
        t4 = time.perf_counter()
        return ipsol, coff, (t1 - t0, t2 - t1, t3 - t2, t4 - t3)

    for _ in range(max(args.warmup, 1)):
        ipsol, coff, _ = step()
    phases = np.zeros(4)
    for _ in range(args.steps):
        ipsol, coff, ph = step()
        phases += ph
    phases *= 1e3 / args.steps
    ms_step = float(phases.sum())
    nblocks = m.nblocks
    value = nblocks / (ms_step * 1e-3)
    # algorithmic work per phase
    nmo = int(np.max(nort))
    kk = [int(np.count_nonzero(iprm[i * ms:(i + 1) * ms])) for i in range(nch)]
    ipl = [int(x) for x in ipsol]
    written = [(i, j) for i in range(2, nch + 1) for j in range(1, i) if coff[i - 1, j - 1] >= 0]
    work = {
        "orth_diag": (nch * 8.0 * nmo * ms * ms, nch * 2.0 * ms * ms * 8),
        "diag_channels": (sum(7.0 * k ** 3 for k in kk), sum(2.0 * k * k * 8 + ms * ipl[i + 1] * 8
                                                              for i, k in enumerate(kk))),
        "orth_offdiag": ((nblocks - nch) * 8.0 * nmo * ms * ms, (nblocks - nch) * 2.0 * ms * ms * 8),
        "record_matrix": (sum(2.0 * ipl[i] * ms * ms + 2.0 * ipl[i] * ipl[j] * ms for i, j in written),
                          sum(ms * ms * 8.0 + ipl[i] * ipl[j] * 8.0 for i, j in written)),
    }
    names = list(work)
    phases_ms = {k: round(float(v), 4) for k, v in zip(names, phases)}
    dom = names[int(np.argmax(phases))]
    fl, by = work[dom]
    dt = phases_ms[dom] * 1e-3
    if fl / by > PEAK_FP64_TFS * 1e12 / (PEAK_HBM_GBS * 1e9):
        roof = {"bound": "mfma", "achieved": fl / dt / 1e12, "peak": PEAK_FP64_TFS, "unit": "TFLOP/s"}
    else:
        roof = {"bound": "hbm", "achieved": by / dt / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s"}
    roof.update({"frac": roof["achieved"] / roof["peak"], "traffic": None, "phase": dom,
                 "note": "host-timed phases (each call returns after its stream drains); flops: rank-2n "
                         "projection 8 n ms^2 per block, eigen-solve 7 kk^3 per channel (dpotrf, dsygst, "
                         "dsytrd, dstedc, dormtr, dtrsm), transform 2 ni ms^2 + 2 ni nj ms per record; bytes: "
                         "each block read and written once"})
    # parity and the CPU baseline: the oracle on channels 1..nch_s, one thread
    nch_s = P["cpu_channels"]
    from threadpoolctl import threadpool_limits
    t0 = time.perf_counter()
    with threadpool_limits(1):
        h_o, d_o, ip_o, recs_o, icc_s, iprm_s = config7_oracle(case, nch_s, ns, d.nsp, d.nsq)
    secs = time.perf_counter() - t0
    nb_s = nch_s * (nch_s + 1) // 2
    cpu = {"value": nb_s / secs, "unit": "channel blocks/s", "cores": 1, "kind": "port",
           "sample": f"the oracle's DBS_ORTH, Diag_channels (LAPACK dsygv through scipy), DBS_ORTH and "
                     f"Record_matrix on channels 1..{nch_s} ({nb_s} blocks), {secs:.1f} s, one thread"}
    gh = m.hch.cpu().numpy().T
    gd = dg.cpu().numpy().T
    errs = {}
    if not np.array_equal(ipsol[:nch_s + 1], ip_o):
        raise SystemExit("config 7: ipsol differs from the oracle")
    icc_f = m.icc
    errs["projected_blocks"] = max(
        float(np.max(np.abs(gh[:, :, icc_f[i, j] - 1] - h_o[:, :, icc_s[i, j] - 1]))
              / np.max(np.abs(h_o[:, :, icc_s[i, j] - 1])))
        for i in range(nch_s) for j in range(i))
    # energies against the channel's largest |E| (the solvers' accuracy scale)
    errs["energies"] = max(
        float(np.max(np.abs(gd[:ip_o[c + 1], ns - 1, c] - d_o[:ip_o[c + 1], ns - 1, c]))
              / np.max(np.abs(d_o[:ip_o[c + 1], ns - 1, c]))) for c in range(nch_s) if ip_o[c + 1] > 0)
    # the records against the oracle's transform of the GPU's own projected
    # blocks and solutions (the eigen-solvers' differences do not enter)
    icc_g = np.zeros((nch_s, nch_s), dtype=np.int32, order="F")
    for i in range(nch_s):
        for j in range(nch_s):
            icc_g[i, j] = icc_f[i, j]
    recs_g, _ = O.record_matrix(gh, icc_g, gd, ipsol[:nch_s + 1], 1)
    c = bufs["c"].cpu().numpy()
    er = 0.0
    for (key, r) in recs_g[:-1]:
        i, j = key
        o = coff[i - 1, j - 1]
        v = c[o:o + r.size].reshape(r.shape, order="F")
        er = max(er, float(np.max(np.abs(v - r)) / np.max(np.abs(r))))
    errs["records"] = er
    if not (errs["projected_blocks"] <= 1e-12 and errs["energies"] <= 1e-12 and er <= 1e-12):
        raise SystemExit(f"config 7: results differ from the oracle {errs}")
    workload = {"workload": "MAT3 channel post-processing of one partial wave: 40 channels, ns=150, ks=9/8/9, "
                            "820 blocks of 300x300 in HBM; DBS_ORTH (diagonal), Diag_channels (Edmax = 1e4 au), "
                            "DBS_ORTH (off-diagonal), Record_matrix (records left in HBM); one step = all four",
                "channels": nch, "blocks": int(nblocks), "ns": ns, "nsol": int(np.sum(ipsol[1:])),
                "conditions_per_channel": [int(x) for x in nort]}
    return config_line(7, value, "channel blocks/s", ms_step, args.steps, max(args.warmup, 1), workload,
                       {"roofline": roof, "phases_ms": phases_ms, "cpu_baseline": cpu,
                        "batched_vs_cpu": value / cpu["value"], "max_rel_err_sample": errs,
                        "phase_work": {k: {"flops": v[0], "bytes": v[1]} for k, v in work.items()}})


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
    ap.add_argument("--memory-budget-mb", type=float, default=0.0,
                    help="config 5: bytes of features + gradients held at once (MiB; 0 = all groups)")
    ap.add_argument("--config", type=int, default=4, choices=[1, 2, 3, 4, 5, 6, 7],
                    help="integral-set configuration (4: the headline workload)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine shards by instance/row blocks "
              "but no multi-GPU figure is produced in this build (use --gpus 1)", file=sys.stderr)
        return 2
    if args.config != 4:
        run = {1: run_config1, 2: run_config2, 3: run_config3, 5: run_config5, 6: run_config6,
               7: run_config7}[args.config]
        print(json.dumps(run(args)), flush=True)
        return 0

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
    pk, pv, pd = pair_kernel(rows, tuning)
    kname = PHASE_KERNEL.get(dom, dom).format(w=tuning[2], pair=pk, v=pv, d=pd)
    roof["traffic"], roof["traffic_tuning"] = measured_traffic(args.traffic_file, kname, tuning)
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

    cpu, cpu_extra = (None, {}) if args.no_cpu else cpu_lines(t, ns, ks, ksp, ksq, pq, rows, args.cpu_seconds,
                                                               vals)
    out_extra.update(cpu_extra)
    if cpu:
        out_extra["batched_vs_cpu"] = value / cpu["value"]
    if "cpu_all_cores" in out_extra:
        out_extra["batched_vs_cpu_all_cores"] = value / out_extra["cpu_all_cores"]["value"]
    out_extra["step_work"] = step_work(eng, prof, ms)

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
