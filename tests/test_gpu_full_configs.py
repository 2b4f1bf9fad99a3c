"""GPU parity at the full sizes of the integral-set configurations.

* config 4 (the headline): 30 orbitals, ns=200, ks=9/8/9, k=0..8, 5e4 seeded
  rows -- every row against the plain GPU path (materialised arrays + convol
  + SUM_AmB, itself bit-exact to the oracle), a sample of rows of every
  multipole against the oracle, the graph replay against the eager run (bit
  for bit), E = sum c*R, central differences of E on 200 seeded coefficients
  at h = 1e-6, and, at the same shape, the adjoint against hf_matrix's
  direct+exchange potentials (DBSR_HF/hf_matrix.f90) for an a(i,j,k) /
  b(i,j,k) energy over the 30 orbitals;
* config 2: 2e5 seeded MAT3 terms through the library's Add_integral and
  c_data, R_data for atype 0..3 into the 820 channel blocks of 40 channels
  (ns=150, k=0..6) -- seeded blocks against the oracle's R_data, every block
  by contraction with channel functions U, W (U^T H W = sum c*R^k with U, W in
  the channel slots, evaluated by the engine, an independent GPU path), and
  the asymptotic coefficients of every block against the oracle.

Integral errors are reported per test as the floored relative error, the
number of entries under the floor with their true relative error, and the
error against each integral's Cauchy-Schwarz bound (conftest.err_report).
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import cs_bound, err_report

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _normalised(g, pq):
    out = np.asfortranarray(pq.copy())
    for o in range(out.shape[2]):
        out[:, :, o] /= math.sqrt(g.quadr(out[:, :, o], out[:, :, o], 0))
    return out


@pytest.fixture(scope="module")
def config4(gpu_available):
    from bench import make_workload
    from dbsr3_amd import DBS, Engine
    ns, ks, ksp, ksq, norb = 200, 9, 8, 9, 30
    t, pq0, rows, coef = make_workload()
    g = O.Grid(t, ns, ks, ksp, ksq)
    pq = _normalised(g, pq0)
    d = DBS(t, ns, ks, ksp, ksq)
    print(f"[config4 ns={ns} norb={norb} rows={len(rows)}]", flush=True)
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq)
    eng.set_coef(0, coef)
    return dict(t=t, g=g, d=d, pq=pq, rows=rows, coef=coef, eng=eng, norb=norb)


def test_config4_values(config4):
    c = config4
    eng, rows, coef = c["eng"], c["rows"], c["coef"]
    eng.graph_capture(with_adjoint=True)
    eng.graph_launch()
    eng.sync()
    vals, grad, E = eng.values(), eng.grad(), eng.energy()
    eng.forward_adjoint()
    assert np.array_equal(eng.values(), vals) and np.array_equal(eng.grad(), grad), "graph != eager"
    assert eng.energy() == E
    assert abs(E - vals @ coef) <= TOL * np.sum(np.abs(vals * coef))
    # every row against the plain path
    eng.forward_plain()
    plain = eng.values()
    cs = cs_bound(c["d"], c["norb"], c["pq"], rows)
    assert np.all(np.abs(plain) <= cs * (1 + 1e-12))
    r = err_report("config4 vs plain, 5e4 rows", vals, plain, scale=cs)
    assert r["floored"] < TOL and r["scaled"] < TOL
    # rows of every multipole against the oracle
    fs = [np.asfortranarray(c["pq"][:, :, o]) for o in range(c["norb"])]
    sel = np.concatenate([np.nonzero(rows[:, 0] == k)[0][:6] for k in range(9)])
    ref = O.rk_rows(c["g"], fs, rows[sel])
    r = err_report("config4 vs oracle, 6 rows per k", vals[sel], ref, scale=cs[sel])
    assert r["floored"] < TOL and r["scaled"] < TOL


def test_config4_finite_differences(config4):
    """central differences of E at h = 1e-6 on 200 seeded coefficients"""
    c = config4
    eng, pq, g = c["eng"], c["pq"], c["g"]
    eng.set_orbitals(0, pq)
    eng.forward_adjoint()
    grad = eng.grad()
    rng = np.random.default_rng(20261203 + 44)
    comp = rng.integers(0, 2, 200)
    orb = rng.integers(0, c["norb"], 200)
    idx = np.where(comp == 0, rng.integers(0, g.nsp, 200), rng.integers(0, g.nsq, 200))
    h = 1e-6
    fd = np.zeros(200)
    for n in range(200):
        e_pm = []
        for sgn in (1.0, -1.0):
            p2 = pq.copy(order="F")
            p2[idx[n], comp[n], orb[n]] += sgn * h
            eng.set_orbitals(0, p2)
            eng.forward()
            e_pm.append(eng.energy())
        fd[n] = (e_pm[0] - e_pm[1]) / (2 * h)
    eng.set_orbitals(0, pq)
    an = grad[idx, comp, orb]
    gs = np.max(np.abs(grad))
    worst = np.max(np.abs(fd - an)) / gs
    print(f"  [config4 FD, 200 coefficients, h=1e-6] max |fd - grad| / max|grad| = {worst:.2e}", flush=True)
    assert worst < 1e-6


def test_config4_adjoint_is_hf_matrix(config4):
    """At config 4's shape (30 orbitals, ns=200, k=0..8): with E the sum of
    a(i,j,k) F^k(i,j) (j <= i) and b(i,j,k) G^k(i,j) (j > i) under the
    selection rules of DBSR_HF/hf_MOD_energy.f90:20-57, the engine's
    dE/d(P_i,Q_i) equals 2 qsum(i) hfm_i (P_i;Q_i) with hfm_i hf_matrix's
    two-electron potential (dbs_hf_matrix_dev), for every orbital."""
    from dbsr3_amd import Engine
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    from test_oracle_hf_matrix import coefs_ab
    c = config4
    d, pq, n, kmax = c["d"], c["pq"], c["norb"], 8
    rng = np.random.default_rng(20261203 + 45)
    kappa = rng.choice([-1, 1, -2, 2, -3, 3, -4], n)
    lbs = np.array([k if k > 0 else -k - 1 for k in kappa], dtype=np.int32)
    jbs = np.array([2 * abs(k) - 1 for k in kappa], dtype=np.int32)
    coef = np.asfortranarray(rng.uniform(-1, 1, (n, n, kmax + 1)))
    qsum = rng.uniform(0.5, 2.0, n)
    rows, cs = [], []
    AB = [coefs_ab(i, coef, lbs, jbs, kmax) for i in range(1, n + 1)]
    for i in range(1, n + 1):
        A, B = AB[i - 1]
        for k in range(kmax + 1):
            for j in range(1, i + 1):
                if A[j - 1, k]:
                    rows.append((k, i, j, i, j)); cs.append(A[j - 1, k])
            for j in range(i + 1, n + 1):
                if B[j - 1, k]:
                    rows.append((k, i, j, j, i)); cs.append(B[j - 1, k])
    print(f"  [hf_matrix reading: {len(rows)} F/G rows]", flush=True)
    eng = Engine([(d, n, np.array(rows, dtype=np.int32))])
    eng.set_orbitals(0, pq)
    eng.set_coef(0, np.array(cs))
    eng.forward_adjoint()
    grad = eng.grad()
    eng.close()
    pqd = fortran_to_device(pq)
    scale = np.max(np.abs(grad))
    worst = 0.0
    for i in range(1, n + 1):
        A, B = AB[i - 1]
        h = device_to_fortran(d.hf_matrix(i, pqd, A, B, kmax, qsum[i - 1]))
        v = np.concatenate([pq[:, 0, i - 1], pq[:, 1, i - 1]])
        pred = 2 * qsum[i - 1] * (h @ v)
        eg = np.concatenate([grad[:, 0, i - 1], grad[:, 1, i - 1]])
        worst = max(worst, np.max(np.abs(pred - eg)) / scale)
    print(f"  [config4 adjoint vs hf_matrix, 30 orbitals] max err / max|grad| = {worst:.2e}", flush=True)
    assert worst <= 1e-11


def test_config2_channel_blocks(gpu_available):
    import torch
    from bench import CONFIG2, config2_lists, config2_sample
    from dbsr3_amd import DBS, Engine, MatBlocks
    from dbsr3_amd.dbs import fortran_to_device
    P = CONFIG2
    ns, ks, ksp, ksq, nch, nb, kmax, ibo = (P[x] for x in ("ns", "ks", "ksp", "ksq", "nch", "nbound", "kmax",
                                                            "ibo"))
    t, pq0, lists = config2_lists()
    g = O.Grid(t, ns, ks, ksp, ksq)
    pq = _normalised(g, pq0)
    nent = sum(len(l[2][0]) for l in lists)
    print(f"[config2 {len(lists)} lists, {nent} merged entries]", flush=True)
    d = DBS(t, ns, ks, ksp, ksq)
    qpq = np.zeros((nb, nb, kmax + 1), order="F")
    for i in range(nb):
        for j in range(nb):
            for m in range(kmax + 1):
                qpq[i, j, m] = g.quadr(pq[:, :, i], pq[:, :, j], m)
    m = MatBlocks(d, nch, 0, mk=kmax)
    pq_dev = fortran_to_device(pq)
    for it, kk, lst in lists:
        for atype in range(4):
            m.R_data(it, kk, atype, lst, pq_dev, nb, ibo=ibo, quadr_pq=qpq)
    torch.cuda.synchronize()
    # asymptotic coefficients of every block (itype 3, atype 0 and 1 each add CA)
    acf_ref = np.zeros((m.nblocks, kmax + 1), order="F")
    for it, kk, lst in lists:
        if it == 3:
            for _ in (0, 1):
                assert O.rdata_acf(lst, ibo, kk, kmax, qpq[:, :, kk], m.icc, acf_ref) == 0
    assert np.array_equal(m.acf.cpu().numpy().T, acf_ref), "acf"
    # seeded blocks against the oracle
    rng = np.random.default_rng(20261203 + 22)
    blocks = sorted(set(rng.choice(m.nblocks, 6, replace=False) + 1) | {int(m.icc[3, 3]), int(m.icc[39, 0])})
    icc_s = np.zeros((nch, nch), dtype=np.int32, order="F")
    for i, b in enumerate(blocks):
        icc_s[m.icc == b] = i + 1
    href = np.zeros((2 * ns, 2 * ns, len(blocks)), order="F")
    for it, kk, lst in config2_sample(lists, blocks, m.icc, ibo):
        for atype in range(4):
            O.rdata_34(g, it, atype, kk, lst[:4], pq, ibo, icc_s, href)
    for i, b in enumerate(blocks):
        got = m.hch[b - 1].cpu().numpy().T
        s = np.max(np.abs(href[:, :, i]))
        assert s > 0
        assert np.max(np.abs(got - href[:, :, i])) <= TOL * s, b
    # every block: U^T H_b W (U on both sides of a diagonal block) = sum over
    # the block's entries of c * R^k with U, W in the channel slots: direct
    # Rk(U, K2; W, K3), exchange Rk(U, K3; K2, W), the four P/Q types summed
    rng = np.random.default_rng(20261203 + 23)
    UW = np.zeros((ns, 2, 2), order="F")
    for o in range(2):
        UW[1:g.nsp - 1, 0, o] = rng.standard_normal(g.nsp - 2)
        UW[1:g.nsq - 1, 1, o] = rng.standard_normal(g.nsq - 2)
    pqx = np.asfortranarray(np.concatenate([pq, UW], axis=2))
    U, W = nb + 1, nb + 2
    rows, cc, blk = [], [], []
    for it, kk, lst in lists:
        c, K1, K2, K3 = (np.asarray(x) for x in lst[:4])
        ich, jch = K1 // ibo, K1 % ibo
        second = np.where(ich == jch, U, W)
        if it == 3:
            r = np.stack([np.full(len(c), kk), np.full(len(c), U), K2, second, K3], axis=1)
        else:
            r = np.stack([np.full(len(c), kk), np.full(len(c), U), K3, K2, second], axis=1)
        rows.append(r)
        cc.append(c)
        blk.append(m.icc[ich - 1, jch - 1])
    rows = np.concatenate(rows).astype(np.int32)
    cc, blk = np.concatenate(cc), np.concatenate(blk)
    eng = Engine([(d, nb + 2, rows)])
    eng.set_orbitals(0, pqx)
    eng.set_coef(0, cc)
    eng.forward()
    vals = eng.values()
    eng.close()
    csb = cs_bound(d, nb + 2, pqx, rows)
    want = np.bincount(blk - 1, weights=cc * vals, minlength=m.nblocks)
    scale = np.bincount(blk - 1, weights=np.abs(cc) * csb, minlength=m.nblocks)
    Ud = torch.tensor(np.concatenate([UW[:, 0, 0], UW[:, 1, 0]]), device="cuda")
    Wd = torch.tensor(np.concatenate([UW[:, 0, 1], UW[:, 1, 1]]), device="cuda")
    # tensor m.hch[b] holds hch(:,:,b) transposed: U^T H W = W . (hch_t U)
    HU = torch.matmul(m.hch, Ud).cpu().numpy()
    diag = np.zeros(m.nblocks, dtype=bool)
    for i in range(nch):
        diag[m.icc[i, i] - 1] = True
    got = np.where(diag, HU @ Ud.cpu().numpy(), HU @ Wd.cpu().numpy())
    err = np.abs(got - want) / np.maximum(scale, 1e-300)
    print(f"  [config2 block contractions, {m.nblocks} blocks] max |U^T H W - sum c R| / sum |c| CS = "
          f"{err.max():.2e}", flush=True)
    assert np.all(scale > 0)
    assert err.max() < TOL
