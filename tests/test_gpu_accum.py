"""GPU parity of list accumulation against the oracle.

Rk_evaluate / Sk_evaluate (DBSR_CI3/rk_evaluate.f90, sk_evaluate.f90): lists
built by Add_integral (CI3 form, with Breit terms) and c_data; every entry's
integral times its coefficient is added into HM/DM.  R_data itype 3/4
(DBSR_MAT3/R_data.f90:214-277) + UPDATE_HS: channel-channel lists built by
the MAT3 Add_integral/Idef_Rtype, accumulated into channel blocks.
Bound: 1e-12 of the largest output magnitude (index handling is exact; the
values are integrals summed in another order).
"""
import numpy as np
import pytest

import oracle as O
from conftest import fe_grid, random_orbitals, small_grid

pytestmark = pytest.mark.gpu
TOL = 1e-12
IBI = 2 ** 15


@pytest.fixture(scope="module", params=["small", "fe"])
def setup(request, gpu_available):
    from dbsr3_amd import DBS
    t, ns, ks, ksp, ksq = small_grid() if request.param == "small" else fe_grid()
    print(f"[accum grid {request.param}]", flush=True)
    return O.Grid(t, ns, ks, ksp, ksq), DBS(t, ns, ks, ksp, ksq)


def test_rk_evaluate(setup):
    import torch
    from dbsr3_amd import CData
    from dbsr3_amd.cdata import rk_evaluate
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    g, d = setup
    norb = 6
    fs = random_orbitals(g, norb, seed=12)
    pq = np.stack(fs, axis=2)
    rng = np.random.default_rng(12)
    kappa = np.array([0] + [int(x) for x in rng.choice([-1, 1, -2, 2, -3], norb)], dtype=np.int32)
    ncj, nzero = 9, 7
    lists = []
    cd = CData(6, 0, 4, 200, 60, 6, 1e-12, lambda it, k, l: lists.append((it, k, l)))
    for _ in range(500):
        k = int(rng.integers(0, 4))
        ic, jc = sorted(int(x) for x in rng.integers(1, ncj + 1, 2))
        i = [int(x) for x in rng.integers(1, norb + 1, 4)]
        cd.add_integral_ci(k, ic, jc, *i, float(rng.uniform(-1, 1)), kappa, ibi=IBI, mbreit=1)
    for it in range(1, 7):
        for k in range(5):
            cd.add_matrix(it, k)
    assert lists
    HM0 = np.zeros((ncj, nzero), order="F")
    DM0 = np.zeros(ncj)
    HM_ref, DM_ref = HM0.copy(order="F"), DM0.copy()
    HM_dev, DM_dev = fortran_to_device(HM0), fortran_to_device(DM0)
    pq_dev = fortran_to_device(pq)
    for it, k, lst in lists:
        if len(lst[0]) == 0:
            continue
        O.rk_evaluate(g, it, k, lst, pq, IBI, HM_ref, nzero, DM_ref)
        rk_evaluate(d, it, k, lst, pq_dev, norb, HM_dev, ncj, nzero, DM_dev)
    HM, DM = device_to_fortran(HM_dev), DM_dev.cpu().numpy()
    s = max(np.max(np.abs(HM_ref)), np.max(np.abs(DM_ref)))
    assert s > 0
    assert np.max(np.abs(HM - HM_ref)) <= TOL * s
    assert np.max(np.abs(DM - DM_ref)) <= TOL * s
    # every list queued without waiting (the tables of consecutive lists in
    # the context's two pinned buffers in turn), one wait at the end: the
    # same sums in the same order
    HM_q, DM_q = fortran_to_device(HM0), fortran_to_device(DM0)
    nq = 0
    for it, k, lst in lists:
        if len(lst[0]) == 0:
            continue
        rk_evaluate(d, it, k, lst, pq_dev, norb, HM_q, ncj, nzero, DM_q, sync=False)
        nq += 1
    assert nq >= 3
    d.sync()
    assert np.array_equal(device_to_fortran(HM_q), HM) and np.array_equal(DM_q.cpu().numpy(), DM)
    torch.cuda.synchronize()


def test_r_data_channel_blocks(setup):
    g, d = setup
    _check_r_data(g, d)


def test_r_data_many_blocks(setup):
    """9 channels (45 blocks): lists of more groups than one convolution
    chunk holds (4 band groups, 8 exchange groups), with a partial last
    chunk"""
    g, d = setup
    _check_r_data(g, d, nch=9, nent=900, min_groups=17)


def test_r_data_wide_band(gpu_available):
    """ks = 10: the exchange convolution takes its untiled kernel (the LDS
    tiles hold halos for ks <= 9 only)"""
    from dbsr3_amd import DBS
    ns, ks = 24, 10
    t = O.mkgrid_02(ns, ks, 0.25, 20.0, 4.0)
    print(f"[accum grid ns={ns} ks={ks}]", flush=True)
    _check_r_data(O.Grid(t, ns, ks, ks - 1, ks), DBS(t, ns, ks, ks - 1, ks))


def _check_r_data(g, d, nch=4, nent=400, min_groups=1):
    from dbsr3_amd import CData
    from dbsr3_amd.cdata import r_data_34
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    nbound = 5
    norb = nbound + nch
    fs = random_orbitals(g, nbound, seed=33)
    pq = np.stack(fs, axis=2)
    rng = np.random.default_rng(33)
    ipbs = np.zeros(norb + 1, dtype=np.int32)
    ipbs[nbound + 1:] = np.arange(1, nch + 1)
    lorb = np.array([0] + [int(x) for x in rng.integers(0, 3, norb)], dtype=np.int32)
    # block numbering icc(ich,jch) for ich >= jch
    icc = np.zeros((nch, nch), dtype=np.int32, order="F")
    nblk = 0
    for jch in range(1, nch + 1):
        for ich in range(jch, nch + 1):
            nblk += 1
            icc[ich - 1, jch - 1] = nblk
    lists = []
    cd = CData(4, 0, 3, 300, 40, 5, 1e-12, lambda it, k, l: lists.append((it, k, l)))
    for _ in range(nent):
        k = int(rng.integers(0, 4))
        ch = [int(x) for x in rng.integers(nbound + 1, norb + 1, 2)]
        bd = [int(x) for x in rng.integers(1, nbound + 1, 2)]
        pos = [(0, 2), (1, 3), (0, 3), (1, 2)][int(rng.integers(0, 4))]
        j = [0] * 4
        j[pos[0]], j[pos[1]] = ch
        rest = [p for p in range(4) if p not in pos]
        j[rest[0]], j[rest[1]] = bd
        cd.add_integral_mat(k, *j, float(rng.uniform(-1, 1)), lorb, ipbs, mk=3)
    for it in (3, 4):
        for k in range(4):
            cd.add_matrix(it, k)
    ms = 2 * g.ns
    hch_ref = np.zeros((ms, ms, nblk), order="F")
    hch_dev = fortran_to_device(hch_ref)
    pq_dev = fortran_to_device(pq)
    n_done = 0
    ng_max = 0  # most channel blocks (groups) in one list
    for it, k, lst in lists:
        if it not in (3, 4) or len(lst[0]) == 0:
            continue
        ng_max = max(ng_max, len(np.unique(np.asarray(lst[1]))))
        for atype in range(4):
            O.rdata_34(g, it, atype, k, lst[:4], pq, IBI, icc, hch_ref)
            r_data_34(d, it, atype, k, lst, pq_dev, nbound, icc, hch_dev, nblk)
            n_done += 1
    assert n_done >= 4
    assert ng_max >= min_groups, ng_max
    hch = device_to_fortran(hch_dev)
    s = np.max(np.abs(hch_ref))
    assert s > 0
    assert np.max(np.abs(hch - hch_ref)) <= TOL * s
