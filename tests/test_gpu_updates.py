"""GPU parity of the overlap, unfolding and matrix-update entry points
(dbs_updates.hip) with the oracle, bit for bit:

* dbs_bspline_overlaps_dev = ZINTYM fpbs/fqbs/fppqq (Def_pbsp, Def_qbsp,
  Def_pq_arrays) on the small, Fe.knot and Z=26 recipe grids;
* dbs_full_mat_sym_dev = Full_mat_sym for 's','l','n','x';
* dbs_update_hx/hw/hv_dev = UPDATE_HX/HW/HV for every channel order, and a
  block that is not stored;
* dbs_update_hb/acf_dev and dbs_overlap_ci_dev = sequential additions in
  list order;
* dbs_o_data_dev = O_data itype 1..4 on random lists with repeated targets;
* the reference's Stop conditions come back as errors with its messages.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import fe_grid, small_grid

pytestmark = pytest.mark.gpu
IBO = 2 ** 15


def _recipe_grid():
    from dbsr3_amd.grid import mkgrid_01
    t, ns = mkgrid_01(9, 26.0, hi=1e-3, he=math.exp(0.125) - 1.0, hmax=0.25, rmax=30.0)
    return t, ns, 9, 8, 9


GRIDS = {"small": small_grid, "fe": fe_grid, "recipe": _recipe_grid}


def _dbs(spec):
    from dbsr3_amd import DBS
    t, ns, ks, ksp, ksq = spec()
    return DBS(t, ns, ks, ksp, ksq), O.Grid(t, ns, ks, ksp, ksq)


@pytest.mark.parametrize("name", ["small", "fe", "recipe"])
def test_bspline_overlaps(gpu_available, name):
    from dbsr3_amd.dbs import device_to_fortran
    print(f"[overlaps {name}]", flush=True)
    d, g = _dbs(GRIDS[name])
    fp, fq, fpq = (device_to_fortran(x) for x in d.bspline_overlaps())
    rp, rq, rpq = O.overlaps(g)
    assert np.array_equal(fp, rp) and np.array_equal(fq, rq) and np.array_equal(fpq, rpq)


def test_full_mat_sym(gpu_available):
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    d, g = _dbs(fe_grid)
    ns, ks = g.ns, g.ks
    rng = np.random.default_rng(1)
    for sym in "slnx":
        nc = ns if sym == "x" else (2 * ks - 1 if sym == "n" else ks)
        x = np.asfortranarray(rng.standard_normal((ns, nc)))
        y = device_to_fortran(d.Full_mat_sym(fortran_to_device(x), sym))
        assert np.array_equal(y, O.full_mat_sym(ns, ks, x, sym)), sym


def _blocks(d, nch, npert, mk=4, idiag=0):
    from dbsr3_amd import MatBlocks
    return MatBlocks(d, nch, npert, mk=mk, idiag=idiag)


def _hch(m):
    return np.asfortranarray(m.hch.cpu().numpy().transpose(2, 1, 0))


def test_update_hx_hw_hv(gpu_available):
    from dbsr3_amd.dbs import fortran_to_device
    d, g = _dbs(small_grid)
    ms = g.ms
    rng = np.random.default_rng(2)
    m = _blocks(d, 3, 2)
    ref = np.zeros((ms, ms, m.nblocks), order="F")
    for rep in range(3):
        for ich, jch in ((1, 1), (2, 1), (1, 2), (3, 3), (3, 2), (2, 3)):
            b = m.icc[ich - 1, jch - 1] - 1
            X = np.asfortranarray(rng.standard_normal((ms, ms)))
            v, w = rng.standard_normal(ms), rng.standard_normal(ms)
            C1, C2 = rng.uniform(-2, 2, 2)
            m.UPDATE_HX(ich, jch, fortran_to_device(X), C1)
            blk = np.asfortranarray(ref[:, :, b])
            O.update_hx(blk, ich, jch, X, C1)
            m.UPDATE_HW(ich, jch, fortran_to_device(v), fortran_to_device(w), C2)
            O.update_hw(blk, ich, jch, v, w, C2)
            ref[:, :, b] = blk
    assert np.array_equal(_hch(m), ref)
    hcp = np.zeros((ms, m.ncb), order="F")
    for ich, ip in ((1, 1), (3, 2), (1, 1), (2, 1)):
        v = rng.standard_normal(ms)
        C_ = float(rng.uniform(-1, 1))
        m.UPDATE_HV(ich, ip, fortran_to_device(v), C_)
        col = np.ascontiguousarray(hcp[:, m.icb[ich - 1, ip - 1] - 1])
        O.update_hv(col, v, C_)
        hcp[:, m.icb[ich - 1, ip - 1] - 1] = col
    assert np.array_equal(m.hcp.cpu().numpy().T, hcp)
    # a block that is not stored (diagonal blocks only): the update is dropped
    m1 = _blocks(d, 2, 0, idiag=1)
    m1.UPDATE_HX(2, 1, fortran_to_device(np.ones((ms, ms))), 1.0)
    assert not m1.hch.abs().sum().item()


def test_update_hb_acf_and_ci_overlap(gpu_available):
    import torch
    from dbsr3_amd.matrix import overlap_ci
    d, g = _dbs(small_grid)
    rng = np.random.default_rng(3)
    m = _blocks(d, 4, 5, mk=6)
    n = 3000
    ic, jc = rng.integers(1, 6, n), rng.integers(1, 6, n)
    c = rng.standard_normal(n)
    m.UPDATE_HB(ic, jc, c)
    hp = np.zeros(m.nbb)
    for e in range(n):
        hp[m.ibb[ic[e] - 1, jc[e] - 1] - 1] += c[e]
    assert np.array_equal(m.hp.cpu().numpy(), hp)
    k, ich, jch = rng.integers(0, 7, n), rng.integers(1, 5, n), rng.integers(1, 5, n)
    m.UPDATE_ACF(k, ich, jch, c)
    acf = np.zeros((m.nblocks, 7), order="F")
    for e in range(n):
        acf[m.icc[ich[e] - 1, jch[e] - 1] - 1, k[e]] += c[e]
    assert np.array_equal(m.acf.cpu().numpy().T, acf)
    # CI overlap SM(ldsm, nzero): jc > nzero and |cc| < eps_C are skipped
    ldsm, nzero = 40, 30
    ii, jj = rng.integers(1, ldsm + 1, n), rng.integers(1, 36, n)
    cc = rng.standard_normal(n)
    cc[::11] = 1e-13
    SM = torch.zeros((nzero, ldsm), dtype=torch.float64, device="cuda")
    overlap_ci(d, (cc, ii, jj), SM, nzero)
    ref = np.zeros((ldsm, nzero), order="F")
    for e in range(n):
        if abs(cc[e]) >= 1e-10 and jj[e] <= nzero:
            ref[ii[e] - 1, jj[e] - 1] += cc[e]
    assert np.array_equal(SM.cpu().numpy().T, ref)


def test_o_data(gpu_available):
    import torch
    from dbsr3_amd.dbs import fortran_to_device
    d, g = _dbs(small_grid)
    ms, nch, npert, nvch = g.ms, 4, 3, 6
    rng = np.random.default_rng(4)
    vch = np.asfortranarray(rng.standard_normal((ms, nvch)))
    fppqq = np.asfortranarray(O.overlaps(g)[2])
    # orbitals 1..4 are channel orbitals of channels 1..4; 5..8 bound
    ipbs = np.array([0, 1, 2, 3, 4, 0, 0, 0, 0], dtype=np.int32)
    iptar = np.array([0, 1, 2, 2, 3], dtype=np.int32)
    kb = lambda i1, i2: (3 * i1 + i2) % nvch + 1
    n = 500
    c = rng.standard_normal(n)
    c[::13] = 1e-12
    lists = {
        1: (rng.integers(1, npert + 1, n), rng.integers(1, npert + 1, n)),
        2: (rng.integers(1, 5, n) * IBO + rng.integers(5, 9, n), rng.integers(1, npert + 1, n)),
        3: (rng.integers(1, 5, n) * IBO + rng.integers(5, 9, n), rng.integers(1, 5, n) * IBO + rng.integers(5, 9, n)),
        4: (rng.integers(1, 5, n) * IBO + rng.integers(1, 5, n), np.zeros(n, dtype=np.int64)),
    }
    for itype, (k1, k2) in lists.items():
        print(f"[O_data itype {itype}]", flush=True)
        m = _blocks(d, nch, npert)
        mats = dict(icc=m.icc, icb=m.icb, ibb=m.ibb, hch=np.zeros((ms, ms, m.nblocks), order="F"),
                    hcp=np.zeros((ms, m.ncb), order="F"), hp=np.zeros(m.nbb))
        ot_gpu, ot_ref = np.zeros(6), np.zeros(6)  # iptar <= 3: ij <= 6
        m.O_data(itype, (c, k1, k2), eps_c=1e-10, ibo=IBO, ipbs=ipbs[1:], kbort=kb,
                 vch=fortran_to_device(vch), fppqq=fortran_to_device(fppqq), iitar=2, iptar=iptar[1:],
                 otarg=ot_gpu)
        assert O.o_data(ms, itype, (c, k1, k2), 1e-10, IBO, ipbs, kb, vch, fppqq, 2, iptar, ot_ref, mats) == 0
        assert np.array_equal(m.hp.cpu().numpy(), mats["hp"])
        assert np.array_equal(m.hcp.cpu().numpy().T, mats["hcp"])
        assert np.array_equal(_hch(m), mats["hch"])
        assert np.array_equal(ot_gpu, ot_ref)
    torch.cuda.synchronize()


def test_stop_conditions(gpu_available):
    from dbsr3_amd import DBSError
    from dbsr3_amd.dbs import fortran_to_device
    d, g = _dbs(small_grid)
    m = _blocks(d, 2, 2, mk=3)
    X = fortran_to_device(np.zeros((g.ms, g.ms)))
    with pytest.raises(DBSError, match="UPDATE_HX: ich > nch"):
        m.UPDATE_HX(3, 1, X, 1.0)
    with pytest.raises(DBSError, match="UPDATE_HB: index is out of range"):
        m.UPDATE_HB([1, 3], [1, 1], [1.0, 1.0])
    with pytest.raises(DBSError, match="Update_cf: k > kmk"):
        m.UPDATE_ACF([4], [1], [1], [1.0])
    with pytest.raises(DBSError, match="UPDATE_HV: bound index out is of range"):
        m.UPDATE_HV(1, 3, X, 1.0)
    with pytest.raises(DBSError, match="O: KBORT=0"):
        m.O_data(2, ([1.0], [1 * IBO + 3], [1]), ibo=IBO, ipbs=np.array([1, 2, 0]),
                 kbort=lambda a, b: 0, vch=X)
    with pytest.raises(DBSError, match="Full_mat_sym: unknown storage mode"):
        d.Full_mat_sym(X, "q")
    assert not m.hp.abs().sum().item() and not m.hch.abs().sum().item()
