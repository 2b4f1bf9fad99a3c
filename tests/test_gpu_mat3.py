"""GPU parity of MAT3's structure drivers with the oracle: R_data itype
1..4 x atype 0..3 (dbs_rdata_dev) and S_data itype 1..9 x atype 0/1
(dbs_sdata_dev) on random merged lists that hit every sub-structure
(bound-bound, bound-channel, channel-channel, target; vectors by V_ch and by
BAV), repeated and transposed channel pairs included.  Structure-1 integrals
come from the engine (FP64, other summation order), so the matrices agree
to <= 1e-12 of their largest entry; the rest is the same arithmetic.
"""
import numpy as np
import pytest

import oracle as O
from conftest import fe_grid, random_orbitals, small_grid

pytestmark = pytest.mark.gpu
IBO = 2 ** 15
NORB, NCH, NPERT, NVCH = 7, 3, 3, 5
TOL = 1e-12
# orbitals 1..3 are the channel orbitals of channels 1..3, 4..7 bound
IPBS = np.array([0, 1, 2, 3, 0, 0, 0, 0], dtype=np.int32)
IPTAR = np.array([0, 1, 2, 2], dtype=np.int32)


def kbort(i1, i2):
    return (3 * i1 + i2) % NVCH + 1


def _setup(spec, seed):
    from dbsr3_amd import DBS
    from dbsr3_amd.dbs import fortran_to_device
    t, ns, ks, ksp, ksq = spec()
    g = O.Grid(t, ns, ks, ksp, ksq)
    d = DBS(t, ns, ks, ksp, ksq)
    fs = random_orbitals(g, NORB, seed)
    pq = np.asfortranarray(np.stack(fs, axis=2))
    rng = np.random.default_rng(seed)
    vch = np.asfortranarray(rng.standard_normal((g.ms, NVCH)))
    fppqq = O.overlaps(g)[2]
    dev = dict(pq=fortran_to_device(pq), vch=fortran_to_device(vch), fppqq=fortran_to_device(fppqq))
    return g, d, pq, vch, fppqq, dev, rng


def _list(rng, stype, itype, n):
    c = rng.uniform(-1, 1, n)
    pair = lambda: rng.integers(1, NORB + 1, n) * IBO + rng.integers(1, NORB + 1, n)
    bound = lambda: rng.integers(1, 4, n) * IBO + rng.integers(4, NORB + 1, n)  # (channel orbital, bound)
    if itype == 1:
        K1, K2 = pair(), pair()
        kind = rng.integers(0, 4, n)
        ic, jc = rng.integers(1, NPERT + 1, n), rng.integers(1, NPERT + 1, n)
        io, jo = bound(), bound()
        tgt = rng.integers(1, 4, n) * IBO + rng.integers(1, 4, n)
        K3 = np.where(kind == 0, -ic, np.where(kind == 3, tgt, io))
        K4 = np.where(kind == 0, -jc, np.where(kind == 1, -jc, np.where(kind == 2, jo, 0)))
    elif (stype == 0 and itype == 2) or (stype == 1 and itype <= 5):
        K1, K2 = pair(), rng.integers(1, NORB + 1, n)
        K3 = rng.integers(1, NCH + 1, n)
        K4 = np.where(rng.integers(0, 2, n) == 0, bound(), -rng.integers(1, NPERT + 1, n))
    else:
        ich, jch = rng.integers(1, NCH + 1, n), rng.integers(1, NCH + 1, n)
        if stype == 0:  # R_data 3/4 keys have ich >= jch (Idef_Rtype); keep a few transposed
            sw = (ich < jch) & (rng.integers(0, 4, n) > 0)
            ich, jch = np.where(sw, jch, ich), np.where(sw, ich, jch)
        K1 = ich * IBO + jch
        K2, K3 = rng.integers(1, NORB + 1, n), rng.integers(1, NORB + 1, n)
        K4 = np.zeros(n, dtype=np.int64)
    order = np.lexsort((K2, K1))
    # keep a tail out of order: runs of a K1 that repeat later
    order = np.concatenate([order[: n - 5], order[n - 5:][::-1]])
    return tuple(np.asarray(x)[order] for x in (c, K1, K2, K3, K4))


def _hch(m):
    return np.asfortranarray(m.hch.cpu().numpy().transpose(2, 1, 0))


def _check(stype, itype, atype, spec, seed, k=2, n=120):
    from dbsr3_amd import MatBlocks
    g, d, pq, vch, fppqq, dev, rng = _setup(spec, seed)
    lst = _list(rng, stype, itype, n)
    m = MatBlocks(d, NCH, NPERT)
    htarg_gpu, htarg_ref = np.zeros(3), np.zeros(3)
    kw = dict(ibo=IBO, ipbs=IPBS[1:], kbort=kbort, vch=dev["vch"], fppqq=dev["fppqq"], iitar=1,
              iptar=IPTAR[1:], htarg=htarg_gpu)
    # QUADR_PQ of the host program: the oracle's quadr on the orbitals (the
    # library only sums and places what the callback returns)
    qpq = np.array([[g.quadr(pq[:, :, i], pq[:, :, j], k) for j in range(NORB)] for i in range(NORB)])
    if stype == 0:
        m.R_data(itype, k, atype, lst, dev["pq"], NORB, quadr_pq=lambda i1, i2, mm: qpq[i1 - 1, i2 - 1], **kw)
    else:
        m.S_data(atype * 10 + itype, k, lst, dev["pq"], NORB, **kw)
    mats = dict(icc=m.icc, icb=m.icb, ibb=m.ibb, hch=np.zeros((g.ms, g.ms, m.nblocks), order="F"),
                hcp=np.zeros((g.ms, m.ncb), order="F"), hp=np.zeros(m.nbb))
    assert O.rs_data(g, stype, itype, atype, k, lst, pq, IBO, IPBS, kbort, vch, fppqq, 1, IPTAR, htarg_ref,
                     mats) == 0
    got = dict(hp=m.hp.cpu().numpy(), hcp=m.hcp.cpu().numpy().T, hch=_hch(m))
    touched = 0
    for key in ("hp", "hcp", "hch"):
        ref = mats[key]
        scale = np.max(np.abs(ref))
        touched += scale > 0
        assert np.max(np.abs(got[key] - ref)) <= TOL * max(scale, 1e-300), key
    assert np.max(np.abs(htarg_gpu - htarg_ref)) <= TOL * max(np.max(np.abs(htarg_ref)), 1e-300)
    assert touched > 0
    # asymptotic coefficients: only R_data itype 3 with atype 0/1 produces them
    acf_ref = np.zeros((m.nblocks, m.mk + 1), order="F")
    if stype == 0 and itype == 3 and atype <= 1:
        assert O.rdata_acf(lst, IBO, k, m.mk, qpq, m.icc, acf_ref) == 0
        assert np.max(np.abs(acf_ref)) > 0
    acf_gpu = m.acf.cpu().numpy().T
    assert np.array_equal(acf_gpu, acf_ref), "acf"
    if stype == 0 and itype >= 3:
        # the same list queued three times without waiting (the tables in
        # the context's two pinned buffers in turn), one wait at the end:
        # three times the waited-for call's contribution, bit for bit
        m2 = MatBlocks(d, NCH, NPERT)
        ref2 = MatBlocks(d, NCH, NPERT)
        for _ in range(3):
            m2.R_data(itype, k, atype, lst, dev["pq"], NORB, quadr_pq=lambda i1, i2, mm: qpq[i1 - 1, i2 - 1],
                      sync=False, **kw)
            ref2.R_data(itype, k, atype, lst, dev["pq"], NORB, quadr_pq=lambda i1, i2, mm: qpq[i1 - 1, i2 - 1],
                        **kw)
        d.sync()
        assert np.array_equal(_hch(m2), _hch(ref2)) and np.array_equal(m2.acf.cpu().numpy(), ref2.acf.cpu().numpy())


@pytest.mark.parametrize("itype", [1, 2, 3, 4])
@pytest.mark.parametrize("atype", [0, 1, 2, 3])
def test_rdata(gpu_available, itype, atype):
    print(f"[R_data itype={itype} atype={atype}]", flush=True)
    _check(0, itype, atype, small_grid, 1000 + 10 * itype + atype)


@pytest.mark.parametrize("itype", list(range(1, 10)))
@pytest.mark.parametrize("atype", [0, 1])
def test_sdata(gpu_available, itype, atype):
    print(f"[S_data itype={itype} atype={atype}]", flush=True)
    _check(1, itype, atype, small_grid, 2000 + 10 * itype + atype)


@pytest.mark.parametrize("stype,itype", [(0, 1), (0, 2), (0, 3), (0, 4), (1, 1), (1, 3), (1, 6), (1, 9)])
def test_fe_grid(gpu_available, stype, itype):
    print(f"[Fe.knot stype={stype} itype={itype}]", flush=True)
    _check(stype, itype, 1, fe_grid, 3000 + 10 * itype + stype, k=3, n=300)


def test_bav(gpu_available):
    from dbsr3_amd.capi import check, lib
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    import ctypes as C
    import torch
    g, d, pq, vch, fppqq, dev, rng = _setup(small_grid, 4)
    ns, ks = g.ns, g.ks
    b = rng.standard_normal(ns)
    for sym in "slnx":
        nc = ns if sym == "x" else (2 * ks - 1 if sym == "n" else ks)
        a = np.asfortranarray(rng.standard_normal((ns, nc)))
        for dr in "rl":
            out = torch.zeros(ns, dtype=torch.float64, device="cuda")
            ad, bd = fortran_to_device(a), fortran_to_device(b)
            check(lib().dbs_bav_dev(d.ctx, C.c_void_p(ad.data_ptr()), C.c_void_p(bd.data_ptr()), sym.encode(),
                                    dr.encode(), C.c_void_p(out.data_ptr())))
            check(lib().dbs_ctx_sync(d.ctx))
            assert np.array_equal(out.cpu().numpy(), O.bav(ns, ks, a, b, sym, dr)), (sym, dr)
