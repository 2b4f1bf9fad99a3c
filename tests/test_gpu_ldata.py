"""GPU L_data (DBSR_MAT3/L_data.f90:2-145) against the oracle's restatement
(pinned in tests/test_oracle_ldata.py), for lists routed by Idef_Ltype:
structures 1 and 3 bit for bit (the same products in the same order per
element), structure 2 (Gen_Lvec, a device matrix-vector product) to 1e-13 of
the largest entry."""
import numpy as np
import pytest

import oracle as O
from conftest import random_orbitals, small_grid
from test_oracle_ldata import IPB, KAPPA, NORB, l_terms
from test_oracle_updates import IBO

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("iitar,ct", [(0, 1), (2, 1)])
def test_ldata_vs_oracle(gpu_available, iitar, ct):
    from dbsr3_amd import DBS
    from dbsr3_amd.dbs import fortran_to_device
    from dbsr3_amd.matrix import MatBlocks
    t, ns, ks, ksp, ksq = small_grid()
    d = DBS(t, ns, ks, ksp, ksq)
    g = O.Grid(t, ns, ks, ksp, ksq)
    ms, nch, npert, nvch = g.ms, 3, 3, 5
    rng = np.random.default_rng(60 + iitar)
    vch = np.asfortranarray(rng.standard_normal((ms, nvch)))
    fppqq = np.asfortranarray(O.overlaps(g)[2])
    kap_list = np.array([-1, 1], dtype=np.int32)
    nuc = (0, 4.0, 0.0, 0.0, 0.0)
    hd = np.asfortranarray(np.stack([O.dhl_pq(g, *nuc, int(k)) for k in kap_list], axis=2))
    pq = np.asfortranarray(np.stack(random_orbitals(g, NORB, 61), axis=2))
    lval = np.asfortranarray(rng.standard_normal((NORB, NORB)))
    iptar = np.array([0, 1, 2, 2], dtype=np.int32)
    kb = lambda i1, i2: (i1 + 2 * i2) % nvch + 1
    lists = {1: [], 2: [], 3: []}
    for (ii, jj, c, ic, jc, io, jo) in l_terms(rng, 900, valid_only=True):
        it, key = O.idef_ltype(ii, jj, c, ic, jc, io, jo, IPB, 1e-10, ct)
        if it:
            lists[it].append((c,) + key)
    for itype, lst in lists.items():
        arr = np.array(lst)
        cols = (arr[:, 0], *[arr[:, q].astype(np.int64) for q in range(1, 5)])
        print(f"[L_data itype {itype}: {len(lst)} entries]", flush=True)
        m = MatBlocks(d, nch, npert)
        mats = dict(icc=m.icc, icb=m.icb, ibb=m.ibb, hch=np.zeros((ms, ms, m.nblocks), order="F"),
                    hcp=np.zeros((ms, m.ncb), order="F"), hp=np.zeros(m.nbb))
        hg, og, hr, orr = np.zeros(3), np.zeros(3), np.zeros(3), np.zeros(3)
        m.L_data(itype, cols, NORB, lval, KAPPA[1:], kap_list, fortran_to_device(hd),
                 pq=fortran_to_device(pq), eps_c=1e-10, ibo=IBO, ipbs=IPB[1:], kbort=kb, vch=fortran_to_device(vch),
                 fppqq=fortran_to_device(fppqq), iitar=iitar, iptar=iptar[1:], htarg=hg, otarg=og, check_target=ct)
        O.l_data(ms, itype, cols, 1e-10, IBO, IPB, kb, vch, fppqq, iitar, iptar, hr, orr, ct, lval, KAPPA, kap_list,
                 hd, pq, mats)
        hch = np.asfortranarray(m.hch.cpu().numpy().transpose(2, 1, 0))
        hcp = m.hcp.cpu().numpy().T
        if itype == 2:
            for a, b in ((hch, mats["hch"]), (hcp, mats["hcp"])):
                assert np.max(np.abs(a - b)) <= 1e-13 * max(np.max(np.abs(b)), 1e-300)
        else:
            assert np.array_equal(hch, mats["hch"]) and np.array_equal(hcp, mats["hcp"])
        assert np.array_equal(m.hp.cpu().numpy(), mats["hp"])
        assert np.array_equal(hg, hr) and np.array_equal(og, orr)
        assert np.any(hch) or np.any(hcp) or np.any(m.hp.cpu().numpy())
