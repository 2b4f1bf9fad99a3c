"""GPU parity of the plain path against the oracle (oracle/dbs_oracle.c, itself
pinned bit-exact to the compiled reference by tests/test_oracle_ref.py).

Tables, moment tiles, diagonal cells and the materialised arrays keep the
reference's operation order per element and must match bit for bit.
Contractions (convol, SUM_AmB, rk, quadr) reorder sums across threads and
are checked at 1e-12 relative to the largest output magnitude.
"""
import numpy as np
import pytest

import oracle as O
from conftest import fe_grid, random_orbitals, small_grid

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    s = np.max(np.abs(b))
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / (s if s > 0 else 1.0)


@pytest.fixture(scope="module", params=["small", "fe"])
def grids(request, gpu_available):
    from dbsr3_amd import DBS
    t, ns, ks, ksp, ksq = small_grid() if request.param == "small" else fe_grid()
    print(f"[grid {request.param}: ns={ns} ks={ks}]", flush=True)
    return O.Grid(t, ns, ks, ksp, ksq), DBS(t, ns, ks, ksp, ksq)


def test_tables_bit_exact(grids):
    g, d = grids
    for a, b in zip(d.tables(), g.tables()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_moments_bit_exact(grids, which):
    g, d = grids
    for k in (0, 2, -1, -7):
        assert np.array_equal(d.moments(which, k), g.moments(which, k))


@pytest.mark.parametrize("itype", [1, 2, 3, 4, 5, 6])
def test_diag_cells(grids, itype):
    g, d = grids
    for k in (0, 1, 4):
        rkd = d.diag_cells(itype, k)
        kk = g.ks * g.ks
        ref = np.stack([g.triang(itype, k, iv) for iv in range(g.nv)], axis=2)
        # only the used pair block is defined in the reference
        assert rel(rkd, ref) == 0.0, (itype, k)


@pytest.mark.parametrize("itype", [1, 2, 3, 4, 5, 6])
def test_arrays_bit_exact(grids, itype):
    g, d = grids
    for k in (0, 3):
        {1: d.mrk_pppp, 2: d.mrk_qqqq, 3: d.mrk_pqpq, 4: d.mrk_qpqp,
         5: d.msk_ppqq, 6: d.msk_pqqp}[itype](k)
        a = d.rkb()
        b = g.mrk(itype, k)
        assert np.array_equal(a, b), (itype, k, rel(a, b))


def test_density_all_types(grids):
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    g, d = grids
    rng = np.random.default_rng(3)
    p1, p2 = rng.standard_normal(g.ns), rng.standard_normal(g.ns)
    for typ in "slnx":
        a = device_to_fortran(d.density(fortran_to_device(p1), fortran_to_device(p2), typ))
        assert np.array_equal(a, O.density(g.ns, g.ks, p1, p2, typ)), typ


@pytest.mark.parametrize("itype", [1, 3, 5, 6])
@pytest.mark.parametrize("icase", [1, 2, 3, 4])
def test_convol_all_cases(grids, itype, icase):
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    g, d = grids
    ns, ks = g.ns, g.ks
    rng = np.random.default_rng(10 * itype + icase)
    combos = [("s", "s"), ("l", "l")] + ([("n", "n"), ("l", "n"), ("n", "l")] if itype >= 5 else [])
    k = 2
    sel = {1: d.mrk_pppp, 3: d.mrk_pqpq, 5: d.msk_ppqq, 6: d.msk_pqqp}[itype]
    sel(k)
    rkb = g.mrk(itype, k)
    for si, sj in combos:
        if icase <= 2:
            dd = np.asfortranarray(rng.standard_normal((ns, 2 * ks - 1)))
        else:
            dd = np.asfortranarray(rng.standard_normal((ns, ns)))
        ref = O.convol(ns, ks, rkb, dd, icase, si, sj)
        got = device_to_fortran(d.convol(fortran_to_device(dd), icase, si, sj))
        ncol = min(ref.shape[1], got.shape[1])
        if icase <= 2:
            # compare the band the reference defines
            mask = np.zeros((ns, ncol), bool)
            s_out = si if icase == 1 else sj
            for c in range(ncol):
                f = c + 1
                if s_out == "s":
                    lo, hi = 1, ns - f + 1
                elif s_out == "l":
                    lo, hi = ks + 1 - f, ns
                else:
                    lo, hi = max(1, 1 + ks - f), min(ns, ns + ks - f)
                mask[lo - 1:hi, c] = True
            e = rel(got[:, :ncol][mask], ref[:, :ncol][mask])
        else:
            e = rel(got, ref)
        assert e < TOL, (itype, icase, si, sj, e)


def test_sum_amb_rk_quadr(grids):
    from dbsr3_amd.dbs import fortran_to_device
    g, d = grids
    ns, ks = g.ns, g.ks
    rng = np.random.default_rng(7)
    for sym in "slnx":
        nc = ns if sym == "x" else (2 * ks - 1 if sym == "n" else ks)
        A = np.asfortranarray(rng.standard_normal((ns, nc)))
        B = np.asfortranarray(rng.standard_normal((ns, nc)))
        x = d.SUM_AmB(fortran_to_device(A), fortran_to_device(B), sym)
        r = O.sum_amb(ns, ks, A, B, sym)
        assert abs(x - r) <= TOL * np.sum(np.abs(A * B)), sym
    fs = random_orbitals(g, 4, seed=11)
    fd = [fortran_to_device(f) for f in fs]
    for k in (0, 1, 2, 5):
        x = d.rk(*fd, k)
        r = g.rk(*fs, k)
        assert abs(x - r) <= TOL * abs(r), (k, x, r)
    for m in (0, 1, -1, 2):
        x = d.quadr(fd[0], fd[1], m)
        r = g.quadr(fs[0], fs[1], m)
        assert abs(x - r) <= TOL * max(abs(r), 1e-300) * 10, (m, x, r)


def test_update_hs(grids):
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    g, d = grids
    ns, ks = g.ns, g.ks
    rng = np.random.default_rng(5)
    for sym in "snx":
        nc = ns if sym == "x" else (2 * ks - 1 if sym == "n" else ks)
        dd = np.asfortranarray(rng.standard_normal((ns, nc)))
        for ii, jj in ((1, 1), (1, 2), (2, 1), (2, 2)):
            h0 = np.asfortranarray(rng.standard_normal((2 * ns, 2 * ns)))
            ref = h0.copy(order="F")
            O.update_hs_lib(ref, ii, jj, ns, ks, dd, sym)
            got = device_to_fortran(d.UPDATE_HS(fortran_to_device(h0), ii, jj,
                                                fortran_to_device(dd), sym))
            assert np.array_equal(got, ref), (sym, ii, jj)


def test_rk_under_small_cache_limit(grids):
    """rk() needs four arrays of one multipole; with the cache limited to
    about one array every build evicts the previous one.  Each array must be
    built just before its convolution so that no kernel reads an evicted
    (freed) array: the values must still equal the oracle's."""
    import ctypes as C
    from dbsr3_amd.capi import check, lib
    from dbsr3_amd.dbs import fortran_to_device
    g, d = grids
    one = g.ns * g.ns * g.ks * g.ks * 8
    check(lib().dbs_ctx_set_cache_limit(d.ctx, C.c_size_t(one + one // 2)))
    try:
        fs = random_orbitals(g, 4, seed=23)
        fd = [fortran_to_device(f) for f in fs]
        for k in (3, 1, 3):
            x = d.rk(*fd, k)
            r = g.rk(*fs, k)
            assert abs(x - r) <= TOL * abs(r), (k, x, r)
    finally:
        check(lib().dbs_ctx_set_cache_limit(d.ctx, C.c_size_t(8 << 30)))
