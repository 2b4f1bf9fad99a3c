"""Pins the oracle (oracle/dbs_oracle.c) to the reference's own compiled
Fortran (oracle/_ref/libdbsref.so, built from DBSR_HF/dbsr_lib_dbs.f90,
dbsr_lib_zcom.f90, hf_energy_BR.f90 and DBSR_MAT3/MOD_c_data.f90 where they
lie).  Everything here is compared bit for bit.  Skipped where the reference
build is absent.
"""
import numpy as np
import pytest

import oracle as O
from conftest import fe_grid, small_grid

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")


def setup_both(spec):
    t, ns, ks, ksp, ksq = spec
    R = O.ref()
    tt = O.f64(t)
    R.ref_setup(ns, ks, ksp, ksq, O._d(tt))
    return R, O.Grid(t, ns, ks, ksp, ksq)


GRIDS = {"small": small_grid, "fe": fe_grid}


@pytest.mark.parametrize("name", ["small", "fe"])
def test_tables_and_moments(name):
    R, g = setup_both(GRIDS[name]())
    mine = g.tables()
    theirs = [np.zeros_like(a) for a in mine]
    R.ref_tables(*[O._d(a) for a in theirs])
    for a, b in zip(mine, theirs):
        assert np.array_equal(a, b)
    for w in range(4):
        for k in (0, 3, -1, -6):
            a = g.moments(w, k)
            b = np.zeros_like(a)
            R.ref_moments(w, k, O._d(b))
            assert np.array_equal(a, b), (w, k)


@pytest.mark.parametrize("name", ["small", "fe"])
@pytest.mark.parametrize("itype", [1, 2, 3, 4, 5, 6])
def test_arrays(name, itype):
    R, g = setup_both(GRIDS[name]())
    for k in (0, 1, 4):
        a = g.mrk(itype, k, cache=False)
        b = np.zeros_like(a)
        R.ref_mrk(itype, k, O._d(b))
        assert np.array_equal(a, b), (itype, k)
    if name == "small":
        for iv in range(g.nv):
            a = g.triang(itype, 2, iv)
            b = np.zeros_like(a)
            R.ref_triang(itype, 2, iv + 1, O._d(b))
            assert np.array_equal(a, b), (itype, iv)


@pytest.mark.parametrize("name", ["small", "fe"])
def test_contractions(name):
    R, g = setup_both(GRIDS[name]())
    ns, ks = g.ns, g.ks
    rng = np.random.default_rng(1)
    p1, p2 = rng.standard_normal(ns), rng.standard_normal(ns)
    for typ in "slnx":
        a = O.density(ns, ks, p1, p2, typ)
        b = np.zeros_like(a)
        R.ref_density(O._d(p1), O._d(p2), typ.encode(), O._d(b))
        assert np.array_equal(a, b), typ
    for itype in (1, 3, 5, 6):
        rkb = g.mrk(itype, 2)
        for icase in (1, 2, 3, 4):
            combos = [("s", "s"), ("l", "l")] + ([("n", "n"), ("l", "n"), ("n", "l")] if itype >= 5 else [])
            for si, sj in combos:
                if icase <= 2:
                    d = O.f64(rng.standard_normal((ns, 2 * ks - 1)))
                else:
                    d = O.f64(rng.standard_normal((ns, ns)))
                a = O.convol(ns, ks, rkb, d, icase, si, sj)
                b = np.zeros((ns, a.shape[1]), order="F")
                R.ref_convol(itype, 2, icase, si.encode(), sj.encode(), O._d(d), O._d(b))
                if icase <= 2:
                    nc = ks if "n" not in (si, sj) else 2 * ks - 1
                    # the reference zeroes only columns 1..ks; compare defined band entries
                    out_sym = si if icase == 1 else sj
                    for col in range(a.shape[1]):
                        f = col + 1
                        if out_sym == "s":
                            lo, hi = 1, ns - f + 1
                        elif out_sym == "l":
                            lo, hi = ks + 1 - f, ns
                        else:
                            lo, hi = max(1, 1 + ks - f), min(ns, ns + ks - f)
                        if f > (2 * ks - 1 if out_sym == "n" else ks):
                            continue
                        assert np.array_equal(a[lo - 1:hi, col], b[lo - 1:hi, col]), (itype, icase, si, sj, col)
                else:
                    assert np.array_equal(a, b), (itype, icase, si, sj)
    for sym in "slnx":
        nc = ns if sym == "x" else (2 * ks - 1 if sym == "n" else ks)
        A = O.f64(rng.standard_normal((ns, nc)))
        B = O.f64(rng.standard_normal((ns, nc)))
        assert O.sum_amb(ns, ks, A, B, sym) == R.ref_sum_amb(O._d(A), O._d(B), sym.encode())
    fs = [O.f64(rng.standard_normal((ns, 2))) for _ in range(4)]
    for k in (0, 2, 5):
        assert g.rk(*fs, k) == R.ref_rk(*[O._d(f) for f in fs], k)
        assert g.quadr(fs[0], fs[1], k) == R.ref_quadr(O._d(fs[0]), O._d(fs[1]), k)
    for sym in "snx":
        nc = ns if sym == "x" else (2 * ks - 1 if sym == "n" else ks)
        d = O.f64(rng.standard_normal((ns, nc)))
        for ii, jj in ((1, 1), (1, 2), (2, 1), (2, 2)):
            h1 = O.f64(rng.standard_normal((2 * ns, 2 * ns)))
            h2 = h1.copy(order="F")
            O.update_hs_lib(h1, ii, jj, ns, ks, d, sym)
            R.ref_update_hs(O._d(h2), ii, jj, O._d(d), sym.encode())
            assert np.array_equal(h1, h2), (sym, ii, jj)


def test_gauleg():
    R = O.ref()
    for n in range(2, 17):
        x1, w1 = O.gauleg(n)
        x2, w2 = np.zeros(n), np.zeros(n)
        R.ref_gauleg(n, O._d(x2), O._d(w2))
        assert np.array_equal(x1, x2) and np.array_equal(w1, w2), n


def _cdata_run(seed, n, ntype, kp2, mb, nb, kb, eps, nkey):
    R = O.ref()
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n):
        c = float(rng.uniform(-1, 1)) if rng.random() > 0.05 else 0.0
        rows.append((c, int(rng.integers(0, kp2 + 1)), *[int(x) for x in rng.integers(1, nkey, 4)],
                     int(rng.integers(1, ntype + 1))))
    log = []
    cd = O.CData(ntype, 0, kp2, mb, nb, kb, eps, lambda it, k, l: log.append((it, k, l)))
    for (c, k, j1, j2, j3, j4, it) in rows:
        cd.add_coef(c, k, j1, j2, j3, j4, it)
    for it in range(1, ntype + 1):
        for k in range(0, kp2 + 1):
            cd.add_matrix(it, k)
    R.ref_cdata_alloc(ntype, 0, kp2, mb, nb, kb, eps, n * 4 + 10)
    for (c, k, j1, j2, j3, j4, it) in rows:
        R.ref_cdata_add(c, k, j1, j2, j3, j4, it)
    for it in range(1, ntype + 1):
        for k in range(0, kp2 + 1):
            R.ref_cdata_flush(it, k)
    m = R.ref_cdata_nlog()
    o = [np.zeros(m, dtype=np.int32) for _ in range(7)]
    oc = np.zeros(m)
    R.ref_cdata_get(m, *[O._i(x) for x in o[:6]], O._d(oc), O._i(o[6]))
    ours = [(it, k, float(c), int(a), int(b), int(cc), int(dd))
            for (it, k, (cv, a1, a2, a3, a4)) in log for (c, a, b, cc, dd) in zip(cv, a1, a2, a3, a4)]
    theirs = [(int(a), int(b), float(c), int(d), int(e), int(f), int(g))
              for a, b, c, d, e, f, g in zip(o[0], o[1], oc, o[2], o[3], o[4], o[5])]
    return ours, theirs


@pytest.mark.parametrize("args", [
    (1, 3000, 4, 2, 2000, 1000, 100, 1e-12, 30),   # single blocks
    (2, 5000, 4, 2, 50, 40, 5, 1e-12, 12),          # block overflow, k-way merges
    (3, 20000, 2, 3, 64, 20, 3, 1e-3, 8),           # early Add_matrix + eps drop
    (4, 20000, 6, 1, 37, 25, 2, 1e-12, 6),
])
def test_cdata_bit_exact(args):
    ours, theirs = _cdata_run(*args)
    assert ours == theirs


def test_smu():
    R = O.ref()
    import ctypes as C
    R.ref_smu.restype = C.c_double
    R.ref_smu.argtypes = [C.c_int] * 6 + [O._dp]
    L = O.lib()
    L.orc_smu.restype = C.c_double
    L.orc_smu.argtypes = [C.c_int] * 6 + [O._dp]
    kappas = [k for k in range(-5, 6) if k != 0]
    rng = np.random.default_rng(2)
    for _ in range(3000):
        ka, kb, kc, kd = [int(x) for x in rng.choice(kappas, 4)]
        Lk = int(rng.integers(0, 7))
        v = Lk + int(rng.integers(-1, 2))
        s1, s2 = np.zeros(8), np.zeros(8)
        a = L.orc_smu(ka, kb, kc, kd, Lk, v, O._d(s1))
        b = R.ref_smu(ka, kb, kc, kd, Lk, v, O._d(s2))
        assert a == b and np.array_equal(s1, s2), (ka, kb, kc, kd, Lk, v)


@pytest.mark.parametrize("name", ["small", "fe"])
def test_overlaps_and_full_mat_sym(name):
    """ZINTYM overlaps (Def_pbsp/Def_qbsp/Def_pq_arrays) and Full_mat_sym"""
    R, g = setup_both(GRIDS[name]())
    fp, fq, fpq = O.overlaps(g)
    rp, rq = np.zeros_like(fp), np.zeros_like(fq)
    rpq = np.zeros_like(fpq)
    R.ref_overlaps(O._d(rp), O._d(rq), O._d(rpq))
    assert np.array_equal(fp, rp) and np.array_equal(fq, rq) and np.array_equal(fpq, rpq)
    ns, ks = g.ns, g.ks
    rng = np.random.default_rng(11)
    for sym in "slnx":
        nc = ns if sym == "x" else (2 * ks - 1 if sym == "n" else ks)
        x = O.f64(rng.standard_normal((ns, nc)))
        y = np.zeros((ns, ns), order="F")
        R.ref_full_mat_sym(ns, ks, O._d(x), O._d(y), sym.encode())
        assert np.array_equal(O.full_mat_sym(ns, ks, x, sym), y), sym


GRID_CASES = [
    (1, "point", 26.0, 0.25, 0.25, 1.0, 50.0, 1, 0.0, 0.0, 0),
    (1, "point", 37.0, 0.25, 0.25, 0.65, 50.0, 2, 0.0, 0.0, 0),
    (1, "Fermi", 26.0, 0.25, 0.25, 1.0, 50.0, 1, 9.890591209340532e-06, 0.0, 0),
    (1, "uniform", 54.0, 0.3, 0.2, 0.8, 40.0, 1, 0.0, 1.2e-4, 0),
    (2, "point", 26.0, 0.25, 0.25, 1.0, 30.0, 1, 0.0, 0.0, 150),
    (2, "Fermi", 2.0, 0.25, 0.25, 1.0, 25.0, 1, 6.520314073119532e-06, 0.0, 90),
    (2, "uniform", 10.0, 0.25, 0.25, 1.0, 30.0, 1, 0.0, 3e-5, 75),
]


@pytest.mark.parametrize("case", GRID_CASES)
def test_mkgrid_matches_reference(case):
    """mkgrid_01 / mkgrid_02 (DBSR_HF/dbsr_lib_dbs.f90:724-823) with each
    nucleus model: the oracle and the host grid module equal the
    reference's own routines bit for bit"""
    from dbsr3_amd import grid as G
    gt, nuc, z, hi, he, hmax, rmax, ml, af, ru, ns = case
    for ks in (7, 8, 9):
        tr, nr = O.ref_mkgrid(gt, ks, hi, rmax, z, nuc, af, ru, ml, he, hmax, ns)
        to, no = O.mkgrid(gt, ks, hi, rmax, z, nuc, af, ru, ml, he, hmax, ns)
        assert nr == no and np.array_equal(tr, to), (case, ks)
        if gt == 1:
            tg, ng = G.mkgrid_01(ks, z, hi, he, hmax, rmax, ml, nuc, af, ru)
        else:
            tg, ng = G.mkgrid_02(ns, ks, z, hi, rmax, nuc, af, ru), ns
        assert ng == nr and np.array_equal(tg, tr), (case, ks)


@pytest.mark.parametrize("name", ["Fe.knot", "He.knot", "Rb_cc27_knot.dat", "Be_092m1.knot"])
def test_read_grid_matches_reference(name):
    """read_grid (DBSR_HF/dbsr_lib_dbs.f90:611-631) on the reference's own
    knot files: the host reader lists the same knots bit for bit"""
    import os
    from conftest import GOLDEN
    from dbsr3_amd.grid import read_knot
    path = os.path.join(GOLDEN, "knots", name)
    p = read_knot(path)
    t = O.ref_read_grid(path, p["ns"], p["ks"])
    assert np.array_equal(p["points"], t)


def test_cpu_baselines_match_reference_rk():
    """the CPU baselines: the reference's own density/convol/SUM_AmB in
    Rk_evaluate's organisation (ref_rk_rows) and the oracle's OpenMP port of
    it (orc_rk_rows_par, any thread count) give every row's rk() bit for bit"""
    from conftest import random_orbitals
    t, ns, ks, ksp, ksq = fe_grid()
    R, g = setup_both((t, ns, ks, ksp, ksq))
    fs = random_orbitals(g, 5, 77)
    pq = np.asfortranarray(np.stack(fs, axis=2))
    rng = np.random.default_rng(78)
    rows = np.zeros((150, 5), np.int32)
    rows[:, 0] = rng.integers(0, 3, 150)
    rows[:, 1:] = rng.integers(1, 6, (150, 4))
    rs, _ = O.sort_rows_by_left_pair(rows)
    ref = O.ref_rk_rows(pq, rs)
    per_row = np.array([R.ref_rk(*[O._d(fs[x - 1]) for x in r[1:]], int(r[0])) for r in rs])
    assert np.array_equal(ref, per_row)
    for nt in (1, 3, 8):
        assert np.array_equal(O.rk_rows_par(g, pq, rs, nt), ref), nt
