"""CPU checks of the C-ABI library: it loads, exports every function that
include/dbsr3_amd.h declares, and its host-side integer work (c_data lists,
Add_integral, Idef_Rtype, SMU, Decode_int) matches the oracle bit for bit.
No GPU work is started here."""
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dbsr3_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dbs_[a-z0-9_]+)\s*\(", text)) - {"dbs_gen_matrix_fn"})


def test_library_exports_every_declared_symbol():
    from dbsr3_amd.capi import lib, PROTOTYPES
    L = lib()
    missing = [f for f in declared_functions() if not hasattr(L, f)]
    assert not missing, missing
    # the Python bindings cover the header
    assert set(declared_functions()) <= set(PROTOTYPES) | {"dbs_last_error", "dbs_quadr_pq_table"}


def test_quadr_pq_table_lookup():
    """the tabulated QUADR_PQ callback reads q(i1,i2,m) column-major, 1-based,
    and returns 0 outside the table (host code, no GPU)"""
    import ctypes as C
    from dbsr3_amd.capi import QpqTableC, lib
    norb, mmax = 3, 2
    q = np.asfortranarray(np.arange(norb * norb * (mmax + 1), dtype=np.float64).reshape(norb, norb, mmax + 1,
                                                                                            order="F"))
    t = QpqTableC(norb, mmax, q.ctypes.data_as(C.POINTER(C.c_double)))
    f = lib().dbs_quadr_pq_table
    for i in range(1, norb + 1):
        for j in range(1, norb + 1):
            for m in range(mmax + 1):
                assert f(C.byref(t), i, j, m) == q[i - 1, j - 1, m]
    assert f(C.byref(t), 0, 1, 0) == 0.0 and f(C.byref(t), 1, 4, 0) == 0.0 and f(C.byref(t), 1, 1, 3) == 0.0


def test_product_has_no_oracle_dependency():
    """the product sources never load, link or import the oracle"""
    pkg = os.path.join(ROOT, "dbsr3_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h", ".f90")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.replace("the oracle", ""), f


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    import dbsr3_amd.capi as capi
    monkeypatch.setattr(capi, "_lib", None)
    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(capi.DBSError):
        capi.lib()


def _run_lists(make, rows, ntype, kp2, mb, nb, kb, eps):
    log = []
    cd = make(ntype, 0, kp2, mb, nb, kb, eps, lambda it, k, l: log.append((it, k, l)))
    for (c, k, j1, j2, j3, j4, it) in rows:
        cd.add_coef(c, k, j1, j2, j3, j4, it)
    for it in range(1, ntype + 1):
        for k in range(kp2 + 1):
            cd.add_matrix(it, k)
    return [(it, k, float(c), int(a), int(b), int(cc), int(d))
            for (it, k, (cv, a1, a2, a3, a4)) in log for (c, a, b, cc, d) in zip(cv, a1, a2, a3, a4)]


@pytest.mark.parametrize("args", [
    (1, 3000, 4, 2, 2000, 1000, 100, 1e-12, 30),
    (2, 5000, 4, 2, 50, 40, 5, 1e-12, 12),
    (3, 20000, 2, 3, 64, 20, 3, 1e-3, 8),
    (4, 20000, 6, 1, 37, 25, 2, 1e-12, 6),
])
def test_cdata_matches_oracle(args):
    from dbsr3_amd import CData
    seed, n, ntype, kp2, mb, nb, kb, eps, nkey = args
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n):
        c = float(rng.uniform(-1, 1)) if rng.random() > 0.05 else 0.0
        rows.append((c, int(rng.integers(0, kp2 + 1)), *[int(x) for x in rng.integers(1, nkey, 4)],
                     int(rng.integers(1, ntype + 1))))
    a = _run_lists(CData, rows, ntype, kp2, mb, nb, kb, eps)
    b = _run_lists(O.CData, rows, ntype, kp2, mb, nb, kb, eps)
    assert a == b and len(a) > 0


def test_smu_matches_oracle():
    import ctypes as C
    from dbsr3_amd.cdata import smu
    L = O.lib()
    L.orc_smu.restype = C.c_double
    L.orc_smu.argtypes = [C.c_int] * 6 + [O._dp]
    kappas = [k for k in range(-5, 6) if k != 0]
    rng = np.random.default_rng(2)
    for _ in range(2000):
        ka, kb, kc, kd = [int(x) for x in rng.choice(kappas, 4)]
        Lk = int(rng.integers(0, 7))
        v = Lk + int(rng.integers(-1, 2))
        s2 = np.zeros(8)
        b = L.orc_smu(ka, kb, kc, kd, Lk, v, O._d(s2))
        a, s1 = smu(ka, kb, kc, kd, Lk, v)
        assert a == b and np.array_equal(s1, s2)


def test_add_integral_ci_matches_oracle():
    """parity rule, four R^k types and the eight Breit S^k terms"""
    import ctypes as C
    from dbsr3_amd import CData
    rng = np.random.default_rng(4)
    norb = 7
    kappa = np.array([0] + [int(x) for x in rng.choice([-1, 1, -2, 2, -3, 3], norb)], dtype=np.int32)
    terms = [(int(rng.integers(0, 4)), int(rng.integers(1, 5)), int(rng.integers(1, 5)),
              *[int(x) for x in rng.integers(1, norb + 1, 4)], float(rng.uniform(-1, 1)))
             for _ in range(3000)]
    L = O.lib()
    L.orc_add_integral_ci.argtypes = [C.c_void_p] + [C.c_int] * 7 + [C.c_double, O._ip, C.c_int, C.c_int]
    for mbreit in (0, 1):
        log_a, log_b = [], []
        a = CData(6, 0, 5, 64, 40, 6, 1e-12, lambda it, k, l: log_a.append((it, k, l)))
        b = O.CData(6, 0, 5, 64, 40, 6, 1e-12, lambda it, k, l: log_b.append((it, k, l)))
        for (k, ic, jc, i1, i2, i3, i4, c) in terms:
            a.add_integral_ci(k, ic, jc, i1, i2, i3, i4, c, kappa, mbreit=mbreit)
            L.orc_add_integral_ci(b.h, k, ic, jc, i1, i2, i3, i4, c, kappa.ctypes.data_as(O._ip),
                                  2 ** 15, mbreit)
        for it in range(1, 7):
            for k in range(6):
                a.add_matrix(it, k)
                b.add_matrix(it, k)
        flat = lambda log: [(it, k, float(c), *map(int, ks)) for (it, k, l) in log
                            for (c, *ks) in zip(*l)]
        assert flat(log_a) == flat(log_b) and len(flat(log_a)) > 0, mbreit


def test_idef_rtype_matches_oracle():
    from dbsr3_amd import CData
    rng = np.random.default_rng(6)
    nb_orb, nch = 8, 5
    norb = nb_orb + nch
    ipbs = np.zeros(norb + 1, dtype=np.int32)
    ipbs[nb_orb + 1:] = np.arange(1, nch + 1)
    lorb = np.array([0] + [int(x) for x in rng.integers(0, 3, norb)], dtype=np.int32)
    log = []
    cd = CData(4, 0, 6, 500, 50, 5, 1e-12, lambda it, k, l: log.append((it, k, l)))
    expect = []
    for _ in range(2000):
        k = int(rng.integers(0, 7))
        ch = [int(x) for x in rng.integers(nb_orb + 1, norb + 1, 2)]
        bd = [int(x) for x in rng.integers(1, nb_orb + 1, 2)]
        pattern = int(rng.integers(0, 4))  # channels at (1,3), (2,4), (1,4), (2,3)
        pos = [(0, 2), (1, 3), (0, 3), (1, 2)][pattern]
        j = [0] * 4
        j[pos[0]], j[pos[1]] = ch
        rest = [p for p in range(4) if p not in pos]
        j[rest[0]], j[rest[1]] = bd
        c = float(rng.uniform(-1, 1))
        cd.add_integral_mat(k, *j, c, lorb, ipbs, mk=6)
        m = (lorb[j[0]] + lorb[j[2]] + k) % 2 + (lorb[j[1]] + lorb[j[3]] + k) % 2
        if m == 0:
            it, keys = O.idef_rtype(*j, ipbs, 2 ** 15)
            expect.append((it, k, keys, c))
    for it in (3, 4):
        for k in range(7):
            cd.add_matrix(it, k)
    got = {(it, k, int(a), int(b), int(c)) for (it, k, l) in log for (cc, a, b, c, d) in zip(*l)}
    want = {(it, k, keys[0], keys[1], keys[2]) for (it, k, keys, c) in expect}
    assert got == want


def test_decode_int_matches_oracle():
    from dbsr3_amd import decode_int
    rng = np.random.default_rng(1)
    codes = rng.integers(0, 2 ** 31 - 1, 500)
    outs = decode_int(codes)
    for q, code in enumerate(codes):
        assert tuple(int(o[q]) for o in outs) == O.decode_int(int(code))
