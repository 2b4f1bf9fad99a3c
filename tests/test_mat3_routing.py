"""MAT3 coefficient routing (DBSR_MAT3/add_integral.f90, idef_type.f90):
Add_integral with icase 0 (Idef_Otype), 2 (parity rule + Idef_Rtype, all
16 structures) and 3 (SMU + Idef_itype, eight permuted S terms).

* the library's host code (dbs_add_integral_mat3, through the C-ABI, no GPU
  needed) produces the same merged c_data lists as the oracle, bit for bit,
  and stops on the same inputs;
* the oracle's R routing is pinned by reading the integral back: every key
  (k1..k4, itype) of a term C*R^k(i1,i2;i3,i4) denotes, in R_data's reading
  of that structure, an integral equal to the original one (total rk() of
  the orbitals, channel orbitals taken at their channel slots).
"""
import numpy as np
import pytest

import oracle as O
from conftest import random_orbitals, small_grid

IBO = 2 ** 15
NORB = 8
# orbitals 1..3: channel orbitals of channels 1..3; 4..8 bound
IPB = np.array([0, 1, 2, 3, 0, 0, 0, 0, 0], dtype=np.int32)   # element 0 unused
KAPPA = np.array([0, -1, 1, -2, -1, 2, -2, 1, -3], dtype=np.int32)
LBS = np.array([0] + [k if k > 0 else -k - 1 for k in KAPPA[1:]], dtype=np.int32)


def _terms(rng, n):
    """random (i1..i4, C, ic, jc, io, jo) covering every structure"""
    out = []
    bound = lambda: int(rng.integers(4, NORB + 1))
    chan = lambda: int(rng.integers(1, 4))
    for _ in range(n):
        s = int(rng.integers(0, 6))
        orbs = [bound() for _ in range(4)]
        ic = jc = io = jo = 0
        if s == 0:                      # bound-bound configurations
            ic, jc = int(rng.integers(1, 5)), int(rng.integers(1, 5))
        elif s == 1:                    # one channel orbital, configuration ic
            orbs[int(rng.integers(0, 4))] = chan()
            ic = int(rng.integers(1, 5))
        elif s == 2:                    # <i|.> ic
            ic, io = int(rng.integers(1, 5)), chan() * IBO + bound()
        elif s == 3:                    # two channel orbitals: direct or exchange pair
            p = [(0, 2), (1, 3), (0, 3), (1, 2)][int(rng.integers(0, 4))]
            orbs[p[0]], orbs[p[1]] = chan(), chan()
        elif s == 4:                    # <i|.> with or without a channel orbital
            if rng.integers(0, 2):
                orbs[int(rng.integers(0, 4))] = chan()
            io = chan() * IBO + bound()
        else:                           # <i|.> <.|j>
            io, jo = chan() * IBO + bound(), chan() * IBO + bound()
        out.append((orbs, float(rng.uniform(-1, 1)), ic, jc, io, jo))
    return out


def _run(make, icase, terms, kmax=3, check_target=1):
    lists = []
    cd = make(lambda it, k, l: lists.append((it, k, tuple(np.asarray(x).tolist() for x in l))))
    status = []
    for (orbs, c, ic, jc, io, jo) in terms:
        for k in range(kmax + 1):
            status.append(cd.add(icase, k, *orbs, c, ic, jc, io, jo, check_target))
    for it in range(1, 20):
        for k in range(0, kmax + 2):
            cd.flush(it, k)
    return lists, status


class _Lib:
    def __init__(self, gen, kmax):
        from dbsr3_amd import CData
        self.cd = CData(19, 0, kmax + 1, 64, 40, 8, 1e-14, gen)

    def add(self, icase, k, i1, i2, i3, i4, c, ic, jc, io, jo, ct):
        from dbsr3_amd import DBSError
        try:
            self.cd.add_integral_mat3(icase, k, i1, i2, i3, i4, c, ic, jc, io, jo, LBS[1:], KAPPA[1:], IPB[1:],
                                      mk=3, eps_c=1e-10, check_target=ct)
            return 0
        except DBSError:
            return -1

    def flush(self, it, k):
        self.cd.add_matrix(it, k)


class _Orc:
    def __init__(self, gen, kmax):
        self.cd = O.CData(19, 0, kmax + 1, 64, 40, 8, 1e-14, gen)

    def add(self, icase, k, i1, i2, i3, i4, c, ic, jc, io, jo, ct):
        r = self.cd.add_integral_mat3(icase, k, i1, i2, i3, i4, c, ic, jc, io, jo, LBS, KAPPA, IPB, mk=3,
                                      eps_c=1e-10, check_target=ct)
        return -1 if r < 0 else 0

    def flush(self, it, k):
        self.cd.add_matrix(it, k)


@pytest.mark.parametrize("icase", [0, 2, 3])
@pytest.mark.parametrize("check_target", [0, 1])
def test_library_routing_equals_oracle(icase, check_target):
    rng = np.random.default_rng(100 + icase + 10 * check_target)
    terms = _terms(rng, 300)
    a, sa = _run(lambda g: _Lib(g, 4), icase, terms, check_target=check_target)
    b, sb = _run(lambda g: _Orc(g, 4), icase, terms, check_target=check_target)
    assert sa == sb
    assert a == b
    assert sum(len(l[2][0]) for l in a) > 50


def test_routing_reads_back_the_integral():
    """every R key denotes the original integral (total rk, the sum over
    the four P/Q types that R_data evaluates atype by atype)"""
    t, ns, ks, ksp, ksq = small_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    fs = random_orbitals(g, NORB, 7)
    f = lambda o: fs[o - 1]
    chan_orb = {1: 1, 2: 2, 3: 3}  # channel -> its orbital
    rng = np.random.default_rng(8)
    k = 2
    checked = set()
    for (orbs, c, ic, jc, io, jo) in _terms(rng, 400):
        it, key = O.idef_rtype_full(*orbs, c, ic, jc, io, jo, IPB, IBO)
        if it <= 0:
            continue
        k1, k2, k3, k4 = key
        ref = g.rk(f(orbs[0]), f(orbs[1]), f(orbs[2]), f(orbs[3]), k)
        if it == 1:      # K1 = (a,c) left pair, K2 = (b,d) right pair
            val = g.rk(f(k1 // IBO), f(k2 // IBO), f(k1 % IBO), f(k2 % IBO), k)
            assert (k3, k4) in ((-ic, -jc), (io, -ic), (io, jo), (io, 0))
        elif it == 2:    # K1 = bound pair on r2, K2 = bound partner of the channel orbital on r1
            ch = chan_orb[k3]
            val = g.rk(f(ch), f(k1 // IBO), f(k2), f(k1 % IBO), k)
            assert k4 in (-ic, io)
        elif it == 3:    # direct: channel pair (ich, jch) on r1, (K2, K3) on r2
            val = g.rk(f(chan_orb[k1 // IBO]), f(k2), f(chan_orb[k1 % IBO]), f(k3), k)
            assert k1 // IBO >= k1 % IBO
        else:            # exchange: R(ch_i, K3; K2, ch_j)
            val = g.rk(f(chan_orb[k1 // IBO]), f(k3), f(k2), f(chan_orb[k1 % IBO]), k)
            assert k1 // IBO >= k1 % IBO
        assert abs(val - ref) <= 1e-13 * max(abs(ref), 1e-300), (orbs, ic, jc, io, jo, it, key)
        checked.add(it)
    assert checked == {1, 2, 3, 4}

