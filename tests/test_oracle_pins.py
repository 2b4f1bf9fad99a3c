"""Pins of the oracle's routing restatements by reading their output back
as the integrals and overlaps it denotes:

* CI3 Add_integral, Breit branch (DBSR_CI3/add_integral.f90:31-40): for
  energy coefficients coef(i,j,k), the S^k lists (types 5/6) it builds,
  evaluated by Sk_evaluate into HM(1,1), equal the reference's own
  Breit_energy (DBSR_HF/hf_energy_BR.f90:2-50, compiled into oracle/_ref),
  which forms the same energy from sk_ppqq alone with permuted orbital
  indices -- another route through the same physics;
* Idef_itype (DBSR_MAT3/idef_type.f90:306-437): every key, read as the
  reference's structure table states (2.0 Sk(i . . .): k1=(i2,i4), k2=i3,
  k3=ich, ...), gives back the original S^k(i1,i2;i3,i4) by the compiled
  sk_ppqq / sk_pqqp, bit for bit (same routine, same arguments);
* Idef_Otype (idef_type.f90:2-50) through O_data (DBSR_MAT3/O_data.f90):
  the matrices built from the overlap keys, contracted with channel
  functions, equal the coefficients times the overlaps <ch|bound> and
  <ch_i|ch_j> computed with quadr.
"""
import numpy as np
import pytest

import oracle as O
from conftest import random_orbitals, small_grid

IBO = 2 ** 15
need_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")


def _grid():
    t, ns, ks, ksp, ksq = small_grid()
    return t, O.Grid(t, ns, ks, ksp, ksq)


@need_ref
@pytest.mark.parametrize("seed", [1, 2])
def test_ci3_breit_branch_equals_breit_energy(seed):
    t, g = _grid()
    kappa = np.array([-1, 1, -2, 2, -3], dtype=np.int32)
    n, kmax = len(kappa), 3
    lbs = np.array([k if k > 0 else -k - 1 for k in kappa], dtype=np.int32)
    jbs = np.array([2 * abs(k) - 1 for k in kappa], dtype=np.int32)
    fs = random_orbitals(g, n, 300 + seed)
    pq = np.asfortranarray(np.stack(fs, axis=2))
    rng = np.random.default_rng(310 + seed)
    coef = np.asfortranarray(rng.uniform(-1, 1, (n, n, kmax + 1)) * (rng.uniform(0, 1, (n, n, kmax + 1)) < 0.7))
    R = O.ref()
    R.ref_setup(g.ns, g.ks, g.ksp, g.ksq, O._d(O.f64(t)))
    z = np.zeros(1, dtype=np.int32)
    R.ref_hf_setup(n, kmax, 1, O._i(kappa), O._i(lbs), O._i(jbs), O._d(np.ones(n)), O._d(coef), O._d(pq), 0,
                   O._i(z), O._i(z), O._i(z), O._i(z), O._d(np.zeros(1)))
    e_ref = R.ref_breit_energy(0.0)
    lists = []
    cd = O.CData(6, 0, kmax + 1, 500, 400, 100, 0.0, lambda it, k, l: lists.append((it, k, l)))
    kap = np.concatenate([[0], kappa]).astype(np.int32)
    for i in range(1, n + 1):
        for j in range(1, n + 1):
            quad = (i, j, i, j) if i >= j else (i, j, j, i)  # as Breit_energy forms it
            for k in range(kmax + 1):
                if coef[i - 1, j - 1, k] != 0.0:
                    assert O.add_integral_ci(cd, k, 1, 1, *quad, float(coef[i - 1, j - 1, k]), kap, IBO, 1) == 0
    for it in range(1, 7):
        for k in range(kmax + 2):
            cd.add_matrix(it, k)
    HM, DM = np.zeros((1, 1), order="F"), np.zeros(1)
    nent = 0
    for it, k, lst in lists:
        if it >= 5 and len(lst[0]):
            O.rk_evaluate(g, it, k, lst, pq, IBO, HM, 1, DM, eps_c=0.0)
            nent += len(lst[0])
    assert nent > 50
    print(f"  Breit: {nent} S^k list entries, E = {HM[0, 0]:.15e}, reference {e_ref:.15e}")
    assert abs(HM[0, 0] - e_ref) <= 1e-12 * abs(e_ref)


# S^k(a,b;c,d) read from a key of structure s (the reference's table,
# DBSR_MAT3/idef_type.f90:333-356); ch(x) is the orbital of channel x
def _s_read(s, key, ch):
    k1, k2, k3, _ = key
    a, c = divmod(k1, IBO)
    if s == 1:
        b, d = divmod(k2, IBO)
        return a, b, c, d
    if s == 2:
        return ch(k3), a, k2, c
    if s == 3:
        return a, ch(k3), c, k2
    if s == 4:
        return k2, a, ch(k3), c
    if s == 5:
        return a, k2, c, ch(k3)
    if s == 6:
        return ch(a), k2, ch(c), k3
    if s == 7:
        return k2, ch(a), k3, ch(c)
    if s == 8:
        return ch(a), k3, k2, ch(c)
    return k2, ch(c), ch(a), k3  # 9: k1 = (ipbs(i3), ipbs(i2))


@need_ref
def test_idef_itype_reads_back_the_integral():
    from test_mat3_routing import IPB, NORB, _terms
    t, g = _grid()
    fs = random_orbitals(g, NORB, 41)
    R = O.ref()
    R.ref_setup(g.ns, g.ks, g.ksp, g.ksq, O._d(O.f64(t)))
    pq = np.asfortranarray(np.stack(fs, axis=2))
    R.ref_set_orbitals(NORB, O._d(pq))
    ch = lambda x: x  # channel x's orbital is orbital x (IPB)
    rng = np.random.default_rng(42)
    seen = set()
    k = 1
    for (orbs, c, ic, jc, io, jo) in _terms(rng, 400):
        for atype in (0, 1):
            it, key = O.idef_itype(atype, *orbs, c, ic, jc, io, jo, IPB, IBO)
            if it <= 0:
                continue
            assert it // 10 == atype
            s = it % 10
            want = R.ref_sk(5 + atype, *orbs, k)
            got = R.ref_sk(5 + atype, *_s_read(s, key, ch), k)
            assert got == want, (orbs, ic, jc, io, jo, it, key)
            seen.add(s)
    assert seen == set(range(1, 10))


def test_idef_otype_reads_back_the_overlaps():
    """overlap terms C<ic|jc> (configurations), C<i|.>ic, C<i|.><.|j> and
    the target term C<i|j>: their keys, put into the matrices by O_data with
    V_ch = fppqq f_bound and iitar = 2, contract with channel functions U_i
    to C, C<U_i|f>, C<U_i|f><f'|U_j> and C<U_i|U_j> (quadr)"""
    t, g = _grid()
    ns, ms = g.ns, g.ms
    nch, nbound, npert = 3, 4, 2
    ipbs = np.array([0, 1, 2, 3] + [0] * nbound, dtype=np.int32)  # orbitals 1..3 channel orbitals
    iptar = np.array([0, 1, 2, 3], dtype=np.int32)
    fs = random_orbitals(g, nbound, 51)
    rng = np.random.default_rng(52)
    U = [np.asfortranarray(rng.standard_normal((ns, 2))) for _ in range(nch)]
    for u in U:
        u[0, :] = 0.0
    vec = lambda f: np.concatenate([f[:, 0], f[:, 1]])
    fppqq = O.overlaps(g)[2]
    vch = np.asfortranarray(np.stack([fppqq @ vec(f) for f in fs], axis=1))
    kbort = lambda i1, i2: i2 - nch  # column of bound orbital i2 (the channel's own basis is shared)
    ov = lambda a, b: g.quadr(a, b, 0)
    terms = []
    for _ in range(200):
        s = int(rng.integers(1, 5))
        c = float(rng.uniform(-1, 1))
        i1, j1 = (int(x) for x in rng.integers(1, nch + 1, 2))
        i2, j2 = (int(x) for x in rng.integers(nch + 1, nch + nbound + 1, 2))
        if s == 1:
            terms.append((c, int(rng.integers(1, npert + 1)), int(rng.integers(1, npert + 1)), 0, 0))
        elif s == 2:
            terms.append((c, int(rng.integers(1, npert + 1)), 0, i1 * IBO + i2, 0))
        elif s == 3:
            terms.append((c, 0, 0, i1 * IBO + i2, j1 * IBO + j2))
        else:
            terms.append((c, 0, 0, i1 * IBO + j1, 0))
    lists = []
    cd = O.CData(4, 0, 0, 200, 40, 10, 1e-14, lambda it, k, l: lists.append((it, k, l)))
    for (c, ic, jc, io, jo) in terms:
        it, key = O.idef_otype(c, ic, jc, io, jo)
        assert it > 0
        cd.add_coef(c, 0, *key, it)
    for it in range(1, 5):
        cd.add_matrix(it, 0)
    from test_oracle_mat3 import _mats
    m = _mats(nch, npert, ms)
    otarg = np.zeros(nch * (nch + 1) // 2)
    for it, _, lst in lists:
        assert O.o_data(ms, it, lst[:3], 1e-10, IBO, ipbs, kbort, vch, fppqq, 2, iptar, otarg, m) == 0
    hp, hcp, blk, tg = np.zeros_like(m["hp"]), np.zeros((nch, npert)), {}, np.zeros_like(otarg)
    for (c, ic, jc, io, jo) in terms:
        if jc > 0:
            hp[m["ibb"][ic - 1, jc - 1] - 1] += c
        elif ic > 0:
            i1, i2 = divmod(io, IBO)
            hcp[ipbs[i1] - 1, ic - 1] += c * ov(U[ipbs[i1] - 1], fs[i2 - nch - 1])
        elif jo > 0:
            (i1, i2), (j1, j2) = divmod(io, IBO), divmod(jo, IBO)
            a, b = ipbs[i1], ipbs[j1]
            x = c * ov(U[a - 1], fs[i2 - nch - 1]) * ov(fs[j2 - nch - 1], U[b - 1])
            blk[(max(a, b), min(a, b))] = blk.get((max(a, b), min(a, b)), 0.0) + x
        else:
            i1, j1 = divmod(io, IBO)
            a, b = ipbs[i1], ipbs[j1]
            tg[(max(a, b) - 1) * max(a, b) // 2 + min(a, b) - 1] += c
            if a != b:
                blk[(max(a, b), min(a, b))] = blk.get((max(a, b), min(a, b)), 0.0) + c * ov(U[a - 1], U[b - 1])
    assert np.max(np.abs(m["hp"] - hp)) <= 1e-14 * np.max(np.abs(hp))
    got_hcp = np.array([[vec(U[ich]) @ m["hcp"][:, m["icb"][ich, ic] - 1] for ic in range(npert)]
                        for ich in range(nch)])
    assert np.max(np.abs(got_hcp - hcp)) <= 1e-13 * np.max(np.abs(hcp))
    assert np.max(np.abs(otarg - tg)) <= 1e-14 * np.max(np.abs(tg))
    for (a, b), want in blk.items():
        got = vec(U[a - 1]) @ m["hch"][:, :, m["icc"][a - 1, b - 1] - 1] @ vec(U[b - 1])
        assert abs(got - want) <= 1e-13 * max(abs(want), 1e-300), (a, b)
