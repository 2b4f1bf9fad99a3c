"""Pins of the oracle's R_data / S_data restatement (DBSR_MAT3/R_data.f90,
S_data.f90), whose reference sources need module dbsr_mat and cannot be
compiled here.  Each check reaches the matrices by another route:

* structure 1: every entry's integral from the scalar rk() with the unused
  P/Q components zeroed (R types) or from the compiled reference's own
  sk_ppqq / sk_pqqp (DBSR_HF/hf_sk.f90, S types), times c, summed naively
  into hp / hcp (V_ch columns) / hch (rank-1) / htarg + fppqq;
* structure 2 (R): the channel-bound vectors contracted with a random
  channel function u equal sum c * R^k with u in the channel slot, by rk();
* structures 6-9 (S): the channel block contracted with random channel
  functions u, w equals sum c * S^k with u, w in the channel slots, by the
  reference's sk_ppqq / sk_pqqp;
* structures 2-5 (S, the BAV vectors, which use BAV's 'n' storage and both
  directions): the vectors contracted with a channel function equal sum c *
  S^k with it in the channel slot, by the reference's sk_ppqq / sk_pqqp.
"""
import numpy as np
import pytest

import oracle as O
from conftest import random_orbitals, small_grid

IBO = 2 ** 15
NORB = 6


def _grid():
    t, ns, ks, ksp, ksq = small_grid()
    return O.Grid(t, ns, ks, ksp, ksq)


def _mats(nch, npert, ms):
    icc = np.zeros((nch, nch), dtype=np.int32, order="F")
    i = 0
    for ich in range(nch):
        i += 1
        icc[ich, ich] = i
    for ich in range(1, nch):
        for jch in range(ich):
            i += 1
            icc[ich, jch] = icc[jch, ich] = i
    icb = np.asfortranarray(np.arange(1, nch * npert + 1, dtype=np.int32).reshape((nch, npert)))
    ibb = np.zeros((npert, npert), dtype=np.int32, order="F")
    j = 0
    for ic in range(npert):
        for jc in range(ic + 1):
            j += 1
            ibb[ic, jc] = ibb[jc, ic] = j
    return dict(icc=icc, icb=icb, ibb=ibb, hch=np.zeros((ms, ms, i), order="F"),
                hcp=np.zeros((ms, nch * npert), order="F"), hp=np.zeros(j))


def _masked(f, comp):
    g = np.zeros_like(f, order="F")
    g[:, comp] = f[:, comp]
    return g


# components (left pair, right pair) of the R types by atype
R_COMPS = {0: (0, 0), 1: (1, 1), 2: (0, 1), 3: (1, 0)}


def _rk_type(g, fs, atype, a, b, c, d, k):
    """single-type R^k: left density (a,c), right (b,d), by rk()"""
    lc, rc = R_COMPS[atype]
    return g.rk(_masked(fs[a], lc), _masked(fs[b], rc), _masked(fs[c], lc), _masked(fs[d], rc), k)


def _pq(fs):
    return np.asfortranarray(np.stack(fs, axis=2))


@pytest.mark.parametrize("atype", [0, 1, 2, 3])
def test_rdata_structure1_is_sum_of_integrals(atype):
    g = _grid()
    ns, ms = g.ns, g.ms
    fs = random_orbitals(g, NORB, 31 + atype)
    rng = np.random.default_rng(40 + atype)
    nch, npert, nvch = 3, 2, 4
    vch = np.asfortranarray(rng.standard_normal((ms, nvch)))
    fppqq = O.overlaps(g)[2]
    ipbs = np.array([0, 1, 2, 3, 0, 0, 0], dtype=np.int32)  # orbitals 1..3 are channel orbitals
    iptar = np.array([0, 1, 2, 2], dtype=np.int32)
    kb = lambda i1, i2: (i1 * 3 + i2) % nvch + 1
    n, k = 60, 2
    K1 = rng.integers(1, NORB + 1, n) * IBO + rng.integers(1, NORB + 1, n)
    K2 = rng.integers(1, NORB + 1, n) * IBO + rng.integers(1, NORB + 1, n)
    kind = rng.integers(0, 4, n)
    ic, jc = rng.integers(1, npert + 1, n), rng.integers(1, npert + 1, n)
    io, jo = rng.integers(1, 4, n) * IBO + rng.integers(4, 7, n), rng.integers(1, 4, n) * IBO + rng.integers(4, 7, n)
    iot = rng.integers(1, 4, n) * IBO + rng.integers(1, 4, n)
    K3 = np.where(kind == 0, -ic, np.where(kind == 3, iot, io))
    K4 = np.where(kind == 0, -jc, np.where(kind == 1, -jc, np.where(kind == 2, jo, 0)))
    c = rng.uniform(-1, 1, n)
    order = np.lexsort((K2, K1))
    lst = tuple(x[order] for x in (c, K1, K2, K3, K4))
    m = _mats(nch, npert, ms)
    htarg = np.zeros(3)
    assert O.rs_data(g, 0, 1, atype, k, lst, _pq(fs), IBO, ipbs, kb, vch, fppqq, 1, iptar, htarg, m) == 0
    hp, hcp, hch, ht = np.zeros_like(m["hp"]), np.zeros_like(m["hcp"]), np.zeros_like(m["hch"]), np.zeros(3)
    for cc, k1, k2, k3, k4 in zip(*lst):
        S = _rk_type(g, fs, atype, k1 // IBO - 1, k2 // IBO - 1, k1 % IBO - 1, k2 % IBO - 1, k)
        C = S * cc
        ic_, jc_ = -k3, -k4
        if ic_ > 0 and jc_ > 0:
            hp[m["ibb"][ic_ - 1, jc_ - 1] - 1] += C
        elif ic_ < 0 and jc_ > 0:
            i1, i2 = divmod(int(k3), IBO)
            hcp[:, m["icb"][ipbs[i1] - 1, jc_ - 1] - 1] += C * vch[:, kb(i1, i2) - 1]
        elif ic_ < 0 and jc_ < 0:
            i1, i2 = divmod(int(k3), IBO)
            j1, j2 = divmod(int(k4), IBO)
            ich, jch = ipbs[i1], ipbs[j1]
            x = np.outer(vch[:, kb(i1, i2) - 1], vch[:, kb(j1, j2) - 1])
            xx = C * x if ich > jch else (C * x.T if jch > ich else (C / 2) * (x + x.T))
            hch[:, :, m["icc"][ich - 1, jch - 1] - 1] += xx
        else:
            i1, i2 = divmod(int(k3), IBO)
            ich, jch = ipbs[i1], ipbs[i2]
            it, jt = iptar[ich], iptar[jch]
            ht[(max(it, jt) - 1) * max(it, jt) // 2 + min(it, jt) - 1] += C
            if ich != jch:
                hch[:, :, m["icc"][ich - 1, jch - 1] - 1] += C * (fppqq if ich >= jch else fppqq.T)
    for a, b in ((m["hp"], hp), (m["hcp"], hcp), (m["hch"], hch), (htarg, ht)):
        assert np.max(np.abs(a - b)) <= 1e-13 * max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("atype", [0, 1, 2, 3])
def test_rdata_structure2_vectors_contract_to_integrals(atype):
    g = _grid()
    ns, ms = g.ns, g.ms
    fs = random_orbitals(g, NORB, 50 + atype)
    rng = np.random.default_rng(60 + atype)
    nch, npert = 2, 3
    ipbs = np.array([0, 1, 2, 0, 0, 0, 0], dtype=np.int32)
    n, k = 40, 1
    K1 = rng.integers(1, NORB + 1, n) * IBO + rng.integers(1, NORB + 1, n)
    K2 = rng.integers(1, NORB + 1, n)
    K3 = rng.integers(1, nch + 1, n)
    K4 = -rng.integers(1, npert + 1, n)  # UPDATE_HV: Rk( i . . .) ic
    c = rng.uniform(-1, 1, n)
    order = np.lexsort((K2, K1))
    lst = tuple(x[order] for x in (c, K1, K2, K3, K4))
    m = _mats(nch, npert, ms)
    vch = np.zeros((ms, 1), order="F")
    assert O.rs_data(g, 0, 2, atype, k, lst, _pq(fs), IBO, ipbs, lambda a, b: 1, vch, vch, 0, np.zeros(3, np.int32),
                     np.zeros(1), m) == 0
    lc, rc = R_COMPS[atype]
    half = slice(0, ns) if lc == 0 else slice(ns, ms)
    u = np.zeros((ns, 2), order="F")
    u[:, lc] = rng.standard_normal(ns)
    ref = np.zeros(m["hcp"].shape[1])
    for cc, k1, k2, k3, k4 in zip(*lst):
        # channel function u in slot 1: left density (u, ii), right (j1, j2)
        S = g.rk(u, _masked(fs[k1 // IBO - 1], rc), _masked(fs[k2 - 1], lc), _masked(fs[k1 % IBO - 1], rc), k)
        ref[m["icb"][k3 - 1, -k4 - 1] - 1] += cc * S
    got = u[:, lc] @ m["hcp"][half, :]
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))
    other = slice(ns, ms) if lc == 0 else slice(0, ns)
    assert not np.any(m["hcp"][other, :])


def _ref_setup(g, fs):
    R = O.ref()
    t, ns, ks, ksp, ksq = small_grid()
    R.ref_setup(ns, ks, ksp, ksq, O._d(O.f64(t)))
    pq = _pq(fs)
    R.ref_set_orbitals(pq.shape[2], O._d(pq))
    return R


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")
@pytest.mark.parametrize("atype", [0, 1])
def test_sdata_structure1_matches_reference_sk(atype):
    g = _grid()
    fs = random_orbitals(g, NORB, 70 + atype)
    R = _ref_setup(g, fs)
    rng = np.random.default_rng(80 + atype)
    npert, n, k = 3, 50, 1
    K1 = rng.integers(1, NORB + 1, n) * IBO + rng.integers(1, NORB + 1, n)
    K2 = rng.integers(1, NORB + 1, n) * IBO + rng.integers(1, NORB + 1, n)
    K3, K4 = -rng.integers(1, npert + 1, n), -rng.integers(1, npert + 1, n)
    c = rng.uniform(-1, 1, n)
    order = np.lexsort((K2, K1))
    lst = tuple(x[order] for x in (c, K1, K2, K3, K4))
    m = _mats(1, npert, g.ms)
    z = np.zeros((g.ms, 1), order="F")
    assert O.rs_data(g, 1, 1, atype, k, lst, _pq(fs), IBO, np.zeros(NORB + 1, np.int32), lambda a, b: 1, z, z, 0,
                     np.zeros(2, np.int32), np.zeros(1), m) == 0
    hp = np.zeros_like(m["hp"])
    for cc, k1, k2, k3, k4 in zip(*lst):
        S = R.ref_sk(5 + atype, int(k1 // IBO), int(k2 // IBO), int(k1 % IBO), int(k2 % IBO), k)
        hp[m["ibb"][-k3 - 1, -k4 - 1] - 1] += S * cc
    assert np.array_equal(m["hp"], hp)


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")
@pytest.mark.parametrize("atype,itype", [(0, 6), (0, 8), (1, 6), (1, 8)])
def test_sdata_channel_blocks_contract_to_reference_sk(atype, itype):
    g = _grid()
    ns, ms = g.ns, g.ms
    fs = random_orbitals(g, NORB, 90 + itype + atype)
    rng = np.random.default_rng(95 + itype + atype)
    # two channel functions u (channel 2) and w (channel 1), appended as orbitals 7, 8
    u, w = (np.asfortranarray(rng.standard_normal((ns, 2))) for _ in range(2))
    R = _ref_setup(g, fs + [u, w])
    n, k = 12, 1
    a2, a4 = rng.integers(1, NORB + 1, n), rng.integers(1, NORB + 1, n)
    c = rng.uniform(-1, 1, n)
    lst = (c, np.full(n, 2 * IBO + 1), a2, a4, np.zeros(n, dtype=np.int64))
    m = _mats(2, 1, ms)
    z = np.zeros((ms, 1), order="F")
    assert O.rs_data(g, 1, itype, atype, k, lst, _pq(fs), IBO, np.zeros(NORB + 1, np.int32), lambda a, b: 1, z, z,
                     0, np.zeros(3, np.int32), np.zeros(1), m) == 0
    blk = m["hch"][:, :, m["icc"][1, 0] - 1]  # (ich, jch) = (2, 1): += y
    # slot assignment of the structure (Idef_itype, DBSR_MAT3/idef_type.f90:401-409)
    # 6: Sk(i . j .): k2=i2, k3=i4;  8: Sk(i . . j): k2=i3, k3=i2
    ref = 0.0
    for cc, i_a, i_b in zip(c, a2, a4):
        if itype == 6:
            S = R.ref_sk(5 + atype, NORB + 1, int(i_a), NORB + 2, int(i_b), k)
        else:
            S = R.ref_sk(5 + atype, NORB + 1, int(i_b), int(i_a), NORB + 2, k)
        ref += cc * S
    # the sub-block (ip1, ip2) of the structure, contracted with the channel
    # functions' components
    tab = {(0, 6): (1, 2), (0, 8): (1, 2), (1, 6): (1, 2), (1, 8): (1, 1)}
    ip1, ip2 = tab[(atype, itype)]
    rows = slice((ip1 - 1) * ns, ip1 * ns)
    cols = slice((ip2 - 1) * ns, ip2 * ns)
    got = u[:, ip1 - 1] @ blk[rows, cols] @ w[:, ip2 - 1]
    assert abs(got - ref) <= 1e-13 * max(abs(ref), 1e-300)


def _vec(f):
    return np.concatenate([f[:, 0], f[:, 1]])


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")
@pytest.mark.parametrize("atype", [0, 1])
@pytest.mark.parametrize("itype", [6, 7, 8, 9])
def test_sdata_channel_blocks_full_contraction(atype, itype):
    """every channel-channel S structure: U^T H(2,1) W over the whole ms x ms
    block (both channel functions with P and Q parts) equals sum c * S^k with
    U, W in the slots the reference's structure table gives them
    (DBSR_MAT3/idef_type.f90:350-356): 6 Sk(i . j .): S(U,k2;W,k3); 7 Sk(. i
    . j): S(k2,U;k3,W); 8 Sk(i . . j): S(U,k3;k2,W); 9 Sk(. i j .), k1 =
    (ipbs(i3), ipbs(i2)): S(k2,W;U,k3)"""
    g = _grid()
    ns, ms = g.ns, g.ms
    fs = random_orbitals(g, NORB, 120 + 10 * itype + atype)
    rng = np.random.default_rng(130 + 10 * itype + atype)
    u, w = (np.asfortranarray(rng.standard_normal((ns, 2))) for _ in range(2))
    R = _ref_setup(g, fs + [u, w])
    U, W = NORB + 1, NORB + 2
    n, k = 14, 2
    a2, a3 = rng.integers(1, NORB + 1, n), rng.integers(1, NORB + 1, n)
    c = rng.uniform(-1, 1, n)
    lst = (c, np.full(n, 2 * IBO + 1), a2, a3, np.zeros(n, dtype=np.int64))
    m = _mats(2, 1, ms)
    z = np.zeros((ms, 1), order="F")
    assert O.rs_data(g, 1, itype, atype, k, lst, _pq(fs), IBO, np.zeros(NORB + 1, np.int32), lambda a, b: 1, z, z,
                     0, np.zeros(3, np.int32), np.zeros(1), m) == 0
    blk = m["hch"][:, :, m["icc"][1, 0] - 1]  # (ich, jch) = (2, 1): channel 2 <-> U, 1 <-> W
    slots = {6: lambda p, q: (U, p, W, q), 7: lambda p, q: (p, U, q, W),
             8: lambda p, q: (U, q, p, W), 9: lambda p, q: (p, W, U, q)}[itype]
    ref = sum(cc * R.ref_sk(5 + atype, *(int(x) for x in slots(p, q)), k) for cc, p, q in zip(c, a2, a3))
    got = _vec(u) @ blk @ _vec(w)
    assert abs(got - ref) <= 1e-13 * max(abs(ref), 1e-300), (got, ref)


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")
@pytest.mark.parametrize("atype", [0, 1])
@pytest.mark.parametrize("itype", [2, 3, 4, 5])
def test_sdata_vectors_contract_to_reference_sk(atype, itype):
    """the bound-channel S structures 2.1-5.1 (BAV vectors into hcp,
    DBSR_MAT3/S_data.f90:111-236, including BAV's 'n' storage and 'l'
    direction): U^T hcp(:, icb(ich, ic)) equals sum c * S^k with the channel
    function U in the channel slot the structure table gives
    (idef_type.f90:339-342): 2 S(U,a;k2,c), 3 S(a,U;c,k2), 4 S(k2,a;U,c),
    5 S(a,k2;c,U), with (a,c) the pair K1.  The integral uses U's component
    ip1 (the channel's), and the reference stores the vector in the ip2 half
    of the channel column (S_data.f90:213-215, kept as written), so U's ip1
    component is contracted with that half; the other half stays zero."""
    g = _grid()
    ns, ms = g.ns, g.ms
    fs = random_orbitals(g, NORB, 140 + 10 * itype + atype)
    rng = np.random.default_rng(150 + 10 * itype + atype)
    u = np.asfortranarray(rng.standard_normal((ns, 2)))
    R = _ref_setup(g, fs + [u])
    U = NORB + 1
    n, k, npert = 16, 1, 2
    K1 = rng.integers(1, NORB + 1, n) * IBO + rng.integers(1, NORB + 1, n)
    K2 = rng.integers(1, NORB + 1, n)
    K4 = -rng.integers(1, npert + 1, n)
    c = rng.uniform(-1, 1, n)
    order = np.lexsort((K2, K1))
    lst = tuple(x[order] for x in (c, K1, K2, np.ones(n, dtype=np.int64), K4))
    m = _mats(1, npert, ms)
    z = np.zeros((ms, 1), order="F")
    assert O.rs_data(g, 1, itype, atype, k, lst, _pq(fs), IBO, np.zeros(NORB + 1, np.int32), lambda a, b: 1, z, z,
                     0, np.zeros(3, np.int32), np.zeros(1), m) == 0
    slots = {2: lambda a, cc, q: (U, a, q, cc), 3: lambda a, cc, q: (a, U, cc, q),
             4: lambda a, cc, q: (q, a, U, cc), 5: lambda a, cc, q: (a, q, cc, U)}[itype]
    ref = np.zeros(npert)
    for cc, k1, k2, _, k4 in zip(*lst):
        a, c2 = divmod(int(k1), IBO)
        ref[-k4 - 1] += cc * R.ref_sk(5 + atype, *slots(a, c2, int(k2)), k)
    ip1, ip2 = {(0, 2): (1, 2), (0, 3): (1, 2), (0, 4): (2, 1), (0, 5): (2, 1),
                (1, 2): (1, 2), (1, 3): (2, 1), (1, 4): (2, 1), (1, 5): (1, 2)}[(atype, itype)]
    half = slice((ip2 - 1) * ns, ip2 * ns)
    other = slice((2 - ip2) * ns, (3 - ip2) * ns)
    got = np.array([u[:, ip1 - 1] @ m["hcp"][half, m["icb"][0, ic] - 1] for ic in range(npert)])
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)), (got, ref)
    assert not np.any(m["hcp"][other, :])
