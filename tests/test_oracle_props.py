"""Pins of the oracle's list and channel-block functions, which no compiled
reference covers (their reference sources need modules that are not in the
reference tree).  Each check reaches the answer by another route:

* Decode_int (DBSR_CI3/setup_matrix.f90:206-223): the packing
  ((((I1*32+I2)*32+I3)*32+I4)*1024+k)*4+met inverted for random fields.
* Rk_evaluate (DBSR_CI3/rk_evaluate.f90:31-56): every list entry's integral
  from the scalar rk() (pinned bit-exact to the reference) with the
  orbitals' unused P/Q component zeroed, so that only the entry's type
  survives; S^k entries from per-entry density/convol/SUM_AmB; added
  naively to HM(ic,jc) (jc <= nzero) or DM(jc) (ic = jc).
* R_data itype 3/4 + MAT3 UPDATE_HS (DBSR_MAT3/R_data.f90:214-277,
  updates.f90:110-179): the channel block contracted with random channel
  vectors u, w equals the sum over the list of c times the integral with
  u, w in the channel slots, each evaluated by SUM_AmB against the pinned
  density of (u, w).
"""
import numpy as np
import pytest

import oracle as O
from conftest import random_orbitals, small_grid

IBI = 2 ** 15


def test_decode_int_inverts_the_packing():
    rng = np.random.default_rng(5)
    for _ in range(2000):
        met, k = int(rng.integers(0, 4)), int(rng.integers(0, 1024))
        i2, i3, i4 = (int(x) for x in rng.integers(0, 32, 3))
        i1 = int(rng.integers(0, 16))  # the code is a default (32-bit) Integer
        code = ((((i1 * 32 + i2) * 32 + i3) * 32 + i4) * 1024 + k) * 4 + met
        assert O.decode_int(code) == (met, k, i1, i2, i3, i4)


def masked(f, comp):
    g = np.zeros_like(f, order="F")
    g[:, comp] = f[:, comp]
    return g


# (r1 components, r2 components) of the list types, as Rk_evaluate uses them
TYPE_COMPS = {1: ((0, 0), (0, 0)), 2: ((1, 1), (1, 1)), 3: ((0, 0), (1, 1)), 4: ((1, 1), (0, 0)),
              5: ((0, 1), (0, 1)), 6: ((0, 1), (1, 0))}


def entry_integral(g, itype, k, fs, i1, i2, j1, j2):
    (a1, a2), (b1, b2) = TYPE_COMPS[itype]
    if itype <= 4:
        return g.rk(masked(fs[i1 - 1], a1), masked(fs[j1 - 1], b1),
                    masked(fs[i2 - 1], a2), masked(fs[j2 - 1], b2), k)
    ns, ks = g.ns, g.ks
    di = O.density(ns, ks, fs[i1 - 1][:, a1], fs[i2 - 1][:, a2], "n")
    dj = O.density(ns, ks, fs[j1 - 1][:, b1], fs[j2 - 1][:, b2], "n")
    conv = O.convol(ns, ks, g.mrk(itype, k), di, 2, "n", "n")
    return O.sum_amb(ns, ks, conv, dj, "n")


@pytest.mark.parametrize("itype", [1, 2, 3, 4, 5, 6])
def test_rk_evaluate_brute_force(itype):
    t, ns, ks, ksp, ksq = small_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    norb, ncj, nzero, k = 5, 6, 4, 2
    fs = random_orbitals(g, norb, seed=40 + itype)
    pq = np.stack(fs, axis=2)
    rng = np.random.default_rng(itype)
    n = 120
    i = rng.integers(1, norb + 1, (n, 4))
    K1 = i[:, 0] * IBI + i[:, 1]
    K2 = i[:, 2] * IBI + i[:, 3]
    ic = rng.integers(1, ncj + 1, n)
    jc = np.where(rng.random(n) < 0.3, ic, rng.integers(1, ncj + 1, n))
    c = rng.uniform(-1, 1, n)
    c[::17] = 1e-12  # below eps_c: skipped
    order = np.lexsort((K2, K1))  # list order: runs of equal K1, then K2
    lst = (c[order], K1[order], K2[order], ic[order], jc[order])
    HM = np.zeros((ncj, nzero), order="F")
    DM = np.zeros(ncj)
    O.rk_evaluate(g, itype, k, lst, pq, IBI, HM, nzero, DM, eps_c=1e-10)
    HMb = np.zeros((ncj, nzero))
    DMb = np.zeros(ncj)
    for cc, k1, k2, a, b in zip(*lst):
        if abs(cc) < 1e-10:
            continue
        S = entry_integral(g, itype, k, fs, k1 // IBI, k1 % IBI, k2 // IBI, k2 % IBI)
        if b <= nzero:
            HMb[a - 1, b - 1] += cc * S
        elif a == b:
            DMb[b - 1] += cc * S
    scale = max(np.max(np.abs(HMb)), np.max(np.abs(DMb)))
    assert scale > 0
    assert np.max(np.abs(HM - HMb)) <= 1e-13 * scale
    assert np.max(np.abs(DM - DMb)) <= 1e-13 * scale


@pytest.mark.parametrize("itype", [3, 4])
@pytest.mark.parametrize("atype", [0, 1, 2, 3])
def test_rdata_34_channel_block_contraction(itype, atype):
    t, ns, ks, ksp, ksq = small_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    nbound, nch, k, ibo = 4, 3, 1, 2 ** 15
    fs = random_orbitals(g, nbound, seed=60 + 10 * itype + atype)
    pq = np.stack(fs, axis=2)
    icc = np.zeros((nch, nch), dtype=np.int32, order="F")
    nblk = 0
    for jch in range(1, nch + 1):
        for ich in range(jch, nch + 1):
            nblk += 1
            icc[ich - 1, jch - 1] = nblk
    rng = np.random.default_rng(100 + 10 * itype + atype)
    # entries per channel pair (ich >= jch, as Idef_Rtype orders them), runs of equal K1
    ents = []
    for ich in range(1, nch + 1):
        for jch in range(1, ich + 1):
            for _ in range(3):
                j1, j2 = (int(x) for x in rng.integers(1, nbound + 1, 2))
                ents.append((float(rng.uniform(-1, 1)), ich * ibo + jch, j1, j2))
    c = np.array([e[0] for e in ents])
    K1 = np.array([e[1] for e in ents], dtype=np.int32)
    K2 = np.array([e[2] for e in ents], dtype=np.int32)
    K3 = np.array([e[3] for e in ents], dtype=np.int32)
    ms = 2 * ns
    hch = np.zeros((ms, ms, nblk), order="F")
    O.rdata_34(g, itype, atype, k, (c, K1, K2, K3), pq, ibo, icc, hch)
    # components: channel slots (ip1, ip2), bound orbitals (jp1, jp2)
    ip1, ip2, jp1, jp2 = {0: (1, 1, 1, 1), 1: (2, 2, 2, 2), 2: (1, 1, 2, 2), 3: (2, 2, 1, 1)}[atype]
    if itype == 4 and atype == 2:
        ip1, ip2, jp1, jp2 = 1, 2, 1, 2
    if itype == 4 and atype == 3:
        ip1, ip2, jp1, jp2 = 2, 1, 2, 1
    rkb = g.mrk({0: O.PPPP, 1: O.QQQQ, 2: O.PQPQ, 3: O.QPQP}[atype], k)
    for ich in range(1, nch + 1):
        for jch in range(1, ich + 1):
            u = rng.standard_normal(ns)
            w = rng.standard_normal(ns)
            blk = hch[:, :, icc[ich - 1, jch - 1] - 1]
            sub = blk[(ip1 - 1) * ns:ip1 * ns, (ip2 - 1) * ns:ip2 * ns]
            got = u @ sub @ w
            want = 0.0
            for cc, kk1, j1, j2 in ents:
                if kk1 != ich * ibo + jch:
                    continue
                p1 = pq[:, jp1 - 1, j1 - 1]
                p2 = pq[:, jp2 - 1, j2 - 1]
                if itype == 3:
                    y = O.convol(ns, ks, rkb, O.density(ns, ks, p1, p2, "s"), 1, "s", "s")
                    # u^T Y w with Y the unfolded symmetric band
                    val = O.sum_amb(ns, ks, y, O.density(ns, ks, u, w, "s"), "s")
                    valt = val
                else:
                    y = O.convol(ns, ks, rkb, O.density(ns, ks, p1, p2, "x"), 3, "s", "s")
                    val = O.sum_amb(ns, ks, y, O.density(ns, ks, u, w, "x"), "x")   # u^T Y w
                    valt = O.sum_amb(ns, ks, y, O.density(ns, ks, w, u, "x"), "x")  # u^T Y^T w
                if ich > jch:
                    want += cc * val
                elif ich == jch:
                    if ip1 == ip2:
                        want += cc * 0.5 * (val + valt)
                    else:
                        want += cc * 0.5 * val
            scale = abs(want) + 1e-300
            assert abs(got - want) <= 1e-12 * max(scale, 1e-6), (ich, jch, got, want)


def test_idef_rtype_keeps_the_integral():
    """Idef_Rtype (DBSR_MAT3/idef_type.f90:142-302) only renames a
    channel-channel term: with real functions on every orbital, the integral
    read back from the canonical key -- direct (itype 3): R^k(ch_i, K2;
    ch_j, K3); exchange (itype 4): R^k(ch_i, K3; K2, ch_j), the forms that
    R_data's convol case 1 of density(K2,K3,'s') / case 3 of
    density(K2,K3,'x') contract (checked below) -- equals the original
    R^k(j1, j2; j3, j4)."""
    t, ns, ks, ksp, ksq = small_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    nbound, nch, ibo = 4, 4, 2 ** 15
    norb = nbound + nch
    fs = random_orbitals(g, norb, seed=91)
    ipbs = np.zeros(norb + 1, dtype=np.int32)
    ipbs[nbound + 1:] = np.arange(1, nch + 1)
    orb_of = {int(ipbs[o]): o for o in range(1, norb + 1) if ipbs[o] > 0}
    rng = np.random.default_rng(91)
    seen = {3: 0, 4: 0}
    for _ in range(300):
        ch = [int(x) for x in rng.integers(nbound + 1, norb + 1, 2)]
        bd = [int(x) for x in rng.integers(1, nbound + 1, 2)]
        pos = [(0, 2), (1, 3), (0, 3), (1, 2)][int(rng.integers(0, 4))]
        j = [0] * 4
        j[pos[0]], j[pos[1]] = ch
        rest = [p for p in range(4) if p not in pos]
        j[rest[0]], j[rest[1]] = bd
        it, (k1, k2, k3, k4) = O.idef_rtype(*j, ipbs, ibo)
        assert it in (3, 4) and k4 == 0
        ich, jch = k1 // ibo, k1 % ibo
        assert ich >= jch
        k = int(rng.integers(0, 4))
        want = g.rk(*[fs[x - 1] for x in j], k)
        a, b = fs[orb_of[ich] - 1], fs[orb_of[jch] - 1]
        if it == 3:
            got = g.rk(a, fs[k2 - 1], b, fs[k3 - 1], k)
        else:
            got = g.rk(a, fs[k3 - 1], fs[k2 - 1], b, k)
        assert abs(got - want) <= 1e-12 * max(abs(want), 1e-3), (j, it, k1, k2, k3)
        seen[it] += 1
    assert seen[3] > 50 and seen[4] > 50


def test_exchange_block_form():
    """u^T Y w for Y = convol case 3 of density(p1,p2,'x') is R^k(u,p2;p1,w)
    and u^T Y w for Y = convol case 1 of density(p1,p2,'s') is
    R^k(u,p1;w,p2): the integrals R_data's blocks carry (P components)."""
    t, ns, ks, ksp, ksq = small_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    u, w, p1, p2 = [masked(f, 0) for f in random_orbitals(g, 4, seed=3)]
    for k in (0, 1, 3):
        rkb = g.mrk(O.PPPP, k)
        Y = O.convol(ns, ks, rkb, O.density(ns, ks, p1[:, 0], p2[:, 0], "x"), 3, "s", "s")
        x = u[:, 0] @ Y @ w[:, 0]
        r = g.rk(u, p2, p1, w, k)
        assert abs(x - r) <= 1e-13 * abs(r)
        assert abs(g.rk(u, p1, p2, w, k) - r) > 1e-6 * abs(r)  # the other order differs
        Y1 = O.convol(ns, ks, rkb, O.density(ns, ks, p1[:, 0], p2[:, 0], "s"), 1, "s", "s")
        x1 = O.sum_amb(ns, ks, Y1, O.density(ns, ks, u[:, 0], w[:, 0], "s"), "s")
        r1 = g.rk(u, p1, w, p2, k)
        assert abs(x1 - r1) <= 1e-13 * abs(r1)


def test_add_integral_ci_keeps_the_energy():
    """CI3 Add_integral without Breit terms (DBSR_CI3/add_integral.f90:2-43)
    followed by c_data merging: for every matrix element (ic,jc) the lists'
    entries, each evaluated by the masked scalar rk() of its type, add up to
    sum C * R^k(i1,i2;i3,i4) over the terms that pass the parity rule
    (l(i1)+l(i3)+k and l(i2)+l(i4)+k even)."""
    import ctypes as C
    t, ns, ks, ksp, ksq = small_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    norb, ncj = 5, 4
    fs = random_orbitals(g, norb, seed=17)
    rng = np.random.default_rng(17)
    kappa = np.array([0] + [int(x) for x in rng.choice([-1, 1, -2, 2, -3], norb)], dtype=np.int32)
    lq = {o: (int(kappa[o]) if kappa[o] > 0 else -int(kappa[o]) - 1) for o in range(1, norb + 1)}
    L = O.lib()
    L.orc_add_integral_ci.argtypes = [C.c_void_p] + [C.c_int] * 7 + [C.c_double, O._ip, C.c_int, C.c_int]
    lists = []
    cd = O.CData(6, 0, 3, 40, 30, 6, 1e-12, lambda it, k, l: lists.append((it, k, l)))
    want = np.zeros((ncj + 1, ncj + 1))
    npass = 0
    for _ in range(300):
        k = int(rng.integers(0, 4))
        ic, jc = (int(x) for x in rng.integers(1, ncj + 1, 2))
        i1, i2, i3, i4 = (int(x) for x in rng.integers(1, norb + 1, 4))
        c = float(rng.uniform(-1, 1))
        L.orc_add_integral_ci(cd.h, k, ic, jc, i1, i2, i3, i4, c, kappa.ctypes.data_as(O._ip), IBI, 0)
        if (lq[i1] + lq[i3] + k) % 2 == 0 and (lq[i2] + lq[i4] + k) % 2 == 0:
            want[ic, jc] += c * g.rk(fs[i1 - 1], fs[i2 - 1], fs[i3 - 1], fs[i4 - 1], k)
            npass += 1
    for it in range(1, 7):
        for k in range(4):
            cd.add_matrix(it, k)
    got = np.zeros_like(want)
    for it, k, (c, K1, K2, K3, K4) in lists:
        assert it <= 4
        for cc, k1, k2, a, b in zip(c, K1, K2, K3, K4):
            got[a, b] += cc * entry_integral(g, it, k, fs, k1 // IBI, k1 % IBI, k2 // IBI, k2 % IBI)
    assert npass > 30
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
