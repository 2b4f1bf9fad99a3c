"""MAT3 one-electron integrals: Idef_Ltype routing (DBSR_MAT3/
idef_type.f90:54-138, called by Add_integral icase 1, add_integral.f90:24-26)
and L_data (DBSR_MAT3/L_data.f90:2-145), on the CPU.

* the library's host routing (dbs_add_integral_mat3 with icase 1, through the
  C-ABI) builds the same merged c_data lists as the oracle's Idef_Ltype, and
  stops on the same inputs;
* L_data's structure 1 is O_data's structures 1-4 with the coefficient
  scaled by Lval and the target term in htarg (iitar > 0 instead of > 1 for
  the channel-block term): checked against the oracle's O_data, itself
  pinned by dense sums (tests/test_oracle_updates.py);
* routing + L_data read back the integral: each term C L(ii,jj) with its
  configuration/overlap structure lands, by the table at the head of
  Idef_Ltype (idef_type.f90:56-76), where the dense restatement of that table
  puts it -- bound-bound C L(i,j), channel-bound C L(i,j) V_ch / C hd P_j,
  channel-channel C L(i,j) v w^T / C hd, in exact small-integer arithmetic.

Lval and Gen_Lvec are in the un-vendored Libraries_03: Lval is a table of
the caller; Gen_Lvec(i, v) is taken as v = hd(kappa of i) (P_i, Q_i).
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_updates import IBO, _ints, _mat3

NORB = 8
# orbitals 1..3: channel orbitals of channels 1..3; 4..8 bound
IPB = np.array([0, 1, 2, 3, 0, 0, 0, 0, 0], dtype=np.int32)
KAPPA = np.array([0, -1, 1, -1, -1, 1, -1, 1, -1], dtype=np.int32)
LBS = np.array([0] + [k if k > 0 else -k - 1 for k in KAPPA[1:]], dtype=np.int32)


def l_terms(rng, n, valid_only=False, ibo=IBO):
    """(ii, jj, C, ic, jc, io, jo) over every structure of Idef_Ltype, some
    of them invalid (the reference stops on those) unless valid_only; the
    overlaps io, jo packed with ibo"""
    bound = lambda: int(rng.integers(4, NORB + 1))
    chan = lambda: int(rng.integers(1, 4))
    out = []
    while len(out) < n:
        s = int(rng.integers(0, 9))
        ii, jj = bound(), bound()
        ic = jc = io = jo = 0
        if s == 0:                     # 1.1 L(..) ic,jc
            ic, jc = int(rng.integers(1, 4)), int(rng.integers(1, 4))
        elif s == 1:                   # 1.2 L(..) <i|.> ic
            ic, io = int(rng.integers(1, 4)), chan() * ibo + bound()
        elif s == 2:                   # 1.3 L(..) <i|.> <.|j>
            io, jo = chan() * ibo + bound(), chan() * ibo + bound()
        elif s == 3:                   # 1.4 L(..) <i|j>
            io = chan() * ibo + chan()
        elif s == 4:                   # 2.1 L(i.) ic
            ii = chan()
            ic = int(rng.integers(1, 4))
        elif s == 5:                   # 2.2 L(i.) <j|.>
            jj = chan()
            io = chan() * ibo + bound()
        elif s == 6:                   # 3.0 L(ij): channel orbitals of one kappa
            ii, jj = int(rng.choice([1, 3])), int(rng.choice([1, 3]))
        elif s == 7:                   # invalid: ic > 0, io = 0, two bound orbitals
            if valid_only:
                continue
            ic = 1
        else:                          # invalid: jc > 0 with ic = 0
            if valid_only:
                continue
            jc = 1
        if rng.integers(0, 2):
            ii, jj = jj, ii
        c = float(rng.integers(-6, 7)) * 0.25 if rng.integers(0, 8) else 1e-12
        out.append((ii, jj, c, ic, jc, io, jo))
    return out


def _lists(add, terms, ct):
    lists = []
    status = []
    cd = add(lambda it, k, l: lists.append((it, k, tuple(np.asarray(x).tolist() for x in l))))
    for t in terms:
        status.append(cd[0](t, ct))
    for it in range(1, 4):
        cd[1](it)
    return lists, status


def _lib(gen):
    from dbsr3_amd import CData, DBSError
    cd = CData(3, 0, 0, 64, 40, 8, 1e-14, gen)

    def add(t, ct):
        ii, jj, c, ic, jc, io, jo = t
        try:
            cd.add_integral_mat3(1, 0, ii, 0, jj, 0, c, ic, jc, io, jo, LBS[1:], KAPPA[1:], IPB[1:], mk=3,
                                 eps_c=1e-10, check_target=ct)
            return 0
        except DBSError:
            return -1
    return add, lambda it: cd.add_matrix(it, 0)


def _orc(gen):
    cd = O.CData(3, 0, 0, 64, 40, 8, 1e-14, gen)

    def add(t, ct):
        ii, jj, c, ic, jc, io, jo = t
        try:
            it, key = O.idef_ltype(ii, jj, c, ic, jc, io, jo, IPB, 1e-10, ct)
        except ValueError:
            return -1
        if it:
            cd.add_coef(c, 0, *key, it)
        return 0
    return add, lambda it: cd.add_matrix(it, 0)


@pytest.mark.parametrize("check_target", [0, 1])
def test_library_l_routing_equals_oracle(check_target):
    terms = l_terms(np.random.default_rng(30 + check_target), 400)
    a, sa = _lists(_lib, terms, check_target)
    b, sb = _lists(_orc, terms, check_target)
    assert sa == sb and -1 in sa
    assert a == b
    assert {l[0] for l in a} == {1, 2, 3} and sum(len(l[2][0]) for l in a) > 100


def _setup(rng):
    ms, nch, npert, nvch = 8, 3, 3, 5
    vch = _ints(rng, ms, nvch)
    fppqq = _ints(rng, ms, ms)
    hd = np.stack([_ints(rng, ms, ms), _ints(rng, ms, ms)], axis=2)  # kappa -1, 1
    kap_list = np.array([-1, 1], dtype=np.int32)
    pq = np.asfortranarray(_ints(rng, ms // 2, 2 * NORB).reshape((ms // 2, 2, NORB), order="F"))
    lval = _ints(rng, NORB, NORB)
    iptar = np.array([0, 1, 2, 2], dtype=np.int32)
    kb = lambda i1, i2: (i1 + 2 * i2) % nvch + 1
    return ms, nch, npert, vch, fppqq, hd, kap_list, pq, lval, iptar, kb


def _run_ldata(ms, mats, itype, lst, s, iitar, ct, htarg, otarg):
    _, nch, npert, vch, fppqq, hd, kap_list, pq, lval, iptar, kb = s
    O.l_data(ms, itype, lst, 1e-10, IBO, IPB, kb, vch, fppqq, iitar, iptar, htarg, otarg, ct, lval, KAPPA,
             kap_list, hd, pq, mats)


def test_ldata_structure1_is_o_data_scaled():
    rng = np.random.default_rng(40)
    s = _setup(rng)
    ms, nch, npert, vch, fppqq, hd, kap_list, pq, lval, iptar, kb = s
    n = 60
    c = rng.integers(-6, 7, n) * 0.25
    c[::9] = 1e-12
    i1, i2 = rng.integers(4, NORB + 1, n), rng.integers(4, NORB + 1, n)
    cs = np.array([c[e] * lval[i1[e] - 1, i2[e] - 1] if abs(c[e]) >= 1e-10 else c[e] for e in range(n)])
    io = rng.integers(1, 4, n) * IBO + rng.integers(4, NORB + 1, n)
    jo = rng.integers(1, 4, n) * IBO + rng.integers(4, NORB + 1, n)
    icj = rng.integers(1, npert + 1, n), rng.integers(1, npert + 1, n)
    io4 = rng.integers(1, 4, n) * IBO + rng.integers(1, 4, n)
    z = np.zeros(n, dtype=np.int64)
    cases = [((icj[0], icj[1]), 1, (icj[0], icj[1])), ((-io, icj[1]), 2, (io, icj[1])), ((-io, -jo), 3, (io, jo)),
             ((-io4, z), 4, (io4, io4))]
    for (k3, k4), ot, (o1, o2) in cases:
        ma, mb = _mat3(nch, npert, ms), _mat3(nch, npert, ms)
        ha, ob = np.zeros(3), np.zeros(3)
        _run_ldata(ms, ma, 1, (c, i1, i2, k3, k4), s, 2, 1, ha, None)
        assert O.o_data(ms, ot, (cs, o1, o2), 1e-10, IBO, IPB, kb, vch, fppqq, 2, iptar, ob, mb) == 0
        for key in ("hp", "hcp", "hch"):
            assert np.array_equal(ma[key], mb[key]), (ot, key)
        assert np.array_equal(ha, ob)
        assert np.any(ma["hp"]) or np.any(ma["hcp"]) or np.any(ma["hch"]) or np.any(ha)


@pytest.mark.parametrize("iitar,ct", [(0, 1), (2, 1), (1, 0)])
def test_routing_and_ldata_read_back_the_integral(iitar, ct):
    rng = np.random.default_rng(50 + iitar)
    s = _setup(rng)
    ms, nch, npert, vch, fppqq, hd, kap_list, pq, lval, iptar, kb = s
    v = lambda o: np.concatenate([pq[:, 0, o - 1], pq[:, 1, o - 1]])
    hdk = lambda o: hd[:, :, list(kap_list).index(KAPPA[o])]
    for t in l_terms(rng, 300, valid_only=True):
        ii, jj, c, ic, jc, io, jo = t
        it, key = O.idef_ltype(ii, jj, c, ic, jc, io, jo, IPB, 1e-10, ct)
        got = _mat3(nch, npert, ms)
        gh, go = np.zeros(3), np.zeros(3)
        if it:
            _run_ldata(ms, got, it, tuple(np.array([x]) for x in (c,) + key), s, iitar, ct, gh, go)
        # the dense meaning of the term
        ref = _mat3(nch, npert, ms)
        rh, ro = np.zeros(3), np.zeros(3)
        i, j = min(ii, jj), max(ii, jj)
        L = c * lval[i - 1, j - 1]
        blk = lambda a, b: ref["icc"][a - 1, b - 1] - 1
        tri = lambda a, b: (max(iptar[a], iptar[b]) - 1) * max(iptar[a], iptar[b]) // 2 + min(iptar[a], iptar[b]) - 1

        def hw(a, b, x, cc):
            if ref["icc"][a - 1, b - 1]:
                ref["hch"][:, :, blk(a, b)] += cc * x if a > b else (cc * x.T if b > a else (cc / 2) * (x + x.T))

        def hx(a, b, d, cc):
            if ref["icc"][a - 1, b - 1]:
                ref["hch"][:, :, blk(a, b)] += cc * (d if a >= b else d.T)
        chan_io = lambda q: (IPB[q // IBO], vch[:, kb(q // IBO, q % IBO) - 1])
        if abs(c) < 1e-10:
            pass
        elif jc > 0:                                   # bound-bound
            ref["hp"][ref["ibb"][ic - 1, jc - 1] - 1] += L
        elif ic > 0 and io > 0:                        # bound-channel via <i|.>
            ch, w = chan_io(io)
            ref["hcp"][:, ref["icb"][ch - 1, ic - 1] - 1] += L * w
        elif ic > 0:                                   # L(i.) ic: C hd P_bound
            ch, b = (IPB[i], j) if IPB[i] else (IPB[j], i)
            ref["hcp"][:, ref["icb"][ch - 1, ic - 1] - 1] += c * (hdk(b) @ v(b))
        elif io > 0 and jo > 0:                        # channel-channel via overlaps
            (a, x1), (b, x2) = chan_io(io), chan_io(jo)
            hw(a, b, np.outer(x1, x2), L)
        elif io > 0 and (IPB[i] or IPB[j]):            # L(i.) <j|.>
            ch, b = (IPB[i], j) if IPB[i] else (IPB[j], i)
            a, w = chan_io(io)
            hw(ch, a, np.outer(hdk(b) @ v(b), w), c)
        elif io > 0:                                   # L(..) <i|j>: target term
            if ct:
                a, b = IPB[io // IBO], IPB[io % IBO]
                if ref["icc"][a - 1, b - 1]:
                    rh[tri(a, b)] += L
                if iitar > 0 and a != b:
                    hx(a, b, fppqq, L)
        else:                                          # L(ij): channel-channel
            a, b = IPB[j], IPB[i]
            if a == b or iitar > 1:
                hx(a, b, hdk(j), c)
            elif ct and ref["icc"][a - 1, b - 1]:
                ro[tri(a, b)] += c
        for k_ in ("hp", "hcp", "hch"):
            assert np.array_equal(got[k_], ref[k_]), (t, k_)
        assert np.array_equal(gh, rh) and np.array_equal(go, ro), t
