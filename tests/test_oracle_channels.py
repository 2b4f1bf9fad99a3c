"""DBSR_MAT3's channel-block post-processing in the oracle, pinned on the CPU.

* Zero_cond (DBSR_MAT3/zero_cond.f90:2-56): the B-splines it removes at
  r = 0 for a channel of (l, kappa) are exactly those on which DBSR_HF's
  converged Fe orbitals of the same (l, kappa) vanish (Fe.bsw, computed with
  the same rule); at r = R, with ibzero = 1, exactly the one B-spline of the
  clamped P (Q) basis that does not vanish there, found by an independent
  Cox-de Boor evaluation.
* DBS_ORTH (dbs_orth.f90:2-84): with B-orthonormal orbitals p_j and
  q_j = B p_j, (1 - q p^T) and (1 - p q^T) are projectors, and a block is
  fixed by what it does on span(p) and on the complement q^T y = 0: rows
  (columns) of the projected channel are annihilated by p, the complement
  is unchanged, blocks outside idiag's selection are untouched.
* Diag_channels (diag.f90:2-79): the eigen-solve is LAPACK's DSYGV through
  scipy, the routine the reference calls; the chain (zero conditions,
  compaction, window, expansion, sign, eigenvalues in column ns) is pinned
  by the Dirac-Coulomb spectrum of a point nucleus, whose lowest levels the
  B-spline basis reproduces to 1e-10 of the closed form, and by the
  eigen-equations themselves.
* Record_matrix (record_matrix.f90:2-60): blocks built as S u w^T S from
  solution vectors u (channel i, column a) and w (channel j, column b)
  transform to the unit matrix entry (a, b), which a transposed or
  mis-indexed product would not give.
"""
import numpy as np
import pytest

import oracle as O
from conftest import fe_grid
from test_oracle_fe import fe_case


def cox_de_boor(t, k, x):
    """values of the n = len(t) - k B-splines of order k on knots t at x,
    x taken as the right end when it equals t[-1]"""
    n = len(t) - k
    B = np.zeros(len(t) - 1)
    if x >= t[-1]:
        j = max(i for i in range(len(t) - 1) if t[i] < t[i + 1])
        B[j] = 1.0
        x = t[-1]
    else:
        j = max(i for i in range(len(t) - 1) if t[i] <= x < t[i + 1])
        B[j] = 1.0
    for m in range(2, k + 1):
        N = np.zeros(len(t) - m)
        for i in range(len(t) - m):
            a = (x - t[i]) / (t[i + m - 1] - t[i]) * B[i] if t[i + m - 1] > t[i] else 0.0
            b = (t[i + m] - x) / (t[i + m] - t[i + 1]) * B[i + 1] if t[i + m] > t[i + 1] else 0.0
            N[i] = a + b
        B = N
    return B[:n]


def clamped(t, ns, ks, kk):
    """tp / tq of Def_pbsp / Def_qbsp (dbsr_lib_dbs.f90:1000-1006): order kk
    on t(1) kk times, t(ks+1:ns), t(ns+1) kk times"""
    return np.concatenate([np.full(kk, t[0]), t[ks:ns], np.full(kk, t[ns])])


def test_zero_cond_front_is_where_fe_orbitals_vanish():
    b, pq, kbs, lbs, jbs, nuc, labels, iq, w = fe_case()
    ns, ksp, ksq = b["ns"], b["ksp"], b["ksq"]
    g = O.Grid(b["t"], ns, b["ks"], ksp, ksq)
    iprm, nhm = O.zero_cond(len(kbs), ns, g.nsp, g.nsq, ksp, ksq, lbs, kbs, npert=3)
    ms = 2 * ns
    assert nhm == int(iprm.sum()) and np.all(iprm[len(kbs) * ms:] == 1)
    for i, o in enumerate(b["orbitals"]):
        mb = o["mbs"]
        m = iprm[i * ms:(i + 1) * ms]
        for comp in (0, 1):
            removed = m[comp * ns:comp * ns + mb] == 0
            zero = pq[:mb, comp, i] == 0.0
            assert np.array_equal(removed, zero), (o["label"], comp)


def test_zero_cond_boundary_is_the_spline_alive_at_r_max():
    t, ns, ks, ksp, ksq = fe_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    for kk, nb, off in ((ksp, g.nsp, 0), (ksq, g.nsq, ns)):
        tt = clamped(t, ns, ks, kk)
        alive = np.nonzero(cox_de_boor(tt, kk, tt[-1]) != 0.0)[0]
        assert len(alive) == 1 and alive[0] == nb - 1
        iprm, _ = O.zero_cond(1, ns, g.nsp, g.nsq, ksp, ksq, [0], [-1], ilzero=0, jlzero=0)
        tail = np.nonzero(iprm[off:off + ns] == 0)[0]
        assert np.array_equal(tail, np.arange(nb - 1, ns))


def projector_case(seed, nch=5, ns=18, npert=2):
    """random channel blocks and bound orbitals that are B-orthonormal per
    kappa for a random symmetric positive definite B"""
    rng = np.random.default_rng(seed)
    ms = 2 * ns
    A = rng.standard_normal((ms, ms))
    B = A @ A.T / ms + np.eye(ms)
    kap = [-1, 1, -2]
    nbf = 9
    kbs = [kap[i % 3] for i in range(nbf)]
    pv = np.zeros((ms, nbf + nch), order="F")
    for kp in kap:
        idx = [i for i in range(nbf) if kbs[i] == kp]
        for n, i in enumerate(idx):
            v = rng.standard_normal(ms)
            for j in idx[:n]:
                v -= (pv[:, j] @ B @ v) * pv[:, j]
            pv[:, i] = v / np.sqrt(v @ B @ v)
    ipbs = [0] * nbf + list(range(1, nch + 1))
    kbs = kbs + [kap[c % 3] for c in range(nch)]
    ipch = list(range(nbf + 1, nbf + nch + 1))
    qv = np.asfortranarray(B @ pv)
    icc = np.zeros((nch, nch), dtype=np.int32, order="F")
    n = 0
    for ich in range(nch):
        n += 1
        icc[ich, ich] = n
    for ich in range(1, nch):
        for jch in range(ich):
            n += 1
            icc[ich, jch] = icc[jch, ich] = n
    hch = np.asfortranarray(rng.standard_normal((ms, ms, n)))
    for ich in range(nch):
        k = icc[ich, ich] - 1
        hch[:, :, k] = hch[:, :, k] + hch[:, :, k].T
    # alloc_dbsr_matrix's numbering: icb(ich,ip) runs over ip fastest
    icb = np.asfortranarray(np.arange(1, nch * npert + 1, dtype=np.int32).reshape((nch, npert)))
    hcp = np.asfortranarray(rng.standard_normal((ms, nch * npert)))
    # IBORT: orbital 4 (kappa -2) is not imposed on channel 3
    ibort = lambda i, j: 1 if (i == ipch[2] and j == 3) else 0
    orth = O.orth_conditions(nch, ipch, ipbs, kbs, ibort)
    return dict(B=B, pv=pv, qv=qv, icc=icc, hch=hch, icb=icb, hcp=hcp, orth=orth, npert=npert, nch=nch,
                ipch=ipch, ipbs=ipbs, kbs=kbs, ibort=ibort)


@pytest.mark.parametrize("idiag", [1, -1, 0])
def test_dbs_orth_projects(idiag):
    c = projector_case(7)
    nch, icc, pv, qv = c["nch"], c["icc"], c["pv"], c["qv"]
    assert c["orth"][2] == [6, 9] and c["orth"][0] == [1, 4, 7]
    h0, p0 = c["hch"].copy(), c["hcp"].copy()
    h, hp = c["hch"].copy(order="F"), c["hcp"].copy(order="F")
    O.dbs_orth(h, icc, idiag, c["orth"], pv, qv, hp, c["icb"], c["npert"])
    rng = np.random.default_rng(1)

    def comp(ich, n=3):
        """n vectors y with q^T y = 0 for channel ich's conditions"""
        Q = qv[:, [j - 1 for j in c["orth"][ich - 1]]]
        Y = rng.standard_normal((Q.shape[0], n))
        if Q.shape[1]:
            Y -= Q @ np.linalg.solve(Q.T @ Q, Q.T @ Y)
        return Y

    for ich in range(1, nch + 1):
        for jch in range(1, ich + 1):
            k = icc[ich - 1, jch - 1] - 1
            x, x0 = h[:, :, k], h0[:, :, k]
            sel = not (idiag == 1 and ich != jch) and not (idiag == -1 and ich == jch)
            if not sel:
                assert np.array_equal(x, x0)
                continue
            sc = np.max(np.abs(x0)) * 1e-12
            Pr = pv[:, [j - 1 for j in c["orth"][ich - 1]]]
            Pc = pv[:, [j - 1 for j in c["orth"][jch - 1]]]
            assert np.max(np.abs(Pr.T @ x), initial=0.0) <= sc * 10
            assert np.max(np.abs(x @ Pc), initial=0.0) <= sc * 10
            Z, Y = comp(ich), comp(jch)
            assert np.max(np.abs(Z.T @ x @ Y - Z.T @ x0 @ Y)) <= 1e-11 * np.max(np.abs(Z.T @ x0 @ Y))
    for ich in range(1, nch + 1):
        for ip in range(1, c["npert"] + 1):
            col = c["icb"][ich - 1, ip - 1] - 1
            if idiag != 0:
                assert np.array_equal(hp[:, col], p0[:, col])
                continue
            P = pv[:, [j - 1 for j in c["orth"][ich - 1]]]
            assert np.max(np.abs(P.T @ hp[:, col])) <= 1e-12 * np.max(np.abs(p0[:, col])) * 10
            Z = comp(ich)
            assert np.allclose(Z.T @ hp[:, col], Z.T @ p0[:, col], rtol=0, atol=1e-12 * np.max(np.abs(p0)))


def dirac_case(kappas, Z=26.0, grid=None):
    """channel blocks = the Dirac-Coulomb operator of a point nucleus
    (MAT_dhl_pq, no Bloch term) per channel kappa, overlaps fppqq"""
    t, ns, ks, ksp, ksq = grid or fe_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    fppqq = O.overlaps(g)[2]
    nch = len(kappas)
    ms = 2 * ns
    hch = np.zeros((ms, ms, nch), order="F")
    diag = np.zeros((ms, ms, nch), order="F")
    for i, kp in enumerate(kappas):
        hch[:, :, i] = O.dhl_pq(g, 0, Z, 0.0, 0.0, 0.0, kp, bloch=0)
        diag[:, :, i] = fppqq
    icc = np.zeros((nch, nch), dtype=np.int32, order="F")
    for i in range(nch):
        icc[i, i] = i + 1
    lch = [kp if kp > 0 else -kp - 1 for kp in kappas]
    iprm, nhm = O.zero_cond(nch, ns, g.nsp, g.nsq, ksp, ksq, lch, kappas)
    return g, fppqq, hch, diag, icc, iprm, lch


def dirac_energy(n, kappa, Z):
    a = 1.0 / O.C_AU
    gam = np.sqrt(kappa * kappa - (a * Z) ** 2)
    return O.C_AU ** 2 / np.sqrt(1.0 + (a * Z / (n - abs(kappa) + gam)) ** 2) - O.C_AU ** 2


def test_diag_channels_dirac_coulomb_levels():
    kappas = [-1, 1, -2, 2, -3]
    Z = 26.0
    g, fppqq, hch, diag, icc, iprm, lch = dirac_case(kappas, Z)
    ns, ms = g.ns, g.ms
    d = diag.copy(order="F")
    ipsol = O.diag_channels(hch, d, icc, iprm, ns)
    assert ipsol[0] == 0
    for i, kp in enumerate(kappas):
        ksol = ipsol[i + 1]
        assert 0 < ksol < ns
        E = d[:ksol, ns - 1, i]
        assert np.all(np.diff(E) > 0) and E[0] > O.EDMIN and np.all(np.abs(E) >= O.EGAP)
        n0 = lch[i] + 1
        assert abs(E[0] - dirac_energy(n0, kp, Z)) <= 5e-11 * abs(E[0])
        assert abs(E[1] - dirac_energy(n0 + 1, kp, Z)) <= 2e-9 * abs(E[1])
        keep = iprm[i * ms:(i + 1) * ms] != 0
        for s in range(ksol):
            v = d[:, s, i]
            assert np.all(v[~keep] == 0.0)
            assert v[int(np.argmax(np.abs(v)))] > 0.0
            assert abs(v @ fppqq @ v - 1.0) <= 1e-11
            r = hch[:, :, i] @ v - E[s] * (fppqq @ v)
            assert np.max(np.abs(r[keep])) <= 1e-9 * np.max(np.abs(hch[:, :, i]))
        # the rest of column ns and the columns after ksol keep the overlap
        assert np.array_equal(d[ksol:, ns - 1, i], fppqq[ksol:, ns - 1])
        assert np.array_equal(d[:, ksol:ns - 1, i], fppqq[:, ksol:ns - 1])
    # the number of solutions in the window: eigenvalues of another LAPACK
    # driver (dsyevd-based eigh) on the same compacted problems
    import scipy.linalg
    for i in range(len(kappas)):
        keep = np.nonzero(iprm[i * ms:(i + 1) * ms])[0]
        ev = scipy.linalg.eigh(hch[np.ix_(keep, keep, [i])][:, :, 0], fppqq[np.ix_(keep, keep)], eigvals_only=True)
        assert ipsol[i + 1] == int(np.count_nonzero((ev >= O.EDMIN) & (ev <= O.EDMAX) & (np.abs(ev) >= O.EGAP)))


def test_diag_channels_more_solutions_than_ns():
    """every eigenvalue kept: solution ns overwrites the eigenvalues stored
    in column ns before it, and the eigenvalues of solutions ns.. are stored
    over its entries (diag.f90:56-57 writes in this order)"""
    kappas = [-1]
    from conftest import small_grid
    g, fppqq, hch, diag, icc, iprm, lch = dirac_case(kappas, 10.0, grid=small_grid())
    ns, ms = g.ns, g.ms
    d = diag.copy(order="F")
    ipsol = O.diag_channels(hch, d, icc, iprm, ns, edmin=-1e300, edmax=1e300, egap=0.0)
    ksol = ipsol[1]
    keep = np.nonzero(iprm[:ms])[0]
    assert ksol == len(keep) > ns
    from scipy.linalg.lapack import dsygv
    w, v, info = dsygv(np.asfortranarray(hch[np.ix_(keep, keep, [0])][:, :, 0]),
                       np.asfortranarray(fppqq[np.ix_(keep, keep)]), 1, "V", "L")
    for s in range(ksol):
        x = np.zeros(ms)
        x[keep] = v[:, s]
        x = -x if x[int(np.argmax(np.abs(x)))] < 0 else x
        if s != ns - 1:
            assert np.array_equal(d[:, s, 0], x)
        else:
            rows = np.arange(ms)
            eig = (rows >= ns - 1) & (rows < ksol)
            assert np.array_equal(d[~eig, s, 0], x[~eig])
            assert np.array_equal(d[eig, s, 0], w[rows[eig]])


def test_record_matrix_unit_entries():
    kappas = [-1, 1, -2, -1]
    g, fppqq, hch_d, diag, icc_d, iprm, lch = dirac_case(kappas, 20.0)
    ns, ms = g.ns, g.ms
    nch = len(kappas)
    d = diag.copy(order="F")
    ipsol = O.diag_channels(hch_d, d, icc_d, iprm, ns)
    # off-diagonal storage (idiag = -1) and perturbers
    npert = 2
    icc = np.zeros((nch, nch), dtype=np.int32, order="F")
    n = 0
    for ich in range(1, nch):
        for jch in range(ich):
            n += 1
            icc[ich, jch] = icc[jch, ich] = n
    hch = np.zeros((ms, ms, n), order="F")
    want = {}
    rng = np.random.default_rng(3)
    for ich in range(2, nch + 1):
        for jch in range(1, ich):
            if (ich, jch) == (3, 2):
                continue  # a zero block is not recorded
            a = int(rng.integers(0, ipsol[ich]))
            b = int(rng.integers(0, ipsol[jch]))
            while b == a:
                b = int(rng.integers(0, ipsol[jch]))
            u, w = d[:, a, ich - 1], d[:, b, jch - 1]
            hch[:, :, icc[ich - 1, jch - 1] - 1] = 2.5 * np.outer(fppqq @ u, fppqq @ w)
            want[(ich, jch)] = (a, b)
    icb = np.arange(1, nch * npert + 1, dtype=np.int32).reshape((nch, npert), order="F")
    hcp = np.zeros((ms, nch * npert), order="F")
    wv = {}
    for ich in range(1, nch + 1):
        for ip in range(1, npert + 1):
            if (ich, ip) == (2, 1):
                continue
            a = int(rng.integers(0, ipsol[ich]))
            hcp[:, icb[ich - 1, ip - 1] - 1] = -1.5 * (fppqq @ d[:, a, ich - 1])
            wv[(nch + ip, ich)] = a
    ibb = np.array([[1, 2], [2, 3]], dtype=np.int32, order="F")
    hp = np.array([0.5, 0.0, -2.0])
    recs, ov = O.record_matrix(hch, icc, d, ipsol, 0, hcp, icb, hp, ibb, npert)
    keys = [r[0] for r in recs]
    order = [(i, j) for i in range(2, nch + 1) for j in range(1, i) if (i, j) != (3, 2)]
    order += [(nch + ip, i) for i in range(1, nch + 1) for ip in range(1, npert + 1) if (i, ip) != (2, 1)]
    order += [(nch + 1, nch + 1), (nch + 2, nch + 2), (0, 0)]
    assert keys == order
    for (key, val) in recs[:-1]:
        if key in want:
            a, b = want[key]
            e = np.zeros((ipsol[key[0]], ipsol[key[1]]))
            e[a, b] = 2.5
            assert val.shape == e.shape and np.max(np.abs(val - e)) <= 1e-11
            assert abs(ov[key[0] - 1, key[1] - 1] - 2.5) <= 1e-11
        elif key in wv:
            e = np.zeros(ipsol[key[1]])
            e[wv[key]] = -1.5
            assert np.max(np.abs(val - e)) <= 1e-11
            assert abs(ov[key[0] - 1, key[1] - 1] - 1.5) <= 1e-11
    assert recs[-3][1] == 0.5 and recs[-2][1] == -2.0
    assert ov[nch, nch] == 0.5 and ov[nch + 1, nch + 1] == 2.0 and ov[2, 1] == 0.0


def test_product_host_tables_equal_oracle():
    """Zero_cond and the condition lists of the library (host code, no GPU
    work) equal the oracle's for every rule branch"""
    from dbsr3_amd.channels import Zero_cond, orth_conditions
    t, ns, ks, ksp, ksq = fe_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    lch = [0, 1, 1, 2, 3, 7, 8, 0]
    kch = [-1, 1, -2, 2, -4, 7, -9, -1]
    for rule in ((-1, -1, 1, 1), (-1, -1, 2, 2), (0, 0, 1, 1), (2, -1, 0, 3), (-1, 1, 1, 0)):
        a = Zero_cond(len(lch), g, lch, kch, *rule, npert=4)
        o = O.zero_cond(len(lch), ns, g.nsp, g.nsq, ksp, ksq, lch, kch, *rule, npert=4)
        assert np.array_equal(a[0], o[0]) and a[1] == o[1], rule
    c = projector_case(7)
    nort, iort = orth_conditions(c["ipch"], c["ipbs"], c["kbs"], c["ibort"])
    for ich, lst in enumerate(c["orth"]):
        assert list(iort[ich, :nort[ich]]) == lst
