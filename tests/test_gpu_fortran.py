"""The Fortran drop-in layer (dbsr3_amd/fortran/dbs_gpu.f90) driven by a host
program written like the reference's callers (tests/fortran/drv_dropin.f90):
rk(), the Rk_evaluate sequence mrk_*/Density/Convol/SUM_AmB, quadr and a
direct potential, checked against the oracle at 1e-12."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, fe_grid, random_orbitals, small_grid

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "tests", "fortran", "drv_dropin")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_dropin(tmp_path, gpu_available):
    t, ns, ks, ksp, ksq = fe_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    fs = random_orbitals(g, 4, seed=44)
    pq = np.stack(fs, axis=2)
    k = 2
    ms = 2 * ns
    # mchf_matrix lists for orbital 1 (every occurrence pattern), dhl, fppqq, iprm
    rng = np.random.default_rng(45)
    llist = [(1, 1, 1.25), (1, 1, 0.75), (2, 1, 0.3), (1, 3, -0.2)]
    pats = [(1, 1, 1, 1), (1, 2, 1, 3), (2, 1, 3, 1), (1, 2, 3, 1), (1, 1, 2, 3), (2, 1, 1, 3), (2, 3, 1, 1),
            (2, 1, 1, 1), (1, 2, 1, 1), (1, 2, 3, 4), (2, 1, 3, 4), (2, 3, 1, 4), (2, 3, 4, 1)]
    rlist = [(int(rng.integers(0, 3)), *p, float(rng.uniform(-1, 1))) for p in pats]
    dhl = O.dhl_pq(g, 0, 26.0, 0.0, 0.0, 0.0, -1)
    dhl = np.asfortranarray((dhl + dhl.T) / 2)
    fppqq = O.overlaps(g)[2]
    iprm = O.boundary_conditions(ns, g.nsp, g.nsq, ksp, ksq, [0] * 4, [-1] * 4, "point")[:, 0]
    with open(tmp_path / "drv_in.bin", "wb") as f:
        np.array([ns, ks, ksp, ksq], dtype=np.int32).tofile(f)
        np.asarray(t, dtype=np.float64).tofile(f)
        np.array([4, k], dtype=np.int32).tofile(f)
        np.asfortranarray(pq).T.tofile(f)  # column-major bytes
        np.array([len(llist), len(rlist)], dtype=np.int32).tofile(f)
        np.array([r[:2] for r in llist], dtype=np.int32).tofile(f)      # l_i(2,nl)
        np.array([r[2] for r in llist]).tofile(f)
        np.array([r[:5] for r in rlist], dtype=np.int32).tofile(f)      # r_i(5,nr)
        np.array([r[5] for r in rlist]).tofile(f)
        dhl.T.tofile(f)
        np.asfortranarray(fppqq).T.tofile(f)
        np.asarray(iprm, dtype=np.int32).tofile(f)
    print("[fortran drop-in driver]", flush=True)
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "drv_out.bin").exists(), r.stdout + r.stderr
    out = np.fromfile(tmp_path / "drv_out.bin", dtype=np.float64)
    assert out[6] == ns and out[7] == k
    ref = g.rk(fs[0], fs[1], fs[2], fs[3], k)
    assert abs(out[0] - ref) <= 1e-12 * abs(ref)
    assert abs(out[1] - ref) <= 1e-12 * abs(ref)
    for i, m in ((2, 0), (3, 1), (4, -1)):
        q = g.quadr(fs[0], fs[1], m)
        assert abs(out[i] - q) <= 1e-12 * max(abs(q), 1e-300) * 10
    # <d13 | V(d24)> with V from convol case 1 equals the pppp part of rk
    d1 = O.density(ns, ks, fs[0][:, 0], fs[2][:, 0], "s")
    d2 = O.density(ns, ks, fs[1][:, 0], fs[3][:, 0], "s")
    v = O.convol(ns, ks, g.mrk(O.PPPP, k), d2, 1, "s", "s")
    x = O.sum_amb(ns, ks, v, d1, "s")
    assert abs(out[5] - x) <= 1e-12 * abs(x)
    # interleaved mrk_/msk_ selections: convol_sk used ppqq, convol used qqqq
    dn1 = O.density(ns, ks, fs[0][:, 0], fs[2][:, 1], "n")
    dn2 = O.density(ns, ks, fs[1][:, 0], fs[3][:, 1], "n")
    v = O.convol(ns, ks, g.mrk(O.PPQQ, k), dn1, 2, "n", "n")
    x = O.sum_amb(ns, ks, v, dn2, "n")
    assert abs(out[8] - x) <= 1e-12 * np.sum(np.abs(v * dn2)), (out[8], x)
    d1 = O.density(ns, ks, fs[0][:, 1], fs[2][:, 1], "s")
    d2 = O.density(ns, ks, fs[1][:, 1], fs[3][:, 1], "s")
    v = O.convol(ns, ks, g.mrk(O.QQQQ, k), d1, 2, "s", "s")
    x = O.sum_amb(ns, ks, v, d2, "s")
    assert abs(out[9] - x) <= 1e-12 * abs(x), (out[9], x)
    # hf_matrix of orbital 1 with a(1,2,0) = 1, b(1,3,1) = 0.5, qsum = 2
    ms = 2 * ns
    hfm = out[10:10 + ms * ms].reshape((ms, ms), order="F")
    A, B = np.zeros((4, 3), order="F"), np.zeros((4, 3), order="F")
    A[1, 0], B[2, 1] = 1.0, 0.5
    ref = O.hf_matrix(g, 1, np.asfortranarray(pq), A, B, 2, 2.0)
    assert np.max(np.abs(hfm - ref)) <= 1e-12 * np.max(np.abs(ref))
    # the engine: E = R^0(1,2;3,4) - 0.5 R^1(1,1;2,2) + 0.25 R^2(2,3;2,3)
    energy = out[10 + ms * ms]
    rows = np.array([(0, 1, 2, 3, 4), (1, 1, 1, 2, 2), (2, 2, 3, 2, 3)], dtype=np.int32)
    vals = O.rk_rows(g, fs, rows)
    e_ref = vals @ np.array([1.0, -0.5, 0.25])
    assert abs(energy - e_ref) <= 1e-12 * np.sum(np.abs(vals))
    pos = 11 + ms * ms
    grad = out[pos:pos + ns * 8].reshape((ns, 2, 4), order="F")
    assert np.all(np.isfinite(grad)) and np.any(grad != 0)
    pos += ns * 8
    take = lambda n: out[pos:pos + n]
    mh = take(ms * ms).reshape((ms, ms), order="F"); pos += ms * ms
    mr = take(ms); pos += ms
    mx = take(ms * ms).reshape((ms, ms), order="F"); pos += ms * ms
    e_i, q_i = out[pos], out[pos + 1]; pos += 2
    ch = take(ms * ms).reshape((ms, ms), order="F"); pos += ms * ms
    cx = take(ms * ms).reshape((ms, ms), order="F"); pos += ms * ms
    ev = take(ms); pos += ms
    e_v = out[pos]
    assert pos + 1 == len(out)
    # mchf_matrix (newton) and core_matrix (newton, a(1,1,1) = 0.75) of orbital 1
    pqf = np.asfortranarray(pq)
    h0, r0, x0, e0, s0 = O.mchf_matrix(g, 1, pqf, dhl, llist, rlist, 2, newton=1, fppqq=fppqq)
    assert q_i == s0 and abs(e_i - e0) <= 1e-12 * abs(e0)
    for a, b in ((mh, h0), (mr, r0), (mx, x0)):
        assert np.any(b) and np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    qs = np.array([2.0, 1.5, 1.0, 1.0])
    A2 = A.copy(order="F")
    A2[0, 1] = 0.75
    hc, xc = O.core_matrix(g, 1, pqf, dhl, A2, B, 2, qs, newton=1)
    for a, b in ((ch, hc), (cx, xc)):
        assert np.any(b) and np.max(np.abs(a - b)) <= 1e-13 * np.max(np.abs(b))
    # hf_eiv of 2 dhl + the potential, orthogonal to orbital 2, m = 1
    he = O.hf_matrix(g, 1, pqf, A, B, 2, 2.0, hfm=np.asfortranarray(2.0 * dhl))
    vo = np.concatenate([pq[:, 0, 1], pq[:, 1, 1]])
    vr, er = O.hf_eiv(he, [vo], iprm, 1, fppqq)
    import scipy.linalg
    keep = np.nonzero(iprm)[0]
    # tolerances of the generalised eigen-solve: 0.1 ulp of the spectral
    # radius for the eigenvalue, 4 ulps of it over the gap for the vector
    allv = scipy.linalg.eigh(he[np.ix_(keep, keep)], fppqq[np.ix_(keep, keep)], lower=True, eigvals_only=True)
    ulp = np.finfo(float).eps * np.max(np.abs(allv))
    gap = np.min(np.abs(allv[np.abs(allv - er) > 1e-9 * abs(er)] - er))
    assert abs(e_v - er) <= max(1e-12 * abs(er), 0.1 * ulp)
    assert np.max(np.abs(ev - vr)) <= max(1e-10, 4 * ulp / gap) * np.max(np.abs(vr))


def _write(f, *arrs):
    for a in arrs:
        a = np.asarray(a)
        (np.asfortranarray(a).T if a.ndim > 1 else a).tofile(f)


CI3 = os.path.join(ROOT, "tests", "fortran", "drv_ci3")
MAT3 = os.path.join(ROOT, "tests", "fortran", "drv_mat3")


@pytest.mark.skipif(not os.path.exists(CI3), reason="Fortran driver not built (no flang)")
def test_fortran_ci3_setup_matrix(tmp_path, gpu_available):
    """a DBSR_CI3 Setup_matrix-style host loop (Add_integral with Breit terms,
    one-electron HM/DM terms added between them, small c_data blocks so
    lists flush early, Add_matrix at the end) through the drop-in routines
    of libdbsr3_amd_ci3.a; HM and DM against the oracle's replay of the same
    loop (c_data, Add_integral, Rk_evaluate / Sk_evaluate)"""
    t, ns, ks, ksp, ksq = fe_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    norb, ncj, nzero, mbreit, mtype, mpol, mb, nb, kb = 6, 9, 7, 1, 6, 4, 8, 70, 2
    fs = random_orbitals(g, norb, seed=61)
    pq = np.stack(fs, axis=2)
    rng = np.random.default_rng(62)
    kef = np.array([-1, 1, -2, 2, -3, -1], dtype=np.int32)
    nterm = 300
    term = np.zeros((nterm, 7), dtype=np.int32)
    for n in range(nterm):
        i, j = sorted((int(x) for x in rng.integers(1, ncj + 1, 2)), reverse=True)
        term[n] = [rng.integers(0, 4), i, j, *rng.integers(1, norb + 1, 4)]
    C = rng.uniform(-1, 1, nterm)
    nl = 40
    lterm = np.stack([np.sort(rng.integers(1, nterm + 1, nl)), rng.integers(1, ncj + 1, nl),
                      rng.integers(1, ncj + 1, nl)], axis=1).astype(np.int32)
    lterm[:, 2] = np.minimum(lterm[:, 1], lterm[:, 2])
    lval = rng.uniform(-1, 1, nl)
    with open(tmp_path / "drv_ci3_in.bin", "wb") as f:
        _write(f, np.array([ns, ks, ksp, ksq], np.int32), np.asarray(t, np.float64),
               np.array([norb, ncj, nzero, mbreit, mtype, mpol, mb, nb, kb], np.int32), pq, kef,
               np.array([nterm], np.int32), term.T, C, np.array([nl], np.int32), lterm.T, lval)
    print("[fortran CI3 driver]", flush=True)
    r = subprocess.run([CI3], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "drv_ci3_out.bin").exists(), r.stdout + r.stderr
    out = np.fromfile(tmp_path / "drv_ci3_out.bin", dtype=np.float64)
    HM = out[:ncj * nzero].reshape((ncj, nzero), order="F")
    DM = out[ncj * nzero:]
    # the oracle's replay of the same loop
    HMr, DMr = np.zeros((ncj, nzero), order="F"), np.zeros(ncj)
    flushes = []

    def gen(it, k, lst):
        if len(lst[0]):
            O.rk_evaluate(g, it, k, lst, pq, 2 ** 15, HMr, nzero, DMr, eps_c=1e-10)
            flushes.append(it)
    cd = O.CData(mtype, 0, mpol, mb, nb, kb, 1e-12, gen)
    kap = np.concatenate([[0], kef]).astype(np.int32)
    li = 0
    for n in range(nterm):
        while li < nl and lterm[li, 0] == n + 1:
            _, a, b = lterm[li]
            if b <= nzero:
                HMr[a - 1, b - 1] += lval[li]
            elif a == b:
                DMr[b - 1] += lval[li]
            li += 1
        k, i, j, j1, j2, j3, j4 = (int(x) for x in term[n])
        assert O.add_integral_ci(cd, k, i, j, j1, j2, j3, j4, float(C[n]), kap, 2 ** 15, mbreit) == 0
    early = len(flushes)
    for it in range(1, mtype + 1):
        for k in range(mpol + 1):
            cd.add_matrix(it, k)
    assert early > 0 and set(flushes) >= {1, 2, 3, 4, 5, 6}  # lists flushed early and at the end, all types
    s = max(np.max(np.abs(HMr)), np.max(np.abs(DMr)))
    assert np.max(np.abs(HM - HMr)) <= 1e-12 * s
    assert np.max(np.abs(DM - DMr)) <= 1e-12 * s


@pytest.mark.skipif(not os.path.exists(MAT3), reason="Fortran driver not built (no flang)")
def test_fortran_mat3_state_res(tmp_path, gpu_available):
    """a DBSR_MAT3 State_res-style host program (overlap, L, R and S terms
    of every structure through Add_integral, then Add_matrix over all types)
    with its own KBORT and QUADR_PQ, Lval and one-electron operators,
    through libdbsr3_amd_mat3.a; every matrix against the oracle's replay
    (Idef_*, c_data, O_data, L_data, R_data for the four atypes with the
    asymptotic coefficients, S_data)"""
    from test_mat3_routing import IBO, IPB, KAPPA, LBS, NORB, _terms
    from test_oracle_mat3 import _mats
    t, ns, ks, ksp, ksq = small_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    ms, nch, npert, nvch, mk = 2 * ns, 3, 4, 5, 3
    fs = random_orbitals(g, NORB, seed=71)
    pq = np.stack(fs, axis=2)
    rng = np.random.default_rng(72)
    vch = np.asfortranarray(rng.standard_normal((ms, nvch)))
    fppqq = O.overlaps(g)[2]
    iptar = np.array([0, 1, 2, 2], dtype=np.int32)
    kb = np.array([[(3 * i + j) % nvch + 1 for j in range(1, NORB + 1)] for i in range(1, NORB + 1)], np.int32)
    m = _mats(nch, npert, ms)
    terms = []
    for (orbs, c, ic, jc, io, jo) in _terms(rng, 150):
        # overlap terms need a configuration pair or a channel overlap (Idef_Otype)
        ok0 = (ic > 0 and jc > 0) or io > 0
        icase = int(rng.choice([0, 2, 3] if ok0 else [2, 3]))
        # <i|j> alone is the target structure -- for overlaps always, for R and
        # S when no channel orbital is in the integral: io pairs two channel
        # orbitals there
        if io > 0 and jo == 0 and ic == 0 and (icase == 0 or not any(IPB[o] > 0 for o in orbs)):
            io = (io // IBO) * IBO + int(rng.integers(1, 4))
        terms.append([icase, int(rng.integers(0, 3)), *orbs, ic, jc, io, jo])
    # L-integrals (icase 1): every structure of Idef_Ltype; L(i j) between
    # channel orbitals of one kappa
    from test_oracle_ldata import l_terms
    for (ii, jj, c, ic, jc, io, jo) in l_terms(rng, 80, valid_only=True, ibo=IBO):
        if ic == jc == io == jo == 0 and KAPPA[ii] != KAPPA[jj]:
            jj = ii
        terms.append([1, 0, ii, 4, jj, 4, ic, jc, io, jo])
    terms = np.array(terms, dtype=np.int32)
    C = rng.uniform(-1, 1, len(terms))
    ntarg = nch * (nch + 1) // 2
    kap_list = np.array(sorted(set(int(k) for k in KAPPA[1:])), dtype=np.int32)
    hd = np.asfortranarray(np.stack([O.dhl_pq(g, 0, 4.0, 0.0, 0.0, 0.0, int(k)) for k in kap_list], axis=2))
    lval = np.asfortranarray(rng.uniform(-1, 1, (NORB, NORB)))
    with open(tmp_path / "drv_mat3_in.bin", "wb") as f:
        _write(f, np.array([ns, ks, ksp, ksq], np.int32), np.asarray(t, np.float64),
               np.array([NORB, NORB, IBO, mk, nch, npert, nvch, 1, ntarg], np.int32), pq, LBS[1:], KAPPA[1:],
               IPB[1:], kb, m["icc"], m["icb"], m["ibb"], iptar[1:], vch, fppqq,
               np.array([len(terms)], np.int32), terms.T, C, np.array([len(kap_list)], np.int32), lval, kap_list,
               hd)
    print("[fortran MAT3 driver]", flush=True)
    r = subprocess.run([MAT3], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "drv_mat3_out.bin").exists(), r.stdout + r.stderr
    out = np.fromfile(tmp_path / "drv_mat3_out.bin", dtype=np.float64)
    nblk, ncb, nbb = int(m["icc"].max()), int(m["icb"].max()), int(m["ibb"].max())
    sizes = [ms * ms * nblk, nblk * (mk + 1), ms * ncb, nbb, ntarg, ntarg]
    parts = np.split(out, np.cumsum(sizes)[:-1])
    got = dict(hch=parts[0].reshape((ms, ms, nblk), order="F"), acf=parts[1].reshape((nblk, mk + 1), order="F"),
               hcp=parts[2].reshape((ms, ncb), order="F"), hp=parts[3], htarg=parts[4], otarg=parts[5])
    # the oracle's replay
    kbort = lambda i1, i2: int(kb[i1 - 1, i2 - 1])
    qpq = np.array([[[g.quadr(fs[i], fs[j], k) for k in range(mk + 1)] for j in range(NORB)] for i in range(NORB)])
    acf, htarg, otarg = np.zeros((nblk, mk + 1), order="F"), np.zeros(ntarg), np.zeros(ntarg)
    for icase in (0, 1, 2, 3):
        def gen(it, k, lst):
            if not len(lst[0]):
                return
            if icase == 1:
                O.l_data(ms, it, lst, 1e-10, IBO, IPB, kbort, vch, fppqq, 1, iptar, htarg, otarg, 1, lval, KAPPA,
                         kap_list, hd, pq, m)
            elif icase == 0:
                assert O.o_data(ms, it, lst[:3], 1e-10, IBO, IPB, kbort, vch, fppqq, 1, iptar, otarg, m) == 0
            elif icase == 2:
                for atype in range(4):
                    assert O.rs_data(g, 0, it, atype, k, lst, pq, IBO, IPB, kbort, vch, fppqq, 1, iptar, htarg,
                                     m) == 0
                    if it == 3 and atype <= 1:
                        assert O.rdata_acf(lst, IBO, k, mk, qpq[:, :, k], m["icc"], acf) == 0
            else:
                assert O.rs_data(g, 1, it % 10, it // 10, k, lst, pq, IBO, IPB, kbort, vch, fppqq, 1, iptar, htarg,
                                 m) == 0
        cd = O.CData(19, 0, mk + 1, 64, 400, 40, 1e-14, gen)
        for row, c in zip(terms, C):
            if row[0] == icase == 1:
                lt, key = O.idef_ltype(int(row[2]), int(row[4]), float(c), *(int(x) for x in row[6:10]), IPB,
                                       1e-10, 1)
                if lt:
                    cd.add_coef(float(c), 0, *key, lt)
            elif row[0] == icase:
                assert cd.add_integral_mat3(icase, *(int(x) for x in row[1:6]), float(c),
                                            *(int(x) for x in row[6:10]), LBS, KAPPA, IPB, mk=mk, eps_c=1e-10,
                                            check_target=1) == 0
        for k in range(mk + 2):
            for it in range(1, 20):
                cd.add_matrix(it, k)
    want = dict(hch=m["hch"], acf=acf, hcp=m["hcp"], hp=m["hp"], htarg=htarg, otarg=otarg)
    for key, ref in want.items():
        s = np.max(np.abs(ref))
        assert s > 0, key
        assert np.max(np.abs(got[key] - ref)) <= 1e-12 * s, key


CHAN = os.path.join(ROOT, "tests", "fortran", "drv_chan")


def _fortran_records(path):
    """the records of a Fortran unformatted sequential file (4-byte length
    markers around each record), as raw bytes"""
    raw = open(path, "rb").read()
    out, o = [], 0
    while o < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, o)[0])
        assert int(np.frombuffer(raw, np.int32, 1, o + 4 + n)[0]) == n
        out.append(raw[o + 4:o + 4 + n])
        o += n + 8
    return out


def _chan_case(seed):
    """one partial wave's blocks as DBSR_MAT3's SUB1 holds them before the
    post-processing, on the Fe grid: 4 channels (kappa -1, 1, -2, -1) and 2
    perturbers; diagonal blocks = the Dirac-Coulomb operator with the Bloch
    term (Z = 26) plus a seeded symmetric 1e-3 part, overlaps fppqq; the
    other blocks, channel-perturber vectors and perturber elements seeded,
    one of each zero (not recorded); 2 bound orbitals per kappa, zero on the
    B-splines Zero_cond removes, B-orthonormal; IBORT = 0 except one pair"""
    t, ns, ks, ksp, ksq = fe_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    ms = g.ms
    rng = np.random.default_rng(seed)
    kch = [-1, 1, -2, -1]
    nch, npert, nb = len(kch), 2, 2
    lch = [k if k > 0 else -k - 1 for k in kch]
    fppqq = O.overlaps(g)[2]
    kset = sorted(set(kch))
    hd = np.zeros((ms, ms, nch), order="F")
    for i, kp in enumerate(kch):
        R = rng.standard_normal((ms, ms)) * 1e-3
        hd[:, :, i] = O.dhl_pq(g, 0, 26.0, 0.0, 0.0, 0.0, kp, bloch=1) + R + R.T
    ovl = np.asfortranarray(np.repeat(fppqq[:, :, None], nch, axis=2))
    nbound = nb * len(kset)
    norb = nbound + nch
    pv = np.zeros((ms, norb), order="F")
    kbs = []
    for kp in kset:
        first = len(kbs)
        alive = O.zero_cond(1, ns, g.nsp, g.nsq, ksp, ksq, [kp if kp > 0 else -kp - 1], [kp])[0]
        for _ in range(nb):
            v = rng.standard_normal(ms) * alive
            for j in range(first, len(kbs)):
                v -= (pv[:, j] @ fppqq @ v) * pv[:, j]
            pv[:, len(kbs)] = v / np.sqrt(v @ fppqq @ v)
            kbs.append(kp)
    kbs = np.array(kbs + kch, np.int32)
    lbs = np.array([k if k > 0 else -k - 1 for k in kbs], np.int32)
    ipbs = np.array([0] * nbound + list(range(1, nch + 1)), np.int32)
    ipch = np.array(range(nbound + 1, norb + 1), np.int32)
    qv = np.asfortranarray(fppqq @ pv)
    ib = np.zeros((norb, norb), np.int32, order="F")
    ib[ipch[3] - 1, 2] = 1           # channel 4 (kappa -1) not kept orthogonal to orbital 3
    # alloc_dbsr_matrix's numbering (DBSR_MAT3/updates.f90:20-70)
    iccd = np.zeros((nch, nch), np.int32, order="F")
    for i in range(nch):
        iccd[i, i] = i + 1
    icco = np.zeros((nch, nch), np.int32, order="F")
    n = 0
    for i in range(1, nch):
        for j in range(i):
            n += 1
            icco[i, j] = icco[j, i] = n
    icb = np.asfortranarray(np.arange(1, nch * npert + 1, dtype=np.int32).reshape((nch, npert)))
    ibb = np.zeros((npert, npert), np.int32, order="F")
    n = 0
    for i in range(npert):
        for j in range(i + 1):
            n += 1
            ibb[i, j] = ibb[j, i] = n
    ho = np.asfortranarray(rng.standard_normal((ms, ms, int(icco.max()))) * 1e-2)
    ho[:, :, icco[2, 1] - 1] = 0.0
    hcp = np.asfortranarray(rng.standard_normal((ms, int(icb.max()))) * 1e-2)
    hcp[:, icb[1, 0] - 1] = 0.0
    hp = rng.standard_normal(int(ibb.max()))
    hp[ibb[1, 0] - 1] = 0.0
    return dict(t=t, g=g, ns=ns, ks=ks, ksp=ksp, ksq=ksq, kch=np.array(kch, np.int32),
                lch=np.array(lch, np.int32), nch=nch, npert=npert, norb=norb, lbs=lbs, kbs=kbs, ipbs=ipbs,
                ipch=ipch, ib=ib, pv=pv, qv=qv, iccd=iccd, ovl=ovl, hd=hd, icco=icco, icb=icb, ibb=ibb, ho=ho,
                hcp=hcp, hp=hp, edmin=O.EDMIN, edmax=1e4, egap=O.EGAP)


@pytest.mark.skipif(not os.path.exists(CHAN), reason="Fortran driver not built (no flang)")
def test_fortran_mat3_channels(tmp_path, gpu_available):
    """a host program run the way DBSR_MAT3's SUB1 runs one partial wave's
    post-processing (DBS_ORTH, Diag_channels with Zero_cond, DBS_ORTH on the
    other blocks, Record_matrix to an unformatted unit) through the
    reference-named routines of libdbsr3_amd_mat3.a, with its own IBORT,
    Get_pv and Get_qv: iprm, nhm, ipsol, the projected blocks, the
    solutions, every record written and the overlaps against the oracle on
    the same inputs, to the tolerances of the C-ABI channel tests"""
    from test_gpu_channels import _diag_compare
    c = _chan_case(31)
    g, ns, nch, npert, norb = c["g"], c["ns"], c["nch"], c["npert"], c["norb"]
    ms = g.ms
    i32 = lambda *v: np.array(v, np.int32)
    with open(tmp_path / "drv_chan_in.bin", "wb") as f:
        _write(f, i32(ns, c["ks"], c["ksp"], c["ksq"], g.nsp, g.nsq), np.asarray(c["t"], np.float64),
               i32(norb, nch, npert), c["lbs"], c["kbs"], c["ipbs"], c["ipch"], c["lch"], c["kch"], c["ib"],
               c["pv"], c["qv"], c["iccd"], c["ovl"], c["hd"], c["icco"], c["icb"], c["ibb"], c["ho"], c["hcp"],
               c["hp"], np.array([c["edmin"], c["edmax"], c["egap"]]))
    print("[fortran MAT3 channel driver]", flush=True)
    r = subprocess.run([CHAN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert (tmp_path / "drv_chan_out.bin").exists(), r.stdout + r.stderr
    raw = open(tmp_path / "drv_chan_out.bin", "rb").read()
    ni = 1 + nch * ms + npert + nch + 1
    ints = np.frombuffer(raw, np.int32, ni)
    nhm, iprm, ipsol = int(ints[0]), ints[1:1 + nch * ms], ints[1 + nch * ms + npert:]
    nbo, ncb = int(c["icco"].max()), int(c["icb"].max())
    nt = nch + npert
    dbl = np.frombuffer(raw, np.float64, offset=4 * ni)
    sizes = [ms * ms * nch, ms * ms * nch, nt * nt, ms * ms * nbo, ms * ncb]
    assert dbl.size == sum(sizes)
    parts = np.split(dbl, np.cumsum(sizes)[:-1])
    diag = parts[0].reshape((ms, ms, nch), order="F")
    hdo = parts[1].reshape((ms, ms, nch), order="F")
    ov = parts[2].reshape((nt, nt), order="F")
    hout = parts[3].reshape((ms, ms, nbo), order="F")
    hcpo = parts[4].reshape((ms, ncb), order="F")
    # Zero_cond
    iprm_o, nhm_o = O.zero_cond(nch, ns, g.nsp, g.nsq, c["ksp"], c["ksq"], c["lch"], c["kch"])
    assert np.array_equal(iprm, iprm_o[:nch * ms]) and nhm == nhm_o
    # DBS_ORTH on the diagonal blocks
    ibort = lambda i, j: int(c["ib"][i - 1, j - 1])
    orth = O.orth_conditions(nch, c["ipch"], c["ipbs"], c["kbs"], ibort)
    assert orth[3] == [4] and orth[0] == [3, 4] and orth[1] == [5, 6] and orth[2] == [1, 2]
    h1 = c["hd"].copy(order="F")
    O.dbs_orth(h1, c["iccd"], 1, orth, c["pv"], c["qv"])
    for b in range(nch):
        assert np.max(np.abs(hdo[:, :, b] - h1[:, :, b])) <= 1e-12 * np.max(np.abs(h1[:, :, b])), b
    # Diag_channels on the projected blocks
    d_ref = c["ovl"].copy(order="F")
    ip_ref = O.diag_channels(hdo, d_ref, c["iccd"], iprm_o, ns, edmin=c["edmin"], edmax=c["edmax"],
                             egap=c["egap"])
    assert np.array_equal(ipsol, ip_ref) and np.all(ipsol[1:] > 0) and np.all(ipsol[1:] < ns)
    _diag_compare(g, hdo, c["ovl"], diag, d_ref, ipsol, iprm_o, c["iccd"])
    # DBS_ORTH on the other blocks (idiag = -1: channel-perturber vectors kept)
    h2, p2 = c["ho"].copy(order="F"), c["hcp"].copy(order="F")
    O.dbs_orth(h2, c["icco"], -1, orth, c["pv"], c["qv"], p2, c["icb"], npert)
    for b in range(nbo):
        s = np.max(np.abs(h2[:, :, b]))
        assert np.max(np.abs(hout[:, :, b] - h2[:, :, b])) <= 1e-12 * s if s > 0 else not np.any(hout[:, :, b])
    assert np.array_equal(hcpo, c["hcp"])
    # Record_matrix: the file's records against the oracle's on the GPU's
    # own projected blocks and solutions
    rrefs, ovr = O.record_matrix(hout, c["icco"], diag, ipsol, 0, hcpo, c["icb"], c["hp"], c["ibb"], npert)
    recs = _fortran_records(tmp_path / "drv_chan_rec.bin")
    assert len(recs) == 2 * len(rrefs) - 1
    for n, (key, ref) in enumerate(rrefs):
        assert tuple(int(x) for x in np.frombuffer(recs[2 * n], np.int32)) == key, n
        if ref is None:
            break
        v = np.frombuffer(recs[2 * n + 1], np.float64)
        if isinstance(ref, float):
            assert v.size == 1 and v[0] == ref, key
            continue
        ref = np.asarray(ref).reshape(-1, order="F")
        assert v.size == ref.size and np.max(np.abs(v - ref)) <= 1e-12 * np.max(np.abs(ref)), key
    keys = [k for k, _ in rrefs]
    assert (3, 2) not in keys and (nch + 1, 2) not in keys and (nch + 2, nch + 1) not in keys
    assert (nch + 2, nch + 2) in keys and (nch + 1, 1) in keys
    assert np.max(np.abs(ov - ovr)) <= 1e-12 * np.max(np.abs(ovr))
    assert np.array_equal(ov == 0, ovr == 0)
