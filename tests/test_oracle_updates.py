"""Pins of the oracle's overlap, unfolding and MAT3 update functions by
routes other than their own formulas:

* ZINTYM overlaps fpbs/fqbs (DBSR_HF/dbsr_lib_dbs.f90:4472-4497): every
  entry against <B_i|B_j> integrated with scipy's B-splines on the P and Q
  knot sequences and a 24-point Gauss-Legendre rule per interval; the sum of
  all entries equals t(ns+1) - t(1) (partition of unity); fppqq holds them
  as diagonal blocks.  (Bit-exact agreement with the compiled reference is
  in test_oracle_ref.py.)
* Full_mat_sym (DBSR_HF/dbsr_lib_zcom.f90:1247-1294): dense band matrices
  packed into each storage mode come back unchanged.
* UPDATE_HX/HW/HV (DBSR_MAT3/updates.f90:79-106, 183-244) and O_data
  (DBSR_MAT3/O_data.f90:2-93): on small-integer data, where every product
  and sum is exact, the blocks equal the dense algebra (C d, C d^T, C v w^T,
  its transpose, the symmetrised form, the sum of the list), with the
  eps_C cut, the not-stored blocks, iitar and the KBORT=0 stop.
* R_data's asymptotic coefficients (DBSR_MAT3/R_data.f90:214-240): on
  small-integer data, acf(block, kpol) equals the plain sum of c*QUADR_PQ
  over every list entry of that block, a sum with no run logic, so a run
  that is not closed, not reset or closed at the wrong entry fails.
"""
import numpy as np
import pytest

import oracle as O
from conftest import fe_grid, small_grid


def _bspline_matrix(knots, k, x):
    from scipy.interpolate import BSpline
    n = len(knots) - k
    B = np.zeros((len(x), n))
    for i in range(n):
        c = np.zeros(n)
        c[i] = 1.0
        B[:, i] = BSpline(knots, c, k - 1, extrapolate=False)(x)
    return np.nan_to_num(B)


@pytest.mark.parametrize("spec", [small_grid, fe_grid])
def test_overlaps_against_scipy_quadrature(spec):
    t, ns, ks, ksp, ksq = spec()
    g = O.Grid(t, ns, ks, ksp, ksq)
    fp, fq, fpq = O.overlaps(g)
    nv = g.nv
    xg, wg = np.polynomial.legendre.leggauss(24)
    brk = t[ks - 1:ns + 1]  # interval ends t(ks)..t(ns+1)
    x = np.concatenate([0.5 * (brk[i + 1] - brk[i]) * xg + 0.5 * (brk[i + 1] + brk[i]) for i in range(nv)])
    w = np.concatenate([0.5 * (brk[i + 1] - brk[i]) * wg for i in range(nv)])
    for kk, f in ((ksp, fp), (ksq, fq)):
        nsx = nv + kk - 1
        kn = np.concatenate([np.full(kk, t[0]), t[ks:ns], np.full(kk, t[ns])])
        B = _bspline_matrix(kn, kk, x)
        ref = (B * w[:, None]).T @ B
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(f[:nsx, :nsx] - ref)) < 1e-13 * scale
        assert not np.any(f[nsx:, :]) and not np.any(f[:, nsx:])
        assert abs(f.sum() - (t[ns] - t[0])) < 1e-12 * (t[ns] - t[0])
        assert np.max(np.abs(f - f.T)) <= 4e-16 * scale
    assert np.array_equal(fpq[:ns, :ns], fp) and np.array_equal(fpq[ns:, ns:], fq)
    assert not np.any(fpq[:ns, ns:]) and not np.any(fpq[ns:, :ns])


def test_full_mat_sym_round_trips_band_storage():
    rng = np.random.default_rng(3)
    ns, ks = 11, 4
    A = rng.standard_normal((ns, ns))
    i, j = np.indices((ns, ns))
    band = np.abs(i - j) < ks
    N = np.where(band, A, 0.0)
    S = np.triu(N) + np.triu(N, 1).T
    xs = np.zeros((ns, ks), order="F")
    xl = np.zeros((ns, ks), order="F")
    xn = np.zeros((ns, 2 * ks - 1), order="F")
    for a in range(ns):
        for b in range(ns):
            if b >= a and b - a < ks:
                xs[a, b - a] = S[a, b]       # 's': x(i, j-i+1)
            if a >= b and a - b < ks:
                xl[a, b - a + ks - 1] = S[a, b]  # 'l': x(i, j-i+ks)
            if abs(a - b) < ks:
                xn[a, b - a + ks - 1] = N[a, b]  # 'n': x(i, j-i+ks)
    assert np.array_equal(O.full_mat_sym(ns, ks, xs, "s"), S)
    assert np.array_equal(O.full_mat_sym(ns, ks, xl, "l"), S)
    assert np.array_equal(O.full_mat_sym(ns, ks, xn, "n"), N)
    assert np.array_equal(O.full_mat_sym(ns, ks, np.asfortranarray(A), "x"), A)


def _ints(rng, *shape):
    return np.asfortranarray(rng.integers(-8, 9, shape).astype(np.float64))


def test_update_hx_hw_hv_exact_algebra():
    rng = np.random.default_rng(4)
    ms = 10
    d = _ints(rng, ms, ms)
    v, w = _ints(rng, ms), _ints(rng, ms)
    for ich, jch, expect in ((3, 1, 1.5 * d), (1, 3, 1.5 * d.T), (2, 2, 1.5 * d)):
        blk = np.zeros((ms, ms), order="F")
        O.update_hx(blk, ich, jch, d, 1.5)
        assert np.array_equal(blk, expect)
    x = np.outer(v, w)
    for ich, jch, expect in ((3, 1, -2.0 * x), (1, 3, -2.0 * x.T), (2, 2, -(x + x.T))):
        blk = np.asfortranarray(np.ones((ms, ms)))
        O.update_hw(blk, ich, jch, v, w, -2.0)
        assert np.array_equal(blk, 1.0 + expect)
    col = np.asfortranarray(np.arange(ms, dtype=np.float64))
    O.update_hv(col, v, 0.5)
    assert np.array_equal(col, np.arange(ms) + 0.5 * v)


def _mat3(nch, npert, ms, idiag=0):
    """MAT3 storage of one process (alloc_dbsr_matrix, DBSR_MAT3/updates.f90:2-75)"""
    icc = np.zeros((nch, nch), dtype=np.int32, order="F")
    i = 0
    if idiag >= 0:
        for ich in range(nch):
            i += 1
            icc[ich, ich] = i
    if idiag <= 0:
        for ich in range(1, nch):
            for jch in range(ich):
                i += 1
                icc[ich, jch] = icc[jch, ich] = i
    icb = np.arange(1, nch * npert + 1, dtype=np.int32).reshape((nch, npert))
    icb = np.asfortranarray(icb)
    ibb = np.zeros((npert, npert), dtype=np.int32, order="F")
    i = 0
    for ic in range(npert):
        for jc in range(ic + 1):
            i += 1
            ibb[ic, jc] = ibb[jc, ic] = i
    nb = int(icc.max())
    return dict(icc=icc, icb=icb, ibb=ibb, hch=np.zeros((ms, ms, nb), order="F"),
                hcp=np.zeros((ms, nch * npert), order="F"), hp=np.zeros(i))


IBO = 100


def test_o_data_equals_dense_sums():
    rng = np.random.default_rng(6)
    ms, nch, npert, nvch = 8, 3, 2, 4
    vch = _ints(rng, ms, nvch)
    fppqq = _ints(rng, ms, ms)
    # orbitals 1..6: channel orbitals 1,2,3 (channels 1,2,3); bound pairs
    # (i1,i2) -> column (i1 + i2) % nvch + 1
    ipbs = np.array([0, 1, 2, 3, 0, 0, 0], dtype=np.int32)
    iptar = np.array([0, 1, 2, 2], dtype=np.int32)
    kb = lambda i1, i2: (i1 + i2) % nvch + 1
    n = 40
    c = rng.integers(-6, 7, n) * 0.25
    c[::7] = 1e-12  # below eps_C
    # itype 1: bound-bound
    m = _mat3(nch, npert, ms)
    k1, k2 = rng.integers(1, npert + 1, n), rng.integers(1, npert + 1, n)
    assert O.o_data(ms, 1, (c, k1, k2), 1e-10, IBO, ipbs, kb, vch, fppqq, 0, iptar, np.zeros(3), m) == 0
    hp = np.zeros_like(m["hp"])
    for e in range(n):
        if abs(c[e]) >= 1e-10:
            hp[m["ibb"][k1[e] - 1, k2[e] - 1] - 1] += c[e]
    assert np.array_equal(m["hp"], hp)
    # itype 2: <i|.> jc
    m = _mat3(nch, npert, ms)
    io = rng.integers(1, 4, n) * IBO + rng.integers(1, 7, n)
    jc = rng.integers(1, npert + 1, n)
    assert O.o_data(ms, 2, (c, io, jc), 1e-10, IBO, ipbs, kb, vch, fppqq, 0, iptar, np.zeros(3), m) == 0
    hcp = np.zeros_like(m["hcp"])
    for e in range(n):
        if abs(c[e]) >= 1e-10:
            i1, i2 = divmod(int(io[e]), IBO)
            hcp[:, m["icb"][ipbs[i1] - 1, jc[e] - 1] - 1] += c[e] * vch[:, kb(i1, i2) - 1]
    assert np.array_equal(m["hcp"], hcp)
    # itype 3: <i|.><.|j>
    m = _mat3(nch, npert, ms)
    jo = rng.integers(1, 4, n) * IBO + rng.integers(1, 7, n)
    assert O.o_data(ms, 3, (c, io, jo), 1e-10, IBO, ipbs, kb, vch, fppqq, 0, iptar, np.zeros(3), m) == 0
    hch = np.zeros_like(m["hch"])
    for e in range(n):
        if abs(c[e]) >= 1e-10:
            i1, i2 = divmod(int(io[e]), IBO)
            j1, j2 = divmod(int(jo[e]), IBO)
            ich, jch = ipbs[i1], ipbs[j1]
            x = np.outer(vch[:, kb(i1, i2) - 1], vch[:, kb(j1, j2) - 1])
            xx = c[e] * x if ich > jch else (c[e] * x.T if jch > ich else (c[e] / 2) * (x + x.T))
            hch[:, :, m["icc"][ich - 1, jch - 1] - 1] += xx
    assert np.array_equal(m["hch"], hch)
    # itype 4: target overlaps, fppqq for iitar > 1 and ich != jch
    for iitar in (0, 2):
        m = _mat3(nch, npert, ms)
        io4 = rng.integers(1, 4, n) * IBO + rng.integers(1, 4, n)
        otarg = np.zeros(3)
        assert O.o_data(ms, 4, (c, io4, io4), 1e-10, IBO, ipbs, kb, vch, fppqq, iitar, iptar, otarg, m) == 0
        hch = np.zeros_like(m["hch"])
        ot = np.zeros(3)
        for e in range(n):
            if abs(c[e]) < 1e-10:
                continue
            i1, i2 = divmod(int(io4[e]), IBO)
            ich, jch = ipbs[i1], ipbs[i2]
            it, jt = iptar[ich], iptar[jch]
            hi, lo = max(it, jt), min(it, jt)
            ot[(hi - 1) * hi // 2 + lo - 1] += c[e]
            if iitar > 1 and ich != jch:
                hch[:, :, m["icc"][ich - 1, jch - 1] - 1] += c[e] * (fppqq if ich >= jch else fppqq.T)
        assert np.array_equal(otarg, ot)
        assert np.array_equal(m["hch"], hch)
        assert (iitar == 0) == (not np.any(hch))


def test_o_data_drops_unstored_blocks_and_stops_on_kbort0():
    rng = np.random.default_rng(7)
    ms, nch = 6, 2
    vch = _ints(rng, ms, 2)
    ipbs = np.array([0, 1, 2], dtype=np.int32)
    m = _mat3(nch, 1, ms, idiag=1)  # diagonal blocks only
    c = np.array([1.0, 2.0])
    io = np.array([1 * IBO + 1, 2 * IBO + 1])
    jo = np.array([2 * IBO + 1, 2 * IBO + 1])
    assert O.o_data(ms, 3, (c, io, jo), 1e-10, IBO, ipbs, lambda a, b: 1, vch, vch, 0, ipbs,
                    np.zeros(3), m) == 0
    x = np.outer(vch[:, 0], vch[:, 0])
    assert np.array_equal(m["hch"][:, :, 1], 2.0 * x) and not np.any(m["hch"][:, :, 0])
    assert O.o_data(ms, 2, (c, io, np.array([1, 1])), 1e-10, IBO, ipbs, lambda a, b: 0, vch, vch, 0,
                    ipbs, np.zeros(3), m) == 2


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rdata_acf_against_block_sums(seed):
    rng = np.random.default_rng(seed)
    nch, norb, ibo, n, mk, kpol = 5, 6, 2 ** 15, 300, 4, 3
    icc = np.zeros((nch, nch), dtype=np.int32, order="F")
    b = 0
    for ich in range(nch):
        for jch in range(ich + 1):
            if (ich + jch) % 4 == 3:
                continue  # some blocks not stored here
            b += 1
            icc[ich, jch] = icc[jch, ich] = b
    ich, jch = rng.integers(1, nch + 1, n), rng.integers(1, nch + 1, n)
    K1 = np.maximum(ich, jch) * ibo + np.minimum(ich, jch)
    K2, K3 = rng.integers(1, norb + 1, n), rng.integers(1, norb + 1, n)
    order = np.lexsort((K2, K1))
    order = np.concatenate([order[:n - 7], order[n - 7:][::-1]])  # runs that repeat later
    c = rng.integers(-8, 9, n).astype(np.float64)
    lst = (c[order], K1[order], K2[order], K3[order])
    qpq = rng.integers(-5, 6, (norb, norb)).astype(np.float64)
    acf = np.zeros((b, mk + 1), order="F")
    assert O.rdata_acf(lst, ibo, kpol, mk, qpq, icc, acf) == 0
    want = np.zeros((b, mk + 1))
    for cc, k1, k2, k3 in zip(*lst):
        blk = icc[k1 // ibo - 1, k1 % ibo - 1]
        if blk:
            want[blk - 1, kpol] += cc * qpq[k2 - 1, k3 - 1]
    assert np.array_equal(acf, want)
    assert np.any(want != 0)
    assert O.rdata_acf(lst, ibo, mk + 1, mk, qpq, icc, acf) == 1  # 'Update_cf: k > kmk'
