"""Pins of the oracle's mchf_matrix / core_matrix restatement
(DBSR_MCHF/mchf_matrix.f90:2-345), on the CPU.

mchf_matrix and core_matrix cannot be built from the reference: they need
the un-vendored Libraries_03 (module id4_data, c_data of DBSR_MCHF's
library).  They are pinned by what the mathematics fixes:

  * the orbital equation is the energy's gradient: for the energy
    E = sum_L c L(i1,i2) + sum_R c R^k(i1,i2;i3,i4) of the lists,
    hfm v_i - rhs = (1/(2 qsum)) dE/dv_i for lists of the patterns in
    which orbital i occurs once, twice or four times (equal keys merged,
    L(i,j) terms into rhs); the directional derivative is taken with the
    five-point stencil, exact for E's degree-4 polynomial in the step;
  * for i occurring three times (R(a,i;i,i) and its images) the reference
    symmetrises the exchange half (UPDATE_HS 'x'), so only the projection
    v_i (hfm v_i - rhs) = (1/(2 qsum)) v_i dE/dv_i holds (Euler's relation);
    the matrix itself is the sum of two pinned two-occurrence matrices, with
    a copy i' of orbital i: R(a,i;i',i) with c (the direct part) and
    R(i,a;i',i) with c/2 (the exchange part);
  * for lists in which i occurs exactly twice (E quadratic in v_i) the
    whole matrix is fixed: hfm u = (1/(2 qsum)) dE/dv_i at v_i = u for
    random u;
  * the Newton term hx (newton = 1): with one R^k(i,i;i,i) entry of
    coefficient qsum/4, hfm + hx restricted to a P-only (Q-only) direction
    is (1/(2 qsum)) times E's second derivative (the reference's hx takes
    the last density without its coefficient; at 4C/qsum = 1 it equals
    the true term);
  * irhs = 1 moves rhs into hfm: hfm' v = hfm v - rhs (v^T B v);
  * core_matrix's two-electron part equals the compiled reference's
    hf_matrix (oracle/_ref) bit for bit, with a(i,j,k) scaled by
    qsum(i) qsum(j) (qsum(i)(qsum(i)-1)/2 for j = i, which hf_matrix
    doubles) and b(i,j,k) likewise; its hx by the second derivative of
    (c/2) F^k(i,i).

Orbitals: DBSR_MCHF/He/He_n4.bsw (16 orbitals, ns = 90); nucleus from
He.knot (Fermi, Z = 2).
"""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
BSW = os.path.join(HERE, "golden", "he", "He_n4.bsw")
KNOT = os.path.join(HERE, "golden", "knots", "He.knot")


def he_case():
    from dbsr3_amd.bsw import pq_array, read_bsw
    from dbsr3_amd.grid import read_knot
    b = read_bsw(BSW)
    kp = read_knot(KNOT)
    nuc = (2, kp["atomic_number"], 0.0, kp["a_fermi"], kp["c_fermi"])
    kbs = [o["kappa"] for o in b["orbitals"]]
    return b, pq_array(b), kbs, nuc


_G = {}


def grid():
    if "g" not in _G:
        b = he_case()[0]
        _G["g"] = O.Grid(b["t"], b["ns"], b["ks"], b["ksp"], b["ksq"])
    return _G["g"]


def sym_dhl(g, nuc, kappa):
    d = O.dhl_pq(g, *nuc, kappa)
    return np.asfortranarray((d + d.T) / 2.0)


def energy(g, pq, dhl, llist, rlist, S, eps_c=1e-8):
    """E of the lists with the entries mchf_matrix keeps (|c/S| >= eps_c)"""
    fs = [np.asfortranarray(pq[:, :, j]) for j in range(pq.shape[2])]
    e = 0.0
    for i1, i2, c in llist:
        v1 = np.concatenate([fs[i1 - 1][:, 0], fs[i1 - 1][:, 1]])
        v2 = np.concatenate([fs[i2 - 1][:, 0], fs[i2 - 1][:, 1]])
        e += c * (v1 @ dhl @ v2)
    for k, i1, i2, i3, i4, c in rlist:
        if abs(c / S) < eps_c:
            continue
        e += c * g.rk(fs[i1 - 1], fs[i2 - 1], fs[i3 - 1], fs[i4 - 1], k)
    return e


def with_orbital(pq, i, v):
    q = np.array(pq, order="F")
    ns = pq.shape[0]
    q[:, 0, i - 1] = v[:ns]
    q[:, 1, i - 1] = v[ns:]
    return q


def d1(f, h):
    """f'(0), exact for polynomials of degree <= 4"""
    return (f(-2 * h) - 8 * f(-h) + 8 * f(h) - f(2 * h)) / (12 * h)


def d2(f, h):
    """f''(0), exact for polynomials of degree <= 5"""
    return (-f(2 * h) + 16 * f(h) - 30 * f(0.0) + 16 * f(-h) - f(-2 * h)) / (12 * h * h)


def vec(pq, i):
    return np.concatenate([pq[:, 0, i - 1], pq[:, 1, i - 1]])


# orbital 2 = 2s (kappa -1); 1s, 3s, 4s share its kappa
I = 2
A, B, Cc = 7, 4, 12   # 3p-, 2p, 4p


THREE = [(A, I, I, I), (I, A, I, I), (I, I, A, I), (I, I, I, A)]


def all_pattern_lists(rng, kmax=3, three=True):
    """L entries (i,i) summing to qsum, (j,i) and (i,j) with j of i's kappa,
    one not involving i; R entries of the occurrence patterns of i (the
    three-occurrence ones if three), some keys twice"""
    i, a, b, c = I, A, B, Cc
    llist = [(i, i, 1.25), (i, i, 0.75), (1, i, 0.31), (i, 5, -0.22), (1, 1, 0.9)]
    pats = [(i, i, i, i), (i, a, i, b), (a, i, b, i), (i, a, b, i), (i, i, a, b), (a, i, i, b), (a, b, i, i),
            (i, a, b, c), (a, i, b, c), (a, b, i, c), (a, b, c, i), (a, b, a, b)] + (THREE if three else [])
    rlist = []
    for n, p in enumerate(pats):
        for k in sorted(rng.choice(kmax + 1, 2, replace=False)):
            rlist.append((int(k), *p, float(rng.uniform(-1, 1))))
    rlist.append((1, i, a, i, b, 0.37))     # merged with an earlier key when k = 1 was drawn for it
    rlist.append((0, i, i, i, i, 1e-12))    # below eps_c: skipped
    return llist, rlist


@pytest.fixture(scope="module")
def case():
    b, pq, kbs, nuc = he_case()
    g = grid()
    return g, pq, kbs, nuc


def test_he_bsw_matches_log(case):
    """He_n4.bsw's orbital energies are the ones He_n4.log prints
    (DBSR_MCHF/He/He_n4.log:475-490)"""
    b = he_case()[0]
    e = {o["label"]: o["e"] for o in b["orbitals"]}
    assert (b["ns"], b["ks"], b["ksp"], b["ksq"]) == (90, 9, 8, 9) and b["nuclear"] is None
    assert abs(e["1s"] + 0.95880798) <= 5e-9 and abs(e["4f"] + 11.64848121) <= 5e-9


@pytest.mark.parametrize("seed", [1, 2])
@pytest.mark.parametrize("three", [False, True])
def test_mchf_matrix_is_half_the_gradient(case, seed, three):
    g, pq, kbs, nuc = case
    rng = np.random.default_rng(seed)
    llist, rlist = all_pattern_lists(rng, three=three)
    dhl = sym_dhl(g, nuc, kbs[I - 1])
    hfm, rhs, hx, e, S = O.mchf_matrix(g, I, pq, dhl, llist, rlist, 3)
    assert S == 2.0 and np.any(rhs) and not np.any(hx)
    v = vec(pq, I)
    lhs = hfm @ v - rhs
    assert abs(e - v @ lhs) <= 1e-14 * abs(e)
    for n in range(3):
        w = rng.standard_normal(g.ms)
        w *= np.linalg.norm(v) / np.linalg.norm(w)
        if three:
            w = v * (1.0 + n)
        f = lambda t: energy(g, with_orbital(pq, I, v + t * w), dhl, llist, rlist, S)
        gw = d1(f, 0.05)
        print(f"  grad.w {gw:+.15e}  2S (hfm v - rhs).w {2 * S * (lhs @ w):+.15e}")
        assert abs(gw - 2 * S * (lhs @ w)) <= 1e-11 * max(abs(gw), 1.0)


@pytest.mark.parametrize("pat", range(4))
def test_mchf_three_occurrence_matrix_decomposes(case, pat):
    g, pq, kbs, nuc = case
    n = pq.shape[2]
    dup = n + 1
    pq2 = np.asfortranarray(np.concatenate([pq, pq[:, :, I - 1:I]], axis=2))
    dhl = sym_dhl(g, nuc, kbs[I - 1])
    ll = [(I, I, 1.3)]
    for k in (0, 2):
        c = 0.6 - 0.2 * k
        h3, r3, _, _, S = O.mchf_matrix(g, I, pq2, dhl, ll, [(k, *THREE[pat], c)], k)
        hd, rd, _, _, _ = O.mchf_matrix(g, I, pq2, dhl, ll, [(k, A, I, dup, I, c)], k)
        hx_, rx, _, _, _ = O.mchf_matrix(g, I, pq2, dhl, ll, [(k, I, A, dup, I, c / 2)], k)
        assert not np.any(r3) and not np.any(rd) and not np.any(rx)
        ref = hd + hx_ - dhl
        assert np.max(np.abs(h3 - ref)) <= 1e-13 * np.max(np.abs(ref)), (pat, k)


def test_mchf_matrix_quadratic_lists_fix_the_matrix(case):
    g, pq, kbs, nuc = case
    rng = np.random.default_rng(5)
    i, a, b = I, A, B
    llist = [(i, i, 1.5)]
    rlist = [(k, *p, float(rng.uniform(-1, 1))) for k in (0, 1, 2)
             for p in ((i, a, i, b), (a, i, b, i), (i, a, b, i), (i, i, a, b), (a, i, i, b), (a, b, i, i))]
    dhl = sym_dhl(g, nuc, kbs[I - 1])
    hfm, rhs, hx, e, S = O.mchf_matrix(g, I, pq, dhl, llist, rlist, 2)
    assert not np.any(rhs)
    assert np.max(np.abs(hfm - hfm.T)) <= 1e-13 * np.max(np.abs(hfm))
    for _ in range(3):
        u = rng.standard_normal(g.ms)
        w = rng.standard_normal(g.ms)
        f = lambda t: energy(g, with_orbital(pq, I, u + t * w), dhl, llist, rlist, S)
        gw = d1(f, 0.5)
        assert abs(gw - 2 * S * (hfm @ u) @ w) <= 1e-11 * max(abs(gw), 1.0)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_mchf_newton_term_is_the_second_derivative(case, k):
    g, pq, kbs, nuc = case
    S = 1.7
    llist = [(I, I, S)]
    rlist = [(k, I, I, I, I, S / 4)]
    dhl = sym_dhl(g, nuc, kbs[I - 1])
    hfm, rhs, hx, e, S2 = O.mchf_matrix(g, I, pq, dhl, llist, rlist, k, newton=1)
    assert S2 == S and np.any(hx)
    v = vec(pq, I)
    rng = np.random.default_rng(k)
    ns = g.ns
    for comp in (0, 1):
        w = np.zeros(g.ms)
        w[comp * ns:(comp + 1) * ns] = rng.standard_normal(ns)
        w *= np.linalg.norm(v) / np.linalg.norm(w)
        f = lambda t: energy(g, with_orbital(pq, I, v + t * w), dhl, llist, rlist, S)
        h2 = d2(f, 0.05)
        mine = 2 * S * (w @ (hfm + hx) @ w)
        print(f"  comp {comp}: d2E {h2:+.15e}  2S w(hfm+hx)w {mine:+.15e}")
        assert abs(h2 - mine) <= 1e-10 * abs(h2)


def test_mchf_irhs_moves_rhs_into_the_matrix(case):
    g, pq, kbs, nuc = case
    llist, rlist = all_pattern_lists(np.random.default_rng(3))
    dhl = sym_dhl(g, nuc, kbs[I - 1])
    fppqq = O.overlaps(g)[2]
    h0, r0, _, e0, S = O.mchf_matrix(g, I, pq, dhl, llist, rlist, 3)
    h1, r1, _, e1, _ = O.mchf_matrix(g, I, pq, dhl, llist, rlist, 3, irhs=1, fppqq=fppqq)
    v = vec(pq, I)
    assert not np.any(r1) and np.any(r0)
    vbv = v @ fppqq @ v
    ref = h0 @ v - r0 * vbv
    assert np.max(np.abs(h1 @ v - ref)) <= 1e-12 * np.max(np.abs(ref))
    assert abs(e1 - (e0 + (vbv - 1.0) * (v @ r0))) <= 1e-12 * abs(e0)


def test_mchf_qsum_floor_and_tiny_rhs(case):
    """|qsum| < qsum_min takes qsum_min (mchf_matrix.f90:72); sum|rhs| <
    1e-8 drops rhs (:240-241)"""
    g, pq, kbs, nuc = case
    dhl = sym_dhl(g, nuc, kbs[I - 1])
    _, _, _, _, S = O.mchf_matrix(g, I, pq, dhl, [(I, I, 1e-10)], [(0, I, A, I, A, 0.5)], 0)
    assert S == 1e-8
    _, rhs, _, _, _ = O.mchf_matrix(g, I, pq, dhl, [(I, I, 1.0), (1, I, 1e-13)], [(0, I, A, I, A, 0.5)], 0)
    assert not np.any(rhs)


def core_case(seed, kmax):
    g = grid()
    b, pq, kbs, nuc = he_case()
    n = pq.shape[2]
    rng = np.random.default_rng(seed)
    lbs = np.array([k if k > 0 else -k - 1 for k in kbs], dtype=np.int32)
    jbs = np.array([2 * abs(k) - 1 for k in kbs], dtype=np.int32)
    coef = np.asfortranarray(rng.uniform(-1, 1, (n, n, kmax + 1)))
    qsum = rng.uniform(1.0, 4.0, n)
    return g, pq, np.asarray(kbs, dtype=np.int32), lbs, jbs, coef, qsum


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")
def test_core_matrix_two_electron_part_is_reference_hf_matrix():
    from test_oracle_hf_matrix import coefs_ab
    kmax = 2
    g, pq, kbs, lbs, jbs, coef, qsum = core_case(4, kmax)
    n = pq.shape[2]
    R = O.ref()
    b = he_case()[0]
    R.ref_setup(b["ns"], b["ks"], b["ksp"], b["ksq"], O._d(O.f64(b["t"])))
    z = np.zeros(1, np.int32)
    for i in (1, 2, 4, 9):
        A_, B_ = coefs_ab(i, coef, lbs, jbs, kmax)
        # coef' whose a(i,.,.) / b(i,.,.) are core_matrix's scaled coefficients
        c2 = np.array(coef, order="F")
        for j in range(1, n + 1):
            for k in range(kmax + 1):
                if j == i:
                    c2[i - 1, i - 1, k] = coef[i - 1, i - 1, k] * qsum[i - 1] * (qsum[i - 1] - 1) / 2
                else:
                    c2[max(i, j) - 1, min(i, j) - 1, k] = coef[max(i, j) - 1, min(i, j) - 1, k] * qsum[i - 1] * qsum[j - 1]
                    c2[min(i, j) - 1, max(i, j) - 1, k] = coef[min(i, j) - 1, max(i, j) - 1, k] * qsum[i - 1] * qsum[j - 1]
        R.ref_hf_setup(n, kmax, 0, O._i(kbs), O._i(lbs), O._i(jbs), O._d(qsum), O._d(c2), O._d(pq), 0,
                       O._i(z), O._i(z), O._i(z), O._i(z), O._d(np.zeros(1)))
        ref = np.zeros((g.ms, g.ms), order="F")
        R.ref_hf_matrix(i, O._d(ref))
        hfm, hx = O.core_matrix(g, i, pq, np.zeros((g.ms, g.ms)), A_, B_, kmax, qsum)
        assert np.any(ref) and not np.any(hx)
        assert np.array_equal(hfm, ref), (i, np.max(np.abs(hfm - ref)))


@pytest.mark.parametrize("k", [0, 2])
def test_core_newton_term_is_the_second_derivative(k):
    g, pq, kbs, lbs, jbs, coef, qsum = core_case(6, k)
    n = pq.shape[2]
    i = 2
    A_ = np.zeros((n, k + 1), order="F")
    A_[i - 1, k] = 0.8
    B_ = np.zeros((n, k + 1), order="F")
    hfm, hx = O.core_matrix(g, i, pq, np.zeros((g.ms, g.ms)), A_, B_, k, qsum, newton=1)
    q = qsum[i - 1]
    c = 0.8 * q * (q - 1)
    v = vec(pq, i)
    rng = np.random.default_rng(k)
    fs = lambda p: np.asfortranarray(p[:, :, i - 1])
    for comp in (0, 1):
        w = np.zeros(g.ms)
        w[comp * g.ns:(comp + 1) * g.ns] = rng.standard_normal(g.ns)
        w *= np.linalg.norm(v) / np.linalg.norm(w)

        def f(t):
            f1 = fs(with_orbital(pq, i, v + t * w))
            return c / 2 * g.rk(f1, f1, f1, f1, k)
        h2 = d2(f, 0.05)
        mine = 2 * q * (w @ (hfm + hx) @ w)
        assert abs(h2 - mine) <= 1e-10 * abs(h2), comp
