"""GPU mchf_matrix / core_matrix (DBSR_MCHF/mchf_matrix.f90:2-345) on the He
orbitals of DBSR_MCHF/He/He_n4.bsw (16 orbitals, ns = 90): against the
oracle (pinned in tests/test_oracle_mchf.py: gradient and second-derivative
identities, core_matrix bit for bit against the compiled hf_matrix), and
against the engine's adjoint, an independent GPU path: 2 qsum (hfm v - rhs)
is the gradient of the lists' energy for orbital i.

Tolerance: 1e-13 of the largest entry for matrices and vectors (the device
sums the same products in other orders), 1e-12 relative for e(i)."""
import numpy as np
import pytest

import oracle as O
from test_oracle_mchf import A, I, all_pattern_lists, core_case, grid, he_case, sym_dhl, vec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def he(gpu_available):
    from dbsr3_amd import DBS
    from dbsr3_amd.dbs import fortran_to_device
    b, pq, kbs, nuc = he_case()
    d = DBS(b["t"], b["ns"], b["ks"], b["ksp"], b["ksq"])
    return d, grid(), pq, fortran_to_device(pq), kbs, nuc


def close(got, ref, tol=1e-13):
    scale = max(np.max(np.abs(ref)), 1e-300)
    return np.max(np.abs(got - ref)) <= tol * scale


@pytest.mark.parametrize("seed,three,newton,irhs", [(1, False, 0, 0), (2, True, 0, 0), (3, True, 1, 0),
                                                     (4, True, 1, 1), (5, False, 0, 1)])
def test_mchf_matrix_vs_oracle(he, seed, three, newton, irhs):
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    d, g, pq, pqd, kbs, nuc = he
    print(f"[mchf_matrix seed={seed} three={three} newton={newton} irhs={irhs}]", flush=True)
    llist, rlist = all_pattern_lists(np.random.default_rng(seed), three=three)
    dhl = sym_dhl(g, nuc, kbs[I - 1])
    fppqq = O.overlaps(g)[2]
    h0, r0, x0, e0, S0 = O.mchf_matrix(g, I, pq, dhl, llist, rlist, 3, newton=newton, irhs=irhs, fppqq=fppqq)
    hfm, rhs, hx, e, S = d.mchf_matrix(I, pqd, fortran_to_device(dhl), llist, rlist, 3, newton=newton, irhs=irhs,
                                       fppqq=fortran_to_device(fppqq))
    assert S == S0
    assert close(device_to_fortran(hfm), h0)
    r = rhs.cpu().numpy()
    assert close(r, r0) if np.any(r0) else not np.any(r)
    if newton:
        assert np.any(x0) and close(device_to_fortran(hx), x0)
    else:
        assert not np.any(device_to_fortran(hx))
    assert abs(e - e0) <= 1e-12 * abs(e0)


def test_mchf_matrix_is_the_engine_adjoint(he):
    """two independent GPU paths: 2 qsum (hfm v - rhs) = dE/dv_i, E the
    lists' energy (R^k rows through the engine, L terms on the host)"""
    from dbsr3_amd import Engine
    from dbsr3_amd.dbs import fortran_to_device
    d, g, pq, pqd, kbs, nuc = he
    llist, rlist = all_pattern_lists(np.random.default_rng(7), three=False)
    dhl = sym_dhl(g, nuc, kbs[I - 1])
    hfm, rhs, _, _, S = d.mchf_matrix(I, pqd, fortran_to_device(dhl), llist, rlist, 3)
    v = vec(pq, I)
    from dbsr3_amd.dbs import device_to_fortran
    lhs = device_to_fortran(hfm) @ v - rhs.cpu().numpy()
    keep = [r for r in rlist if abs(r[5] / S) >= 1e-8]
    eng = Engine([(d, pq.shape[2], np.array([r[:5] for r in keep], dtype=np.int32))])
    eng.set_orbitals(0, pq)
    eng.set_coef(0, np.array([r[5] for r in keep]))
    eng.forward_adjoint()
    gr = eng.grad()[:, :, I - 1]
    grad = np.concatenate([gr[:, 0], gr[:, 1]])
    for i1, i2, c in llist:
        if i1 == I:
            grad += c * dhl @ vec(pq, i2)
        if i2 == I:
            grad += c * dhl.T @ vec(pq, i1)
    assert close(2 * S * lhs, grad, 1e-12)


@pytest.mark.parametrize("newton", [0, 1])
def test_core_matrix_vs_oracle(he, newton):
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    from test_oracle_hf_matrix import coefs_ab
    d, g, pq, pqd, kbs, nuc = he
    kmax = 2
    g, pq, kb, lbs, jbs, coef, qsum = core_case(4, kmax)
    for i in (1, 3, 10):
        print(f"[core_matrix i={i} newton={newton}]", flush=True)
        A_, B_ = coefs_ab(i, coef, lbs, jbs, kmax)
        hd = sym_dhl(g, nuc, int(kb[i - 1]))
        h0, x0 = O.core_matrix(g, i, pq, hd, A_, B_, kmax, qsum, newton=newton)
        hfm, hx = d.core_matrix(i, pqd, fortran_to_device(hd), A_, B_, kmax, qsum, newton=newton)
        assert close(device_to_fortran(hfm), h0)
        if newton:
            assert close(device_to_fortran(hx), x0)
        else:
            assert not np.any(device_to_fortran(hx))


def test_df_integrals_vs_oracle(he):
    """rk_df and its single-type, S^k and hf_fk / hf_gk relatives as one
    request set, against the oracle's densities, convolutions and SUM_AmB in
    the reference's order"""
    from dbsr3_amd.mchf import DfIntegrals
    from test_gpu_engine import oracle_values
    d, g, pq, pqd, kbs, nuc = he
    n = pq.shape[2]
    rng = np.random.default_rng(8)
    names = ["rk_df", "rk_pppp_df", "rk_qqqq_df", "rk_pqpq_df", "rk_qpqp_df", "sk_ppqq_df", "sk_pqqp_df"]
    reqs = []
    for nm in names:
        for _ in range(12):
            a, b, c, e = (int(x) for x in rng.integers(1, n + 1, 4))
            reqs.append((nm, a, b, c, e, int(rng.integers(0, 4))))
    for _ in range(10):
        i, j = (int(x) for x in rng.integers(1, n + 1, 2))
        k = int(rng.integers(0, 4))
        reqs += [("hf_fk", i, j, k), ("hf_gk", i, j, k)]
    got = DfIntegrals(d, n, reqs).values(pq)
    fs = [np.asfortranarray(pq[:, :, j]) for j in range(n)]
    for q, r in enumerate(reqs):
        if r[0] == "hf_fk":
            ref = O.hf_fk(g, fs, r[1] - 1, r[2] - 1, r[3])
        elif r[0] == "hf_gk":
            ref = O.hf_gk(g, fs, r[1] - 1, r[2] - 1, r[3])
        else:
            kind = names.index(r[0])
            ref = oracle_values(g, fs, [(r[5], r[1], r[2], r[3], r[4])], kind)[0]
        assert abs(got[q] - ref) <= 1e-12 * max(abs(ref), 1e-3 * np.max(np.abs(got))), r
