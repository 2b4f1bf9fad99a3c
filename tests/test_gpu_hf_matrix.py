"""GPU hf_matrix (two-electron part, DBSR_HF/hf_matrix.f90, with the Breit
block of hf_matrix_breit) against the oracle (itself bit-exact to the
reference's compiled hf_matrix), and against the engine's adjoint: with E
the sum of a(i,j,k) F^k(i,j) (j <= i) and b(i,j,k) G^k(i,j) (j > i),
dE/d(P_i,Q_i) = 2 qsum(i) hfm_i (P_i;Q_i) -- two independent GPU paths."""
import numpy as np
import pytest

import oracle as O
from conftest import fe_grid, small_grid
from test_oracle_hf_matrix import KAPPA, coefs_ab, hf_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("breit", [False, True])
@pytest.mark.parametrize("spec", [small_grid, fe_grid])
def test_hf_matrix_vs_oracle(gpu_available, spec, breit):
    from dbsr3_amd import DBS
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    kmax = 3
    g, (t, ns, ks, ksp, ksq), pq, lbs, jbs, coef, qsum, br = hf_case(spec, kmax, 11, breit)
    print(f"[hf_matrix ns={ns} breit={breit}]", flush=True)
    d = DBS(t, ns, ks, ksp, ksq)
    pqd = fortran_to_device(pq)
    for i in range(1, pq.shape[2] + 1):
        A, B = coefs_ab(i, coef, lbs, jbs, kmax)
        got = device_to_fortran(d.hf_matrix(i, pqd, A, B, kmax, qsum[i - 1], breit=br))
        ref = O.hf_matrix(g, i, pq, A, B, kmax, qsum[i - 1], breit=br)
        assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref)), i


@pytest.mark.parametrize("spec", [small_grid, fe_grid])
def test_hf_matrix_is_the_engine_adjoint(gpu_available, spec):
    from dbsr3_amd import DBS, Engine
    from dbsr3_amd.dbs import device_to_fortran, fortran_to_device
    kmax = 4
    g, (t, ns, ks, ksp, ksq), pq, lbs, jbs, coef, qsum, _ = hf_case(spec, kmax, 21, False)
    n = pq.shape[2]
    rows, cs = [], []
    for i in range(1, n + 1):
        A, B = coefs_ab(i, coef, lbs, jbs, kmax)
        for k in range(kmax + 1):
            for j in range(1, i + 1):
                if A[j - 1, k]:
                    rows.append((k, i, j, i, j)); cs.append(A[j - 1, k])
            for j in range(i + 1, n + 1):
                if B[j - 1, k]:
                    rows.append((k, i, j, j, i)); cs.append(B[j - 1, k])
    d = DBS(t, ns, ks, ksp, ksq)
    eng = Engine([(d, n, np.array(rows, dtype=np.int32))])
    eng.set_orbitals(0, pq)
    eng.set_coef(0, np.array(cs))
    eng.forward_adjoint()
    grad = eng.grad()
    pqd = fortran_to_device(pq)
    scale = np.max(np.abs(grad))
    for i in range(1, n + 1):
        A, B = coefs_ab(i, coef, lbs, jbs, kmax)
        h = device_to_fortran(d.hf_matrix(i, pqd, A, B, kmax, qsum[i - 1]))
        v = np.concatenate([pq[:, 0, i - 1], pq[:, 1, i - 1]])
        pred = 2 * qsum[i - 1] * (h @ v)
        eg = np.concatenate([grad[:, 0, i - 1], grad[:, 1, i - 1]])
        # the engine's gradient covers the B-spline coefficients of P (1..nsp)
        # and Q (1..nsq); hfm rows beyond them are zero
        assert np.max(np.abs(pred - eg)) <= 1e-11 * scale, i
