"""Pins the oracle's hf_matrix restatement (two-electron part of
DBSR_HF/hf_matrix.f90, with and without the Breit block of
hf_matrix_breit) bit for bit to the reference's own compiled hf_matrix
(oracle/_ref), with the one-electron operator set to zero there."""
import numpy as np
import pytest

import oracle as O
from conftest import fe_grid, random_orbitals, small_grid

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")

KAPPA = np.array([-1, 1, -2, 2, -3, -1], dtype=np.int32)


def hf_case(spec, kmax, seed, breit):
    """orbitals, coef(i,j,k) and an rk4 Breit list; returns what the
    reference's hf_matrix and the oracle need"""
    t, ns, ks, ksp, ksq = spec()
    g = O.Grid(t, ns, ks, ksp, ksq)
    n = len(KAPPA)
    fs = random_orbitals(g, n, seed)
    pq = np.asfortranarray(np.stack(fs, axis=2))
    lbs = np.array([k if k > 0 else -k - 1 for k in KAPPA], dtype=np.int32)
    jbs = np.array([2 * abs(k) - 1 for k in KAPPA], dtype=np.int32)
    rng = np.random.default_rng(seed)
    coef = np.asfortranarray(rng.uniform(-1, 1, (n, n, kmax + 1)))
    qsum = rng.uniform(0.5, 2.0, n)
    if breit:
        nrk = 60
        kr1 = rng.integers(0, kmax + 1, nrk) + 1000 * rng.integers(0, 2, nrk)
        kr2 = rng.integers(1, 5, nrk)
        kr3, kr4 = rng.integers(1, n + 1, nrk), rng.integers(1, n + 1, nrk)
        crk = rng.uniform(-1, 1, nrk)
        br = (kr1, kr2, kr3, kr4, crk)
    else:
        br = None
    return g, (t, ns, ks, ksp, ksq), pq, lbs, jbs, coef, qsum, br


def coefs_ab(i, coef, lbs, jbs, kmax):
    """a(i,j,k), b(i,j,k) by the selection rules of DBSR_HF/hf_MOD_energy.f90:20-57"""
    n = coef.shape[0]
    A = np.zeros((n, kmax + 1), order="F")
    B = np.zeros((n, kmax + 1), order="F")
    for j in range(1, n + 1):
        for k in range(kmax + 1):
            if k % 2 == 0 and k <= (jbs[i - 1] + jbs[j - 1]) // 2:
                A[j - 1, k] = coef[max(i, j) - 1, min(i, j) - 1, k]
            if (i != j and (k + lbs[i - 1] + lbs[j - 1]) % 2 == 0 and k >= abs(jbs[i - 1] - jbs[j - 1]) // 2
                    and k <= (jbs[i - 1] + jbs[j - 1]) // 2):
                B[j - 1, k] = coef[min(i, j) - 1, max(i, j) - 1, k]
    return A, B


@pytest.mark.parametrize("breit", [False, True])
@pytest.mark.parametrize("spec", [small_grid, fe_grid])
def test_hf_matrix_matches_reference(spec, breit):
    kmax = 3
    g, grid, pq, lbs, jbs, coef, qsum, br = hf_case(spec, kmax, 11, breit)
    t, ns, ks, ksp, ksq = grid
    R = O.ref()
    R.ref_setup(ns, ks, ksp, ksq, O._d(O.f64(t)))
    n = pq.shape[2]
    ci = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    if br is None:
        z = np.zeros(1, np.int32)
        lst = (0, z, z, z, z, np.zeros(1))
    else:
        lst = (len(br[4]), ci(br[0]), ci(br[1]), ci(br[2]), ci(br[3]), np.ascontiguousarray(br[4]))
    R.ref_hf_setup(n, kmax, 2 if breit else 0, O._i(ci(KAPPA)), O._i(ci(lbs)), O._i(ci(jbs)), O._d(qsum),
                   O._d(coef), O._d(pq), lst[0], O._i(lst[1]), O._i(lst[2]), O._i(lst[3]), O._i(lst[4]),
                   O._d(lst[5]))
    for i in range(1, n + 1):
        ref = np.zeros((g.ms, g.ms), order="F")
        R.ref_hf_matrix(i, O._d(ref))
        A, B = coefs_ab(i, coef, lbs, jbs, kmax)
        mine = O.hf_matrix(g, i, pq, A, B, kmax, qsum[i - 1], breit=br)
        assert np.array_equal(mine, ref), i
        assert np.any(mine)
