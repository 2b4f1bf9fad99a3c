"""DBSR_HF's orbital update on the GPU (hf_matrix with the Dirac operator,
the orthogonality projections, the boundary conditions and the generalised
eigen-solve on rocSOLVER) for every orbital of the Fe example, against the
oracle's (scipy's LAPACK) on the same matrices, and the SCF fixed point:
the converged Fe.bsw orbitals come back to 2e-7 (the run's orb_tol, see
tests/test_oracle_hf_solve.py).  Tolerance: two eigen-solvers (rocSOLVER's
and LAPACK's dsygvd) agree on an eigenvalue to a few ulps of the pencil's
spectral radius (the Dirac pencil's eigenvalues reach 1.4e8 on the Fe grid,
so 1 ulp of it is 3e-8 au), here asserted at 1e-12 relative or 0.1 ulp of
the spectral radius, whichever is larger; the eigenvector to 4 ulps of the
spectral radius over the eigenvalue's gap to its neighbours (the first-order
bound of the eigenvector's sensitivity, 1e-9 to 3e-8 here), relative to its
largest coefficient."""
import numpy as np
import pytest

import oracle as O
from test_oracle_hf_solve import fe_hf_inputs, ortho_and_m

pytestmark = pytest.mark.gpu


def test_fe_orbital_update_vs_oracle(gpu_available):
    from dbsr3_amd import DBS
    from dbsr3_amd.dbs import fortran_to_device
    from dbsr3_amd.hf_solve import orbital_update
    b, g, pq, kbs, lbs, nuc, qsum, AB, kmax, iprm, nbs = fe_hf_inputs()
    d = DBS(b["t"], b["ns"], b["ks"], b["ksp"], b["ksq"])
    nu = DBS.nucleus(b["nuclear"], b["atomic_number"], b["r_uniform"], b["a_fermi"], b["c_fermi"])
    fppqq_dev = d.bspline_overlaps()[2]
    fppqq = O.overlaps(g)[2]
    pqd = fortran_to_device(pq)
    for i in range(len(kbs)):
        A, B = AB[i]
        lab = b["orbitals"][i]["label"]
        print(f"[hf orbital update {lab}]", flush=True)
        v, e = orbital_update(d, i + 1, pqd, d.dhl_pq(nu, kbs[i]), A, B, kmax, qsum, kbs, lbs, nbs, iprm, fppqq_dev)
        v = v.cpu().numpy()
        h = O.hf_matrix(g, i + 1, pq, A, B, kmax, qsum[i], hfm=np.asfortranarray(qsum[i] * O.dhl_pq(g, *nuc, kbs[i])))
        vs, m = ortho_and_m(i, kbs, lbs, nbs, pq)
        vo, eo = O.hf_eiv(h, vs, iprm[:, i], m, fppqq)
        print(f"  e = {e!r} oracle {eo!r} diff {abs(e - eo) / abs(eo):.1e}; "
              f"|v - vo| {np.max(np.abs(v - vo)) / np.max(np.abs(vo)):.1e}", flush=True)
        import scipy.linalg
        keep = np.nonzero(iprm[:, i])[0]
        ev = scipy.linalg.eigh(h[np.ix_(keep, keep)], fppqq[np.ix_(keep, keep)], lower=True, eigvals_only=True)
        rad = np.max(np.abs(ev))
        gap = np.min(np.abs(ev[np.abs(ev - eo) > 1e-9 * abs(eo)] - eo))
        ulp = np.finfo(float).eps * rad
        assert abs(e - eo) <= max(1e-12 * abs(eo), 0.1 * ulp), lab
        assert np.max(np.abs(v - vo)) <= max(1e-10, 4 * ulp / gap) * np.max(np.abs(vo)), (lab, gap)
        p = np.concatenate([pq[:, 0, i], pq[:, 1, i]])
        assert np.max(np.abs(v - p)) <= 2e-7 * np.max(np.abs(p)), lab
