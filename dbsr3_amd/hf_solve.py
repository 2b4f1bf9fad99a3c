"""The orbital update of DBSR_HF's SCF loop (solve_HF, DBSR_HF/hf_solve_HF.f90)
with the matrices and the eigen-solve on the GPU.

For orbital i: hfm = qsum(i) dhl(kappa_i) + the direct and exchange
potentials (hf_matrix, DBSR_HF/hf_matrix.f90:2-78; dbs_hf_matrix_dev), the
orthogonality projections to the lower orbitals of the same kappa (the
Lagrange-multiplier flags e(i,j) = 1e-5 for j < i, DBSR_HF/hf_get_case.f90:41-47),
the boundary B-splines removed (Boundary_conditions,
DBSR_HF/hf_boundary.f90:2-32) and the generalised eigenproblem
(hf_eiv, :118-202; dbs_hf_eiv_dev on rocSOLVER).
"""
from __future__ import annotations

import numpy as np


def boundary_conditions(ns, nsp, nsq, ksp, ksq, lbs, kbs, nuclear="Fermi", ilzero=1, ibzero=2):
    """iprm(ms, nbf): 0 where a B-spline is removed (zero conditions at r=0
    and at the boundary)"""
    nbf, ms = len(lbs), 2 * ns
    iprm = np.ones((ms, nbf), dtype=np.int32)
    for i in range(nbf):
        if nuclear == "point" or ilzero == 0:
            iprm[0, i] = 0
            iprm[nsp - ibzero:ns, i] = 0
            iprm[ns, i] = 0
            iprm[nsq - ibzero + ns:ms, i] = 0
            continue
        j = lbs[i] + 1
        if j > ksp - 1:
            j = 1
        iprm[:j, i] = 0
        iprm[nsp - ibzero:ns, i] = 0
        j = lbs[i] + 2 if kbs[i] < 0 else lbs[i]
        if j > ksq - 1:
            j = 1
        iprm[ns:ns + j, i] = 0
        iprm[nsq - ibzero + ns:ms, i] = 0
    return iprm


def orbital_update(dbs, i, pq_dev, dhl_dev, acoef, bcoef, kmax, qsum, kbs, lbs, nbs, iprm, fppqq_dev):
    """(v, e) of orbital i (1-based) for orbitals pq_dev (device tensor
    holding pq(ns,2,nbf)): hf_matrix with the one-electron operator, the
    projections to the lower orbitals of the same kappa, hf_eiv"""
    hfm = (qsum[i - 1] * dhl_dev).contiguous()
    dbs.hf_matrix(i, pq_dev, acoef, bcoef, kmax, qsum[i - 1], hfm=hfm)
    lower = [j for j in range(1, i) if kbs[j - 1] == kbs[i - 1]]
    vorth = [pq_dev[j - 1].reshape(-1) for j in lower]
    m = nbs[i - 1] - lbs[i - 1] - len(lower)
    return dbs.hf_eiv(hfm, vorth, iprm[:, i - 1], m, fppqq_dev)
