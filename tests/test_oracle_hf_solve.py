"""DBSR_HF's orbital update (hf_eiv, DBSR_HF/hf_solve_HF.f90:118-202) on the
converged Fe orbitals, on the CPU.

The orbitals of Fe.bsw are the fixed point of the SCF iterations: rebuilding
each orbital's HF matrix from them (Dirac operator with the Fermi nucleus
times qsum, plus hf_matrix's direct and exchange potentials with DBSR_HF's
a(i,j,k), b(i,j,k)), projecting out the lower orbitals of the same kappa,
removing the boundary B-splines and solving H x = e B x must give back, for
the m-th positive-energy eigenvalue (m = n - l - the projections), the
orbital itself and its energy e(nl) of Fe.log:396-405 (also stored in
Fe.bsw).  Tolerances follow the run's convergence (orb_tol = 1e-7,
Fe.inp:19; the last orbital differences Fe.log:396-405): the orbitals to
2e-7 of their largest coefficient.  e(nl) is the value of the iteration in
which the orbital last moved (1s at iteration 10, Fe.log:264, while 4s
still moved by 5e-5), so it matches to 2e-6 relative, and to 1e-7 for 4s,
solved in the last iteration.  The eigenvector is also B-normalised and
B-orthogonal to the projected orbitals (1e-12; 1e-9, the eigen-solver's accuracy against the matrix norm of order c^2).
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_fe import ENL_LOG, fe_case


def fe_hf_inputs():
    from dbsr3_amd.hf_energy import a_coef, av_energy_coef, b_coef
    b, pq, kbs, lbs, jbs, nuc, labels, iq, w = fe_case()
    g = O.Grid(b["t"], b["ns"], b["ks"], b["ksp"], b["ksq"])
    n = len(kbs)
    wn = np.asarray(w) / np.sum(w)
    qsum = np.asarray(iq).T @ wn
    coef = av_energy_coef(jbs, iq, wn)
    kmax = coef.shape[2] - 1
    AB = []
    for i in range(n):
        A = np.zeros((n, kmax + 1), order="F")
        B = np.zeros((n, kmax + 1), order="F")
        for j in range(n):
            for k in range(kmax + 1):
                A[j, k] = a_coef(coef, jbs, i, j, k)
                B[j, k] = b_coef(coef, lbs, jbs, i, j, k)
        AB.append((A, B))
    iprm = O.boundary_conditions(g.ns, g.nsp, g.nsq, g.ksp, g.ksq, lbs, kbs, b["nuclear"])
    nbs = [o["n"] for o in b["orbitals"]]
    return b, g, pq, kbs, lbs, nuc, qsum, AB, kmax, iprm, nbs


def ortho_and_m(i, kbs, lbs, nbs, pq):
    """the lower orbitals of the same kappa (e(i,j) = 1e-5 for j < i,
    hf_get_case.f90:41-47) and m = n - l - their number"""
    vs = [np.concatenate([pq[:, 0, j], pq[:, 1, j]]) for j in range(i) if kbs[j] == kbs[i]]
    return vs, nbs[i] - lbs[i] - len(vs)


@pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")
def test_boundary_conditions_against_reference_rule():
    """the kept B-splines: P from l+2 (Fermi nucleus, ilzero = 1) to
    nsp-2, Q from l+3 (kappa < 0) or l+1 (kappa > 0) to nsq-2"""
    b, g, pq, kbs, lbs, nuc, qsum, AB, kmax, iprm, nbs = fe_hf_inputs()
    for i in range(len(kbs)):
        keep = np.nonzero(iprm[:, i])[0]
        jp = lbs[i] + 1 if lbs[i] + 1 <= g.ksp - 1 else 1
        jq = (lbs[i] + 2 if kbs[i] < 0 else lbs[i])
        jq = jq if jq <= g.ksq - 1 else 1
        assert keep[0] == jp and keep[keep < g.ns][-1] == g.nsp - 3
        assert keep[keep >= g.ns][0] == g.ns + jq and keep[-1] == g.ns + g.nsq - 3
        # the converged orbital vanishes on the removed B-splines
        v = np.concatenate([pq[:, 0, i], pq[:, 1, i]])
        assert np.all(v[iprm[:, i] == 0] == 0.0)


def test_fe_orbitals_are_the_fixed_point():
    b, g, pq, kbs, lbs, nuc, qsum, AB, kmax, iprm, nbs = fe_hf_inputs()
    fppqq = O.overlaps(g)[2]
    for i in range(len(kbs)):
        A, B = AB[i]
        dhl = O.dhl_pq(g, *nuc, kbs[i])
        h = O.hf_matrix(g, i + 1, pq, A, B, kmax, qsum[i], hfm=np.asfortranarray(qsum[i] * dhl))
        vs, m = ortho_and_m(i, kbs, lbs, nbs, pq)
        v, e = O.hf_eiv(h, vs, iprm[:, i], m, fppqq)
        lab = b["orbitals"][i]["label"]
        p = np.concatenate([pq[:, 0, i], pq[:, 1, i]])
        print(f"  {lab:4s} m={m} e={e:+.10f} Fe.log {ENL_LOG[lab]:+.8f} bsw {b['orbitals'][i]['e']:+.12f}"
              f" |v-p|/|p| {np.max(np.abs(v - p)) / np.max(np.abs(p)):.1e}")
        assert abs(e - b["orbitals"][i]["e"]) <= (1e-7 if lab == "4s" else 2e-6) * abs(e), lab
        assert np.max(np.abs(v - p)) <= 2e-7 * np.max(np.abs(p)), lab
        assert abs(v @ fppqq @ v - 1.0) <= 1e-12
        for u in vs:
            assert abs(u @ fppqq @ v) <= 1e-9


def test_product_boundary_conditions_equal_oracle():
    from dbsr3_amd.hf_solve import boundary_conditions
    b, g, pq, kbs, lbs, nuc, qsum, AB, kmax, iprm, nbs = fe_hf_inputs()
    for nuclear, ilz in (("Fermi", 1), ("point", 1), ("Fermi", 0)):
        a = boundary_conditions(g.ns, g.nsp, g.nsq, g.ksp, g.ksq, lbs, kbs, nuclear, ilz, 2)
        o = O.boundary_conditions(g.ns, g.nsp, g.nsq, g.ksp, g.ksq, lbs, kbs, nuclear, ilz, 2)
        assert np.array_equal(a, o)
