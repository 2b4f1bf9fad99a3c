"""DBSR_MCHF's integral functions and orbital matrices on the GPU.

``DfIntegrals`` evaluates a fixed set of integral requests by the reference's
function names, as one batch on the engine per row kind:

    rk_df(i1,j1,i2,j2,k)        R^k(i1,j1;i2,j2), the four P/Q terms
                                (DBSR_MCHF/rk_df.f90:2-51)
    rk_pppp_df / rk_qqqq_df /   one term each (rk_df.f90:55-140)
    rk_pqpq_df / rk_qpqp_df
    sk_ppqq_df / sk_pqqp_df     S^k (rk_df.f90:144-190)
    hf_fk(i,j,k)                R^k(i,j;i,j) (DBSR_HF/hf_rk.f90:43-80)
    hf_gk(i,j,k)                pppp + qqqq + 2 qpqp of R^k(i,j;j,i)
                                (DBSR_HF/hf_rk.f90:84-113)

rk_df caches the densities of its last left and right pair and the four
convolutions of its last (left pair, k) across calls (rk_df.f90:16-43); the
engine shares every convolution of a distinct (k, left pair) across the
whole request set, so the set is evaluated at once.  The request set is fixed
when the object is made (the integral list of an SCF run does not change
between iterations); ``values(pq)`` re-evaluates it for new orbitals.

``mchf_matrix`` / ``core_matrix`` wrap the device entry points
(DBSR_MCHF/mchf_matrix.f90:2-345) with the reference's dispatch: core
orbitals (i <= ncore) go to core_matrix.
"""
from __future__ import annotations

import numpy as np

# name -> (engine kind, whether the request is (i,j,k) and its rows)
_KIND = {"rk_df": 0, "rk_pppp_df": 1, "rk_qqqq_df": 2, "rk_pqpq_df": 3, "rk_qpqp_df": 4, "sk_ppqq_df": 5,
         "sk_pqqp_df": 6}


class DfIntegrals:
    def __init__(self, dbs, norb, requests):
        """requests: list of (name, i1, j1, i2, j2, k) for the rk_/sk_ names,
        (name, i, j, k) for hf_fk / hf_gk; 1-based orbitals"""
        from .engine import Engine
        self.dbs = dbs
        self.norb = int(norb)
        rows = {kind: [] for kind in range(7)}
        # per request: list of (kind, row index, factor), summed in order
        self.plan = []
        for req in requests:
            name = req[0]
            if name == "hf_fk":
                i, j, k = req[1:]
                self.plan.append([(0, self._row(rows, 0, (k, i, j, i, j)), 1.0)])
            elif name == "hf_gk":
                i, j, k = req[1:]
                r = (k, i, j, j, i)
                self.plan.append([(1, self._row(rows, 1, r), 1.0), (2, self._row(rows, 2, r), 1.0),
                                  (4, self._row(rows, 4, r), 2.0)])
            elif name in _KIND:
                i1, j1, i2, j2, k = req[1:]
                if k < 0:
                    raise ValueError(f"{name}: k < 0")
                kind = _KIND[name]
                self.plan.append([(kind, self._row(rows, kind, (k, i1, j1, i2, j2)), 1.0)])
            else:
                raise ValueError(f"unknown integral function {name!r}")
        for r in rows.values():
            for (k, a, b, c, d) in r:
                if min(a, b, c, d) < 1 or max(a, b, c, d) > self.norb or k < 0:
                    raise ValueError(f"orbital index or k out of range in row {(k, a, b, c, d)}")
        self.engines = {kind: Engine([(dbs, self.norb, np.array(r, dtype=np.int32))], kind=kind)
                        for kind, r in rows.items() if r}
        self._rowmap = rows

    @staticmethod
    def _row(rows, kind, r):
        rows[kind].append(tuple(int(x) for x in r))
        return len(rows[kind]) - 1

    def values(self, pq):
        """the requests' values for orbitals pq (ns,2,norb) (host array)"""
        vals = {}
        for kind, eng in self.engines.items():
            eng.set_orbitals(0, pq)
            eng.forward()
        for kind, eng in self.engines.items():
            vals[kind] = eng.values()
        out = np.zeros(len(self.plan))
        for n, terms in enumerate(self.plan):
            s = 0.0
            for kind, idx, f in terms:
                s = s + f * vals[kind][idx]
            out[n] = s
        return out


def mchf_matrix(dbs, i, ncore, pq_dev, dhl_dev, llist, rlist, kmax, core=None, **kw):
    """mchf_matrix(i): core orbitals (i <= ncore) take core_matrix with core =
    dict(acoef=, bcoef=, qsum=) and dhl_dev as Get_hd(kbs(i)); the rest the
    full matrix.  Returns (hfm, rhs, hx, e, qsum) (rhs, e and qsum None for a
    core orbital)."""
    if i <= ncore:
        if core is None:
            raise ValueError("mchf_matrix: core orbital without core coefficients")
        hfm, hx = dbs.core_matrix(i, pq_dev, dhl_dev, core["acoef"], core["bcoef"], kmax, core["qsum"],
                                  newton=kw.get("newton", 0))
        return hfm, None, hx, None, None
    return dbs.mchf_matrix(i, pq_dev, dhl_dev, llist, rlist, kmax, **kw)
