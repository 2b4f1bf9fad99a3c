"""Python wrapper of the batched integral engine (dbs_engine_* in the C-ABI).

An instance is (grid context, orbitals, rows (k,a,b,c,d) with 1-based orbital
indices).  Row n evaluates R^k(a,b;c,d) as the reference's rk() does
(DBSR_HF/dbsr_lib_dbs.f90:3981-4022), or a single P/Q type, or an S^k type.
The adjoint returns E = sum_n coef_n*R_n and dE/d(pq), the orbital gradient
that hf_matrix's direct+exchange potentials produce (DBSR_HF/hf_matrix.f90).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import check, lib

RK_TOTAL = 0


class Engine:
    def __init__(self, instances, kind=RK_TOTAL):
        """instances: list of (DBS, norb, rows) with rows an (n,5) int array."""
        self._dbs = [inst[0] for inst in instances]
        self.norb = [int(inst[1]) for inst in instances]
        rows = [np.ascontiguousarray(np.asarray(inst[2], dtype=np.int32).reshape(-1, 5))
                for inst in instances]
        self.nrows = [r.shape[0] for r in rows]
        allrows = np.ascontiguousarray(np.concatenate(rows, axis=0) if rows else np.zeros((0, 5), np.int32))
        n = len(instances)
        ctxs = (C.c_void_p * n)(*[d.ctx.value for d in self._dbs])
        norb = np.ascontiguousarray(self.norb, dtype=np.int32)
        nrows = np.ascontiguousarray(self.nrows, dtype=np.int32)
        ip = C.POINTER(C.c_int)
        h = C.c_void_p()
        check(lib().dbs_engine_create(C.byref(h), n, ctxs, norb.ctypes.data_as(ip),
                                      nrows.ctypes.data_as(ip), allrows.ctypes.data_as(ip), kind))
        self.h = h
        self.kind = kind
        self.ninst = n

    def close(self):
        if getattr(self, "h", None):
            lib().dbs_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- buffers ------------------------------------------------------------
    def buffers(self, inst):
        ptrs = [C.c_void_p() for _ in range(5)]
        check(lib().dbs_engine_buffers(self.h, inst, *[C.byref(p) for p in ptrs]))
        return [p.value for p in ptrs]  # pq, coef, values, grad, energy

    def _ctx(self, inst):
        return self._dbs[inst].ctx

    def set_orbitals(self, inst, pq):
        """pq: numpy array (ns,2,norb) (any order; stored Fortran)."""
        d = self._dbs[inst]
        a = np.asfortranarray(np.asarray(pq, dtype=np.float64))
        assert a.shape == (d.ns, 2, self.norb[inst]), a.shape
        p = self.buffers(inst)[0]
        check(lib().dbs_memcpy_h2d(self._ctx(inst), C.c_void_p(p), a.ctypes.data_as(C.c_void_p), a.nbytes))

    def set_coef(self, inst, coef):
        a = np.ascontiguousarray(np.asarray(coef, dtype=np.float64))
        assert a.shape == (self.nrows[inst],)
        p = self.buffers(inst)[1]
        check(lib().dbs_memcpy_h2d(self._ctx(inst), C.c_void_p(p), a.ctypes.data_as(C.c_void_p), a.nbytes))

    def _get(self, inst, which, n):
        out = np.zeros(n)
        p = self.buffers(inst)[which]
        self.sync()
        check(lib().dbs_memcpy_d2h(self._ctx(inst), out.ctypes.data_as(C.c_void_p), C.c_void_p(p), out.nbytes))
        return out

    def values(self, inst=0):
        return self._get(inst, 2, self.nrows[inst])

    def energy(self, inst=0):
        return float(self._get(inst, 4, 1)[0])

    def grad(self, inst=0):
        d = self._dbs[inst]
        g = self._get(inst, 3, d.ns * 2 * self.norb[inst])
        return g.reshape((d.ns, 2, self.norb[inst]), order="F")

    # -- execution ----------------------------------------------------------
    def forward(self):
        check(lib().dbs_engine_forward(self.h))

    def forward_adjoint(self):
        check(lib().dbs_engine_forward_adjoint(self.h))

    def forward_plain(self):
        check(lib().dbs_engine_forward_plain(self.h))

    def sync(self):
        check(lib().dbs_engine_sync(self.h))

    def graph_capture(self, with_adjoint=True):
        check(lib().dbs_engine_graph_capture(self.h, int(with_adjoint)))

    def graph_launch(self):
        check(lib().dbs_engine_graph_launch(self.h))

    def set_tuning(self, params):
        p = np.ascontiguousarray(params, dtype=np.int32)
        check(lib().dbs_engine_set_tuning(self.h, len(p), p.ctypes.data_as(C.POINTER(C.c_int))))

    def set_memory_budget(self, nbytes):
        """bytes of features + gradients held at once (0: no budget); the
        groups then run in batches that fit"""
        check(lib().dbs_engine_set_memory_budget(self.h, int(nbytes)))

    def memory(self):
        """(batches, bytes of features + gradients held)"""
        nb = C.c_int()
        by = C.c_size_t()
        check(lib().dbs_engine_memory(self.h, C.byref(nb), C.byref(by)))
        return nb.value, by.value

    def get_tuning(self):
        p = np.zeros(7, dtype=np.int32)
        check(lib().dbs_engine_get_tuning(self.h, 7, p.ctypes.data_as(C.POINTER(C.c_int))))
        return [int(x) for x in p]

    def autotune(self, with_adjoint=True, reps=3):
        check(lib().dbs_engine_autotune(self.h, int(with_adjoint), reps))
        return self.get_tuning()

    def work(self, with_adjoint=True):
        f = C.c_double()
        b = C.c_double()
        check(lib().dbs_engine_work(self.h, int(with_adjoint), C.byref(f), C.byref(b)))
        return f.value, b.value

    def profile(self, with_adjoint=True, reps=5):
        """per-phase device time (ms) and algorithmic work of one launch:
        {name: (ms, flops, bytes)} in launch order"""
        nmax = 32
        names = C.create_string_buffer(32 * nmax)
        ms, fl, by = np.zeros(nmax), np.zeros(nmax), np.zeros(nmax)
        n = C.c_int()
        dp = C.POINTER(C.c_double)
        check(lib().dbs_engine_profile(self.h, int(with_adjoint), reps, nmax, names,
                                       ms.ctypes.data_as(dp), fl.ctypes.data_as(dp),
                                       by.ctypes.data_as(dp), C.byref(n)))
        raw = names.raw
        out = {}
        for i in range(n.value):
            nm = raw[32 * i:32 * (i + 1)].split(b"\0")[0].decode()
            out[nm] = (float(ms[i]), float(fl[i]), float(by[i]))
        return out
