"""Coefficient lists and Add_integral, the host half of the accumulation path.

``CData`` mirrors module c_data (DBSR_MAT3/MOD_c_data.f90:2-368): Add_coef,
Add_matrix and the Gen_matrix hook, which receives each merged list in IPT
order.  The lists are integer host work inside libdbsr3_amd.so; evaluation
and accumulation of the listed integrals run on the GPU
(``rk_evaluate`` / ``r_data_34`` below).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import GEN_MATRIX_FN, Mat3OrbsC, check, lib

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


class CData:
    def __init__(self, ntype, kpol1, kpol2, mblock, nblock, kblock, eps_c, gen):
        """gen(itype, kpol, (c, K1, K2, K3, K4)) is the Gen_matrix hook."""

        def _cb(user, itype, kpol, nc, c, k1, k2, k3, k4):
            if nc > 0:
                arrs = (np.ctypeslib.as_array(c, (nc,)).copy(),) + tuple(
                    np.ctypeslib.as_array(p, (nc,)).copy() for p in (k1, k2, k3, k4))
            else:
                arrs = (np.zeros(0),) + tuple(np.zeros(0, np.int32) for _ in range(4))
            gen(itype, kpol, arrs)

        self._cb = GEN_MATRIX_FN(_cb)
        h = C.c_void_p()
        check(lib().dbs_cdata_create(C.byref(h), ntype, kpol1, kpol2, mblock, nblock, kblock,
                                     eps_c, self._cb, None))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().dbs_cdata_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_coef(self, c, kpol, j1, j2, j3, j4, itype):
        check(lib().dbs_cdata_add_coef(self.h, c, kpol, j1, j2, j3, j4, itype))

    def add_matrix(self, itype, kpol):
        check(lib().dbs_cdata_add_matrix(self.h, itype, kpol))

    def add_integral_ci(self, kpol, ic, jc, i1, i2, i3, i4, c, kappa, ibi=2 ** 15, mbreit=0):
        k = _i32(kappa)
        check(lib().dbs_add_integral_ci(self.h, kpol, ic, jc, i1, i2, i3, i4, c,
                                        k.ctypes.data_as(_ip), ibi, mbreit))

    def add_integral_mat3(self, icase, kpol, i1, i2, i3, i4, c, ic, jc, io, jo, lbs, kbs, ipbs,
                          ibo=2 ** 15, mk=20, eps_c=1e-10, check_target=0):
        """MAT3 Add_integral (icase 0 overlaps, 2 R, 3 S); lbs/kbs/ipbs hold
        orbital o at element o-1"""
        a, b, d = _i32(lbs), _i32(kbs), _i32(ipbs)
        o = Mat3OrbsC(len(d), a.ctypes.data_as(_ip), b.ctypes.data_as(_ip), d.ctypes.data_as(_ip), ibo, mk,
                      eps_c, check_target)
        check(lib().dbs_add_integral_mat3(self.h, C.byref(o), icase, kpol, i1, i2, i3, i4, c, ic, jc, io, jo))

    def add_integral_mat(self, kpol, i1, i2, i3, i4, c, lorb, ipbs, ibo=2 ** 15, mk=20, eps_c=1e-10):
        lo, ipb = _i32(lorb), _i32(ipbs)
        check(lib().dbs_add_integral_mat(self.h, kpol, i1, i2, i3, i4, c, lo.ctypes.data_as(_ip),
                                         ipb.ctypes.data_as(_ip), ibo, mk, eps_c))


def decode_int(codes):
    codes = _i32(codes)
    n = len(codes)
    outs = [np.zeros(n, np.int32) for _ in range(6)]
    check(lib().dbs_decode_int(n, codes.ctypes.data_as(_ip), *[o.ctypes.data_as(_ip) for o in outs]))
    return tuple(outs)  # met, k, i1, i2, i3, i4


def smu(ka, kb, kc, kd, L, v):
    S = np.zeros(8)
    s = C.c_double()
    check(lib().dbs_smu(ka, kb, kc, kd, L, v, S.ctypes.data_as(_dp), C.byref(s)))
    return s.value, S


def rk_evaluate(dbs, itype, k, lst, pq_dev, norb, HM_dev, ldhm, nzero, DM_dev, eps_c=1e-10, ibi=2 ** 15,
                sync=True):
    """Rk_evaluate/Sk_evaluate of one merged list (c, K1..K4) on the GPU.
    pq_dev, HM_dev, DM_dev: torch float64 device tensors (Fortran layouts).
    sync=False queues the list on the context's stream and returns at once
    (the next list's tables are built while the device runs this one); then
    nothing may touch HM/DM from another stream before dbs.sync() (or a
    device-wide synchronize)."""
    import torch
    c, K1, K2, K3, K4 = lst
    c = np.ascontiguousarray(c, dtype=np.float64)
    K = [_i32(x) for x in (K1, K2, K3, K4)]
    if sync:
        torch.cuda.synchronize()
    dm = C.c_void_p(DM_dev.data_ptr()) if DM_dev is not None else None
    check(lib().dbs_rk_evaluate_dev(dbs.ctx, itype, k, len(c), c.ctypes.data_as(_dp),
                                    *[x.ctypes.data_as(_ip) for x in K], eps_c, ibi,
                                    C.c_void_p(pq_dev.data_ptr()), norb,
                                    C.c_void_p(HM_dev.data_ptr()), ldhm, nzero, dm))
    if sync:
        check(lib().dbs_ctx_sync(dbs.ctx))


def r_data_34(dbs, itype, atype, k, lst, pq_dev, norb, icc, hch_dev, nblocks, ibo=2 ** 15):
    """R_data itype 3/4 for one atype on the GPU; hch_dev holds (ms,ms,nblocks)."""
    import torch
    c, K1, K2, K3 = lst[:4]
    c = np.ascontiguousarray(c, dtype=np.float64)
    K = [_i32(x) for x in (K1, K2, K3)]
    iccf = np.asfortranarray(np.asarray(icc, dtype=np.int32))
    torch.cuda.synchronize()
    check(lib().dbs_rdata_34_dev(dbs.ctx, itype, atype, k, len(c), c.ctypes.data_as(_dp),
                                 *[x.ctypes.data_as(_ip) for x in K], ibo,
                                 C.c_void_p(pq_dev.data_ptr()), norb, iccf.ctypes.data_as(_ip),
                                 iccf.shape[0], C.c_void_p(hch_dev.data_ptr()), nblocks))
    check(lib().dbs_ctx_sync(dbs.ctx))
