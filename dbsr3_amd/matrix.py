"""MAT3 matrix storage and its updates on the GPU.

Mirrors module dbsr_mat's channel-block storage (DBSR_MAT3/mod_dbsr_mat.f90:
134-147) as allocated by alloc_dbsr_matrix for one process
(DBSR_MAT3/updates.f90:2-75): diagonal blocks icc(ich,ich) = 1..nch, then
the lower pairs icc(ich,jch) = icc(jch,ich) for ich > jch; icb(ich,ip) in
column order of (ich, ip); ibb(ic,jc) = ibb(jc,ic) for ic >= jc.  The matrices
hch(ms,ms,iicc), acf(iicc,0:mk), hcp(ms,iicb), hp(iibb) live on the device as
torch float64 tensors holding the Fortran layouts.  The routines keep the
reference's names, argument meaning and Stop messages (raised as DBSError).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import KBORT_FN, QUADR_PQ_FN, MatBlocksC, Mat3AuxC, QpqTableC, check, lib

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class MatBlocks:
    """Channel-block, channel-bound and bound-bound matrices of one MAT3 run."""

    def __init__(self, dbs, nch, npert, mk=20, idiag=0):
        import torch
        self.dbs, self.nch, self.npert, self.mk = dbs, nch, npert, mk
        ms = dbs.ms
        icc = np.zeros((nch, nch), dtype=np.int32, order="F")
        i = 0
        if idiag >= 0:
            for ich in range(nch):
                i += 1
                icc[ich, ich] = i
        if idiag <= 0:
            for ich in range(1, nch):
                for jch in range(ich):
                    i += 1
                    icc[ich, jch] = icc[jch, ich] = i
        self.icc, self.nblocks = icc, int(icc.max()) if icc.size else 0
        icb = np.zeros((nch, npert), dtype=np.int32, order="F")
        ibb = np.zeros((npert, npert), dtype=np.int32, order="F")
        if npert > 0 and idiag != 1:
            i = 0
            for ich in range(nch):
                for ip in range(npert):
                    i += 1
                    icb[ich, ip] = i
            i = 0
            for ic in range(npert):
                for jc in range(ic + 1):
                    i += 1
                    ibb[ic, jc] = ibb[jc, ic] = i
        self.icb, self.ibb = icb, ibb
        self.ncb, self.nbb = int(icb.max()) if icb.size else 0, int(ibb.max()) if ibb.size else 0
        dev = "cuda"
        z = lambda *shape: torch.zeros(shape, dtype=torch.float64, device=dev)
        # tensors hold the Fortran layouts: hch(ms,ms,nb) is (nb, ms, ms)
        self.hch = z(max(self.nblocks, 1), ms, ms)
        self.acf = z(mk + 1, max(self.nblocks, 1))
        self.hcp = z(max(self.ncb, 1), ms)
        self.hp = z(max(self.nbb, 1))
        self._desc = MatBlocksC(nch, npert, mk, self.icc.ctypes.data_as(_ip), self.nblocks, _vp(self.hch),
                                _vp(self.acf), self.icb.ctypes.data_as(_ip), self.ncb, _vp(self.hcp),
                                self.ibb.ctypes.data_as(_ip), self.nbb, _vp(self.hp))

    def _pre(self):
        import torch
        torch.cuda.synchronize()

    def _post(self):
        check(lib().dbs_ctx_sync(self.dbs.ctx))

    def _call(self, name, *args):
        self._pre()
        check(getattr(lib(), name)(self.dbs.ctx, C.byref(self._desc), *args))
        self._post()

    # -- updates.f90 -------------------------------------------------------
    def UPDATE_HX(self, ich, jch, d, C_):
        self._call("dbs_update_hx_dev", ich, jch, _vp(d), C_)

    def UPDATE_HW(self, ich, jch, v, w, C_):
        self._call("dbs_update_hw_dev", ich, jch, _vp(v), _vp(w), C_)

    def UPDATE_HV(self, ich, ip, v, C_):
        self._call("dbs_update_hv_dev", ich, ip, _vp(v), C_)

    def UPDATE_HB(self, ic, jc, c):
        """list form: UPDATE_HB(ic(n), jc(n), c(n)) applied in order"""
        ic, jc = _i32(np.atleast_1d(ic)), _i32(np.atleast_1d(jc))
        c = np.ascontiguousarray(np.atleast_1d(c), dtype=np.float64)
        self._call("dbs_update_hb_dev", len(c), ic.ctypes.data_as(_ip), jc.ctypes.data_as(_ip),
                   c.ctypes.data_as(_dp))

    def UPDATE_ACF(self, k, ich, jch, c):
        """list form: UPDATE_ACF(k(n), ich(n), jch(n), C(n)) applied in order"""
        k, ich, jch = (_i32(np.atleast_1d(x)) for x in (k, ich, jch))
        c = np.ascontiguousarray(np.atleast_1d(c), dtype=np.float64)
        self._call("dbs_update_acf_dev", len(c), k.ctypes.data_as(_ip), ich.ctypes.data_as(_ip),
                   jch.ctypes.data_as(_ip), c.ctypes.data_as(_dp))

    # -- O_data.f90 --------------------------------------------------------
    def O_data(self, itype, lst, eps_c=1e-10, ibo=2 ** 15, ipbs=None, kbort=None, vch=None,
               fppqq=None, iitar=0, iptar=None, otarg=None):
        """one merged overlap list lst = (c, k1, k2) in IPT order.
        ipbs: 1-based orbital -> channel table (numpy, length >= max orbital);
        kbort(i1, i2) -> 1-based column of vch (device tensor (nvch, ms));
        otarg: numpy array of the packed target overlaps, updated in place."""
        c, k1, k2 = lst
        c = np.ascontiguousarray(c, dtype=np.float64)
        k1, k2 = _i32(k1), _i32(k2)
        ipbs_a = _i32(ipbs) if ipbs is not None else None
        iptar_a = _i32(iptar) if iptar is not None else None
        cb = KBORT_FN(lambda user, i1, i2: int(kbort(i1, i2))) if kbort is not None else KBORT_FN()
        nvch = int(vch.shape[0]) if vch is not None else 0
        if otarg is not None:
            assert otarg.dtype == np.float64 and otarg.flags.c_contiguous
        self._call("dbs_o_data_dev", itype, len(c), k1.ctypes.data_as(_ip), k2.ctypes.data_as(_ip),
                   c.ctypes.data_as(_dp), eps_c, ibo,
                   ipbs_a.ctypes.data_as(_ip) if ipbs_a is not None else None, cb, None, _vp(vch), nvch,
                   _vp(fppqq), iitar, iptar_a.ctypes.data_as(_ip) if iptar_a is not None else None,
                   otarg.ctypes.data_as(_dp) if otarg is not None else None)


    # -- L_data.f90 --------------------------------------------------------
    def L_data(self, itype, lst, norb, lval, kbs, kap_list, hd, pq=None, eps_c=1e-10, ibo=2 ** 15, ipbs=None,
               kbort=None, vch=None, fppqq=None, iitar=0, iptar=None, htarg=None, otarg=None, check_target=1):
        """L_data(itype) of one merged list lst = (c, k1, k2, k3, k4) in IPT
        order.  lval: (norb,norb) host table of Lval(i,j); kbs: kappa of each
        orbital; hd: device tensor (nkap, ms, ms) holding hd(ms,ms,nkap) of
        kap_list; pq: device tensor holding pq(ns,2,norb) (itype 2);
        htarg/otarg: packed target matrices (numpy), updated in place."""
        from .capi import LDataAuxC
        c, k1, k2, k3, k4 = lst
        c = np.ascontiguousarray(c, dtype=np.float64)
        K = [_i32(x) for x in (k1, k2, k3, k4)]
        aux = self._aux(ibo, ipbs, kbort, vch, fppqq, iitar, iptar, htarg)
        lv = np.asfortranarray(lval, dtype=np.float64)
        kb, kl = _i32(kbs), _i32(kap_list)
        if otarg is not None:
            assert otarg.dtype == np.float64 and otarg.flags.c_contiguous
        la = LDataAuxC(C.pointer(aux), otarg.ctypes.data_as(_dp) if otarg is not None else None, check_target,
                       int(norb), lv.ctypes.data_as(_dp), kb.ctypes.data_as(_ip), len(kl), kl.ctypes.data_as(_ip),
                       _vp(hd))
        self._keep += [lv, kb, kl, aux]
        self._pre()
        check(lib().dbs_ldata_dev(self.dbs.ctx, C.byref(self._desc), C.byref(la), itype, len(c),
                                  *[x.ctypes.data_as(_ip) for x in K], c.ctypes.data_as(_dp), eps_c, _vp(pq)))
        self._post()

    # -- R_data.f90 / S_data.f90 ----------------------------------------------
    def _aux(self, ibo, ipbs, kbort, vch, fppqq, iitar, iptar, htarg, quadr_pq=None):
        self._keep = []
        ip = _i32(ipbs) if ipbs is not None else None
        it = _i32(iptar) if iptar is not None else None
        cb = KBORT_FN(lambda user, i1, i2: int(kbort(i1, i2))) if kbort is not None else KBORT_FN()
        quser = None
        if isinstance(quadr_pq, np.ndarray):
            # tabulated QUADR_PQ q(norb,norb,0:mmax): the library's own lookup
            # callback, no Python call per list entry
            q = np.asfortranarray(quadr_pq, dtype=np.float64)
            tab = QpqTableC(q.shape[0], q.shape[2] - 1, q.ctypes.data_as(_dp))
            qb = C.cast(lib().dbs_quadr_pq_table, QUADR_PQ_FN)
            quser = C.cast(C.pointer(tab), C.c_void_p)
            self._keep += [q, tab]
        elif quadr_pq is not None:
            qb = QUADR_PQ_FN(lambda user, i1, i2, m: float(quadr_pq(i1, i2, m)))
        else:
            qb = QUADR_PQ_FN()
        self._keep += [ip, it, cb, qb]
        if htarg is not None:
            assert htarg.dtype == np.float64 and htarg.flags.c_contiguous
        return Mat3AuxC(ibo, ip.ctypes.data_as(_ip) if ip is not None else None, cb, None, _vp(vch),
                        int(vch.shape[0]) if vch is not None else 0, _vp(fppqq), iitar,
                        it.ctypes.data_as(_ip) if it is not None else None,
                        htarg.ctypes.data_as(_dp) if htarg is not None else None, qb, quser)

    def _rs(self, name, tkey, k, lst, pq, norb, ibo, ipbs, kbort, vch, fppqq, iitar, iptar, htarg,
            quadr_pq=None, sync=True):
        c, K1, K2, K3, K4 = lst
        c = np.ascontiguousarray(c, dtype=np.float64)
        K = [_i32(x) for x in (K1, K2, K3, K4)]
        aux = self._aux(ibo, ipbs, kbort, vch, fppqq, iitar, iptar, htarg, quadr_pq)
        if sync:
            self._pre()
        check(getattr(lib(), name)(self.dbs.ctx, C.byref(self._desc), C.byref(aux), *tkey, k, len(c),
                                   c.ctypes.data_as(_dp), *[x.ctypes.data_as(_ip) for x in K], _vp(pq), norb))
        if sync:
            self._post()

    def R_data(self, itype, kpol, btype, lst, pq, norb, ibo=2 ** 15, ipbs=None, kbort=None, vch=None,
               fppqq=None, iitar=0, iptar=None, htarg=None, quadr_pq=None, sync=True):
        """R_data(itype, kpol, btype) of one merged list lst = (c, K1, K2, K3, K4)
        in IPT order; pq: device tensor holding pq(ns,2,norb); vch: device
        tensor (nvch, ms) holding V_ch(ms,nvch); ipbs/iptar 1-based tables
        (element i-1 for orbital/channel i); quadr_pq: the host program's
        QUADR_PQ (needed by itype 3, btype 0/1 for the asymptotic
        coefficients acf), a callable quadr_pq(i1, i2, m) or a table
        q(norb,norb,0:mmax) looked up by the library.  sync=False leaves
        the call's device work queued (itype 3/4 return without waiting):
        nothing may touch the matrices from another stream before
        dbs.sync()."""
        self._rs("dbs_rdata_dev", (itype, btype), kpol, lst, pq, norb, ibo, ipbs, kbort, vch, fppqq, iitar,
                 iptar, htarg, quadr_pq, sync)

    def S_data(self, ktype, kpol, lst, pq, norb, ibo=2 ** 15, ipbs=None, kbort=None, vch=None,
               fppqq=None, iitar=0, iptar=None, htarg=None):
        """S_data(ktype, kpol), ktype = atype*10 + itype"""
        self._rs("dbs_sdata_dev", (ktype,), kpol, lst, pq, norb, ibo, ipbs, kbort, vch, fppqq, iitar, iptar,
                 htarg)


def overlap_ci(dbs, lst, SM, nzero, eps_c=1e-10):
    """CI overlap accumulation (met = 0): SM(ic,jc) += cc for jc <= nzero.
    lst = (cc, ic, jc); SM: device tensor holding SM(ldsm, nzero)."""
    import torch
    cc, ic, jc = lst
    cc = np.ascontiguousarray(cc, dtype=np.float64)
    ic, jc = _i32(ic), _i32(jc)
    torch.cuda.synchronize()
    check(lib().dbs_overlap_ci_dev(dbs.ctx, len(cc), ic.ctypes.data_as(_ip), jc.ctypes.data_as(_ip),
                                   cc.ctypes.data_as(_dp), eps_c, nzero, _vp(SM), int(SM.shape[-1])))
    check(lib().dbs_ctx_sync(dbs.ctx))
