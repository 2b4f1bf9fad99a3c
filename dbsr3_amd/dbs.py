"""Host-side mirror of the reference's DBS library interface.

The reference keeps its state in Fortran modules (DBS_grid, DBS_gauss,
DBS_integrals: DBSR_HF/dbsr_lib_dbs.f90:442-477, 830-891, 1569-1591): the
grid, the Gauss tables and a module pointer ``rkb`` that mrk_*/msk_* point at
the array of (type,k); convol then works on that array.  ``DBS`` mirrors that
interface -- same routine names, same argument meaning and errors -- on top
of the C-ABI of libdbsr3_amd.so.  Device arrays are torch float64 tensors
that hold Fortran (column-major) arrays: a Fortran d(ns,ks) is a tensor of
shape (ks, ns).  torch is used only for device memory.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import check, lib

PPPP, QQQQ, PQPQ, QPQP, PPQQ, PQQP = 1, 2, 3, 4, 5, 6
_TYPE_NAMES = {PPPP: "pppp", QQQQ: "qqqq", PQPQ: "pqpq", QPQP: "qpqp", PPQQ: "ppqq", PQQP: "pqqp"}


def _torch():
    import torch
    return torch


def fortran_to_device(a, device="cuda"):
    """numpy array (any order) -> device tensor holding its Fortran layout."""
    torch = _torch()
    a = np.asfortranarray(np.asarray(a, dtype=np.float64))
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(device)


def device_to_fortran(t):
    """device tensor holding a Fortran array -> numpy Fortran-order array."""
    return np.asfortranarray(t.detach().cpu().numpy().T)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class DBS:
    """One B-spline grid on one GPU (def_grid + alloc_DBS_gauss)."""

    def __init__(self, t, ns, ks, ksp=0, ksq=0, device=0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("dbsr3_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = device
        self.t = np.ascontiguousarray(np.asarray(t, dtype=np.float64))
        h = C.c_void_p()
        check(lib().dbs_ctx_create(C.byref(h), ns, ks, ksp, ksq,
                                   self.t.ctypes.data_as(C.POINTER(C.c_double)), device))
        self.ctx = h
        d = np.zeros(8, dtype=np.int32)
        check(lib().dbs_ctx_dims(self.ctx, d.ctypes.data_as(C.POINTER(C.c_int))))
        (self.ns, self.ks, self.ksp, self.ksq, self.nv, self.nsp, self.nsq, self.ms) = map(int, d)
        self.itype = "aaaa"   # module DBS_integrals: itype / krk of the current rkb
        self.krk = -100
        self._type = None

    def close(self):
        if getattr(self, "ctx", None):
            lib().dbs_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_workspace_limit(self, nbytes):
        """device workspace per batch of blocks of the channel
        post-processing (default 256 MiB)"""
        check(lib().dbs_ctx_set_workspace_limit(self.ctx, int(nbytes)))

    def sync(self):
        """wait for the work queued on the context's stream"""
        check(lib().dbs_ctx_sync(self.ctx))

    # -- plumbing --------------------------------------------------------
    def _pre(self):
        _torch().cuda.synchronize()

    def _post(self):
        check(lib().dbs_ctx_sync(self.ctx))

    def _empty(self, *fshape):
        torch = _torch()
        return torch.zeros(tuple(reversed(fshape)), dtype=torch.float64, device="cuda")

    def kb(self, itype):
        return self.ks if itype <= 4 else 2 * self.ks - 1

    # -- tables ------------------------------------------------------------
    def tables(self):
        nv, ks = self.nv, self.ks
        gr = np.zeros((nv, ks), order="F")
        grw = np.zeros((nv, ks), order="F")
        pb = np.zeros((nv + 1, ks, self.ksp), order="F")
        qb = np.zeros((nv + 1, ks, self.ksq), order="F")
        dp = C.POINTER(C.c_double)
        check(lib().dbs_ctx_tables(self.ctx, gr.ctypes.data_as(dp), grw.ctypes.data_as(dp),
                                   pb.ctypes.data_as(dp), qb.ctypes.data_as(dp)))
        return gr, grw, pb, qb

    def moments(self, which, k):
        """moments_pp/qq/pq/qp(k): which 0 pp, 1 qq, 2 pq, 3 qp -> (ks*ks, nv)"""
        out = self._empty(self.ks * self.ks, self.nv)
        self._pre()
        check(lib().dbs_moments_dev(self.ctx, which, k, _ptr(out)))
        self._post()
        return device_to_fortran(out)

    def diag_cells(self, itype, k):
        """triang_<type>(k, iv) for iv=1..nv -> rkd(ks*ks, ks*ks, nv)"""
        kk = self.ks * self.ks
        out = self._empty(kk, kk, self.nv)
        self._pre()
        check(lib().dbs_diag_cells_dev(self.ctx, itype, k, _ptr(out)))
        self._post()
        return device_to_fortran(out)

    # -- arrays: mrk_* / msk_* select the current rkb ------------------------
    def _select(self, itype, k):
        p = C.c_void_p()
        check(lib().dbs_mrk(self.ctx, itype, k, C.byref(p)))
        self._type = itype
        self.itype = _TYPE_NAMES[itype]
        self.krk = k

    def mrk_pppp(self, k): self._select(PPPP, k)
    def mrk_qqqq(self, k): self._select(QQQQ, k)
    def mrk_pqpq(self, k): self._select(PQPQ, k)
    def mrk_qpqp(self, k): self._select(QPQP, k)
    def msk_ppqq(self, k): self._select(PPQQ, k)
    def msk_pqqp(self, k): self._select(PQQP, k)

    def rkb(self):
        """the current array rkb(ns,ns,kb,kb) as a numpy Fortran array"""
        if self._type is None:
            raise RuntimeError("rkb: no array selected (call mrk_*/msk_* first)")
        kb = self.kb(self._type)
        out = np.zeros((self.ns, self.ns, kb, kb), order="F")
        check(lib().dbs_mrk_host(self.ctx, self._type, self.krk,
                                 out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    # -- contractions ---------------------------------------------------------
    def density(self, p1, p2, type):
        """density(ns,ks,d,p1,p2,type); p1,p2 device tensors of length ns"""
        ncol = {"s": self.ks, "u": self.ks, "l": self.ks, "n": 2 * self.ks - 1}.get(type, self.ns)
        d = self._empty(self.ns, ncol)
        self._pre()
        check(lib().dbs_density_dev(self.ctx, _ptr(p1), _ptr(p2), type.encode(), _ptr(d)))
        self._post()
        return d

    def convol(self, d, icase, sym_i, sym_j):
        """convol(ns,ks,a,d,icase,sym_i,sym_j) on the current rkb -> a"""
        if self._type is None:
            raise RuntimeError("convol: no array selected (call mrk_*/msk_* first)")
        if icase <= 2:
            ncol = 2 * self.ks - 1 if "n" in (sym_i, sym_j) else self.ks
            a = self._empty(self.ns, ncol)
        else:
            a = self._empty(self.ns, self.ns)
        self._pre()
        check(lib().dbs_convol_dev(self.ctx, self._type, self.krk, icase, sym_i.encode(),
                                   sym_j.encode(), _ptr(d), _ptr(a)))
        self._post()
        return a

    def SUM_AmB(self, a, b, sym):
        torch = _torch()
        out = torch.zeros(1, dtype=torch.float64, device="cuda")
        self._pre()
        check(lib().dbs_sum_amb_dev(self.ctx, _ptr(a), _ptr(b), sym.encode(), _ptr(out)))
        self._post()
        return float(out.item())

    def rk(self, f1, f2, f3, f4, k):
        """rk(f1,f2,f3,f4,k); f* device tensors holding (ns,2)"""
        torch = _torch()
        out = torch.zeros(1, dtype=torch.float64, device="cuda")
        self._pre()
        check(lib().dbs_rk_dev(self.ctx, _ptr(f1), _ptr(f2), _ptr(f3), _ptr(f4), k, _ptr(out)))
        self._post()
        return float(out.item())

    def quadr(self, fi, fj, m):
        torch = _torch()
        out = torch.zeros(1, dtype=torch.float64, device="cuda")
        self._pre()
        check(lib().dbs_quadr_dev(self.ctx, _ptr(fi), _ptr(fj), m, _ptr(out)))
        self._post()
        return float(out.item())

    # -- overlaps and unfolding ---------------------------------------------
    def bspline_overlaps(self):
        """fpbs(ns,ns), fqbs(ns,ns), fppqq(ms,ms) as device tensors
        (ZINTYM via Def_pbsp / Def_qbsp / Def_pq_arrays)"""
        fp, fq, fpq = self._empty(self.ns, self.ns), self._empty(self.ns, self.ns), self._empty(self.ms, self.ms)
        self._pre()
        check(lib().dbs_bspline_overlaps_dev(self.ctx, _ptr(fp), _ptr(fq), _ptr(fpq)))
        self._post()
        return fp, fq, fpq

    # -- the Dirac operator -------------------------------------------------
    @staticmethod
    def nucleus(model, z, r_uniform=0.0, a_fermi=0.0, c_fermi=0.0):
        """module DBS_nuclear's charge distribution: model 'point',
        'uniform' or 'Fermi' (or 0, 1, 2), atomic number z, radii in a.u."""
        from .capi import NucleusC
        codes = {"point": 0, "uniform": 1, "fermi": 2}
        m = codes[model.lower()] if isinstance(model, str) else int(model)
        return NucleusC(m, float(z), float(r_uniform), float(a_fermi), float(c_fermi))

    def vnucl(self, nucleus):
        """VR_nucl(nv,ks) = V_nuclear(gr)*grw (Def_Vnucl) as a device tensor"""
        vr = self._empty(self.nv, self.ks)
        self._pre()
        check(lib().dbs_vnucl_dev(self.ctx, C.byref(nucleus), _ptr(vr)))
        self._post()
        return vr

    def dhl_pq(self, nucleus, kappa, bloch=True, rb=0.0, pnu=1.0):
        """MAT_dhl_pq(kappa): the Dirac operator dhl(ms,ms) as a device
        tensor (row-major view of the column-major matrix: dhl[j, i] =
        dhl(i,j)); bloch/rb/pnu are module DBS_dhl_pq's mbloch, RB, pnu"""
        d = self._empty(self.ms, self.ms)
        self._pre()
        check(lib().dbs_dhl_pq_dev(self.ctx, C.byref(nucleus), int(kappa), 1 if bloch else 0, float(rb),
                                   float(pnu), _ptr(d)))
        self._post()
        return d

    def Full_mat_sym(self, x, sym):
        """y(ns,ns) from band/full storage x (device tensor) in mode sym"""
        y = self._empty(self.ns, self.ns)
        self._pre()
        check(lib().dbs_full_mat_sym_dev(self.ctx, _ptr(x), sym.encode(), _ptr(y)))
        self._post()
        return y

    def hf_matrix(self, i, pq, acoef, bcoef, kmax, qsum, hfm=None, breit=None):
        """two-electron part of hf_matrix for orbital i (1-based): pq device
        tensor holding pq(ns,2,norb); acoef/bcoef (norb, kmax+1) host arrays
        of a(i,j,k), b(i,j,k); breit = (kr1, kr2, kr3, kr4, crk) or None.
        hfm: device tensor holding hfm(ms,ms) (zeros if None), updated."""
        norb = int(pq.shape[0])
        if hfm is None:
            hfm = self._empty(self.ms, self.ms)
        A = np.asfortranarray(acoef, dtype=np.float64)
        B = np.asfortranarray(bcoef, dtype=np.float64)
        ci = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        if breit is None:
            lst, mb = (0, None, None, None, None, None), 0
        else:
            kr1, kr2, kr3, kr4, crk = breit
            lst = (len(crk), ci(kr1), ci(kr2), ci(kr3), ci(kr4), np.ascontiguousarray(crk, dtype=np.float64))
            mb = 2
        ip_ = C.POINTER(C.c_int)
        dp_ = C.POINTER(C.c_double)
        arr = lambda a, t: a.ctypes.data_as(t) if a is not None else None
        self._pre()
        check(lib().dbs_hf_matrix_dev(self.ctx, i, norb, _ptr(pq), A.ctypes.data_as(dp_), B.ctypes.data_as(dp_),
                                      kmax, float(qsum), mb, lst[0], arr(lst[1], ip_), arr(lst[2], ip_),
                                      arr(lst[3], ip_), arr(lst[4], ip_), arr(lst[5], dp_), _ptr(hfm)))
        self._post()
        return hfm

    def mchf_matrix(self, i, pq, dhl, llist, rlist, kmax, eps_c=1e-8, qsum_min=1e-8, newton=0, irhs=0,
                    fppqq=None):
        """mchf_matrix for the non-core orbital i (1-based): pq device tensor
        holding pq(ns,2,norb), dhl device tensor holding Get_hd_core(kbs(i))
        (ms,ms); llist rows (i1, i2, c) and rlist rows (k, i1, i2, i3, i4, c)
        with c = WW(ic,jc)*coef per angular coefficient.  Returns (hfm, rhs,
        hx, e(i), qsum(i)), the matrices as device tensors."""
        norb = int(pq.shape[0])
        L = np.asarray(llist, dtype=np.float64).reshape(-1, 3)
        R = np.asarray(rlist, dtype=np.float64).reshape(-1, 6)
        ci = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        cd = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        l1, l2, lc = ci(L[:, 0]), ci(L[:, 1]), cd(L[:, 2])
        rk, r1, r2, r3, r4, rc = ci(R[:, 0]), ci(R[:, 1]), ci(R[:, 2]), ci(R[:, 3]), ci(R[:, 4]), cd(R[:, 5])
        hfm = self._empty(self.ms, self.ms)
        hx = self._empty(self.ms, self.ms)
        rhs = self._empty(self.ms)
        e = C.c_double()
        q = C.c_double()
        ip_ = C.POINTER(C.c_int)
        dp_ = C.POINTER(C.c_double)
        self._pre()
        check(lib().dbs_mchf_matrix_dev(self.ctx, i, norb, _ptr(pq), _ptr(dhl), len(lc), l1.ctypes.data_as(ip_),
                                        l2.ctypes.data_as(ip_), lc.ctypes.data_as(dp_), len(rc),
                                        rk.ctypes.data_as(ip_), r1.ctypes.data_as(ip_), r2.ctypes.data_as(ip_),
                                        r3.ctypes.data_as(ip_), r4.ctypes.data_as(ip_), rc.ctypes.data_as(dp_),
                                        kmax, float(eps_c), float(qsum_min), int(newton), int(irhs),
                                        _ptr(fppqq) if fppqq is not None else None, _ptr(hfm), _ptr(rhs),
                                        _ptr(hx), C.byref(e), C.byref(q)))
        self._post()
        return hfm, rhs, hx, e.value, q.value

    def core_matrix(self, i, pq, hd, acoef, bcoef, kmax, qsum, newton=0):
        """core_matrix for the core orbital i (1-based): hd device tensor
        holding Get_hd(kbs(i)) (ms,ms); acoef/bcoef (norb, kmax+1) host
        a(i,j,k), b(i,j,k); qsum (norb) host.  Returns (hfm, hx) device
        tensors."""
        norb = int(pq.shape[0])
        A = np.asfortranarray(acoef, dtype=np.float64)
        B = np.asfortranarray(bcoef, dtype=np.float64)
        q = np.ascontiguousarray(qsum, dtype=np.float64)
        hfm = self._empty(self.ms, self.ms)
        hx = self._empty(self.ms, self.ms)
        dp_ = C.POINTER(C.c_double)
        self._pre()
        check(lib().dbs_core_matrix_dev(self.ctx, i, norb, _ptr(pq), _ptr(hd), A.ctypes.data_as(dp_),
                                        B.ctypes.data_as(dp_), kmax, q.ctypes.data_as(dp_), int(newton),
                                        _ptr(hfm), _ptr(hx)))
        self._post()
        return hfm, hx

    def hf_eiv(self, hfm, vorth, iprm, m, fppqq):
        """hf_eiv (DBSR_HF/hf_solve_HF.f90:118-202): hfm device tensor holding
        hfm(ms,ms) (overwritten by the projections), vorth a list of device
        tensors (ms) to project out, iprm (ms) host 0/1, m the wanted
        positive-energy eigenvalue, fppqq device tensor (ms,ms).  Returns
        (v device tensor (ms), e)."""
        torch = _torch()
        ms = self.ms
        vo = torch.stack(list(vorth)) if len(vorth) else None
        ip = np.ascontiguousarray(iprm, dtype=np.int32)
        v = self._empty(ms)
        e = C.c_double()
        self._pre()
        check(lib().dbs_hf_eiv_dev(self.ctx, _ptr(hfm), 0 if vo is None else int(vo.shape[0]),
                                   _ptr(vo) if vo is not None else None, ip.ctypes.data_as(C.POINTER(C.c_int)),
                                   int(m), _ptr(fppqq), _ptr(v), C.byref(e)))
        self._post()
        return v, e.value

    def UPDATE_HS(self, hm, ii, jj, d, sym):
        """UPDATE_HS(ms,hm,ii,jj,ns,ks,d,sym): hm device tensor holding (ms,ms)"""
        self._pre()
        check(lib().dbs_update_hs_dev(self.ctx, _ptr(hm), ii, jj, _ptr(d), sym.encode()))
        self._post()
        return hm
