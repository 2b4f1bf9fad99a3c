"""ctypes front end to the oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/liboracle.so`` -- the C restatement of the reference's
B-spline Slater-integral algorithms (``oracle/dbs_oracle.c``) -- and, where it
was built, ``oracle/_ref/libdbsref.so`` -- the reference's own Fortran
(``DBSR_HF/dbsr_lib_dbs.f90``, ``dbsr_lib_zcom.f90``, ``DBSR_MAT3/MOD_c_data.f90``)
compiled from its sources.  Only tests, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg import this module, and only as the checker.

Arrays are numpy float64 in Fortran (column-major) order with the reference's
dimensions.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libdbsref.so")

PPPP, QQQQ, PQPQ, QPQP, PPQQ, PQQP = 1, 2, 3, 4, 5, 6
MOM_PP, MOM_QQ, MOM_PQ, MOM_QP = 0, 1, 2, 3

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


_lib = None
KBORT_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_grid_new.restype = C.c_void_p
        L.orc_grid_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_grid_delete.argtypes = [C.c_void_p]
        L.orc_grid_dims.argtypes = [C.c_void_p, _ip]
        L.orc_grid_tables.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_moments.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.orc_triang.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_mrk.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        L.orc_rkb_size.restype = C.c_size_t
        L.orc_rkb_size.argtypes = [C.c_void_p, C.c_int]
        L.orc_density.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_char]
        L.orc_convol.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, C.c_char, C.c_char]
        L.orc_sum_amb.restype = C.c_double
        L.orc_sum_amb.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_char]
        L.orc_quadr.restype = C.c_double
        L.orc_quadr.argtypes = [C.c_void_p, _dp, _dp, C.c_int]
        L.orc_update_hs_lib.argtypes = [C.c_int, _dp, C.c_int, C.c_int, C.c_int, C.c_int, _dp, C.c_char]
        L.orc_update_hs_mat.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_char, C.c_int, C.c_int]
        L.orc_rk.restype = C.c_double
        L.orc_rk.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_gauleg.argtypes = [C.c_double, C.c_double, _dp, _dp, C.c_int]
        L.orc_mkgrid_01.restype = C.c_int
        L.orc_mkgrid_01.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                    C.c_double, C.c_double, _dp, C.c_int]
        L.orc_mkgrid_02.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_grid_start.restype = C.c_double
        L.orc_grid_start.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.orc_mkgrid_01s.restype = C.c_int
        L.orc_mkgrid_01s.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _dp,
                                     C.c_int]
        L.orc_mkgrid_02s.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_rk_evaluate.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_int, C.c_int, _dp,
                                      _ip, _ip, _ip, _ip, C.c_double, _dp, C.c_int,
                                      _dp, C.c_int, C.c_int, _dp]
        L.orc_rdata_34.argtypes = [C.c_int, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _dp,
                                   _ip, _ip, _ip, _dp, C.c_int, _ip, C.c_int, _dp]
        L.orc_idef_rtype.restype = C.c_int
        L.orc_idef_rtype.argtypes = [C.c_int] * 4 + [_ip, C.c_int, _ip, _ip, _ip, _ip]
        L.orc_decode_int.argtypes = [C.c_int] + [_ip] * 6
        L.orc_full_mat_sym.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_char]
        L.orc_zintym.argtypes = [C.c_int] * 4 + [_dp, _dp, _dp, C.c_int, _dp]
        L.orc_overlaps.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_update_hx.argtypes = [C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_double]
        L.orc_update_hw.argtypes = [C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, C.c_double]
        L.orc_update_hv.argtypes = [C.c_int, _dp, _dp, C.c_double]
        L.orc_o_data.restype = C.c_int
        L.orc_o_data.argtypes = [C.c_int, C.c_int, C.c_int, _ip, _ip, _dp, C.c_double, C.c_int, _ip,
                                 KBORT_FN, C.c_void_p, _dp, _dp, C.c_int, _ip, _dp,
                                 C.c_int, C.c_int, _ip, _dp, _ip, _dp, _ip, _dp]
        L.orc_bav.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_char, C.c_char]
        L.orc_hf_matrix.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int,
                                    C.c_int, _ip, _ip, _ip, _ip, _dp, _dp]
        L.orc_rs_data.restype = C.c_int
        L.orc_rs_data.argtypes = ([C.c_int, C.c_int, _dp] + [C.c_int] * 4 + [_dp] + [_ip] * 4 +
                                  [_dp, C.c_int, _ip, KBORT_FN, C.c_void_p, _dp, _dp, C.c_int, _ip, _dp,
                                   C.c_int, C.c_int, _ip, _dp, _ip, _dp, _ip, _dp])
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# grids
# --------------------------------------------------------------------------
def gauleg(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauleg(0.0, 1.0, _d(x), _d(w), n)
    return x, w


def mkgrid_01(ks, ml, hi, he, hmax, rmax, z):
    ns = lib().orc_mkgrid_01(ks, ml, hi, he, hmax, rmax, z, None, 0)
    t = np.zeros(ns + ks)
    lib().orc_mkgrid_01(ks, ml, hi, he, hmax, rmax, z, _d(t), ns + ks)
    return t, ns


def mkgrid_02(ns, ks, hi, rmax, z):
    t = np.zeros(ns + ks)
    lib().orc_mkgrid_02(ns, ks, hi, rmax, z, _d(t))
    return t


NUCLEAR = {"point": 0, "Fermi": 1, "uniform": 2}


def mkgrid(grid_type, ks, hi, rmax, z, nuclear="point", a_fermi=0.0, r_uniform=0.0, ml=1, he=0.25, hmax=1.0,
           ns=0):
    """mkgrid_01 (grid_type 1) / mkgrid_02 (grid_type 2, given ns) with the
    first step of the nucleus model; returns (t, ns)"""
    L = lib()
    h0 = L.orc_grid_start(NUCLEAR[nuclear], hi, z, a_fermi, r_uniform)
    if grid_type == 1:
        ns = L.orc_mkgrid_01s(ks, ml, h0, he, hmax, rmax, None, 0)
        t = np.zeros(ns + ks)
        L.orc_mkgrid_01s(ks, ml, h0, he, hmax, rmax, _d(t), ns + ks)
    else:
        t = np.zeros(ns + ks)
        L.orc_mkgrid_02s(ns, ks, h0, rmax, _d(t))
    return t, ns


def ref_mkgrid(grid_type, ks, hi, rmax, z, nuclear="point", a_fermi=0.0, r_uniform=0.0, ml=1, he=0.25,
               hmax=1.0, ns=0):
    """the same by the reference's own mkgrid_01/02 (oracle/_ref)"""
    R = ref()
    args = (grid_type, ks, ml, hi, he, hmax, rmax, z, NUCLEAR[nuclear], a_fermi, r_uniform, ns)
    n = R.ref_mkgrid(*args, None, 0)
    t = np.zeros(n + ks)
    R.ref_mkgrid(*args, _d(t), n + ks)
    return t, n


def ref_read_grid(path, ns, ks):
    """the reference's read_grid on a knot file"""
    t = np.zeros(ns + ks)
    b = path.encode()
    assert ref().ref_read_grid(b, len(b), ns, ks, _d(t)) == 0
    return t


class Grid:
    """B-spline grid + Gauss tables (alloc_DBS_gauss restated)."""

    def __init__(self, t, ns, ks, ksp=0, ksq=0):
        self.t = f64(t)
        h = lib().orc_grid_new(ns, ks, ksp, ksq, _d(self.t))
        if not h:
            raise ValueError("bad grid parameters")
        self.h = C.c_void_p(h)
        dims = np.zeros(8, dtype=np.int32)
        lib().orc_grid_dims(self.h, _i(dims))
        (self.ns, self.ks, self.ksp, self.ksq, self.nv, self.nsp, self.nsq, self.ms) = map(int, dims)
        self._cache = {}

    def __del__(self):
        try:
            lib().orc_grid_delete(self.h)
        except Exception:
            pass

    def tables(self):
        nv, ks = self.nv, self.ks
        gr = np.zeros((nv, ks), order="F")
        grw = np.zeros((nv, ks), order="F")
        pb = np.zeros((nv + 1, ks, self.ksp), order="F")
        qb = np.zeros((nv + 1, ks, self.ksq), order="F")
        lib().orc_grid_tables(self.h, _d(gr), _d(grw), _d(pb), _d(qb))
        return gr, grw, pb, qb

    def moments(self, which, k):
        r = np.zeros((self.ks * self.ks, self.nv), order="F")
        lib().orc_moments(self.h, which, k, _d(r))
        return r

    def triang(self, itype, k, iv):
        kk = self.ks * self.ks
        r = np.zeros((kk, kk), order="F")
        lib().orc_triang(self.h, itype, k, iv, _d(r))
        return r

    def kb(self, itype):
        return self.ks if itype <= 4 else 2 * self.ks - 1

    def mrk(self, itype, k, cache=True):
        key = (itype, k)
        if cache and key in self._cache:
            return self._cache[key]
        kb = self.kb(itype)
        r = np.zeros((self.ns, self.ns, kb, kb), order="F")
        lib().orc_mrk(self.h, itype, k, _d(r))
        if cache:
            self._cache[key] = r
        return r

    def quadr(self, fi, fj, m):
        return lib().orc_quadr(self.h, _d(f64(fi)), _d(f64(fj)), m)

    def rk(self, f1, f2, f3, f4, k):
        a = [self.mrk(t, k) for t in (PPPP, QQQQ, PQPQ, QPQP)]
        return lib().orc_rk(self.h, _d(a[0]), _d(a[1]), _d(a[2]), _d(a[3]),
                            _d(f64(f1)), _d(f64(f2)), _d(f64(f3)), _d(f64(f4)))


def density(ns, ks, p1, p2, typ):
    ncol = {"s": ks, "u": ks, "l": ks, "n": 2 * ks - 1}.get(typ, ns)
    d = np.zeros((ns, ncol), order="F")
    lib().orc_density(ns, ks, _d(d), _d(f64(p1)), _d(f64(p2)), typ.encode())
    return d


def convol(ns, ks, rkb, d, icase, si, sj):
    kb = rkb.shape[2]
    if icase <= 2:
        ncol = 2 * ks - 1 if "n" in (si, sj) else ks
        a = np.zeros((ns, max(ncol, ks)), order="F")
    else:
        a = np.zeros((ns, ns), order="F")
    dd = f64(d)
    if dd.shape[1] < a.shape[1] and icase <= 2:
        pad = np.zeros((ns, a.shape[1]), order="F")
        pad[:, : dd.shape[1]] = dd
        dd = pad
    lib().orc_convol(ns, ks, kb, _d(rkb), _d(a), _d(dd), icase, si.encode(), sj.encode())
    return a


def sum_amb(ns, ks, a, b, sym):
    return lib().orc_sum_amb(ns, ks, _d(f64(a)), _d(f64(b)), sym.encode())


def update_hs_lib(hm, ii, jj, ns, ks, d, sym):
    ms = hm.shape[0]
    lib().orc_update_hs_lib(ms, _d(hm), ii, jj, ns, ks, _d(f64(d)), sym.encode())


def update_hs_mat(blk, ns, ks, ich, jch, d, sym, ii, jj):
    lib().orc_update_hs_mat(ns, ks, _d(blk), ich, jch, _d(f64(d)), sym.encode(), ii, jj)


def rk_evaluate(g, itype, k, lst, pq, ibi, HM, nzero, DM, eps_c=1e-10):
    """lst = (c, K1, K2, K3, K4) in IPT order; pq (ns,2,norb); HM/DM updated in place."""
    c, K1, K2, K3, K4 = lst
    rkb = g.mrk(itype, k)
    ci = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    K1, K2, K3, K4 = ci(K1), ci(K2), ci(K3), ci(K4)
    cc = np.ascontiguousarray(c, dtype=np.float64)
    lib().orc_rk_evaluate(g.ns, g.ks, rkb.shape[2], _d(rkb), itype, len(cc), _d(cc),
                          _i(K1), _i(K2), _i(K3), _i(K4), eps_c, _d(f64(pq)), ibi,
                          _d(HM), HM.shape[0], nzero, _d(DM))


def rdata_34(g, itype, atype, k, lst, pq, ibo, icc, hch):
    c, K1, K2, K3 = lst
    rkb = g.mrk({0: PPPP, 1: QQQQ, 2: PQPQ, 3: QPQP}[atype], k)
    ci = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    icc_f = np.asfortranarray(icc, dtype=np.int32)
    cc = np.ascontiguousarray(c, dtype=np.float64)
    lib().orc_rdata_34(g.ns, g.ks, _d(rkb), itype, atype, len(cc), _d(cc), _i(ci(K1)),
                       _i(ci(K2)), _i(ci(K3)), _d(f64(pq)), ibo, _i(icc_f), icc.shape[0], _d(hch))


def full_mat_sym(ns, ks, x, sym):
    """Full_mat_sym (DBSR_HF/dbsr_lib_zcom.f90:1247-1294)"""
    y = np.zeros((ns, ns), order="F")
    lib().orc_full_mat_sym(ns, ks, _d(f64(x)), _d(y), sym.encode())
    return y


def zintym(nv, ks, ks1, ks2, b1, b2, ygr, mm):
    """ZINTYM (DBSR_HF/dbsr_lib_dbs.f90:4472-4497)"""
    ym = np.zeros((mm, mm), order="F")
    lib().orc_zintym(nv, ks, ks1, ks2, _d(f64(b1)), _d(f64(b2)), _d(f64(ygr)), mm, _d(ym))
    return ym


def overlaps(g):
    """fpbs, fqbs (Def_pbsp/Def_qbsp) and fppqq (Def_pq_arrays)"""
    fp = np.zeros((g.ns, g.ns), order="F")
    fq = np.zeros((g.ns, g.ns), order="F")
    fpq = np.zeros((g.ms, g.ms), order="F")
    lib().orc_overlaps(g.h, _d(fp), _d(fq), _d(fpq))
    return fp, fq, fpq


def update_hx(blk, ich, jch, d, C_):
    lib().orc_update_hx(blk.shape[0], _d(blk), ich, jch, _d(f64(d)), C_)


def update_hw(blk, ich, jch, v, w, C_):
    lib().orc_update_hw(blk.shape[0], _d(blk), ich, jch, _d(f64(v)), _d(f64(w)), C_)


def update_hv(col, v, C_):
    lib().orc_update_hv(col.shape[0], _d(col), _d(f64(v)), C_)


def o_data(ms, itype, lst, eps_c, ibo, ipbs, kbort, vch, fppqq, iitar, iptar, otarg, mats):
    """O_data (DBSR_MAT3/O_data.f90:2-93).  lst = (c, k1, k2) in IPT order;
    ipbs/iptar 1-based (element 0 unused); kbort(i1,i2) -> 1-based column of
    vch; mats = dict(icc, hch, icb, hcp, ibb, hp), updated in place.
    Returns the restatement's status (0 = ok)."""
    c, k1, k2 = lst
    ci = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    fi = lambda a: np.asfortranarray(a, dtype=np.int32)
    cb = KBORT_FN(lambda user, i1, i2: int(kbort(i1, i2)))
    icc, icb, ibb = fi(mats["icc"]), fi(mats["icb"]), fi(mats["ibb"])
    nch, npert = icc.shape[0], ibb.shape[0]
    return lib().orc_o_data(ms, itype, len(c), _i(ci(k1)), _i(ci(k2)), _d(np.ascontiguousarray(c, dtype=np.float64)),
                            eps_c, ibo, _i(ci(ipbs)), cb, None, _d(f64(vch)), _d(f64(fppqq)), iitar,
                            _i(ci(iptar)), _d(otarg), nch, npert, _i(icc), _d(mats["hch"]), _i(icb),
                            _d(mats["hcp"]), _i(ibb), _d(mats["hp"]))


def hf_matrix(g, i, pq, acoef, bcoef, kmax, qsum, hfm=None, breit=None):
    """hf_matrix two-electron part for orbital i (1-based); acoef/bcoef
    (norb, kmax+1) = a(i,j,k), b(i,j,k); breit = (kr1, kr2, kr3, kr4, crk)
    or None.  Returns hfm (ms,ms)."""
    norb = pq.shape[2]
    h = np.zeros((g.ms, g.ms), order="F") if hfm is None else hfm
    ci = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    if breit is None:
        z = np.zeros(1, dtype=np.int32)
        br = (0, z, z, z, z, np.zeros(1))
        mb = 0
    else:
        kr1, kr2, kr3, kr4, crk = breit
        br = (len(crk), ci(kr1), ci(kr2), ci(kr3), ci(kr4), np.ascontiguousarray(crk, dtype=np.float64))
        mb = 2
    lib().orc_hf_matrix(g.h, i, norb, _d(f64(pq)), _d(f64(acoef)), _d(f64(bcoef)), kmax, qsum, mb, br[0],
                        _i(br[1]), _i(br[2]), _i(br[3]), _i(br[4]), _d(br[5]), _d(h))
    return h


def bav(ns, ks, a, b, sym, dir_):
    """BAV restated: A b ('r') or A^T b ('l') for A held in storage sym"""
    c = np.zeros(ns)
    lib().orc_bav(ns, ks, _d(f64(a)), _d(f64(b)), _d(c), sym.encode(), dir_.encode())
    return c


def rs_data(g, stype, itype, atype, k, lst, pq, ibo, ipbs, kbort, vch, fppqq, iitar, iptar, htarg, mats):
    """R_data (stype 0) / S_data (stype 1) restated; lst = (c, K1, K2, K3, K4)
    in IPT order; ipbs/iptar 1-based (element 0 unused); mats as o_data."""
    c, K1, K2, K3, K4 = lst
    typ = (5 + atype) if stype else (atype + 1)
    rkb = g.mrk(typ, k)
    ci = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    fi = lambda a: np.asfortranarray(a, dtype=np.int32)
    cb = KBORT_FN(lambda user, i1, i2: int(kbort(i1, i2)))
    icc, icb, ibb = fi(mats["icc"]), fi(mats["icb"]), fi(mats["ibb"])
    return lib().orc_rs_data(g.ns, g.ks, _d(rkb), stype, itype, atype, len(c),
                             _d(np.ascontiguousarray(c, dtype=np.float64)), _i(ci(K1)), _i(ci(K2)), _i(ci(K3)),
                             _i(ci(K4)), _d(f64(pq)), ibo, _i(ci(ipbs)), cb, None, _d(f64(vch)), _d(f64(fppqq)),
                             iitar, _i(ci(iptar)), _d(htarg), icc.shape[0], ibb.shape[0], _i(icc),
                             _d(mats["hch"]), _i(icb), _d(mats["hcp"]), _i(ibb), _d(mats["hp"]))


def rdata_acf(lst, ibo, kpol, mk, qpq, icc, acf):
    """asymptotic coefficients of R_data itype 3, atype 0/1 (R_data.f90:214-240):
    lst = (c, K1, K2, K3) in IPT order; qpq(norb,norb) = QUADR_PQ(i,j,kpol);
    acf (nblocks, mk+1) Fortran order, updated in place; returns the status"""
    c, K1, K2, K3 = lst[:4]
    ci = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    icc_f = np.asfortranarray(icc, dtype=np.int32)
    q = np.asfortranarray(qpq, dtype=np.float64)
    assert acf.flags.f_contiguous and acf.dtype == np.float64
    return lib().orc_rdata_acf(len(c), _d(np.ascontiguousarray(c, dtype=np.float64)), _i(ci(K1)), _i(ci(K2)),
                               _i(ci(K3)), ibo, kpol, mk, _d(q), q.shape[0], _i(icc_f), icc_f.shape[0],
                               acf.shape[0], _d(acf))


def idef_itype(atype, i1, i2, i3, i4, c, ic, jc, io, jo, ipbs, ibo):
    """Idef_itype (DBSR_MAT3/idef_type.f90:306-437); returns (itype + 10*atype,
    key) (itype 0: dropped, < 0: the reference stops)"""
    L = lib()
    L.orc_idef_itype.restype = C.c_int
    L.orc_idef_itype.argtypes = [C.c_int] * 5 + [C.c_double] + [C.c_int] * 4 + [_ip, C.c_int, _ip]
    key = np.zeros(4, dtype=np.int32)
    ip = np.ascontiguousarray(ipbs, dtype=np.int32)
    it = L.orc_idef_itype(atype, i1, i2, i3, i4, c, ic, jc, io, jo, _i(ip), ibo, _i(key))
    return it, tuple(int(x) for x in key)


def idef_otype(c, ic, jc, io, jo, eps_c=1e-10, check_target=1):
    """Idef_Otype (DBSR_MAT3/idef_type.f90:2-50); returns (itype, key)"""
    L = lib()
    L.orc_idef_otype.restype = C.c_int
    L.orc_idef_otype.argtypes = [C.c_double] + [C.c_int] * 4 + [C.c_double, C.c_int, _ip]
    key = np.zeros(4, dtype=np.int32)
    it = L.orc_idef_otype(c, ic, jc, io, jo, eps_c, check_target, _i(key))
    return it, tuple(int(x) for x in key)


def add_integral_ci(cd, kpol, ic, jc, i1, i2, i3, i4, c, kappa, ibi=2 ** 15, mbreit=0):
    """CI3 Add_integral (DBSR_CI3/add_integral.f90:2-43) into the oracle's
    CData cd; kappa indexed by orbital (element 0 unused)"""
    L = lib()
    L.orc_add_integral_ci.restype = C.c_int
    L.orc_add_integral_ci.argtypes = [C.c_void_p] + [C.c_int] * 7 + [C.c_double, _ip, C.c_int, C.c_int]
    k = np.ascontiguousarray(kappa, dtype=np.int32)
    return L.orc_add_integral_ci(cd.h, kpol, ic, jc, i1, i2, i3, i4, c, _i(k), ibi, mbreit)


def idef_rtype(j1, j2, j3, j4, ipbs, ibo):
    ip = np.ascontiguousarray(ipbs, dtype=np.int32)
    out = [np.zeros(1, dtype=np.int32) for _ in range(4)]
    it = lib().orc_idef_rtype(j1, j2, j3, j4, _i(ip), ibo, *[_i(o) for o in out])
    return it, tuple(int(o[0]) for o in out)


def idef_rtype_full(j1, j2, j3, j4, c, ic, jc, io, jo, ipbs, ibo, eps_c=1e-10, check_target=1):
    """Idef_Rtype, all structures; returns (itype, key) (itype <= 0: dropped / stop)"""
    L = lib()
    L.orc_idef_rtype_full.restype = C.c_int
    L.orc_idef_rtype_full.argtypes = [C.c_int] * 4 + [C.c_double] + [C.c_int] * 4 + [_ip, C.c_int, C.c_double,
                                                                                      C.c_int, _ip]
    key = np.zeros(4, dtype=np.int32)
    it = L.orc_idef_rtype_full(j1, j2, j3, j4, c, ic, jc, io, jo, _i(np.ascontiguousarray(ipbs, dtype=np.int32)),
                               ibo, eps_c, check_target, _i(key))
    return it, tuple(int(x) for x in key)


def decode_int(code):
    out = [np.zeros(1, dtype=np.int32) for _ in range(6)]
    lib().orc_decode_int(int(code), *[_i(o) for o in out])
    return tuple(int(o[0]) for o in out)


# --------------------------------------------------------------------------
# c_data (DBSR_MAT3/MOD_c_data.f90 restated in dbs_oracle.c)
# --------------------------------------------------------------------------
GEN_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_int, C.c_int, _dp, _ip, _ip, _ip, _ip)


class CData:
    """Blocked sorted coefficient lists.  ``gen(itype, kpol, (c,K1,K2,K3,K4))``
    is called by Add_matrix with the merged list in IPT order."""

    def __init__(self, ntype, kp1, kp2, mblock, nblock, kblock, eps_c, gen):
        L = lib()
        L.orc_cdata_alloc.restype = C.c_void_p
        L.orc_cdata_alloc.argtypes = [C.c_int] * 6 + [C.c_double, GEN_FN, C.c_void_p]
        L.orc_cdata_free.argtypes = [C.c_void_p]
        L.orc_cdata_add_coef.restype = C.c_int
        L.orc_cdata_add_coef.argtypes = [C.c_void_p, C.c_double] + [C.c_int] * 6
        L.orc_cdata_add_matrix.restype = C.c_int
        L.orc_cdata_add_matrix.argtypes = [C.c_void_p, C.c_int, C.c_int]

        def _cb(user, itype, kpol, nc, c, k1, k2, k3, k4):
            if nc > 0:
                arrs = (np.ctypeslib.as_array(c, (nc,)).copy(),) + tuple(
                    np.ctypeslib.as_array(p, (nc,)).copy() for p in (k1, k2, k3, k4))
            else:
                arrs = (np.zeros(0),) + tuple(np.zeros(0, dtype=np.int32) for _ in range(4))
            gen(itype, kpol, arrs)

        self._cb = GEN_FN(_cb)
        self.h = C.c_void_p(L.orc_cdata_alloc(ntype, kp1, kp2, mblock, nblock, kblock, eps_c,
                                                self._cb, None))

    def add_coef(self, c, kpol, j1, j2, j3, j4, itype):
        r = lib().orc_cdata_add_coef(self.h, c, kpol, j1, j2, j3, j4, itype)
        if r < 0:
            raise RuntimeError(" Add_coef: problems with new block ")

    def add_matrix(self, itype, kpol):
        lib().orc_cdata_add_matrix(self.h, itype, kpol)

    def add_integral_mat3(self, icase, kpol, i1, i2, i3, i4, c, ic, jc, io, jo, lbs, kbs, ipbs,
                          ibo=2 ** 15, mk=20, eps_c=1e-10, check_target=0):
        """MAT3 Add_integral; lbs/kbs/ipbs indexed by orbital (element 0 unused)"""
        L = lib()
        L.orc_add_integral_mat3.restype = C.c_int
        L.orc_add_integral_mat3.argtypes = ([C.c_void_p] + [C.c_int] * 6 + [C.c_double] + [C.c_int] * 4 +
                                            [_ip] * 3 + [C.c_int, C.c_int, C.c_double, C.c_int])
        ci = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        a, b, d = ci(lbs), ci(kbs), ci(ipbs)
        return L.orc_add_integral_mat3(self.h, icase, kpol, i1, i2, i3, i4, c, ic, jc, io, jo, _i(a), _i(b), _i(d),
                                       ibo, mk, eps_c, check_target)

    def __del__(self):
        try:
            lib().orc_cdata_free(self.h)
        except Exception:
            pass


# --------------------------------------------------------------------------
# the compiled reference (oracle/_ref), when present
# --------------------------------------------------------------------------
_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference build missing: " + REF_PATH)
        L = C.CDLL(REF_PATH)
        L.ref_setup.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _dp]
        L.ref_mkgrid.restype = C.c_int
        L.ref_mkgrid.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 5 + [C.c_int, C.c_double,
                                                                                C.c_double, C.c_int, _dp, C.c_int]
        L.ref_rk_rows.argtypes = [C.c_int, _ip, C.c_int, _dp, _dp]
        L.ref_breit_energy.restype = C.c_double
        L.ref_breit_energy.argtypes = [C.c_double]
        L.ref_read_grid.restype = C.c_int
        L.ref_read_grid.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _dp]
        L.ref_tables.argtypes = [_dp, _dp, _dp, _dp]
        L.ref_gauleg.argtypes = [C.c_int, _dp, _dp]
        L.ref_moments.argtypes = [C.c_int, C.c_int, _dp]
        L.ref_triang.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
        L.ref_mrk.argtypes = [C.c_int, C.c_int, _dp]
        L.ref_convol.argtypes = [C.c_int, C.c_int, C.c_int, C.c_char, C.c_char, _dp, _dp]
        L.ref_density.argtypes = [_dp, _dp, C.c_char, _dp]
        L.ref_sum_amb.restype = C.c_double
        L.ref_sum_amb.argtypes = [_dp, _dp, C.c_char]
        L.ref_rk.restype = C.c_double
        L.ref_rk.argtypes = [_dp, _dp, _dp, _dp, C.c_int]
        L.ref_quadr.restype = C.c_double
        L.ref_quadr.argtypes = [_dp, _dp, C.c_int]
        L.ref_update_hs.argtypes = [_dp, C.c_int, C.c_int, _dp, C.c_char]
        L.ref_overlaps.argtypes = [_dp, _dp, _dp]
        L.ref_set_orbitals.argtypes = [C.c_int, _dp]
        L.ref_sk.restype = C.c_double
        L.ref_sk.argtypes = [C.c_int] * 6
        L.ref_hf_setup.argtypes = [C.c_int] * 3 + [_ip] * 3 + [_dp, _dp, _dp, C.c_int] + [_ip] * 4 + [_dp]
        L.ref_hf_matrix.argtypes = [C.c_int, _dp]
        L.ref_full_mat_sym.argtypes = [C.c_int, C.c_int, _dp, _dp, C.c_char]
        L.ref_cdata_alloc.argtypes = [C.c_int] * 6 + [C.c_double, C.c_int]
        L.ref_cdata_add.argtypes = [C.c_double] + [C.c_int] * 6
        L.ref_cdata_flush.argtypes = [C.c_int, C.c_int]
        L.ref_cdata_nlog.restype = C.c_int
        L.ref_cdata_get.argtypes = [C.c_int] + [_ip] * 6 + [_dp, _ip]
        _ref = L
    return _ref


# --------------------------------------------------------------------------
# integral sets the way the reference evaluates them (rk(), Rk_evaluate):
# materialised arrays, one convol case 2 per distinct left density, one
# SUM_AmB per row.  Used as the CPU baseline and as the checker of sets.
# --------------------------------------------------------------------------
_SET_TYPES = {0: [(PPPP, 0, 0), (QQQQ, 1, 1), (QPQP, 1, 0), (PQPQ, 0, 1)],
              1: [(PPPP, 0, 0)], 2: [(QQQQ, 1, 1)], 3: [(PQPQ, 0, 1)], 4: [(QPQP, 1, 0)]}


def sort_rows_by_left_pair(rows):
    """rows with each pair put as (min, max) (R^k(a,b;c,d) = R^k(c,b;a,d) =
    R^k(a,d;c,b): 's' densities are symmetric) and sorted by (k, left pair,
    right pair), the order in which Rk_evaluate shares convolutions;
    returns (sorted rows, permutation)"""
    r = np.array(rows, dtype=np.int32).reshape(-1, 5)
    a, c = np.minimum(r[:, 1], r[:, 3]), np.maximum(r[:, 1], r[:, 3])
    b, d = np.minimum(r[:, 2], r[:, 4]), np.maximum(r[:, 2], r[:, 4])
    can = np.stack([r[:, 0], a, b, c, d], axis=1).astype(np.int32)
    perm = np.lexsort((can[:, 4], can[:, 2], can[:, 3], can[:, 1], can[:, 0]))
    return np.ascontiguousarray(can[perm]), perm


def rk_rows_par(g, pq, rows_sorted, nthreads, budget=0.0):
    """rk() values of rows sorted by sort_rows_by_left_pair, Rk_evaluate's
    organisation, OpenMP on nthreads threads; with a budget (seconds) only
    the leading rows finished in time are returned"""
    L = lib()
    L.orc_rk_rows_par.restype = C.c_int
    L.orc_rk_rows_par.argtypes = [C.c_void_p, C.c_int, _ip, _dp, C.c_int, C.c_double, _dp]
    r = np.ascontiguousarray(rows_sorted, dtype=np.int32)
    out = np.zeros(len(r))
    n = L.orc_rk_rows_par(g.h, len(r), _i(r), _d(f64(pq)), nthreads, budget, _d(out))
    return out[:n]


def ref_rk_rows(pq, rows_sorted, budget=0.0):
    """the same by the reference's own density/convol/SUM_AmB (oracle/_ref),
    one thread, in chunks of whole left pairs until the budget (seconds) is
    spent (0: all rows); ref_setup must have set the grid"""
    import time
    R = ref()
    r = np.ascontiguousarray(rows_sorted, dtype=np.int32)
    p = f64(pq)
    if budget <= 0:
        out = np.zeros(len(r))
        R.ref_rk_rows(len(r), _i(r), p.shape[2], _d(p), _d(out))
        return out
    starts = np.flatnonzero(np.concatenate([[True], np.any(r[1:, [0, 1, 3]] != r[:-1, [0, 1, 3]], axis=1)]))
    ends = np.concatenate([starts[1:], [len(r)]])
    outs = []
    t0 = time.perf_counter()
    for lo, hi in zip(starts, ends):
        sub = np.ascontiguousarray(r[lo:hi])
        o = np.zeros(hi - lo)
        R.ref_rk_rows(hi - lo, _i(sub), p.shape[2], _d(p), _d(o))
        outs.append(o)
        if time.perf_counter() - t0 > budget:
            break
    return np.concatenate(outs) if outs else np.zeros(0)


def rk_rows(g, fs, rows, kind=0, deadline=None):
    """values of rows (k,a,b,c,d), 1-based orbitals, for kind 0 (rk sum) or a
    single R^k type 1..4; fs: list of (ns,2) Fortran arrays.  With a deadline
    (time.perf_counter() value) stops after the row that passes it and
    returns only the rows done."""
    import time
    ns, ks = g.ns, g.ks
    out = np.zeros(len(rows))
    dens = {}
    conv = {}

    def d_of(a, c, comp):
        key = (min(a, c), max(a, c), comp)
        if key not in dens:
            dens[key] = density(ns, ks, fs[a][:, comp], fs[c][:, comp], "s")
        return dens[key]

    for n, (k, a, b, c, d) in enumerate(rows):
        a, b, c, d = int(a) - 1, int(b) - 1, int(c) - 1, int(d) - 1
        s = 0.0
        for t, cl, cr in _SET_TYPES[kind]:
            key = (t, int(k), min(a, c), max(a, c))
            if key not in conv:
                conv[key] = convol(ns, ks, g.mrk(t, int(k)), d_of(a, c, cl), 2, "s", "s")
            s += sum_amb(ns, ks, conv[key], d_of(b, d, cr), "s")
        out[n] = s
        if deadline is not None and time.perf_counter() > deadline:
            return out[:n + 1]
    return out


# --------------------------------------------------------------------------
# DBSR_HF's average-of-configuration energy (term = 'LS'/'AV'), restated:
# the Dirac operator with a finite nucleus (orc_dhl_pq: V_nuclear,
# Def_Vnucl, Def_pq_arrays, MAT_dhl_pq, DBSR_HF/dbsr_lib_dbs.f90:220-310,
# 1078-1206, 2756-2780) and Vp_dhl (:1210-1227); hf_fk / hf_gk
# (DBSR_HF/hf_rk.f90:43-113) with the reference's densities, convolutions
# and SUM_AmB in its order (hf_gk's cross term as 2 qpqp); av_energy_coef,
# WW, a, b, update_int, energy (DBSR_HF/hf_MOD_energy.f90:19-255);
# Conf_energies (DBSR_HF/hf_summry.f90:95-134).  The 3j symbol is Racah's
# formula in floating point (the closed forms of the tests pin it).
# --------------------------------------------------------------------------
def vnuclear(model, z, b, a, c, r):
    L = lib()
    L.orc_vnuclear.restype = C.c_double
    L.orc_vnuclear.argtypes = [C.c_int] + [C.c_double] * 5
    return L.orc_vnuclear(model, z, b, a, c, r)


def vnucl(g, model, z, b, a, c):
    L = lib()
    L.orc_vnucl.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 4 + [_dp]
    vr = np.zeros((g.nv, g.ks), order="F")
    L.orc_vnucl(g.h, model, z, b, a, c, _d(vr))
    return vr


def dhl_pq(g, model, z, b, a, c, kappa, bloch=1, rb=0.0, pnu=1.0):
    L = lib()
    L.orc_dhl_pq.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 4 + [C.c_int, C.c_int, C.c_double,
                                                                        C.c_double, _dp]
    d = np.zeros((g.ms, g.ms), order="F")
    L.orc_dhl_pq(g.h, model, z, b, a, c, kappa, bloch, rb, pnu, _d(d))
    return d


def threej_sq(j1, j2, j3, m1, m2, m3):
    """(j1/2 j2/2 j3/2; m1/2 m2/2 m3/2)^2, doubled arguments, floating point"""
    from math import factorial as fa
    if m1 + m2 + m3 != 0 or j3 > j1 + j2 or j3 < abs(j1 - j2) or (j1 + j2 + j3) % 2:
        return 0.0
    if any(abs(m) > j or (j + m) % 2 for j, m in ((j1, m1), (j2, m2), (j3, m3))):
        return 0.0
    a, b, c = (j1 + j2 - j3) // 2, (j1 - j2 + j3) // 2, (-j1 + j2 + j3) // 2
    tri = fa(a) * fa(b) * fa(c) / fa((j1 + j2 + j3) // 2 + 1)
    mm = 1.0
    for j, m in ((j1, m1), (j2, m2), (j3, m3)):
        mm *= fa((j + m) // 2) * fa((j - m) // 2)
    s = 0.0
    for t in range(0, a + 1):
        args = (t, (j3 - j2 + m1) // 2 + t, (j3 - j1 - m2) // 2 + t, a - t, (j1 - m1) // 2 - t, (j2 + m2) // 2 - t)
        if min(args) < 0:
            continue
        den = 1.0
        for x in args:
            den *= fa(x)
        s += (-1) ** t / den
    return tri * mm * s * s


def cjkj_sq(jj1, k, jj2):
    return threej_sq(jj1, 2 * k, jj2, 1, 0, -1) * (jj1 + 1) * (jj2 + 1)


def av_energy_coef(jbs, iqconf, weight):
    n, kmax = len(jbs), max(jbs)
    iq = np.asarray(iqconf, dtype=float)
    w = np.asarray(weight, dtype=float)
    coef = np.zeros((n, n, kmax + 1))
    for i in range(n):
        for j in range(i + 1):
            if i == j:
                c = float(np.sum(iq[:, i] * (iq[:, i] - 1) * w)) / 2.0
                coef[i, i, 0] = c
                for k in range(1, jbs[i] + 1):
                    coef[i, i, k] = -c * cjkj_sq(jbs[i], k, jbs[i]) / (jbs[i] * (jbs[i] + 1))
            else:
                c = float(np.sum(iq[:, i] * iq[:, j] * w))
                coef[i, j, 0] = c
                for k in range(abs(jbs[i] - jbs[j]) // 2, (jbs[i] + jbs[j]) // 2 + 1):
                    coef[j, i, k] = -c * cjkj_sq(jbs[i], k, jbs[j]) / ((jbs[i] + 1) * (jbs[j] + 1))
    return coef


def hf_fk(g, fs, i, j, k):
    ns, ks = g.ns, g.ks
    ppi, ppj = density(ns, ks, fs[i][:, 0], fs[i][:, 0], "s"), density(ns, ks, fs[j][:, 0], fs[j][:, 0], "s")
    qqi, qqj = density(ns, ks, fs[i][:, 1], fs[i][:, 1], "s"), density(ns, ks, fs[j][:, 1], fs[j][:, 1], "s")
    s = 0.0
    s += sum_amb(ns, ks, convol(ns, ks, g.mrk(PPPP, k), ppi, 2, "s", "s"), ppj, "s")
    s += sum_amb(ns, ks, convol(ns, ks, g.mrk(QQQQ, k), qqi, 2, "s", "s"), qqj, "s")
    s += sum_amb(ns, ks, convol(ns, ks, g.mrk(QPQP, k), qqi, 2, "s", "s"), ppj, "s")
    s += sum_amb(ns, ks, convol(ns, ks, g.mrk(PQPQ, k), ppi, 2, "s", "s"), qqj, "s")
    return s


def hf_gk(g, fs, i, j, k):
    ns, ks = g.ns, g.ks
    pp = density(ns, ks, fs[i][:, 0], fs[j][:, 0], "s")
    qq = density(ns, ks, fs[i][:, 1], fs[j][:, 1], "s")
    s = 0.0
    s += sum_amb(ns, ks, convol(ns, ks, g.mrk(PPPP, k), pp, 2, "s", "s"), pp, "s")
    s += sum_amb(ns, ks, convol(ns, ks, g.mrk(QQQQ, k), qq, 2, "s", "s"), qq, "s")
    s += 2.0 * sum_amb(ns, ks, convol(ns, ks, g.mrk(QPQP, k), qq, 2, "s", "s"), pp, "s")
    return s


def av_energy(g, pq, kbs, lbs, jbs, nucleus, iqconf, weight):
    """(Etotal, E1body, E2body, [configuration energies]) of the orbitals
    pq(ns,2,nbf); nucleus = (model, z, r_uniform, a_fermi, c_fermi)"""
    n = len(kbs)
    fs = [np.asfortranarray(pq[:, :, i]) for i in range(n)]
    valueI = np.zeros(n)
    dhl = {}
    for i in range(n):
        if kbs[i] not in dhl:
            dhl[kbs[i]] = dhl_pq(g, *nucleus, kbs[i])
        v = np.concatenate([fs[i][:, 0], fs[i][:, 1]])
        valueI[i] = float(np.sum(v * (dhl[kbs[i]] @ v)))
    F = {(i, j, k): hf_fk(g, fs, i, j, k) for i in range(n) for j in range(i + 1)
         for k in range(0, 2 * min(lbs[i], lbs[j]) + 1, 2)}
    G = {(i, j, k): hf_gk(g, fs, i, j, k) for i in range(n) for j in range(i)
         for k in range(abs(lbs[i] - lbs[j]), lbs[i] + lbs[j] + 1, 2)}

    def energy(iq, w):
        w = np.asarray(w, dtype=float) / np.sum(w)
        qsum = np.asarray(iq, dtype=float).T @ w
        coef = av_energy_coef(jbs, iq, w)
        et = e1 = e2 = 0.0
        for i in range(n):
            et += qsum[i] * valueI[i]
            e1 += qsum[i] * valueI[i]
            for j in range(i + 1):
                for k in range(0, 2 * min(lbs[i], lbs[j]) + 1, 2):
                    if k % 2 or k > (jbs[i] + jbs[j]) // 2:
                        continue
                    c = coef[i, j, k]
                    if c != 0.0:
                        et += c * F[(i, j, k)]
                        e2 += c * F[(i, j, k)]
            for j in range(i):
                for k in range(abs(lbs[i] - lbs[j]), lbs[i] + lbs[j] + 1, 2):
                    if (k + lbs[i] + lbs[j]) % 2 or k < abs(jbs[i] - jbs[j]) // 2 or k > (jbs[i] + jbs[j]) // 2:
                        continue
                    c = coef[j, i, k]
                    if c != 0.0:
                        et += c * G[(i, j, k)]
                        e2 += c * G[(i, j, k)]
        return et, e1, e2

    iq = np.asarray(iqconf)
    et, e1, e2 = energy(iq, weight)
    confs = []
    for ic in range(len(iq)):
        w = np.zeros(len(iq))
        w[ic] = 1.0
        confs.append(energy(iq, w)[0])
    return et, e1, e2, confs, valueI, F, G


def ref_av_energy(pq, kbs, lbs, jbs, nucleus, iqconf, weight):
    """the same by the compiled reference (ref_setup must have set the
    grid): ref_hf_setup for the orbitals, ref_set_nucleus, then DBSR_HF's
    av_energy_coef / update_int / energy / Conf_energies"""
    R = ref()
    R.ref_set_nucleus.argtypes = [C.c_int] + [C.c_double] * 4
    R.ref_av_energy.argtypes = [C.c_int, _ip, _dp, _dp]
    n = len(kbs)
    kmax = max(jbs)
    coef0 = np.zeros((n, n, kmax + 1), order="F")
    z = np.zeros(1, dtype=np.int32)
    R.ref_hf_setup(n, kmax, 0, _i(np.asarray(kbs, dtype=np.int32)), _i(np.asarray(lbs, dtype=np.int32)),
                   _i(np.asarray(jbs, dtype=np.int32)), _d(np.ones(n)), _d(coef0), _d(f64(pq)), 0,
                   _i(z), _i(z), _i(z), _i(z), _d(np.zeros(1)))
    R.ref_set_nucleus(*nucleus)
    iq = np.asfortranarray(np.asarray(iqconf, dtype=np.int32))
    w = np.asarray(weight, dtype=float)
    out = np.zeros(3 + len(w))
    R.ref_av_energy(len(w), _i(iq), _d(w), _d(out))
    return out[0], out[1], out[2], list(out[3:])


def ref_dhl_pq(kappa, nucleus, ms):
    R = ref()
    R.ref_set_nucleus.argtypes = [C.c_int] + [C.c_double] * 4
    R.ref_dhl_pq.argtypes = [C.c_int, _dp]
    R.ref_set_nucleus(*nucleus)
    d = np.zeros((ms, ms), order="F")
    R.ref_dhl_pq(kappa, _d(d))
    return d


def ref_vnuclear(nucleus, r):
    R = ref()
    R.ref_set_nucleus.argtypes = [C.c_int] + [C.c_double] * 4
    R.ref_vnuclear.restype = C.c_double
    R.ref_vnuclear.argtypes = [C.c_double]
    R.ref_set_nucleus(*nucleus)
    return R.ref_vnuclear(r)


# --------------------------------------------------------------------------
# DBSR_MCHF's orbital matrices, restated with the primitives above:
# mchf_matrix (DBSR_MCHF/mchf_matrix.f90:2-248) and core_matrix (:253-345).
# The integral list of the reference (i1_int..i4_int, ip_int, Rk_coef with
# the level factor WW(ic,jc), :89-95) arrives flattened: one entry per
# angular coefficient, c = WW*coef.  The collection list id4_data lives in
# the un-vendored Libraries_03; it is restated as a list in order of first
# insertion whose equal keys (k, id2, id3, id4) sum their coefficients.
# Only the Newton term depends on that order (below).
# --------------------------------------------------------------------------
_MCHF_BLOCKS = ((1, 1, PPPP), (1, 2, PQPQ), (2, 1, QPQP), (2, 2, QQQQ))


def id4_lists(i, rlist, S, eps_c, kmax):
    """per k: the (id2, id3, id4, cid) list mchf_matrix collects for orbital
    i (mchf_matrix.f90:89-144); flags -1 direct, -2 exchange, -3 Newton, >0
    the rhs orbital"""
    out = []
    for k in range(kmax + 1):
        keys, vals = [], {}

        def add(id2, id3, id4, c):
            key = (id2, id3, id4)
            if key in vals:
                vals[key] += c
            else:
                keys.append(key)
                vals[key] = c

        for kk, i1, i2, i3, i4, c0 in rlist:
            if kk != k:
                continue
            C_ = c0 / S
            if abs(C_) < eps_c:
                continue
            ii = (1000 if i == i1 else 0) + (100 if i == i2 else 0) + (10 if i == i3 else 0) + (1 if i == i4 else 0)
            if ii == 1111:
                add(-1, i1, i3, 2 * C_); add(-3, i1, i3, 4 * C_)
            elif ii == 1010:
                add(-1, i2, i4, C_)
            elif ii == 101:
                add(-1, i1, i3, C_)
            elif ii == 1001:
                add(-2, i2, i3, C_)
            elif ii == 1100:
                add(-2, i3, i4, C_)
            elif ii == 110:
                add(-2, i1, i4, C_)
            elif ii == 11:
                add(-2, i1, i2, C_)
            elif ii in (111, 1101):
                add(-1, i1, i3, C_); add(-2, i1, i3, C_ / 2); add(-3, i1, i3, 3 * C_ / 2)
            elif ii in (1011, 1110):
                add(-1, i2, i4, C_); add(-2, i2, i4, C_ / 2); add(-3, i2, i4, 3 * C_ / 2)
            elif ii == 1000:
                add(i3, i2, i4, C_ / 2)
            elif ii == 100:
                add(i4, i1, i3, C_ / 2)
            elif ii == 10:
                add(i1, i2, i4, C_ / 2)
            elif ii == 1:
                add(i2, i1, i3, C_ / 2)
        out.append([(a, b, c, vals[(a, b, c)]) for (a, b, c) in keys])
    return out


def mchf_qsum(i, llist, norb, qsum_min=1e-8):
    """S = qsum(i) and CI(io) from the L-integral coefficients
    (mchf_matrix.f90:59-72)"""
    S = 0.0
    CI = np.zeros(norb)
    for i1, i2, c in llist:
        if i1 == i and i2 == i:
            S = S + c
        elif i1 == i:
            CI[i2 - 1] += c
        elif i2 == i:
            CI[i1 - 1] += c
    if abs(S) < qsum_min:
        S = qsum_min
    return S, CI


def mchf_matrix(g, i, pq, dhl, llist, rlist, kmax, eps_c=1e-8, qsum_min=1e-8, newton=0, irhs=0, fppqq=None):
    """mchf_matrix for orbital i (1-based): llist (i1, i2, c), rlist
    (k, i1, i2, i3, i4, c) with c = WW*coef; dhl (ms,ms) = Get_hd_core(kbs(i)).
    Returns hfm, rhs, hx, e(i), qsum(i).

    The Newton term hx of a (k, block) is, as the reference computes it
    (mchf_matrix.f90:220-234), the case-4 convolution of the LAST collected
    -3 density, without its coefficient: Convol's output and input are
    xx and x there, so the accumulated xx is overwritten."""
    ns, ks, ms = g.ns, g.ks, g.ms
    norb = pq.shape[2]
    hfm = np.zeros((ms, ms), order="F")
    hx = np.zeros((ms, ms), order="F")
    rhs = np.zeros(ms)
    S, CI = mchf_qsum(i, llist, norb, qsum_min)
    v_of = lambda o: np.concatenate([pq[:, 0, o - 1], pq[:, 1, o - 1]])
    if np.sum(np.abs(CI)) != 0.0:
        vv = np.zeros(ms)
        for io in range(1, norb + 1):
            if CI[io - 1] == 0.0:
                continue
            vv = vv + v_of(io) * CI[io - 1] / S / 2.0
        rhs = rhs + np.asarray(dhl) @ vv
    lists = id4_lists(i, rlist, S, eps_c, kmax)
    for k in range(kmax + 1):
        lst = lists[k]
        if not lst:
            continue
        for ii, jj, typ in _MCHF_BLOCKS:
            rkb = g.mrk(typ, k)
            dd, met = np.zeros((ns, ks), order="F"), False
            for id2, id3, id4, cid in lst:
                if id2 != -1 or cid == 0.0:
                    continue
                dd = dd + cid * density(ns, ks, pq[:, jj - 1, id3 - 1], pq[:, jj - 1, id4 - 1], "s")
                met = True
            if met:
                d = convol(ns, ks, rkb, dd, 1, "s", "s")
                update_hs_lib(hfm, ii, ii, ns, ks, d, "s")
            xx, met = np.zeros((ns, ns), order="F"), False
            for id2, id3, id4, cid in lst:
                if id2 != -2 or cid == 0.0:
                    continue
                xx = xx + cid * density(ns, ks, pq[:, ii - 1, id3 - 1], pq[:, jj - 1, id4 - 1], "x")
                met = True
            if met:
                x = convol(ns, ks, rkb, xx, 4, "s", "s")
                update_hs_lib(hfm, ii, jj, ns, ks, x, "x")
            aa = {}
            for id2, id3, id4, cid in lst:
                if id2 <= 0:
                    continue
                d = density(ns, ks, pq[:, jj - 1, id3 - 1], pq[:, jj - 1, id4 - 1], "s")
                aa[id2] = (aa[id2] + d * cid) if id2 in aa else (np.zeros((ns, ks), order="F") + d * cid)
            for j in sorted(aa):
                dd = convol(ns, ks, rkb, aa[j], 1, "s", "s")
                y = full_mat_sym(ns, ks, dd, "s")
                v1 = y @ pq[:, ii - 1, j - 1]
                rhs[(ii - 1) * ns:ii * ns] = rhs[(ii - 1) * ns:ii * ns] + v1
            if newton:
                last = None
                for id2, id3, id4, cid in lst:
                    if id2 != -3 or cid == 0.0:
                        continue
                    last = (id3, id4)
                if last is not None:
                    x = density(ns, ks, pq[:, jj - 1, last[0] - 1], pq[:, jj - 1, last[1] - 1], "x")
                    xx = convol(ns, ks, rkb, x, 4, "s", "s")
                    update_hs_lib(hx, ii, jj, ns, ks, xx, "x")
    srhs = np.sum(np.abs(rhs))
    if srhs < 1e-8:
        rhs[:] = 0.0
        srhs = 0.0
    if irhs > 0 and srhs > 0.0:
        v = v_of(i)
        B = np.asarray(fppqq)
        SS = -np.sum(rhs * v)
        hfm = hfm + SS * B
        vv = B @ v
        hfm = hfm + np.outer(rhs, vv) + np.outer(vv, rhs)
        rhs[:] = 0.0
    hfm = np.asfortranarray(np.asarray(dhl) + hfm)
    rhs = -rhs
    v = v_of(i)
    e = float(np.dot(v, hfm @ v - rhs))
    return hfm, rhs, hx, e, S


def core_matrix(g, i, pq, hd, acoef, bcoef, kmax, qsum, newton=0):
    """core_matrix for core orbital i (1-based): hd (ms,ms) = Get_hd(kbs(i)),
    acoef/bcoef (norb, kmax+1) = a(i,j,k), b(i,j,k), qsum (norb).  Returns
    hfm, hx."""
    ns, ks, ms = g.ns, g.ks, g.ms
    norb = pq.shape[2]
    q = qsum[i - 1]
    hfm = np.asfortranarray(q * np.asarray(hd))
    hx = np.zeros((ms, ms), order="F")
    for ii, jj, typ in _MCHF_BLOCKS:
        for k in range(kmax + 1):
            rkb = g.mrk(typ, k)
            dd, met = np.zeros((ns, ks), order="F"), False
            for j in range(1, norb + 1):
                c = acoef[j - 1, k]
                c = c * q * qsum[j - 1] if i != j else c * q * (q - 1)
                if c == 0.0:
                    continue
                dd = dd + c * density(ns, ks, pq[:, jj - 1, j - 1], pq[:, jj - 1, j - 1], "s")
                met = True
            if met:
                d = convol(ns, ks, rkb, dd, 1, "s", "s")
                update_hs_lib(hfm, ii, ii, ns, ks, d, "s")
            xx, met = np.zeros((ns, ns), order="F"), False
            for j in range(1, norb + 1):
                c = bcoef[j - 1, k]
                c = c * q * qsum[j - 1] if i != j else c * q * (q - 1)
                if c == 0.0:
                    continue
                xx = xx + density(ns, ks, pq[:, ii - 1, j - 1], pq[:, jj - 1, j - 1], "x") * c
                met = True
            if met:
                x = convol(ns, ks, rkb, xx, 4, "s", "s")
                update_hs_lib(hfm, ii, jj, ns, ks, x, "x")
            if newton:
                c = 2.0 * acoef[i - 1, k]
                if c == 0.0:
                    continue
                c = c * q * (q - 1)
                x = density(ns, ks, pq[:, ii - 1, i - 1], pq[:, jj - 1, i - 1], "x")
                x = convol(ns, ks, rkb, c * x, 4, "s", "s")
                update_hs_lib(hx, ii, jj, ns, ks, x, "x")
    return np.asfortranarray(hfm / q), np.asfortranarray(hx / q)


# --------------------------------------------------------------------------
# One-electron (L) integrals of DBSR_MAT3, restated: Idef_Ltype
# (DBSR_MAT3/idef_type.f90:54-138) and L_data (DBSR_MAT3/L_data.f90:2-145).
# Lval and Gen_Lvec are in the un-vendored Libraries_03: Lval comes as a
# table lval[i-1, j-1]; Gen_Lvec(i, v) is restated as v = hd(kappa of i)
# (P_i, Q_i), the one-electron operator of the orbital's kappa applied to it.
# Tables follow o_data's conventions: ipbs / iptar 1-based (element 0 unused).
# --------------------------------------------------------------------------
def idef_ltype(ii, jj, C_, ic, jc, io, jo, ipbs, eps_c=1e-10, check_target=1):
    """(itype, (k1, k2, k3, k4)), itype 0 when dropped; raises on the
    reference's Stop conditions"""
    if abs(C_) < eps_c:
        return 0, None
    if jc > 0 and ic <= 0:
        raise ValueError("jc>0, ic<=0")
    if jo > 0 and io <= 0:
        raise ValueError("jo>0, io<=0")
    if io > 0 and jc != 0:
        raise ValueError("jc>0, io>0")
    if jo > 0 and ic != 0:
        raise ValueError("jc>0, jo>0")
    i, j = min(ii, jj), max(ii, jj)
    ich, jch = ipbs[i], ipbs[j]
    t, key = 0, None
    if jc > 0:
        t, key = 1, (i, j, ic, jc)
    elif ic > 0:
        if io == 0:
            if ich != 0 and jch == 0:
                t, key = 2, (j, ich, ic, 0)
            elif ich == 0 and jch != 0:
                t, key = 2, (i, jch, ic, 0)
        elif io > 0:
            t, key = 1, (i, j, -io, ic)
    elif io == 0:
        t, key = 3, (j, i, 0, 0)
    elif io > 0 and jo == 0:
        if ich != 0 and jch == 0:
            t, key = 2, (j, ich, -io, 0)
        elif jch != 0 and ich == 0:
            t, key = 2, (i, jch, -io, 0)
        elif ich == 0 and jch == 0:
            t, key = 1, (i, j, -io, 0)
            if check_target == 0:
                return 0, None
    elif io > 0 and jo > 0:
        t, key = 1, (i, j, -io, -jo)
    if t == 0:
        raise ValueError("Idef_Ltype: itype=0")
    return t, key


def _target(mats, iptar, tgt, ich, jch, C_):
    if mats["icc"][ich - 1, jch - 1] == 0 or tgt is None:
        return
    it, jt = iptar[ich], iptar[jch]
    hi, lo = max(it, jt), min(it, jt)
    tgt[(hi - 1) * hi // 2 + lo - 1] += C_


def l_data(ms, itype, lst, eps_c, ibo, ipbs, kbort, vch, fppqq, iitar, iptar, htarg, otarg, check_target,
           lval, kbs, kap_list, hd, pq, mats):
    """L_data(itype) of lst = (c, k1, k2, k3, k4) in IPT order into mats =
    dict(icc, hch, icb, hcp, ibb, hp) (updated in place); hd (ms,ms,nkap);
    pq (ns,2,norb); kbs 1-based (element 0 unused)"""
    c, k1, k2, k3, k4 = lst
    hch, hcp, hp = mats["hch"], mats["hcp"], mats["hp"]
    icc, icb, ibb = mats["icc"], mats["icb"], mats["ibb"]

    def blk_add(ich, jch, x):
        b = icc[ich - 1, jch - 1]
        if b:
            hch[:, :, b - 1] += x

    def hx(ich, jch, d, C_):
        blk_add(ich, jch, C_ * (d if ich >= jch else d.T))

    def hw(ich, jch, v, w, C_):
        x = np.outer(v, w)
        blk_add(ich, jch, C_ * x if ich > jch else (C_ * x.T if jch > ich else (C_ / 2) * (x + x.T)))

    def hv(ich, ip, v, C_):
        col = icb[ich - 1, ip - 1]
        if col:
            hcp[:, col - 1] += C_ * v

    def vcol(io):
        i1, i2 = divmod(int(io), ibo)
        k = kbort(i1, i2)
        if k == 0:
            raise ValueError("O: KBORT=0")
        return i1, vch[:, k - 1]

    def kap(o):
        return list(kap_list).index(kbs[o]) if kbs[o] in list(kap_list) else -1

    for e in range(len(c)):
        C_ = c[e]
        if abs(C_) < eps_c:
            continue
        if itype == 1:
            i1, i2, ic, jc = int(k1[e]), int(k2[e]), int(k3[e]), int(k4[e])
            C_ = C_ * lval[i1 - 1, i2 - 1]
            if ic > 0 and jc > 0:
                b = ibb[ic - 1, jc - 1]
                if b:
                    hp[b - 1] += C_
            elif ic < 0 < jc:
                o1, v = vcol(-ic)
                hv(ipbs[o1], jc, v, C_)
            elif ic < 0 and jc < 0:
                o1, v = vcol(-ic)
                o2, w = vcol(-jc)
                hw(ipbs[o1], ipbs[o2], v, w, C_)
            elif ic < 0 and jc == 0:
                o1, o2 = divmod(-ic, ibo)
                ich, jch = ipbs[o1], ipbs[o2]
                if iitar > 0 and ich != jch:
                    hx(ich, jch, fppqq, C_)
                _target(mats, iptar, htarg, ich, jch, C_)
            else:
                raise ValueError("L_data:  wrong combination of ic,jc for itype=1")
        elif itype == 2:
            o1, ich, jc = int(k1[e]), int(k2[e]), int(k3[e])
            q = kap(o1)
            if q < 0:
                raise ValueError("L_data: unknown kappa, itype=2")
            v = hd[:, :, q] @ np.concatenate([pq[:, 0, o1 - 1], pq[:, 1, o1 - 1]])
            if jc > 0:
                hv(ich, jc, v, C_)
            elif jc < 0:
                j1, w = vcol(-jc)
                hw(ich, ipbs[j1], v, w, C_)
            else:
                raise ValueError("L_data:  wrong value jc = 0 for itype=2")
        else:
            o1, o2 = int(k1[e]), int(k2[e])
            ich, jch = ipbs[o1], ipbs[o2]
            if kbs[o1] != kbs[o2]:
                raise ValueError("L_data: different kappa, itype=3")
            q = kap(o1)
            if q < 0:
                raise ValueError("L_data: unknown kappa, itype=3")
            if ich == jch or iitar > 1:
                hx(ich, jch, hd[:, :, q], C_)
            elif check_target == 1:
                _target(mats, iptar, otarg, ich, jch, C_)


# --------------------------------------------------------------------------
# DBSR_HF's orbital update, restated: Boundary_conditions
# (DBSR_HF/hf_boundary.f90:2-32), orthogonality (hf_orthogonality.f90:2-29)
# and hf_eiv (hf_solve_HF.f90:118-202), the generalised eigenproblem by
# LAPACK through scipy (dsygvd; the reference calls dsygv: the same
# problem, another algorithm).
# --------------------------------------------------------------------------
C_AU = 137.03599976


def boundary_conditions(ns, nsp, nsq, ksp, ksq, lbs, kbs, nuclear="Fermi", ilzero=1, ibzero=2):
    """iprm(ms, nbf) of 0/1 (0 = B-spline removed)"""
    nbf = len(lbs)
    ms = 2 * ns
    iprm = np.ones((ms, nbf), dtype=np.int32)
    for i in range(nbf):
        if nuclear == "point" or ilzero == 0:
            iprm[0, i] = 0
            iprm[nsp - ibzero:ns, i] = 0
            iprm[ns, i] = 0
            iprm[nsq - ibzero + ns:ms, i] = 0
        else:
            j = lbs[i] + 1
            if j > ksp - 1:
                j = 1
            iprm[0:j, i] = 0
            iprm[nsp - ibzero:ns, i] = 0
            j = lbs[i] + 2 if kbs[i] < 0 else lbs[i]
            if j > ksq - 1:
                j = 1
            iprm[ns:ns + j, i] = 0
            iprm[nsq - ibzero + ns:ms, i] = 0
    return iprm


def hf_eiv(hfm, vorth, iprm, m, fppqq):
    """(v, e): the orbital update of hf_eiv for the matrix hfm (ms,ms), the
    vectors to project out vorth (list of ms), iprm (ms) and m"""
    import scipy.linalg
    h = np.array(hfm, dtype=np.float64)
    B = np.asarray(fppqq)
    ms = h.shape[0]
    for v in vorth:
        w = B @ v
        c = -np.outer(w, v) + np.eye(ms)
        h = c @ h
        h = h @ c.T
    keep = np.nonzero(np.asarray(iprm))[0]
    aa = h[np.ix_(keep, keep)]
    ss = B[np.ix_(keep, keep)]
    ev, z = scipy.linalg.eigh(aa, ss, lower=True)
    zz = -1.999 * C_AU ** 2
    mm = 0
    for j in range(1, len(ev) + 1):
        mm = j
        if ev[j - 1] > zz:
            break
    mm = m + mm - 1
    v = np.zeros(ms)
    v[keep] = z[:, mm - 1]
    if v[int(np.argmax(np.abs(v)))] < 0.0:
        v = -v
    return v, float(ev[mm - 1])


# --------------------------------------------------------------------------
# DBSR_MAT3's channel-block post-processing, restated: Zero_cond
# (DBSR_MAT3/zero_cond.f90:2-56), the orthogonality conditions of DBS_ORTH
# (dbs_orth.f90:2-84), Diag_channels (diag.f90:2-79; the eigen-solve is
# LAPACK's DSYGV(1,'V','L') through scipy, the routine the reference calls)
# and Record_matrix (record_matrix.f90:2-60).  Matrices hold the reference's
# layouts: hch(ms,ms,nblocks), diag(ms,ms,nd), hcp(ms,ncb), hp(nbb);
# icc/icb/ibb 1-based block numbers, 0 = not stored.
# --------------------------------------------------------------------------
EDMIN = -2 * C_AU * C_AU   # mod_dbsr_mat.f90:81
EDMAX = 1.0e10             # :80
EGAP = 0.001               # :82


def zero_cond(nch, ns, nsp, nsq, ksp, ksq, lch, kch, ilzero=-1, jlzero=-1, ibzero=1, jbzero=1, npert=0):
    """(iprm(mhm) of 0/1, nhm): zero conditions at r = 0 and at the boundary
    per channel (lch, kch: l and kappa of channel ich at element ich-1)"""
    ms = 2 * ns
    iprm = np.ones(nch * ms + npert, dtype=np.int32)
    ip = jp = 0
    for ich in range(nch):
        mi = 0
        if ilzero >= 0:
            il = ilzero
        else:
            il = lch[ich] + 1
            if il > ksp - 1:
                il, mi = 1, 1
        mj = 0
        if jlzero >= 0:
            jl = jlzero
        else:
            if kch[ich] < 0:
                jl = lch[ich] + 2
            elif kch[ich] >= 1:
                jl = lch[ich]
            else:
                raise ValueError("zero_cond: kappa = 0")
            if jl > ksq - 1:
                jl, mj = 1, 1
        if mi == 1 or mj == 1:
            il, jl = 1, 1
        for lo, mb in ((il, nsp + 1 - ibzero), (jl, nsq + 1 - jbzero)):
            for i in range(1, ns + 1):
                ip += 1
                if i <= lo:
                    iprm[ip - 1] = 0
                if i >= mb:
                    iprm[ip - 1] = 0
                jp += iprm[ip - 1]
    return iprm, jp + npert


def orth_conditions(nch, ipch, ipbs, kbs, ibort):
    """per channel the 1-based bound orbitals j (ascending) it is kept
    orthogonal to: ipbs(j) = 0, kbs(j) = kbs(ipch(ich)), IBORT(i,j) = 0
    (dbs_orth.f90:36-42); ipch/ipbs/kbs indexed by orbital - 1"""
    out = []
    for ich in range(nch):
        i = ipch[ich]
        out.append([j for j in range(1, len(kbs) + 1)
                    if ipbs[j - 1] <= 0 and kbs[j - 1] == kbs[i - 1] and ibort(i, j) == 0])
    return out


def dbs_orth(hch, icc, idiag, orth, pv, qv, hcp=None, icb=None, npert=0):
    """hch (and hcp) updated in place; orth[ich-1] = the orbitals of channel
    ich (orth_conditions); pv[:, j-1], qv[:, j-1] = Get_pv / Get_qv of
    orbital j"""
    nch = icc.shape[0]
    ms = hch.shape[0]
    for ich in range(1, nch + 1):
        if not orth[ich - 1]:
            continue
        aa = np.zeros((ms, ms))
        bb = np.zeros((ms, ms))
        for j in orth[ich - 1]:
            S = np.outer(qv[:, j - 1], pv[:, j - 1])
            aa += S
            bb += S.T
        for jch in range(1, nch + 1):
            if idiag == 1 and ich != jch:
                continue
            if idiag == -1 and ich == jch:
                continue
            ij = icc[ich - 1, jch - 1]
            if ij == 0:
                continue
            x = hch[:, :, ij - 1].copy()
            if ich > jch:
                x = x - aa @ x
            elif ich == jch:
                x = x - aa @ x
                x = x - x @ bb
            else:
                x = x - x @ bb
            hch[:, :, ij - 1] = x
        if npert > 0 and idiag == 0:
            for ic in range(1, npert + 1):
                jc = icb[ich - 1, ic - 1]
                if jc == 0:
                    continue
                vv = hcp[:, jc - 1] @ bb
                hcp[:, jc - 1] = hcp[:, jc - 1] - vv


def diag_channels(hch, diag, icc, iprm, ns, edmin=EDMIN, edmax=EDMAX, egap=EGAP):
    """diag (ms,ms,nd) updated in place: per diagonal block k = icc(ich,ich)
    the solutions of H x = E S x on the B-splines with iprm = 1 (S = the
    block's overlap in diag(:,:,k)), those with edmin <= E <= edmax and
    |E| >= egap in columns 1..ksol (largest entry positive), E in
    diag(ksol,ns,k).  Returns ipsol(0:nch) with ipsol[0] = 0."""
    from scipy.linalg.lapack import dsygv
    nch = icc.shape[0]
    ms = 2 * ns
    ipsol = np.zeros(nch + 1, dtype=np.int32)
    for ich in range(1, nch + 1):
        k = icc[ich - 1, ich - 1]
        if k == 0:
            continue
        keep = np.nonzero(np.asarray(iprm[(ich - 1) * ms:ich * ms]))[0]
        a = np.asfortranarray(hch[np.ix_(keep, keep, [k - 1])][:, :, 0])
        b = np.asfortranarray(diag[np.ix_(keep, keep, [k - 1])][:, :, 0])
        w, v, info = dsygv(a, b, itype=1, jobz="V", uplo="L")
        if info != 0:
            raise RuntimeError("channel diagonalization failed")
        ksol = 0
        for i in range(len(keep)):
            if w[i] < edmin or w[i] > edmax or abs(w[i]) < egap:
                continue
            ksol += 1
            x = np.zeros(ms)
            x[keep] = v[:, i]
            if x[int(np.argmax(np.abs(x)))] < 0.0:
                x = -x
            diag[:, ksol - 1, k - 1] = x
            diag[ksol - 1, ns - 1, k - 1] = w[i]
        ipsol[ich] = ksol
    return ipsol


def record_matrix(hch, icc, diag, ipsol, icase, hcp=None, icb=None, hp=None, ibb=None, npert=0):
    """(records, overlaps): the records in the order Record_matrix writes
    them, ((i, j), values) with values c(1:ni,1:nj), w(1:ni) or hp(k), ending
    with ((0, 0), None); overlaps (nch+npert, nch+npert), set for icase 0.
    diag(:,:,ich) holds channel ich's solutions."""
    nch = icc.shape[0]
    recs = []
    ov = np.zeros((nch + npert, nch + npert))
    for ich in range(2, nch + 1):
        ni = ipsol[ich]
        for jch in range(1, ich):
            nj = ipsol[jch]
            k = icc[ich - 1, jch - 1]
            if k == 0:
                continue
            h = hch[:, :, k - 1]
            if np.max(np.abs(h)) < 1e-20:
                continue
            w = diag[:, :ni, ich - 1].T @ h
            c = w @ diag[:, :nj, jch - 1]
            recs.append(((ich, jch), c))
            if icase == 0:
                ov[ich - 1, jch - 1] = np.max(np.abs(c)) if c.size else 0.0
    if npert > 0:
        for ich in range(1, nch + 1):
            ni = ipsol[ich]
            for ip in range(1, npert + 1):
                k = icb[ich - 1, ip - 1]
                if k == 0:
                    continue
                if np.max(np.abs(hcp[:, k - 1])) < 1e-20:
                    continue
                w = diag[:, :ni, ich - 1].T @ hcp[:, k - 1]
                if icase == 0:
                    ov[nch + ip - 1, ich - 1] = np.max(np.abs(w)) if w.size else 0.0
                recs.append(((nch + ip, ich), w))
        for ip in range(1, npert + 1):
            for jp in range(1, ip + 1):
                k = ibb[ip - 1, jp - 1]
                if k == 0 or hp[k - 1] == 0.0:
                    continue
                recs.append(((ip + nch, jp + nch), hp[k - 1]))
                if icase == 0:
                    ov[nch + ip - 1, nch + jp - 1] = abs(hp[k - 1])
    recs.append(((0, 0), None))
    return recs, ov
