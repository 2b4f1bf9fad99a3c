"""DBSR_MAT3's channel-block post-processing with the matrices on the GPU.

The host mirror of the steps that follow the accumulation of one partial
wave (DBSR_MAT3/sub1.f90:133-170, 258-265), with the reference's names and
argument meaning:

* ``Zero_cond`` (DBSR_MAT3/zero_cond.f90:2-56): the B-splines removed at
  r = 0 and at the boundary, iprm(mhm) and nhm;
* ``DBS_ORTH`` (dbs_orth.f90:2-84): the orthogonality conditions of the
  channels to the bound orbitals of their kappa (IBORT(i,j) = 0), imposed on
  the blocks as (1 - B p p') H (1 - p p' B);
* ``Diag_channels`` (diag.f90:2-79): the diagonal blocks' generalised
  eigen-problems on the kept B-splines, the solutions in the energy window
  stored in diag with their energies in column ns, ipsol;
* ``Record_matrix`` (record_matrix.f90:2-60): the other blocks transformed
  to the channel solutions, returned as the records the reference writes.

All matrix work runs in libdbsr3_amd.so (dbs_channels.hip: rocSOLVER batched
eigen-solves, rocBLAS batched products, HIP kernels); this module only
prepares host tables.  Device arrays are torch float64 tensors holding the
Fortran layouts (a Fortran a(ms,ms,n) is a tensor of shape (n, ms, ms)).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import IBORT_FN, check, lib

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_longlong)

C_AU = 137.03599976            # zconst c_au
EDMIN = -2.0 * C_AU * C_AU     # mod_dbsr_mat.f90:81
EDMAX = 1.0e10                 # :80
EGAP = 0.001                   # :82


def _vp(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def Zero_cond(nch, grid, lch, kch, ilzero=-1, jlzero=-1, ibzero=1, jbzero=1, npert=0):
    """(iprm, nhm): iprm(nch*ms + npert) of 0/1 for the channels' (l, kappa);
    grid: anything with the grid's ns, nsp, nsq, ksp, ksq (a DBS)"""
    lch, kch = _i32(lch), _i32(kch)
    assert len(lch) >= nch and len(kch) >= nch
    iprm = np.zeros(nch * 2 * grid.ns + npert, dtype=np.int32)
    nhm = C.c_int(0)
    check(lib().dbs_zero_cond(nch, grid.ns, grid.nsp, grid.nsq, grid.ksp, grid.ksq, lch.ctypes.data_as(_ip),
                              kch.ctypes.data_as(_ip), ilzero, jlzero, ibzero, jbzero, npert,
                              iprm.ctypes.data_as(_ip), C.byref(nhm)))
    return iprm, nhm.value


def orth_conditions(ipch, ipbs, kbs, ibort, ldo=None):
    """(nort, iort): per channel the bound orbitals (1-based, ascending) it is
    kept orthogonal to; ipch/ipbs/kbs are indexed by channel/orbital - 1,
    ibort(i, j) is the host program's IBORT"""
    ipch, ipbs, kbs = _i32(ipch), _i32(ipbs), _i32(kbs)
    nch, nbf = len(ipch), len(kbs)
    ldo = ldo or max(nbf, 1)
    nort = np.zeros(nch, dtype=np.int32)
    iort = np.zeros(nch * ldo, dtype=np.int32)
    cb = IBORT_FN(lambda user, i, j: int(ibort(i, j)))
    check(lib().dbs_orth_conditions(nch, ipch.ctypes.data_as(_ip), nbf, ipbs.ctypes.data_as(_ip),
                                    kbs.ctypes.data_as(_ip), cb, None, ldo, nort.ctypes.data_as(_ip),
                                    iort.ctypes.data_as(_ip)))
    return nort, iort.reshape(nch, ldo)


def DBS_ORTH(m, idiag, nort, iort, pv, qv):
    """m: MatBlocks; (nort, iort) from orth_conditions; pv, qv: device tensors
    (nbf, ms) holding Get_pv / Get_qv of every orbital (Fortran (ms, nbf))"""
    nort = _i32(nort)
    iort = _i32(iort)
    ldo = iort.shape[1] if iort.ndim == 2 else max(len(iort) // max(len(nort), 1), 1)
    m._pre()
    check(lib().dbs_orth_dev(m.dbs.ctx, C.byref(m._desc), idiag, nort.ctypes.data_as(_ip),
                             iort.ctypes.data_as(_ip), ldo, _vp(pv), _vp(qv), int(pv.shape[0])))
    m._post()


def Diag_channels(m, diag, iprm, edmin=EDMIN, edmax=EDMAX, egap=EGAP):
    """diag: device tensor (nd, ms, ms) holding diag(ms,ms,nd), the overlaps of
    the diagonal blocks in, the solutions out; returns ipsol(0:nch)"""
    iprm = _i32(iprm)
    ipsol = np.zeros(m.nch + 1, dtype=np.int32)
    m._pre()
    check(lib().dbs_diag_channels_dev(m.dbs.ctx, C.byref(m._desc), _vp(diag), int(diag.shape[0]),
                                      iprm.ctypes.data_as(_ip), edmin, edmax, egap, ipsol.ctypes.data_as(_ip)))
    m._post()
    return ipsol


def record_sizes(m, ipsol):
    """capacities (in doubles) the channel and perturber records need at most"""
    nc = sum(int(ipsol[i]) * int(ipsol[j]) for i in range(2, m.nch + 1) for j in range(1, i)
             if m.icc[i - 1, j - 1] != 0)
    nw = sum(int(ipsol[i]) for i in range(1, m.nch + 1) for p in range(1, m.npert + 1)
             if m.npert and m.icb[i - 1, p - 1] != 0)
    return nc, nw


def record_matrix_device(m, diag, ipsol, cbuf, wbuf, overlaps=True):
    """Record_matrix with the records left in the device buffers cbuf / wbuf
    (float64 tensors of at least record_sizes' capacities): returns the host
    offset tables coff(nch,nch), woff(nch,npert) (-1 = not written) and the
    overlaps (nch+npert)^2 (None when overlaps is False)"""
    ipsol = _i32(ipsol)
    nch, npert = m.nch, m.npert
    coff = np.full(nch * nch, -1, dtype=np.int64)
    woff = np.full(max(nch * npert, 1), -1, dtype=np.int64)
    nt = nch + npert
    ov = np.zeros(nt * nt, dtype=np.float64) if overlaps else None
    m._pre()
    check(lib().dbs_record_matrix_dev(m.dbs.ctx, C.byref(m._desc), _vp(diag), int(diag.shape[0]),
                                      ipsol.ctypes.data_as(_ip), _vp(cbuf), int(cbuf.numel()),
                                      coff.ctypes.data_as(_lp), _vp(wbuf), int(wbuf.numel()),
                                      woff.ctypes.data_as(_lp), ov.ctypes.data_as(_dp) if overlaps else None))
    m._post()
    return (coff.reshape((nch, nch), order="F"), woff[:nch * npert].reshape((nch, npert), order="F"),
            ov.reshape((nt, nt), order="F") if overlaps else None)


def Record_matrix(m, diag, ipsol, icase=0, hp=None):
    """the records of Record_matrix in write order, ((i, j), values) with
    values a numpy array (c(1:ni,1:nj), w(1:ni)) or hp(k), ending with
    ((0, 0), None), and overlaps (nch+npert)^2 (set when icase = 0).
    diag: device tensor (nd >= nch, ms, ms) with the solutions by channel.
    hp: the host copy of the perturber-perturber matrix (read from m when
    None)."""
    import torch
    ipsol = _i32(ipsol)
    nch, npert = m.nch, m.npert
    nc, nw = record_sizes(m, ipsol)
    cbuf = torch.zeros(max(nc, 1), dtype=torch.float64, device="cuda")
    wbuf = torch.zeros(max(nw, 1), dtype=torch.float64, device="cuda")
    coff, woff, ov = record_matrix_device(m, diag, ipsol, cbuf, wbuf, overlaps=icase == 0)
    if ov is None:
        ov = np.zeros((nch + npert, nch + npert))
    c = cbuf.cpu().numpy()
    w = wbuf.cpu().numpy()
    recs = []
    for ich in range(2, nch + 1):
        ni = int(ipsol[ich])
        for jch in range(1, ich):
            o = coff[ich - 1, jch - 1]
            if o < 0:
                continue
            nj = int(ipsol[jch])
            recs.append(((ich, jch), c[o:o + ni * nj].reshape((ni, nj), order="F")))
    if npert > 0:
        for ich in range(1, nch + 1):
            ni = int(ipsol[ich])
            for ip in range(1, npert + 1):
                o = woff[ich - 1, ip - 1]
                if o >= 0:
                    recs.append(((nch + ip, ich), w[o:o + ni].copy()))
        hpv = hp if hp is not None else m.hp.cpu().numpy()
        for ip in range(1, npert + 1):
            for jp in range(1, ip + 1):
                k = m.ibb[ip - 1, jp - 1]
                if k == 0 or hpv[k - 1] == 0.0:
                    continue
                recs.append(((ip + nch, jp + nch), float(hpv[k - 1])))
    recs.append(((0, 0), None))
    return recs, ov
