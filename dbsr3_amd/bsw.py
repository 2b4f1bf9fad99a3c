"""B-spline wave-function files (.bsw) of DBSR_HF and DBSR_MCHF, read on
the host.

Layout, as write_pqbs writes it (DBSR_HF/hf_read_pqbs.f90:61-96) and
Read_pqbs / read_grid_bsw read it (:2-57; DBSR_HF/dbsr_lib_dbs.f90:635-651):
Fortran sequential unformatted records, each framed by its byte length in
4-byte integers before and after.

    record 1:  grid_type, ns, ks (int32), t(1:ns+ks), ksp, ksq (int32),
               atomic_number, atomic_weight, nuclear (20 characters),
               r_uniform, a_fermi, c_fermi, t_fermi, rrms (float64)
               -- DBSR_MCHF's Write_pqbs stops after ksp, ksq
               (DBSR_MCHF/read_pqbs.f90:84); the nucleus keys are then None
    per orbital, three records:
               label (5 characters), mbs (int32), e (float64)
               P(1:mbs)
               Q(1:mbs)

The coefficients past mbs are zero.  Labels are decoded as EL_nljk does
(DBSR_HF/dbsr_lib_zconfjj.f90:4112-4175): n, l, 2j = 2l+1 (or 2l-1 with a
trailing '-'), kappa = (2l - 2j)(2j + 1)/2 (kappa_lj,
DBSR_HF/dbsr_lib_zcom.f90:487-496).
"""
from __future__ import annotations

import struct

import numpy as np

_L_LETTERS = "spdfghiklmnoqrtuvwxyz"


def _records(buf):
    out = []
    pos = 0
    while pos < len(buf):
        if pos + 4 > len(buf):
            raise ValueError("bsw: truncated record marker")
        n = struct.unpack_from("<i", buf, pos)[0]
        end = pos + 4 + n
        if n < 0 or end + 4 > len(buf) or struct.unpack_from("<i", buf, end)[0] != n:
            raise ValueError("bsw: record framing does not match")
        out.append(buf[pos + 4:end])
        pos = end + 4
    return out


def el_nljk(label):
    """(n, kappa, l, 2j) of a spectroscopic orbital label such as '2p-'"""
    s = label.strip()
    if not s:
        raise ValueError("empty orbital label")
    minus = s.endswith("-")
    core = s[:-1] if s[-1] in "-+" else s
    i = 0
    while i < len(core) and core[i].isdigit():
        i += 1
    if i == 0 or i >= len(core):
        raise ValueError(f"can not decode orbital label {label!r}")
    n = int(core[:i])
    l = _L_LETTERS.index(core[i].lower())
    jj = 2 * l - 1 if minus else 2 * l + 1
    kappa = (2 * l - jj) * (jj + 1) // 2
    return n, kappa, l, jj


def read_bsw(path):
    """dict: grid_type, ns, ks, ksp, ksq, t (ns+ks), atomic_number,
    atomic_weight, nuclear, r_uniform, a_fermi, c_fermi, t_fermi, rrms and
    'orbitals': list of dicts (label, n, kappa, l, jj, mbs, e, p(ns), q(ns))"""
    recs = _records(open(path, "rb").read())
    if not recs:
        raise ValueError("bsw: no records")
    h = recs[0]
    grid_type, ns, ks = struct.unpack_from("<3i", h, 0)
    off = 12
    t = np.frombuffer(h, dtype="<f8", count=ns + ks, offset=off).copy()
    off += 8 * (ns + ks)
    ksp, ksq = struct.unpack_from("<2i", h, off)
    off += 8
    if off == len(h):
        nucl = dict(atomic_number=None, atomic_weight=None, nuclear=None, r_uniform=None, a_fermi=None,
                    c_fermi=None, t_fermi=None, rrms=None)
        return _orbitals(dict(grid_type=grid_type, ns=ns, ks=ks, ksp=ksp, ksq=ksq, t=t, **nucl, orbitals=[]),
                         recs[1:])
    z, atw = struct.unpack_from("<2d", h, off)
    off += 16
    nuclear = h[off:off + 20].decode("ascii").strip()
    off += 20
    r_uniform, a_fermi, c_fermi, t_fermi, rrms = struct.unpack_from("<5d", h, off)
    off += 40
    if off != len(h):
        raise ValueError("bsw: unexpected header length")
    out = dict(grid_type=grid_type, ns=ns, ks=ks, ksp=ksp, ksq=ksq, t=t, atomic_number=z, atomic_weight=atw,
               nuclear=nuclear, r_uniform=r_uniform, a_fermi=a_fermi, c_fermi=c_fermi, t_fermi=t_fermi,
               rrms=rrms, orbitals=[])
    return _orbitals(out, recs[1:])


def _orbitals(out, body):
    ns = out["ns"]
    if len(body) % 3:
        raise ValueError("bsw: orbital records come in threes")
    for q in range(0, len(body), 3):
        hd = body[q]
        label = hd[:5].decode("ascii")
        mbs, e = struct.unpack_from("<id", hd, 5)
        p = np.zeros(ns)
        qq = np.zeros(ns)
        pw = np.frombuffer(body[q + 1], dtype="<f8")
        qw = np.frombuffer(body[q + 2], dtype="<f8")
        if len(pw) != mbs or len(qw) != mbs or mbs > ns:
            raise ValueError(f"bsw: orbital {label!r}: {len(pw)} coefficients for mbs = {mbs}")
        p[:mbs] = pw
        qq[:mbs] = qw
        n, kappa, l, jj = el_nljk(label)
        out["orbitals"].append(dict(label=label.strip(), n=n, kappa=kappa, l=l, jj=jj, mbs=mbs, e=e, p=p, q=qq))
    return out


def pq_array(bsw):
    """pq(ns, 2, norb) of the file's orbitals, in file order"""
    orbs = bsw["orbitals"]
    pq = np.zeros((bsw["ns"], 2, len(orbs)), order="F")
    for i, o in enumerate(orbs):
        pq[:, 0, i] = o["p"]
        pq[:, 1, i] = o["q"]
    return pq
