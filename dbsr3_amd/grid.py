"""Knot sequences of the reference's grids (host side).

mkgrid_01 (DBSR_HF/dbsr_lib_dbs.f90:724-783): ml equal steps hs, then an
exponential region with ratio 1+he while the step stays below hmax, then
equal steps hmax up to rmax; ns follows from the parameters.
mkgrid_02 (:787-823): a purely exponential grid for given ns and rmax.
The first step hs depends on the nucleus model (:737-745, :800-807):
hi/z for a point nucleus, 4 ln3 a_fermi hi for a Fermi distribution,
r_uniform hi for a uniform sphere.  ks-fold knots at 0 and at the end, as
the reference lays them out.

read_knot / knot_grid: the knot.dat files of def_grid (:480-596) --
parameters read as Read_ipar / Read_rpar / Read_apar find them ('name =
value' anywhere in the file), the grid rebuilt by mkgrid_01/02 for
grid_type 1/2 or the listed points read as read_grid does (:611-631) for
grid_type -1.
"""
from __future__ import annotations

import math
import re

import numpy as np


def grid_start(hi, z, nuclear="point", a_fermi=0.0, r_uniform=0.0):
    if nuclear == "point":
        if z <= 0:
            raise ValueError("mkgrid_01: Z <= 0")
        return hi / z
    if nuclear == "Fermi":
        ts = 4.0 * math.log(3.0) * a_fermi
        return ts * hi
    return r_uniform * hi


def mkgrid_01(ks, z, hi=0.25, he=0.25, hmax=1.0, rmax=50.0, ml=1, nuclear="point", a_fermi=0.0,
              r_uniform=0.0):
    hs = grid_start(hi, z, nuclear, a_fermi, r_uniform)
    ti = hs * ml
    ns = ks + ml
    ht = 1.0 + he
    ne = ns
    while True:
        tj = ti * ht
        if tj - ti > hmax:
            break
        ns += 1
        ti = tj
        ne = ns
        if ti > rmax:
            ns -= 1
            break
    if ti < rmax:
        while True:
            ti = ti + hmax
            if ti >= rmax:
                break
            ns += 1
    t = np.zeros(ns + ks)
    for i in range(ks + 1, ks + ml + 1):
        t[i - 1] = hs * (i - ks)
    for i in range(ks + ml + 1, ne + 1):
        t[i - 1] = t[i - 2] * ht
    if t[ne - 1] < rmax:
        for i in range(ne + 1, ns + 2):
            t[i - 1] = t[i - 2] + hmax
    t[ns + 1:ns + ks] = t[ns]
    return t, ns


def mkgrid_02(ns, ks, z, hi=0.25, rmax=30.0, nuclear="point", a_fermi=0.0, r_uniform=0.0):
    ti = grid_start(hi, z, nuclear, a_fermi, r_uniform)
    he = math.exp((math.log(rmax) - math.log(ti)) / (ns - ks))
    t = np.zeros(ns + ks)
    t[ks] = ti
    for i in range(ks + 2, ns + 2):
        t[i - 1] = t[i - 2] * he
    t[ns:ns + ks] = rmax
    return t


_INT = ("grid_type", "ns", "ks", "nv", "ml", "ksp", "ksq")
_REAL = ("hi", "he", "hmax", "rmax", "atomic_number", "atomic_weight", "rrms", "a_fermi", "c_fermi",
         "r_uniform", "t_fermi")
_DEFAULTS = dict(grid_type=1, ks=9, ns=0, ml=1, hi=0.25, he=0.25, hmax=1.0, rmax=50.0, ksp=8, ksq=9,
                 atomic_number=0.0, a_fermi=0.0, r_uniform=0.0, nuclear="Fermi")


def _fnum(s):
    return float(s.replace("D", "E").replace("d", "e"))


def read_knot(path):
    """parameters and listed grid points of a knot.dat file: a dict with the
    DBS_grid / DBS_nuclear names and 'points' (the t(1:ns+ks) listed after
    'grid points:', or None)"""
    text = open(path).read()
    par = dict(_DEFAULTS)
    for name in _INT + _REAL:
        m = re.search(r"(?m)^\s*" + re.escape(name) + r"\s*=\s*([-+0-9.DdEe]+)", text)
        if m:
            par[name] = int(m.group(1)) if name in _INT else _fnum(m.group(1))
    m = re.search(r"(?m)^\s*nuclear\s*=\s*(\S+)", text)
    if m:
        par["nuclear"] = m.group(1)
    par["points"] = None
    i = text.find("grid points:")
    if i >= 0:
        lines = text[i:].splitlines()[2:]
        n = par["ns"] + par["ks"]
        t = np.zeros(n)
        for line in lines[:n]:
            j, v = line.split()
            t[int(j) - 1] = _fnum(v)
        par["points"] = t
    return par


def knot_grid(path, z=0.0):
    """the grid def_grid sets up from a knot file: (t, ns, ks, ksp, ksq)"""
    p = read_knot(path)
    ks = p["ks"]
    ksp = p["ksp"] if 0 < p["ksp"] <= ks else ks - 1
    ksq = p["ksq"] if 0 < p["ksq"] <= ks else ks
    zz = z if z > 0 else p["atomic_number"]
    kw = dict(nuclear=p["nuclear"], a_fermi=p["a_fermi"], r_uniform=p["r_uniform"])
    if p["grid_type"] == 1:
        t, ns = mkgrid_01(ks, zz, p["hi"], p["he"], p["hmax"], p["rmax"], p["ml"], **kw)
    elif p["grid_type"] == 2:
        ns = p["ns"]
        t = mkgrid_02(ns, ks, zz, p["hi"], p["rmax"], **kw)
    elif p["grid_type"] == -1:
        if p["ns"] <= 0:
            raise ValueError("Stop in read_grid:  ns = 0")
        if p["points"] is None:
            raise ValueError("Stop in read_grid:  can not find grid points")
        t, ns = p["points"], p["ns"]
    else:
        raise ValueError("Unknown grid_type")
    return t, ns, ks, ksp, ksq
