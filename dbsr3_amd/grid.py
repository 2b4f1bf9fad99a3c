"""Knot sequences of the reference's grid recipes (host side, point nucleus).

mkgrid_01 (DBSR_HF/dbsr_lib_dbs.f90:724-783): ml equal steps hi/z, then an
exponential region with ratio 1+he while the step stays below hmax, then
equal steps hmax up to rmax; ns follows from the parameters.
mkgrid_02 (:787-823): a purely exponential grid for given ns and rmax.
ks-fold knots at 0 and at the end, as the reference lays them out.
"""
from __future__ import annotations

import math

import numpy as np


def mkgrid_01(ks, z, hi=0.25, he=0.25, hmax=1.0, rmax=50.0, ml=1):
    hs = hi / z
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


def mkgrid_02(ns, ks, z, hi=0.25, rmax=30.0):
    ti = hi / z
    he = math.exp((math.log(rmax) - math.log(ti)) / (ns - ks))
    t = np.zeros(ns + ks)
    t[ks] = ti
    for i in range(ks + 2, ns + 2):
        t[i - 1] = t[i - 2] * he
    t[ns:ns + ks] = rmax
    return t
