"""Average-of-configuration energy of DBSR_HF, with the integrals on the GPU.

The energy is Etotal = sum_i qsum(i) <i|h_D|i> + sum_{i>=j,k} a(i,j,k) F^k(i,j)
+ sum_{i>j,k} b(i,j,k) G^k(i,j), in DBSR_HF's term = 'LS'/'AV' scheme:

* Read_confs (DBSR_HF/hf_get_case.f90:428-565): configurations from a
  conf-file (core subshells, peel subshells, one CSF per line), occupation
  numbers iqconf(ic,io), weights by eal (1 equal, 5 statistical: the product
  over the line's subshells of ndets_jq = C(2j+1, q),
  DBSR_HF/dbsr_lib_zconfjj.f90:4859-4871; 9 read from the line),
  normalised; qsum(io) = sum_ic iqconf(ic,io) weight(ic).
* av_energy_coef, WW, a, b (DBSR_HF/hf_MOD_energy.f90:19-170): the angular
  coefficients, with Cjkj (DBSR_HF/dbsr_lib_zcom.f90:435-458) computed here
  from the exact rational 3j symbol.
* update_int (:174-208): <i|h_D|i> by Vp_dhl with MAT_dhl_pq(kappa) built
  on the device (DBS.dhl_pq); F^k(i,j) = hf_fk(i,j,k) = R^k(i,j;i,j) and
  G^k(i,j) = hf_gk(i,j,k) = R^k(i,j;j,i) (DBSR_HF/hf_rk.f90:43-113) as rows
  of the integral engine in one batch.  hf_gk forms its P/Q cross term as
  2 qpqp; the engine sums qpqp and pqpq, which are the same quadrature with
  the roles of r1 and r2 exchanged, so the two agree to rounding.
* energy (:212-255), in the reference's summation order; Conf_energies
  (DBSR_HF/hf_summry.f90:95-134): each configuration alone with weight 1.
"""
from __future__ import annotations

import re
from fractions import Fraction
from math import factorial

import numpy as np

from .bsw import el_nljk


# ---------------------------------------------------------------------------
# angular algebra
# ---------------------------------------------------------------------------
def threej_sq(j1, j2, j3, m1, m2, m3):
    """square of the 3j symbol (j1/2 j2/2 j3/2; m1/2 m2/2 m3/2), arguments
    doubled, as an exact Fraction (Racah's formula)"""
    if m1 + m2 + m3 != 0:
        return Fraction(0)
    for j, m in ((j1, m1), (j2, m2), (j3, m3)):
        if abs(m) > j or (j + m) % 2:
            return Fraction(0)
    if j3 > j1 + j2 or j3 < abs(j1 - j2) or (j1 + j2 + j3) % 2:
        return Fraction(0)
    h = lambda x: x // 2  # doubled -> integer (all sums below are even)
    delta = Fraction(factorial(h(j1 + j2 - j3)) * factorial(h(j1 - j2 + j3)) * factorial(h(-j1 + j2 + j3)),
                     factorial(h(j1 + j2 + j3) + 1))
    f = 1
    for j, m in ((j1, m1), (j2, m2), (j3, m3)):
        f *= factorial(h(j + m)) * factorial(h(j - m))
    s = Fraction(0)
    tmin = max(0, h(j2 - j3 - m1), h(j1 - j3 + m2))
    tmax = min(h(j1 + j2 - j3), h(j1 - m1), h(j2 + m2))
    for t in range(tmin, tmax + 1):
        den = (factorial(t) * factorial(h(j3 - j2 + m1) + t) * factorial(h(j3 - j1 - m2) + t)
               * factorial(h(j1 + j2 - j3) - t) * factorial(h(j1 - m1) - t) * factorial(h(j2 + m2) - t))
        s += Fraction((-1) ** t, den)
    return delta * f * s * s


def cjkj_sq(jj1, k, jj2):
    """Cjkj(j1,k,j2)^2 = (j1 k j2; 1/2 0 -1/2)^2 (2j1+1)(2j2+1), j's doubled"""
    return threej_sq(jj1, 2 * k, jj2, 1, 0, -1) * (jj1 + 1) * (jj2 + 1)


def ndets_jq(jj, q):
    """number of determinants of subshell j^q: C(2j+1, q)"""
    if q > jj + 1:
        raise ValueError("ndets_jq: q > q_max")
    return factorial(jj + 1) // (factorial(q) * factorial(jj + 1 - q))


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
_CSF = re.compile(r"([0-9]+[a-zA-Z][-+]?)\s*\(\s*([0-9]+)\s*\)")


def read_confs(path, eal=5):
    """(labels, jj, iqconf(nconf, nwf), weight(nconf)) of a conf-file: core
    subshells first (full), then the peel subshells, in file order"""
    lines = open(path).read().splitlines()

    def after(tag):
        for q, line in enumerate(lines):
            if line.startswith(tag):
                return q + 1
        raise ValueError(f"Read_confs: no {tag!r} flag in {path}")

    core = lines[after("Core subshells:")].split()
    peel = lines[after("Peel subshells:")].split()
    labels = core + peel
    jj = [el_nljk(x)[3] for x in labels]
    index = {x: q for q, x in enumerate(labels)}
    confs, weights = [], []
    for line in lines[after("CSF(s):"):]:
        if line.startswith("*"):
            break
        if "(" not in line:
            continue
        iq = [jj[q] + 1 if q < len(core) else 0 for q in range(len(labels))]
        w = 1
        end = 0
        for m in _CSF.finditer(line):
            lab, q = m.group(1), int(m.group(2))
            if lab not in index:
                raise ValueError(f"unknown orbital in conf-file: {lab}")
            iq[index[lab]] = q
            w *= ndets_jq(jj[index[lab]], q)
            end = m.end()
        if eal == 5:
            weights.append(float(w))
        elif eal == 9:
            weights.append(float(line[end:].split()[0]))
        else:
            weights.append(1.0)
        confs.append(iq)
    weight = np.array(weights, dtype=float)
    if weight.sum() == 0.0:
        weight[:] = 1.0
    weight = weight / weight.sum()
    return labels, jj, np.array(confs, dtype=np.int64), weight


# ---------------------------------------------------------------------------
# coefficients
# ---------------------------------------------------------------------------
def ww(iqconf, weight, i, j):
    """WC(ic) WC(jc)-like product of occupations summed over configurations"""
    if i == j:
        return float(sum(iqconf[ic, i] * (iqconf[ic, i] - 1) * weight[ic] for ic in range(len(weight))))
    return float(sum(iqconf[ic, i] * iqconf[ic, j] * weight[ic] for ic in range(len(weight))))


def av_energy_coef(jbs, iqconf, weight):
    """coef(nbf, nbf, kmax+1): i>=j direct F^k, i<j exchange G^k"""
    nbf = len(jbs)
    kmax = max(jbs)
    coef = np.zeros((nbf, nbf, kmax + 1))
    for i in range(nbf):
        for j in range(i + 1):
            if i == j:
                c = ww(iqconf, weight, i, i) / 2.0
                coef[i, i, 0] = c
                for k in range(1, jbs[i] + 1):
                    coef[i, i, k] = -c * float(cjkj_sq(jbs[i], k, jbs[i])) / (jbs[i] * (jbs[i] + 1))
            else:
                c = ww(iqconf, weight, i, j)
                coef[i, j, 0] = c
                for k in range(abs(jbs[i] - jbs[j]) // 2, (jbs[i] + jbs[j]) // 2 + 1):
                    coef[j, i, k] = -c * float(cjkj_sq(jbs[i], k, jbs[j])) / ((jbs[i] + 1) * (jbs[j] + 1))
    return coef


def a_coef(coef, jbs, i, j, k):
    """coefficient of F^k(i,j) (Function a)"""
    if k % 2 or k < 0 or k > (jbs[i] + jbs[j]) // 2:
        return 0.0
    return coef[max(i, j), min(i, j), k]


def b_coef(coef, lbs, jbs, i, j, k):
    """coefficient of G^k(i,j) (Function b)"""
    if i == j or (k + lbs[i] + lbs[j]) % 2 or k < abs(jbs[i] - jbs[j]) // 2 or k > (jbs[i] + jbs[j]) // 2:
        return 0.0
    return coef[min(i, j), max(i, j), k]


# ---------------------------------------------------------------------------
# integrals on the GPU and the energy
# ---------------------------------------------------------------------------
class AverageEnergy:
    """The integrals of DBSR_HF's energy for one orbital set: orbitals given
    as (kappa, l, jj) with jj = 2j, the grid's DBS context and the nucleus
    (DBS.nucleus)."""

    def __init__(self, dbs, kappa, lbs, jbs, nucleus):
        self.dbs, self.nucleus = dbs, nucleus
        self.kbs, self.lbs, self.jbs = list(kappa), list(lbs), list(jbs)
        n = self.nbf = len(self.kbs)
        # F^k: j <= i, k = 0, 2, .., 2 min(l_i, l_j); G^k: j < i,
        # k = |l_i - l_j| .. l_i + l_j step 2 (update_int's loops)
        self.fkeys = [(i, j, k) for i in range(n) for j in range(i + 1)
                      for k in range(0, 2 * min(self.lbs[i], self.lbs[j]) + 1, 2)]
        self.gkeys = [(i, j, k) for i in range(n) for j in range(i)
                      for k in range(abs(self.lbs[i] - self.lbs[j]), self.lbs[i] + self.lbs[j] + 1, 2)]
        rows = [(k, i + 1, j + 1, i + 1, j + 1) for (i, j, k) in self.fkeys]
        rows += [(k, i + 1, j + 1, j + 1, i + 1) for (i, j, k) in self.gkeys]
        from .engine import Engine
        self.rows = np.array(rows, dtype=np.int32)
        self.engine = Engine([(dbs, n, self.rows)])
        self.valueI = np.zeros(n)
        self.fk, self.gk = {}, {}

    def update_int(self, pq):
        """one-electron <i|h_D|i> and every F^k, G^k of the orbitals pq(ns,2,nbf)"""
        import torch
        pq = np.asarray(pq, dtype=np.float64)
        self.engine.set_orbitals(0, pq)
        self.engine.forward()
        vals = self.engine.values()
        nf = len(self.fkeys)
        self.fk = {key: float(vals[q]) for q, key in enumerate(self.fkeys)}
        self.gk = {key: float(vals[nf + q]) for q, key in enumerate(self.gkeys)}
        dhl = {}
        for i in range(self.nbf):
            kap = self.kbs[i]
            if kap not in dhl:
                dhl[kap] = self.dbs.dhl_pq(self.nucleus, kap)  # [c, r] = dhl(r, c)
            v = torch.from_numpy(np.concatenate([pq[:, 0, i], pq[:, 1, i]])).to(dhl[kap].device)
            w = dhl[kap].T @ v  # MATMUL(dhl, v)
            self.valueI[i] = float((v * w).sum())
        return self

    def energy(self, iqconf, weight):
        """(Etotal, E1body, E2body) for configurations iqconf(nconf,nbf) with
        weights (normalised here), in ENERGY's summation order"""
        iqconf = np.asarray(iqconf)
        weight = np.asarray(weight, dtype=float)
        weight = weight / weight.sum()
        qsum = iqconf.T @ weight
        coef = av_energy_coef(self.jbs, iqconf, weight)
        et = e1 = e2 = 0.0
        l, jb = self.lbs, self.jbs
        for i in range(self.nbf):
            s = self.valueI[i]
            et = et + qsum[i] * s
            e1 = e1 + qsum[i] * s
            for j in range(i + 1):
                for k in range(0, 2 * min(l[i], l[j]) + 1, 2):
                    c = a_coef(coef, jb, i, j, k)
                    if c == 0.0:
                        continue
                    s = self.fk[(i, j, k)]
                    et = et + c * s
                    e2 = e2 + c * s
            for j in range(i):
                for k in range(abs(l[i] - l[j]), l[i] + l[j] + 1, 2):
                    c = b_coef(coef, l, jb, i, j, k)
                    if c == 0.0:
                        continue
                    s = self.gk[(i, j, k)]
                    et = et + c * s
                    e2 = e2 + c * s
        return et, e1, e2

    def conf_energies(self, iqconf):
        """each configuration's energy alone (Conf_energies)"""
        iqconf = np.asarray(iqconf)
        out = []
        for ic in range(len(iqconf)):
            w = np.zeros(len(iqconf))
            w[ic] = 1.0
            out.append(self.energy(iqconf, w)[0])
        return out
