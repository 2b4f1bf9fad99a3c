"""The Fe example of DBSR_HF as a known answer, on the CPU.

Fixtures (tests/golden/fe, copied from DBSR_HF/examples/Fe): Fe.bsw, the
converged orbitals the run wrote, and Fe.conf, its five 3d^6 4s^2
jj-configurations (term = LS, eal = 5).  The grid is tests/golden/knots/Fe.knot.
Known answers from DBSR_HF/examples/Fe/Fe.log:
  * :414 Coulomb energy -1271.392047659758646 au (the statistically
    weighted average energy);
  * :418-422 the five configuration energies;
  * :396-405 the orbital energies e(nl), which the .bsw header repeats.

The oracle (oracle.av_energy) rebuilds the energy from the orbitals: the
Dirac operator with the Fermi nucleus (V_nuclear, Def_Vnucl, MAT_dhl_pq),
hf_fk / hf_gk with the reference's densities, convolutions and SUM_AmB,
av_energy_coef and ENERGY.  It is pinned by the known answers, by the
compiled reference's tables (oracle/_ref), and the 3j algebra by closed
forms.

The compiled reference's Cjkj takes sqrt(real((j1+1)*(j2+1))) in default
REAL, single precision under gfortran or flang without real promotion
(DBSR_HF/dbsr_lib_zcom.f90:455-456), so its angular coefficients differ
from the exact ones by up to 1.7e-7.  Fe.log matches the exact
coefficients to 2e-13 au.  The build that produced Fe.log evaluated that
square root in double precision, so the oracle and the product follow the
exact values.  The tests check the compiled build's coefficients against the
single-precision square root.
"""
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FE = os.path.join(HERE, "golden", "fe")
KNOT = os.path.join(HERE, "golden", "knots", "Fe.knot")
need_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build (oracle/_ref) absent")

E_LOG = -1271.392047659758646  # Fe.log:414
# Fe.log:418-422, in Fe.conf's configuration order
CONF_LOG = {"3d-(4)3d(2)": -1271.414086871778181, "3d-(3)3d(3)": -1271.393154798735395,
            "3d-(2)3d(4)": -1271.386537856753193, "3d-(1)3d(5)": -1271.394236045834759,
            "3d(6)": -1271.416249365976000}
CONF_ORDER = ["3d-(4)3d(2)", "3d-(3)3d(3)", "3d-(2)3d(4)", "3d-(1)3d(5)", "3d(6)"]
# Fe.log:396-405
ENL_LOG = {"1s": -263.83043503, "2s": -32.47592950, "2p-": -27.84899719, "2p": -27.38260382,
           "3s": -4.26968011, "3p-": -2.81557004, "3p": -2.75494465, "3d-": -0.59745589,
           "3d": -0.59146973, "4s": -0.26371538}


def fe_case():
    from dbsr3_amd.bsw import pq_array, read_bsw
    from dbsr3_amd.hf_energy import read_confs
    b = read_bsw(os.path.join(FE, "Fe.bsw"))
    labels, jj, iq, w = read_confs(os.path.join(FE, "Fe.conf"))
    orbs = b["orbitals"]
    kbs = [o["kappa"] for o in orbs]
    lbs = [o["l"] for o in orbs]
    jbs = [o["jj"] for o in orbs]
    nuc = (2, b["atomic_number"], b["r_uniform"], b["a_fermi"], b["c_fermi"])
    return b, pq_array(b), kbs, lbs, jbs, nuc, labels, iq, w


def test_bsw_reader_matches_knot_and_log():
    from dbsr3_amd.grid import knot_grid, read_knot
    b = fe_case()[0]
    t, ns, ks, ksp, ksq = knot_grid(KNOT)
    assert (b["ns"], b["ks"], b["ksp"], b["ksq"]) == (ns, ks, ksp, ksq)
    assert np.array_equal(b["t"], t)
    kp = read_knot(KNOT)
    assert b["nuclear"] == "Fermi" and b["atomic_number"] == 26.0
    assert b["a_fermi"] == kp["a_fermi"] and b["c_fermi"] == kp["c_fermi"]
    labels = [o["label"] for o in b["orbitals"]]
    assert labels == list(ENL_LOG)
    for o in b["orbitals"]:
        assert abs(o["e"] - ENL_LOG[o["label"]]) <= 5e-9, o["label"]
        assert o["mbs"] < b["ns"] and np.all(o["p"][o["mbs"]:] == 0.0)
    # kappa and 2j of the labels (EL_nljk, kappa_lj)
    expect = {"1s": (-1, 0, 1), "2p-": (1, 1, 1), "2p": (-2, 1, 3), "3d-": (2, 2, 3), "3d": (-3, 2, 5)}
    for o in b["orbitals"]:
        if o["label"] in expect:
            assert (o["kappa"], o["l"], o["jj"]) == expect[o["label"]]


def test_read_confs_statistical_weights():
    labels, iq, w = fe_case()[6:]
    assert labels[:7] == ["1s", "2s", "2p-", "2p", "3s", "3p-", "3p"]
    assert labels[7:] == ["3d-", "3d", "4s"]
    # product of C(2j+1, q) over the peel subshells, normalised
    assert np.allclose(w * 210, [15, 80, 90, 24, 1], rtol=0, atol=1e-12)
    # the weights Fe.conf prints (4 decimals)
    printed = [float(line.split()[-1]) for line in open(os.path.join(FE, "Fe.conf"))
               if "(" in line and not line.startswith("CSF")]
    assert np.max(np.abs(np.array(printed) - w)) <= 5e-5
    assert np.all(iq[:, :7] == [2, 2, 2, 4, 2, 2, 4])
    assert np.all(iq[:, 7] + iq[:, 8] == 6) and np.all(iq[:, 9] == 2)


@pytest.mark.parametrize("impl", ["product", "oracle"])
def test_cjkj_closed_forms(impl):
    """C(j,0,j)^2 = 2j+1, and sum_k (2k+1) (j1 k j2; 1/2 0 -1/2)^2 = 1
    (orthogonality of the 3j symbols over the coupled moment)"""
    if impl == "product":
        from dbsr3_amd.hf_energy import cjkj_sq
        f = lambda a, k, b: float(cjkj_sq(a, k, b))
    else:
        f = O.cjkj_sq
    for j1 in (1, 3, 5, 7, 9):
        assert abs(f(j1, 0, j1) - (j1 + 1)) <= 1e-14 * (j1 + 1)
        for j2 in (1, 3, 5, 7, 9):
            s = sum((2 * k + 1) * f(j1, k, j2) / ((j1 + 1) * (j2 + 1))
                    for k in range(abs(j1 - j2) // 2, (j1 + j2) // 2 + 1))
            assert abs(s - 1.0) <= 1e-14, (j1, j2, s)


def test_product_coefficients_equal_oracle():
    from dbsr3_amd.hf_energy import av_energy_coef
    jbs, iq, w = fe_case()[4], fe_case()[7], fe_case()[8]
    a = av_energy_coef(jbs, iq, w)
    b = O.av_energy_coef(jbs, iq, w)
    assert a.shape == b.shape
    assert np.max(np.abs(a - b)) <= 1e-14 * np.max(np.abs(b))


def test_fe_energy_known_answer_oracle():
    b, pq, kbs, lbs, jbs, nuc, labels, iq, w = fe_case()
    g = O.Grid(b["t"], b["ns"], b["ks"], b["ksp"], b["ksq"])
    et, e1, e2, confs, *_ = O.av_energy(g, pq, kbs, lbs, jbs, nuc, iq, w)
    print(f"  oracle E = {et!r}, Fe.log {E_LOG!r}, diff {et - E_LOG:.2e}")
    assert abs(et - E_LOG) <= 1e-12 * abs(E_LOG)
    assert abs(e1 + e2 - et) <= 1e-12 * abs(et)
    for name, e in zip(CONF_ORDER, confs):
        assert abs(e - CONF_LOG[name]) <= 1e-12 * abs(CONF_LOG[name]), name


@need_ref
def test_dirac_operator_and_potential_vs_reference():
    b, pq, kbs, lbs, jbs, nuc = fe_case()[:6]
    g = O.Grid(b["t"], b["ns"], b["ks"], b["ksp"], b["ksq"])
    R = O.ref()
    R.ref_setup(b["ns"], b["ks"], b["ksp"], b["ksq"], O._d(O.f64(b["t"])))
    gr = g.tables()[0]
    for r in list(gr[:3, 0]) + [nuc[4] * 0.5, nuc[4], nuc[4] * 2.0, 1.0, 30.0]:
        assert O.vnuclear(*nuc, r) == O.ref_vnuclear(nuc, r), r
    for kap in sorted(set(kbs)):
        mine = O.dhl_pq(g, *nuc, kap)
        theirs = O.ref_dhl_pq(kap, nuc, g.ms)
        scale = np.max(np.abs(theirs))
        assert np.max(np.abs(mine - theirs)) <= 1e-15 * scale, kap
    # the point and uniform models of V_nuclear as well
    for model, bb in ((0, 0.0), (1, 5e-5)):
        n2 = (model, 26.0, bb, 0.0, 0.0)
        for r in (1e-6, 4e-5, 1e-3, 2.0):
            assert O.vnuclear(*n2, r) == O.ref_vnuclear(n2, r), (model, r)


@need_ref
def test_fe_energy_tables_vs_reference():
    """the compiled reference's update_int tables (valueI, F^k, G^k) equal
    the oracle's; its coefficients are the oracle's with Cjkj's square root
    in single precision"""
    import ctypes as C
    b, pq, kbs, lbs, jbs, nuc, labels, iq, w = fe_case()
    g = O.Grid(b["t"], b["ns"], b["ks"], b["ksp"], b["ksq"])
    R = O.ref()
    R.ref_setup(b["ns"], b["ks"], b["ksp"], b["ksq"], O._d(O.f64(b["t"])))
    et, e1, e2, confs, valueI, F, G = O.av_energy(g, pq, kbs, lbs, jbs, nuc, iq, w)
    rt = O.ref_av_energy(pq, kbs, lbs, jbs, nuc, iq, w)
    n, km = len(kbs), max(jbs)
    vals = np.zeros((n, n, km + 1), order="F")
    vali = np.zeros(n)
    co = np.zeros((n, n, km + 1), order="F")
    R.ref_av_tables.argtypes = [C.c_int, O._dp, O._dp, O._dp]
    R.ref_av_tables(km, O._d(vals), O._d(vali), O._d(co))
    assert np.max(np.abs(vali - valueI)) <= 1e-15 * np.max(np.abs(valueI))
    assert rt[1] == e1
    for (i, j, k), v in F.items():
        if vals[i, j, k] != 0.0:
            assert abs(vals[i, j, k] - v) <= 1e-15 * abs(v), (i, j, k)
    for (i, j, k), v in G.items():
        if vals[j, i, k] != 0.0:
            assert abs(vals[j, i, k] - v) <= 1e-15 * abs(v), (i, j, k)
    # coefficients of the last configuration (Conf_energies leaves them)
    e_last = np.eye(len(w))[-1]
    exact = O.av_energy_coef(jbs, iq, e_last)
    orig = O.cjkj_sq
    try:
        O.cjkj_sq = lambda a, k, c: O.threej_sq(a, 2 * k, c, 1, 0, -1) * float(
            np.sqrt(np.float32((a + 1) * (c + 1)))) ** 2
        single = O.av_energy_coef(jbs, iq, e_last)
    finally:
        O.cjkj_sq = orig
    assert np.max(np.abs(co - single)) <= 1e-14 * np.max(np.abs(co))
    assert 0 < np.max(np.abs(co - exact)) <= 1e-6 * np.max(np.abs(co))
    print(f"  compiled reference E = {rt[0]!r} (single-precision Cjkj), exact {et!r}, Fe.log {E_LOG!r}")
