"""The Fe example of DBSR_HF end to end on the GPU: Fe.bsw's orbitals and
Fe.knot's grid, the Dirac operator with the Fermi nucleus built on the
device (dbs_dhl_pq_dev), every F^k and G^k of the energy as one batch of
engine rows, and DBSR_HF's average-of-configuration energy over Fe.conf's
five configurations.  Known answers: DBSR_HF/examples/Fe/Fe.log:414 (the
average energy, -1271.392047659758646 au) and :418-422 (the configuration
energies); the device matrices and integrals are checked against the oracle,
which reproduces those energies to 2e-16 (tests/test_oracle_fe.py).

Tolerance: 1e-12 relative on every energy (the path's per-result bound).  The engine sums each integral's quadrature in another order than
hf_fk / hf_gk (G^k's P/Q cross term as qpqp + pqpq rather than 2 qpqp),
and <i|h_D|i> is a device matrix product; each differs from the oracle by a
few 1e-16 of its magnitude.
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_fe import CONF_LOG, CONF_ORDER, E_LOG, fe_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fe(gpu_available):
    from dbsr3_amd import DBS
    case = fe_case()
    b = case[0]
    d = DBS(b["t"], b["ns"], b["ks"], b["ksp"], b["ksq"])
    g = O.Grid(b["t"], b["ns"], b["ks"], b["ksp"], b["ksq"])
    return case, d, g


def test_dirac_operator_on_device(fe):
    (b, pq, kbs, lbs, jbs, nuc, *_), d, g = fe
    from dbsr3_amd import DBS
    nu = DBS.nucleus("Fermi", b["atomic_number"], a_fermi=b["a_fermi"], c_fermi=b["c_fermi"])
    vr = d.vnucl(nu).cpu().numpy().T
    vr_o = O.vnucl(g, *nuc)
    assert np.max(np.abs(vr - vr_o)) <= 1e-14 * np.max(np.abs(vr_o))
    for kap in sorted(set(kbs)):
        dh = d.dhl_pq(nu, kap).cpu().numpy().T  # column-major (ms,ms)
        ref = O.dhl_pq(g, *nuc, kap)
        err = np.max(np.abs(dh - ref)) / np.max(np.abs(ref))
        print(f"  dhl kappa={kap}: max |diff| / max = {err:.2e}", flush=True)
        assert err <= 1e-14, kap


def test_fe_energy_known_answer(fe):
    (b, pq, kbs, lbs, jbs, nuc, labels, iq, w), d, g = fe
    from dbsr3_amd import DBS
    from dbsr3_amd.hf_energy import AverageEnergy
    nu = DBS.nucleus(b["nuclear"], b["atomic_number"], b["r_uniform"], b["a_fermi"], b["c_fermi"])
    print(f"[Fe end to end: ns={b['ns']} orbitals={len(kbs)} configurations={len(w)}]", flush=True)
    ae = AverageEnergy(d, kbs, lbs, jbs, nu).update_int(pq)
    et, e1, e2 = ae.energy(iq, w)
    print(f"  E = {et!r}  Fe.log {E_LOG!r}  diff {et - E_LOG:.2e}  (E1 {e1:.12f}, E2 {e2:.12f})", flush=True)
    assert abs(et - E_LOG) <= 1e-12 * abs(E_LOG)
    confs = ae.conf_energies(iq)
    for name, e in zip(CONF_ORDER, confs):
        print(f"  {name:12s} {e!r}  Fe.log {CONF_LOG[name]!r}", flush=True)
        assert abs(e - CONF_LOG[name]) <= 1e-12 * abs(CONF_LOG[name]), name
    # every integral and one-electron term against the oracle's hf_fk / hf_gk / Vp_dhl
    oet, oe1, oe2, oconfs, valueI, F, G = O.av_energy(g, pq, kbs, lbs, jbs, nuc, iq, w)
    assert np.max(np.abs(ae.valueI - valueI)) <= 1e-13 * np.max(np.abs(valueI))
    for key, v in F.items():
        assert abs(ae.fk[key] - v) <= 1e-12 * abs(v), ("F", key)
    for key, v in G.items():
        assert abs(ae.gk[key] - v) <= 1e-12 * max(abs(v), 1e-300), ("G", key)
