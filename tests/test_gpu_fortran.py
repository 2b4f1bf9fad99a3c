"""The Fortran drop-in layer (dbsr3_amd/fortran/dbs_gpu.f90) driven by a host
program written like the reference's callers (tests/fortran/drv_dropin.f90):
rk(), the Rk_evaluate sequence mrk_*/Density/Convol/SUM_AmB, quadr and a
direct potential, checked against the oracle at 1e-12."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, fe_grid, random_orbitals

pytestmark = pytest.mark.gpu
DRIVER = os.path.join(ROOT, "tests", "fortran", "drv_dropin")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built (no flang)")
def test_fortran_dropin(tmp_path, gpu_available):
    t, ns, ks, ksp, ksq = fe_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    fs = random_orbitals(g, 4, seed=44)
    pq = np.stack(fs, axis=2)
    k = 2
    with open(tmp_path / "drv_in.bin", "wb") as f:
        np.array([ns, ks, ksp, ksq], dtype=np.int32).tofile(f)
        np.asarray(t, dtype=np.float64).tofile(f)
        np.array([4, k], dtype=np.int32).tofile(f)
        np.asfortranarray(pq).T.tofile(f)  # column-major bytes
    print("[fortran drop-in driver]", flush=True)
    r = subprocess.run([DRIVER], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(tmp_path / "drv_out.bin", dtype=np.float64)
    assert out[6] == ns and out[7] == k
    ref = g.rk(fs[0], fs[1], fs[2], fs[3], k)
    assert abs(out[0] - ref) <= 1e-12 * abs(ref)
    assert abs(out[1] - ref) <= 1e-12 * abs(ref)
    for i, m in ((2, 0), (3, 1), (4, -1)):
        q = g.quadr(fs[0], fs[1], m)
        assert abs(out[i] - q) <= 1e-12 * max(abs(q), 1e-300) * 10
    # <d13 | V(d24)> with V from convol case 1 equals the pppp part of rk
    d1 = O.density(ns, ks, fs[0][:, 0], fs[2][:, 0], "s")
    d2 = O.density(ns, ks, fs[1][:, 0], fs[3][:, 0], "s")
    v = O.convol(ns, ks, g.mrk(O.PPPP, k), d2, 1, "s", "s")
    x = O.sum_amb(ns, ks, v, d1, "s")
    assert abs(out[5] - x) <= 1e-12 * abs(x)
    # interleaved mrk_/msk_ selections: convol_sk used ppqq, convol used qqqq
    dn1 = O.density(ns, ks, fs[0][:, 0], fs[2][:, 1], "n")
    dn2 = O.density(ns, ks, fs[1][:, 0], fs[3][:, 1], "n")
    v = O.convol(ns, ks, g.mrk(O.PPQQ, k), dn1, 2, "n", "n")
    x = O.sum_amb(ns, ks, v, dn2, "n")
    assert abs(out[8] - x) <= 1e-12 * np.sum(np.abs(v * dn2)), (out[8], x)
    d1 = O.density(ns, ks, fs[0][:, 1], fs[2][:, 1], "s")
    d2 = O.density(ns, ks, fs[1][:, 1], fs[3][:, 1], "s")
    v = O.convol(ns, ks, g.mrk(O.QQQQ, k), d1, 2, "s", "s")
    x = O.sum_amb(ns, ks, v, d2, "s")
    assert abs(out[9] - x) <= 1e-12 * abs(x), (out[9], x)
