import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


GOLDEN = os.path.join(ROOT, "tests", "golden")


def fe_grid():
    """Fe.knot grid of the reference example (DBSR_HF/examples/Fe/Fe.knot):
    ns=113, ks=9, ksp=8, ksq=9."""
    t = np.loadtxt(os.path.join(GOLDEN, "fe_knot_t.txt"))
    return t, 113, 9, 8, 9


def small_grid():
    """exponential grid (mkgrid_02 recipe) with ns=14, ks=5, ksp=4, ksq=5"""
    import oracle as O
    return O.mkgrid_02(14, 5, 0.25, 20.0, 4.0), 14, 5, 4, 5


def random_orbitals(g, n, seed):
    """seeded orbitals: P,Q ~ N(0,1) on 2..nsp-1 / 2..nsq-1, first and last
    zero (bdcwf_pq convention), normalised so quadr(f,f,0) = 1."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        f = np.zeros((g.ns, 2), order="F")
        f[1:g.nsp - 1, 0] = rng.standard_normal(g.nsp - 2)
        f[1:g.nsq - 1, 1] = rng.standard_normal(g.nsq - 2)
        f /= np.sqrt(g.quadr(f, f, 0))
        out.append(np.asfortranarray(f))
    return out


@pytest.fixture(scope="session")
def gpu_available():
    try:
        import torch
        ok = torch.cuda.is_available()
    except Exception:
        ok = False
    if not ok:
        pytest.skip("no GPU")
    return True
