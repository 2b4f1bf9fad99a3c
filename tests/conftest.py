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


def err_report(name, x, ref, scale=None, floor_frac=1e-3):
    """Error statistics of x against ref, printed and returned:
    'floored' the largest |x-ref| / max(|ref|, floor_frac*max|ref|); 'under'
    the number of entries whose |ref| lies below that floor and 'under_true'
    their largest true relative error |x-ref|/|ref| (the floor hides it);
    with `scale` (a per-entry magnitude bound, e.g. the Cauchy-Schwarz bound
    of an integral) 'scaled' the largest |x-ref|/scale."""
    x, ref = np.asarray(x, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    d = np.abs(x - ref)
    floor = floor_frac * np.max(np.abs(ref)) if ref.size else 0.0
    out = {"n": int(ref.size), "floored": float(np.max(d / np.maximum(np.abs(ref), floor))) if ref.size else 0.0}
    low = np.abs(ref) < floor
    out["under"] = int(np.count_nonzero(low))
    nz = low & (ref != 0)
    out["under_true"] = float(np.max(d[nz] / np.abs(ref[nz]))) if np.any(nz) else 0.0
    out["zero_ref_abs"] = float(np.max(d[ref == 0])) if np.any(ref == 0) else 0.0
    if scale is not None:
        out["scaled"] = float(np.max(d / np.asarray(scale)))
    print(f"  [{name}] " + " ".join(f"{k}={v:.3g}" if isinstance(v, float) else f"{k}={v}"
                                     for k, v in out.items()), flush=True)
    return out


def cs_bound(dbs, norb, pq, rows):
    """sqrt(R^k(a,a;c,c) * R^k(b,b;d,d)) for every row (k,a,b,c,d): the
    multipole kernel r<^k/r>^(k+1) is positive definite, so this bounds
    |R^k(a,b;c,d)| (Cauchy-Schwarz) and is the natural size against which the
    rounding error of one integral is measured.  The self integrals come
    from the engine (kind 0, the rk() sum)."""
    from dbsr3_amd import Engine
    rows = np.asarray(rows, dtype=np.int64)
    k, a, b, c, d = rows.T
    L = np.stack([k, np.minimum(a, c), np.maximum(a, c)], axis=1)
    R = np.stack([k, np.minimum(b, d), np.maximum(b, d)], axis=1)
    pairs, inv = np.unique(np.concatenate([L, R]), axis=0, return_inverse=True)
    inv = inv.ravel()
    self_rows = np.stack([pairs[:, 0], pairs[:, 1], pairs[:, 1], pairs[:, 2], pairs[:, 2]], axis=1)
    eng = Engine([(dbs, norb, self_rows.astype(np.int32))])
    eng.set_orbitals(0, pq)
    eng.set_coef(0, np.ones(len(self_rows)))
    eng.forward()
    v = eng.values()
    eng.close()
    assert np.all(v > 0), "self integrals must be positive"
    n = len(rows)
    return np.sqrt(v[inv[:n]] * v[inv[n:]])
