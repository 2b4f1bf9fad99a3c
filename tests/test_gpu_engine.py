"""GPU parity of the integral engine (dbs_engine_*) against the oracle.

Values: every row R^k(a,b;c,d) (rk() sum, single P/Q types, S^k types) is
compared with the oracle's materialised-array evaluation (mrk/msk +
convol case 2 + SUM_AmB, the reference's rk()/Rk_evaluate path).  The
engine sums the same quadrature in another order; the bound is 1e-12
relative, with the error of values smaller than 1e-3 of the set's largest
magnitude measured against that floor: such values come out of cancelling
terms of the set's scale, where rounding of any summation order is ~1e-16 of
that scale (an absolute bound of 1e-15 of the scale).

Adjoint: dE/d(pq) of E = sum c_n R_n against the oracle's potentials --
convol case 1 (direct, over r2) and case 2 (over r1) of every row's
densities, unfolded by Full_mat_sym and applied to the partner orbital, as
hf_matrix does (DBSR_HF/hf_matrix.f90:41-59) -- and against central finite
differences of the engine's own energy.
"""
import numpy as np
import pytest

import oracle as O
from conftest import fe_grid, random_orbitals, small_grid

pytestmark = pytest.mark.gpu
TOL = 1e-12

TYPES_OF_KIND = {0: [(O.PPPP, 0, 0), (O.QQQQ, 1, 1), (O.QPQP, 1, 0), (O.PQPQ, 0, 1)],
                 1: [(O.PPPP, 0, 0)], 2: [(O.QQQQ, 1, 1)], 3: [(O.PQPQ, 0, 1)], 4: [(O.QPQP, 1, 0)]}


def relerr(x, ref):
    ref = np.asarray(ref)
    floor = 1e-3 * np.max(np.abs(ref))
    return np.abs(np.asarray(x) - ref) / np.maximum(np.abs(ref), floor)


def oracle_values(g, fs, rows, kind):
    ns, ks = g.ns, g.ks
    out = np.zeros(len(rows))
    dcache = {}

    def dens(a, c, ca, cc, typ):
        key = (a, c, ca, cc, typ)
        if key not in dcache:
            dcache[key] = O.density(ns, ks, fs[a][:, ca], fs[c][:, cc], typ)
        return dcache[key]

    for n, (k, a, b, c, d) in enumerate(rows):
        a, b, c, d = a - 1, b - 1, c - 1, d - 1
        if kind <= 4:
            s = 0.0
            for t, cl, cr in TYPES_OF_KIND[kind]:
                rkb = g.mrk(t, k)
                conv = O.convol(ns, ks, rkb, dens(a, c, cl, cl, "s"), 2, "s", "s")
                s += O.sum_amb(ns, ks, conv, dens(b, d, cr, cr, "s"), "s")
            out[n] = s
        else:
            t = kind
            rkb = g.mrk(t, k)
            dl = dens(a, c, 0, 1, "n")
            dr = dens(b, d, 0, 1, "n") if t == O.PPQQ else dens(b, d, 1, 0, "n")
            conv = O.convol(ns, ks, rkb, dl, 2, "n", "n")
            out[n] = O.sum_amb(ns, ks, conv, dr, "n")
    return out


def unfold_s(a, ns, ks):
    y = np.zeros((ns, ns))
    for jp in range(ks):
        for i in range(ns - jp):
            y[i, i + jp] = a[i, jp]
            y[i + jp, i] = a[i, jp]
    return y


def oracle_gradient(g, fs, rows, coef):
    """dE/d(pq) for E = sum c*rk via convol case 1/2 potentials (kind 0)."""
    ns, ks = g.ns, g.ks
    grad = np.zeros((ns, 2, len(fs)), order="F")
    for (k, a, b, c, d), cf in zip(rows, coef):
        a, b, c, d = a - 1, b - 1, c - 1, d - 1
        for t, cl, cr in TYPES_OF_KIND[0]:
            rkb = g.mrk(t, k)
            dl = O.density(ns, ks, fs[a][:, cl], fs[c][:, cl], "s")
            dr = O.density(ns, ks, fs[b][:, cr], fs[d][:, cr], "s")
            vl = unfold_s(O.convol(ns, ks, rkb, dr, 1, "s", "s"), ns, ks)  # dR/dDL
            vr = unfold_s(O.convol(ns, ks, rkb, dl, 2, "s", "s"), ns, ks)  # dR/dDR
            grad[:, cl, a] += cf * vl @ fs[c][:, cl]
            grad[:, cl, c] += cf * vl @ fs[a][:, cl]
            grad[:, cr, b] += cf * vr @ fs[d][:, cr]
            grad[:, cr, d] += cf * vr @ fs[b][:, cr]
    return grad


def make_rows(rng, norb, kmax, n):
    rows = np.zeros((n, 5), dtype=np.int32)
    rows[:, 0] = rng.integers(0, kmax + 1, n)
    rows[:, 1:] = rng.integers(1, norb + 1, (n, 4))
    return rows


@pytest.fixture(scope="module")
def setups(gpu_available):
    from dbsr3_amd import DBS
    out = {}
    for name, spec in (("small", small_grid()), ("fe", fe_grid())):
        t, ns, ks, ksp, ksq = spec
        g = O.Grid(t, ns, ks, ksp, ksq)
        out[name] = (g, DBS(t, ns, ks, ksp, ksq))
    return out


def pq_of(fs):
    return np.stack(fs, axis=2)


@pytest.mark.parametrize("name", ["small", "fe"])
@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4, 5, 6])
def test_values(setups, name, kind):
    from dbsr3_amd import Engine
    g, d = setups[name]
    norb = 5
    fs = random_orbitals(g, norb, seed=100 + kind)
    rng = np.random.default_rng(kind)
    rows = make_rows(rng, norb, 4 if name == "fe" else 3, 60 if name == "fe" else 120)
    print(f"[engine values {name} kind={kind} rows={len(rows)}]", flush=True)
    eng = Engine([(d, norb, rows)], kind=kind)
    eng.set_orbitals(0, pq_of(fs))
    eng.forward()
    got = eng.values()
    ref = oracle_values(g, fs, rows, kind)
    e = relerr(got, ref)
    print(f"  max rel {e.max():.3e} median {np.median(e):.3e}")
    assert e.max() < TOL, (name, kind, e.max())
    for path in (0, 1):  # every group on the sparse, then on the dense MFMA path
        eng.set_tuning([path])
        eng.forward()
        ep = relerr(eng.values(), ref)
        assert ep.max() < TOL, (name, kind, "path", path, ep.max())
    # plain path (materialised arrays) on the device
    eng.forward_plain()
    gp = eng.values()
    ep = relerr(gp, ref)
    assert ep.max() < TOL, ("plain", name, kind, ep.max())


@pytest.mark.parametrize("name", ["small", "fe"])
def test_dense_path_matches_sparse(setups, name):
    from dbsr3_amd import Engine
    g, d = setups[name]
    norb = 6
    fs = random_orbitals(g, norb, seed=7)
    # all pairs a<=c, b<=d at two k: a dense set
    pairs = [(a, c) for a in range(1, norb + 1) for c in range(a, norb + 1)]
    rows = np.array([(k, a, b, c, dd) for k in (0, 2) for (a, c) in pairs for (b, dd) in pairs],
                    dtype=np.int32)
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq_of(fs))
    eng.set_tuning([0])
    eng.forward()
    sparse = eng.values()
    eng.set_tuning([1])
    eng.forward()
    dense = eng.values()
    assert np.max(relerr(dense, sparse)) < 1e-13
    ref = oracle_values(g, fs, rows[::37], 0)
    assert np.max(relerr(dense[::37], ref)) < TOL


@pytest.mark.parametrize("name", ["small", "fe"])
def test_adjoint_vs_oracle_potentials(setups, name):
    from dbsr3_amd import Engine
    g, d = setups[name]
    norb = 4
    fs = random_orbitals(g, norb, seed=21)
    rng = np.random.default_rng(21)
    rows = make_rows(rng, norb, 3, 40 if name == "fe" else 80)
    coef = rng.uniform(-1, 1, len(rows))
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq_of(fs))
    eng.set_coef(0, coef)
    eng.forward_adjoint()
    E = eng.energy()
    G = eng.grad()
    vals = oracle_values(g, fs, rows, 0)
    assert abs(E - vals @ coef) <= TOL * np.sum(np.abs(vals * coef))
    Gref = oracle_gradient(g, fs, rows, coef)
    scale = np.max(np.abs(Gref))
    err = np.max(np.abs(G - Gref)) / scale
    print(f"[adjoint {name}] max |dG|/max|G| = {err:.3e}")
    assert err < TOL


def test_adjoint_finite_differences(setups):
    from dbsr3_amd import Engine
    g, d = setups["fe"]
    norb = 4
    fs = random_orbitals(g, norb, seed=5)
    rng = np.random.default_rng(5)
    rows = make_rows(rng, norb, 4, 50)
    coef = rng.uniform(-1, 1, len(rows))
    eng = Engine([(d, norb, rows)])
    pq = pq_of(fs)
    eng.set_orbitals(0, pq)
    eng.set_coef(0, coef)
    eng.forward_adjoint()
    G = eng.grad()
    h = 1e-6
    worst = 0.0
    for _ in range(24):
        i = int(rng.integers(1, g.nsp - 1))
        comp = int(rng.integers(0, 2))
        o = int(rng.integers(0, norb))
        p2 = pq.copy(order="F")
        p2[i, comp, o] += h
        eng.set_orbitals(0, p2)
        eng.forward()
        ep = eng.energy()
        p2[i, comp, o] -= 2 * h
        eng.set_orbitals(0, p2)
        eng.forward()
        em = eng.energy()
        fd = (ep - em) / (2 * h)
        worst = max(worst, abs(fd - G[i, comp, o]) / max(abs(G[i, comp, o]), 1e-3 * np.max(np.abs(G))))
    print(f"[finite differences] worst rel {worst:.3e}")
    assert worst < 1e-6


def test_batched_instances_match_single(setups):
    """uneven instances (different grids, orders, norb, k ranges) in one engine"""
    from dbsr3_amd import DBS, Engine
    specs = []
    rng = np.random.default_rng(9)
    for n, (ns, ks) in enumerate(((20, 5), (31, 7), (26, 6), (40, 8))):
        t = O.mkgrid_02(ns, ks, 0.25, 25.0, 3.0 + n)
        g = O.Grid(t, ns, ks, ks - 1, ks)
        d = DBS(t, ns, ks, ks - 1, ks)
        norb = 3 + n
        fs = random_orbitals(g, norb, seed=30 + n)
        rows = make_rows(rng, norb, 2 + n, 30 + 10 * n)
        coef = rng.uniform(-1, 1, len(rows))
        specs.append((g, d, norb, fs, rows, coef))
    eng = Engine([(d, norb, rows) for (g, d, norb, fs, rows, coef) in specs])
    for i, (g, d, norb, fs, rows, coef) in enumerate(specs):
        eng.set_orbitals(i, pq_of(fs))
        eng.set_coef(i, coef)
    eng.forward_adjoint()
    for i, (g, d, norb, fs, rows, coef) in enumerate(specs):
        single = Engine([(d, norb, rows)])
        single.set_orbitals(0, pq_of(fs))
        single.set_coef(0, coef)
        single.forward_adjoint()
        assert np.array_equal(eng.values(i), single.values()), i
        assert np.array_equal(eng.grad(i), single.grad()), i
        ref = oracle_values(g, fs, rows, 0)
        assert relerr(eng.values(i), ref).max() < TOL


def test_graph_replay_matches_eager(setups):
    import torch
    from dbsr3_amd import Engine
    g, d = setups["fe"]
    norb = 5
    fs = random_orbitals(g, norb, seed=3)
    rng = np.random.default_rng(3)
    rows = make_rows(rng, norb, 4, 80)
    coef = rng.uniform(-1, 1, len(rows))
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq_of(fs))
    eng.set_coef(0, coef)
    eng.forward_adjoint()
    v0, g0, e0 = eng.values(), eng.grad(), eng.energy()
    eng.graph_capture(with_adjoint=True)
    for rep in range(4):
        # eager work of the program between replays
        x = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()
        eng.graph_launch()
        eng.sync()
        assert np.array_equal(eng.values(), v0), rep
        assert np.array_equal(eng.grad(), g0), rep
        assert eng.energy() == e0
        del x


def test_autotune_keeps_results(setups):
    from dbsr3_amd import Engine
    g, d = setups["small"]
    norb = 5
    fs = random_orbitals(g, norb, seed=8)
    rng = np.random.default_rng(8)
    rows = make_rows(rng, norb, 3, 200)
    coef = rng.uniform(-1, 1, len(rows))
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq_of(fs))
    eng.set_coef(0, coef)
    eng.forward_adjoint()
    v0, g0 = eng.values(), eng.grad()
    best = eng.autotune(with_adjoint=True, reps=2)
    assert len(best) == 7
    eng.forward_adjoint()
    assert np.max(relerr(eng.values(), v0)) < 1e-13
    assert np.max(np.abs(eng.grad() - g0)) <= 1e-13 * np.max(np.abs(g0))


@pytest.mark.parametrize("kind", [0, 5])
def test_pair_panel_shapes(setups, kind):
    """every pair-panel launch shape (panel width, workgroup size, panels
    per workgroup) and the dense path give the same values, energy and
    gradient"""
    from dbsr3_amd import Engine
    g, d = setups["fe"]
    norb = 8
    fs = random_orbitals(g, norb, seed=300 + kind)
    rng = np.random.default_rng(300 + kind)
    rows = make_rows(rng, norb, 3, 300)
    coef = rng.uniform(-1, 1, len(rows))
    print(f"[pair panel shapes kind={kind} rows={len(rows)}]", flush=True)
    eng = Engine([(d, norb, rows)], kind=kind)
    eng.set_orbitals(0, pq_of(fs))
    eng.set_coef(0, coef)
    eng.set_tuning([0, 256, 16, 4096, 256, 1])
    eng.forward_adjoint()
    v0, g0, e0 = eng.values(), eng.grad(), eng.energy()
    ref = oracle_values(g, fs, rows[::7], kind)
    assert relerr(v0[::7], ref).max() < TOL
    gs = np.max(np.abs(g0))
    assert gs > 0
    for t in ([0, 128, 4, 4096, 64, 1], [0, 256, 8, 4096, 512, 3], [0, 128, 16, 64, 128, 64],
              [0, 256, 8, 4096, 192, 2], [1, 256, 16, 4096, 256, 1], [2, 256, 8, 4096, 256, 4],
              [0, 256, 16, 4096, 256, 4, 2]):
        print(f"  tuning {t}", flush=True)
        eng.set_tuning(t)
        eng.forward_adjoint()
        assert np.max(relerr(eng.values(), v0)) < 1e-13, t
        assert np.max(np.abs(eng.grad() - g0)) <= 1e-13 * gs, t
        assert abs(eng.energy() - e0) <= 1e-13 * np.sum(np.abs(coef * v0)), t
        eng.forward()
        assert np.max(relerr(eng.values(), v0)) < 1e-13, t


@pytest.mark.parametrize("norb,width,fblock", [(24, 16, 256), (24, 8, 64), (40, 8, 256)])
def test_group_panel_variants(setups, norb, width, fblock):
    """one multipole with 300 (24 orbitals) or 820 (40 orbitals) total
    densities: the pair panels of the rotated kernel (the group's panel in
    LDS) or of the chunked kernel (partners staged chunk by chunk); the features
    by the matrix-core Gram kernel or (64-thread blocks, more than four
    densities per thread) the per-density kernel; 40 rows carry
    coefficients and are checked with their gradient against the oracle"""
    from dbsr3_amd import Engine
    g, d = setups["small"]
    fs = random_orbitals(g, norb, seed=norb + width)
    rng = np.random.default_rng(norb + width)
    rows = make_rows(rng, norb, 0, 4000)
    rows[:, 0] = 2
    sel = rng.choice(len(rows), 40, replace=False)
    coef = np.zeros(len(rows))
    coef[sel] = rng.uniform(-1, 1, len(sel))
    print(f"[panel variants norb={norb} width={width} fblock={fblock}]", flush=True)
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq_of(fs))
    eng.set_coef(0, coef)
    eng.set_tuning([0, fblock, width, 4096, 256, 2])
    eng.forward_adjoint()
    ref = oracle_values(g, fs, rows[sel], 0)
    assert relerr(eng.values()[sel], ref).max() < TOL
    Gref = oracle_gradient(g, fs, rows[sel], coef[sel])
    assert np.max(np.abs(eng.grad() - Gref)) < TOL * np.max(np.abs(Gref))
    v1, g1 = eng.values(), eng.grad()
    eng.forward()
    assert np.max(relerr(eng.values(), v1)) < 1e-13
    # owner ranges of the rotated kernel (2, 4, 8 workgroups per slot range,
    # several panels each): the same values and gradient
    for sp, npn in ((2, 4), (4, 2), (8, 1)):
        eng.set_tuning([0, fblock, width, 4096, 256, npn, sp])
        eng.forward_adjoint()
        assert np.max(relerr(eng.values(), v1)) < 1e-13, sp
        assert np.max(np.abs(eng.grad() - g1)) <= 1e-13 * np.max(np.abs(g1)), sp
        eng.forward()
        assert np.max(relerr(eng.values(), v1)) < 1e-13, sp


@pytest.mark.parametrize("kind", [0, 5])
def test_chunked_panels_match_rotated(setups, kind):
    """one multipole with 1830 densities (60 orbitals): at panel width 4 the
    group's panel fits in LDS (rotated kernel, two densities per thread);
    at widths 8 and 16 its partners are staged in LDS chunk by chunk
    (chunked kernel; owner ranges of 256 or 512 densities, one to three
    panels per workgroup).  Values, energy and gradient agree to 1e-13
    across the launch shapes; 30 rows are checked against the oracle."""
    from dbsr3_amd import Engine
    g, d = setups["small"]
    norb = 60
    fs = random_orbitals(g, norb, seed=600 + kind)
    rng = np.random.default_rng(600 + kind)
    rows = make_rows(rng, norb, 0, 30000)
    rows[:, 0] = 2
    coef = rng.uniform(-1, 1, len(rows))
    print(f"[chunked panels kind={kind} rows={len(rows)}]", flush=True)
    eng = Engine([(d, norb, rows)], kind=kind)
    eng.set_orbitals(0, pq_of(fs))
    eng.set_coef(0, coef)
    eng.set_tuning([0, 256, 4, 4096, 512, 1])
    eng.forward_adjoint()
    v0, g0, e0 = eng.values(), eng.grad(), eng.energy()
    sel = rng.choice(len(rows), 30, replace=False)
    ref = oracle_values(g, fs, rows[sel], kind)
    assert relerr(v0[sel], ref).max() < TOL
    gs = np.max(np.abs(g0))
    assert gs > 0
    for t in ([0, 256, 8, 4096, 512, 1], [0, 256, 16, 4096, 512, 1], [0, 256, 16, 4096, 256, 3],
              [0, 256, 8, 64, 256, 2], [0, 256, 4, 4096, 512, 2, 2], [0, 256, 4, 4096, 512, 4, 4]):
        print(f"  tuning {t}", flush=True)
        eng.set_tuning(t)
        eng.forward_adjoint()
        assert np.max(relerr(eng.values(), v0)) < 1e-13, t
        assert np.max(np.abs(eng.grad() - g0)) <= 1e-13 * gs, t
        assert abs(eng.energy() - e0) <= 1e-13 * np.sum(np.abs(coef * v0)), t
        eng.forward()
        assert np.max(relerr(eng.values(), v0)) < 1e-13, t


@pytest.mark.parametrize("kind", [0, 3])
def test_features_beyond_64_orbitals(setups, kind):
    """more than 64 orbitals: the features come from the per-density kernel
    instead of the matrix-core Gram kernel"""
    from dbsr3_amd import Engine
    g, d = setups["small"]
    norb = 70
    fs = random_orbitals(g, norb, seed=70 + kind)
    rng = np.random.default_rng(70 + kind)
    rows = make_rows(rng, norb, 2, 200)
    coef = rng.uniform(-1, 1, len(rows))
    print(f"[features fallback kind={kind}]", flush=True)
    eng = Engine([(d, norb, rows)], kind=kind)
    eng.set_orbitals(0, pq_of(fs))
    eng.set_coef(0, coef)
    eng.forward_adjoint()
    ref = oracle_values(g, fs, rows, kind)
    assert relerr(eng.values(), ref).max() < TOL
    if kind == 0:
        Gref = oracle_gradient(g, fs, rows, coef)
        assert np.max(np.abs(eng.grad() - Gref)) < TOL * np.max(np.abs(Gref))


def test_large_group_global_panels(setups):
    """one multipole with ~2080 total densities (64 orbitals): the panel of
    the group exceeds LDS and the pair panels stage its partners in LDS
    chunk by chunk.  Only 40 rows carry a coefficient, so the gradient is
    checked against the oracle's potentials of those rows."""
    from dbsr3_amd import Engine
    g, d = setups["small"]
    norb = 64
    fs = random_orbitals(g, norb, seed=64)
    rng = np.random.default_rng(64)
    rows = make_rows(rng, norb, 0, 8000)
    rows[:, 0] = 1
    sel = rng.choice(len(rows), 40, replace=False)
    coef = np.zeros(len(rows))
    coef[sel] = rng.uniform(-1, 1, len(sel))
    print(f"[large group rows={len(rows)}]", flush=True)
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq_of(fs))
    eng.set_coef(0, coef)
    eng.set_tuning([0])
    eng.forward_adjoint()
    ref = oracle_values(g, fs, rows[sel], 0)
    assert relerr(eng.values()[sel], ref).max() < TOL
    Gref = oracle_gradient(g, fs, rows[sel], coef[sel])
    assert np.max(np.abs(eng.grad() - Gref)) < TOL * np.max(np.abs(Gref))
    assert abs(eng.energy() - coef[sel] @ ref) <= TOL * np.sum(np.abs(coef[sel] * ref))


@pytest.mark.parametrize("kind", [0, 3, 5])
def test_many_orbitals_fallback_paths(setups, kind):
    """60 orbitals: the LDS-staged kernels fall back to their global-memory
    forms where the per-interval tiles exceed LDS; values and adjoint stay
    on the oracle"""
    from dbsr3_amd import Engine
    g, d = setups["small"]
    norb = 60
    fs = random_orbitals(g, norb, seed=77)
    rng = np.random.default_rng(77)
    rows = make_rows(rng, norb, 3, 150)
    coef = rng.uniform(-1, 1, len(rows))
    eng = Engine([(d, norb, rows)], kind=kind)
    eng.set_orbitals(0, pq_of(fs))
    eng.set_coef(0, coef)
    eng.forward_adjoint()
    ref = oracle_values(g, fs, rows, kind)
    assert relerr(eng.values(), ref).max() < TOL
    if kind == 0:
        Gref = oracle_gradient(g, fs, rows, coef)
        assert np.max(np.abs(eng.grad() - Gref)) < TOL * np.max(np.abs(Gref))


@pytest.mark.parametrize("name,norb", [("small", 12), ("fe", 12), ("small", 28), ("fe", 40)])
def test_complete_groups_register_adjoint(setups, name, norb):
    """every multipole lists every orbital pair (a <= c), so the adjoint
    sums each pair's gradients over the groups in registers (78 to 820
    densities: one to four per thread); gradient against the oracle and
    against central differences of E"""
    from dbsr3_amd import Engine
    g, d = setups[name]
    fs = random_orbitals(g, norb, seed=7 * norb)
    rng = np.random.default_rng(norb)
    pairs = [(a, c) for a in range(1, norb + 1) for c in range(a, norb + 1)]
    rows = []
    for k in range(4):
        for (a, c) in pairs:
            b, dd = pairs[int(rng.integers(len(pairs)))]
            rows.append((k, a, b, c, dd))
    rows = np.array(rows, dtype=np.int32)
    coef = rng.uniform(-1, 1, len(rows))
    print(f"[complete groups {name} norb={norb} rows={len(rows)}]", flush=True)
    eng = Engine([(d, norb, rows)])
    eng.set_orbitals(0, pq_of(fs))
    eng.set_coef(0, coef)
    eng.forward_adjoint()
    G = eng.grad()
    sel = rng.choice(len(rows), 60, replace=False)
    c2 = np.zeros(len(rows))
    c2[sel] = coef[sel]
    eng.set_coef(0, c2)
    eng.forward_adjoint()
    Gref = oracle_gradient(g, fs, rows[sel], coef[sel])
    assert np.max(np.abs(eng.grad() - Gref)) < TOL * np.max(np.abs(Gref))
    # central differences of the full E on two coefficients
    eng.set_coef(0, coef)
    pq = pq_of(fs)
    for (j, comp, o) in [(3, 0, 1), (g.ns // 2, 1, norb - 1)]:
        e_pm = []
        for sgn in (1, -1):
            p2 = pq.copy(order="F")
            p2[j, comp, o] += sgn * 1e-5
            eng.set_orbitals(0, p2)
            eng.forward()
            e_pm.append(eng.energy())
        fd = (e_pm[0] - e_pm[1]) / 2e-5
        assert abs(fd - G[j, comp, o]) <= 1e-6 * max(abs(fd), 1e-3 * np.max(np.abs(G)))


@pytest.mark.parametrize("kind", [0, 1, 5])
def test_memory_budget_is_bit_exact(setups, kind):
    """the large-ns path: under a memory budget the (instance, k) groups
    run in batches whose features and gradients fit, the buffers reused;
    two instances of different grids (fe and small knots) so batches cut
    across an instance.  Values, energies and gradients equal the path that
    holds every group's features at once, bit for bit (the same adjoint
    kernels, a batch continuing the orbital-point sums of the earlier ones),
    in eager runs and in graph replays; against the default path (which may
    take the register-summed adjoint kernel) within 1e-13."""
    from dbsr3_amd import Engine
    (gf, df), (gs, ds) = setups["fe"], setups["small"]
    rng = np.random.default_rng(700 + kind)
    insts, pqs, coefs = [], [], []
    for (g, d, norb, kmax, n) in ((gf, df, 9, 5, 900), (gs, ds, 12, 6, 1500)):
        fs = random_orbitals(g, norb, seed=701 + kind + norb)
        rows = make_rows(rng, norb, kmax, n)
        insts.append((d, norb, rows))
        pqs.append(pq_of(fs))
        coefs.append(rng.uniform(-1, 1, n))
    print(f"[memory budget kind={kind}]", flush=True)
    eng = Engine(insts, kind=kind)
    for i in range(2):
        eng.set_orbitals(i, pqs[i])
        eng.set_coef(i, coefs[i])

    def run():
        eng.forward_adjoint()
        return ([eng.values(i) for i in range(2)], [eng.grad(i) for i in range(2)],
                [eng.energy(i) for i in range(2)])

    v0, g0, e0 = run()                       # default
    eng.set_memory_budget(1 << 40)           # one batch, no register-summed adjoint
    nb1, by1 = eng.memory()
    v1, g1, e1 = run()
    assert nb1 == 1
    for i in range(2):
        assert np.max(relerr(v1[i], v0[i])) < 1e-13
        assert np.max(np.abs(g1[i] - g0[i])) <= 1e-13 * np.max(np.abs(g0[i]))
    for frac in (0.45, 0.2, 1e-9):           # several batches; one group per batch
        eng.set_memory_budget(max(1, int(by1 * frac)))
        nb, by = eng.memory()
        print(f"  budget {frac}: {nb} batches, {by} bytes of {by1}", flush=True)
        assert nb > 1 and (by <= by1 * frac or frac < 1e-6)
        v2, g2, e2 = run()
        for i in range(2):
            assert np.array_equal(v2[i], v1[i]), (frac, i)
            assert np.array_equal(g2[i], g1[i]), (frac, i)
            assert e2[i] == e1[i], (frac, i)
        eng.forward()
        for i in range(2):
            assert np.array_equal(eng.values(i), v1[i])
    # graph replays under a budget
    eng.graph_capture(with_adjoint=True)
    for _ in range(3):
        eng.graph_launch()
        eng.sync()
        for i in range(2):
            assert np.array_equal(eng.grad(i), g1[i]) and np.array_equal(eng.values(i), v1[i])
    eng.set_memory_budget(0)
    v3, g3, e3 = run()
    for i in range(2):
        assert np.array_equal(v3[i], v0[i]) and np.array_equal(g3[i], g0[i])
