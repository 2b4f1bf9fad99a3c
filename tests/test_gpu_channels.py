"""GPU parity of DBSR_MAT3's channel-block post-processing with the oracle
(dbs_channels.hip: DBS_ORTH, Diag_channels, Record_matrix).

Tolerances, stated per quantity:
* DBS_ORTH: the same products as the reference's matmul sequence summed in
  another order (batched rocBLAS): every block entry to 1e-12 of the
  block's largest entry; blocks and columns outside idiag's selection
  bit-identical.
* Diag_channels: rocSOLVER dsygvd against LAPACK dsygv, two backward-stable
  solvers of the same problem: ipsol identical; eigenvalues to 1e-12 of the
  channel's largest |E| (the norm of the reduced matrix L^-1 H L^-T, which
  bounds both solvers' eigenvalue errors: the B-spline overlaps here have
  condition numbers near 2e8, so |E| reaches 1e8 while |H| stays near 1e4);
  eigenvectors: backward error 1e-13, S-normalised, largest entry positive,
  and equal to the oracle's to (1e-13 + 2e-14 max|E|/gap) of their largest
  entry, the perturbation bound for an eigenvalue separated by gap from the
  rest of the spectrum; the removed B-splines exactly zero and the rest of
  diag bit-identical.
* Record_matrix: the same two products per block in another summation
  order: record keys and offsets identical, values to 1e-12 of each
  record's largest entry (computed from the GPU's own solutions, so the
  eigen-solver's differences do not enter), overlaps likewise.
"""
import numpy as np
import pytest

import oracle as O
from conftest import fe_grid, small_grid
from test_oracle_channels import dirac_case, projector_case

pytestmark = pytest.mark.gpu


def _blocks(d, nch, npert, idiag, hch, hcp=None, hp=None):
    import torch
    from dbsr3_amd import MatBlocks
    m = MatBlocks(d, nch, npert, idiag=idiag)
    m.hch.copy_(torch.from_numpy(np.ascontiguousarray(hch.T)))
    if hcp is not None:
        m.hcp.copy_(torch.from_numpy(np.ascontiguousarray(hcp.T)))
    if hp is not None:
        m.hp.copy_(torch.from_numpy(np.ascontiguousarray(hp)))
    return m


def _f(t):
    return np.asfortranarray(t.cpu().numpy().T)


@pytest.mark.parametrize("idiag", [1, -1, 0])
def test_dbs_orth_matches_oracle(gpu_available, idiag):
    from dbsr3_amd import DBS
    from dbsr3_amd.channels import DBS_ORTH, orth_conditions
    from dbsr3_amd.dbs import fortran_to_device
    c = projector_case(7, nch=5, ns=18, npert=2)
    ms = c["pv"].shape[0]
    ns = ms // 2
    t = O.mkgrid_02(ns, 5, 0.25, 20.0, 4.0)
    d = DBS(t, ns, 5, 4, 5)
    m = _blocks(d, c["nch"], c["npert"], 0, c["hch"], c["hcp"])
    assert np.array_equal(m.icc, c["icc"]) and np.array_equal(m.icb, c["icb"])
    nort, iort = orth_conditions(c["ipch"], c["ipbs"], c["kbs"], c["ibort"])
    DBS_ORTH(m, idiag, nort, iort, fortran_to_device(c["pv"]), fortran_to_device(c["qv"]))
    h, hp = c["hch"].copy(order="F"), c["hcp"].copy(order="F")
    O.dbs_orth(h, c["icc"], idiag, c["orth"], c["pv"], c["qv"], hp, c["icb"], c["npert"])
    gh, gp = _f(m.hch), _f(m.hcp)
    for b in range(h.shape[2]):
        if np.array_equal(h[:, :, b], c["hch"][:, :, b]):
            assert np.array_equal(gh[:, :, b], h[:, :, b]), b
        else:
            assert np.max(np.abs(gh[:, :, b] - h[:, :, b])) <= 1e-12 * np.max(np.abs(h[:, :, b])), b
    for col in range(hp.shape[1]):
        if idiag != 0:
            assert np.array_equal(gp[:, col], hp[:, col])
        else:
            assert np.max(np.abs(gp[:, col] - hp[:, col])) <= 1e-12 * np.max(np.abs(hp[:, col]))


def _perturbed_dirac(kappas, Z, grid, seed, eps=1e-3):
    g, fppqq, hch, diag, icc, iprm, lch = dirac_case(kappas, Z, grid)
    rng = np.random.default_rng(seed)
    for i in range(len(kappas)):
        R = rng.standard_normal(hch.shape[:2])
        hch[:, :, i] += eps * (R + R.T)
    return g, fppqq, hch, diag, icc, iprm, lch


def _diag_compare(g, hch, ovl, d_gpu, d_ref, ipsol, iprm, icc):
    """GPU solutions against the oracle's: each GPU pair (E, v) solves the
    compacted problem to a backward error of 1e-13 (|H v - E S v| against
    (|H| + |E| |S|) |v|, infinity norms; LAPACK's own is 4e-16 here), is
    S-normalised and has its largest entry positive; E agrees to 1e-12 of
    the channel's largest |E| and v to (1e-13 + 2e-14 max|E| / gap) of its
    largest entry, the perturbation bound of a backward-stable solver for an
    eigenvalue separated by gap from the rest of the spectrum"""
    from scipy.linalg.lapack import dsygv
    ns, ms = g.ns, g.ms
    for ich in range(1, icc.shape[0] + 1):
        k = icc[ich - 1, ich - 1] - 1
        ksol = ipsol[ich]
        keep = np.nonzero(iprm[(ich - 1) * ms:ich * ms])[0]
        # the problem both solvers see: the lower triangles (uplo = 'L')
        A = np.tril(hch[np.ix_(keep, keep, [k])][:, :, 0])
        A = A + np.tril(A, -1).T
        B = np.tril(ovl[np.ix_(keep, keep, [k])][:, :, 0])
        B = B + np.tril(B, -1).T
        w_all = dsygv(np.asfortranarray(A), np.asfortranarray(B), 1, "N", "L")[0]
        emax = np.max(np.abs(w_all))
        nA, nB = np.max(np.sum(np.abs(A), 1)), np.max(np.sum(np.abs(B), 1))
        E_gpu = np.empty(ksol)
        E_ref = np.empty(ksol)
        for s in range(ksol):
            E_ref[s] = d_ref[s, ns - 1, k] if (ksol < ns or s >= ns - 1) else np.nan
            E_gpu[s] = d_gpu[s, ns - 1, k] if (ksol < ns or s >= ns - 1) else np.nan
        cols = [s for s in range(ksol) if not (ksol >= ns and s == ns - 1)]
        for s in cols:
            v, r = d_gpu[:, s, k], d_ref[:, s, k]
            x = v[keep]
            assert x[int(np.argmax(np.abs(x)))] > 0.0, (ich, s)
            assert abs(x @ B @ x - 1.0) <= 1e-12, (ich, s)
            if not np.isnan(E_gpu[s]):
                res = A @ x - E_gpu[s] * (B @ x)
                assert np.max(np.abs(res)) <= 1e-13 * (nA + abs(E_gpu[s]) * nB) * np.max(np.abs(x)), (ich, s)
                j = int(np.argmin(np.abs(w_all - E_ref[s])))
                gap = np.min(np.abs(np.delete(w_all, j) - w_all[j]))
                tol = 1e-13 + 2e-14 * emax / gap
                assert np.max(np.abs(v - r)) <= tol * np.max(np.abs(r)), (ich, s)
        assert np.all(d_gpu[np.setdiff1d(np.arange(ms), keep)][:, cols, k] == 0.0)
        ok = ~np.isnan(E_ref)
        assert np.max(np.abs(E_gpu[ok] - E_ref[ok]), initial=0.0) <= 1e-12 * emax, ich
        if ksol < ns:
            assert np.array_equal(d_gpu[ksol:, ns - 1, k], d_ref[ksol:, ns - 1, k])
        rest = [c for c in range(ms) if c >= ksol and c != ns - 1]
        assert np.array_equal(d_gpu[:, rest, k], d_ref[:, rest, k])


def test_diag_channels_matches_oracle(gpu_available):
    import torch
    from dbsr3_amd import DBS
    from dbsr3_amd.channels import Diag_channels
    kappas = [-1, 1, -2, 2, -3, -1]
    g, fppqq, hch, diag, icc, iprm, lch = _perturbed_dirac(kappas, 26.0, None, 11)
    t, ns, ks, ksp, ksq = fe_grid()
    d = DBS(t, ns, ks, ksp, ksq)
    m = _blocks(d, len(kappas), 0, 1, hch)
    dg = torch.from_numpy(np.ascontiguousarray(diag.T)).cuda()
    ipsol = Diag_channels(m, dg, iprm)
    d_ref = diag.copy(order="F")
    ip_ref = O.diag_channels(hch, d_ref, icc, iprm, ns)
    assert np.array_equal(ipsol, ip_ref)
    _diag_compare(g, hch, diag, _f(dg), d_ref, ipsol, iprm, icc)


def test_diag_channels_every_solution(gpu_available):
    """the window open: more solutions than ns, so solution ns overwrites
    the stored energies and later energies overwrite it, as in the
    reference; channels of two kept sizes in two batches"""
    import torch
    from dbsr3_amd import DBS
    from dbsr3_amd.channels import Diag_channels
    kappas = [-1, 2, -1]
    g, fppqq, hch, diag, icc, iprm, lch = _perturbed_dirac(kappas, 10.0, small_grid(), 5)
    t, ns, ks, ksp, ksq = small_grid()
    d = DBS(t, ns, ks, ksp, ksq)
    m = _blocks(d, len(kappas), 0, 1, hch)
    dg = torch.from_numpy(np.ascontiguousarray(diag.T)).cuda()
    ipsol = Diag_channels(m, dg, iprm, edmin=-1e300, edmax=1e300, egap=0.0)
    d_ref = diag.copy(order="F")
    ip_ref = O.diag_channels(hch, d_ref, icc, iprm, ns, edmin=-1e300, edmax=1e300, egap=0.0)
    assert np.array_equal(ipsol, ip_ref) and np.all(ipsol[1:] > ns)
    dgh = _f(dg)
    _diag_compare(g, hch, diag, dgh, d_ref, ipsol, iprm, icc)
    for ich in range(1, len(kappas) + 1):
        k = icc[ich - 1, ich - 1] - 1
        ksol = ipsol[ich]
        rows = np.arange(g.ms)
        eig = (rows >= ns - 1) & (rows < ksol)
        Er = d_ref[eig, ns - 1, k]
        assert np.max(np.abs(dgh[eig, ns - 1, k] - Er)) <= 1e-12 * np.max(np.abs(Er))
        # the rest of column ns is solution ns's vector, to the bound of
        # _diag_compare for its eigenvalue
        from scipy.linalg.lapack import dsygv
        keep = np.nonzero(iprm[(ich - 1) * g.ms:ich * g.ms])[0]
        w_all = dsygv(np.asfortranarray(hch[np.ix_(keep, keep, [k])][:, :, 0]),
                      np.asfortranarray(diag[np.ix_(keep, keep, [k])][:, :, 0]), 1, "N", "L")[0]
        gap = min(w_all[ns - 1] - w_all[ns - 2], w_all[ns] - w_all[ns - 1])
        r = d_ref[~eig, ns - 1, k]
        tol = 1e-13 + 2e-14 * np.max(np.abs(w_all)) / gap
        assert np.max(np.abs(dgh[~eig, ns - 1, k] - r)) <= tol * np.max(np.abs(r))


def test_diag_channels_rejects_indefinite_overlap(gpu_available):
    import torch
    from dbsr3_amd import DBS, DBSError
    from dbsr3_amd.channels import Diag_channels
    g, fppqq, hch, diag, icc, iprm, lch = dirac_case([-1], 10.0, small_grid())
    t, ns, ks, ksp, ksq = small_grid()
    d = DBS(t, ns, ks, ksp, ksq)
    m = _blocks(d, 1, 0, 1, hch)
    dg = torch.from_numpy(np.ascontiguousarray((-diag).T)).cuda()
    with pytest.raises(DBSError, match="channel diagonalization failed"):
        Diag_channels(m, dg, iprm)


def test_record_matrix_matches_oracle(gpu_available):
    import torch
    from dbsr3_amd import DBS
    from dbsr3_amd.channels import Diag_channels, Record_matrix
    kappas = [-1, 1, -2, -1]
    nch, npert = len(kappas), 3
    g, fppqq, hch_d, diag, icc_d, iprm, lch = _perturbed_dirac(kappas, 20.0, None, 2)
    t, ns, ks, ksp, ksq = fe_grid()
    d = DBS(t, ns, ks, ksp, ksq)
    md = _blocks(d, nch, 0, 1, hch_d)
    dg = torch.from_numpy(np.ascontiguousarray(diag.T)).cuda()
    ipsol = Diag_channels(md, dg, iprm)
    sol = _f(dg)
    rng = np.random.default_rng(4)
    from dbsr3_amd import MatBlocks
    m0 = MatBlocks(d, nch, npert, idiag=-1)
    nb, ms = m0.nblocks, g.ms
    hch = np.asfortranarray(rng.standard_normal((ms, ms, nb)))
    hch[:, :, m0.icc[2, 1] - 1] = 0.0  # not recorded
    hch[:, :, m0.icc[3, 0] - 1] *= 1e-22  # below 1e-20: not recorded
    hcp = np.asfortranarray(rng.standard_normal((ms, m0.ncb)))
    hcp[:, m0.icb[1, 0] - 1] = 0.0
    hp = rng.standard_normal(m0.nbb)
    hp[m0.ibb[1, 0] - 1] = 0.0
    m = _blocks(d, nch, npert, -1, hch, hcp, hp)
    recs, ov = Record_matrix(m, dg, ipsol, icase=0)
    rrefs, ovr = O.record_matrix(hch, m.icc, sol, ipsol, 0, hcp, m.icb, hp, m.ibb, npert)
    assert [r[0] for r in recs] == [r[0] for r in rrefs]
    for (key, v), (_, r) in zip(recs[:-1], rrefs[:-1]):
        if isinstance(r, float):
            assert v == r
            continue
        assert v.shape == r.shape and np.max(np.abs(v - r)) <= 1e-12 * np.max(np.abs(r)), key
    assert np.max(np.abs(ov - ovr)) <= 1e-12 * np.max(np.abs(ovr))
    assert np.array_equal(ov == 0, ovr == 0)


def test_channel_post_processing_config2_size(gpu_available):
    """config 2's partial wave (40 channels, ns=150, 820 blocks of 300x300;
    bench.config7_case) through the four steps, each stage against the
    oracle run on the same input: DBS_ORTH on the diagonal blocks, then
    Diag_channels on the GPU's projected blocks, DBS_ORTH on the other
    blocks, Record_matrix on the GPU's projected blocks and solutions"""
    import torch
    from bench import CONFIG7, config7_case
    from dbsr3_amd import DBS, MatBlocks
    from dbsr3_amd.channels import DBS_ORTH, Diag_channels, Record_matrix, Zero_cond, orth_conditions
    from dbsr3_amd.dbs import fortran_to_device
    from dbsr3_amd.grid import mkgrid_02
    P = CONFIG7
    ns, ks, ksp, ksq, nch = P["ns"], P["ks"], P["ksp"], P["ksq"], P["nch"]
    t = mkgrid_02(ns, ks, P["z"], 0.25, 30.0)
    d = DBS(t, ns, ks, ksp, ksq)
    g = O.Grid(t, ns, ks, ksp, ksq)
    case = config7_case(d)
    print(f"[channels nch={nch} ns={ns}]", flush=True)
    m = _blocks(d, nch, 0, 0, case["hch"])
    ovl = np.asfortranarray(np.repeat(case["fppqq"][:, :, None], nch, axis=2))
    dg = torch.from_numpy(np.ascontiguousarray(ovl.T)).cuda()
    iprm, nhm = Zero_cond(nch, d, case["lch"], case["kap"])
    iprm_o, nhm_o = O.zero_cond(nch, ns, g.nsp, g.nsq, ksp, ksq, case["lch"], case["kap"])
    assert np.array_equal(iprm, iprm_o) and nhm == nhm_o
    nort, iort = orth_conditions(case["ipch"], case["ipbs"], case["kbs"], case["ibort"])
    orth = O.orth_conditions(nch, case["ipch"], case["ipbs"], case["kbs"], case["ibort"])
    assert all(list(iort[i, :nort[i]]) == orth[i] for i in range(nch)) and max(nort) > 0
    pv, qv = fortran_to_device(case["pv"]), fortran_to_device(case["qv"])
    icc = m.icc
    # DBS_ORTH, diagonal blocks
    DBS_ORTH(m, 1, nort, iort, pv, qv)
    h1 = _f(m.hch)
    ref = case["hch"].copy(order="F")
    O.dbs_orth(ref, icc, 1, orth, case["pv"], case["qv"])
    err = max(float(np.max(np.abs(h1[:, :, b] - ref[:, :, b])) / np.max(np.abs(ref[:, :, b])))
              for b in range(nch))
    print(f"  orth(+1) max rel {err:.2e}", flush=True)
    assert err <= 1e-12 and np.array_equal(h1[:, :, nch:], case["hch"][:, :, nch:])
    # Diag_channels on the projected diagonal blocks
    ipsol = Diag_channels(m, dg, iprm, edmax=P["edmax"])
    d_ref = ovl.copy(order="F")
    ip_ref = O.diag_channels(h1, d_ref, icc, iprm, ns, edmax=P["edmax"])
    assert np.array_equal(ipsol, ip_ref) and np.all(ipsol[1:] < ns)
    sol = _f(dg)
    _diag_compare(g, h1, ovl, sol, d_ref, ipsol, iprm, icc)
    # DBS_ORTH, the other blocks
    DBS_ORTH(m, -1, nort, iort, pv, qv)
    h2 = _f(m.hch)
    O.dbs_orth(ref, icc, -1, orth, case["pv"], case["qv"])
    err = max(float(np.max(np.abs(h2[:, :, b] - ref[:, :, b])) / np.max(np.abs(ref[:, :, b])))
              for b in range(nch, m.nblocks))
    print(f"  orth(-1) max rel {err:.2e}", flush=True)
    assert err <= 1e-12 and np.array_equal(h2[:, :, :nch], h1[:, :, :nch])
    # Record_matrix on the GPU's projected blocks and solutions
    recs, ov = Record_matrix(m, dg, ipsol, icase=0)
    rrefs, ovr = O.record_matrix(h2, icc, sol, ipsol, 0)
    assert [r[0] for r in recs] == [r[0] for r in rrefs] and len(recs) == m.nblocks - nch + 1
    err = max(float(np.max(np.abs(v - r)) / np.max(np.abs(r))) for (_, v), (_, r) in zip(recs[:-1], rrefs[:-1]))
    print(f"  records max rel {err:.2e}", flush=True)
    assert err <= 1e-12
    assert np.max(np.abs(ov - ovr)) <= 1e-12 * np.max(np.abs(ovr))


def test_post_processing_in_small_batches(gpu_available):
    """a 64 KiB workspace limit splits DBS_ORTH (36 blocks, 18 per batch),
    the eigen-solves (one channel per batch) and Record_matrix (one record
    per batch) into batches, the pointer and host tables rewritten between
    them: the same results as the oracle"""
    import torch
    from dbsr3_amd import DBS, MatBlocks
    from dbsr3_amd.channels import DBS_ORTH, Diag_channels, Record_matrix, orth_conditions
    from dbsr3_amd.dbs import fortran_to_device
    # projections: 36 blocks of 36x36 with 3 conditions, 3.5 KB of workspace each
    c = projector_case(9, nch=8, ns=18, npert=2)
    ns = c["pv"].shape[0] // 2
    d = DBS(O.mkgrid_02(ns, 5, 0.25, 20.0, 4.0), ns, 5, 4, 5)
    d.set_workspace_limit(64 << 10)
    m = _blocks(d, c["nch"], c["npert"], 0, c["hch"], c["hcp"])
    nort, iort = orth_conditions(c["ipch"], c["ipbs"], c["kbs"], c["ibort"])
    DBS_ORTH(m, 0, nort, iort, fortran_to_device(c["pv"]), fortran_to_device(c["qv"]))
    h, hp = c["hch"].copy(order="F"), c["hcp"].copy(order="F")
    O.dbs_orth(h, c["icc"], 0, c["orth"], c["pv"], c["qv"], hp, c["icb"], c["npert"])
    gh = _f(m.hch)
    for b in range(h.shape[2]):
        assert np.max(np.abs(gh[:, :, b] - h[:, :, b])) <= 1e-12 * np.max(np.abs(h[:, :, b])), b
    # eigen-solves (0.8 MB per channel) and records (0.3 MB each)
    kappas = [-1, 1, -2, 2, -1]
    g, fppqq, hch_d, diag, icc_d, iprm, lch = _perturbed_dirac(kappas, 26.0, None, 3)
    t, ns, ks, ksp, ksq = fe_grid()
    d2 = DBS(t, ns, ks, ksp, ksq)
    d2.set_workspace_limit(64 << 10)
    md = _blocks(d2, len(kappas), 0, 1, hch_d)
    dg = torch.from_numpy(np.ascontiguousarray(diag.T)).cuda()
    ipsol = Diag_channels(md, dg, iprm)
    d_ref = diag.copy(order="F")
    assert np.array_equal(ipsol, O.diag_channels(hch_d, d_ref, icc_d, iprm, ns))
    sol = _f(dg)
    _diag_compare(g, hch_d, diag, sol, d_ref, ipsol, iprm, icc_d)
    mo = MatBlocks(d2, len(kappas), 0, idiag=-1)
    rng = np.random.default_rng(8)
    hch = np.asfortranarray(rng.standard_normal((g.ms, g.ms, mo.nblocks)))
    m2 = _blocks(d2, len(kappas), 0, -1, hch)
    recs, ov = Record_matrix(m2, dg, ipsol, icase=0)
    rrefs, ovr = O.record_matrix(hch, m2.icc, sol, ipsol, 0)
    assert [r[0] for r in recs] == [r[0] for r in rrefs]
    for (key, v), (_, r) in zip(recs[:-1], rrefs[:-1]):
        assert np.max(np.abs(v - r)) <= 1e-12 * np.max(np.abs(r)), key
