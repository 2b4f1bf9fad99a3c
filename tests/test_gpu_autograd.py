"""The adjoint as a torch.autograd.Function (test harness only: the library
itself has no torch types in its interface).

RkEnergy.apply(pq) = E = sum_n c_n R^k_n(pq) through the engine's forward;
its backward is the engine's adjoint dE/d(pq).  Checked:
  * torch.autograd.gradcheck: the adjoint against central finite
    differences of the engine's own forward, coordinate by coordinate;
  * the reference potentials: with the energy of DBSR_HF's coefficients
    (E = sum a(i,j,k) F^k(i,j) + b(i,j,k) G^k(i,j)) the gradient for
    orbital i is 2 qsum(i) hfm_i v_i, hfm_i the oracle's hf_matrix
    (DBSR_HF/hf_matrix.f90:41-59, bit-exact to the compiled reference);
  * the large-ns path: the same gradient under a memory budget (groups in
    batches), bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from conftest import random_orbitals, small_grid
from test_oracle_hf_matrix import coefs_ab, hf_case

pytestmark = pytest.mark.gpu


def rk_energy_fn(eng):
    import torch

    class RkEnergy(torch.autograd.Function):
        @staticmethod
        def forward(ctx, pq):
            x = pq.detach().cpu().numpy()
            eng.set_orbitals(0, np.asfortranarray(x))
            eng.forward_adjoint()
            ctx.save_for_backward(torch.from_numpy(np.ascontiguousarray(eng.grad())).to(pq.device))
            return torch.tensor(eng.energy(), dtype=torch.float64, device=pq.device)

        @staticmethod
        def backward(ctx, gout):
            (g,) = ctx.saved_tensors
            return gout * g

    return RkEnergy.apply


def test_gradcheck(gpu_available):
    import torch
    from dbsr3_amd import DBS, Engine
    t, ns, ks, ksp, ksq = small_grid()
    g = O.Grid(t, ns, ks, ksp, ksq)
    norb = 4
    fs = random_orbitals(g, norb, 41)
    rng = np.random.default_rng(41)
    rows = np.zeros((60, 5), dtype=np.int32)
    rows[:, 0] = rng.integers(0, 4, 60)
    rows[:, 1:] = rng.integers(1, norb + 1, (60, 4))
    eng = Engine([(DBS(t, ns, ks, ksp, ksq), norb, rows)])
    eng.set_coef(0, rng.uniform(-1, 1, 60))
    f = rk_energy_fn(eng)
    x = torch.tensor(np.stack(fs, axis=2), dtype=torch.float64, requires_grad=True)
    print(f"[gradcheck ns={ns} norb={norb}: {x.numel()} coordinates]", flush=True)
    assert torch.autograd.gradcheck(f, (x,), eps=1e-6, atol=1e-8, rtol=1e-6)


@pytest.mark.parametrize("budget", [0, 1])
def test_gradient_is_the_reference_potential(gpu_available, budget):
    import torch
    from dbsr3_amd import DBS, Engine
    kmax = 3
    g, (t, ns, ks, ksp, ksq), pq, lbs, jbs, coef, qsum, _ = hf_case(small_grid, kmax, 23, False)
    n = pq.shape[2]
    rows, cs = [], []
    for i in range(1, n + 1):
        A, B = coefs_ab(i, coef, lbs, jbs, kmax)
        for k in range(kmax + 1):
            for j in range(1, i + 1):
                if A[j - 1, k]:
                    rows.append((k, i, j, i, j)); cs.append(A[j - 1, k])
            for j in range(i + 1, n + 1):
                if B[j - 1, k]:
                    rows.append((k, i, j, j, i)); cs.append(B[j - 1, k])
    eng = Engine([(DBS(t, ns, ks, ksp, ksq), n, np.array(rows, dtype=np.int32))])
    eng.set_coef(0, np.array(cs))
    if budget:
        eng.set_memory_budget(1)   # one (instance, k) group per batch
        assert eng.memory()[0] == kmax + 1
    f = rk_energy_fn(eng)
    x = torch.tensor(pq, dtype=torch.float64, requires_grad=True)
    E = f(x)
    (gx,) = torch.autograd.grad(E, x)
    gx = gx.numpy()
    scale = np.max(np.abs(gx))
    print(f"[reference potentials budget={budget}: E = {E.item():.15e}]", flush=True)
    for i in range(1, n + 1):
        A, B = coefs_ab(i, coef, lbs, jbs, kmax)
        h = O.hf_matrix(g, i, pq, A, B, kmax, qsum[i - 1])
        v = np.concatenate([pq[:, 0, i - 1], pq[:, 1, i - 1]])
        pred = 2 * qsum[i - 1] * (h @ v)
        got = np.concatenate([gx[:, 0, i - 1], gx[:, 1, i - 1]])
        assert np.max(np.abs(pred - got)) <= 1e-11 * scale, i
    if budget:
        eng.set_memory_budget(1 << 40)  # every group held at once, same adjoint kernels
        (g2,) = torch.autograd.grad(f(x), x)
        assert np.array_equal(g2.numpy(), gx)
