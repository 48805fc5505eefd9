"""Batched ILU(0)-GMRES over many uneven systems and several RHS per system
(the 512-system workload's path, at small scale): every (system, RHS) pair
must match the oracle's single-system GMRES run -- iterations within one
step, FP64 true relative residual at the tolerance -- and replays of the
per-cycle graph must reproduce the same bits."""
from __future__ import annotations

import numpy as np
import pytest

import par_ilu0_cpp_amd as P
from par_ilu0_cpp_amd import generators as gen

pytestmark = pytest.mark.gpu


def make_batch(sides, seed):
    rng = np.random.default_rng(seed)
    return [gen.diffusion2d(rng.lognormal(0.0, 1.0, size=(s, s))) for s in sides]


def to_dev(*arrs):
    import torch

    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


@pytest.mark.parametrize("nrhs,bits", [(4, 64), (8, 64), (2, 32), (1, 64)])
def test_batched_matches_per_system_oracle(oracle, nrhs, bits):
    import torch

    mats = make_batch([10, 17, 23, 31, 12, 40, 45], seed=nrhs)
    rp, ci, va, sp = P.block_diagonal(mats)
    n = rp.size - 1
    rng = np.random.default_rng(100 + nrhs)
    B = rng.uniform(-1, 1, size=(n, nrhs))
    d_rp, d_ci, d_va, d_b = to_dev(rp, ci, va, B)
    d_x = torch.zeros((n, nrhs), dtype=torch.float64, device="cuda")
    bg = P.BatchGmres()
    bg.SetFactorPrecision(bits)
    assert bg.Setup(n, d_rp.data_ptr(), d_ci.data_ptr(), 0, sp, nrhs) == ci.size
    assert bg.Factorize(d_va.data_ptr()) == 0
    m, tol = 30, 1e-8
    cyc, its, conv, rel = bg.Solve(d_rp.data_ptr(), d_ci.data_ptr(), d_va.data_ptr(), d_b.data_ptr(),
                                   d_x.data_ptr(), tol, m, 500)
    X = d_x.cpu().numpy()
    assert conv.all() and np.all(rel <= 1.01 * tol)
    for s, a in enumerate(mats):
        fac, dg, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, a.value)
        if bits == 32:
            fac = fac.astype(np.float32).astype(np.float64)
        for r in range(nrhs):
            b = B[sp[s]:sp[s + 1], r]
            ok_o, its_o, x_o, _ = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, tol * np.linalg.norm(b), m, 500,
                                               factors=(a.row_ptr, a.col_idx, dg, fac))
            assert ok_o and abs(int(its[s, r]) - its_o) <= 1, (s, r, its[s, r], its_o)
            x = X[sp[s]:sp[s + 1], r]
            assert np.linalg.norm(b - a.matvec(x)) <= 1.01 * tol * np.linalg.norm(b)
    # replay the captured cycle graph: identical bits
    d_x2 = torch.zeros_like(d_x)
    cyc2, its2, conv2, rel2 = bg.Solve(d_rp.data_ptr(), d_ci.data_ptr(), d_va.data_ptr(), d_b.data_ptr(),
                                       d_x2.data_ptr(), tol, m, 500)
    assert cyc2 == cyc and np.array_equal(its2, its) and torch.equal(d_x2, d_x)


def test_batched_factors_are_per_system_factors(oracle):
    import torch

    mats = make_batch([9, 14, 20], seed=3)
    rp, ci, va, sp = P.block_diagonal(mats)
    n = rp.size - 1
    d_rp, d_ci, d_va = to_dev(rp, ci, va)
    ilu = P.IluSolver(0)
    assert ilu.SymbolicDevice(n, d_rp.data_ptr(), d_ci.data_ptr(), 0) == ci.size
    assert ilu.FactorizeDevice(d_va.data_ptr(), True) == 0
    fac = ilu.DeviceFactors()
    off = 0
    for a in mats:
        ofac, _, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, a.value)
        assert np.array_equal(fac[off:off + a.nnz], ofac)
        off += a.nnz


def test_batched_restart_boundary_and_iteration_cap(oracle):
    import torch

    mats = make_batch([30, 25], seed=5)
    rp, ci, va, sp = P.block_diagonal(mats)
    n = rp.size - 1
    B = np.random.default_rng(5).uniform(-1, 1, size=(n, 2))
    d_rp, d_ci, d_va, d_b = to_dev(rp, ci, va, B)
    bg = P.BatchGmres()
    assert bg.Setup(n, d_rp.data_ptr(), d_ci.data_ptr(), 0, sp, 2) > 0
    assert bg.Factorize(d_va.data_ptr()) == 0
    for maxits, expect_ok, expect_its in ((20, True, 20), (25, False, 26)):
        d_x = torch.zeros((n, 2), dtype=torch.float64, device="cuda")
        cyc, its, conv, rel = bg.Solve(d_rp.data_ptr(), d_ci.data_ptr(), d_va.data_ptr(), d_b.data_ptr(),
                                       d_x.data_ptr(), 1e-30, 10, maxits)
        assert np.all(its == expect_its) and np.all(conv == expect_ok)


def _solve_batch(mats, B, nrhs, m=30, tol=1e-8, maxits=800):
    import torch

    rp, ci, va, sp = P.block_diagonal(mats)
    n = rp.size - 1
    d_rp, d_ci, d_va, d_b = to_dev(rp, ci, va, B)
    d_x = torch.zeros((n, nrhs), dtype=torch.float64, device="cuda")
    bg = P.BatchGmres()
    assert bg.Setup(n, d_rp.data_ptr(), d_ci.data_ptr(), 0, sp, nrhs) == ci.size
    assert bg.Factorize(d_va.data_ptr()) == 0
    cyc, its, conv, rel = bg.Solve(d_rp.data_ptr(), d_ci.data_ptr(), d_va.data_ptr(), d_b.data_ptr(),
                                   d_x.data_ptr(), tol, m, maxits)
    return cyc, its, conv, d_x.cpu().numpy(), sp


def test_batched_systems_are_independent_of_their_neighbours():
    """A system that converges in a few steps sits next to one that needs
    many cycles: once the fast one has finished, its rows are skipped by the
    sweeps, the SpMV and the Gram-Schmidt kernels.  Each system's solution
    and step counts must be bit-identical to solving it alone, in either
    position of the batch."""
    rng = np.random.default_rng(5)
    fast = gen.diffusion2d(np.ones((6, 6)))
    slow = gen.diffusion2d(rng.lognormal(0.0, 2.0, size=(48, 48)))
    nrhs = 4
    Bf = rng.uniform(-1, 1, size=(fast.n, nrhs))
    Bs = rng.uniform(-1, 1, size=(slow.n, nrhs))
    _, its_f, conv_f, Xf, _ = _solve_batch([fast], Bf, nrhs)
    cyc_s, its_s, conv_s, Xs, _ = _solve_batch([slow], Bs, nrhs)
    assert conv_f.all() and conv_s.all() and cyc_s >= 3 and its_f.max() < 30
    for order in ((fast, slow), (slow, fast)):
        B = np.vstack([Bf if a is fast else Bs for a in order])
        _, its, conv, X, sp = _solve_batch(list(order), B, nrhs)
        assert conv.all()
        for k, a in enumerate(order):
            ref_x, ref_its = (Xf, its_f[0]) if a is fast else (Xs, its_s[0])
            assert np.array_equal(its[k], ref_its)
            assert np.array_equal(X[sp[k]:sp[k + 1]], ref_x)
