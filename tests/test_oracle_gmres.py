"""Pins the oracle's restarted GMRES (oracle.c: orc_gmres, following
src/gmres.cpp:138-288).

Invariants and special cases that fix its behaviour:
  * with M = I and m >= n, GMRES terminates in at most n steps;
  * the Givens residual estimate never increases inside a cycle;
  * on success the true residual ||b - A x|| is at the (absolute) tolerance;
  * A = c I gives an exact breakdown (w = 0) at the first step, after which
    the back-solve with the un-rotated column still returns x = b / c;
  * max_iterations equal to a multiple of m returns success without
    convergence (the reference's restart-boundary quirk);
  * a NaN right-hand side returns failure;
  * the solution agrees with scipy's direct sparse solve.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from par_ilu0_cpp_amd import generators as gen
from tests.conftest import dominant_values, random_pattern


def factors_of(oracle, a, lof=0):
    frp, fci, _, fval, _ = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, lof)
    fac, dg, _ = oracle.ilu_factorize(frp, fci, fval)
    return frp, fci, dg, fac


def test_identity_preconditioner_terminates_within_n(oracle):
    rp, ci = random_pattern(40, 6, 39, seed=2)
    va = dominant_values(rp, ci, 2, dominance=1.2)
    b = np.random.default_rng(2).uniform(-1, 1, 40)
    ok, its, x, hist = oracle.gmres((rp, ci, va), b, 1e-10 * np.linalg.norm(b), 60, 200)
    assert ok and its <= 40


def test_estimate_monotone_within_cycle(oracle):
    a = gen.convdiff3d(10)
    b = gen.rhs_uniform(a.n, seed=1)
    m = 15
    ok, its, x, hist = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, 1e-9 * np.linalg.norm(b), m, 300,
                                    factors=factors_of(oracle, a))
    assert ok
    for c in range(0, hist.size, m):
        cyc = hist[c:c + m]
        assert np.all(np.diff(cyc) <= 1e-12 * cyc[0])


@pytest.mark.parametrize("lof", [0, 1])
def test_true_residual_at_tolerance(oracle, lof):
    a = gen.poisson2d(40)
    b = a.matvec(np.ones(a.n))
    tol = 1e-8
    ok, its, x, hist = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, tol, 30, 500,
                                    factors=factors_of(oracle, a, lof))
    assert ok
    assert hist[-1] <= tol
    assert np.linalg.norm(b - a.matvec(x)) <= 10 * tol
    assert np.max(np.abs(x - 1.0)) < 1e-6


def test_breakdown_with_scaled_identity(oracle):
    # b = ones(16): ||b|| = 4 and (v0, v0) = 1 exactly, so w = v0 - 1 * v0 = 0
    n = 16
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    va = np.full(n, 2.0)
    b = np.ones(n)
    fac = (rp, ci, np.arange(n, dtype=np.int32), va.copy())
    ok, its, x, hist = oracle.gmres((rp, ci, va), b, 1e-12, 5, 20, factors=fac)
    assert ok and its == 1 and hist.size == 0  # no estimate: the step broke down
    assert np.allclose(x, b / 2, rtol=1e-15)


def test_maxits_at_restart_boundary_reports_success(oracle):
    a = gen.convdiff3d(10)
    b = gen.rhs_uniform(a.n, seed=3)
    ok, its, x, hist = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, 1e-30, 10, 20)
    assert ok and its == 20  # not converged, still "success"
    ok2, its2, _, _ = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, 1e-30, 10, 25)
    assert not ok2 and its2 == 26  # mid-cycle exhaustion: its = maxits + 1, failure


def test_nan_rhs_fails(oracle):
    a = gen.poisson2d(8)
    b = np.ones(a.n)
    b[3] = np.nan
    ok, its, x, hist = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, 1e-8, 10, 50)
    assert not ok


def test_solution_matches_direct_solve(oracle):
    a = gen.convdiff3d(12)
    b = gen.rhs_uniform(a.n, seed=4)
    ok, its, x, hist = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, 1e-11 * np.linalg.norm(b), 40, 400,
                                    factors=factors_of(oracle, a, 1))
    ref = spla.spsolve(a.to_scipy().tocsc(), b)
    assert ok
    assert np.linalg.norm(x - ref) <= 1e-7 * np.linalg.norm(ref)


@pytest.mark.parametrize("sched", [0, 1])
def test_openmp_baseline_matches_serial_oracle(oracle, sched):
    """The threaded CPU baseline (ILU + GMRES on level sets, sched 0, or on
    the reference's subtree / queue schedule, sched 1) solves the same
    systems as the serial oracle: the same factors bit for bit (identical
    per-row arithmetic in a dependency-respecting order), iterations within
    one, true residual at the tolerance."""
    from oracle import py_oracle as O

    a = gen.convdiff3d(12)
    b = gen.rhs_uniform(a.n, seed=3)
    tol = 1e-8 * np.linalg.norm(b)
    frp, fci, _, fval, _ = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    fac, dg, _ = oracle.ilu_factorize(frp, fci, fval)
    ok_s, its_s, _, _ = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, tol, 30, 500, factors=(frp, fci, dg, fac))
    for threads in (1, 4):
        O.baseline_threads(threads)
        ret, its, x, times, f = O.baseline_ilu_gmres((a.row_ptr, a.col_idx, a.value), (frp, fci, fval), b, tol, 30,
                                                     500, sched=sched, return_factors=True)
        assert ret == 1 and ok_s and abs(its - its_s) <= 1
        assert np.array_equal(f, fac)
        assert np.linalg.norm(b - a.matvec(x)) <= 1.01 * tol
        assert np.all(times >= 0)


def _forest(k, side):
    from par_ilu0_cpp_amd import block_diagonal

    rp, ci, va, _ = block_diagonal([gen.poisson2d(side) for _ in range(k)])
    return rp, ci, va


@pytest.mark.parametrize("threads", [2, 3, 8])
def test_subtree_schedule_is_valid(oracle, threads):
    """The subtree / queue schedule of the baseline (the reference's
    src/subtree.hpp): every row appears exactly once in each of the three
    schedules, and no row of a thread's subtrees depends on another thread's
    rows or on a queued row, on grids (a chain-like elimination tree: mostly
    stripped levels and queue), on rows with >= 40 entries (parallel sweep
    queue), and on a forest of independent systems, where whole systems
    become subtrees and almost nothing is queued."""
    from oracle import py_oracle as O
    from tests.conftest import random_pattern

    O.baseline_threads(threads)
    cases = []
    a = gen.poisson2d(40)
    cases.append(oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 0)[:2])
    a = gen.convdiff3d(12)
    cases.append(oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)[:2])
    rp, ci = random_pattern(2000, 50, 400, seed=5)
    cases.append((rp, ci))
    rp, ci, va = _forest(24, 12)
    forest = oracle.iluk_symbolic(rp, ci, va, 1)[:2]
    cases.append(forest)
    for frp, fci in cases:
        info = O.baseline_subtree_info(frp, fci)
        assert info["ok"], info
    dense = O.baseline_subtree_info(*cases[2])
    assert dense["parallel_sweep_queue"] == (threads > 1)
    info = O.baseline_subtree_info(*forest)
    n = len(forest[0]) - 1
    assert max(info["queue"].values()) <= 0.1 * n, info
