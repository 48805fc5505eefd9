"""Pins the oracle's numeric ILU, sweeps and SpMV (oracle.c:
orc_ilu_factorize / orc_substitute / orc_spmv, following
src/ilu_solver.cpp:553-589,775-796 and the SpMV of src/gmres.cpp).

Checks come from the mathematics, not from the oracle's loops:
  * (L U)_ij = A_ij on every position of pattern(F) (the defining property
    of an incomplete LU on that pattern);
  * the tridiagonal closed form u_1 = 2, u_i = 2 - 1/u_{i-1};
  * with full fill (lof >= bandwidth) the factors equal scipy's dense LU of
    a column-diagonally-dominant matrix (no row exchanges happen);
  * the sweeps equal scipy's dense triangular solves;
  * SpMV equals exactly rounded math.fsum row sums to a few ulps.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.linalg as sla

from par_ilu0_cpp_amd import generators as gen
from tests.conftest import dominant_values, random_pattern


def dense(rp, ci, va, n):
    m = np.zeros((n, n))
    for i in range(n):
        m[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
    return m


def split_lu(rp, ci, fac, n):
    m = dense(rp, ci, fac, n)
    L = np.tril(m, -1) + np.eye(n)
    U = np.triu(m)
    return L, U


@pytest.mark.parametrize("lof", [0, 1, 2])
def test_lu_reproduces_a_on_pattern(oracle, lof):
    rp, ci = random_pattern(80, 4, 10, seed=lof)
    va = dominant_values(rp, ci, seed=lof)
    frp, fci, _, fval, _ = oracle.iluk_symbolic(rp, ci, va, lof)
    fac, dg, z = oracle.ilu_factorize(frp, fci, fval)
    assert z == 0
    n = rp.size - 1
    L, U = split_lu(frp, fci, fac, n)
    A = dense(rp, ci, va, n)
    LU = L @ U
    mask = dense(frp, fci, np.ones(fci.size), n) > 0
    assert np.max(np.abs(LU[mask] - A[mask])) <= 1e-13 * np.max(np.abs(A))


def test_tridiagonal_closed_form(oracle):
    n = 50
    rp = np.zeros(n + 1, dtype=np.int32)
    cols, vals = [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                cols.append(j)
                vals.append(v)
        rp[i + 1] = len(cols)
    fac, dg, _ = oracle.ilu_factorize(rp, np.array(cols, dtype=np.int32), np.array(vals))
    u = [2.0]
    for _ in range(n - 1):
        u.append(2.0 - 1.0 / u[-1])
    assert np.allclose(fac[dg], u, rtol=1e-15, atol=0)
    # multipliers l_{i,i-1} = -1/u_{i-1}
    lower = np.array([fac[rp[i]] for i in range(1, n)])
    assert np.allclose(lower, [-1.0 / x for x in u[:-1]], rtol=1e-15)


def test_full_fill_equals_dense_lu(oracle):
    n, band = 40, 3
    rp, ci = random_pattern(n, 4, band, seed=9)
    rng = np.random.default_rng(9)
    va = rng.uniform(-1, 1, ci.size)
    A = dense(rp, ci, va, n)
    A += np.diag(np.abs(A).sum(axis=0) + 1.0)  # column dominance: no pivoting
    for i in range(n):
        s, e = rp[i], rp[i + 1]
        va[s:e] = A[i, ci[s:e]]
    frp, fci, _, fval, _ = oracle.iluk_symbolic(rp, ci, va, 4 * band)
    fac, _, _ = oracle.ilu_factorize(frp, fci, fval)
    L, U = split_lu(frp, fci, fac, n)
    P, Ls, Us = sla.lu(A)
    assert np.array_equal(P, np.eye(n))
    assert np.allclose(L, Ls, rtol=1e-12, atol=1e-14)
    assert np.allclose(U, Us, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("lof", [0, 2])
def test_sweeps_equal_dense_triangular_solves(oracle, lof):
    a = gen.convdiff3d(5)
    frp, fci, _, fval, _ = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, lof)
    fac, dg, _ = oracle.ilu_factorize(frp, fci, fval)
    L, U = split_lu(frp, fci, fac, a.n)
    b = np.random.default_rng(1).uniform(-1, 1, a.n)
    x = oracle.substitute(frp, fci, dg, fac, b)
    ref = sla.solve_triangular(U, sla.solve_triangular(L, b, lower=True, unit_diagonal=True))
    assert np.allclose(x, ref, rtol=1e-12, atol=1e-14)
    # in place (b aliases x in the reference) gives the same
    assert np.array_equal(oracle.substitute(frp, fci, dg, fac, b), x)


def test_spmv_matches_exact_sums(oracle):
    rp, ci = random_pattern(200, 8, 50, seed=5)
    va = dominant_values(rp, ci, 5)
    x = np.random.default_rng(5).uniform(-1, 1, 200)
    y0 = np.random.default_rng(6).uniform(-1, 1, 200)
    y = oracle.spmv(rp, ci, va, x, alpha=-1.0, beta=1.0, y=y0)
    for i in range(200):
        ex = y0[i] - math.fsum(va[p] * x[ci[p]] for p in range(rp[i], rp[i + 1]))
        assert abs(y[i] - ex) <= 1e-13 * max(1.0, abs(ex))


def test_zero_pivot_is_counted(oracle):
    rp = np.array([0, 2, 4], dtype=np.int32)
    ci = np.array([0, 1, 0, 1], dtype=np.int32)
    _, _, z = oracle.ilu_factorize(rp, ci, np.array([0.0, 1.0, 1.0, 1.0]))
    assert z == 1


@pytest.mark.parametrize("lof", [0, 1])
def test_compensated_factorization_row_sums(oracle, lof):
    """Modified ILU with omega = 1 preserves row sums: (L U) 1 = A 1; with
    omega = 0 it is plain ILU, bit for bit."""
    a = gen.convdiff3d(7)
    frp, fci, _, fval, _ = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, lof)
    n = a.n
    fac1, dg = oracle.ilu_factorize_comp(frp, fci, fval, np.ones(n, dtype=np.float32))
    L, U = split_lu(frp, fci, fac1, n)
    A = dense(a.row_ptr, a.col_idx, a.value, n)
    assert np.allclose((L @ U).sum(axis=1), A.sum(axis=1), rtol=0, atol=1e-13)
    fac0, _ = oracle.ilu_factorize_comp(frp, fci, fval, np.zeros(n, dtype=np.float32))
    plain, _, _ = oracle.ilu_factorize(frp, fci, fval)
    assert np.array_equal(fac0, plain)
    if lof == 0:
        # the compensation actually changed something (ILU(0) drops fill here)
        assert not np.array_equal(fac1, plain)


def _fma(a, b, c):
    """Correctly rounded a*b + c (exact rational arithmetic, one rounding)."""
    from fractions import Fraction

    return float(Fraction(a) * Fraction(b) + Fraction(c))


def _lane_sweep_reference(rp, ci, dg, fac, b, g):
    """Pure-Python restatement of the g-lane sweep order: lane t takes
    entries t, t+g, ... with one rounding per fma, lane 0 starts from the
    right-hand side, then lane t += lane t+off for off = g/2, ..., 1."""
    n = rp.size - 1
    x = [float(v) for v in b]

    def row(init, p0, p1):
        part = [0.0] * g
        part[0] = init
        for p in range(p0, p1):
            t = (p - p0) % g
            part[t] = _fma(-x[ci[p]], float(fac[p]), part[t])
        off = g // 2
        while off:
            for t in range(off):
                part[t] = part[t] + part[t + off]
            off //= 2
        return part[0]

    for i in range(n):
        x[i] = row(x[i], rp[i], dg[i])
    for i in range(n - 1, -1, -1):
        x[i] = row(x[i], dg[i] + 1, rp[i + 1]) / float(fac[dg[i]])
    return np.array(x)


@pytest.mark.parametrize("g", [1, 2, 4, 8, 16, 32, 64])
def test_lane_order_sweep_pinned(oracle, g):
    """orc_substitute_lanes equals an exact-arithmetic restatement of the
    lane order for every g, and orc_substitute itself at g = 1."""
    rp, ci = random_pattern(60, 30, 59, seed=g)
    va = dominant_values(rp, ci, g)
    frp, fci, _, fval, _ = oracle.iluk_symbolic(rp, ci, va, 2)
    fac, dg, _ = oracle.ilu_factorize(frp, fci, fval)
    b = np.random.default_rng(g).uniform(-1, 1, rp.size - 1)
    x = oracle.substitute_lanes(frp, fci, dg, fac, b, g)
    assert np.array_equal(x, _lane_sweep_reference(frp, fci, dg, fac, b, g))
    if g == 1:
        assert np.array_equal(x, oracle.substitute(frp, fci, dg, fac, b))
    else:
        # a different rounding order: close to, and here not identical with, the serial sweep
        xs = oracle.substitute(frp, fci, dg, fac, b)
        assert np.allclose(x, xs, rtol=1e-12, atol=1e-14)
