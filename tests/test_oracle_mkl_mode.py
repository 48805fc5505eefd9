"""The oracle's restatement of the reference's MKL-mode path (USE_MKL_ILU /
USE_MKL_SV: dcsrilu0 and two mkl_sparse_d_trsv calls,
src/ilu_solver.cpp:248-259,631-639,753-773) pinned without MKL (closed and
absent here): the pivot policy by hand-worked closed forms, the triangular
solves by their composition into the already-pinned substitution and by
scipy's triangular solves."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.linalg import spsolve_triangular

from par_ilu0_cpp_amd import generators as gen
from tests.conftest import dominant_values, random_pattern


def csr(dense):
    m = sp.csr_matrix(np.asarray(dense, dtype=np.float64))
    m.sort_indices()
    return m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data.copy()


def csr_pattern(dense_pattern, values):
    """CSR over an explicit pattern (zeros kept as entries)."""
    rows = [np.nonzero(r)[0] for r in np.asarray(dense_pattern)]
    rp = np.zeros(len(rows) + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.int32)
    va = np.array([values[i][j] for i, r in enumerate(rows) for j in r], dtype=np.float64)
    return rp, ci, va


def test_policy_zero_is_the_plain_factorization(oracle):
    rp, ci = random_pattern(400, 6, 40, seed=3)
    va = dominant_values(rp, ci, seed=3)
    f0, dg, st = oracle.ilu0_guarded(rp, ci, va, 0)
    f1, _, _ = oracle.ilu_factorize(rp, ci, va)
    assert st == 0 and np.array_equal(f0, f1)
    f2, _, st2 = oracle.ilu0_guarded(rp, ci, va, 1)  # no small pivot: nothing changes
    assert st2 == 0 and np.array_equal(f2, f1)


def test_stop_codes_closed_form(oracle):
    # u_22 = 1 - 1*1 = 0 exactly
    rp, ci, va = csr([[1, 1], [1, 1]])
    assert oracle.ilu0_guarded(rp, ci, va, 1)[2] == -102
    # u_22 = 1.0005 - 1 = 5e-4 < 1e-3
    rp, ci, va = csr([[1, 1], [1, 1.0005]])
    assert oracle.ilu0_guarded(rp, ci, va, 1, tiny=1e-3)[2] == -103
    assert oracle.ilu0_guarded(rp, ci, va, 1, tiny=1e-4)[2] == 0
    assert oracle.ilu0_guarded(rp, ci, va, 0, tiny=1e-3)[2] == 0
    # the first failing row decides: row 1 small, row 3 zero -> -103; swapped -> -102
    pat = np.eye(4, dtype=bool)
    for small, zero, want in ((1, 3, -103), (3, 1, -102)):
        vals = np.eye(4)
        vals[small, small] = 1e-20
        vals[zero, zero] = 0.0
        rp, ci, va = csr_pattern(pat, vals)
        assert oracle.ilu0_guarded(rp, ci, va, 1)[2] == want
    # a negative pivot of large magnitude is not small
    rp, ci, va = csr([[-3.0]])
    assert oracle.ilu0_guarded(rp, ci, va, 1, tiny=1.0)[2] == 0


def test_replaced_pivot_closed_form(oracle):
    # A = [[1,1,0],[1,1,1],[0,1,2]]: u11 = 1, l21 = 1, u22 = 1 - 1 = 0 -> r,
    # u23 = 1, l32 = 1 / r, u33 = 2 - l32 * 1
    r = 1e-10
    rp, ci, va = csr([[1, 1, 0], [1, 1, 1], [0, 1, 2]])
    f, dg, st = oracle.ilu0_guarded(rp, ci, va, 2, 1e-16, r)
    assert st == 0
    assert np.array_equal(f, np.array([1, 1, 1, r, 1, 1 / r, 2 - 1 / r]))
    # without the policy the zero pivot is divided by
    f0, _, z = oracle.ilu_factorize(rp, ci, va)
    assert z == 1 and np.isinf(f0[5])


def test_input_checks(oracle):
    rp, ci, va = csr_pattern([[1, 1], [1, 0]], [[2.0, 1.0], [1.0, 0.0]])  # row 1 has no diagonal
    assert oracle.ilu0_guarded(rp, ci, va, 1)[2] == -101
    rp = np.array([0, 2, 4], np.int32)
    ci = np.array([0, 1, 1, 0], np.int32)  # row 1 descending
    va = np.array([2.0, 1.0, 2.0, 1.0])
    assert oracle.ilu0_guarded(rp, ci, va, 1)[2] == -106


@pytest.mark.parametrize("g", [1, 4, 8])
def test_trsv_composes_to_substitution(oracle, g):
    a = gen.convdiff3d(9)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    fac, dg, _ = oracle.ilu_factorize(o[0], o[1], o[3])
    b = np.random.default_rng(g).uniform(-1, 1, a.n)
    y = oracle.trsv_lanes(o[0], o[1], dg, fac, b, False, 1.0, g)
    x = oracle.trsv_lanes(o[0], o[1], dg, fac, y, True, 1.0, g)
    assert np.array_equal(x, oracle.substitute_lanes(o[0], o[1], dg, fac, b, g))


def test_trsv_solves_each_triangle(oracle):
    rp, ci = random_pattern(300, 5, 30, seed=5)
    va = dominant_values(rp, ci, seed=5)
    fac, dg, _ = oracle.ilu_factorize(rp, ci, va)
    n = rp.size - 1
    m = sp.csr_matrix((fac, ci, rp), shape=(n, n))
    low = (sp.tril(m, -1) + sp.eye(n)).tocsr()
    up = sp.triu(m).tocsr()
    x = np.random.default_rng(1).uniform(-1, 1, n)
    for upper, t in ((False, low), (True, up)):
        for alpha in (1.0, -0.75):
            y = oracle.trsv_lanes(rp, ci, dg, fac, x, upper, alpha, 4)
            ys = spsolve_triangular(t, x, lower=not upper, unit_diagonal=not upper) * alpha
            assert np.max(np.abs(y - ys)) <= 1e-12 * np.max(np.abs(ys))
            assert np.max(np.abs(t @ y - alpha * x)) <= 1e-12 * np.max(np.abs(x))
        y1 = oracle.trsv_lanes(rp, ci, dg, fac, x, upper, 1.0, 4)
        assert np.array_equal(oracle.trsv_lanes(rp, ci, dg, fac, x, upper, 2.5, 4), y1 * 2.5)
