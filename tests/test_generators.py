"""Synthetic workload generators: sizes, closed-form entry counts and the
structural properties each workload is defined by."""
from __future__ import annotations

import numpy as np
import pytest

from par_ilu0_cpp_amd import generators as gen


def offdiag_sums(a):
    n = a.n
    diag = np.zeros(n)
    off = np.zeros(n)
    rows = np.repeat(np.arange(n), np.diff(a.row_ptr))
    on = a.col_idx == rows
    diag[rows[on]] = a.value[on]
    np.add.at(off, rows[~on], np.abs(a.value[~on]))
    return diag, off


def test_poisson64_sizes_and_values():
    a = gen.poisson2d(64)
    assert a.n == 4096 and a.nnz == 20224
    sp = a.to_scipy()
    assert (sp != sp.T).nnz == 0
    d, off = offdiag_sums(a)
    assert np.all(d == 4.0) and np.all(off <= 4.0)
    assert np.linalg.norm(a.matvec(np.ones(a.n))) == pytest.approx(np.sqrt(264.0))


@pytest.mark.parametrize("s", [4, 9, 16])
def test_convdiff_counts_and_coefficients(s):
    a = gen.convdiff3d(s)
    assert a.n == s**3 and a.nnz == 7 * s**3 - 6 * s**2
    h, eps = 1.0 / (s + 1), 1.0 / 50
    d, off = offdiag_sums(a)
    assert np.allclose(d, 6 * eps + h * 1.75)
    # interior row: west/south/down carry the upwind term
    i = (s // 2) + s * (s // 2) + s * s * (s // 2)
    row = dict(zip(a.col_idx[a.row_ptr[i]:a.row_ptr[i + 1]], a.value[a.row_ptr[i]:a.row_ptr[i + 1]]))
    assert row[i - 1] == pytest.approx(-eps - h * 1.0)
    assert row[i - s] == pytest.approx(-eps - h * 0.5)
    assert row[i - s * s] == pytest.approx(-eps - h * 0.25)
    assert row[i + 1] == row[i + s] == row[i + s * s] == pytest.approx(-eps)
    assert np.all(d >= off - 1e-15)  # weakly diagonally dominant M-matrix


def test_convdiff_128_sizes():
    a = gen.convdiff3d(128)
    assert a.n == 2_097_152 and a.nnz == 14_581_760


def test_random_banded_structure():
    a = gen.random_banded(3000, 12, 500, seed=1)
    rows = np.repeat(np.arange(a.n), np.diff(a.row_ptr))
    assert np.all(np.diff(a.row_ptr) == 13)
    off = a.col_idx != rows
    assert np.all(np.abs(a.col_idx[off] - rows[off]) <= 500)
    assert np.all(np.abs(a.value[off]) <= 1.0)
    d, s = offdiag_sums(a)
    assert np.allclose(d, 1.05 * s)
    for i in range(a.n):  # sorted, distinct
        c = a.col_idx[a.row_ptr[i]:a.row_ptr[i + 1]]
        assert np.all(np.diff(c) > 0)
    b = gen.random_banded(3000, 12, 500, seed=1)
    assert np.array_equal(a.col_idx, b.col_idx) and np.array_equal(a.value, b.value)


def test_diffusion2d_spd_and_harmonic_faces():
    kappa = np.random.default_rng(0).lognormal(0, 1, size=(7, 5))
    a = gen.diffusion2d(kappa)
    sp = a.to_scipy().toarray()
    assert np.allclose(sp, sp.T)
    assert np.all(np.linalg.eigvalsh(sp) > 0)
    i = 2 + 7 * 3
    hm = 2 * kappa[2, 3] * kappa[3, 3] / (kappa[2, 3] + kappa[3, 3])
    assert sp[i, i + 1] == pytest.approx(-hm)


def test_batched_sizes():
    mats = gen.batched_diffusion2d(20, seed=3)
    sides = [int(round(np.sqrt(m.n))) for m in mats]
    assert len(mats) == 20 and all(45 <= s <= 256 for s in sides)
    assert all(m.n == s * s and m.nnz == 5 * s * s - 4 * s for m, s in zip(mats, sides))


def test_grf_unit_variance_and_correlation():
    g = gen.gaussian_random_field((48, 48, 48), corr_len=8.0, seed=0)
    assert abs(g.mean()) < 1e-12 and g.std() == pytest.approx(1.0)
    near = np.mean(g[:-1] * g[1:])
    far = np.mean(g[:-24] * g[24:])
    assert near > 0.9 and abs(far) < near / 4
    a = gen.grf_diffusion3d(8, seed=1)
    assert a.n == 512 and a.nnz == 7 * 512 - 6 * 64
