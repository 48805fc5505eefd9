"""Pins the oracle's symbolic ILU(k) (oracle.c: orc_iluk_symbolic, following
src/iluk.cpp:5-54 of the reference).

Checks, none of which re-types the oracle's own loop:
  * hand-derived known-answer tests of the first-found level rule (n = 7 at
    lof 2; n = 9 at lof 1, 2 and 3, where the first-found and the
    minimum-level rules differ in a level and in two entries), so that a
    pivot-order or level mistake in the oracle fails a check derived by
    hand;
  * golden fixtures written by an independent Python restatement of the
    rule (tests/golden/make_golden.py: a dict per row and a sorted key list
    re-searched after every merge, the way iterating an ordered map visits
    keys inserted during the walk), and a live comparison with that
    restatement on random patterns;
  * closed-form entry counts of stencil patterns;
  * ILU(0) = pattern of A, ILU(1) = textbook ILU(1), monotone growth in k.
The reference's own iluk.cpp cannot be built here (it needs the
reference's csr_matrix.cpp, which needs MKL and SuiteSparse).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from par_ilu0_cpp_amd import generators as gen
from tests.conftest import dominant_values, random_pattern
from tests.golden.make_golden import first_found_rows

GOLDEN = Path(__file__).resolve().parent / "golden"
ROOT = Path(__file__).resolve().parent.parent


def kat7():
    rows = {0: [0, 5], 1: [1], 2: [0, 2], 3: [3, 5], 4: [2, 3, 4], 5: [5], 6: [4, 6]}
    rp = np.zeros(8, dtype=np.int32)
    rp[1:] = np.cumsum([len(rows[i]) for i in range(7)])
    ci = np.array([c for i in range(7) for c in rows[i]], dtype=np.int32)
    return rp, ci


def row_dict(frp, fci, flev, i):
    return {int(c): int(l) for c, l in zip(fci[frp[i]:frp[i + 1]], flev[frp[i]:frp[i + 1]])}


def test_kat_first_found_level(oracle):
    """(4,5) is first reached through pivot 2 at level 0+1+1 = 2; pivot 3
    would give 1 but the first level is kept, so with lof = 2 row 6 gets no
    (6,5) (0+2+1 = 3 > 2): nnz(F) = 15.  The minimum-level rule would give
    (4,5)@1, add (6,5)@2 and 16 entries."""
    rp, ci = kat7()
    frp, fci, flev, fval, nz = oracle.iluk_symbolic(rp, ci, np.arange(1.0, 14.0), 2)
    assert frp[-1] == 15
    assert row_dict(frp, fci, flev, 2) == {0: 0, 2: 0, 5: 1}
    assert row_dict(frp, fci, flev, 4) == {2: 0, 3: 0, 4: 0, 5: 2}
    assert row_dict(frp, fci, flev, 6) == {4: 0, 6: 0}
    # A's values land at the level-0 slots, fill is zero
    assert np.array_equal(fval[nz], np.arange(1.0, 14.0))
    assert np.count_nonzero(fval) == 13


def minimum_level_iluk(rp, ci, lof):
    """Textbook ILU(k) (minimum level over all pivots), for contrast."""
    n = rp.size - 1
    rows = []
    for i in range(n):
        lev = {int(c): 0 for c in ci[rp[i]:rp[i + 1]]}
        changed = True
        while changed:
            changed = False
            for k in sorted(c for c in lev if c < i):
                for j, lkj in rows[k].items():
                    if j <= k:
                        continue
                    L = lev[k] + lkj + 1
                    if L <= lof and L < lev.get(j, lof + 1):
                        lev[j] = L
                        changed = True
        rows.append(lev)
    return sum(len(r) for r in rows)


def test_kat_differs_from_minimum_level_rule():
    rp, ci = kat7()
    assert minimum_level_iluk(rp, ci, 2) == 16


def kat9():
    rows = {0: [0, 6], 1: [0, 1], 2: [2, 6], 3: [1, 2, 3], 4: [3, 4], 5: [3, 5], 6: [6, 8], 7: [4, 7], 8: [8]}
    rp = np.zeros(10, dtype=np.int32)
    rp[1:] = np.cumsum([len(rows[i]) for i in range(9)])
    ci = np.array([c for i in range(9) for c in rows[i]], dtype=np.int32)
    return rp, ci


@pytest.mark.parametrize("lof,nnz,fill", [
    # Row 1 gets (1,6)@1 through pivot 0.  Row 3 meets column 6 first
    # through pivot 1 (0 + 1 + 1 = 2) and then through pivot 2 (0 + 0 + 1 =
    # 1); the first level found within lof is kept.  Rows 4 and 5 reach 6
    # only through pivot 3 at lev(3,6) + 1; row 7 only through pivot 4 at
    # lev(4,6) + 1.
    (1, 20, {1: {6: 1}, 3: {6: 1}}),                    # via pivot 1: 2 > 1, so pivot 2's level 1
    (2, 20, {1: {6: 1}, 3: {6: 2}}),                    # (4,6), (5,6) would be 3 > 2
    (3, 22, {1: {6: 1}, 3: {6: 2}, 4: {6: 3}, 5: {6: 3}}),  # (7,6) would be 4 > 3
])
def test_kat9_first_found_levels(oracle, lof, nnz, fill):
    """Hand-derived first-found patterns and levels (every entry of A at
    level 0, plus `fill`); the minimum-level rule gives (3,6)@1 and two more
    entries at lof 2 ((4,6)@2, (5,6)@2), and (7,6)@3 on top at lof 3."""
    rp, ci = kat9()
    frp, fci, flev, fval, nz = oracle.iluk_symbolic(rp, ci, np.arange(1.0, ci.size + 1), lof)
    assert frp[-1] == nnz
    for i in range(9):
        want = {int(c): 0 for c in ci[rp[i]:rp[i + 1]]}
        want.update(fill.get(i, {}))
        assert row_dict(frp, fci, flev, i) == want, i
    assert np.array_equal(fval[nz], np.arange(1.0, ci.size + 1))
    assert minimum_level_iluk(rp, ci, lof) == {1: 20, 2: 22, 3: 23}[lof]


@pytest.mark.parametrize("lof", [0, 1, 2, 3])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_matches_ordered_map_restatement(oracle, lof, seed):
    rp, ci = random_pattern(120, 3 + seed, 15, seed=seed)
    frp, fci, flev, _, _ = oracle.iluk_symbolic(rp, ci, dominant_values(rp, ci, seed), lof)
    rows = first_found_rows(rp, ci, lof)
    for i, r in enumerate(rows):
        assert row_dict(frp, fci, flev, i) == r, i


@pytest.mark.parametrize("path", sorted(GOLDEN.glob("iluk_*.npz")), ids=lambda p: p.stem)
def test_matches_golden(oracle, path):
    d = np.load(path)
    for lof in d["lofs"]:
        frp, fci, flev, fval, nz = oracle.iluk_symbolic(d["row_ptr"], d["col_idx"], d["value"], int(lof))
        assert np.array_equal(frp, d[f"f{lof}_row_ptr"]), lof
        assert np.array_equal(fci, d[f"f{lof}_col_idx"]), lof
        assert np.array_equal(nz, d[f"f{lof}_nzmap"]), lof
        assert np.array_equal(fval[nz], d[f"f{lof}_level0_values"]), lof
        assert np.all(flev[nz] == 0) and np.count_nonzero(flev == 0) == nz.size


def test_ilu0_is_pattern_of_a(oracle):
    a = gen.convdiff3d(8)
    frp, fci, flev, fval, nz = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 0)
    assert np.array_equal(frp, a.row_ptr) and np.array_equal(fci, a.col_idx)
    assert np.array_equal(fval, a.value) and np.array_equal(nz, np.arange(a.nnz))


@pytest.mark.parametrize("s", [8, 16, 24])
def test_stencil_closed_forms(oracle, s):
    a3 = gen.convdiff3d(s)
    f = oracle.iluk_symbolic(a3.row_ptr, a3.col_idx, None, 1)
    assert f[0][-1] == 13 * s**3 - 18 * s**2 + 6 * s
    assert a3.nnz == 7 * s**3 - 6 * s**2


def test_poisson64_ilu1_count(oracle):
    a = gen.poisson2d(64)
    assert a.nnz == 20224
    f = oracle.iluk_symbolic(a.row_ptr, a.col_idx, None, 1)
    assert f[0][-1] == 28162


@pytest.mark.parametrize("seed", [3, 4])
def test_ilu1_equals_textbook_and_grows(oracle, seed):
    rp, ci = random_pattern(90, 4, 12, seed=seed)
    counts = [oracle.iluk_symbolic(rp, ci, None, k)[0][-1] for k in range(5)]
    assert counts[1] == minimum_level_iluk(rp, ci, 1)
    assert all(x <= y for x, y in zip(counts, counts[1:]))


def test_repeated_column_is_rejected(oracle):
    rp = np.array([0, 2, 4], dtype=np.int32)
    ci = np.array([0, 0, 0, 1], dtype=np.int32)
    assert oracle.iluk_symbolic(rp, ci, np.ones(4), 1) is None
