"""GPU symbolic ILU(k) (IlukSymbolic through the C ABI) against the oracle,
the reference-generated golden fixtures and closed forms: pattern, fill
levels, value placement and nzmap must be bit-exact."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

import par_ilu0_cpp_amd as P
from par_ilu0_cpp_amd import generators as gen
from tests.conftest import dominant_values, random_pattern

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def gpu_symbolic(rp, ci, va, lof):
    return P.iluk_symbolic_levels(P.HostCsr(rp, ci, va), lof)


@pytest.mark.parametrize("path", sorted(GOLDEN.glob("iluk_*.npz")), ids=lambda p: p.stem)
def test_matches_golden(path):
    d = np.load(path)
    for lof in d["lofs"]:
        frp, fci, fva, flev, nz = gpu_symbolic(d["row_ptr"], d["col_idx"], d["value"], int(lof))
        assert np.array_equal(frp, d[f"f{lof}_row_ptr"]), lof
        assert np.array_equal(fci, d[f"f{lof}_col_idx"]), lof
        assert np.array_equal(nz, d[f"f{lof}_nzmap"]), lof
        assert np.array_equal(fva[nz], d[f"f{lof}_level0_values"]), lof
        fill = np.ones(fci.size, dtype=bool)
        fill[nz] = False
        assert np.all(fva[fill] == 0.0) and np.all(flev[nz] == 0) and np.all(flev[fill] > 0)


@pytest.mark.parametrize("lof", [0, 1, 2, 3, 4, 6])
@pytest.mark.parametrize("shape", [(300, 4, 20), (2000, 6, 300), (500, 12, 499)])
def test_matches_oracle_random(oracle, lof, shape):
    n, per, band = shape
    rp, ci = random_pattern(n, per, band, seed=lof + n)
    va = dominant_values(rp, ci, seed=lof)
    frp, fci, fva, flev, nz = gpu_symbolic(rp, ci, va, lof)
    o = oracle.iluk_symbolic(rp, ci, va, lof)
    assert np.array_equal(frp, o[0]) and np.array_equal(fci, o[1])
    assert np.array_equal(flev.astype(np.int32), o[2])
    assert np.array_equal(fva, o[3]) and np.array_equal(nz, o[4])


@pytest.mark.parametrize("lof", [0, 1])
@pytest.mark.parametrize("case", ["random", "convdiff", "poisson", "unsorted", "nodiag", "small", "random_big",
                                  "convdiff_big"])
def test_rows_from_a_match_oracle(oracle, case, lof):
    """With lof <= 1 and sorted rows of A the pass reads pivot rows from A
    (no row waits for another); with a row of A out of column order or
    without its diagonal it waits for the published rows.  Either way:
    bit-exact."""
    if case == "random":
        rp, ci = random_pattern(6000, 6, 300, seed=11)
    elif case == "convdiff":
        a = gen.convdiff3d(18)
        rp, ci = a.row_ptr, a.col_idx
    elif case == "poisson":
        a = gen.poisson2d(80)
        rp, ci = a.row_ptr, a.col_idx
    elif case == "nodiag":  # every third row without its diagonal; lof 1 fills some in
        rp, ci = random_pattern(5000, 5, 200, seed=13)
        keep = np.ones(ci.size, dtype=bool)
        for i in range(0, 5000, 3):
            keep[rp[i] + np.nonzero(ci[rp[i]:rp[i + 1]] == i)[0]] = False
        ci = ci[keep]
        rp = np.concatenate([[0], np.cumsum([keep[rp[i]:rp[i + 1]].sum() for i in range(5000)])]).astype(np.int32)
    elif case == "small":
        rp, ci = random_pattern(300, 12, 299, seed=14)
    elif case == "random_big":  # more rows than waves: each wave strides over several rows
        rp, ci = random_pattern(40000, 6, 500, seed=15)
    elif case == "convdiff_big":
        a = gen.convdiff3d(30)
        rp, ci = a.row_ptr, a.col_idx
    else:
        rp, ci = random_pattern(5000, 5, 200, seed=12)
        ci = ci.copy()
        for i in range(0, 5000, 7):  # every 7th row reversed
            ci[rp[i]:rp[i + 1]] = ci[rp[i]:rp[i + 1]][::-1]
    va = dominant_values(rp, ci, seed=lof)
    o = oracle.iluk_symbolic(rp, ci, va, lof)
    if case == "nodiag":
        # IlukSymbolic alone: the solver's analysis (which also returns the
        # levels) refuses rows without a diagonal
        a = P.HostCsr(rp, ci, va)
        f = P.HostCsr()
        nz = np.zeros(ci.size, dtype=np.int32)
        assert P.IlukSymbolic(a, lof, f.mat, nz) == o[0][-1]
        frp, fci, fva = f.arrays()
        assert np.array_equal(frp, o[0]) and np.array_equal(fci, o[1])
        assert np.array_equal(fva, o[3]) and np.array_equal(nz, o[4])
        return
    frp, fci, fva, flev, nz = gpu_symbolic(rp, ci, va, lof)
    assert np.array_equal(frp, o[0]) and np.array_equal(fci, o[1])
    assert np.array_equal(flev.astype(np.int32), o[2])
    assert np.array_equal(fva, o[3]) and np.array_equal(nz, o[4])


def test_unsorted_rows_follow_reference_value_placement(oracle):
    # A's row given out of order: the reference sorts the pattern (std::map)
    # but hands out A's values in A's entry order to the level-0 slots.
    rp = np.array([0, 2, 5, 7], dtype=np.int32)
    ci = np.array([1, 0, 2, 0, 1, 2, 1], dtype=np.int32)
    va = np.array([10.0, 11.0, 12.0, 13.0, 14.0, 15.0, 16.0])
    frp, fci, fva, flev, nz = gpu_symbolic(rp, ci, va, 1)
    o = oracle.iluk_symbolic(rp, ci, va, 1)
    assert np.array_equal(fci, o[1]) and np.array_equal(fva, o[3]) and np.array_equal(nz, o[4])
    assert list(fci[:2]) == [0, 1] and list(fva[:2]) == [10.0, 11.0]


def test_repeated_column_returns_error():
    rp = np.array([0, 3, 4], dtype=np.int32)
    ci = np.array([1, 0, 1, 1], dtype=np.int32)
    f = P.HostCsr()
    assert P.IlukSymbolic(P.HostCsr(rp, ci, np.ones(4)), 1, f.mat) == -1


def test_tiny_and_diagonal_matrices(oracle):
    for n in (1, 2, 5):
        rp = np.arange(n + 1, dtype=np.int32)
        ci = np.arange(n, dtype=np.int32)
        frp, fci, fva, flev, nz = gpu_symbolic(rp, ci, np.arange(1.0, n + 1), 3)
        assert np.array_equal(fci, ci) and np.array_equal(fva, np.arange(1.0, n + 1))


@pytest.mark.parametrize("s,lof", [(20, 1), (20, 2), (12, 3)])
def test_stencil_matches_oracle(oracle, s, lof):
    a = gen.convdiff3d(s)
    frp, fci, fva, flev, nz = gpu_symbolic(a.row_ptr, a.col_idx, a.value, lof)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, lof)
    assert np.array_equal(fci, o[1]) and np.array_equal(flev.astype(np.int32), o[2])


def test_full_size_128_cube_ilu1_closed_form():
    """The 128^3 ILU(1) pattern at full size: nnz(F) = 13 s^3 - 18 s^2 + 6 s,
    levels of the L and U graphs 6 s - 5 each."""
    import torch

    s = 128
    a = gen.convdiff3d(s)
    ilu = P.IluSolver(0)
    d_rp = torch.from_numpy(a.row_ptr).cuda()
    d_ci = torch.from_numpy(a.col_idx).cuda()
    nnz = ilu.SymbolicDevice(a.n, d_rp.data_ptr(), d_ci.data_ptr(), 1)
    assert nnz == 13 * s**3 - 18 * s**2 + 6 * s == 26_968_832
    assert ilu.Levels() == (6 * s - 5, 6 * s - 5)
    lev = ilu.FillLevels()
    assert lev.size == nnz and int(np.count_nonzero(lev == 0)) == a.nnz and int(lev.max()) == 1


@pytest.mark.parametrize("lof", [1, 2, 3])
def test_kat9_first_found_levels(lof):
    """The hand-derived n = 9 case (tests/test_oracle_symbolic.py) on the GPU:
    the first level found within lof is kept (row 3's (3,6) at 2, not 1),
    so (4,6), (5,6) appear only at lof 3 and (7,6) never; levels and
    pattern equal the independent restatement's."""
    from tests.golden.make_golden import first_found_rows
    from tests.test_oracle_symbolic import kat9

    rp, ci = kat9()
    frp, fci, fva, flev, nz = gpu_symbolic(rp, ci, np.arange(1.0, ci.size + 1), lof)
    want = first_found_rows(rp, ci, lof)
    assert frp[-1] == {1: 20, 2: 20, 3: 22}[lof]
    for i, r in enumerate(want):
        got = {int(c): int(l) for c, l in zip(fci[frp[i]:frp[i + 1]], flev[frp[i]:frp[i + 1]])}
        assert got == r, i
