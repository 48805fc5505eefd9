"""GPU numeric ILU and triangular sweeps against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

import par_ilu0_cpp_amd as P
from par_ilu0_cpp_amd import generators as gen
from tests.conftest import dominant_values, random_pattern

pytestmark = pytest.mark.gpu

LANES = (1, 2, 4, 8, 16, 32, 64)


def factored(frp, fci, fva, bits=64):
    ilu = P.IluSolver(0)
    ilu.SetFactorPrecision(bits)
    ilu.SetMatrix(frp, fci, fva)
    assert ilu.SetupMatrix() == 0
    assert ilu.Factorize() == 0
    return ilu


def host_factors(ilu, nnz):
    return P._view(ilu.GetMatrix().value, nnz, np.float64).copy()


def relmax(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300)))


@pytest.mark.parametrize("case", ["conv12_ilu1", "conv10_ilu2", "rand_ilu0", "rand_ilu3", "poisson_ilu0"])
def test_factors_match_oracle(oracle, case):
    if case.startswith("conv"):
        s, lof = (12, 1) if case == "conv12_ilu1" else (10, 2)
        a = gen.convdiff3d(s)
        rp, ci, va = a.row_ptr, a.col_idx, a.value
    elif case == "poisson_ilu0":
        a = gen.poisson2d(64)
        rp, ci, va, lof = a.row_ptr, a.col_idx, a.value, 0
    else:
        lof = 0 if case == "rand_ilu0" else 3
        rp, ci = random_pattern(1500, 6, 80, seed=lof)
        va = dominant_values(rp, ci, seed=lof)
    frp, fci, _, _, fval = None, None, None, None, None
    o = oracle.iluk_symbolic(rp, ci, va, lof)
    frp, fci, fval = o[0], o[1], o[3]
    ilu = factored(frp, fci, fval)
    fac = host_factors(ilu, fci.size)
    ofac, dg, z = oracle.ilu_factorize(frp, fci, fval)
    assert z == 0
    # same elimination order and the same fma per update: expect bit equality
    assert relmax(fac, ofac) <= 1e-12
    assert np.array_equal(fac, ofac)
    assert np.array_equal(ilu.DeviceFactors(), fac)


def test_long_rows_use_wide_staging(oracle):
    # rows of ~300 entries exercise the 64-lane / large-LDS variant
    rp, ci = random_pattern(1200, 150, 400, seed=7)
    va = dominant_values(rp, ci, seed=7)
    ilu = factored(rp, ci, va)
    ofac, dg, _ = oracle.ilu_factorize(rp, ci, va)
    assert relmax(host_factors(ilu, ci.size), ofac) <= 1e-12


@pytest.mark.parametrize("bits", [64, 32])
def test_sweeps_all_lane_shapes(oracle, bits):
    a = gen.convdiff3d(14)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    frp, fci, fval = o[0], o[1], o[3]
    ilu = factored(frp, fci, fval, bits)
    ofac, dg, _ = oracle.ilu_factorize(frp, fci, fval)
    rng = np.random.default_rng(bits)
    for lanes_l in LANES:
        lanes_u = LANES[(LANES.index(lanes_l) + 3) % len(LANES)]
        ilu.SetSweepLanes(lanes_l, lanes_u)
        for rep in range(2):  # the second apply checks the re-armed buffers
            b = rng.uniform(-1, 1, a.n)
            x = ilu.Substitute(b)
            if bits == 64:
                xo = oracle.substitute(frp, fci, dg, ofac, b)
                assert relmax(x, xo) <= 1e-12 or np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))
                if lanes_l == 1 and lanes_u == 1:
                    assert np.array_equal(x, xo)
            else:
                xo = oracle.substitute(frp, fci, dg, ofac.astype(np.float32).astype(np.float64), b)
                assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("mode,blk", [(0, 0), (1, 512), (2, 512), (2, 4096), (3, 512), (3, 4096), (4, 0)])
def test_sweeps_bit_exact_in_lane_order(oracle, bits, mode, blk):
    """Every sweep kernel, at every lanes-per-row shape, reproduces the
    oracle's lane-order sweep bit for bit (FP32 factors: the oracle gets the
    same float-rounded values)."""
    a = gen.convdiff3d(14)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    frp, fci, fval = o[0], o[1], o[3]
    ilu = factored(frp, fci, fval, bits)
    ofac, dg, _ = oracle.ilu_factorize(frp, fci, fval)
    if bits == 32:
        ofac = ofac.astype(np.float32).astype(np.float64)
    if mode:
        ilu.SetSweepMode(mode, mode, blk)
    rng = np.random.default_rng(mode * 10 + bits)
    for g in (LANES if mode in (0, 1, 4) else (1, 2, 4, 8, 16)):
        ilu.SetSweepLanes(g, g)
        for _ in range(2):
            b = rng.uniform(-1, 1, a.n)
            assert np.array_equal(ilu.Substitute(b), oracle.substitute_lanes(frp, fci, dg, ofac, b, g)), g


def test_sweep_lane_one_bit_exact(oracle):
    a = gen.poisson2d(50)
    ilu = factored(a.row_ptr, a.col_idx, a.value)
    ofac, dg, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, a.value)
    ilu.SetSweepLanes(1, 1)
    b = np.random.default_rng(3).uniform(-1, 1, a.n)
    assert np.array_equal(ilu.Substitute(b), oracle.substitute(a.row_ptr, a.col_idx, dg, ofac, b))


def test_fp32_factor_error_bound(oracle):
    a = gen.convdiff3d(16)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    ilu64 = factored(o[0], o[1], o[3], 64)
    ilu32 = factored(o[0], o[1], o[3], 32)
    b = np.random.default_rng(1).uniform(-1, 1, a.n)
    x64, x32 = ilu64.Substitute(b), ilu32.Substitute(b)
    assert np.max(np.abs(x32 - x64)) <= 1e-4 * np.max(np.abs(x64))
    assert not np.array_equal(x32, x64)


def test_missing_diagonal_is_minus_two():
    rp = np.array([0, 2, 3, 5], dtype=np.int32)
    ci = np.array([0, 2, 0, 1, 2], dtype=np.int32)  # row 1 has no (1,1)
    ilu = P.IluSolver(0)
    ilu.SetMatrix(rp, ci, np.ones(5))
    assert ilu.SetupMatrix() == -2


def test_unsorted_pattern_is_rejected():
    rp = np.array([0, 2, 4], dtype=np.int32)
    ci = np.array([1, 0, 0, 1], dtype=np.int32)
    ilu = P.IluSolver(0)
    ilu.SetMatrix(rp, ci, np.ones(4))
    assert ilu.SetupMatrix() == -1


def test_column_out_of_range_is_rejected():
    rp = np.array([0, 1, 3], dtype=np.int32)
    ci = np.array([0, 1, 7], dtype=np.int32)
    ilu = P.IluSolver(0)
    ilu.SetMatrix(rp, ci, np.ones(3))
    assert ilu.SetupMatrix() == -1


def test_zero_pivot_reports_failure():
    rp = np.array([0, 2, 4], dtype=np.int32)
    ci = np.array([0, 1, 0, 1], dtype=np.int32)
    ilu = P.IluSolver(0)
    ilu.SetMatrix(rp, ci, np.array([0.0, 1.0, 1.0, 1.0]))
    assert ilu.SetupMatrix() == 0
    assert ilu.Factorize() == -1
    # the reference's factors: row 1 divides by the zero pivot
    assert np.isinf(P._view(ilu.GetMatrix().value, 4, np.float64)[2])


def test_refactorize_new_values(oracle):
    a = gen.poisson2d(30)
    ilu = factored(a.row_ptr, a.col_idx, a.value)
    va2 = a.value * np.linspace(1.0, 2.0, a.nnz)
    P.fill_csr(ilu.GetMatrix(), a.row_ptr, a.col_idx, va2)
    assert ilu.Factorize() == 0
    ofac, _, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, va2)
    assert np.array_equal(host_factors(ilu, a.nnz), ofac)


def test_autotune_keeps_results(oracle):
    a = gen.convdiff3d(18)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    ilu = factored(o[0], o[1], o[3])
    b = np.random.default_rng(9).uniform(-1, 1, a.n)
    ilu.SetSweepLanes(8, 8)
    x0 = ilu.Substitute(b)
    cfg, table = ilu.Autotune(reps=2)
    assert table.shape[0] >= 7 and np.all(table[:, 6:] > 0)
    assert cfg["lanes_l"] in LANES and cfg["lanes_u"] in LANES
    x1 = ilu.Substitute(b)
    assert np.max(np.abs(x1 - x0)) <= 1e-12 * np.max(np.abs(x0))


def test_autotune_with_grid_times_tiles(oracle):
    """With a declared grid the autotuner also times tiles spanning the x
    extent and several planes; whatever it keeps reproduces the sweep."""
    s = 20
    a = gen.convdiff3d(s)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    ilu = factored(o[0], o[1], o[3])
    ilu.SetGridShape(s, s, s)
    b = np.random.default_rng(11).uniform(-1, 1, a.n)
    x0 = ilu.Substitute(b)
    cfg, table = ilu.Autotune(reps=1)
    tiles = table[table[:, 2] > 0]
    assert {(s, 4, 8), (s, 8, 6), (15, 4, 8)} <= {tuple(int(v) for v in r[2:5]) for r in tiles}
    assert np.all(table[:, 6:] > 0)
    ofac, dg, _ = oracle.ilu_factorize(o[0], o[1], o[3])
    x1 = ilu.Substitute(b)
    assert np.max(np.abs(x1 - x0)) <= 1e-12 * np.max(np.abs(x0))
    if cfg["lanes_l"] == cfg["lanes_u"]:
        assert np.array_equal(x1, oracle.substitute_lanes(o[0], o[1], dg, ofac, b, cfg["lanes_l"]))


@pytest.mark.parametrize("mode", [1, 2, 3])
@pytest.mark.parametrize("blk", [64, 1000, 4096, 100000])
@pytest.mark.parametrize("lanes", [1, 4, 8, 16])
def test_block_sweep_bit_identical_to_level_sweep(oracle, mode, blk, lanes):
    """The block-local sweeps (CSR or packed factors) sum each row in the
    same order as the level-scheduled sweep with the same lanes per row:
    identical bits, in either direction, applied repeatedly."""
    a = gen.convdiff3d(15)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    ilu = factored(o[0], o[1], o[3])
    b = np.random.default_rng(lanes).uniform(-1, 1, a.n)
    ilu.SetSweepLanes(lanes, lanes)
    x_level = ilu.Substitute(b)
    blk = min(blk, 16384)
    ilu.SetSweepMode(mode, mode, blk)
    for _ in range(2):
        assert np.array_equal(ilu.Substitute(b), x_level)
    ilu.SetSweepMode(mode, 0, blk)
    assert np.array_equal(ilu.Substitute(b), x_level)
    ilu.SetSweepMode(0, mode, blk)
    assert np.array_equal(ilu.Substitute(b), x_level)
    if lanes == 1:
        ofac, dg, _ = oracle.ilu_factorize(o[0], o[1], o[3])
        assert np.array_equal(x_level, oracle.substitute(o[0], o[1], dg, ofac, b))


def test_packed_sweep_follows_refactorization(oracle):
    a = gen.convdiff3d(12)
    ilu = factored(a.row_ptr, a.col_idx, a.value)
    ilu.SetSweepMode(1, 2, 512)
    va2 = a.value * np.linspace(1.0, 1.5, a.nnz)
    P.fill_csr(ilu.GetMatrix(), a.row_ptr, a.col_idx, va2)
    assert ilu.Factorize() == 0
    b = np.random.default_rng(4).uniform(-1, 1, a.n)
    ofac, dg, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, va2)
    xo = oracle.substitute(a.row_ptr, a.col_idx, dg, ofac, b)
    assert np.max(np.abs(ilu.Substitute(b) - xo)) <= 1e-12 * np.max(np.abs(xo))


def test_block_sweep_random_pattern(oracle):
    rp, ci = random_pattern(5000, 8, 700, seed=21)
    va = dominant_values(rp, ci, seed=21)
    ilu = factored(rp, ci, va)
    ofac, dg, _ = oracle.ilu_factorize(rp, ci, va)
    b = np.random.default_rng(2).uniform(-1, 1, 5000)
    xo = oracle.substitute(rp, ci, dg, ofac, b)
    ilu.SetSweepLanes(4, 4)
    for mode in (1, 2, 3):
        ilu.SetSweepMode(mode, mode, 512)
        x = ilu.Substitute(b)
        assert np.max(np.abs(x - xo)) <= 1e-12 * np.max(np.abs(xo))


def test_packed_sweep_long_rows(oracle):
    """Rows longer than the packed sweep's register slots (at most 4 per
    lane) finish in the kernel's tail loop: rows of ~150 entries per
    direction, bit-identical to the oracle's lane-order sweep and to the
    level sweep with the same lanes, in both packed forms."""
    rp, ci = random_pattern(1200, 150, 400, seed=7)
    va = dominant_values(rp, ci, seed=7)
    ilu = factored(rp, ci, va)
    ofac, dg, _ = oracle.ilu_factorize(rp, ci, va)
    b = np.random.default_rng(3).uniform(-1, 1, rp.size - 1)
    for g in (1, 4, 16):
        ilu.SetSweepLanes(g, g)
        ilu.SetSweepMode(0, 0, 0)
        x_level = ilu.Substitute(b)
        assert np.array_equal(x_level, oracle.substitute_lanes(rp, ci, dg, ofac, b, g)), g
        for mode in (2, 3):
            ilu.SetSweepMode(mode, mode, 512)
            for _ in range(2):
                assert np.array_equal(ilu.Substitute(b), x_level), (g, mode)
    assert ilu.Schedule()["pack_l"]["max_width"] > 64


@pytest.mark.parametrize("lof,closed_form", [(0, lambda s: 3 * s - 2), (1, lambda s: 6 * s - 5)])
def test_schedule_levels_closed_form(oracle, lof, closed_form):
    """Levels of the L and U dependency graphs of a 3D 7-point stencil:
    3s-2 for ILU(0) and 6s-5 for ILU(1) (the fill couples (x+1, y-1) and
    similar diagonals, doubling the critical path)."""
    s = 11
    a = gen.convdiff3d(s)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, lof)
    ilu = factored(o[0], o[1], o[3])
    sch = ilu.Schedule()
    assert sch["n"] == a.n and sch["nnz"] == o[0][-1]
    assert sch["levels_l"] == closed_form(s) and sch["levels_u"] == closed_form(s)
    assert sch["pack_l"] is None
    ilu.SetSweepMode(2, 2, 256)
    b = np.ones(a.n)
    ilu.Substitute(b)
    sch = ilu.Schedule()
    assert sch["mode_l"] == 2 and sch["blk_rows"] == 256 and sch["blocks"] == -(-a.n // 256)
    assert sch["pack_l"]["entries"] >= (o[0][-1] - a.n) // 2 and sch["pack_l"]["max_passes_per_block"] % 8 == 0


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("tile", [(8, 8, 4), (20, 5, 3), (7, 9, 11), (20, 20, 1), (20, 4, 8)])
@pytest.mark.parametrize("lof", [0, 1])
def test_tiled_sweeps_bit_exact(oracle, tile, lof, bits):
    """Block-local sweeps over 3D tiles of the grid (including tiles that do
    not divide it, and the benchmark's x-extent x 4 x 8 shape) reproduce the
    oracle's lane-order sweep bit for bit, in either direction and mixed
    with the level-scheduled sweep, with FP64 factors and with FP32 factors
    (the oracle then sweeps the same float-rounded factors)."""
    s = 20
    a = gen.convdiff3d(s)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, lof)
    frp, fci = o[0], o[1]
    ilu = factored(frp, fci, o[3], bits)
    ofac, dg, _ = oracle.ilu_factorize(frp, fci, o[3])
    if bits == 32:
        ofac = ofac.astype(np.float32).astype(np.float64)
    ilu.SetGridShape(s, s, s)
    ilu.SetSweepTiles(2, 2, *tile)
    sch = ilu.Schedule()
    # ILU(0) tiles are plain boxes; ILU(1)'s fill needs u = x+y+z, v = y+z
    assert sch["tile"] == list(tile) and sch["skew"] == ([0, 0, 0] if lof == 0 else [1, 1, 1])
    assert sch["max_block_rows"] <= np.prod(tile)
    if lof == 0:
        assert sch["blocks"] == np.prod([-(-s // t) for t in tile])
    rng = np.random.default_rng(sum(tile) + lof)
    for g in (1, 2, 8, 16):
        ilu.SetSweepLanes(g, g)
        b = rng.uniform(-1, 1, a.n)
        xo = oracle.substitute_lanes(frp, fci, dg, ofac, b, g)
        assert np.array_equal(ilu.Substitute(b), xo), g
        assert np.array_equal(ilu.Substitute(b), xo), g  # re-armed buffers
    ilu.SetSweepTiles(0, 2, *tile)
    assert np.array_equal(ilu.Substitute(b), xo)
    ilu.SetSweepTiles(3, 3, *tile)  # barrier-free form: LDS value-flags
    assert np.array_equal(ilu.Substitute(b), xo)
    assert ilu.SweepConfig()["tile"] == list(tile)


def test_tiles_need_a_grid_and_follow_the_dependencies():
    a = gen.convdiff3d(6)
    ilu = factored(a.row_ptr, a.col_idx, a.value)
    with pytest.raises(RuntimeError):
        ilu.SetSweepTiles(2, 2, 2, 2, 2)  # no grid declared
    with pytest.raises(RuntimeError):
        ilu.SetGridShape(6, 6, 5)  # 180 != 216
    ilu.SetGridShape(6, 6, 6)
    with pytest.raises(RuntimeError):
        ilu.SetSweepTiles(1, 1, 2, 2, 2)  # the CSR block sweep needs contiguous blocks
    with pytest.raises(RuntimeError):
        ilu.SetSweepTiles(2, 2, 64, 64, 8)  # more than 16384 rows per tile
    # A 4 x 2 grid: row 2 needs row 1 and row 4 = (0, 1) needs row 3 =
    # (3, 0).  Plain 2 x 2 boxes would depend on each other both ways; the
    # offset (+3, -1) makes the tiles skew by u = x + 3y, which breaks the
    # cycle.
    rp = np.array([0, 1, 2, 4, 5, 7, 8, 9, 10], dtype=np.int32)
    ci = np.array([0, 1, 1, 2, 3, 3, 4, 5, 6, 7], dtype=np.int32)
    va = np.where(ci == np.repeat(np.arange(8), np.diff(rp)), 4.0, -1.0)
    ilu2 = factored(rp, ci, va)
    ilu2.SetGridShape(4, 2, 1)
    ilu2.SetSweepTiles(2, 2, 2, 2, 1)
    assert ilu2.Schedule()["skew"] == [3, 0, 0]
    b = np.arange(8.0)
    ilu2.SetSweepLanes(1, 1)
    x = ilu2.Substitute(b)
    L = np.eye(8)
    U = np.zeros((8, 8))
    for i in range(8):
        for p in range(rp[i], rp[i + 1]):
            j = ci[p]
            if j < i:
                L[i, j] = va[p] / 4.0  # ILU(0) of this matrix: l_ij = a_ij / u_jj, u_jj = 4
            else:
                U[i, j] = va[p]
    assert np.allclose(L @ (U @ x), b, rtol=0, atol=1e-12)


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("case", ["poisson64", "rand_deep", "rand_long"])
def test_level_walk_sweep_bit_exact(oracle, bits, case):
    """The single-workgroup level walk (mode 4) against the oracle's
    lane-order sweep, bit for bit, at every lanes-per-row shape: levels
    wider than the workgroup's row slots (2D Poisson, up to 64 rows per
    level), a deep random ILU(2) pattern (~13 rows per level, rows of ~100
    entries: the 4-entry-per-lane head and the tail), and rows of ~150
    entries per direction.  Mixed with the other modes per direction and
    applied repeatedly (re-armed buffers)."""
    if case == "poisson64":
        a = gen.poisson2d(64)
        rp, ci, va = a.row_ptr, a.col_idx, a.value
    elif case == "rand_deep":
        rp0, ci0 = random_pattern(6000, 12, 300, seed=5)
        va0 = dominant_values(rp0, ci0, seed=5)
        o = oracle.iluk_symbolic(rp0, ci0, va0, 2)
        rp, ci, va = o[0], o[1], o[3]
    else:
        rp, ci = random_pattern(1200, 150, 400, seed=7)
        va = dominant_values(rp, ci, seed=7)
    n = rp.size - 1
    ilu = factored(rp, ci, va, bits)
    ofac, dg, _ = oracle.ilu_factorize(rp, ci, va)
    if bits == 32:
        ofac = ofac.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(bits + len(case))
    for g in LANES:
        ilu.SetSweepLanes(g, g)
        ilu.SetSweepMode(4, 4, 0)
        assert ilu.SweepConfig()["mode_l"] == 4 and ilu.SweepConfig()["mode_u"] == 4
        for _ in range(2):
            b = rng.uniform(-1, 1, n)
            assert np.array_equal(ilu.Substitute(b), oracle.substitute_lanes(rp, ci, dg, ofac, b, g)), g
    b = rng.uniform(-1, 1, n)
    for g in (4, 16):
        ilu.SetSweepLanes(g, g)
        ilu.SetSweepMode(0, 0, 0)
        x_level = ilu.Substitute(b)
        for ml, mu, blk in ((4, 0, 0), (0, 4, 0), (4, 2, 512), (2, 4, 512), (4, 4, 0)):
            ilu.SetSweepMode(ml, mu, blk)
            for _ in range(2):
                assert np.array_equal(ilu.Substitute(b), x_level), (g, ml, mu)


def test_level_walk_follows_new_pattern_and_values(oracle):
    """Mode 4's per-level descriptors are rebuilt for a new pattern and the
    sweep reads refactorized values."""
    ilu = P.IluSolver(0)
    for s, scale in ((30, 1.0), (41, 1.3)):
        a = gen.poisson2d(s)
        va = a.value * np.linspace(1.0, scale, a.nnz)
        ilu.SetMatrix(a.row_ptr, a.col_idx, va)
        assert ilu.SetupMatrix() == 0 and ilu.Factorize() == 0
        ilu.SetSweepLanes(8, 8)
        ilu.SetSweepMode(4, 4, 0)
        ofac, dg, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, va)
        b = np.random.default_rng(s).uniform(-1, 1, a.n)
        assert np.array_equal(ilu.Substitute(b), oracle.substitute_lanes(a.row_ptr, a.col_idx, dg, ofac, b, 8))
        va2 = va * 1.1
        P.fill_csr(ilu.GetMatrix(), a.row_ptr, a.col_idx, va2)
        assert ilu.Factorize() == 0
        ofac2, _, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, va2)
        assert np.array_equal(ilu.Substitute(b), oracle.substitute_lanes(a.row_ptr, a.col_idx, dg, ofac2, b, 8))


def test_autotune_tries_level_walk_on_deep_patterns(oracle):
    """On a pattern with few rows per level the autotuner times mode 4 and
    whatever it picks keeps the results within 1e-12 of the oracle."""
    rp0, ci0 = random_pattern(6000, 12, 300, seed=9)
    va0 = dominant_values(rp0, ci0, seed=9)
    o = oracle.iluk_symbolic(rp0, ci0, va0, 2)
    rp, ci, va = o[0], o[1], o[3]
    ilu = factored(rp, ci, va)
    nl, nu = ilu.Levels()
    assert (rp.size - 1) / max(nl, nu) <= 256
    cfg, table = ilu.Autotune(reps=1)
    assert any(int(r[0]) == 4 for r in table)
    ofac, dg, _ = oracle.ilu_factorize(rp, ci, va)
    b = np.random.default_rng(1).uniform(-1, 1, rp.size - 1)
    xo = oracle.substitute(rp, ci, dg, ofac, b)
    assert np.max(np.abs(ilu.Substitute(b) - xo)) <= 1e-12 * np.max(np.abs(xo))


@pytest.mark.parametrize("case", ["poisson128", "two_copies"])
def test_level_walk_cache_paths_bit_exact(oracle, case):
    """Mode 4 keeps recent x in an LDS ring of 8192 tagged slots.  2D Poisson
    128^2 (n = 16384): the ring is used and reads of rows whose slot was
    reused go to memory.  Two copies of an 8192-row tridiagonal system side
    by side: rows i and i + 8192 share a level and a slot, so no level uses
    the ring.
    Both bit-identical to the oracle's lane-order sweep, applied twice."""
    if case == "poisson128":
        a = gen.poisson2d(128)
        rp, ci, va = a.row_ptr, a.col_idx, a.value
    else:
        import scipy.sparse as sp
        m = sp.diags([-1.0, 2.5, -1.0], [-1, 0, 1], shape=(8192, 8192), format="csr")
        a2 = sp.block_diag([m, m], format="csr")
        a2.sort_indices()
        rp = a2.indptr.astype(np.int32)
        ci = a2.indices.astype(np.int32)
        va = a2.data.astype(np.float64)
    rp, ci, va = (np.ascontiguousarray(v) for v in (rp, ci, va))
    n = rp.size - 1
    ilu = factored(rp, ci, va)
    ofac, dg, _ = oracle.ilu_factorize(rp, ci, va)
    rng = np.random.default_rng(n)
    for g in (4, 16, 64):
        ilu.SetSweepLanes(g, g)
        ilu.SetSweepMode(4, 4, 0)
        for _ in range(2):
            b = rng.uniform(-1, 1, n)
            assert np.array_equal(ilu.Substitute(b), oracle.substitute_lanes(rp, ci, dg, ofac, b, g)), g
        sch = ilu.Schedule()
        nl, nu = ilu.Levels()
        assert sch["lv_cached_levels"] == ([nl, nu] if case == "poisson128" else [0, 0])


def test_watchdog_zero_refuses_work(oracle):
    """A watchdog limit of 0 makes every bounded dependency wait give up at
    once: the kernels drain normally and the call that launched them fails.
    A solver whose factorization or analysis did not complete refuses to
    sweep or solve, and works again once both are redone with a limit."""
    a = gen.convdiff3d(20)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    frp, fci, fval = o[0], o[1], o[3]
    ilu = factored(frp, fci, fval)
    b = np.random.default_rng(12).uniform(-1, 1, a.n)
    ofac, dg, _ = oracle.ilu_factorize(frp, fci, fval)
    x_ok = ilu.Substitute(b)
    assert np.max(np.abs(x_ok - oracle.substitute(frp, fci, dg, ofac, b))) <= 1e-12 * np.max(np.abs(x_ok))
    ilu.SetWatchdog(0.0)
    P.fill_csr(ilu.GetMatrix(), frp, fci, fval)  # Factorize left the factors in F
    assert ilu.Factorize() == -1
    with pytest.raises(RuntimeError):
        ilu.Substitute(b)
    g = P.PreconditionedGmres()
    g.SetIluPreconditioner(ilu)
    ok, its = g.Solve(P.HostCsr.from_csr(a), b, np.zeros(a.n))
    assert not ok
    # the analysis waits too: with a zero limit Setup fails and leaves
    # nothing to factorize, even with the limit raised again
    assert ilu.SetupMatrix() == -1
    ilu.SetWatchdog(10.0)
    P.fill_csr(ilu.GetMatrix(), frp, fci, fval)
    assert ilu.Factorize() == -1
    with pytest.raises(RuntimeError):
        ilu.Substitute(b)
    assert ilu.SetupMatrix() == 0
    assert ilu.Factorize() == 0
    assert np.array_equal(host_factors(ilu, fci.size), ofac)
    assert np.array_equal(ilu.Substitute(b), x_ok)
    with pytest.raises(RuntimeError):
        ilu.SetWatchdog(-1.0)


def test_level_walk_refuses_oversized_descriptor_table():
    """The level walk keeps one descriptor per (level, row slot): a chain
    (one row per level) with 1 lane per row (1024 slots) would need 1024
    descriptors per row.  It is refused with an error instead of an
    allocation that grows with levels x slots; 64 lanes (16 slots) fit."""
    import scipy.sparse as sp

    n = 1 << 17
    m = sp.diags([-1.0, 2.5], [-1, 0], shape=(n, n), format="csr")
    m.sort_indices()
    rp, ci, va = (np.ascontiguousarray(v) for v in (m.indptr.astype(np.int32), m.indices.astype(np.int32), m.data))
    ilu = factored(rp, ci, va)
    ilu.SetSweepLanes(1, 1)
    with pytest.raises(RuntimeError):
        ilu.SetSweepMode(4, 4, 0)
        ilu.Substitute(np.ones(n))
    ilu.SetSweepLanes(64, 64)
    ilu.SetSweepMode(4, 4, 0)
    x = ilu.Substitute(np.ones(n))
    assert np.all(np.isfinite(x))


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("part", [("tile", (20, 4, 8)), ("tile", (8, 8, 4)), ("tile", (7, 9, 11)), ("blk", 512),
                                  ("blk", 4096)])
@pytest.mark.parametrize("lof", [0, 1])
def test_wave_specialised_sweep_bit_exact(oracle, part, lof, bits):
    """Mode 5 (solver, stream-loader and importer waves per block) over
    tiles and contiguous blocks: bit for bit the oracle's lane-order sweep
    at every lanes-per-row count it supports, applied repeatedly (the face
    buffers are re-armed by the opposite sweep), and mixed per direction
    with the level-scheduled and packed sweeps."""
    s = 20
    a = gen.convdiff3d(s)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, lof)
    frp, fci = o[0], o[1]
    ilu = factored(frp, fci, o[3], bits)
    ilu.SetWatchdog(2.0)
    ofac, dg, _ = oracle.ilu_factorize(frp, fci, o[3])
    if bits == 32:
        ofac = ofac.astype(np.float32).astype(np.float64)

    def set_modes(ml, mu):
        if part[0] == "tile":
            ilu.SetSweepTiles(ml, mu, *part[1])
        else:
            ilu.SetSweepMode(ml, mu, part[1])

    ilu.SetGridShape(s, s, s)
    rng = np.random.default_rng(lof * 7 + bits)
    for g in (1, 2, 4, 8, 16):
        ilu.SetSweepLanes(g, g)
        set_modes(5, 5)
        assert ilu.SweepConfig()["mode_l"] == 5
        for _ in range(3):
            b = rng.uniform(-1, 1, a.n)
            assert np.array_equal(ilu.Substitute(b), oracle.substitute_lanes(frp, fci, dg, ofac, b, g)), g
    b = rng.uniform(-1, 1, a.n)
    ilu.SetSweepLanes(4, 4)
    xo = oracle.substitute_lanes(frp, fci, dg, ofac, b, 4)
    for ml, mu in ((5, 0), (0, 5), (5, 2), (3, 5), (5, 5)):
        set_modes(ml, mu)
        for _ in range(2):
            assert np.array_equal(ilu.Substitute(b), xo), (ml, mu)


def test_wave_specialised_sweep_long_rows_and_refactorization(oracle):
    """Rows of ~150 entries: with 16 lanes per row a record holds 4 rows and
    fits the ring; with one lane per row it would not, and the sweep is
    refused.  After new values are factorized, the record streams carry the
    new factors."""
    rp, ci = random_pattern(1200, 150, 400, seed=7)
    va = dominant_values(rp, ci, seed=7)
    ilu = factored(rp, ci, va)
    ilu.SetWatchdog(2.0)
    ofac, dg, _ = oracle.ilu_factorize(rp, ci, va)
    n = rp.size - 1
    b = np.random.default_rng(5).uniform(-1, 1, n)
    ilu.SetSweepLanes(16, 16)
    ilu.SetSweepMode(5, 5, 256)
    for _ in range(2):
        assert np.array_equal(ilu.Substitute(b), oracle.substitute_lanes(rp, ci, dg, ofac, b, 16))
    va2 = va * np.linspace(1.0, 1.2, va.size)
    P.fill_csr(ilu.GetMatrix(), rp, ci, va2)
    assert ilu.Factorize() == 0
    ofac2, _, _ = oracle.ilu_factorize(rp, ci, va2)
    assert np.array_equal(ilu.Substitute(b), oracle.substitute_lanes(rp, ci, dg, ofac2, b, 16))
    ilu.SetSweepLanes(1, 1)
    with pytest.raises(RuntimeError):
        ilu.SetSweepMode(5, 5, 256)
        ilu.Substitute(b)


def test_gmres_with_wave_specialised_sweeps(oracle):
    """GMRES(50) on 3D convection-diffusion ILU(1) with mode 5 on tiles takes
    the same steps to the same bits as with the level-scheduled sweep (same
    lanes), within one step of the oracle."""
    s = 24
    a = gen.convdiff3d(s)
    b = gen.rhs_uniform(a.n, seed=6)
    tol = 1e-8 * float(np.linalg.norm(b))
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    am = P.HostCsr.from_csr(a)
    runs = []
    for mode in (0, 5):
        ilu = factored(o[0], o[1], o[3])
        ilu.SetWatchdog(2.0)
        ilu.SetSweepLanes(2, 2)
        if mode:
            ilu.SetGridShape(s, s, s)
            ilu.SetSweepTiles(5, 5, s, 4, 8)
        g = P.PreconditionedGmres()
        g.SetIluPreconditioner(ilu)
        prm = g.param()
        prm.tolerance, prm.krylov_subspace_dimension, prm.max_iterations = tol, 50, 1000
        x = np.zeros(a.n)
        ok, its = g.Solve(am, b, x)
        assert ok
        runs.append((its, x))
    assert runs[0][0] == runs[1][0] and np.array_equal(runs[0][1], runs[1][1])
    fac, dg, _ = oracle.ilu_factorize(o[0], o[1], o[3])
    ok_o, its_o, _, _ = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, tol, 50, 1000, factors=(o[0], o[1], dg, fac))
    assert ok_o and abs(runs[1][0] - its_o) <= 1
