"""The MKL-mode equivalents on the GPU (the reference's USE_MKL_ILU /
USE_MKL_SV builds, src/ilu_solver.cpp:248-259,631-639,753-773): the
dcsrilu0-style pivot policy of the factorization and single triangular
solves, against the oracle's restatement (orc_ilu0_guarded,
orc_trsv_lanes), bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

import par_ilu0_cpp_amd as P
from par_ilu0_cpp_amd import generators as gen
from tests.conftest import dominant_values, random_pattern

pytestmark = pytest.mark.gpu


def solver(rp, ci, va, bits=64, policy=None):
    ilu = P.IluSolver(0)
    ilu.SetFactorPrecision(bits)
    ilu.SetMatrix(rp, ci, va)
    assert ilu.SetupMatrix() == 0
    if policy is not None:
        ilu.SetPivotPolicy(*policy)
    return ilu


def host_factors(ilu, nnz):
    return P._view(ilu.GetMatrix().value, nnz, np.float64).copy()


def conv_ilu1(oracle, s=14):
    a = gen.convdiff3d(s)
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    return o[0], o[1], o[3]


SHAPES = [("mode", 0, 0), ("mode", 1, 512), ("mode", 2, 512), ("mode", 3, 512), ("mode", 4, 0),
          ("tile", 2, (14, 4, 4)), ("tile", 3, (14, 4, 4)), ("tile", 5, (14, 4, 4))]


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: f"{s[0]}{s[1]}")
def test_trsv_bit_exact(oracle, shape, bits):
    import torch

    frp, fci, fval = conv_ilu1(oracle)
    n = frp.size - 1
    ilu = solver(frp, fci, fval, bits)
    assert ilu.Factorize() == 0
    ofac, dg, _ = oracle.ilu_factorize(frp, fci, fval)
    if bits == 32:
        ofac = ofac.astype(np.float32).astype(np.float64)
    kind, m, arg = shape
    if kind == "mode":
        if m:
            ilu.SetSweepMode(m, m, arg)
    else:
        ilu.SetGridShape(14, 14, 14)
        ilu.SetSweepTiles(m, m, *arg)
    rng = np.random.default_rng(17 + m + bits)
    for g in (1, 8):
        ilu.SetSweepLanes(g, g)
        for upper in (False, True):
            for alpha in (1.0, -0.75):
                x = rng.uniform(-1, 1, n)
                want = oracle.trsv_lanes(frp, fci, dg, ofac, x, upper, alpha, g)
                xt = torch.from_numpy(x).cuda()
                yt = torch.empty_like(xt)
                ilu.TrsvDevice(upper, alpha, xt.data_ptr(), yt.data_ptr())
                assert np.array_equal(yt.cpu().numpy(), want), (g, upper, alpha)
                ilu.TrsvDevice(upper, alpha, xt.data_ptr(), xt.data_ptr())  # in place
                assert np.array_equal(xt.cpu().numpy(), want), (g, upper, alpha, "in place")
        # two solves compose to the preconditioner apply, and the apply
        # still works after standalone solves
        b = rng.uniform(-1, 1, n)
        bt = torch.from_numpy(b).cuda()
        yt, xt, st = torch.empty_like(bt), torch.empty_like(bt), torch.empty_like(bt)
        ilu.TrsvDevice(False, 1.0, bt.data_ptr(), yt.data_ptr())
        ilu.TrsvDevice(True, 1.0, yt.data_ptr(), xt.data_ptr())
        ilu.SubstituteDevice(bt.data_ptr(), st.data_ptr())
        want = oracle.substitute_lanes(frp, fci, dg, ofac, b, g)
        assert np.array_equal(xt.cpu().numpy(), want) and np.array_equal(st.cpu().numpy(), want), g


def test_trsv_host_vectors_and_refusal(oracle):
    rp, ci = random_pattern(3000, 8, 200, seed=4)
    va = dominant_values(rp, ci, seed=4)
    ilu = solver(rp, ci, va)
    x = np.random.default_rng(2).uniform(-1, 1, rp.size - 1)
    with pytest.raises(RuntimeError):
        ilu.Trsv(False, 1.0, x)  # not factorized yet
    assert ilu.Factorize() == 0
    ofac, dg, _ = oracle.ilu_factorize(rp, ci, va)
    ilu.SetSweepLanes(4, 4)
    for upper in (False, True):
        assert np.array_equal(ilu.Trsv(upper, 2.0, x), oracle.trsv_lanes(rp, ci, dg, ofac, x, upper, 2.0, 4))


def planted(rp, ci, va, rows_vals):
    """Rows whose entries left of the diagonal are zero and whose diagonal
    is set: their final pivot is exactly that value."""
    va = va.copy()
    for r, d in rows_vals:
        for p in range(rp[r], rp[r + 1]):
            if ci[p] < r:
                va[p] = 0.0
            elif ci[p] == r:
                va[p] = d
    return va


@pytest.mark.parametrize("case", ["short", "long"])
def test_stop_policy_codes(oracle, case):
    if case == "short":  # 5-point rows: the short-row factorization kernel
        a = gen.poisson2d(40)
        rp, ci, va = a.row_ptr, a.col_idx, a.value
    else:  # ~40-entry rows: the general kernel
        rp, ci = random_pattern(2000, 40, 300, seed=9)
        va = dominant_values(rp, ci, seed=9)
    for rows_vals, want in (([(100, 1e-20), (900, 0.0)], -103), ([(100, 0.0), (900, 1e-20)], -102),
                            ([(500, 5e-17)], -103)):
        v = planted(rp, ci, va, rows_vals)
        assert oracle.ilu0_guarded(rp, ci, v, 1)[2] == want
        ilu = solver(rp, ci, v, policy=(1, 1e-16, 1e-10))
        assert ilu.Factorize() == -1 and ilu.FactorStatus() == want, rows_vals
        # policy 0: no check, no status
        ilu.SetPivotPolicy(0)
        ilu.Factorize()
        assert ilu.FactorStatus() == 0
    ok = solver(rp, ci, va, policy=(1, 1e-16, 1e-10))
    assert ok.Factorize() == 0 and ok.FactorStatus() == 0


@pytest.mark.parametrize("bits", [64, 32])
@pytest.mark.parametrize("case", ["short", "long"])
def test_replace_policy_bit_exact(oracle, case, bits):
    rng = np.random.default_rng(5)
    if case == "short":
        a = gen.poisson2d(48)
        rp, ci = a.row_ptr, a.col_idx
        va = a.value * rng.uniform(0.5, 1.5, a.value.size)
    else:
        rp, ci = random_pattern(1500, 40, 200, seed=6)
        va = dominant_values(rp, ci, seed=6)
    n = rp.size - 1
    dg0 = oracle.diag_ptr(rp, ci)
    va = va.copy()
    va[dg0] *= rng.uniform(0.05, 1.2, n)  # weaken diagonals so pivots cross the threshold
    tiny = float(np.quantile(np.abs(oracle.ilu_factorize(rp, ci, va)[0][dg0]), 0.1))
    ofac, dg, st = oracle.ilu0_guarded(rp, ci, va, 2, tiny, 1.0)
    assert st == 0 and np.count_nonzero(ofac[dg] == 1.0) >= 10
    ilu = solver(rp, ci, va, bits, policy=(2, tiny, 1.0))
    assert ilu.Factorize() == 0 and ilu.FactorStatus() == 0
    assert np.array_equal(host_factors(ilu, ci.size), ofac)
    if bits == 32:
        ofac = ofac.astype(np.float32).astype(np.float64)
    ilu.SetSweepLanes(4, 4)
    b = rng.uniform(-1, 1, n)
    assert np.array_equal(ilu.Substitute(b), oracle.substitute_lanes(rp, ci, dg, ofac, b, 4))


def test_policy_arguments_checked():
    ilu = P.IluSolver(0)
    for bad in ((3, 1e-16, 1e-10), (1, -1.0, 1e-10), (2, 1e-16, float("inf"))):
        with pytest.raises(RuntimeError):
            ilu.SetPivotPolicy(*bad)
