"""First GPU parity checks: symbolic pattern, factors, sweeps and one solve
against the CPU oracle on small inputs."""
from __future__ import annotations

import numpy as np
import pytest

import par_ilu0_cpp_amd as P
from par_ilu0_cpp_amd import generators as gen
from tests.conftest import dominant_values, random_pattern

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lof", [0, 1, 2, 3])
def test_symbolic_matches_oracle_random(oracle, lof):
    rp, ci = random_pattern(300, 4, 20, seed=lof)
    va = dominant_values(rp, ci, seed=lof)
    a = P.HostCsr(rp, ci, va)
    frp, fci, fva, flev, nzmap = P.iluk_symbolic_levels(a, lof)
    o = oracle.iluk_symbolic(rp, ci, va, lof)
    assert np.array_equal(frp, o[0])
    assert np.array_equal(fci, o[1])
    assert np.array_equal(flev.astype(np.int32), o[2])
    assert np.array_equal(fva, o[3])
    assert np.array_equal(nzmap, o[4])


def test_factor_and_sweeps_match_oracle(oracle):
    a = gen.convdiff3d(12)
    frp, fci, fva, flev, nzmap = P.iluk_symbolic_levels(P.HostCsr.from_csr(a), 1)
    ilu = P.IluSolver(0)
    ilu.SetMatrix(frp, fci, fva)
    assert ilu.SetupMatrix() == 0
    assert ilu.Factorize() == 0
    fac = P._view(ilu.GetMatrix().value, fci.size, np.float64).copy()
    ofac, dg, z = oracle.ilu_factorize(frp, fci, fva)
    assert z == 0
    rel = np.max(np.abs(fac - ofac) / np.maximum(np.abs(ofac), 1e-300))
    assert rel <= 1e-12, rel
    b = np.random.default_rng(0).uniform(-1, 1, a.n)
    for lanes in (1, 2, 4, 8, 16, 32, 64):
        ilu.SetSweepLanes(lanes, lanes)
        x = ilu.Substitute(b)
        xo = oracle.substitute(frp, fci, dg, ofac, b)
        err = np.max(np.abs(x - xo)) / np.max(np.abs(xo))
        assert err <= 1e-12, (lanes, err)
        if lanes == 1:
            assert np.array_equal(x, xo)


def test_gmres_poisson_matches_oracle(oracle):
    a = gen.poisson2d(64)
    rhs = a.matvec(np.ones(a.n))
    ilu = P.IluSolver(0)
    ilu.SetMatrix(a.row_ptr, a.col_idx, a.value)
    assert ilu.SetupMatrix() == 0
    assert ilu.Factorize() == 0
    g = P.PreconditionedGmres()
    g.SetIluPreconditioner(ilu)
    p = g.param()
    p.tolerance, p.krylov_subspace_dimension, p.max_iterations = 1e-8, 30, 200
    x = np.zeros(a.n)
    ok, its = g.Solve(P.HostCsr.from_csr(a), rhs, x)
    fac, dg, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, a.value)
    ok_o, its_o, x_o, hist_o = oracle.gmres((a.row_ptr, a.col_idx, a.value), rhs, 1e-8, 30, 200,
                                           factors=(a.row_ptr, a.col_idx, dg, fac))
    assert ok and ok_o
    assert abs(its - its_o) <= 1, (its, its_o)
    assert np.linalg.norm(rhs - a.matvec(x)) <= 1e-8
    hist = g.ResidualHistory()
    assert hist.size == its
