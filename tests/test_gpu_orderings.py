"""ILU(k)-GMRES on symmetrically permuted systems (AMD and nested
dissection, as the reference benchmark's --amd / --nd do,
test/benchmark.cpp:44-80).

For each ordering the device path is compared with the oracle on the same
permuted matrix: the level schedules the library builds (levels of L and U)
against a longest-path count over the oracle's ILU(k) pattern, the fill
nnz(F), and the GMRES iteration count (+-1) and solution."""
from __future__ import annotations

import numpy as np
import pytest

import par_ilu0_cpp_amd as P
from par_ilu0_cpp_amd import generators as gen

pytestmark = pytest.mark.gpu


def schedule_depths(rp, ci):
    """Levels of the forward (L) and backward (U) sweeps of the pattern F:
    row i waits for every row its strictly lower (upper) part names."""
    n = len(rp) - 1
    lo = np.zeros(n, np.int64)
    for i in range(n):
        c = ci[rp[i]:rp[i + 1]]
        c = c[c < i]
        lo[i] = lo[c].max() + 1 if c.size else 0
    up = np.zeros(n, np.int64)
    for i in range(n - 1, -1, -1):
        c = ci[rp[i]:rp[i + 1]]
        c = c[c > i]
        up[i] = up[c].max() + 1 if c.size else 0
    return int(lo.max()) + 1, int(up.max()) + 1


def permuted(a, name):
    if name == "natural":
        return a.row_ptr, a.col_idx, a.value
    order = P.CsrAmdOrder if name == "amd" else P.CsrNdOrder
    rc, p, ip = order(P.HostCsr(a.row_ptr, a.col_idx, a.value))
    assert rc == (0 if name == "amd" else 1)
    return P.permute_symmetric(a.row_ptr, a.col_idx, a.value, ip)


@pytest.mark.parametrize("name", ["natural", "amd", "nd"])
@pytest.mark.parametrize("lof", [0, 1])
def test_ordered_solve_matches_oracle(oracle, name, lof):
    import torch

    a = gen.convdiff3d(16)
    rp, ci, va = permuted(a, name)
    n = a.n
    b = gen.rhs_uniform(n, seed=3)
    o = oracle.iluk_symbolic(rp, ci, va, lof)
    fac, dg, _ = oracle.ilu_factorize(o[0], o[1], o[3])
    ok_o, its_o, x_o, _ = oracle.gmres((rp, ci, va), b, 1e-8 * np.linalg.norm(b), 30, 1000,
                                       factors=(o[0], o[1], dg, fac))
    assert ok_o
    d_rp, d_ci, d_va, d_b = (torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (rp, ci, va, b))
    ilu = P.IluSolver(0)
    assert ilu.SymbolicDevice(n, d_rp.data_ptr(), d_ci.data_ptr(), lof) == o[0][-1]
    js = ilu.Schedule()
    assert (js["levels_l"], js["levels_u"]) == schedule_depths(o[0], o[1])
    assert ilu.FactorizeDevice(d_va.data_ptr(), True) == 0
    g = P.PreconditionedGmres()
    g.SetIluPreconditioner(ilu)
    prm = g.param()
    prm.tolerance, prm.krylov_subspace_dimension, prm.max_iterations = 1e-8 * np.linalg.norm(b), 30, 1000
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    ok, its = g.SolveDevice(n, d_rp.data_ptr(), d_ci.data_ptr(), d_va.data_ptr(), d_b.data_ptr(), x.data_ptr())
    assert ok and abs(its - its_o) <= 1
    xh = x.cpu().numpy()
    assert np.max(np.abs(xh - x_o)) <= 1e-6 * np.max(np.abs(x_o))
