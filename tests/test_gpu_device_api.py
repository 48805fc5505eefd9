"""Device-resident extension path (LLJBASH_*Ex): symbolic straight into the
solver, factorization from A's values through the kept map, device solve --
same results as the host drop-in path."""
from __future__ import annotations

import numpy as np
import pytest

import par_ilu0_cpp_amd as P
from par_ilu0_cpp_amd import generators as gen

pytestmark = pytest.mark.gpu


def test_device_path_equals_host_path(oracle):
    import torch

    a = gen.convdiff3d(20)
    b = gen.rhs_uniform(a.n, seed=7)
    tol = 1e-8 * np.linalg.norm(b)
    # host path: IlukSymbolic -> F, SetupMatrix, Factorize, Solve
    f = P.HostCsr()
    nz = np.zeros(a.nnz, dtype=np.int32)
    am = P.HostCsr.from_csr(a)
    nnzf = P.IlukSymbolic(am, 1, f.mat, nz)
    ilu_h = P.IluSolver(0)
    ilu_h.SetMatrix(*f.arrays())
    assert ilu_h.SetupMatrix() == 0 and ilu_h.Factorize() == 0
    gh = P.PreconditionedGmres()
    gh.SetIluPreconditioner(ilu_h)
    gh.param().tolerance, gh.param().krylov_subspace_dimension = tol, 50
    xh = np.zeros(a.n)
    okh, itsh = gh.Solve(am, b, xh)
    # device path
    t = {k: torch.from_numpy(v).cuda() for k, v in (("rp", a.row_ptr), ("ci", a.col_idx), ("va", a.value),
                                                     ("b", b))}
    x = torch.zeros(a.n, dtype=torch.float64, device="cuda")
    ilu = P.IluSolver(0)
    assert ilu.SymbolicDevice(a.n, t["rp"].data_ptr(), t["ci"].data_ptr(), 1) == nnzf
    assert ilu.FactorizeDevice(t["va"].data_ptr(), True) == 0
    fac = ilu.DeviceFactors()
    assert np.array_equal(fac, P._view(ilu_h.GetMatrix().value, nnzf, np.float64))
    g = P.PreconditionedGmres()
    g.SetIluPreconditioner(ilu)
    g.param().tolerance, g.param().krylov_subspace_dimension = tol, 50
    ok, its = g.SolveDevice(a.n, t["rp"].data_ptr(), t["ci"].data_ptr(), t["va"].data_ptr(), t["b"].data_ptr(),
                            x.data_ptr())
    assert ok == okh and its == itsh
    assert np.array_equal(x.cpu().numpy(), xh)
    # fill levels kept by the device symbolic pass
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, 1)
    assert np.array_equal(ilu.FillLevels().astype(np.int32), o[2])


def test_substitute_device_in_place(oracle):
    import torch

    a = gen.poisson2d(40)
    ilu = P.IluSolver(0)
    ilu.SetMatrix(a.row_ptr, a.col_idx, a.value)
    assert ilu.SetupMatrix() == 0 and ilu.Factorize() == 0
    b = np.random.default_rng(0).uniform(-1, 1, a.n)
    tb = torch.from_numpy(b.copy()).cuda()
    ilu.SubstituteDevice(tb.data_ptr(), tb.data_ptr())
    torch.cuda.synchronize()
    fac, dg, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, a.value)
    xo = oracle.substitute(a.row_ptr, a.col_idx, dg, fac, b)
    assert np.max(np.abs(tb.cpu().numpy() - xo)) <= 1e-12 * np.max(np.abs(xo))


def test_bad_device_matrix_is_rejected_before_use():
    import torch

    a = gen.poisson2d(10)
    ilu = P.IluSolver(0)
    ilu.SetMatrix(a.row_ptr, a.col_idx, a.value)
    assert ilu.SetupMatrix() == 0 and ilu.Factorize() == 0
    g = P.PreconditionedGmres()
    g.SetIluPreconditioner(ilu)
    ci = a.col_idx.copy()
    ci[7] = a.n + 5  # out of range
    t = [torch.from_numpy(v).cuda() for v in (a.row_ptr, ci, a.value, np.ones(a.n))]
    x = torch.zeros(a.n, dtype=torch.float64, device="cuda")
    ok, its = g.SolveDevice(a.n, t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), t[3].data_ptr(), x.data_ptr())
    assert not ok and its == 0
