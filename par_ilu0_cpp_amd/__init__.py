"""MI355X (gfx950) ILU(k)-preconditioned restarted GMRES.

Python host mirror of the reference interface of lljbash/par_ilu0_cpp over
this project's C ABI (include/*.h).  Names follow the reference's C++ API
(include/ilu_solver.hpp:8-30, include/gmres.hpp:9-22, include/iluk.h:15) so
that tests read like the reference's own drivers (test/test.c,
test/benchmark.cpp).  All numerical work runs in the HIP libraries built from
``csrc/``; nothing here computes a result.
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import _lib
from ._lib import CsrMatrix, GmresParameters, GmresStat, fn, lib

__all__ = [
    "CsrMatrix",
    "GmresParameters",
    "GmresStat",
    "HostCsr",
    "IluSolver",
    "PreconditionedGmres",
    "BatchGmres",
    "block_diagonal",
    "IlukSymbolic",
    "iluk_symbolic_levels",
    "CsrAmdOrder",
    "CsrNdOrder",
    "permute_symmetric",
    "device_available",
    "lib",
]


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _view(ptr, count: int, dtype) -> np.ndarray:
    if count <= 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(count,)).view(dtype)


def csr_to_numpy(m: CsrMatrix):
    """Copies (row_ptr, col_idx, value) out of a library CsrMatrix."""
    n = m.size
    rp = _view(m.row_ptr, n + 1, np.int32).copy()
    nnz = int(rp[-1])
    return rp, _view(m.col_idx, nnz, np.int32).copy(), _view(m.value, nnz, np.float64).copy()


def fill_csr(m: CsrMatrix, row_ptr, col_idx, value) -> None:
    """SetupCsrMatrix + copy of numpy arrays into library-owned storage."""
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col_idx, dtype=np.int32)
    va = np.ascontiguousarray(value, dtype=np.float64)
    n = rp.size - 1
    nnz = int(rp[-1])
    fn("SetupCsrMatrix")(C.byref(m), n, nnz)
    C.memmove(m.row_ptr, rp.ctypes.data, rp.nbytes)
    if nnz:
        C.memmove(m.col_idx, ci.ctypes.data, nnz * 4)
        C.memmove(m.value, va.ctypes.data, nnz * 8)


class HostCsr:
    """A library-owned host CsrMatrix (released by DestroyCsrMatrix)."""

    def __init__(self, row_ptr=None, col_idx=None, value=None):
        self.mat = CsrMatrix(-1, 0, None, None, None)
        if row_ptr is not None:
            fill_csr(self.mat, row_ptr, col_idx, value)

    @classmethod
    def from_csr(cls, csr) -> "HostCsr":
        return cls(csr.row_ptr, csr.col_idx, csr.value)

    @classmethod
    def read_mm1(cls, path: str) -> "HostCsr":
        h = cls()
        fn("ReadCsrMatrixMM1")(C.byref(h.mat), path.encode())
        return h

    @property
    def size(self) -> int:
        return self.mat.size

    @property
    def nnz(self) -> int:
        return int(self.mat.row_ptr[self.mat.size]) if self.mat.size >= 0 and self.mat.row_ptr else 0

    def arrays(self):
        return csr_to_numpy(self.mat)

    def matvec(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.size, dtype=np.float64)
        fn("CsrMatVec")(C.byref(self.mat), _ptr(x, C.c_double), _ptr(y, C.c_double))
        return y

    def destroy(self) -> None:
        if self.mat.size >= 0 or self.mat.row_ptr:
            fn("DestroyCsrMatrix")(C.byref(self.mat))

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def IlukSymbolic(a: HostCsr, lof: int, f: CsrMatrix, nzmap: np.ndarray | None = None) -> int:
    """include/iluk.h (reference include/iluk.h:15): returns nnz(F)."""
    p = _ptr(nzmap, C.c_int) if nzmap is not None else None
    return int(fn("IlukSymbolic")(C.byref(a.mat), int(lof), C.byref(f), p))


def _order(name: str, a: HostCsr):
    n = a.size
    p = np.zeros(max(n, 1), dtype=np.int32)
    ip = np.zeros(max(n, 1), dtype=np.int32)
    rc = int(fn(name)(C.byref(a.mat), _ptr(p, C.c_int), _ptr(ip, C.c_int)))
    return rc, p[:n].copy(), ip[:n].copy()


def CsrAmdOrder(a: HostCsr):
    """include/csr_matrix.h CsrAmdOrder (reference include/csr_matrix.h:47,
    src/csr_matrix.cpp:173-195): approximate minimum degree ordering of the
    pattern of A + A^T, on the host.  Returns (rc, p, ip) with p[k] the old
    index of new row k and ip[p[k]] = k; rc is 0 on success, -1 on failure."""
    return _order("CsrAmdOrder", a)


def CsrNdOrder(a: HostCsr):
    """include/csr_matrix.h CsrNdOrder (reference include/csr_matrix.h:48,
    src/csr_matrix.cpp:197-225): nested dissection ordering of the pattern of
    A + A^T, on the host.  Returns (rc, p, ip); rc is 1 on success, 0 on
    failure (the reference returns CHOLMOD's status test)."""
    return _order("CsrNdOrder", a)


def permute_symmetric(row_ptr, col_idx, value, ip):
    """B = P A P^T with B[ip[i], ip[j]] = A[i, j], columns sorted in each row:
    the permutation the reference benchmark applies (test/benchmark.cpp:44-80)."""
    rp = np.asarray(row_ptr, dtype=np.int64)
    ci = np.asarray(col_idx, dtype=np.int64)
    ip = np.asarray(ip, dtype=np.int64)
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    nr, nc = ip[rows], ip[ci]
    order = np.lexsort((nc, nr))
    out_rp = np.zeros(n + 1, dtype=np.int32)
    np.add.at(out_rp, nr + 1, 1)
    return (np.cumsum(out_rp).astype(np.int32), nc[order].astype(np.int32),
            np.asarray(value, dtype=np.float64)[order])


def iluk_symbolic_levels(a: HostCsr, lof: int):
    """Symbolic ILU(lof) on the GPU returning (row_ptr, col_idx, value,
    levels, nzmap) of F as numpy arrays."""
    f = HostCsr()
    nzmap = np.zeros(max(a.nnz, 1), dtype=np.int32)
    nnzf = IlukSymbolic(a, lof, f.mat, nzmap)
    if nnzf < 0:
        raise RuntimeError("IlukSymbolic failed")
    lev = np.zeros(max(nnzf, 1), dtype=np.int8)
    rc = fn("IlukSymbolicLevelsEx")(
        C.byref(a.mat), int(lof), C.byref(f.mat), _ptr(nzmap, C.c_int), _ptr(lev, C.c_byte), lev.size
    )
    if rc != nnzf:
        raise RuntimeError(f"IlukSymbolicLevelsEx returned {rc}, expected {nnzf}")
    rp, ci, va = f.arrays()
    return rp, ci, va, lev[:nnzf].copy(), nzmap[: a.nnz].copy()


class IluSolver:
    """lljbash::IluSolver through the C ABI (include/ilu_solver_c.h)."""

    def __init__(self, nthreads: int = 0):
        self.hdl = fn("IluSolverCreate")(int(nthreads))

    def GetMatrix(self) -> CsrMatrix:
        return fn("IluSolverGetMatrix")(self.hdl).contents

    def SetMatrix(self, row_ptr, col_idx, value) -> None:
        """Fill F (the reference's CopyCsrMatrix(GetMatrix(), &a))."""
        fill_csr(self.GetMatrix(), row_ptr, col_idx, value)

    def SetupMatrix(self) -> int:
        return int(fn("IluSolverSetup")(self.hdl))

    def Factorize(self) -> int:
        return int(fn("IluSolverFactorize")(self.hdl))

    def Substitute(self, b: np.ndarray) -> np.ndarray:
        """x = U^-1 L^-1 b through the device sweeps (host vectors)."""
        import torch

        bt = torch.as_tensor(np.ascontiguousarray(b, dtype=np.float64)).cuda()
        xt = torch.empty_like(bt)
        self.SubstituteDevice(bt.data_ptr(), xt.data_ptr())
        torch.cuda.synchronize()
        return xt.cpu().numpy()

    def SubstituteDevice(self, d_b: int, d_x: int) -> None:
        if fn("IluSolverSubstituteDeviceEx")(self.hdl, C.c_void_p(d_b), C.c_void_p(d_x)) != 0:
            raise RuntimeError("IluSolverSubstituteDeviceEx failed")

    def SubstituteNoAlloc(self, d_b: int, d_x: int) -> None:
        """SubstituteDevice without device allocation (static per-device
        scratch; the sweep structures must already exist)."""
        if fn("IluSolverSubstituteNoAllocEx")(self.hdl, C.c_void_p(d_b), C.c_void_p(d_x)) != 0:
            raise RuntimeError("IluSolverSubstituteNoAllocEx failed")

    def SetFactorPrecision(self, bits: int) -> None:
        if fn("IluSolverSetFactorPrecisionEx")(self.hdl, int(bits)) != 0:
            raise RuntimeError("bad precision")

    def SetSweepLanes(self, lanes_l: int, lanes_u: int) -> None:
        if fn("IluSolverSetSweepLanesEx")(self.hdl, int(lanes_l), int(lanes_u)) != 0:
            raise RuntimeError("bad lanes")

    def SetDiagonalCompensation(self, d_omega: int) -> None:
        """Device pointer to n float32 relaxation factors (0 disables)."""
        if fn("IluSolverSetDiagonalCompensationEx")(self.hdl, C.c_void_p(d_omega or None)) != 0:
            raise RuntimeError("SetDiagonalCompensation failed")

    def SetPivotPolicy(self, policy: int, tiny: float = 1e-16, replacement: float = 1e-10) -> None:
        """dcsrilu0-style pivot check of the next factorizations: 0 none,
        1 fail on a pivot below `tiny`, 2 replace it by `replacement`."""
        if fn("IluSolverSetPivotPolicyEx")(self.hdl, int(policy), float(tiny), float(replacement)) != 0:
            raise RuntimeError("SetPivotPolicy failed")

    def FactorStatus(self) -> int:
        return int(fn("IluSolverFactorStatusEx")(self.hdl))

    def TrsvDevice(self, upper: bool, alpha: float, d_x: int, d_y: int) -> None:
        if fn("IluSolverTrsvDeviceEx")(self.hdl, int(bool(upper)), float(alpha), C.c_void_p(d_x),
                                       C.c_void_p(d_y)) != 0:
            raise RuntimeError("IluSolverTrsvDeviceEx failed")

    def Trsv(self, upper: bool, alpha: float, x: np.ndarray) -> np.ndarray:
        """y = alpha T^-1 x for T = L (unit lower) or U (host vectors)."""
        xx = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(xx)
        if fn("IluSolverTrsvEx")(self.hdl, int(bool(upper)), float(alpha), _ptr(xx, C.c_double),
                                 _ptr(y, C.c_double)) != 0:
            raise RuntimeError("IluSolverTrsvEx failed")
        return y

    def SetWatchdog(self, seconds: float) -> None:
        """Wall-clock limit of every bounded dependency wait (0..3600 s)."""
        if fn("IluSolverSetWatchdogEx")(self.hdl, float(seconds)) != 0:
            raise RuntimeError("SetWatchdog failed")

    def Autotune(self, reps: int = 3):
        """Times every sweep shape and keeps the fastest per direction.
        -> (config dict, table[rows, 8]: mode, block rows, tile x, y, z,
        lanes, fwd ms, bwd ms)."""
        table = np.zeros((128, 8), dtype=np.float64)
        rc = fn("IluSolverAutotuneEx")(self.hdl, int(reps), _ptr(table, C.c_double), 128)
        if rc < 0:
            raise RuntimeError("IluSolverAutotuneEx failed")
        return self.SweepConfig(), table[: min(rc, 128)].copy()

    def SetSweepMode(self, mode_l: int, mode_u: int, blk_rows: int = 0) -> None:
        if fn("IluSolverSetSweepModeEx")(self.hdl, int(mode_l), int(mode_u), int(blk_rows)) != 0:
            raise RuntimeError("SetSweepMode failed")

    def SetGridShape(self, nx: int, ny: int, nz: int) -> None:
        if fn("IluSolverSetGridShapeEx")(self.hdl, int(nx), int(ny), int(nz)) != 0:
            raise RuntimeError("SetGridShape failed")

    def SetSweepTiles(self, mode_l: int, mode_u: int, tx: int, ty: int, tz: int) -> None:
        if fn("IluSolverSetSweepTilesEx")(self.hdl, int(mode_l), int(mode_u), int(tx), int(ty), int(tz)) != 0:
            raise RuntimeError("SetSweepTiles failed")

    def Schedule(self) -> dict:
        """Schedules and sweep configuration (IluSolver::DescribeSchedule)."""
        need = fn("IluSolverScheduleJsonEx")(self.hdl, None, 0)
        if need < 0:
            raise RuntimeError("IluSolverScheduleJsonEx failed")
        buf = C.create_string_buffer(need + 1)
        fn("IluSolverScheduleJsonEx")(self.hdl, buf, need + 1)
        return json.loads(buf.value.decode())

    def SweepConfig(self) -> dict:
        cfg = np.zeros(8, dtype=np.int32)
        fn("IluSolverGetSweepConfigEx")(self.hdl, _ptr(cfg, C.c_int))
        return dict(mode_l=int(cfg[0]), lanes_l=int(cfg[1]), mode_u=int(cfg[2]), lanes_u=int(cfg[3]),
                    blk_rows=int(cfg[4]), tile=[int(cfg[5]), int(cfg[6]), int(cfg[7])])

    def Levels(self) -> tuple[int, int]:
        a, b = C.c_int(0), C.c_int(0)
        fn("IluSolverGetLevelsEx")(self.hdl, C.byref(a), C.byref(b))
        return a.value, b.value

    def SymbolicDevice(self, n: int, d_row_ptr: int, d_col_idx: int, lof: int) -> int:
        m = CsrMatrix(n, 0, C.cast(C.c_void_p(d_row_ptr), C.POINTER(C.c_int)),
                      C.cast(C.c_void_p(d_col_idx), C.POINTER(C.c_int)), None)
        return int(fn("IluSolverSymbolicDeviceEx")(self.hdl, C.byref(m), int(lof)))

    def FactorizeDevice(self, d_values: int, values_are_a: bool) -> int:
        return int(fn("IluSolverFactorizeDeviceEx")(self.hdl, C.c_void_p(d_values), int(values_are_a)))

    def Nonzeros(self) -> int:
        return int(fn("IluSolverNonzerosEx")(self.hdl))

    def DeviceFactors(self) -> np.ndarray:
        out = np.empty(max(self.Nonzeros(), 1), dtype=np.float64)
        rc = fn("IluSolverCopyFactorsEx")(self.hdl, _ptr(out, C.c_double))
        if rc < 0:
            raise RuntimeError("IluSolverCopyFactorsEx failed")
        return out[:rc]

    def FillLevels(self) -> np.ndarray:
        out = np.empty(max(self.Nonzeros(), 1), dtype=np.int8)
        rc = fn("IluSolverCopyFillLevelsEx")(self.hdl, _ptr(out, C.c_byte))
        return out[:rc]

    def close(self) -> None:
        if self.hdl:
            fn("IluSolverDestroy")(self.hdl)
            self.hdl = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PreconditionedGmres:
    """lljbash::PreconditionedGmres through the C ABI (include/gmres_c.h)."""

    def __init__(self):
        self.hdl = fn("GmresCreate")()
        self._ilu = None

    def param(self) -> GmresParameters:
        return fn("GmresGetParameters")(self.hdl).contents

    def stat(self) -> GmresStat:
        return fn("GmresGetStat")(self.hdl).contents

    def SetIluPreconditioner(self, ilu: IluSolver) -> None:
        self._ilu = ilu
        fn("GmresSetPreconditioner")(self.hdl, ilu.hdl)

    def Solve(self, a: HostCsr, rhs: np.ndarray, sol: np.ndarray) -> tuple[bool, int]:
        """sol (float64, contiguous) is the initial guess and receives x."""
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        assert sol.dtype == np.float64 and sol.flags.c_contiguous
        it = C.c_int(0)
        ok = fn("GmresSolve")(self.hdl, C.byref(a.mat), _ptr(rhs, C.c_double), _ptr(sol, C.c_double), C.byref(it))
        return bool(ok), it.value

    def SolveDevice(self, n: int, d_row_ptr: int, d_col_idx: int, d_value: int, d_rhs: int, d_sol: int):
        m = CsrMatrix(n, 0, C.cast(C.c_void_p(d_row_ptr), C.POINTER(C.c_int)),
                      C.cast(C.c_void_p(d_col_idx), C.POINTER(C.c_int)),
                      C.cast(C.c_void_p(d_value), C.POINTER(C.c_double)))
        it = C.c_int(0)
        ok = fn("GmresSolveDeviceEx")(self.hdl, C.byref(m), C.c_void_p(d_rhs), C.c_void_p(d_sol), C.byref(it))
        return bool(ok), it.value

    def LastTrueResidual(self) -> float:
        return float(fn("GmresLastTrueResidualEx")(self.hdl))

    def ResidualHistory(self, cap: int = 100000) -> np.ndarray:
        out = np.empty(cap, dtype=np.float64)
        k = fn("GmresResidualHistoryEx")(self.hdl, _ptr(out, C.c_double), cap)
        return out[: max(k, 0)].copy()

    def close(self) -> None:
        if self.hdl:
            fn("GmresDestroy")(self.hdl)
            self.hdl = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _dev_csr(n: int, d_row_ptr: int, d_col_idx: int, d_value: int = 0) -> CsrMatrix:
    return CsrMatrix(n, 0, C.cast(C.c_void_p(d_row_ptr), C.POINTER(C.c_int)),
                     C.cast(C.c_void_p(d_col_idx), C.POINTER(C.c_int)),
                     C.cast(C.c_void_p(d_value), C.POINTER(C.c_double)) if d_value else None)


class BatchGmres:
    """Batched ILU(k)-GMRES over a block-diagonal device matrix
    (LLJBASH_BatchGmres*Ex): every (system, RHS) pair runs the reference
    iteration with its own tolerance rel_tol * ||b||."""

    def __init__(self):
        self.hdl = fn("BatchGmresCreateEx")()
        self.nsys = 0
        self.nrhs = 0

    def Setup(self, n: int, d_row_ptr: int, d_col_idx: int, lof: int, sys_ptr, nrhs: int) -> int:
        sp = np.ascontiguousarray(sys_ptr, dtype=np.int32)
        self.nsys, self.nrhs, self.n = sp.size - 1, int(nrhs), int(n)
        m = _dev_csr(n, d_row_ptr, d_col_idx)
        return int(fn("BatchGmresSetupEx")(self.hdl, C.byref(m), int(lof), self.nsys, _ptr(sp, C.c_int), self.nrhs))

    def SetFactorPrecision(self, bits: int) -> None:
        if fn("BatchGmresSetFactorPrecisionEx")(self.hdl, int(bits)) != 0:
            raise RuntimeError("bad precision")

    def Factorize(self, d_avalues: int) -> int:
        return int(fn("BatchGmresFactorizeEx")(self.hdl, C.c_void_p(d_avalues)))

    def Solve(self, d_row_ptr: int, d_col_idx: int, d_value: int, d_rhs: int, d_sol: int, rel_tol: float, m: int,
              maxits: int):
        """-> (cycles, its[nsys, nrhs], converged[nsys, nrhs], true_rel_res[nsys, nrhs])."""
        ni = self.nsys * self.nrhs
        its = np.zeros(ni, dtype=np.int32)
        conv = np.zeros(ni, dtype=np.int32)
        rel = np.zeros(ni, dtype=np.float64)
        mat = _dev_csr(self.n, d_row_ptr, d_col_idx, d_value)
        cyc = fn("BatchGmresSolveEx")(self.hdl, C.byref(mat), C.c_void_p(d_rhs), C.c_void_p(d_sol), float(rel_tol),
                                      int(m), int(maxits), _ptr(its, C.c_int), _ptr(conv, C.c_int),
                                      _ptr(rel, C.c_double))
        if cyc < 0:
            raise RuntimeError("BatchGmresSolveEx failed")
        shape = (self.nsys, self.nrhs)
        return int(cyc), its.reshape(shape), conv.reshape(shape).astype(bool), rel.reshape(shape)

    def close(self) -> None:
        if self.hdl:
            fn("BatchGmresDestroyEx")(self.hdl)
            self.hdl = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def block_diagonal(mats):
    """Assemble host CSR systems into one block-diagonal CSR; returns
    (row_ptr, col_idx, value, sys_ptr)."""
    sys_ptr = np.zeros(len(mats) + 1, dtype=np.int64)
    sys_ptr[1:] = np.cumsum([m.n for m in mats])
    nnz_ptr = np.zeros(len(mats) + 1, dtype=np.int64)
    nnz_ptr[1:] = np.cumsum([m.nnz for m in mats])
    rp = np.empty(sys_ptr[-1] + 1, dtype=np.int64)
    rp[-1] = nnz_ptr[-1]
    ci = np.empty(nnz_ptr[-1], dtype=np.int32)
    va = np.empty(nnz_ptr[-1], dtype=np.float64)
    for k, m in enumerate(mats):
        rp[sys_ptr[k]:sys_ptr[k + 1]] = m.row_ptr[:-1] + nnz_ptr[k]
        ci[nnz_ptr[k]:nnz_ptr[k + 1]] = m.col_idx + sys_ptr[k]
        va[nnz_ptr[k]:nnz_ptr[k + 1]] = m.value
    return rp.astype(np.int32), ci, va, sys_ptr.astype(np.int32)


def device_available() -> bool:
    return bool(fn("DeviceAvailableEx")())
