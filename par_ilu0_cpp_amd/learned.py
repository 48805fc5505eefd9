"""Learned diagonal compensation: host side of the network in csrc/gnn.hip.

The network predicts per-row relaxation factors omega_i for the compensated
ILU factorization (row i adds omega_i times its dropped fill to its pivot).
It is trained on the GPU by the library's own forward/backward kernels on
targets the solver itself produces: for every training matrix the solver is
run with constant omega over a grid, and the omega with the fewest GMRES
steps (ties: the smallest omega) becomes that matrix's target.  Plain ILU(k)
(omega = 0) is the fallback whenever the learned factorization misses its
step budget (GmresSolveLearnedDeviceEx).  No counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from . import IluSolver, PreconditionedGmres, _dev_csr, _ptr, fn

WEIGHTS = Path(__file__).resolve().parent / "weights" / "gnn_compensation.json"
OMEGA_GRID = (0.0, 0.5, 0.8, 0.9, 0.95, 0.98, 1.0)


def transpose_map(rp, ci):
    """Pattern of A^T and, for each of its entries, the index of the
    corresponding entry of A -> (rpt, cit, tmap), all int32."""
    rp = np.asarray(rp)
    ci = np.asarray(ci)
    n = rp.size - 1
    rows = np.repeat(np.arange(n), np.diff(rp))
    order = np.lexsort((rows, ci))
    cit = rows[order].astype(np.int32)
    rpt = np.zeros(n + 1, dtype=np.int64)
    np.add.at(rpt, ci + 1, 1)
    return np.cumsum(rpt).astype(np.int32), cit, order.astype(np.int32)


class Gnn:
    """Handle on the device network (LLJBASH_Gnn*Ex)."""

    def __init__(self, params: np.ndarray | None = None):
        self.hdl = fn("GnnCreateEx")()
        if not self.hdl:
            raise RuntimeError("GnnCreateEx failed")
        self.count = int(fn("GnnParameterCountEx")())
        if params is not None:
            self.set_params(params)

    def params(self) -> np.ndarray:
        out = np.empty(self.count, dtype=np.float32)
        if fn("GnnGetParametersEx")(self.hdl, _ptr(out, C.c_float)) != 0:
            raise RuntimeError("GnnGetParametersEx failed")
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.count
        if fn("GnnSetParametersEx")(self.hdl, _ptr(p, C.c_float)) != 0:
            raise RuntimeError("GnnSetParametersEx failed")

    def predict(self, n, d_rp, d_ci, d_va, d_omega) -> None:
        m = _dev_csr(n, d_rp, d_ci, d_va)
        if fn("GnnPredictEx")(self.hdl, C.byref(m), C.c_void_p(d_omega)) != 0:
            raise RuntimeError("GnnPredictEx failed")

    def gradients(self, n, d_rp, d_ci, d_va, d_rpt, d_cit, d_tmap, d_target):
        a = _dev_csr(n, d_rp, d_ci, d_va)
        at = _dev_csr(n, d_rpt, d_cit)
        g = np.empty(self.count, dtype=np.float32)
        loss = C.c_double(0)
        if fn("GnnGradientsEx")(self.hdl, C.byref(a), C.byref(at), C.c_void_p(d_tmap), C.c_void_p(d_target),
                                _ptr(g, C.c_float), C.byref(loss)) != 0:
            raise RuntimeError("GnnGradientsEx failed")
        return loss.value, g

    def train_step(self, n, d_rp, d_ci, d_va, d_rpt, d_cit, d_tmap, d_target, lr: float) -> float:
        a = _dev_csr(n, d_rp, d_ci, d_va)
        at = _dev_csr(n, d_rpt, d_cit)
        loss = C.c_double(0)
        if fn("GnnTrainStepEx")(self.hdl, C.byref(a), C.byref(at), C.c_void_p(d_tmap), C.c_void_p(d_target),
                                float(lr), C.byref(loss)) != 0:
            raise RuntimeError("GnnTrainStepEx failed")
        return loss.value

    def save(self, path: Path = WEIGHTS, meta: dict | None = None) -> None:
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps({"params": [float(x) for x in self.params()], "meta": meta or {}}))

    @classmethod
    def load(cls, path: Path = WEIGHTS) -> "Gnn":
        d = json.loads(Path(path).read_text())
        return cls(np.array(d["params"], dtype=np.float32))

    def close(self) -> None:
        if self.hdl:
            fn("GnnDestroyEx")(self.hdl)
            self.hdl = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceSystem:
    """A matrix and right-hand side resident on the device (torch tensors
    as allocator), with its symbolic ILU(lof) pattern set up."""

    def __init__(self, csr, rhs, lof: int, m: int, rel_tol: float = 1e-8, maxits: int = 2000, bits: int = 64):
        import torch

        self.csr = csr
        self.n = csr.n
        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda()
                  for k, v in (("rp", csr.row_ptr), ("ci", csr.col_idx), ("va", csr.value), ("b", rhs))}
        self.x = torch.zeros(self.n, dtype=torch.float64, device="cuda")
        self.omega = torch.zeros(self.n, dtype=torch.float32, device="cuda")
        self.ilu = IluSolver(0)
        self.ilu.SetFactorPrecision(bits)
        if self.ilu.SymbolicDevice(self.n, self.ptr("rp"), self.ptr("ci"), lof) < 0:
            raise RuntimeError("symbolic failed")
        self.g = PreconditionedGmres()
        self.g.SetIluPreconditioner(self.ilu)
        p = self.g.param()
        p.tolerance = rel_tol * float(np.linalg.norm(rhs))
        p.krylov_subspace_dimension = m
        p.max_iterations = maxits
        self.bnorm = float(np.linalg.norm(rhs))

    def ptr(self, k):
        return self.t[k].data_ptr()

    def solve_with_omega(self, omega_value: float | None):
        """Constant omega (None = plain ILU) -> (ok, its, true_rel_res)."""
        if omega_value is None or omega_value == 0.0:
            self.ilu.SetDiagonalCompensation(0)
        else:
            self.omega.fill_(float(omega_value))
            self.ilu.SetDiagonalCompensation(self.omega.data_ptr())
        rc = self.ilu.FactorizeDevice(self.ptr("va"), True)
        self.ilu.SetDiagonalCompensation(0)
        if rc != 0:
            return False, 0, float("inf")
        self.x.zero_()
        ok, its = self.g.SolveDevice(self.n, self.ptr("rp"), self.ptr("ci"), self.ptr("va"), self.ptr("b"),
                                     self.x.data_ptr())
        return ok, its, self.g.LastTrueResidual() / self.bnorm

    def solve_learned(self, gnn: Gnn, budget: int):
        """-> (ok, its, true_rel_res, used_fallback)."""
        self.x.zero_()
        it = C.c_int(0)
        fb = C.c_int(0)
        a = _dev_csr(self.n, self.ptr("rp"), self.ptr("ci"), self.ptr("va"))
        ok = fn("GmresSolveLearnedDeviceEx")(self.g.hdl, self.ilu.hdl, gnn.hdl, C.byref(a), C.c_void_p(self.ptr("b")),
                                             C.c_void_p(self.x.data_ptr()), int(budget), C.byref(it), C.byref(fb))
        return bool(ok), it.value, self.g.LastTrueResidual() / self.bnorm, bool(fb.value)


def best_omega(sys: DeviceSystem, grid=OMEGA_GRID):
    """Solver-generated label: the omega of `grid` with the fewest steps."""
    results = []
    for w in grid:
        ok, its, res = sys.solve_with_omega(w)
        results.append((its if ok and res <= 1.01e-8 else 10**9, w))
    results.sort()
    return results[0][1], dict((w, i) for i, w in results)
