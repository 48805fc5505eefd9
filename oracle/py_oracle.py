"""ctypes bindings of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
_LIBS: dict[str, C.CDLL] = {}

_I = C.POINTER(C.c_int)
_D = C.POINTER(C.c_double)


def _load(fast: bool = False) -> C.CDLL:
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name in _LIBS:
        return _LIBS[name]
    path = HERE / name
    if not path.exists():
        raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
    lib = C.CDLL(str(path))
    lib.orc_iluk_symbolic.restype = C.c_int
    lib.orc_iluk_symbolic.argtypes = [C.c_int, _I, _I, _D, C.c_int, C.POINTER(_I), C.POINTER(_I),
                                      C.POINTER(_I), C.POINTER(_D), _I]
    lib.orc_free.argtypes = [C.c_void_p]
    lib.orc_diag_ptr.argtypes = [C.c_int, _I, _I, _I]
    lib.orc_ilu_factorize.restype = C.c_int
    lib.orc_ilu_factorize.argtypes = [C.c_int, _I, _I, _I, _D]
    lib.orc_ilu_factorize_comp.restype = C.c_int
    lib.orc_ilu_factorize_comp.argtypes = [C.c_int, _I, _I, _I, _D, C.POINTER(C.c_float)]
    lib.orc_substitute.argtypes = [C.c_int, _I, _I, _I, _D, _D, _D]
    lib.orc_ilu0_guarded.restype = C.c_int
    lib.orc_ilu0_guarded.argtypes = [C.c_int, _I, _I, _I, _D, C.c_int, C.c_double, C.c_double]
    lib.orc_trsv_lanes.argtypes = [C.c_int, _I, _I, _I, _D, C.c_int, C.c_double, _D, _D, C.c_int]
    lib.orc_substitute_lanes.argtypes = [C.c_int, _I, _I, _I, _D, _D, _D, C.c_int]
    lib.orc_spmv.argtypes = [C.c_int, _I, _I, _D, C.c_double, _D, C.c_double, _D]
    lib.orc_gmres.restype = C.c_int
    lib.orc_gmres.argtypes = [C.c_int, _I, _I, _D, _I, _I, _I, _D, _D, _D, C.c_double, C.c_int, C.c_int,
                              _I, _D, C.c_int, _I]
    _LIBS[name] = lib
    return lib


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _pi(a):
    return a.ctypes.data_as(_I)


def _pd(a):
    return a.ctypes.data_as(_D)


def iluk_symbolic(row_ptr, col_idx, value, lof: int, fast: bool = False):
    """-> (frp, fci, flev, fval, nzmap) or None when nnz(F) would not be
    defined (repeated column)."""
    lib = _load(fast)
    rp, ci = _i(row_ptr), _i(col_idx)
    va = _d(value) if value is not None else None
    n = rp.size - 1
    frp, fci, flev = _I(), _I(), _I()
    fval = _D()
    nzmap = np.zeros(max(int(rp[-1]), 1), dtype=np.int32)
    nnz = lib.orc_iluk_symbolic(n, _pi(rp), _pi(ci), _pd(va) if va is not None else None, int(lof),
                                C.byref(frp), C.byref(fci), C.byref(flev), C.byref(fval), _pi(nzmap))
    if nnz < 0:
        return None
    out = (
        np.ctypeslib.as_array(frp, shape=(n + 1,)).copy(),
        np.ctypeslib.as_array(fci, shape=(max(nnz, 1),))[:nnz].copy(),
        np.ctypeslib.as_array(flev, shape=(max(nnz, 1),))[:nnz].copy(),
        np.ctypeslib.as_array(fval, shape=(max(nnz, 1),))[:nnz].copy(),
        nzmap[: int(rp[-1])].copy(),
    )
    for p in (frp, fci, flev, fval):
        lib.orc_free(C.cast(p, C.c_void_p))
    return out


def diag_ptr(row_ptr, col_idx):
    lib = _load()
    rp, ci = _i(row_ptr), _i(col_idx)
    n = rp.size - 1
    dg = np.empty(n, dtype=np.int32)
    lib.orc_diag_ptr(n, _pi(rp), _pi(ci), _pi(dg))
    return dg


def ilu_factorize(row_ptr, col_idx, value, fast: bool = False):
    """-> (factors, diag, zero_pivots)."""
    lib = _load(fast)
    rp, ci = _i(row_ptr), _i(col_idx)
    val = _d(value).copy()
    dg = diag_ptr(rp, ci)
    z = lib.orc_ilu_factorize(rp.size - 1, _pi(rp), _pi(ci), _pi(dg), _pd(val))
    return val, dg, z


def ilu_factorize_comp(row_ptr, col_idx, value, omega, fast: bool = False):
    """Factorization with per-row diagonal compensation -> (factors, diag)."""
    lib = _load(fast)
    rp, ci = _i(row_ptr), _i(col_idx)
    val = _d(value).copy()
    dg = diag_ptr(rp, ci)
    om = np.ascontiguousarray(omega, dtype=np.float32)
    lib.orc_ilu_factorize_comp(rp.size - 1, _pi(rp), _pi(ci), _pi(dg), _pd(val), om.ctypes.data_as(C.POINTER(C.c_float)))
    return val, dg


def ilu0_guarded(row_ptr, col_idx, value, policy: int, tiny: float = 1e-16, repl: float = 1e-10):
    """dcsrilu0-contract factorization (orc_ilu0_guarded) -> (factors, diag, status)."""
    lib = _load()
    rp, ci = _i(row_ptr), _i(col_idx)
    val = _d(value).copy()
    dg = diag_ptr(rp, ci)
    st = lib.orc_ilu0_guarded(rp.size - 1, _pi(rp), _pi(ci), _pi(dg), _pd(val), int(policy), float(tiny),
                              float(repl))
    return val, dg, st


def trsv_lanes(row_ptr, col_idx, diag, factors, x, upper: bool, alpha: float = 1.0, lanes: int = 1):
    """y = alpha L^-1 x (upper False, L unit lower) or alpha U^-1 x (upper
    True) with row sums in the g-lane order of the GPU sweeps."""
    lib = _load()
    rp, ci, dg, fv = _i(row_ptr), _i(col_idx), _i(diag), _d(factors)
    xx = _d(x)
    y = np.empty_like(xx)
    lib.orc_trsv_lanes(rp.size - 1, _pi(rp), _pi(ci), _pi(dg), _pd(fv), int(bool(upper)), float(alpha), _pd(xx),
                       _pd(y), int(lanes))
    return y


def substitute(row_ptr, col_idx, diag, factors, b, fast: bool = False):
    lib = _load(fast)
    rp, ci, dg, fv = _i(row_ptr), _i(col_idx), _i(diag), _d(factors)
    bb = _d(b)
    x = np.empty_like(bb)
    lib.orc_substitute(rp.size - 1, _pi(rp), _pi(ci), _pi(dg), _pd(fv), _pd(bb), _pd(x))
    return x


def substitute_lanes(row_ptr, col_idx, diag, factors, b, lanes: int):
    """orc_substitute with each row's sum split over `lanes` lanes in the
    GPU sweeps' order (bit-exact reference for any lanes per row)."""
    lib = _load()
    rp, ci, dg, fv = _i(row_ptr), _i(col_idx), _i(diag), _d(factors)
    bb = _d(b)
    x = np.empty_like(bb)
    lib.orc_substitute_lanes(rp.size - 1, _pi(rp), _pi(ci), _pi(dg), _pd(fv), _pd(bb), _pd(x), int(lanes))
    return x


def spmv(row_ptr, col_idx, value, x, alpha=1.0, beta=0.0, y=None):
    lib = _load()
    rp, ci, va, xx = _i(row_ptr), _i(col_idx), _d(value), _d(x)
    yy = np.zeros(rp.size - 1) if y is None else _d(y).copy()
    lib.orc_spmv(rp.size - 1, _pi(rp), _pi(ci), _pd(va), float(alpha), _pd(xx), float(beta), _pd(yy))
    return yy


def gmres(a, rhs, tol, m, maxits, factors=None, x0=None, fast: bool = False):
    """Restarted GMRES(m) of the reference.  a = (row_ptr, col_idx, value);
    factors = (row_ptr, col_idx, diag, value) of the ILU or None (M = I).
    -> (ok, its, x, history)."""
    lib = _load(fast)
    arp, aci, ava = _i(a[0]), _i(a[1]), _d(a[2])
    n = arp.size - 1
    b = _d(rhs)
    x = np.zeros(n) if x0 is None else _d(x0).copy()
    its = C.c_int(0)
    nh = C.c_int(0)
    hist = np.zeros(maxits + 2)
    if factors is not None:
        frp, fci, fdg, fva = _i(factors[0]), _i(factors[1]), _i(factors[2]), _d(factors[3])
        fp = (_pi(frp), _pi(fci), _pi(fdg), _pd(fva))
    else:
        fp = (None, None, None, None)
    ok = lib.orc_gmres(n, _pi(arp), _pi(aci), _pd(ava), *fp, _pd(b), _pd(x), float(tol), int(m), int(maxits),
                       C.byref(its), _pd(hist), hist.size, C.byref(nh))
    return bool(ok == 1), its.value, x, hist[: nh.value].copy()


_OMP = None


def _load_omp() -> C.CDLL:
    global _OMP
    if _OMP is None:
        lib = C.CDLL(str(HERE / "liboracle_omp.so"))
        lib.baseline_threads.restype = C.c_int
        lib.baseline_set_threads.argtypes = [C.c_int]
        lib.baseline_ilu_gmres.restype = C.c_int
        lib.baseline_ilu_gmres.argtypes = [C.c_int, _I, _I, _D, _I, _I, _I, _D, _D, _D, C.c_double, C.c_int, C.c_int,
                                           C.POINTER(C.c_int), _D, C.c_int]
        lib.baseline_subtree_info.restype = C.c_int
        lib.baseline_subtree_info.argtypes = [C.c_int, _I, _I, _I, _I, C.c_int]
        _OMP = lib
    return _OMP


def baseline_threads(k: int | None = None) -> int:
    """Threads of the OpenMP baseline (set them first when k is given)."""
    lib = _load_omp()
    if k is not None:
        lib.baseline_set_threads(int(k))
    return int(lib.baseline_threads())


def baseline_ilu_gmres(a, factors_pattern, rhs, tol, m, maxits, sched: int = 1, return_factors: bool = False):
    """OpenMP ILU + GMRES(m) (baseline_omp.c); sched 1: the reference's
    subtree / queue schedule, 0: level sets.
    factors_pattern = (frp, fci, fval before factorization).
    -> (ret, its, x, times[schedule, factorize, gmres])."""
    lib = _load_omp()
    arp, aci, aval = _i(a[0]), _i(a[1]), _d(a[2])
    frp, fci, fval = _i(factors_pattern[0]), _i(factors_pattern[1]), _d(factors_pattern[2]).copy()
    dg = diag_ptr(frp, fci)
    b = _d(rhs)
    x = np.zeros_like(b)
    its = C.c_int(0)
    times = np.zeros(3)
    ret = lib.baseline_ilu_gmres(arp.size - 1, _pi(arp), _pi(aci), _pd(aval), _pi(frp), _pi(fci), _pi(_i(dg)),
                                 _pd(fval), _pd(b), _pd(x), float(tol), int(m), int(maxits), C.byref(its), _pd(times),
                                 int(sched))
    if return_factors:
        return ret, its.value, x, times, fval
    return ret, its.value, x, times


def baseline_subtree_info(frp, fci, check: bool = True):
    """Shape of the subtree / queue schedule (baseline_omp.c) on the current
    threads: dict of queued rows and stripped levels per schedule, whether
    the sweeps queue in parallel, and (check) whether every row appears once
    and no subtree row depends on another thread's or a queued row."""
    lib = _load_omp()
    rp, ci = _i(frp), _i(fci)
    dg = _i(diag_ptr(rp, ci))
    info = np.zeros(7, dtype=np.int32)
    rc = lib.baseline_subtree_info(rp.size - 1, _pi(rp), _pi(ci), _pi(dg), _pi(info), int(check))
    if rc < 0:
        raise RuntimeError("baseline_subtree_info failed")
    return dict(ok=bool(rc) if check else None, queue=dict(fact=int(info[0]), l=int(info[1]), u=int(info[2])),
                stripped_levels=dict(fact=int(info[3]), l=int(info[4]), u=int(info[5])),
                parallel_sweep_queue=bool(info[6]))
