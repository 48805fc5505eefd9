"""ctypes view of the two drop-in shared libraries.

Loads ``lib/libpar-ilu0-gmres_c.so`` (which links ``lib/libpar-ilu0-gmres.so``)
built in-tree by ``csrc/Makefile`` and declares every exported symbol of
include/csr_matrix.h, include/iluk.h, include/ilu_solver_c.h,
include/gmres_c.h and include/par_ilu_ex.h.  There is no fallback: if the
libraries are missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"
CORE_NAME = "libpar-ilu0-gmres.so"
CABI_NAME = "libpar-ilu0-gmres_c.so"


class CsrMatrix(C.Structure):
    """struct LLJBASH_CsrMatrix (include/csr_matrix.h; reference include/csr_matrix.h:10-16)."""

    _fields_ = [
        ("size", C.c_int),
        ("max_nnz", C.c_int),
        ("row_ptr", C.POINTER(C.c_int)),
        ("col_idx", C.POINTER(C.c_int)),
        ("value", C.POINTER(C.c_double)),
    ]


class GmresParameters(C.Structure):
    """struct LLJBASH_GmresParameters (reference include/gmres_param.h:10-14)."""

    _fields_ = [
        ("tolerance", C.c_double),
        ("krylov_subspace_dimension", C.c_int),
        ("max_iterations", C.c_int),
    ]


class GmresStat(C.Structure):
    """struct LLJBASH_GmresStat (reference include/gmres_param.h:18-23)."""

    _fields_ = [
        ("total_init_time", C.c_double),
        ("total_fgmr_time", C.c_double),
        ("total_mv_time", C.c_double),
        ("total_precon_time", C.c_double),
    ]


P = C.POINTER
_vp = C.c_void_p
_int = C.c_int
_dbl = C.c_double

# (name without prefix, restype, argtypes)
SIGNATURES = [
    # include/csr_matrix.h
    ("SetupCsrMatrix", None, [P(CsrMatrix), _int, _int]),
    ("CopyCsrMatrix", None, [P(CsrMatrix), P(CsrMatrix)]),
    ("CopyCsrMatrixValues", None, [P(CsrMatrix), P(CsrMatrix)]),
    ("DestroyCsrMatrix", None, [P(CsrMatrix)]),
    ("ReadCsrMatrixMM1", None, [P(CsrMatrix), C.c_char_p]),
    ("CsrMatVec", None, [P(CsrMatrix), P(_dbl), P(_dbl)]),
    ("CsrAmdOrder", _int, [P(CsrMatrix), P(_int), P(_int)]),
    ("CsrNdOrder", _int, [P(CsrMatrix), P(_int), P(_int)]),
    # include/iluk.h
    ("IlukSymbolic", _int, [P(CsrMatrix), _int, P(CsrMatrix), P(_int)]),
    # include/ilu_solver_c.h
    ("IluSolverCreate", _vp, [_int]),
    ("IluSolverDestroy", None, [_vp]),
    ("IluSolverGetMatrix", P(CsrMatrix), [_vp]),
    ("IluSolverSetup", _int, [_vp]),
    ("IluSolverFactorize", _int, [_vp]),
    # include/gmres_c.h
    ("GmresCreate", _vp, []),
    ("GmresDestroy", None, [_vp]),
    ("GmresGetParameters", P(GmresParameters), [_vp]),
    ("GmresGetStat", P(GmresStat), [_vp]),
    ("GmresSetPreconditioner", None, [_vp, _vp]),
    ("GmresSolve", C.c_bool, [_vp, P(CsrMatrix), P(_dbl), P(_dbl), P(_int)]),
    # include/par_ilu_ex.h
    ("IluSolverSetFactorPrecisionEx", _int, [_vp, _int]),
    ("IluSolverSetSweepLanesEx", _int, [_vp, _int, _int]),
    ("IluSolverGetLevelsEx", _int, [_vp, P(_int), P(_int)]),
    ("IluSolverScheduleJsonEx", _int, [_vp, C.c_char_p, _int]),
    ("IluSolverSetWatchdogEx", _int, [_vp, _dbl]),
    ("IluSolverSetPivotPolicyEx", _int, [_vp, _int, _dbl, _dbl]),
    ("IluSolverFactorStatusEx", _int, [_vp]),
    ("IluSolverTrsvDeviceEx", _int, [_vp, _int, _dbl, _vp, _vp]),
    ("IluSolverTrsvEx", _int, [_vp, _int, _dbl, P(_dbl), P(_dbl)]),
    ("IluSolverAutotuneEx", _int, [_vp, _int, P(_dbl), _int]),
    ("IluSolverSetSweepModeEx", _int, [_vp, _int, _int, _int]),
    ("IluSolverSetGridShapeEx", _int, [_vp, _int, _int, _int]),
    ("IluSolverSetSweepTilesEx", _int, [_vp, _int, _int, _int, _int, _int]),
    ("IluSolverGetSweepConfigEx", _int, [_vp, P(_int)]),
    ("IluSolverSetDiagonalCompensationEx", _int, [_vp, _vp]),
    ("IluSolverSymbolicDeviceEx", _int, [_vp, P(CsrMatrix), _int]),
    ("IluSolverFactorizeDeviceEx", _int, [_vp, _vp, _int]),
    ("IluSolverSubstituteDeviceEx", _int, [_vp, _vp, _vp]),
    ("IluSolverSubstituteNoAllocEx", _int, [_vp, _vp, _vp]),
    ("IluSolverCopyFactorsEx", _int, [_vp, P(_dbl)]),
    ("IluSolverCopyFillLevelsEx", _int, [_vp, P(C.c_byte)]),
    ("IluSolverNonzerosEx", _int, [_vp]),
    ("IlukSymbolicLevelsEx", _int, [P(CsrMatrix), _int, P(CsrMatrix), P(_int), P(C.c_byte), C.c_long]),
    ("GmresSolveDeviceEx", C.c_bool, [_vp, P(CsrMatrix), _vp, _vp, P(_int)]),
    ("GmresLastTrueResidualEx", _dbl, [_vp]),
    ("GmresResidualHistoryEx", _int, [_vp, P(_dbl), _int]),
    ("DeviceAvailableEx", _int, []),
    ("GnnCreateEx", _vp, []),
    ("GnnDestroyEx", None, [_vp]),
    ("GnnParameterCountEx", _int, []),
    ("GnnSetParametersEx", _int, [_vp, P(C.c_float)]),
    ("GnnGetParametersEx", _int, [_vp, P(C.c_float)]),
    ("GnnPredictEx", _int, [_vp, P(CsrMatrix), _vp]),
    ("GnnGradientsEx", _int, [_vp, P(CsrMatrix), P(CsrMatrix), _vp, _vp, P(C.c_float), P(_dbl)]),
    ("GnnTrainStepEx", _int, [_vp, P(CsrMatrix), P(CsrMatrix), _vp, _vp, C.c_float, P(_dbl)]),
    ("GmresSolveLearnedDeviceEx", C.c_bool, [_vp, _vp, _vp, P(CsrMatrix), _vp, _vp, _int, P(_int), P(_int)]),
    ("BatchGmresCreateEx", _vp, []),
    ("BatchGmresDestroyEx", None, [_vp]),
    ("BatchGmresSetupEx", _int, [_vp, P(CsrMatrix), _int, _int, P(_int), _int]),
    ("BatchGmresSetFactorPrecisionEx", _int, [_vp, _int]),
    ("BatchGmresFactorizeEx", _int, [_vp, _vp]),
    ("BatchGmresSolveEx", _int, [_vp, P(CsrMatrix), _vp, _vp, _dbl, _int, _int, P(_int), P(_int), P(_dbl)]),
]

DATA_SYMBOLS = ["CSR_MATRIX_DEFAULT"]

_LIB = None


def lib_paths() -> tuple[Path, Path]:
    return LIB_DIR / CORE_NAME, LIB_DIR / CABI_NAME


def lib():
    """The loaded C ABI library with typed LLJBASH_* functions."""
    global _LIB
    if _LIB is not None:
        return _LIB
    core, cabi = lib_paths()
    if not cabi.exists() or not core.exists():
        raise RuntimeError(
            f"par_ilu0_cpp_amd: native libraries not built ({cabi}); run __graft_entry__.build() "
            "or `make -C par_ilu0_cpp_amd/csrc`"
        )
    C.CDLL(str(core), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    handle = C.CDLL(str(cabi), mode=os.RTLD_NOW | C.RTLD_GLOBAL)
    for name, res, args in SIGNATURES:
        fn = getattr(handle, "LLJBASH_" + name)
        fn.restype = res
        fn.argtypes = args
    _LIB = handle
    return handle


def fn(name: str):
    return getattr(lib(), "LLJBASH_" + name)
