"""The drop-in boundary without a GPU: both libraries load, every entry point
declared in include/*.h is exported under its LLJBASH_ name, and the host-only
parts of the interface (CSR container, MatrixMarket reader, y = A x, handle
lifetime, parameter defaults, error codes) behave as the reference's
(include/csr_matrix.h, src/csr_matrix.cpp, include_c/*.h, src_c/*.cpp)."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
import sys
import textwrap
from pathlib import Path

import numpy as np
import pytest

import par_ilu0_cpp_amd as P
from par_ilu0_cpp_amd import _lib
from par_ilu0_cpp_amd import generators as gen

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        for m in re.finditer(r"#define\s+(\w+)\s+LLJBASH_DECL\((\w+)\)", h.read_text()):
            assert m.group(1) == m.group(2)
            names.add(m.group(1))
    return names


def test_libraries_load_and_export_every_declared_symbol():
    core, cabi = _lib.lib_paths()
    assert core.exists() and cabi.exists()
    P.lib()
    handles = [C.CDLL(str(core)), C.CDLL(str(cabi))]
    typedefs = {"CsrMatrix", "GmresParameters", "GmresStat", "IluSolverHdl", "GmresHdl", "BatchGmresHdl", "GnnHdl"}
    inline = {"GetCsrNonzeros"}  # static inline in the header, as in the reference
    missing = []
    for name in sorted(declared_symbols() - typedefs - inline):
        if not any(hasattr(h, "LLJBASH_" + name) for h in handles):
            missing.append(name)
    assert not missing, missing
    # the typed table in _lib covers every function
    assert {s[0] for s in _lib.SIGNATURES} | set(_lib.DATA_SYMBOLS) == declared_symbols() - typedefs - inline


def test_csr_default_constant():
    d = _lib.CsrMatrix.in_dll(C.CDLL(str(_lib.lib_paths()[0])), "LLJBASH_CSR_MATRIX_DEFAULT")
    assert d.size == -1 and d.max_nnz == 0
    assert not d.row_ptr and not d.col_idx and not d.value


def test_setup_copy_destroy():
    a = gen.poisson2d(9)
    h = P.HostCsr.from_csr(a)
    assert h.size == a.n and h.nnz == a.nnz and h.mat.max_nnz == a.nnz
    for p in (h.mat.row_ptr, h.mat.col_idx, h.mat.value):
        assert C.cast(p, C.c_void_p).value % 64 == 0
    rp, ci, va = h.arrays()
    assert np.array_equal(rp, a.row_ptr) and np.array_equal(ci, a.col_idx) and np.array_equal(va, a.value)
    c = P.HostCsr()
    P.fn("CopyCsrMatrix")(C.byref(c.mat), C.byref(h.mat))
    assert np.array_equal(c.arrays()[2], va)
    h2 = P.HostCsr(a.row_ptr, a.col_idx, np.zeros(a.nnz))
    P.fn("CopyCsrMatrixValues")(C.byref(h2.mat), C.byref(h.mat))
    assert np.array_equal(h2.arrays()[2], va)
    P.fn("DestroyCsrMatrix")(C.byref(c.mat))
    assert c.mat.size == -1 and c.mat.max_nnz == 0 and not c.mat.row_ptr


def test_csr_matvec():
    a = gen.convdiff3d(7)
    x = np.random.default_rng(0).uniform(-1, 1, a.n)
    y = P.HostCsr.from_csr(a).matvec(x)
    assert np.allclose(y, a.to_scipy() @ x, rtol=1e-14, atol=1e-15)


def write_mm(path: Path, n: int, entries):
    lines = ["%%MatrixMarket matrix coordinate real general", "% comment", f"{n} {n} {len(entries)}"]
    lines += [f"{i} {j} {v!r}" for i, j, v in entries]
    path.write_text("\n".join(lines) + "\n")


def test_read_mm_row_matching(tmp_path):
    # row 1 (1-based) has no diagonal: the only perfect matching on |a|>1e-12
    # swaps rows 1 and 2.
    entries = [(1, 2, 3.0), (2, 1, 4.0), (2, 2, 0.5), (3, 3, 2.0), (3, 1, -1.0)]
    f = tmp_path / "a.mtx"
    write_mm(f, 3, entries)
    h = P.HostCsr.read_mm1(str(f))
    rp, ci, va = h.arrays()
    assert h.mat.max_nnz == len(entries) + 3
    dense = np.zeros((3, 3))
    for i in range(3):
        dense[i, ci[rp[i]:rp[i + 1]]] = va[rp[i]:rp[i + 1]]
        assert list(ci[rp[i]:rp[i + 1]]) == sorted(ci[rp[i]:rp[i + 1]])
    assert np.array_equal(dense, np.array([[4.0, 0.5, 0], [0, 3.0, 0], [-1.0, 0, 2.0]]))


def test_read_mm_keeps_rows_with_diagonal(tmp_path):
    a = gen.poisson2d(5)
    ents = []
    for i in range(a.n):
        for p in range(a.row_ptr[i], a.row_ptr[i + 1]):
            ents.append((i + 1, int(a.col_idx[p]) + 1, float(a.value[p])))
    f = tmp_path / "p.mtx"
    write_mm(f, a.n, ents[::-1])
    rp, ci, va = P.HostCsr.read_mm1(str(f)).arrays()
    assert np.array_equal(rp, a.row_ptr) and np.array_equal(ci, a.col_idx) and np.array_equal(va, a.value)


def test_read_mm_nnz_mismatch_exits(tmp_path):
    f = tmp_path / "bad.mtx"
    f.write_text("3 3 4\n1 1 1.0\n2 2 1.0\n")
    code = textwrap.dedent(f"""
        import sys; sys.path.insert(0, {str(ROOT)!r})
        import par_ilu0_cpp_amd as P
        P.HostCsr.read_mm1({str(f)!r})
    """)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "nnz number error" in r.stdout


def test_ordering_return_conventions():
    """The reference's codes (src/csr_matrix.cpp:194, :216-224): CsrAmdOrder
    0 on success and -1 on failure, CsrNdOrder 1 on success and 0 on
    failure; ip may be null."""
    a = P.HostCsr.from_csr(gen.poisson2d(4))
    p = np.zeros(16, dtype=np.int32)
    ip = np.zeros(16, dtype=np.int32)
    pp = p.ctypes.data_as(C.POINTER(C.c_int))
    args = (C.byref(a.mat), pp, ip.ctypes.data_as(C.POINTER(C.c_int)))
    assert P.fn("CsrAmdOrder")(*args) == 0 and np.array_equal(ip[p], np.arange(16))
    assert P.fn("CsrNdOrder")(*args) == 1 and np.array_equal(ip[p], np.arange(16))
    assert P.fn("CsrAmdOrder")(C.byref(a.mat), pp, None) == 0
    assert P.fn("CsrAmdOrder")(C.byref(a.mat), None, None) == -1
    assert P.fn("CsrNdOrder")(C.byref(a.mat), None, None) == 0


def test_handles_and_defaults_without_device():
    g = P.PreconditionedGmres()
    prm = g.param()
    assert (prm.tolerance, prm.krylov_subspace_dimension, prm.max_iterations) == (1e-8, 60, 200)
    prm.max_iterations = 77
    assert g.param().max_iterations == 77  # mutable in place, as GmresGetParameters
    st = g.stat()
    assert (st.total_init_time, st.total_fgmr_time, st.total_mv_time, st.total_precon_time) == (0, 0, 0, 0)
    ilu = P.IluSolver(4)
    m = ilu.GetMatrix()
    assert m.size == -1 and not m.row_ptr
    g.SetIluPreconditioner(ilu)
    # an empty F is rejected before any device work
    assert ilu.SetupMatrix() == -1
    assert ilu.Factorize() == -1
    g.close()
    ilu.close()


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "_LIB", None)
    monkeypatch.setattr(_lib, "LIB_DIR", tmp_path)
    with pytest.raises(RuntimeError, match="not built"):
        _lib.lib()
