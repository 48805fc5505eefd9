"""Host programs against the drop-in boundary, compiled with the system
compilers against include/ only and linked to the libraries, the way an
existing caller of the reference's libraries would be:
  * tests/c_abi/drop_in.c (gcc): the reference's C-driver flow (test/test.c:
    ILU(0) from CopyCsrMatrix, b = A * 1, x0 = 0, default GMRES parameters),
    compared with the oracle;
  * tests/c_abi/reference_layout.cpp (g++): the C++ classes keep the
    reference's object layout (static checks at compile time), and a solve
    whose preconditioner is set the way the reference header's inline code
    sets it (precon_ only) agrees with the device path.
The CPU tests check that both compile and link; the GPU tests run them."""
from __future__ import annotations

import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from par_ilu0_cpp_amd import generators as gen
from tests.test_benchmark_cli import write_mm

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "c_abi" / "drop_in.c"
LIB = ROOT / "par_ilu0_cpp_amd" / "lib"


def build(out: Path) -> Path:
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = out / "drop_in"
    cmd = [cc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-O2", f"-I{ROOT / 'include'}", str(SRC), "-o", str(exe),
           f"-L{LIB}", "-lpar-ilu0-gmres_c", "-lpar-ilu0-gmres", f"-Wl,-rpath,{LIB}", "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_program_compiles_and_links(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr


def fields(out: str) -> dict:
    d = {}
    for line in out.splitlines():
        k, *v = line.split()
        d[k] = v
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["poisson64", "convdiff12"])
def test_c_program_solves_like_the_reference_driver(oracle, tmp_path, case):
    a = gen.poisson2d(64) if case == "poisson64" else gen.convdiff3d(12)
    mtx = tmp_path / "a.mtx"
    write_mm(mtx, a)
    exe = build(tmp_path)
    r = subprocess.run([str(exe), str(mtx)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr)
    f = fields(r.stdout)
    assert f["setup"] == ["0"] and f["factorize"] == ["0"] and f["converged"] == ["1"]
    its = int(f["iterations"][0])
    # the reference's defaults: absolute tolerance 1e-8, GMRES(60), 200 steps
    b = a.matvec(np.ones(a.n))
    fac, dg, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, a.value)
    ok_o, its_o, x_o, _ = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, 1e-8, 60, 200,
                                       factors=(a.row_ptr, a.col_idx, dg, fac))
    assert ok_o and abs(its - its_o) <= 1, (its, its_o)
    assert float(f["max_error"][0]) < 1e-6
    assert all(float(t) >= 0 for t in f["timers"]) and float(f["timers"][3]) > 0


@pytest.mark.gpu
def test_c_program_mkl_mode(oracle, tmp_path):
    """The USE_MKL flow from C: dcsrilu0-style pivot check, two single
    triangular solves composing to the preconditioner, then the solve."""
    a = gen.convdiff3d(12)
    mtx = tmp_path / "a.mtx"
    write_mm(mtx, a)
    exe = build(tmp_path)
    r = subprocess.run([str(exe), str(mtx), "mkl"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr)
    f = fields(r.stdout)
    assert f["factorize"] == ["0"] and f["factor_status"] == ["0"] and f["converged"] == ["1"]
    b = a.matvec(np.ones(a.n))
    fac, dg, _ = oracle.ilu_factorize(a.row_ptr, a.col_idx, a.value)
    y = oracle.substitute(a.row_ptr, a.col_idx, dg, fac, b)
    got = [float(v) for v in f["trsv"]]
    want = [y[0], y[a.n // 2], y[-1], np.max(np.abs(y))]
    assert np.allclose(got, want, rtol=1e-12, atol=0), (got, want)
    # [[1, 1], [1, 1]] in the leading block (nonzero diagonal, so the
    # reader keeps the rows): u_22 = 1 - 1 * 1 = 0 stops the factorization
    # with dcsrilu0's -102
    rp = np.array([0, 2, 4] + list(range(5, 13)), dtype=np.int32)
    ci = np.array([0, 1, 0, 1] + list(range(2, 10)), dtype=np.int32)
    va = np.array([1.0, 1.0, 1.0, 1.0] + [2.0] * 8)
    z = tmp_path / "z.mtx"
    write_mm(z, gen.Csr(10, rp, ci, va))
    r = subprocess.run([str(exe), str(z), "mkl"], capture_output=True, text=True, timeout=300)
    f = fields(r.stdout)
    assert r.returncode == 4 and f["factorize"] == ["-1"] and f["factor_status"] == ["-102"], r.stdout


CPP = ROOT / "tests" / "c_abi" / "reference_layout.cpp"


def build_cpp(out: Path) -> Path:
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no C++ compiler")
    exe = out / "reference_layout"
    cmd = [cxx, "-std=c++17", "-Wall", "-Werror", "-O2", f"-I{ROOT / 'include'}", str(CPP), "-o", str(exe),
           f"-L{LIB}", "-lpar-ilu0-gmres", f"-Wl,-rpath,{LIB}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_cpp_layout_program_compiles(tmp_path):
    build_cpp(tmp_path)


@pytest.mark.gpu
def test_cpp_reference_layout_precon_paths_agree(tmp_path):
    exe = build_cpp(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout, r.stderr)
    f = fields(r.stdout)
    its = [int(v) for v in f["iterations"]]
    # the host-callback path applies the same sweeps to the same vectors
    assert its[0] == its[1] == its[2] and its[0] > 0, its
    assert float(f["max_diff"][0]) == 0.0
