"""The command-line benchmark (par_ilu0_cpp_amd/csrc/benchmark.cpp), which
keeps the options, phase protocol and summary lines of the reference's
test/benchmark.cpp:16-213.  CPU tests cover option handling; the GPU tests
run the reference's default flow (ILU(k) through IlukSymbolic, b = A * 1,
x0 = 0, absolute tolerance) and compare the iteration count with the
oracle's."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from par_ilu0_cpp_amd import generators as gen

BIN = Path(__file__).resolve().parent.parent / "par_ilu0_cpp_amd" / "lib" / "par-ilu0-benchmark"


def write_mm(path: Path, a) -> None:
    rows = np.repeat(np.arange(a.n), np.diff(a.row_ptr))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{a.n} {a.n} {a.nnz}\n")
        for i, j, v in zip(rows, a.col_idx, a.value):
            f.write(f"{i + 1} {j + 1} {float(v):.17g}\n")


def run(*args, timeout=300):
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built")
    return subprocess.run([str(BIN), *map(str, args)], capture_output=True, text=True, timeout=timeout)


def field(out: str, name: str) -> float:
    m = re.search(rf"^{re.escape(name)}:\s*(\S+)$", out, re.M)
    assert m, f"{name} missing from output:\n{out}"
    return float(m.group(1))


def test_help_lists_reference_options():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--mat", "--lof", "--amd", "--nd", "--eps", "--maxit", "--restart", "--threads", "--rep",
                "--warmup", "--onlyfact", "--onlysubs"):
        assert opt in r.stdout


def test_missing_matrix_is_an_error():
    r = run()
    assert r.returncode == 1 and "--mat" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("flag", ["--amd", "--nd"])
def test_ordered_run_matches_oracle(oracle, tmp_path, flag):
    """--amd / --nd permute A symmetrically before ILU(k) (reference
    test/benchmark.cpp:44-80); the oracle solves the same permuted system."""
    import par_ilu0_cpp_amd as P
    a = gen.convdiff3d(14)
    p = tmp_path / "a.mtx"
    write_mm(p, a)
    r = run("-m", p, "-k", 1, "-r", 30, "-i", 500, "-n", 2, "-w", flag)
    assert r.returncode == 0, r.stderr
    its = [int(v) for v in re.findall(r"^iter: (\d+)$", r.stdout, re.M)]
    assert len(its) == 3 and len(set(its)) == 1
    order = P.CsrAmdOrder if flag == "--amd" else P.CsrNdOrder
    rc, perm, ip = order(P.HostCsr(a.row_ptr, a.col_idx, a.value))
    assert rc == (0 if flag == "--amd" else 1)
    brp, bci, bva = P.permute_symmetric(a.row_ptr, a.col_idx, a.value, ip)
    b = oracle.spmv(brp, bci, bva, np.ones(a.n))
    o = oracle.iluk_symbolic(brp, bci, bva, 1)
    fac, dg, _ = oracle.ilu_factorize(o[0], o[1], o[3])
    ok_o, its_o, _, _ = oracle.gmres((brp, bci, bva), b, 1e-8, 30, 500, factors=(o[0], o[1], dg, fac))
    assert ok_o and abs(its[0] - its_o) <= 1
    assert field(r.stdout, "knnz") == o[0][-1]
    assert field(r.stdout, "max_err") < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("lof,restart", [(0, 30), (1, 20)])
def test_gmres_run_matches_oracle(oracle, tmp_path, lof, restart):
    a = gen.convdiff3d(14)
    p = tmp_path / "a.mtx"
    write_mm(p, a)
    r = run("-m", p, "-k", lof, "-r", restart, "-i", 500, "-n", 2, "-w")
    assert r.returncode == 0, r.stderr
    out = r.stdout
    its = [int(v) for v in re.findall(r"^iter: (\d+)$", out, re.M)]
    assert len(its) == 3 and len(set(its)) == 1  # one warm-up + two timed runs
    b = a.matvec(np.ones(a.n))
    o = oracle.iluk_symbolic(a.row_ptr, a.col_idx, a.value, lof)
    fac, dg, _ = oracle.ilu_factorize(o[0], o[1], o[3])
    ok_o, its_o, _, _ = oracle.gmres((a.row_ptr, a.col_idx, a.value), b, 1e-8, restart, 500,
                                     factors=(o[0], o[1], dg, fac))
    assert ok_o and abs(its[0] - its_o) <= 1
    assert field(out, "knnz") == o[0][-1]
    assert field(out, "max_err") < 1e-6
    assert field(out, "reps") == 2 and field(out, "k") == lof
    assert field(out, "total gmres time (s)") > 0 and field(out, "total precon time (s)") > 0


@pytest.mark.gpu
def test_onlyfact_and_onlysubs(tmp_path):
    p = tmp_path / "a.mtx"
    write_mm(p, gen.poisson2d(40))
    r = run("-m", p, "-k", 1, "--onlyfact", "-n", 3)
    assert r.returncode == 0, r.stderr
    assert field(r.stdout, "total factorize time (s)") > 0
    r = run("-m", p, "--onlysubs", "-n", 3, "--fp32", "--autotune")
    assert r.returncode == 0, r.stderr
    assert field(r.stdout, "total precon time (s)") > 0
