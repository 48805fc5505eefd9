"""Writes the symbolic-ILU(k) golden fixtures (tests/golden/iluk_*.npz).

Each fixture holds an input pattern A (row_ptr, col_idx, value) and the
output of the reference's own src/iluk.cpp for it -- compiled from the
reference sources into oracle/_ref/libref_iluk.so by `make -C oracle ref`
-- at one level of fill: the pattern of F, the nzmap, and F's values at the
level-0 slots (the reference leaves fill slots uninitialised, so only the
level-0 slots are stored, as a mask plus values).

Run from the repository root where /root/reference is present:
    make -C oracle ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from par_ilu0_cpp_amd import HostCsr  # noqa: E402
from par_ilu0_cpp_amd import generators as gen  # noqa: E402
from par_ilu0_cpp_amd._lib import CsrMatrix  # noqa: E402
from tests.conftest import dominant_values, random_pattern  # noqa: E402

OUT = Path(__file__).resolve().parent


def ref_lib():
    lib = C.CDLL(str(ROOT / "oracle" / "_ref" / "libref_iluk.so"))
    f = lib.LLJBASH_IlukSymbolic
    f.argtypes = [C.POINTER(CsrMatrix), C.c_int, C.POINTER(CsrMatrix), C.POINTER(C.c_int)]
    f.restype = C.c_int
    return f


def run_ref(fn, rp, ci, va, lof):
    a = HostCsr(rp, ci, va)
    f = HostCsr()
    nz = np.zeros(max(ci.size, 1), dtype=np.int32)
    k = fn(C.byref(a.mat), int(lof), C.byref(f.mat), nz.ctypes.data_as(C.POINTER(C.c_int)))
    frp, fci, fva = f.arrays()
    assert k == fci.size
    return frp, fci, fva, nz[: ci.size]


def cases():
    # hand-derived first-found case (n=7): see tests/test_oracle_symbolic.py
    rows = {0: [0, 5], 1: [1], 2: [0, 2], 3: [3, 5], 4: [2, 3, 4], 5: [5], 6: [4, 6]}
    rp = np.zeros(8, dtype=np.int32)
    rp[1:] = np.cumsum([len(rows[i]) for i in range(7)])
    ci = np.array([c for i in range(7) for c in rows[i]], dtype=np.int32)
    yield "kat7", rp, ci, np.arange(1.0, ci.size + 1), [0, 1, 2, 3]
    for seed, (n, per, band) in enumerate([(60, 3, 8), (200, 4, 25), (400, 6, 60), (150, 10, 149)]):
        rp, ci = random_pattern(n, per, band, seed=100 + seed)
        yield f"rand{n}_{per}_{band}", rp, ci, dominant_values(rp, ci, seed), [0, 1, 2, 3, 4]
    a = gen.poisson2d(12)
    yield "poisson12", a.row_ptr, a.col_idx, a.value, [0, 1, 2, 3]
    a = gen.convdiff3d(6)
    yield "convdiff6", a.row_ptr, a.col_idx, a.value, [0, 1, 2, 3]
    a = gen.random_banded(500, 12, 40, seed=3)
    yield "randband500", a.row_ptr, a.col_idx, a.value, [0, 1, 2]


def main():
    fn = ref_lib()
    for name, rp, ci, va, lofs in cases():
        data = {"row_ptr": rp, "col_idx": ci, "value": va}
        for lof in lofs:
            frp, fci, fva, nz = run_ref(fn, rp, ci, va, lof)
            level0 = np.zeros(fci.size, dtype=bool)
            level0[nz] = True
            data[f"f{lof}_row_ptr"] = frp
            data[f"f{lof}_col_idx"] = fci
            data[f"f{lof}_nzmap"] = nz
            data[f"f{lof}_level0_values"] = fva[level0]
        data["lofs"] = np.array(lofs, dtype=np.int32)
        np.savez_compressed(OUT / f"iluk_{name}.npz", **data)
        print(name, {lof: int(data[f'f{lof}_row_ptr'][-1]) for lof in lofs})


if __name__ == "__main__":
    main()
