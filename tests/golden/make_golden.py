"""Writes the symbolic-ILU(k) golden fixtures (tests/golden/iluk_*.npz).

Each fixture holds an input pattern A (row_ptr, col_idx, value) and, for
several levels of fill, the pattern of F, the nzmap (index in F of every
entry of A) and F's values at the level-0 slots.  They are computed by
`first_found_rows` below: a restatement of the reference's level-of-fill
rule (src/iluk.cpp:14-28) written independently of the C oracle -- a dict
per row plus a sorted key list, re-searched after every merge so that
columns inserted during the pivot walk are visited, as iterating a
std::map does, and a level recorded the first time a column is reached
(try_emplace).  The fixtures fix the oracle's and the GPU pass's outputs on
these inputs; the rule itself is pinned by the hand-derived known-answer
tests in tests/test_oracle_symbolic.py.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import bisect
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

OUT = Path(__file__).resolve().parent


def first_found_rows(rp, ci, lof):
    """Rows of F as {column: level} dicts under the first-found rule."""
    n = len(rp) - 1
    rows = []
    for i in range(n):
        lev = {int(c): 0 for c in ci[rp[i]:rp[i + 1]]}
        keys = sorted(lev)
        pos = 0
        while pos < len(keys):
            k = keys[pos]
            if k >= i:
                break
            if lev[k] < lof and k in rows[k]:
                for j in sorted(rows[k]):
                    if j < k:
                        continue
                    level = lev[k] + rows[k][j] + 1
                    if level <= lof and j not in lev:
                        lev[j] = level
                        bisect.insort(keys, j)
            pos = bisect.bisect_right(keys, k)
        rows.append(lev)
    return rows


def symbolic(rp, ci, va, lof):
    """(row_ptr, col_idx, nzmap, level-0 values) of F from first_found_rows."""
    rows = first_found_rows(rp, ci, lof)
    frp = np.zeros(len(rows) + 1, dtype=np.int32)
    frp[1:] = np.cumsum([len(r) for r in rows])
    fci = np.array([c for r in rows for c in sorted(r)], dtype=np.int32)
    nz = np.empty(len(ci), dtype=np.int32)
    for i in range(len(rows)):
        cols = fci[frp[i]:frp[i + 1]]
        for p in range(rp[i], rp[i + 1]):
            nz[p] = frp[i] + int(np.searchsorted(cols, ci[p]))
    order = np.argsort(nz)
    return frp, fci, nz, np.asarray(va, dtype=np.float64)[order]


def cases():
    from par_ilu0_cpp_amd import generators as gen
    from tests.conftest import dominant_values, random_pattern

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
    for name, rp, ci, va, lofs in cases():
        data = {"row_ptr": rp, "col_idx": ci, "value": va}
        for lof in lofs:
            frp, fci, nz, level0 = symbolic(rp, ci, va, lof)
            data[f"f{lof}_row_ptr"] = frp
            data[f"f{lof}_col_idx"] = fci
            data[f"f{lof}_nzmap"] = nz
            data[f"f{lof}_level0_values"] = level0
        data["lofs"] = np.array(lofs, dtype=np.int32)
        np.savez_compressed(OUT / f"iluk_{name}.npz", **data)
        print(name, {lof: int(data[f'f{lof}_row_ptr'][-1]) for lof in lofs})


if __name__ == "__main__":
    main()
