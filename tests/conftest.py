"""Shared fixtures: repo on sys.path, the `gpu` marker, small helpers."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def random_pattern(n: int, per_row: int, band: int, seed: int, with_diag: bool = True):
    """Random sorted CSR (int32) with `per_row` off-diagonals in a band."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        lo, hi = max(0, i - band), min(n - 1, i + band)
        cand = np.arange(lo, hi + 1)
        cand = cand[cand != i]
        k = min(per_row, cand.size)
        cols = set(rng.choice(cand, size=k, replace=False).tolist()) if k else set()
        if with_diag:
            cols.add(i)
        rows.append(sorted(cols))
    rp = np.zeros(n + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([c for r in rows for c in r], dtype=np.int32)
    return rp, ci


def dominant_values(rp, ci, seed: int, dominance: float = 1.5):
    """Values with a diagonal `dominance` times the row's off-diagonal sum."""
    rng = np.random.default_rng(seed + 7)
    n = rp.size - 1
    va = rng.uniform(-1.0, 1.0, size=ci.size)
    for i in range(n):
        s, e = rp[i], rp[i + 1]
        off = ci[s:e] != i
        d = np.nonzero(ci[s:e] == i)[0]
        if d.size:
            va[s + d[0]] = dominance * (np.abs(va[s:e][off]).sum() + 1.0)
    return va


@pytest.fixture(scope="session")
def oracle():
    from oracle import py_oracle

    return py_oracle
