"""Seeded synthetic sparse systems for the five benchmark workloads.

The reference ships no generators: its drivers read MatrixMarket files
(test/benchmark.cpp:40, test/test.c:15).  These build the same kind of
matrices directly in CSR form (0-based, int32 indices, float64 values,
column-sorted rows that always hold the diagonal), vectorised with numpy.
Every generator is a pure function of its arguments and seed
(numpy.random.default_rng / PCG64).

Row numbering of grids: i = x + s*y (2D) and i = x + s*y + s*s*z (3D), x
fastest, Dirichlet boundaries (neighbours outside the grid are dropped).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Csr:
    """Host CSR triple plus its size; arrays are contiguous numpy arrays."""

    n: int
    row_ptr: np.ndarray  # int32[n+1]
    col_idx: np.ndarray  # int32[nnz]
    value: np.ndarray  # float64[nnz]

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])

    def matvec(self, x: np.ndarray) -> np.ndarray:
        rows = np.repeat(np.arange(self.n), np.diff(self.row_ptr))
        return np.bincount(rows, weights=self.value * x[self.col_idx], minlength=self.n)

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.value, self.col_idx, self.row_ptr), shape=(self.n, self.n))


def _assemble(n: int, rows: np.ndarray, cols: np.ndarray, vals: np.ndarray) -> Csr:
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(row_ptr, rows + 1, 1)
    row_ptr = np.cumsum(row_ptr)
    return Csr(n, row_ptr.astype(np.int32), cols.astype(np.int32), np.ascontiguousarray(vals, dtype=np.float64))


def _stencil(shape, coef_diag: np.ndarray, offsets) -> Csr:
    """Assemble a structured-grid operator.

    `offsets` is a list of (axis, step, coef_array) where coef_array has the
    grid shape and gives the coupling of each cell to its neighbour at
    +step along axis (entries leaving the grid are dropped)."""
    dims = len(shape)
    n = int(np.prod(shape))
    strides = [1, shape[0], shape[0] * shape[1]][:dims]
    idx = np.arange(n).reshape(shape[::-1]).T  # idx[x, y(, z)] = x + s*y + s*s*z
    rows = [idx.ravel()]
    cols = [idx.ravel()]
    vals = [coef_diag.ravel()]
    for axis, step, coef in offsets:
        sl_src = [slice(None)] * dims
        if step > 0:
            sl_src[axis] = slice(0, shape[axis] - step)
        else:
            sl_src[axis] = slice(-step, shape[axis])
        src = idx[tuple(sl_src)]
        rows.append(src.ravel())
        cols.append((src + step * strides[axis]).ravel())
        vals.append(coef[tuple(sl_src)].ravel())
    return _assemble(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def poisson2d(s: int = 64) -> Csr:
    """2D 5-point Poisson, unscaled stencil (4, -1), natural order.
    s=64: n=4096, nnz=20224 (the 64x64 ILU(0) GMRES(30) workload)."""
    shape = (s, s)
    diag = np.full(shape, 4.0)
    m1 = np.full(shape, -1.0)
    offs = [(0, -1, m1), (0, 1, m1), (1, -1, m1), (1, 1, m1)]
    return _stencil(shape, diag, offs)


def convdiff3d(s: int = 128, peclet: float = 50.0, velocity=(1.0, 0.5, 0.25)) -> Csr:
    """3D 7-point convection-diffusion on the unit cube, s^3 interior cells,
    h = 1/(s+1), diffusion eps = max|v| / Pe, first-order upwind convection,
    scaled by h^2:
      diag        6 eps + h (vx + vy + vz)
      west/south/down (upwind side)  -eps - h v_axis
      east/north/up                   -eps
    s=128: n=2,097,152, nnz=14,581,760 (= 7 s^3 - 6 s^2)."""
    shape = (s, s, s)
    h = 1.0 / (s + 1)
    eps = max(abs(v) for v in velocity) / peclet
    diag = np.full(shape, 6.0 * eps + h * float(sum(velocity)))
    offs = []
    for axis, v in enumerate(velocity):
        offs.append((axis, -1, np.full(shape, -eps - h * v)))
        offs.append((axis, 1, np.full(shape, -eps)))
    return _stencil(shape, diag, offs)


def random_banded(n: int = 1_000_000, per_row: int = 12, band: int = 5000, seed: int = 0) -> Csr:
    """Random nonsymmetric matrix: `per_row` distinct off-diagonal columns per
    row at offsets uniform in [-band, band] \\ {0} (redrawn while out of range
    or repeated), values uniform[-1, 1], diagonal = 1.05 * row abs-sum."""
    rng = np.random.default_rng(seed)
    k = min(per_row, n - 1)
    band = min(band, n - 1)
    rows = np.arange(n)
    offs = rng.integers(1, band + 1, size=(n, k)) * np.where(rng.random((n, k)) < 0.5, -1, 1)
    cols = rows[:, None] + offs
    while True:
        bad = (cols < 0) | (cols >= n)
        srt = np.sort(cols, axis=1)
        dup_sorted = np.zeros_like(bad)
        dup_sorted[:, 1:] = srt[:, 1:] == srt[:, :-1]
        if dup_sorted.any():
            # mark the later copy of each repeated column in original order
            order = np.argsort(cols, axis=1, kind="stable")
            dup = np.zeros_like(bad)
            np.put_along_axis(dup, order, dup_sorted, axis=1)
            bad |= dup
        if not bad.any():
            break
        r, c = np.nonzero(bad)
        o = rng.integers(1, band + 1, size=r.size) * np.where(rng.random(r.size) < 0.5, -1, 1)
        cols[r, c] = r + o
    vals = rng.uniform(-1.0, 1.0, size=(n, k))
    diag = 1.05 * np.abs(vals).sum(axis=1)
    all_rows = np.concatenate([rows, np.repeat(rows, k)])
    all_cols = np.concatenate([rows, cols.ravel()])
    all_vals = np.concatenate([diag, vals.ravel()])
    return _assemble(n, all_rows, all_cols, all_vals)


def _harmonic(a, b):
    return 2.0 * a * b / (a + b)


def diffusion2d(kappa: np.ndarray) -> Csr:
    """2D 5-point variable-coefficient diffusion, cell coefficients
    kappa[x, y]; interior faces use the harmonic mean of the two cells,
    boundary faces (Dirichlet) the cell's own coefficient."""
    sx, sy = kappa.shape
    diag = np.zeros_like(kappa)
    offs = []
    for axis in (0, 1):
        n_ax = kappa.shape[axis]
        lo = [slice(None)] * 2
        hi = [slice(None)] * 2
        lo[axis] = slice(0, n_ax - 1)
        hi[axis] = slice(1, n_ax)
        face = _harmonic(kappa[tuple(lo)], kappa[tuple(hi)])
        plus = np.zeros_like(kappa)
        minus = np.zeros_like(kappa)
        plus[tuple(lo)] = -face
        minus[tuple(hi)] = -face
        offs.append((axis, -1, minus))
        offs.append((axis, 1, plus))
        diag[tuple(lo)] += face
        diag[tuple(hi)] += face
        first = [slice(None)] * 2
        last = [slice(None)] * 2
        first[axis] = 0
        last[axis] = n_ax - 1
        diag[tuple(first)] += kappa[tuple(first)]
        diag[tuple(last)] += kappa[tuple(last)]
    return _stencil((sx, sy), diag, offs)


def diffusion3d(kappa: np.ndarray) -> Csr:
    """3D 7-point variable-coefficient diffusion (same face rule as 2D)."""
    diag = np.zeros_like(kappa)
    offs = []
    for axis in (0, 1, 2):
        n_ax = kappa.shape[axis]
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis] = slice(0, n_ax - 1)
        hi[axis] = slice(1, n_ax)
        face = _harmonic(kappa[tuple(lo)], kappa[tuple(hi)])
        plus = np.zeros_like(kappa)
        minus = np.zeros_like(kappa)
        plus[tuple(lo)] = -face
        minus[tuple(hi)] = -face
        offs.append((axis, -1, minus))
        offs.append((axis, 1, plus))
        diag[tuple(lo)] += face
        diag[tuple(hi)] += face
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[axis] = 0
        last[axis] = n_ax - 1
        diag[tuple(first)] += kappa[tuple(first)]
        diag[tuple(last)] += kappa[tuple(last)]
    return _stencil(kappa.shape, diag, offs)


def batched_diffusion2d(count: int = 512, lo: int = 45, hi: int = 256, seed: int = 0):
    """`count` 2D diffusion systems, side s = round(exp(U[ln lo, ln hi])),
    cell coefficients lognormal(0, 1).  Returns a list of Csr."""
    rng = np.random.default_rng(seed)
    sides = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), size=count))).astype(int)
    out = []
    for s in sides:
        kappa = rng.lognormal(0.0, 1.0, size=(int(s), int(s)))
        out.append(diffusion2d(kappa))
    return out


def gaussian_random_field(shape, corr_len: float = 8.0, seed: int = 0) -> np.ndarray:
    """Unit-variance stationary Gaussian field with a Gaussian covariance of
    correlation length `corr_len` cells, drawn by FFT on a periodic
    embedding twice the grid size and cropped."""
    rng = np.random.default_rng(seed)
    big = tuple(2 * d for d in shape)
    freqs = np.meshgrid(*[np.fft.fftfreq(d) for d in big], indexing="ij")
    k2 = sum(f * f for f in freqs)
    spec = np.exp(-2.0 * (np.pi * corr_len) ** 2 * k2)
    noise = rng.standard_normal(big)
    field = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.sqrt(spec)))
    field = field[tuple(slice(0, d) for d in shape)]
    field -= field.mean()
    sd = field.std()
    return field / (sd if sd > 0 else 1.0)


def grf_diffusion3d(s: int = 64, sigma: float = 1.5, corr_len: float = 8.0, seed: int = 0) -> Csr:
    """3D diffusion with kappa = exp(sigma * g), g a unit-variance GRF."""
    g = gaussian_random_field((s, s, s), corr_len, seed)
    return diffusion3d(np.exp(sigma * g))


def rhs_uniform(n: int, seed: int = 0, count: int | None = None) -> np.ndarray:
    """Right-hand side(s) uniform in [-1, 1]."""
    rng = np.random.default_rng(seed)
    shape = (n,) if count is None else (count, n)
    return rng.uniform(-1.0, 1.0, size=shape)
