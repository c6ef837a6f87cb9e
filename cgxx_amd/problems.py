"""Synthetic sparse SPD systems for the CG path, and a Matrix Market writer.

Every generator returns a :class:`CRS` whose rows list their entries in
ascending column order.  Written as a ``general`` Matrix Market file with the
entries in that order, the reference's reader (Matrix.cpp:13-111) and its
stable COO->CRS conversion (Matrix.cpp:203-228) rebuild exactly these arrays,
so in-memory and file-based runs see the same per-row summation order.

Workloads (sizes are the ones the solver is measured on):

* ``laplace2d(256)``   5-point Dirichlet Laplacian, values 4/-1, N = 65,536.
* ``laplace3d(128)``   7-point Dirichlet Laplacian, values 6/-1, N = 2,097,152.
* ``diffusion3d(256, sigma=1.0, corr=8, seed)`` 7-point variable-coefficient
  diffusion with log-normal cell coefficients, N = 16,777,216.
* ``zipf_spd(1_000_000, seed)`` heavy-tailed random SPD matrix.
* ``anisotropic_family(seed)`` one member of the anisotropic log-normal
  diffusion family on 64^3..160^3 grids.
"""
from __future__ import annotations

import dataclasses
import hashlib
import os

import numpy as np


@dataclasses.dataclass
class CRS:
    """Compressed row storage with int32 indices and float64 values."""

    n: int
    ptr: np.ndarray  # int32[n+1]
    index: np.ndarray  # int32[nnz]
    value: np.ndarray  # float64[nnz]
    rhs: np.ndarray | None = None  # float64[n]; None means k = A*1
    info: dict = dataclasses.field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.ptr).astype(np.int32)

    def diagonal(self) -> np.ndarray:
        rows = np.repeat(np.arange(self.n, dtype=np.int64), self.row_lengths())
        on = self.index == rows
        d = np.zeros(self.n)
        d[rows[on]] = self.value[on]
        return d

    def ones_rhs(self) -> np.ndarray:
        """k = A*1 with the row sums accumulated left to right (CG.cpp:256-261).

        Rows are summed in stored order, entry by entry, exactly as the
        reference does in COO order for a row-sorted file.
        """
        k = np.zeros(self.n)
        lens = self.row_lengths()
        width = int(lens.max()) if self.n else 0
        starts = self.ptr[:-1].astype(np.int64)
        for j in range(width):
            live = lens > j
            k[live] += self.value[starts[live] + j]
        return k

    def to_ell(self):
        """(width, length, index, data) in column-major ELL, slot j*n + i,
        as Matrix.cpp:339-361 builds it from CRS (padding slots zero)."""
        n = self.n
        lens = self.row_lengths().astype(np.int32)
        w = int(lens.max()) if n else 0
        idx = np.zeros(n * w, np.int32)
        dat = np.zeros(n * w, np.float64)
        starts = self.ptr[:-1].astype(np.int64)
        rows = np.arange(n, dtype=np.int64)
        for j in range(w):
            live = lens > j
            idx[j * n + rows[live]] = self.index[starts[live] + j]
            dat[j * n + rows[live]] = self.value[starts[live] + j]
        return w, lens, idx, dat

    def matvec(self, x: np.ndarray) -> np.ndarray:
        """Plain float64 product (any order); for residual checks only."""
        rows = np.repeat(np.arange(self.n, dtype=np.int64), self.row_lengths())
        return np.bincount(rows, weights=self.value * x[self.index], minlength=self.n)


def _assemble(n: int, cols: np.ndarray, vals: np.ndarray, live: np.ndarray) -> CRS:
    """Compress a dense (n, w) stencil table whose columns are ascending."""
    counts = live.sum(axis=1).astype(np.int64)
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    if ptr[-1] >= 2**31:
        raise ValueError("nnz exceeds int32 indexing")
    index = cols[live].astype(np.int32)
    value = vals[live].astype(np.float64)
    return CRS(n=n, ptr=ptr.astype(np.int32), index=index, value=value)


def laplace2d(side: int) -> CRS:
    """5-point Dirichlet Laplacian on side x side, row i = y*side + x."""
    n = side * side
    i = np.arange(n, dtype=np.int64)
    y, x = i // side, i % side
    cols = np.stack([i - side, i - 1, i, i + 1, i + side], axis=1)
    live = np.stack([y > 0, x > 0, np.ones(n, bool), x < side - 1, y < side - 1], axis=1)
    vals = np.where(cols == i[:, None], 4.0, -1.0)
    m = _assemble(n, cols, vals, live)
    m.info = {"kind": "laplace2d", "side": side}
    return m


def _stencil3d(side: int, diag: np.ndarray, faces: list[np.ndarray]) -> CRS:
    """7-point matrix from per-row diagonal and the six face couplings.

    ``faces`` = off-diagonal values towards -z, -y, -x, +x, +y, +z (positive
    numbers; stored negated), each of shape (n,), ignored on the boundary.
    """
    n = side**3
    s2 = side * side
    shape = (side, side, side)
    # Entries per row: 7 minus one per boundary face.
    cnt = np.full(shape, 7, dtype=np.int64)
    cnt[0] -= 1
    cnt[-1] -= 1
    cnt[:, 0] -= 1
    cnt[:, -1] -= 1
    cnt[:, :, 0] -= 1
    cnt[:, :, -1] -= 1
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(cnt.ravel(), out=ptr[1:])
    del cnt
    nnz = int(ptr[-1])
    if nnz >= 2**31:
        raise ValueError("nnz exceeds int32 indexing")
    index = np.empty(nnz, dtype=np.int32)
    value = np.empty(nnz, dtype=np.float64)
    pos = ptr[:-1].copy()  # next free slot of each row
    rows = np.arange(n, dtype=np.int32)

    def live_mask(axis: int, lower: bool) -> np.ndarray:
        m = np.ones(shape, dtype=bool)
        sl = [slice(None)] * 3
        sl[axis] = 0 if lower else side - 1
        m[tuple(sl)] = False
        return m.ravel()

    # Ascending column order: -z, -y, -x, diagonal, +x, +y, +z.
    dirs = [(-s2, live_mask(0, True), faces[0], -1.0), (-side, live_mask(1, True), faces[1], -1.0),
            (-1, live_mask(2, True), faces[2], -1.0), (0, None, diag, 1.0),
            (1, live_mask(2, False), faces[3], -1.0), (side, live_mask(1, False), faces[4], -1.0),
            (s2, live_mask(0, False), faces[5], -1.0)]
    for delta, live, vals, sign in dirs:
        if live is None:
            p = pos
            index[p] = rows
            value[p] = vals
            pos += 1
        else:
            p = pos[live]
            index[p] = rows[live] + np.int32(delta)
            value[p] = sign * vals[live]
            pos[live] += 1
    return CRS(n=n, ptr=ptr.astype(np.int32), index=index, value=value)


def laplace3d(side: int) -> CRS:
    """7-point Dirichlet Laplacian on side^3, values 6/-1."""
    n = side**3
    one = np.ones(n)
    m = _stencil3d(side, 6.0 * one, [one] * 6)
    m.info = {"kind": "laplace3d", "side": side}
    return m


def gaussian_field(side: int, corr: float, seed: int) -> np.ndarray:
    """Unit-variance Gaussian random field on side^3 cells.

    White noise smoothed by a Gaussian kernel of standard deviation
    corr/sqrt(2) cells (so the field's correlation is exp(-r^2 / (2 corr^2))),
    then renormalised to zero mean and unit variance.  Seeded.
    """
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    g = rng.standard_normal((side, side, side))
    g = gaussian_filter(g, sigma=corr / np.sqrt(2.0), mode="wrap")
    g -= g.mean()
    g /= g.std()
    return g


def _diffusion_from_kappa(kappa: np.ndarray, axis_weight=(1.0, 1.0, 1.0)) -> CRS:
    """Cell-centred 7-point diffusion matrix for coefficient ``kappa[z,y,x]``.

    Interior faces carry the harmonic mean 2 k1 k2 / (k1 + k2) times the
    axis weight; a Dirichlet boundary face carries the cell's own kappa times
    the axis weight.  The diagonal is the sum of the six face weights.
    """
    side = kappa.shape[0]
    n = side**3
    k = kappa.astype(np.float64)
    wz, wy, wx = axis_weight

    def harm(a, b):
        return 2.0 * a * b / (a + b)

    fz = np.zeros((side + 1, side, side))
    fz[1:-1] = harm(k[:-1], k[1:]) * wz
    fz[0] = k[0] * wz
    fz[-1] = k[-1] * wz
    fy = np.zeros((side, side + 1, side))
    fy[:, 1:-1] = harm(k[:, :-1], k[:, 1:]) * wy
    fy[:, 0] = k[:, 0] * wy
    fy[:, -1] = k[:, -1] * wy
    fx = np.zeros((side, side, side + 1))
    fx[:, :, 1:-1] = harm(k[:, :, :-1], k[:, :, 1:]) * wx
    fx[:, :, 0] = k[:, :, 0] * wx
    fx[:, :, -1] = k[:, :, -1] * wx
    mz, pz = fz[:-1].ravel(), fz[1:].ravel()
    my, py = fy[:, :-1].ravel(), fy[:, 1:].ravel()
    mx, px = fx[:, :, :-1].ravel(), fx[:, :, 1:].ravel()
    # Row sum of the six face weights, left to right in a fixed order.
    diag = ((((mz + my) + mx) + px) + py) + pz
    del fz, fy, fx
    return _stencil3d(side, diag, [mz, my, mx, px, py, pz])


def diffusion3d(side: int = 256, sigma: float = 1.0, corr: float = 8.0, seed: int = 20261202) -> CRS:
    """Variable-coefficient 7-point diffusion, kappa = exp(sigma * g)."""
    g = gaussian_field(side, corr, seed)
    m = _diffusion_from_kappa(np.exp(sigma * g))
    m.info = {"kind": "diffusion3d", "side": side, "sigma": sigma, "corr": corr, "seed": seed}
    return m


def anisotropic_family(seed: int, sides=(64, 80, 96, 112, 128, 144, 160)) -> CRS:
    """One member of the anisotropic log-normal diffusion family.

    Grid side drawn from ``sides``; sigma ~ U[0.5, 1.5]; correlation length
    ~ U[4, 16] cells; two axis-weight ratios log-uniform in [1, 1e3] (the x
    axis has weight 1).  Everything is drawn from ``seed``.
    """
    rng = np.random.default_rng(seed)
    side = int(rng.choice(sides))
    sigma = float(rng.uniform(0.5, 1.5))
    corr = float(rng.uniform(4.0, 16.0))
    ry, rz = 10.0 ** rng.uniform(0.0, 3.0, size=2)
    g = gaussian_field(side, corr, int(rng.integers(2**31)))
    m = _diffusion_from_kappa(np.exp(sigma * g), axis_weight=(float(rz), float(ry), 1.0))
    m.info = {
        "kind": "anisotropic",
        "seed": seed,
        "side": side,
        "sigma": sigma,
        "corr": corr,
        "ratio_y": float(ry),
        "ratio_z": float(rz),
    }
    return m


def splitmix_uniform(n: int, seed: int) -> np.ndarray:
    """Uniform doubles in [0, 1): entry i is SplitMix64's output for state
    `seed` after i + 1 steps (Steele, Lea & Flood, OOPSLA 2014), top 53 bits.
    The generator of the CG_RHS=random:<seed> right-hand side of the C++
    driver (k_i = 2 u_i - 1, cgxx_amd/host/CG.cpp), bit for bit."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def random_rhs(n: int, seed: int) -> np.ndarray:
    """The CG_RHS=random:<seed> right-hand side: k_i = 2 u_i - 1 in [-1, 1)."""
    return 2.0 * splitmix_uniform(n, seed) - 1.0


def zipf_spd(n: int = 1_000_000, alpha: float = 2.0, cap: int = 2000, seed: int = 20261202) -> CRS:
    """Heavy-tailed random SPD matrix with a symmetric pattern.

    Row i draws d_i ~ Zipf(alpha) capped at ``cap`` partners, uniformly among
    the other rows; the pattern is symmetrised and de-duplicated.  Each edge
    gets one value U(-1, 0) used at (i, j) and (j, i); the diagonal is
    1.05 * |off-diagonal row sum|, so A is strictly diagonally dominant with a
    positive diagonal, hence SPD.  The right-hand side is U(-1, 1).  The
    realised mean and maximum row lengths are stored in ``info``.
    """
    rng = np.random.default_rng(seed)
    d = np.minimum(rng.zipf(alpha, size=n), cap).astype(np.int64)
    src = np.repeat(np.arange(n, dtype=np.int64), d)
    dst = rng.integers(0, n - 1, size=src.size, dtype=np.int64)
    dst += dst >= src  # uniform over the other rows
    lo, hi = np.minimum(src, dst), np.maximum(src, dst)
    key = np.unique(lo * n + hi)
    lo, hi = key // n, key % n
    w = -rng.uniform(0.0, 1.0, size=key.size)
    rows = np.concatenate([lo, hi, np.arange(n, dtype=np.int64)])
    cols = np.concatenate([hi, lo, np.arange(n, dtype=np.int64)])
    vals = np.concatenate([w, w, np.zeros(n)])
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    counts = np.bincount(rows, minlength=n)
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    diag_pos = np.flatnonzero(rows == cols)
    offsum = np.bincount(rows, weights=np.abs(vals), minlength=n)
    vals[diag_pos] = 1.05 * offsum[rows[diag_pos]]
    # Isolated rows (no partner) would get a zero diagonal; give them 1.
    vals[diag_pos[vals[diag_pos] == 0.0]] = 1.0
    m = CRS(n=n, ptr=ptr.astype(np.int32), index=cols.astype(np.int32), value=vals)
    m.rhs = rng.uniform(-1.0, 1.0, size=n)
    lens = np.diff(ptr)
    m.info = {
        "kind": "zipf_spd",
        "n": n,
        "alpha": alpha,
        "cap": cap,
        "seed": seed,
        "mean_row": float(lens.mean()),
        "max_row": int(lens.max()),
        "mean_offdiag": float(lens.mean() - 1),
    }
    return m


def random_spd_dense(n: int, density: float, seed: int) -> CRS:
    """Small random SPD matrix (symmetric pattern, dominant diagonal)."""
    rng = np.random.default_rng(seed)
    a = np.where(rng.random((n, n)) < density, rng.uniform(-1.0, 0.0, (n, n)), 0.0)
    a = np.triu(a, 1)
    a = a + a.T
    np.fill_diagonal(a, 1.05 * np.abs(a).sum(axis=1) + 0.5)
    r, c = np.nonzero(a)
    counts = np.bincount(r, minlength=n)
    ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    m = CRS(n=n, ptr=ptr.astype(np.int32), index=c.astype(np.int32), value=a[r, c])
    m.info = {"kind": "random_spd_dense", "n": n, "density": density, "seed": seed}
    return m


def fmt_value(v: float) -> str:
    """Shortest decimal that reads back to the same double."""
    s = repr(float(v))
    return s[:-2] if s.endswith(".0") else s


def write_mm(path: str | os.PathLike, m: CRS, symmetric: bool = False) -> str:
    """Write ``m`` as Matrix Market coordinate real; returns the sha256.

    ``general``: every entry, rows ascending, entries in stored order.
    ``symmetric``: the lower triangle only, in the same order; the reader
    mirrors each off-diagonal entry right after reading it (Matrix.cpp:94-103),
    so the per-row order of a symmetric file differs from the general one.
    """
    rows = np.repeat(np.arange(m.n, dtype=np.int64), m.row_lengths())
    cols = m.index.astype(np.int64)
    vals = m.value
    if symmetric:
        keep = cols <= rows
        rows, cols, vals = rows[keep], cols[keep], vals[keep]
    h = hashlib.sha256()
    with open(path, "w") as f:
        head = "%%MatrixMarket matrix coordinate real {}\n{} {} {}\n".format(
            "symmetric" if symmetric else "general", m.n, m.n, rows.size
        )
        f.write(head)
        h.update(head.encode())
        chunk = 1 << 20
        for s in range(0, rows.size, chunk):
            e = min(rows.size, s + chunk)
            lines = "".join(
                f"{r + 1} {c + 1} {fmt_value(v)}\n"
                for r, c, v in zip(rows[s:e].tolist(), cols[s:e].tolist(), vals[s:e].tolist())
            )
            f.write(lines)
            h.update(lines.encode())
    return h.hexdigest()
