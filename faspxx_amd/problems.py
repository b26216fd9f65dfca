"""Synthetic sparse systems for the parity tests and the benchmark.

Every matrix is CSRx as FASP++ stores it (src/MAT.hxx:43-47): u32 ``row_ptr``
and ``col_ind``, f64 ``values``, u32 ``diag_ptr``; rows are lexicographic
(x fastest) and columns ascend within a row, so the diagonal is always
stored and ``diag_ptr`` is the first column-equal entry (src/MAT.cxx:917-928).

Random numbers come from a counter-based generator defined here so that the
same stream can be recomputed anywhere:

    key(seed, stream) = mix(mix(seed * G) + stream)
    u64(seed, stream, i) = mix(key + (i + 1) * G)      G = 0x9E3779B97F4A7C15
    mix = the SplitMix64 finaliser (shifts 30/27/31, multipliers
          0xBF58476D1CE4E5B9 / 0x94D049BB133111EB)
    U[0,1)  = (u64 >> 11) * 2**-53         U[-1,1) = 2 U[0,1) - 1
    N(0,1)  = Box-Muller on pairs (u_{2j}, u_{2j+1}):
              sqrt(-2 ln(1 - u_{2j})) * cos(2 pi u_{2j+1})
"""
from __future__ import annotations

import dataclasses

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix(z: np.ndarray) -> np.ndarray:
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def rng_u64(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """The generator's raw 64-bit outputs for counters ``idx``."""
    with np.errstate(over="ignore"):
        key = _mix(np.uint64(seed) * _G)
        key = _mix(key + np.uint64(stream))
        ctr = np.asarray(idx, dtype=np.uint64) + np.uint64(1)
        return _mix(key + ctr * _G)


def rng_uniform(seed: int, stream: int, n: int, start: int = 0) -> np.ndarray:
    """U[0,1) doubles for counters start .. start+n-1."""
    u = rng_u64(seed, stream, np.arange(start, start + n, dtype=np.uint64))
    return (u >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def rng_sym(seed: int, stream: int, n: int) -> np.ndarray:
    """U[-1,1) doubles."""
    return 2.0 * rng_uniform(seed, stream, n) - 1.0


def rng_normal(seed: int, stream: int, n: int) -> np.ndarray:
    """N(0,1) doubles by Box-Muller on consecutive uniform pairs."""
    m = (n + 1) // 2
    u = rng_uniform(seed, stream, 2 * m)
    r = np.sqrt(-2.0 * np.log1p(-u[0::2]))
    th = 2.0 * np.pi * u[1::2]
    return np.stack([r * np.cos(th), r * np.sin(th)], axis=1).reshape(-1)[:n]


@dataclasses.dataclass
class CSR:
    nrow: int
    mcol: int
    row_ptr: np.ndarray  # u32[nrow+1]
    col_ind: np.ndarray  # u32[nnz]
    values: np.ndarray  # f64[nnz]
    diag_ptr: np.ndarray  # u32[min(nrow, mcol)]

    @property
    def nnz(self) -> int:
        return int(self.col_ind.shape[0])

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix(
            (self.values, self.col_ind.astype(np.int64), self.row_ptr.astype(np.int64)),
            shape=(self.nrow, self.mcol),
        )

    def matvec(self, x: np.ndarray) -> np.ndarray:
        """Row sums in index order (used only to form right-hand sides)."""
        return self.to_scipy() @ x


def form_diag_ptr(row_ptr: np.ndarray, col_ind: np.ndarray, nrow: int) -> np.ndarray:
    """First entry of row j whose column is j (src/MAT.cxx:917-928); 0 if none."""
    rows = np.repeat(np.arange(nrow, dtype=np.int64), np.diff(row_ptr.astype(np.int64)))
    hit = np.nonzero(col_ind.astype(np.int64) == rows)[0]
    dp = np.zeros(nrow, dtype=np.uint32)
    # reversed assignment keeps the first hit in each row
    dp[rows[hit][::-1]] = hit[::-1].astype(np.uint32)
    return dp


def _from_dense_stencil(n: int, cols: np.ndarray, vals: np.ndarray, mask: np.ndarray) -> CSR:
    """(n, s) candidate columns/values, s in ascending-offset order."""
    counts = mask.sum(axis=1)
    rp = np.zeros(n + 1, dtype=np.uint32)
    np.cumsum(counts, out=rp[1:])
    ci = cols[mask].astype(np.uint32)
    va = np.ascontiguousarray(vals[mask], dtype=np.float64)
    dp = form_diag_ptr(rp, ci, n)
    return CSR(n, n, rp, ci, va, dp)


def poisson2d(m: int, diag: float = 4.0) -> CSR:
    """5-point Laplacian, m x m interior grid, Dirichlet-zero boundary."""
    n = m * m
    i = np.arange(n, dtype=np.int64)
    ix, iy = i % m, i // m
    offs = [-m, -1, 0, 1, m]
    ok = [iy > 0, ix > 0, np.ones(n, bool), ix < m - 1, iy < m - 1]
    cols = np.stack([i + o for o in offs], axis=1)
    vals = np.stack([np.full(n, -1.0)] * 2 + [np.full(n, diag)] + [np.full(n, -1.0)] * 2, 1)
    return _from_dense_stencil(n, cols, vals, np.stack(ok, axis=1))


def poisson3d(m: int) -> CSR:
    """7-point Laplacian, m^3 interior grid, Dirichlet-zero boundary."""
    n = m * m * m
    i = np.arange(n, dtype=np.int64)
    ix, iy, iz = i % m, (i // m) % m, i // (m * m)
    offs = [-m * m, -m, -1, 0, 1, m, m * m]
    ok = [iz > 0, iy > 0, ix > 0, np.ones(n, bool), ix < m - 1, iy < m - 1, iz < m - 1]
    cols = np.stack([i + o for o in offs], axis=1)
    vals = np.full((n, 7), -1.0)
    vals[:, 3] = 6.0
    return _from_dense_stencil(n, cols, vals, np.stack(ok, axis=1))


def convdiff2d(m: int, eps: float = 1e-2, w=(1.0, 0.5)) -> CSR:
    """-eps Lap u + w.grad u, first-order upwind (w > 0), h = 1/(m+1)."""
    n = m * m
    h = 1.0 / (m + 1)
    e = eps / (h * h)
    i = np.arange(n, dtype=np.int64)
    ix, iy = i % m, i // m
    offs = [-m, -1, 0, 1, m]
    cvals = [-e - w[1] / h, -e - w[0] / h, 4 * e + (w[0] + w[1]) / h, -e, -e]
    ok = [iy > 0, ix > 0, np.ones(n, bool), ix < m - 1, iy < m - 1]
    cols = np.stack([i + o for o in offs], axis=1)
    vals = np.stack([np.full(n, v) for v in cvals], axis=1)
    return _from_dense_stencil(n, cols, vals, np.stack(ok, axis=1))


def varcoef_fv(m: int, seed: int = 4):
    """-div(kappa grad u) on an m x m cell grid, kappa = exp(0.5 z), z ~ N(0,1).

    Interior faces use the harmonic mean 2 kP kN / (kP + kN); a boundary face
    (Dirichlet zero at half-cell distance) contributes 2 kP to the diagonal.
    Unscaled (h = 1).  Returns (CSR, kappa)."""
    n = m * m
    kap = np.exp(0.5 * rng_normal(seed, 0, n))
    i = np.arange(n, dtype=np.int64)
    ix, iy = i % m, i // m

    def face(j_ok, nb):
        kn = kap[np.where(j_ok, nb, i)]
        return np.where(j_ok, 2.0 * kap * kn / (kap + kn), 2.0 * kap)

    okS, okW, okE, okN = iy > 0, ix > 0, ix < m - 1, iy < m - 1
    fS, fW, fE, fN = face(okS, i - m), face(okW, i - 1), face(okE, i + 1), face(okN, i + m)
    diag = ((fS + fW) + fE) + fN
    cols = np.stack([i - m, i - 1, i, i + 1, i + m], axis=1)
    vals = np.stack([-fS, -fW, diag, -fE, -fN], axis=1)
    mask = np.stack([okS, okW, np.ones(n, bool), okE, okN], axis=1)
    return _from_dense_stencil(n, cols, vals, mask), kap


def powerlaw_spd(n: int = 1 << 20, seed: int = 3, kmin: int = 2, kmax: int = 2048,
                 expo: float = 2.2) -> CSR:
    """Irregular SPD M-matrix: row i of B gets k_i ~ P(k) ~ k^-expo on
    [kmin, kmax] distinct uniform columns != i with values -U(0,1); the matrix
    is B + B^T (coinciding entries summed) plus a diagonal equal to the row's
    sum of |off-diagonal| + 1.

    Streams: 1 degrees, 2 columns, 3 values, 5+r column redraws (round r)."""
    ks = np.arange(kmin, kmax + 1, dtype=np.float64)
    pk = ks ** (-expo)
    cdf = np.cumsum(pk) / pk.sum()
    k = np.searchsorted(cdf, rng_uniform(seed, 1, n), side="right") + kmin
    k = np.minimum(k, kmax).astype(np.int64)
    rows = np.repeat(np.arange(n, dtype=np.int64), k)
    m = rows.shape[0]

    def draw(stream, cnt, start=0):
        c = np.floor(rng_uniform(seed, stream, cnt, start) * (n - 1)).astype(np.int64)
        return c

    cols = draw(2, m)
    cols += cols >= rows  # skip the diagonal
    rnd = 0
    while True:
        key = rows * n + cols
        order = np.argsort(key, kind="stable")
        ks_sorted = key[order]
        dup = np.zeros(m, dtype=bool)
        dup[order[1:]] = ks_sorted[1:] == ks_sorted[:-1]
        nd = int(dup.sum())
        if nd == 0:
            break
        nc = draw(5 + rnd, nd)
        nc += nc >= rows[dup]
        cols[dup] = nc
        rnd += 1
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    vals = -rng_uniform(seed, 3, m)
    # S = B + B^T, coinciding entries summed
    r2 = np.concatenate([rows, cols])
    c2 = np.concatenate([cols, rows])
    v2 = np.concatenate([vals, vals])
    key = r2 * n + c2
    order = np.argsort(key, kind="stable")
    key, v2 = key[order], v2[order]
    first = np.ones(key.shape[0], dtype=bool)
    first[1:] = key[1:] != key[:-1]
    starts = np.nonzero(first)[0]
    uv = np.add.reduceat(v2, starts)
    uk = key[starts]
    ur, uc = uk // n, uk % n
    # diagonal = sum |offdiag| + 1 (row sums in column order)
    cnt = np.bincount(ur, minlength=n)
    offrp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(cnt, out=offrp[1:])
    absum = np.add.reduceat(np.abs(uv), offrp[:-1]) if uv.size else np.zeros(n)
    absum = np.where(cnt > 0, absum, 0.0)
    dval = absum + 1.0
    # merge the diagonal at its sorted position
    allr = np.concatenate([ur, np.arange(n, dtype=np.int64)])
    allc = np.concatenate([uc, np.arange(n, dtype=np.int64)])
    allv = np.concatenate([uv, dval])
    order = np.lexsort((allc, allr))
    allr, allc, allv = allr[order], allc[order], allv[order]
    rp = np.zeros(n + 1, dtype=np.uint32)
    np.cumsum(np.bincount(allr, minlength=n), out=rp[1:])
    ci = allc.astype(np.uint32)
    dp = form_diag_ptr(rp, ci, n)
    return CSR(n, n, rp, ci, np.ascontiguousarray(allv), dp)


@dataclasses.dataclass
class Problem:
    name: str
    A: CSR
    b: np.ndarray
    xstar: np.ndarray | None


def config(cid: int, scale: int | None = None) -> Problem:
    """The five benchmark systems.  ``scale`` shrinks the grid (side length, or
    n for the power-law matrix) for CPU-sized parity runs."""
    if cid == 1:
        m = scale or 256
        A = poisson2d(m)
        xs = rng_sym(0, 0, A.nrow)
        return Problem(f"poisson2d_{m}", A, A.matvec(xs), xs)
    if cid == 2:
        m = scale or 128
        A = poisson3d(m)
        xs = rng_sym(1, 0, A.nrow)
        return Problem(f"poisson3d_{m}", A, A.matvec(xs), xs)
    if cid == 3:
        m = scale or 1024
        A = convdiff2d(m)
        xs = rng_sym(2, 0, A.nrow)
        return Problem(f"convdiff2d_{m}", A, A.matvec(xs), xs)
    if cid == 4:
        n = scale or (1 << 20)
        A = powerlaw_spd(n)
        return Problem(f"powerlaw_{n}", A, rng_sym(3, 4, n), None)
    if cid == 5:
        m = scale or 512
        A, _ = varcoef_fv(m)
        return Problem(f"varcoef_fv_{m}", A, rng_sym(4, 1, A.nrow), rng_sym(4, 2, A.nrow))
    raise ValueError(f"unknown config {cid}")
