"""Operand matrices for the sparse-algebra parity tests (shared by the CPU
tests of the oracle and the GPU tests of the device kernels)."""
import numpy as np
import scipy.sparse as sp

from csrfile import read_csr
from oracle import Matrix


def from_scipy(oracle, S, values=True):
    S = sp.csr_matrix(S)
    S.sort_indices()
    n, m = S.shape
    rp = S.indptr.astype(np.uint32)
    ci = S.indices.astype(np.uint32)
    va = S.data.astype(np.float64) if values else None
    dp = oracle.form_diagptr(n, rp, ci) if n <= m else None
    return Matrix(n, m, rp, ci, va, dp)


def to_scipy(M):
    va = M.va if M.va is not None else np.ones(M.nnz)
    return sp.csr_matrix((va, M.ci.astype(np.int64), M.rp.astype(np.int64)),
                         shape=(M.nrow, M.mcol))


def random_with_diag(oracle, n, density, seed, unsorted=False):
    """A random square matrix with a full diagonal (so CSRx holds), entries
    U(-1, 1) off the diagonal and n on it; `unsorted` shuffles each row but
    the first (the reference's row sort fails on an entry that must move to
    the front of row 0)."""
    rng = np.random.default_rng(seed)
    S = sp.random(n, n, density=density, random_state=rng, format="csr")
    S.data = rng.uniform(-1, 1, S.nnz)
    S = (S + sp.eye(n) * n).tocsr()
    S.sort_indices()
    M = from_scipy(oracle, S)
    if unsorted:
        ci, va = M.ci.copy(), M.va.copy()
        for i in range(1, n):
            s, e = M.rp[i], M.rp[i + 1]
            perm = rng.permutation(e - s)
            ci[s:e] = ci[s:e][perm]
            va[s:e] = va[s:e][perm]
        M = Matrix(n, n, M.rp, ci, va, oracle.form_diagptr(n, M.rp, ci))
    return M


def aggregation(oracle, n, block, seed):
    """An n x n_c prolongation: each fine row in one aggregate of `block`
    consecutive rows, weight U(0.5, 1)."""
    rng = np.random.default_rng(seed)
    nc = (n + block - 1) // block
    rows = np.arange(n)
    P = sp.csr_matrix((rng.uniform(0.5, 1.0, n), (rows, rows // block)), shape=(n, nc))
    return from_scipy(oracle, P)


def fixture(oracle, name):
    n, rp, ci, va = read_csr(name)
    return Matrix(n, n, rp, ci, va, oracle.form_diagptr(n, rp, ci))


def missing_diag(oracle, n, m, density, seed, row0_diag=True):
    """A random n x m matrix (entries U(-1, 1), columns sorted) in which about
    a third of the rows 0..min(n, m)-1 carry their diagonal entry and the
    rest do not (nor do rows >= m); diagPtr formed as FormDiagPtr does, so
    the rows without a diagonal hold 0.  row0_diag: whether row 0 has its
    diagonal (then values[0] is that entry, which the other rows' shifts also
    land on)."""
    rng = np.random.default_rng(seed)
    S = sp.random(n, m, density=density, random_state=rng, format="lil")
    k = min(n, m)
    for i in range(k):
        S[i, i] = 0.0
    keep = rng.random(k) < 1 / 3
    keep[0] = row0_diag
    for i in np.flatnonzero(keep):
        S[i, i] = 2.0 + rng.random()
    S = S.tocsr()
    S.eliminate_zeros()
    S.sort_indices()
    S.data = np.where(S.data == 0, 0.5, S.data)
    rp = S.indptr.astype(np.uint32)
    ci = S.indices.astype(np.uint32)
    va = S.data.astype(np.float64)
    if row0_diag:
        assert ci[0] == 0
    return Matrix(n, m, rp, ci, va, oracle.form_diagptr(n, rp, ci))
