"""The oracle's setup-time sparse algebra (orc_mat_tri / orc_mat_add /
orc_mat_mult / orc_mat_shift: MAT::GetLowerTri, GetUpperTri, Add, Mult,
Shift, src/MAT.cxx:290-365, 652-800) pinned on the CPU: bit for bit against the compiled reference
(oracle/_ref, where it builds) on stencil, fixture, random and rectangular
operands; against SciPy's sparse algebra (patterns exactly, values to
rounding); and the product's column order (first appearance) against a
brute-force restatement on small inputs."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import Matrix
from spalg_cases import (aggregation, fixture, from_scipy, missing_diag, random_with_diag,
                         to_scipy)


def same(a, b):
    assert a[0] == b[0] and a[1] == b[1]
    for x, y in zip(a[2:5], b[2:5]):
        if x is None or y is None:
            assert x is None and y is None
        else:
            assert np.array_equal(x, y)


def cases(oracle, problems):
    out = [Matrix.of(problems(c, s).A) for c, s in ((1, 16), (2, 6), (3, 16), (5, 12))]
    out += [fixture(oracle, "fdm_10X10.csr"), fixture(oracle, "fem_small.csr")]
    out += [random_with_diag(oracle, 300, 0.02, 1), random_with_diag(oracle, 200, 0.05, 2, True)]
    return out


def test_against_compiled_reference(oracle, reference, problems):
    for M in cases(oracle, problems):
        same(oracle.lower_tri(M), reference.tri(M, False))
        same(oracle.upper_tri(M), reference.tri(M, True))
        same(oracle.mat_mult(M, M), reference.mat_mult(M, M))
        same(oracle.mat_add(0.1, M, -2.5, M), reference.mat_add(0.1, M, -2.5, M))
    A = random_with_diag(oracle, 400, 0.01, 3)
    B = random_with_diag(oracle, 400, 0.015, 4)
    same(oracle.mat_add(1.0, A, 1.0, B), reference.mat_add(1.0, A, 1.0, B))
    same(oracle.mat_add(-0.7, A, 1.3, B), reference.mat_add(-0.7, A, 1.3, B))
    same(oracle.mat_mult(A, B), reference.mat_mult(A, B))
    P = aggregation(oracle, 400, 4, 5)
    R = from_scipy(oracle, to_scipy(P).T)
    AP = oracle.mat_mult(A, P)
    same(AP, reference.mat_mult(A, P))
    same(oracle.mat_mult(R, A), reference.mat_mult(R, A))


def test_against_scipy(oracle, problems):
    for M in cases(oracle, problems):
        S = to_scipy(M)
        rows = np.repeat(np.arange(M.nrow), np.diff(M.rp.astype(np.int64)))
        if np.all(np.diff(M.ci.astype(np.int64))[np.diff(rows) == 0] > 0):  # sorted rows
            for got, want, keep in ((oracle.lower_tri(M), sp.tril(S), M.ci <= rows),
                                    (oracle.upper_tri(M), sp.triu(S), M.ci >= rows)):
                G = sp.csr_matrix((got[4], got[3].astype(np.int64), got[2].astype(np.int64)),
                                  shape=got[:2])
                assert (G != want).nnz == 0 and G.nnz == int(keep.sum())  # explicit zeros kept
        got = oracle.mat_add(0.1, M, -2.5, M)
        G = sp.csr_matrix((got[4], got[3].astype(np.int64), got[2].astype(np.int64)),
                          shape=got[:2])
        assert np.all(np.diff(got[3].astype(np.int64))[np.diff(np.repeat(
            np.arange(got[0]), np.diff(got[2].astype(np.int64)))) == 0] > 0)  # rows ascending
        W = 0.1 * S - 2.5 * S
        assert abs(G - W).max() <= 1e-15 * abs(W).max()
        got = oracle.mat_mult(M, M)
        G = sp.csr_matrix((got[4], got[3].astype(np.int64), got[2].astype(np.int64)),
                          shape=got[:2])
        W = (S @ S).tocsr()
        W.eliminate_zeros()
        assert G.nnz >= W.nnz and abs(G - W).max() <= 1e-13 * abs(W).max()
        Gp = G.copy()
        Gp.data[:] = 1
        Wp = sp.csr_matrix((np.ones(len(S.indices)), S.indices, S.indptr), shape=S.shape)
        Wp = (Wp @ Wp).tocsr()
        assert (Gp != (Wp != 0).astype(float)).nnz == 0  # the structural product


def test_add_pattern_union_and_empty(oracle):
    A = random_with_diag(oracle, 120, 0.03, 6)
    B = random_with_diag(oracle, 120, 0.04, 7, True)
    got = oracle.mat_add(2.0, A, -1.0, B)
    G = sp.csr_matrix((got[4], got[3].astype(np.int64), got[2].astype(np.int64)), shape=got[:2])
    SA, SB = to_scipy(A), to_scipy(B)
    U = ((SA != 0).astype(int) + (SB != 0).astype(int)) != 0
    assert G.nnz == U.nnz
    assert abs(G - (2.0 * SA - SB)).max() <= 1e-15 * abs(2.0 * SA - SB).max()
    empty = Matrix(120, 120, np.zeros(121, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    assert oracle.mat_add(1.0, A, 1.0, empty) is None
    assert oracle.mat_add(1.0, empty, 1.0, B) is None


def brute_mult(L, R):
    rp, ci = [0], []
    for i in range(L.nrow):
        seen = []
        for k in range(L.rp[i], L.rp[i + 1]):
            r = L.ci[k]
            for l in range(R.rp[r], R.rp[r + 1]):
                if R.ci[l] not in seen:
                    seen.append(int(R.ci[l]))
        ci += seen
        rp.append(len(ci))
    return np.array(rp), np.array(ci)


def test_mult_first_appearance_order(oracle):
    L = random_with_diag(oracle, 40, 0.1, 8, True)
    R = random_with_diag(oracle, 40, 0.1, 9, True)
    got = oracle.mat_mult(L, R)
    rp, ci = brute_mult(L, R)
    assert np.array_equal(got[2], rp) and np.array_equal(got[3], ci)
    # not sorted in general
    assert any(np.any(np.diff(ci[rp[i]:rp[i + 1]].astype(int)) < 0) for i in range(40))
    # structure-only operands: structure-only product with the same pattern
    Ls = Matrix(L.nrow, L.mcol, L.rp, L.ci, None, L.dp)
    s = oracle.mat_mult(Ls, R)
    assert s[4] is None and np.array_equal(s[3], ci)


def shift_cases(oracle, problems):
    return [Matrix.of(problems(1, 16).A), random_with_diag(oracle, 300, 0.02, 1),
            missing_diag(oracle, 200, 200, 0.03, 11), missing_diag(oracle, 150, 150, 0.04, 12, False),
            missing_diag(oracle, 180, 120, 0.05, 13), missing_diag(oracle, 90, 140, 0.05, 14)]


def test_shift_against_compiled_reference(oracle, reference, problems):
    """Bit for bit, including rows without a diagonal entry (each adds to
    values[0]) and rectangular matrices."""
    for M in shift_cases(oracle, problems):
        for a in (0.1, -3.75, 1e-3):
            assert np.array_equal(oracle.mat_shift(M, a), reference.mat_shift(M, a))


def test_shift_closed_form(oracle, problems):
    """A full diagonal: exactly the diagonal entries change, each by a.
    Missing diagonals: values[0] takes one sequential add of a per row whose
    diagPtr is 0, and nothing else off the diagonal changes."""
    M = random_with_diag(oracle, 300, 0.02, 1)
    got = oracle.mat_shift(M, 0.25)
    want = M.va.copy()
    want[M.dp] += 0.25
    assert np.array_equal(got, want)
    for M in shift_cases(oracle, problems)[2:]:
        a = 0.1
        got = oracle.mat_shift(M, a)
        rows = np.repeat(np.arange(M.nrow), np.diff(M.rp.astype(np.int64)))
        isdiag = M.ci.astype(np.int64) == rows
        hits0 = int(np.sum(M.dp == 0))
        v0 = M.va[0]
        for _ in range(hits0):
            v0 = v0 + a
        assert got[0] == v0
        want = M.va.copy()
        want[isdiag] += a
        assert np.array_equal(got[1:], want[1:])
