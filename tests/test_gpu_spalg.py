"""Device setup-time sparse algebra (fx_csr_lower_tri / upper_tri / add /
mult / shift_diag, faspxx_amd/csrc/fx_spalg.hip) against the oracle's
restatement of MAT::GetLowerTri, GetUpperTri, Add, Mult and Shift (itself pinned bit for bit
against the compiled reference by tests/test_oracle_spalg.py).  Bar: the
structure (rowPtr, colInd, diagPtr) and the values bit for bit; at full size
(3D Poisson 128^3 squared) against SciPy's product, pattern exactly and
values to rounding."""
import numpy as np
import pytest
import scipy.sparse as sp

from faspxx_amd import fx
from oracle import Matrix
from spalg_cases import (aggregation, fixture, from_scipy, missing_diag, random_with_diag,
                         to_scipy)

pytestmark = pytest.mark.gpu


def dev(ctx, M):
    return fx.Csr(ctx, M.nrow, M.mcol, M.rp, M.ci, M.va, M.dp)


def check(D, want, values=True):
    nrow, mcol, rp, ci, va, dp = want
    assert (D.nrow, D.mcol, D.nnz) == (nrow, mcol, len(ci))
    grp, gci, gva, gdp = D.download()
    assert np.array_equal(grp, rp)
    assert np.array_equal(gci, ci)
    if values:
        assert np.array_equal(gva, va)
    nd = min(nrow, mcol)
    assert np.array_equal(gdp[:nd], dp[:nd])


def operands(oracle, problems):
    out = [Matrix.of(problems(c, s).A) for c, s in ((1, 64), (2, 16), (3, 64), (5, 32), (4, 3000))]
    out += [fixture(oracle, "fdm_10X10.csr"), fixture(oracle, "fem_small.csr")]
    out += [random_with_diag(oracle, 2000, 0.005, 1), random_with_diag(oracle, 1500, 0.01, 2, True)]
    return out


def test_triangular_parts(ctx, oracle, problems):
    for M in operands(oracle, problems):
        D = dev(ctx, M)
        check(D.lower_tri(), oracle.lower_tri(M))
        check(D.upper_tri(), oracle.upper_tri(M))


def test_add(ctx, oracle, problems):
    for M in operands(oracle, problems):
        D = dev(ctx, M)
        check(fx.Csr.add(0.1, D, -2.5, D), oracle.mat_add(0.1, M, -2.5, M))
    A = random_with_diag(oracle, 3000, 0.003, 3)
    B = random_with_diag(oracle, 3000, 0.004, 4, True)
    for a, b in ((1.0, 1.0), (-0.7, 1.3), (3.0, 0.0)):
        check(fx.Csr.add(a, dev(ctx, A), b, dev(ctx, B)), oracle.mat_add(a, A, b, B))


def test_add_with_empty_operand(ctx, oracle):
    """The C ABI returns the other operand scaled (the reference leaves its
    target unchanged; the drop-in MAT::Add keeps that)."""
    A = random_with_diag(oracle, 500, 0.01, 5)
    E = fx.Csr(ctx, 500, 500, np.zeros(501, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
    got = fx.Csr.add(1.0, E, -0.3, dev(ctx, A))
    rp, ci, va, _ = got.download()
    assert np.array_equal(rp, A.rp) and np.array_equal(ci, A.ci)
    assert np.array_equal(va, A.va * -0.3)


def test_mult(ctx, oracle, problems):
    for M in operands(oracle, problems)[:-1] + [random_with_diag(oracle, 800, 0.01, 2, True)]:
        D = dev(ctx, M)
        check(fx.Csr.mult(D, D), oracle.mat_mult(M, M))
    A = random_with_diag(oracle, 2000, 0.004, 6, True)
    B = random_with_diag(oracle, 2000, 0.003, 7, True)
    check(fx.Csr.mult(dev(ctx, A), dev(ctx, B)), oracle.mat_mult(A, B))


def test_galerkin_triple_product(ctx, oracle, problems):
    """R A P with an aggregation P (n x n/4) and R = P^T: rectangular
    operands, each step against the oracle."""
    A = Matrix.of(problems(1, 64).A)
    P = aggregation(oracle, A.nrow, 4, 8)
    R = from_scipy(oracle, to_scipy(P).T)
    dA, dP, dR = dev(ctx, A), dev(ctx, P), dev(ctx, R)
    AP = oracle.mat_mult(A, P)
    dAP = fx.Csr.mult(dA, dP)
    check(dAP, AP)
    APm = Matrix(AP[0], AP[1], AP[2], AP[3], AP[4], AP[5][:min(AP[0], AP[1])])
    check(fx.Csr.mult(dR, dAP), oracle.mat_mult(R, APm))


def test_structure_only_product(ctx, oracle):
    M = random_with_diag(oracle, 600, 0.01, 9, True)
    S = Matrix(M.nrow, M.mcol, M.rp, M.ci, None, M.dp)
    got = fx.Csr.mult(dev(ctx, S), dev(ctx, M))
    check(got, oracle.mat_mult(S, M), values=False)


def test_poisson3d_128_squared(ctx, problems):
    """A A on the full 3D Poisson 128^3 matrix (n = 2,097,152): the
    product's pattern (25 points per interior row) and values against SciPy."""
    pr = problems(2, None)
    D = fx.Csr.from_problem(ctx, pr.A)
    P = fx.Csr.mult(D, D)
    rp, ci, va, dp = P.download()
    S = pr.A.to_scipy().tocsr()
    W = (S @ S).tocsr()
    W.sort_indices()
    G = sp.csr_matrix((va, ci.astype(np.int64), rp.astype(np.int64)), shape=S.shape)
    G.sort_indices()
    assert G.nnz == W.nnz
    assert np.array_equal(G.indptr, W.indptr) and np.array_equal(G.indices, W.indices)
    assert np.max(np.abs(G.data - W.data)) <= 1e-13 * np.max(np.abs(W.data))


def test_shift(ctx, oracle, problems):
    """MAT::Shift on matrices with and without their diagonal entries
    (rows without one add to values[0]), square and rectangular, against
    the oracle bit for bit; twice in a row on the same matrix."""
    cases = [Matrix.of(problems(2, 16).A), random_with_diag(oracle, 2000, 0.005, 1),
             missing_diag(oracle, 3000, 3000, 0.002, 11),
             missing_diag(oracle, 2500, 2500, 0.003, 12, False),
             missing_diag(oracle, 3000, 1800, 0.003, 13), missing_diag(oracle, 1500, 2600, 0.003, 14)]
    for M in cases:
        D = dev(ctx, M)
        D.shift_diag(0.1)
        want = oracle.mat_shift(M, 0.1)
        assert np.array_equal(D.download()[2], want)
        D.shift_diag(-3.75)
        M2 = Matrix(M.nrow, M.mcol, M.rp, M.ci, want, M.dp)
        assert np.array_equal(D.download()[2], oracle.mat_shift(M2, -3.75))
