"""The oracle's MG<MAT> (oracle/fx_oracle.cpp orc_mg, a restatement of
src/MG.cxx:16-229) pinned by what the mathematics fixes, CPU only.

With Jacobi smoothers and coarse solvers doing one sweep each (SOLParams
maxIter 1, tolerances 0), the error of one cycle is multiplied by a
polynomial in S = I - w D^-1 A.  Level l restricts its residual with T,
solves A e = T r from zero with what lies below it (nu cycles of level l+1,
or on the last level one coarse sweep, error operator S) and adds T e, so
with F the error operator of what lies below, level l maps the error by
E_l = S (I - T (I - F) A^-1 T A) S.  Identity transfers give S^(2L+1) for
L levels and one cycle per level, S^18 for three levels and two cycles; a
random nonsymmetric transfer T catches a transposed or misplaced operand."""
import numpy as np
import pytest

from csrfile import read_csr
from oracle import JACOBI, Matrix

ONE = dict(maxIter=1, relTol=0.0, absTol=0.0)


def fdm(oracle):
    n, rp, ci, va = read_csr("fdm_10X10.csr")
    return Matrix(n, n, rp, ci, va, oracle.form_diagptr(n, rp, ci))


def dense(M):
    A = np.zeros((M.nrow, M.mcol))
    for i in range(M.nrow):
        for k in range(M.rp[i], M.rp[i + 1]):
            A[i, M.ci[k]] = M.va[k]
    return A


@pytest.mark.parametrize("nlev,nu,sweeps", [(1, 1, 3), (2, 1, 5), (3, 1, 7), (3, 2, 18),
                                            (4, 1, 9)])
def test_identity_levels_are_jacobi_sweeps(oracle, nlev, nu, sweeps):
    M = fdm(oracle)
    rng = np.random.default_rng(nlev * 10 + nu)
    b, x0 = rng.standard_normal(M.nrow), rng.standard_normal(M.nrow)
    for w in (1.0, 0.8):
        r = oracle.mg(M, b, x0, [(None, ONE, ONE)] * nlev, ncycles=nu, weight=w, maxIter=1,
                      relTol=0.0, absTol=0.0)
        j = oracle.solve(JACOBI, M, b, x0, weight=w, maxIter=sweeps, relTol=0.0, absTol=0.0)
        assert r["numIter"] == 1
        assert np.linalg.norm(r["x"] - j["x"]) <= 1e-13 * np.linalg.norm(j["x"])


def level_error(A, S, T, nlev, nu, lev=0):
    """E_l = S (I - T (I - F) A^-1 T A) S, F = E_{l+1}^nu or S on the last level."""
    n = A.shape[0]
    F = S if lev + 1 == nlev else np.linalg.matrix_power(level_error(A, S, T, nlev, nu, lev + 1),
                                                         nu)
    return S @ (np.eye(n) - T @ (np.eye(n) - F) @ np.linalg.solve(A, T @ A)) @ S


@pytest.mark.parametrize("nlev,nu", [(1, 1), (2, 1), (3, 2)])
def test_nonsymmetric_transfer(oracle, nlev, nu):
    M = fdm(oracle)
    n = M.nrow
    A = dense(M)
    rng = np.random.default_rng(3 + nlev)
    Td = np.eye(n) * 0.7 + (rng.random((n, n)) < 0.03) * rng.uniform(-0.2, 0.2, (n, n))
    rows, cols = np.nonzero(Td)
    rp = np.searchsorted(rows, np.arange(n + 1)).astype(np.uint32)
    T = Matrix(n, n, rp, cols.astype(np.uint32), Td[rows, cols],
               oracle.form_diagptr(n, rp, cols.astype(np.uint32)))
    S = np.eye(n) - np.diag(1.0 / np.diag(A)) @ A
    E = level_error(A, S, Td, nlev, nu)
    xs = rng.standard_normal(n)
    b = A @ xs
    x0 = rng.standard_normal(n)
    r = oracle.mg(M, b, x0, [(T, ONE, ONE)] * nlev, ncycles=nu, maxIter=1, relTol=0.0,
                  absTol=0.0)
    want = xs - E @ (xs - x0)
    assert np.linalg.norm(r["x"] - want) <= 1e-10 * np.linalg.norm(want)
    # the transposed transfer gives a different answer
    Et = level_error(A, S, Td.T, nlev, nu)
    assert np.linalg.norm(xs - Et @ (xs - x0) - want) > 1e-3 * np.linalg.norm(want)


def test_solver_loop_matches_jacobi_history(oracle):
    """As a solver, iteration k of three-level identity MG sits at Jacobi
    sweep 7k: the checked residual norms agree, and the loop stops at the
    first k whose residual meets relTol."""
    M = fdm(oracle)
    rng = np.random.default_rng(7)
    b = rng.standard_normal(M.nrow)
    x0 = np.zeros(M.nrow)
    r = oracle.mg(M, b, x0, [(None, ONE, ONE)] * 3, maxIter=100, relTol=1e-9, absTol=0.0,
                  hist_cap=100)
    j = oracle.solve(JACOBI, M, b, x0, maxIter=7 * 100, relTol=0.0, absTol=0.0, hist_cap=701)
    k = r["numIter"]
    assert 0 < k < 100
    h = r["hist"][:k]
    np.testing.assert_allclose(h, j["hist"][0:7 * k:7], rtol=1e-11, atol=1e-13 * h[0])
    assert h[-1] / h[0] >= 1e-9
    nxt = np.linalg.norm(b - dense(M) @ r["x"])
    assert nxt / h[0] < 1e-9 and abs(nxt - r["norm2"]) <= 1e-12 * h[0]


def test_testmg_sample_run(oracle):
    """example/TestMG.cxx with its sample options (-maxIter 200 -minIter 100),
    x0 = 1, b = 0, default smoother parameters (maxIter 1, relTol 1e-6, absTol
    1e-8): no check before iteration 100, where ||r|| is already below absTol,
    so the loop stops there."""
    M = fdm(oracle)
    r = oracle.mg(M, np.zeros(M.nrow), np.ones(M.nrow), [(None, None, None)] * 3, maxIter=200,
                  minIter=100)
    assert r["retcode"] == 0 and r["numIter"] == 100
    assert r["norm2"] < 1e-8 and np.all(np.abs(r["x"]) < 1e-8)


def test_no_levels_copies_b(oracle):
    M = fdm(oracle)
    b = np.arange(M.nrow, dtype=float)
    r = oracle.mg(M, b, np.zeros(M.nrow), [], maxIter=1, relTol=0.0, absTol=0.0)
    assert np.array_equal(r["x"], b)
