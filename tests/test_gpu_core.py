"""GPU parity of the MAT / VEC layer against the oracle, through the C ABI.

Bars: elementwise VEC ops and the one-thread-per-row SpMV / Residual are
bit-exact (same FMA placement as the reference's contracted build);
reductions (Dot, Norm2) agree to 1e-13 relative (tree vs 4-way order);
SpMV rows summed by lane groups (long rows) agree to 1e-12 relative; CSR
structure (diagPtr, transpose) is bit-exact."""
import json
import pathlib

import numpy as np
import pytest

from conftest import rel_err
from csrfile import read_csr
from faspxx_amd import fx
from oracle import Matrix

pytestmark = pytest.mark.gpu
GOLD = json.loads((pathlib.Path(__file__).parent / "golden" / "unit_tests.json").read_text())


def dev(ctx, a, dt=fx.F64):
    return fx.Vec.from_numpy(ctx, np.asarray(a), dt)


def test_spmv_unit_golden(ctx):
    g = GOLD["mat_apply"]
    A = fx.Csr(ctx, 4, 4, g["rowPtr"], g["colInd"], g["values"], g["diagPtr"])
    w = fx.Vec(ctx, 4)
    A.Apply(dev(ctx, np.full(4, g["x_value"])), w)
    assert np.all(np.abs(w.numpy() - g["expected"]) < g["tol"])


def test_vec_unit_golden(ctx):
    v = np.array(GOLD["vec_dot"]["v"], dtype=np.float64)
    dv = dev(ctx, v)
    assert dv.Dot(dv) == GOLD["vec_dot"]["dot_vv"]
    assert abs(dv.Norm2() - np.sqrt(GOLD["vec_dot"]["dot_vv"])) < 1e-10
    assert dv.NormInf() == 100.0


@pytest.mark.parametrize("n", [1, 5, 1000, 65537, 1 << 20])
def test_vec_elementwise_bitexact(ctx, oracle, n):
    rng = np.random.default_rng(n)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    for a, b in [(1.0, 1.0), (1.0, -0.37), (2.5, 1.0), (-0.7, 1.9)]:
        d = dev(ctx, y); d.AXPY(a, dev(ctx, x))
        assert np.array_equal(d.numpy(), oracle.axpy(a, x, y))
        d = dev(ctx, y); d.XPAY(a, dev(ctx, x))
        assert np.array_equal(d.numpy(), oracle.xpay(a, x, y))
        d = dev(ctx, y); d.AXPBY(a, b, dev(ctx, x))
        assert np.array_equal(d.numpy(), oracle.axpby(a, b, x, y))
        w = fx.Vec(ctx, n); w.WAXPBY(a, dev(ctx, x), b, dev(ctx, y))
        assert np.array_equal(w.numpy(), oracle.waxpby(a, x, b, y))
    d = dev(ctx, y); d.Scale(0.3)
    assert np.array_equal(d.numpy(), y * 0.3)
    d = dev(ctx, y); d.PointwiseMult(dev(ctx, x))
    assert np.array_equal(d.numpy(), y * x)
    d = dev(ctx, y); d.Reciprocal()
    assert np.array_equal(d.numpy(), 1.0 / y)
    d = dev(ctx, y); d += dev(ctx, x)
    assert np.array_equal(d.numpy(), y + x)


@pytest.mark.parametrize("n", [1, 3, 4099, 1 << 21])
def test_vec_reductions(ctx, oracle, n):
    rng = np.random.default_rng(n + 1)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    dx, dy = dev(ctx, x), dev(ctx, y)
    exact_dot = float(np.dot(x.astype(np.longdouble), y.astype(np.longdouble)))
    assert abs(dx.Dot(dy) - exact_dot) <= 1e-13 * np.dot(np.abs(x), np.abs(y))
    assert abs(dx.Dot(dy) - oracle.dot(x, y)) <= 1e-13 * np.dot(np.abs(x), np.abs(y))
    assert abs(dx.Norm2() - oracle.norm2(x)) <= 1e-14 * oracle.norm2(x)
    assert dx.NormInf() == oracle.norminf(x)
    # deterministic run to run
    assert dx.Dot(dy) == dx.Dot(dy)


@pytest.mark.parametrize("n", [1, 4099, 1 << 20])
def test_vec_fused_updates(ctx, oracle, n):
    """The fused one-pass forms: the updated vector equals the reference's
    AXPY / XPAY bit for bit, the reduction matches Norm2 / Dot of it."""
    rng = np.random.default_rng(n + 3)
    x, y, z = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    a = -0.37
    d = dev(ctx, y)
    nrm = d.AXPY_Norm2(a, dev(ctx, x))
    want = oracle.axpy(a, x, y)
    assert np.array_equal(d.numpy(), want)
    assert abs(nrm - oracle.norm2(want)) <= 1e-14 * oracle.norm2(want)
    d = dev(ctx, y)
    nrm = d.XPAY_Norm2(a, dev(ctx, x))
    want = oracle.xpay(a, x, y)
    assert np.array_equal(d.numpy(), want)
    assert abs(nrm - oracle.norm2(want)) <= 1e-14 * oracle.norm2(want)
    d = dev(ctx, y)
    dt = d.AXPY_Dot(a, dev(ctx, x), dev(ctx, z))
    want = oracle.axpy(a, x, y)
    assert np.array_equal(d.numpy(), want)
    assert abs(dt - oracle.dot(want, z)) <= 1e-13 * np.dot(np.abs(want), np.abs(z))


PARITY = [(1, None), (2, 32), (2, None), (3, 256), (4, 16384), (5, 64), (5, None)]


@pytest.mark.parametrize("cid,scale", PARITY)
def test_spmv_residual_parity(ctx, oracle, problems, cid, scale):
    pr = problems(cid, scale)
    M = Matrix.of(pr.A)
    A = fx.Csr.from_problem(ctx, pr.A)
    x = np.random.default_rng(cid).uniform(-1, 1, M.nrow)
    w = fx.Vec(ctx, M.nrow)
    A.Apply(dev(ctx, x), w)
    ref = oracle.spmv(M, x)
    r = fx.Vec(ctx, M.nrow)
    A.Residual(dev(ctx, pr.b), dev(ctx, x), r)
    rref = oracle.residual(M, pr.b, x)
    if cid == 4:  # rows in short blocks are summed by lane groups
        assert rel_err(w.numpy(), ref) < 1e-12
        assert rel_err(r.numpy(), rref) < 1e-12
    else:
        assert np.array_equal(w.numpy(), ref)
        assert np.array_equal(r.numpy(), rref)
    # x = 0 gives r = b exactly (MAT::Residual shortcut)
    A.Residual(dev(ctx, pr.b), dev(ctx, np.zeros(M.nrow)), r)
    assert np.array_equal(r.numpy(), pr.b)
    # structure download is bit-exact
    rp, ci, va, dp = A.download()
    assert np.array_equal(rp, pr.A.row_ptr) and np.array_equal(ci, pr.A.col_ind)
    assert np.array_equal(va, pr.A.values) and np.array_equal(dp, pr.A.diag_ptr)


@pytest.mark.parametrize("cid,scale,nbytes", [(1, None, 2), (2, 32, 2), (3, 256, 2), (5, 64, 2),
                                               (4, 16384, 2), (4, 200000, 4)])
def test_column_delta_format(ctx, problems, monkeypatch, cid, scale, nbytes):
    """16-bit column deltas (banded matrices) and u32 columns give the same
    bits; the power-law matrix's random columns keep u32 once n exceeds the
    16-bit reach (at n = 16384 every column is within 32K of every row)."""
    pr = problems(cid, scale)
    monkeypatch.setenv("FASPXX_FX_NOPAT", "1")  # row patterns off: deltas where they fit
    monkeypatch.setenv("FASPXX_FX_NOSELL", "1")  # no sliced or sorted rows where deltas do not fit
    monkeypatch.setenv("FASPXX_FX_NOSORT", "1")
    A = fx.Csr.from_problem(ctx, pr.A)
    monkeypatch.setenv("FASPXX_FX_COL32", "1")
    A32 = fx.Csr.from_problem(ctx, pr.A)
    monkeypatch.delenv("FASPXX_FX_COL32")
    monkeypatch.delenv("FASPXX_FX_NOPAT")
    monkeypatch.delenv("FASPXX_FX_NOSELL")
    monkeypatch.delenv("FASPXX_FX_NOSORT")
    assert A.index_bytes == nbytes and A32.index_bytes == 4
    x = dev(ctx, np.random.default_rng(7).uniform(-1, 1, pr.A.nrow))
    w, w32 = fx.Vec(ctx, pr.A.nrow), fx.Vec(ctx, pr.A.nrow)
    A.Apply(x, w)
    A32.Apply(x, w32)
    assert np.array_equal(w.numpy(), w32.numpy())


def test_column_delta_fallback(ctx, oracle, monkeypatch):
    """One column 39,999 rows away from its block start: the whole matrix
    keeps u32 columns, and the SpMV stays bit-exact."""
    from faspxx_amd import problems as P
    monkeypatch.setenv("FASPXX_FX_NOPAT", "1")
    B = P.poisson2d(200)
    n = B.nrow
    rp = B.row_ptr.astype(np.int64)
    ci = list(B.col_ind[: rp[1]]) + [n - 1] + list(B.col_ind[rp[1]:])
    va = list(B.values[: rp[1]]) + [-0.5] + list(B.values[rp[1]:])
    rp2 = np.concatenate([[0], rp[1:] + 1]).astype(np.uint32)
    ci, va = np.array(ci, np.uint32), np.array(va)
    dp = oracle.form_diagptr(n, rp2, ci)
    A = fx.Csr(ctx, n, n, rp2, ci, va, None)
    assert A.index_bytes == 4
    x = np.random.default_rng(8).uniform(-1, 1, n)
    w = fx.Vec(ctx, n)
    A.Apply(dev(ctx, x), w)
    assert np.array_equal(w.numpy(), oracle.spmv(Matrix(n, n, rp2, ci, va, dp), x))


def _formats(ctx, A64, monkeypatch):
    """The same matrix as row patterns (default), 16-bit deltas and u32."""
    A = fx.Csr.from_problem(ctx, A64)
    monkeypatch.setenv("FASPXX_FX_NOPAT", "1")
    Ad = fx.Csr.from_problem(ctx, A64)
    monkeypatch.setenv("FASPXX_FX_COL32", "1")
    Au = fx.Csr.from_problem(ctx, A64)
    monkeypatch.delenv("FASPXX_FX_COL32")
    monkeypatch.delenv("FASPXX_FX_NOPAT")
    return A, Ad, Au


@pytest.mark.parametrize("cid,scale", [(1, None), (2, 32), (3, 256), (5, 64), (2, None)])
def test_row_pattern_format(ctx, problems, oracle, monkeypatch, cid, scale):
    """Stencil matrices repeat a few column-offset patterns (27 for the 3D
    7-point stencil): they are stored as one pattern id per row, and
    Apply (f64 and f32) and Residual give the same bits as with 16-bit
    deltas and u32 columns, and as the oracle's SpMV."""
    pr = problems(cid, scale)
    A, Ad, Au = _formats(ctx, pr.A, monkeypatch)
    assert (A.index_bytes, Ad.index_bytes, Au.index_bytes) == (0, 2, 4)
    n = pr.A.nrow
    xh = np.random.default_rng(11).uniform(-1, 1, n)
    x = dev(ctx, xh)
    outs = []
    for M in (A, Ad, Au):
        w = fx.Vec(ctx, n)
        M.Apply(x, w)
        outs.append(w.numpy())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    if n <= 1 << 17:
        assert np.array_equal(outs[0], oracle.spmv(Matrix.of(pr.A), xh))
    b = dev(ctx, pr.b)
    res = []
    for M in (A, Au):
        r = fx.Vec(ctx, n)
        M.Residual(b, x, r)
        res.append(r.numpy())
    assert np.array_equal(res[0], res[1])
    f32 = []
    x32 = dev(ctx, xh.astype(np.float32), fx.F32)
    for M in (A, Au):
        M.make_f32()
        w = fx.Vec(ctx, n, fx.F32)
        M.Apply(x32, w)
        f32.append(w.numpy())
    assert np.array_equal(f32[0], f32[1])


def _two_entry_rows(n, K):
    """Row i < n - K - 1: columns i and i + 1 + (i mod K); later rows: i
    alone.  K + 1 distinct row patterns."""
    rp, ci, va = [0], [], []
    for i in range(n):
        cols = [i] if i >= n - K - 1 else [i, i + 1 + (i % K)]
        ci += cols
        va += [4.0] + [-1.0] * (len(cols) - 1)
        rp.append(len(ci))
    return (np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va))


@pytest.mark.parametrize("K,fmt", [(254, 0), (255, 2)])
def test_row_pattern_limits(ctx, oracle, K, fmt):
    """255 distinct row patterns fit the 1-byte ids, 256 do not (the
    matrix keeps 16-bit deltas); either way the SpMV is bit-exact."""
    n = 4000
    rp, ci, va = _two_entry_rows(n, K)
    A = fx.Csr(ctx, n, n, rp, ci, va, None)
    assert A.index_bytes == fmt
    x = np.random.default_rng(12).uniform(-1, 1, n)
    w = fx.Vec(ctx, n)
    A.Apply(dev(ctx, x), w)
    dp = oracle.form_diagptr(n, rp, ci)
    assert np.array_equal(w.numpy(), oracle.spmv(Matrix(n, n, rp, ci, va, dp), x))


def test_row_pattern_long_row(ctx, oracle):
    """A row of 33 entries (longer than the pattern format takes) keeps the
    matrix on 16-bit deltas; the SpMV stays bit-exact."""
    from faspxx_amd import problems as P
    B = P.poisson2d(64)
    n = B.nrow
    rp = B.row_ptr.astype(np.int64)
    extra = [c for c in range(40, 73)]  # row 5 gets 33 entries: columns 40..72
    ci = list(B.col_ind[: rp[5]]) + extra + list(B.col_ind[rp[6]:])
    va = list(B.values[: rp[5]]) + [-0.25] * len(extra) + list(B.values[rp[6]:])
    d = len(extra) - (rp[6] - rp[5])
    rp2 = np.concatenate([rp[:6], rp[6:] + d]).astype(np.uint32)
    ci, va = np.array(ci, np.uint32), np.array(va)
    A = fx.Csr(ctx, n, n, rp2, ci, va, None)
    assert A.index_bytes == 2
    x = np.random.default_rng(13).uniform(-1, 1, n)
    w = fx.Vec(ctx, n)
    A.Apply(dev(ctx, x), w)
    dp = oracle.form_diagptr(n, rp2, ci)
    assert np.array_equal(w.numpy(), oracle.spmv(Matrix(n, n, rp2, ci, va, dp), x))


def test_row_pattern_rectangular_empty_rows(ctx, oracle, monkeypatch):
    """Row patterns on a rectangular matrix (3000 x 3100) whose every 7th row
    is empty, with values and structure-only: the same bits as u32 columns
    and as the oracle; empty rows give 0; a structure-only row is the sum of
    x over its columns, left to right."""
    n, m = 3000, 3100
    rp, ci, va = [0], [], []
    for i in range(n):
        cols = [] if i % 7 == 3 else [c for c in (i - 2, i, i + 50, i + 100) if c >= 0]
        ci += cols
        va += [float(1 + i % 5)] * len(cols[:1]) + [-0.5] * max(len(cols) - 1, 0)
        rp.append(len(ci))
    rp, ci, va = np.array(rp, np.uint32), np.array(ci, np.uint32), np.array(va)
    x = np.random.default_rng(14).uniform(-1, 1, m)
    dp = oracle.form_diagptr(n, rp, ci)
    for vals in (va, None):
        A = fx.Csr(ctx, n, m, rp, ci, vals, None)
        monkeypatch.setenv("FASPXX_FX_COL32", "1")
        Au = fx.Csr(ctx, n, m, rp, ci, vals, None)
        monkeypatch.delenv("FASPXX_FX_COL32")
        assert A.index_bytes == 0 and Au.index_bytes == 4
        w, wu = fx.Vec(ctx, n), fx.Vec(ctx, n)
        A.Apply(dev(ctx, x), w)
        Au.Apply(dev(ctx, x), wu)
        got = w.numpy()
        assert np.array_equal(got, wu.numpy())
        assert np.all(got[3::7] == 0.0)
        if vals is not None:
            assert np.array_equal(got, oracle.spmv(Matrix(n, m, rp, ci, va, dp), x))
        else:
            ref = np.zeros(n)
            for i in range(n):
                t = 0.0
                for k in range(rp[i], rp[i + 1]):
                    t += x[ci[k]]
                ref[i] = t
            assert np.array_equal(got, ref)


def test_spmv_f32(ctx, problems):
    pr = problems(2, 32)
    A = fx.Csr.from_problem(ctx, pr.A)
    A.make_f32()
    x = np.random.default_rng(3).uniform(-1, 1, pr.A.nrow).astype(np.float32)
    w = fx.Vec(ctx, pr.A.nrow, fx.F32)
    A.Apply(dev(ctx, x, fx.F32), w)
    ref = pr.A.to_scipy() @ x.astype(np.float64)
    assert rel_err(w.numpy(), ref) < 1e-6


def test_diagptr_formed_on_device(ctx, problems):
    pr = problems(3, 64)
    A = fx.Csr(ctx, pr.A.nrow, pr.A.mcol, pr.A.row_ptr, pr.A.col_ind, pr.A.values, None)
    _, _, _, dp = A.download()
    assert np.array_equal(dp, pr.A.diag_ptr)
    d = fx.Vec(ctx, pr.A.nrow)
    A.GetDiag(d)
    assert np.array_equal(d.numpy(), pr.A.values[pr.A.diag_ptr])


@pytest.mark.parametrize("cid,scale", [(1, 64), (3, 100), (4, 20000), (5, 64)])
def test_transpose_bitexact(ctx, oracle, problems, cid, scale):
    pr = problems(cid, scale)
    M = Matrix.of(pr.A)
    A = fx.Csr.from_problem(ctx, pr.A)
    T = A.transpose()
    rp, ci, va, dp = T.download()
    orp, oci, ova, odp = oracle.transpose(M)
    assert np.array_equal(rp, orp) and np.array_equal(ci, oci)
    assert np.array_equal(va, ova) and np.array_equal(dp, odp)


def test_data_file_fixtures(ctx, oracle):
    for name in ("fdm_10X10.csr", "fem_small.csr"):
        n, rp, ci, va = read_csr(name)
        dp = oracle.form_diagptr(n, rp, ci)
        M = Matrix(n, n, rp, ci, va, dp)
        A = fx.Csr(ctx, n, n, rp, ci, va, None)
        x = np.random.default_rng(n).standard_normal(n)
        w = fx.Vec(ctx, n)
        A.Apply(dev(ctx, x), w)
        assert np.array_equal(w.numpy(), oracle.spmv(M, x))
        assert np.array_equal(A.download()[3], dp)


def test_empty_and_invalid(ctx):
    A = fx.Csr(ctx, 0, 0, [0], [], [], None)
    assert A.nrow == 0
    with pytest.raises(fx.FxError) as e:
        fx.Csr(ctx, 2, 2, [0, 1, 2], [0, 5], [1.0, 1.0], None)
    assert e.value.code == -17
    with pytest.raises(fx.FxError):
        fx.Csr(ctx, 2, 2, [0, 2, 1], [0, 1], [1.0, 1.0], None)
    B = fx.Csr(ctx, 2, 2, [0, 1, 2], [0, 1], [2.0, 3.0], None)
    with pytest.raises(fx.FxError) as e:
        B.Apply(fx.Vec(ctx, 3), fx.Vec(ctx, 2))
    assert e.value.code == -16


def _sliced_pair(ctx, Acsr, monkeypatch):
    """The matrix with sliced rows (forced) and with plain u32-column blocks."""
    monkeypatch.setenv("FASPXX_FX_SELL", "1")
    S = fx.Csr.from_problem(ctx, Acsr)
    monkeypatch.delenv("FASPXX_FX_SELL")
    monkeypatch.setenv("FASPXX_FX_COL32", "1")
    U = fx.Csr.from_problem(ctx, Acsr)
    monkeypatch.delenv("FASPXX_FX_COL32")
    return S, U


def _check_sliced(ctx, oracle, Acsr, S, U, seed):
    """Apply (f64, f32), Residual and the x = 0 shortcut: rows of at most 32
    entries (sliced, one lane each, left to right) equal the oracle bit for
    bit; longer rows (a wave each) agree to 1e-12 relative."""
    n, m = Acsr.nrow, Acsr.mcol
    M = Matrix.of(Acsr)
    lens = np.diff(Acsr.row_ptr.astype(np.int64))
    short = lens <= 32
    xh = np.random.default_rng(seed).uniform(-1, 1, m)
    w = fx.Vec(ctx, n)
    S.Apply(dev(ctx, xh), w)
    ref = oracle.spmv(M, xh)
    got = w.numpy()
    assert np.array_equal(got[short], ref[short])
    if not short.all():
        assert rel_err(got[~short], ref[~short]) < 1e-12
    if n == m:
        r = fx.Vec(ctx, n)
        S.Residual(dev(ctx, Acsr_b(Acsr, seed)), dev(ctx, xh), r)
        rref = oracle.residual(M, Acsr_b(Acsr, seed), xh)
        assert np.array_equal(r.numpy()[short], rref[short])
        if not short.all():
            assert rel_err(r.numpy()[~short], rref[~short]) < 1e-12
        S.Residual(dev(ctx, Acsr_b(Acsr, seed)), dev(ctx, np.zeros(n)), r)
        assert np.array_equal(r.numpy(), Acsr_b(Acsr, seed))
    # f32: short rows equal the plain u32 blocks' thread-per-row sums
    S.make_f32()
    U.make_f32()
    x32 = dev(ctx, xh.astype(np.float32), fx.F32)
    a, b = fx.Vec(ctx, n, fx.F32), fx.Vec(ctx, n, fx.F32)
    S.Apply(x32, a)
    U.Apply(x32, b)
    if lens.max() <= 32:  # (longer rows change the plain blocks' lane groups)
        assert np.array_equal(a.numpy()[short], b.numpy()[short])
    assert rel_err(a.numpy(), b.numpy()) < 1e-6


def Acsr_b(Acsr, seed):
    return np.random.default_rng(seed + 100).uniform(-1, 1, Acsr.nrow)


@pytest.mark.parametrize("cid,scale", [(1, None), (2, 32), (3, 256), (5, 64), (4, 16384),
                                       (4, 200000)])
def test_sliced_format(ctx, oracle, problems, monkeypatch, cid, scale):
    """Sliced rows (the default for irregular matrices such as the power-law
    system; forced here for the stencils) against the oracle."""
    pr = problems(cid, scale)
    S, U = _sliced_pair(ctx, pr.A, monkeypatch)
    assert S.index_bytes == 4 and S.sliced and not U.sliced
    _check_sliced(ctx, oracle, pr.A, S, U, cid)
    if cid == 4 and scale == 200000:  # the default form there sorts the rows by length
        D = fx.Csr.from_problem(ctx, pr.A)
        assert D.layout == "sorted"
        monkeypatch.setenv("FASPXX_FX_NOSORT", "1")
        D2 = fx.Csr.from_problem(ctx, pr.A)
        monkeypatch.delenv("FASPXX_FX_NOSORT")
        assert D2.sliced


def test_sliced_rectangular_empty_and_long_rows(ctx, oracle, monkeypatch):
    """A rectangular matrix (n > m and n < m) with empty rows, rows just at
    and past the 32-entry slice limit, rows filling a whole window, and
    windows without any short row; with values and structure-only."""
    rng = np.random.default_rng(5)
    for n, m in ((3000, 2200), (1500, 4000)):
        lens = rng.integers(0, 12, n)
        lens[rng.random(n) < 0.1] = 0
        lens[rng.integers(0, n, 40)] = 32
        lens[rng.integers(0, n, 40)] = 33
        lens[rng.integers(0, n, 12)] = rng.integers(100, min(m, 1500), 12)
        lens[512:768] = 40  # a window of long rows only
        lens = np.minimum(lens, m)
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
        ci = np.concatenate([np.sort(rng.choice(m, k, replace=False)) for k in lens]).astype(
            np.uint32)
        va = rng.uniform(-1, 1, int(rp[-1]))
        for values in (va, None):
            monkeypatch.setenv("FASPXX_FX_SELL", "1")
            S = fx.Csr(ctx, n, m, rp, ci, values, None)
            monkeypatch.delenv("FASPXX_FX_SELL")
            assert S.sliced
            xh = rng.uniform(-1, 1, m)
            w = fx.Vec(ctx, n)
            S.Apply(dev(ctx, xh), w)
            M = Matrix(n, m, rp, ci, values, oracle.form_diagptr(n, rp, ci))
            ref = oracle.spmv(M, xh)
            short = lens <= 32
            assert np.array_equal(w.numpy()[short], ref[short])
            assert rel_err(w.numpy()[~short], ref[~short]) < 1e-12


def _check_sorted(ctx, oracle, Acsr, D, seed):
    """Apply (f64, f32) and Residual of a matrix with rows sorted by length:
    every row is summed by the u32 row-block code, and the blocks keep rows
    of at most 32 entries apart from longer ones, so those rows (one thread
    each, left to right) equal the oracle bit for bit and the rest agree to
    1e-12 relative; the x = 0 shortcut gives b."""
    n, m = Acsr.nrow, Acsr.mcol
    M = Matrix.of(Acsr)
    lens = np.diff(Acsr.row_ptr.astype(np.int64))
    short = lens <= 32
    xh = np.random.default_rng(seed).uniform(-1, 1, m)
    w = fx.Vec(ctx, n)
    D.Apply(dev(ctx, xh), w)
    ref = oracle.spmv(M, xh)
    assert np.array_equal(w.numpy()[short], ref[short])
    assert rel_err(w.numpy(), ref) < 1e-12
    if n == m:
        b = Acsr_b(Acsr, seed)
        r = fx.Vec(ctx, n)
        D.Residual(dev(ctx, b), dev(ctx, xh), r)
        rref = oracle.residual(M, b, xh)
        assert np.array_equal(r.numpy()[short], rref[short])
        assert rel_err(r.numpy(), rref) < 1e-12
        D.Residual(dev(ctx, b), dev(ctx, np.zeros(n)), r)
        assert np.array_equal(r.numpy(), b)
    D.make_f32()
    a = fx.Vec(ctx, n, fx.F32)
    D.Apply(dev(ctx, xh.astype(np.float32), fx.F32), a)
    assert rel_err(a.numpy(), ref) < 1e-5


def test_sorted_rows_powerlaw(ctx, oracle, problems):
    """The power-law system past the 16-bit reach takes rows sorted by
    length by default."""
    pr = problems(4, 200000)
    D = fx.Csr.from_problem(ctx, pr.A)
    assert D.layout == "sorted" and D.index_bytes == 4
    _check_sorted(ctx, oracle, pr.A, D, 4)


def test_sorted_rows_rectangular_empty_and_long(ctx, oracle):
    """Rectangular matrices (n > m and n < m) with empty rows, rows longer
    than a block's 1792 entries and many short rows, with values and
    structure-only (columns too scattered for 16-bit deltas)."""
    rng = np.random.default_rng(11)
    for n, m in ((80000, 50000), (6000, 90000)):  # columns past the 16-bit reach
        lens = rng.integers(0, 12, n)
        lens[rng.random(n) < 0.1] = 0
        lens[rng.integers(0, n, 30)] = rng.integers(33, 300, 30)
        lens[rng.integers(0, n, 4)] = rng.integers(1800, min(m, 4000), 4)
        lens = np.minimum(lens, m)
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint32)
        ci = np.concatenate([np.sort(rng.choice(m, k, replace=False)) for k in lens]).astype(
            np.uint32)
        va = rng.uniform(-1, 1, int(rp[-1]))
        for values in (va, None):
            D = fx.Csr(ctx, n, m, rp, ci, values, None)
            assert D.layout == "sorted"
            xh = rng.uniform(-1, 1, m)
            w = fx.Vec(ctx, n)
            D.Apply(dev(ctx, xh), w)
            M = Matrix(n, m, rp, ci, values, oracle.form_diagptr(n, rp, ci))
            ref = oracle.spmv(M, xh)
            short = lens <= 32
            assert np.array_equal(w.numpy()[short], ref[short])
            assert rel_err(w.numpy(), ref) < 1e-12


def test_sorted_rows_values_refresh(ctx, oracle, problems):
    """New values (set on the device, a diagonal shift) reach the sorted
    copy, in f64 and f32."""
    pr = problems(4, 200000)
    A = fx.Csr.from_problem(ctx, pr.A)
    assert A.layout == "sorted"
    A.make_f32()
    M = Matrix.of(pr.A)
    xh = np.random.default_rng(3).uniform(-1, 1, pr.A.nrow)
    short = np.diff(pr.A.row_ptr.astype(np.int64)) <= 32
    nv = dev(ctx, pr.A.values * -1.5)
    A.set_values_dev(nv.data_ptr)
    A.shift_diag(0.25)
    want_vals = oracle.mat_shift(Matrix(M.nrow, M.mcol, M.rp, M.ci, pr.A.values * -1.5, M.dp), 0.25)
    w = fx.Vec(ctx, pr.A.nrow)
    A.Apply(dev(ctx, xh), w)
    ref = oracle.spmv(Matrix(M.nrow, M.mcol, M.rp, M.ci, want_vals, M.dp), xh)
    assert np.array_equal(w.numpy()[short], ref[short])
    assert rel_err(w.numpy(), ref) < 1e-12
    a = fx.Vec(ctx, pr.A.nrow, fx.F32)
    A.Apply(dev(ctx, xh.astype(np.float32), fx.F32), a)
    assert rel_err(a.numpy(), ref) < 1e-5


def test_sliced_values_refresh(ctx, oracle, problems, monkeypatch):
    """New values (set on the device, a diagonal shift) reach the slices."""
    pr = problems(4, 16384)
    monkeypatch.setenv("FASPXX_FX_SELL", "1")
    A = fx.Csr.from_problem(ctx, pr.A)
    monkeypatch.delenv("FASPXX_FX_SELL")
    assert A.sliced
    M = Matrix.of(pr.A)
    xh = np.random.default_rng(3).uniform(-1, 1, pr.A.nrow)
    lens = np.diff(pr.A.row_ptr.astype(np.int64))
    short = lens <= 32
    nv = dev(ctx, pr.A.values * -1.5)
    A.set_values_dev(nv.data_ptr)
    A.shift_diag(0.25)
    want_vals = oracle.mat_shift(Matrix(M.nrow, M.mcol, M.rp, M.ci, pr.A.values * -1.5, M.dp), 0.25)
    w = fx.Vec(ctx, pr.A.nrow)
    A.Apply(dev(ctx, xh), w)
    ref = oracle.spmv(Matrix(M.nrow, M.mcol, M.rp, M.ci, want_vals, M.dp), xh)
    assert np.array_equal(w.numpy()[short], ref[short])
    assert rel_err(w.numpy(), ref) < 1e-12
