"""ctypes access to the oracle (TEST INFRASTRUCTURE ONLY).

``Oracle`` wraps liboracle.so, the CPU restatement of the reference path
(fx_oracle.cpp).  ``Reference`` wraps _ref/libfaspxx_ref.so, the reference's
own sources compiled by oracle/Makefile (present only where the reference
checkout was available at build time).  Only tests/, the smoke check and the
benchmark's CPU-baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess
import tempfile

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libfaspxx_ref.so"
REF_OMP_SO = HERE / "_ref" / "libfaspxx_ref_omp.so"

CG, BICGSTAB, GMRES, FGMRES, JACOBI = 1, 2, 4, 5, 11


def build(ref: bool = True) -> None:
    """Compile the oracle (and the reference, when its sources exist)."""
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", str(HERE), target], check=True)


class _Csr(C.Structure):
    _fields_ = [("nrow", C.c_uint32), ("mcol", C.c_uint32), ("nnz", C.c_uint32),
                ("rp", C.c_void_p), ("ci", C.c_void_p), ("va", C.c_void_p), ("dp", C.c_void_p)]


class _Params(C.Structure):
    _fields_ = [("maxIter", C.c_uint32), ("minIter", C.c_uint32), ("savIter", C.c_uint32),
                ("restart", C.c_uint32), ("relTol", C.c_double), ("absTol", C.c_double),
                ("maxRestart", C.c_uint32), ("minRestart", C.c_uint32), ("leftPrecond", C.c_int)]


class _Pc(C.Structure):
    _fields_ = [("kind", C.c_int), ("iters", C.c_uint32), ("dinv", C.c_void_p),
                ("color", C.c_void_p), ("ncolors", C.c_uint32), ("weight", C.c_double)]


class _MgLevel(C.Structure):
    _fields_ = [("transfer", C.POINTER(_Csr)), ("smooth", _Params), ("coarse", _Params)]


class _MatOut(C.Structure):
    _fields_ = [("nrow", C.c_uint32), ("mcol", C.c_uint32), ("nnz", C.c_uint32),
                ("rp", C.POINTER(C.c_uint32)), ("ci", C.POINTER(C.c_uint32)),
                ("va", C.POINTER(C.c_double)), ("dp", C.POINTER(C.c_uint32))]


class _RefMatOut(C.Structure):
    _fields_ = [("nrow", C.c_uint32), ("mcol", C.c_uint32), ("nnz", C.c_uint32),
                ("rp", C.POINTER(C.c_uint32)), ("ci", C.POINTER(C.c_uint32)),
                ("va", C.POINTER(C.c_double))]


def _take(o, with_dp):
    """Copy a malloc'd result struct into numpy arrays (rp, ci, va | None, dp | None)."""
    n, z = o.nrow, o.nnz
    rp = np.ctypeslib.as_array(o.rp, (n + 1,)).copy()
    ci = np.ctypeslib.as_array(o.ci, (z,)).copy() if z else np.zeros(0, np.uint32)
    va = (np.ctypeslib.as_array(o.va, (z,)).copy() if z else np.zeros(0)) if o.va else None
    dp = None
    if with_dp:
        dp = np.ctypeslib.as_array(o.dp, (n,)).copy() if n else np.zeros(0, np.uint32)
    return rp, ci, va, dp


class _Cycles(C.Structure):
    _fields_ = [("cap", C.c_uint32), ("count", C.c_uint32), ("len", C.c_void_p),
                ("planned", C.c_void_p), ("est", C.c_void_p), ("rtrue", C.c_void_p)]


class _Result(C.Structure):
    _fields_ = [("retcode", C.c_int), ("numIter", C.c_uint32), ("norm2", C.c_double),
                ("normInf", C.c_double)]


class _RefResult(C.Structure):
    _fields_ = [("retcode", C.c_int), ("numIter", C.c_uint32), ("norm2", C.c_double),
                ("normInf", C.c_double), ("solve_ms", C.c_double)]


class _RefParams(C.Structure):
    _fields_ = [("maxIter", C.c_uint32), ("minIter", C.c_uint32), ("savIter", C.c_uint32),
                ("restart", C.c_uint32), ("relTol", C.c_double), ("absTol", C.c_double)]


def _p(a):
    return None if a is None else a.ctypes.data


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


DEFAULTS = dict(maxIter=1, minIter=0, savIter=100000, restart=30, relTol=1e-6, absTol=1e-8,
                maxRestart=30, minRestart=10, leftPrecond=0)


class Matrix:
    """Holds CSRx arrays alive for ctypes calls."""

    def __init__(self, nrow, mcol, rp, ci, va=None, dp=None):
        self.nrow, self.mcol = int(nrow), int(mcol)
        self.rp, self.ci = _u32(rp), _u32(ci)
        self.va = None if va is None else _f64(va)
        self.dp = None if dp is None else _u32(dp)
        self.nnz = int(self.ci.shape[0])
        self.c = _Csr(self.nrow, self.mcol, self.nnz, _p(self.rp), _p(self.ci), _p(self.va),
                      _p(self.dp))

    @classmethod
    def of(cls, A) -> "Matrix":
        return cls(A.nrow, A.mcol, A.row_ptr, A.col_ind, A.values, A.diag_ptr)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not pathlib.Path(path).exists():
            build(ref=False)
        L = C.CDLL(str(path))
        cs = C.POINTER(_Csr)
        d = C.c_double
        u = C.c_uint32
        v = C.c_void_p
        sig = {
            "orc_dot": (d, [u, v, v]), "orc_norm2": (d, [u, v]), "orc_norminf": (d, [u, v]),
            "orc_axpy": (None, [u, d, v, v]), "orc_xpay": (None, [u, d, v, v]),
            "orc_axpby": (None, [u, d, d, v, v]), "orc_waxpby": (None, [u, d, v, d, v, v]),
            "orc_pmul": (None, [u, v, v]), "orc_scale": (None, [u, d, v]),
            "orc_add": (None, [u, v, v]),
            "orc_spmv": (None, [cs, v, v]), "orc_residual": (None, [cs, v, v, v]),
            "orc_form_diagptr": (None, [u, v, v, v]), "orc_get_diag": (None, [cs, v]),
            "orc_jacobi_dinv": (None, [cs, d, v]), "orc_transpose": (None, [cs, v, v, v, v]),
            "orc_check_csrx": (C.c_int, [cs]),
            "orc_mc_color": (u, [cs, v]),
            "orc_mcgs": (None, [cs, v, u, d, v, v, u, C.c_int]),
            "orc_solve": (C.c_int, [C.c_int, cs, C.POINTER(_Pc), C.POINTER(_Params), v, v,
                                    C.POINTER(_Result), v, u]),
            "orc_mg": (C.c_int, [cs, v, u, C.POINTER(_MgLevel), v, C.POINTER(_Params), v, v,
                                 C.POINTER(_Result), v, u]),
            "orc_mat_free": (None, [C.POINTER(_MatOut)]),
            "orc_mat_tri": (None, [cs, C.c_int, C.POINTER(_MatOut)]),
            "orc_mat_add": (C.c_int, [d, cs, d, cs, C.POINTER(_MatOut)]),
            "orc_mat_mult": (None, [cs, cs, C.POINTER(_MatOut)]),
            "orc_mat_shift": (None, [cs, d, v]),
            "orc_krylov_traced": (C.c_int, [C.c_int, cs, C.POINTER(_Pc), C.POINTER(_Params), v,
                                            v, C.POINTER(_Result), v, u, C.POINTER(_Cycles), v,
                                            v, C.POINTER(C.c_uint32)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        self.L = L

    # VEC
    def dot(self, a, b):
        a, b = _f64(a), _f64(b)
        return self.L.orc_dot(a.shape[0], _p(a), _p(b))

    def norm2(self, a):
        a = _f64(a)
        return self.L.orc_norm2(a.shape[0], _p(a))

    def norminf(self, a):
        a = _f64(a)
        return self.L.orc_norminf(a.shape[0], _p(a))

    def axpy(self, a, x, y):
        x, y = _f64(x), _f64(y).copy()
        self.L.orc_axpy(x.shape[0], a, _p(x), _p(y))
        return y

    def xpay(self, a, x, y):
        x, y = _f64(x), _f64(y).copy()
        self.L.orc_xpay(x.shape[0], a, _p(x), _p(y))
        return y

    def axpby(self, a, b, x, y):
        x, y = _f64(x), _f64(y).copy()
        self.L.orc_axpby(x.shape[0], a, b, _p(x), _p(y))
        return y

    def waxpby(self, a, x, b, y):
        x, y = _f64(x), _f64(y)
        w = np.empty_like(x)
        self.L.orc_waxpby(x.shape[0], a, _p(x), b, _p(y), _p(w))
        return w

    # MAT
    def spmv(self, M: Matrix, v):
        v = _f64(v)
        w = np.empty(M.nrow)
        self.L.orc_spmv(C.byref(M.c), _p(v), _p(w))
        return w

    def residual(self, M: Matrix, b, x):
        b, x = _f64(b), _f64(x)
        r = np.empty(M.nrow)
        self.L.orc_residual(C.byref(M.c), _p(b), _p(x), _p(r))
        return r

    def form_diagptr(self, nrow, rp, ci):
        rp, ci = _u32(rp), _u32(ci)
        dp = np.zeros(nrow, np.uint32)
        self.L.orc_form_diagptr(nrow, _p(rp), _p(ci), _p(dp))
        return dp

    def mc_color(self, M: Matrix):
        """Greedy natural-order coloring of A + A^T: (ncolors, color[nrow])."""
        c = np.empty(M.nrow, np.uint32)
        nc = self.L.orc_mc_color(C.byref(M.c), _p(c))
        return nc, c

    def mcgs(self, M: Matrix, color, ncolors, weight, b, x0, sweeps=1, symmetric=False):
        """Multicolor Gauss-Seidel sweeps from x0; returns the new x."""
        x = _f64(x0).copy()
        color = _u32(color)
        self.L.orc_mcgs(C.byref(M.c), _p(color), ncolors, weight, _p(_f64(b)), _p(x), sweeps,
                        1 if symmetric else 0)
        return x

    def jacobi_dinv(self, M: Matrix, weight):
        d = np.empty(M.nrow)
        self.L.orc_jacobi_dinv(C.byref(M.c), weight, _p(d))
        return d

    def transpose(self, M: Matrix):
        rp = np.zeros(M.mcol + 1, np.uint32)
        ci = np.zeros(M.nnz, np.uint32)
        va = np.zeros(M.nnz) if M.va is not None else None
        dp = np.zeros(M.mcol, np.uint32)
        self.L.orc_transpose(C.byref(M.c), _p(rp), _p(ci), _p(va), _p(dp))
        return rp, ci, va, dp

    def check_csrx(self, M: Matrix) -> int:
        return self.L.orc_check_csrx(C.byref(M.c))

    # solvers
    def solve(self, kind, M: Matrix, b, x0, pc="identity", weight=1.0, pc_iters=1,
              hist_cap=0, **prm):
        P = dict(DEFAULTS)
        P.update(prm)
        params = _Params(**P)
        b = _f64(b)
        x = _f64(x0).copy()
        dinv = None
        if pc == "jacobi" or kind == JACOBI:
            dinv = self.jacobi_dinv(M, weight)
        if pc == "mcgs":
            ncol, color = self.mc_color(M)
            pcs = _Pc(2, pc_iters, None, _p(color), ncol, weight)
        else:
            pcs = _Pc(1 if dinv is not None else 0, pc_iters, _p(dinv), None, 0, weight)
        res = _Result()
        hist = np.zeros(hist_cap) if hist_cap else None
        self.L.orc_solve(kind, C.byref(M.c), C.byref(pcs), C.byref(params), _p(b), _p(x),
                         C.byref(res), _p(hist), hist_cap)
        out = {"retcode": res.retcode, "numIter": res.numIter, "norm2": res.norm2,
               "normInf": res.normInf, "x": x}
        if hist is not None:
            out["hist"] = hist
        return out

    def krylov_traced(self, kind, M: Matrix, b, x0, pc="identity", weight=1.0, z=None,
                      restart_in=0, want_basis=False, cycle_cap=100000, hist_cap=0, **prm):
        """GMRES or FGMRES with the oracle's test instrumentation.  Returns
        the solve() dict plus "cycles" (per cycle: inner iterations, the
        Givens estimate at its end, the true residual norm after its x
        update), "planned" (per cycle, the restart length it started with),
        "V" (the last cycle's basis, when want_basis), "restart"
        (the restart length the solver object keeps for its next Solve;
        `restart_in` is the one it starts with, 0 = freshly set up) and, for
        FGMRES, "Z" (the kept basis on exit; `z` is the stale basis on entry,
        (maxR+1) x n, zeros when None)."""
        P = dict(DEFAULTS)
        P.update(prm)
        params = _Params(**P)
        maxR = min(P["maxRestart"], P["maxIter"])
        n = M.nrow
        b = _f64(b)
        x = _f64(x0).copy()
        dinv = self.jacobi_dinv(M, weight) if pc == "jacobi" else None
        pcs = _Pc(1 if dinv is not None else 0, 1, _p(dinv), None, 0, weight)
        res = _Result()
        ln = np.zeros(cycle_cap, np.uint32)
        planned = np.zeros(cycle_cap, np.uint32)
        est = np.zeros(cycle_cap)
        rtrue = np.zeros(cycle_cap)
        cyc = _Cycles(cycle_cap, 0, _p(ln), _p(planned), _p(est), _p(rtrue))
        V = np.zeros((maxR + 1, n)) if want_basis else None
        Z = None
        if kind == FGMRES:
            Z = np.zeros((maxR + 1, n)) if z is None else np.ascontiguousarray(z, np.float64).copy()
            assert Z.shape == (maxR + 1, n)
        hist = np.zeros(hist_cap) if hist_cap else None
        rs = C.c_uint32(restart_in)
        self.L.orc_krylov_traced(kind, C.byref(M.c), C.byref(pcs), C.byref(params), _p(b),
                                 _p(x), C.byref(res), _p(hist), hist_cap, C.byref(cyc), _p(V),
                                 _p(Z), C.byref(rs))
        k = min(cyc.count, cycle_cap)
        out = {"retcode": res.retcode, "numIter": res.numIter, "norm2": res.norm2,
               "normInf": res.normInf, "x": x,
               "cycles": list(zip(ln[:k].tolist(), est[:k].tolist(), rtrue[:k].tolist())),
               "planned": planned[:k].tolist(),
               "V": V, "Z": Z, "dinv": dinv, "restart": rs.value}
        if hist is not None:
            out["hist"] = hist
        return out

    # setup-time sparse algebra: each returns (nrow, mcol, rp, ci, va | None, dp)
    def _mat(self, fn, *args):
        o = _MatOut()
        rc = fn(*args, C.byref(o))
        if rc == 1:
            return None
        try:
            return (o.nrow, o.mcol) + _take(o, True)
        finally:
            self.L.orc_mat_free(C.byref(o))

    def lower_tri(self, M: Matrix):
        return self._mat(self.L.orc_mat_tri, C.byref(M.c), 0)

    def upper_tri(self, M: Matrix):
        return self._mat(self.L.orc_mat_tri, C.byref(M.c), 1)

    def mat_add(self, a, A: Matrix, b, B: Matrix):
        """a A + b B as MAT::Add; None when the reference leaves its target
        unchanged (an operand without entries)."""
        return self._mat(self.L.orc_mat_add, a, C.byref(A.c), b, C.byref(B.c))

    def mat_mult(self, L: Matrix, R: Matrix):
        return self._mat(self.L.orc_mat_mult, C.byref(L.c), C.byref(R.c))

    def mat_shift(self, M: Matrix, a):
        """MAT::Shift: the shifted values (M's dp must be set)."""
        va = M.va.copy()
        self.L.orc_mat_shift(C.byref(M.c), a, _p(va))
        return va

    def mg(self, M: Matrix, b, x0, levels, ncycles=1, weight=1.0, hist_cap=0, **prm):
        """MG<MAT>::Solve with `levels` = [(transfer, smooth_prm, coarse_prm)]:
        transfer a square Matrix or None (IdentityMatrix), the two dicts the
        SOLParams of the level's Jacobi smoother and coarse solver (defaults
        as SOLParams()).  ncycles: an int for every level or a list."""
        nlev = len(levels)
        keep = []
        arr = (_MgLevel * max(nlev, 1))()
        for i, (T, sp, cp) in enumerate(levels):
            arr[i].transfer = C.pointer(T.c) if T is not None else None
            arr[i].smooth = _params(**(sp or {}))
            arr[i].coarse = _params(**(cp or {}))
            keep.append(T)
        nc = np.full(max(nlev, 1), ncycles, np.uint32) if np.isscalar(ncycles) else \
            _u32(ncycles)
        dinv = self.jacobi_dinv(M, weight)
        b = _f64(b)
        x = _f64(x0).copy()
        res = _Result()
        hist = np.zeros(hist_cap) if hist_cap else None
        rc = self.L.orc_mg(C.byref(M.c), _p(dinv), nlev, arr, _p(nc), C.byref(_params(**prm)),
                           _p(b), _p(x), C.byref(res), _p(hist), hist_cap)
        out = {"retcode": rc, "numIter": res.numIter, "norm2": res.norm2,
               "normInf": res.normInf, "x": x}
        if hist is not None:
            out["hist"] = hist
        return out


def _params(**prm) -> _Params:
    P = dict(DEFAULTS)
    P.update(prm)
    return _Params(**P)


class Reference:
    """The compiled reference (oracle/_ref).  GMRES / FGMRES are not exposed:
    their back-substitution loops do not terminate (src/GMRES.cxx:214)."""

    def __init__(self, path=REF_SO):
        if not pathlib.Path(path).exists():
            raise FileNotFoundError(path)
        L = C.CDLL(str(path))
        d, u, v = C.c_double, C.c_uint32, C.c_void_p
        sig = {
            "ref_spmv": (None, [u, u, u, v, v, v, v, v, v]),
            "ref_residual": (None, [u, u, u, v, v, v, v, v, v, v]),
            "ref_transpose": (None, [u, u, u, v, v, v, v, v, v, v]),
            "ref_check_csrx": (C.c_int, [u, u, u, v, v, v, v]),
            "ref_dot": (d, [u, v, v]), "ref_norm2": (d, [u, v]), "ref_norminf": (d, [u, v]),
            "ref_axpy": (None, [u, d, v, v]), "ref_xpay": (None, [u, d, v, v]),
            "ref_axpby": (None, [u, d, d, v, v]), "ref_waxpby": (None, [u, d, v, d, v, v]),
            "ref_solve": (C.c_int, [C.c_int, u, u, v, v, v, v, C.c_int, u, d,
                                    C.POINTER(_RefParams), v, v, C.POINTER(_RefResult)]),
            "ref_mat_free": (None, [C.POINTER(_RefMatOut)]),
            "ref_mat_tri": (None, [u, u, u, v, v, v, v, C.c_int, C.POINTER(_RefMatOut)]),
            "ref_mat_add": (C.c_int, [d, u, u, u, v, v, v, d, u, v, v, v,
                                      C.POINTER(_RefMatOut)]),
            "ref_mat_mult": (None, [u, u, u, v, v, v, u, u, v, v, v, C.POINTER(_RefMatOut)]),
            "ref_mat_shift": (None, [u, u, u, v, v, v, d, v]),
            "ref_omp_threads": (C.c_int, []),
            "ref_read_mat": (C.c_int, [C.c_char_p, C.POINTER(_RefMatOut)]),
            "ref_read_csr": (C.c_int, [C.c_char_p, C.POINTER(_RefMatOut)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        self.L = L

    @staticmethod
    def available(path=REF_SO) -> bool:
        return pathlib.Path(path).exists()

    def threads(self) -> int:
        """Threads of the reference's OpenMP loops (1 for the serial build)."""
        return self.L.ref_omp_threads()

    def _mat(self, fn, *args):
        o = _RefMatOut()
        rc = fn(*args, C.byref(o))
        if rc == 1:
            return None
        try:
            return (o.nrow, o.mcol) + _take(o, False)[:3]
        finally:
            self.L.ref_mat_free(C.byref(o))

    def tri(self, M: Matrix, upper):
        return self._mat(self.L.ref_mat_tri, M.nrow, M.mcol, M.nnz, _p(M.rp), _p(M.ci), _p(M.va),
                         _p(M.dp), 1 if upper else 0)

    def mat_add(self, a, A: Matrix, b, B: Matrix):
        return self._mat(self.L.ref_mat_add, a, A.nrow, A.mcol, A.nnz, _p(A.rp), _p(A.ci),
                         _p(A.va), b, B.nnz, _p(B.rp), _p(B.ci), _p(B.va))

    def mat_mult(self, L: Matrix, R: Matrix):
        return self._mat(self.L.ref_mat_mult, L.nrow, L.mcol, L.nnz, _p(L.rp), _p(L.ci),
                         _p(L.va), R.mcol, R.nnz, _p(R.rp), _p(R.ci), _p(R.va))

    def _read(self, fn, path):
        path = pathlib.Path(path)
        o = _RefMatOut()
        if path.suffix.lower() == ".csr" and path.exists():
            # ReadCSR loads the file into a buffer of its exact length and
            # scans the values for a NUL that the buffer does not hold
            # (src/ReadData.cxx:243-251,331-345), so it reads past the
            # allocation and its result depends on the heap there.  It is
            # given a copy of the file with a NUL appended: the scan then
            # ends inside the buffer, after the file's own last line.
            with tempfile.TemporaryDirectory() as td:
                tmp = pathlib.Path(td) / path.name
                tmp.write_bytes(path.read_bytes() + b"\0")
                rc = fn(str(tmp).encode(), C.byref(o))
        else:
            rc = fn(str(path).encode(), C.byref(o))
        if rc != 0:
            return rc
        try:
            return (o.nrow, o.mcol) + _take(o, False)[:3]
        finally:
            self.L.ref_mat_free(C.byref(o))

    def read_mat(self, path):
        """ReadMat: (nrow, mcol, rp, ci, va) of the normalised CSRx, or the
        reference's (nonzero) error code."""
        return self._read(self.L.ref_read_mat, path)

    def read_csr(self, path):
        """ReadCSR: the file's raw arrays, or the reference's error code."""
        return self._read(self.L.ref_read_csr, path)

    def mat_shift(self, M: Matrix, a):
        """MAT::Shift with the reference's own FormDiagPtr; the shifted values."""
        out = np.empty(M.nnz)
        self.L.ref_mat_shift(M.nrow, M.mcol, M.nnz, _p(M.rp), _p(M.ci), _p(M.va), a, _p(out))
        return out

    def _m(self, M: Matrix):
        return (M.nrow, M.mcol, M.nnz, _p(M.rp), _p(M.ci), _p(M.va), _p(M.dp))

    def spmv(self, M: Matrix, v):
        v = _f64(v)
        w = np.empty(M.nrow)
        self.L.ref_spmv(*self._m(M), _p(v), _p(w))
        return w

    def residual(self, M: Matrix, b, x):
        b, x = _f64(b), _f64(x)
        r = np.empty(M.nrow)
        self.L.ref_residual(*self._m(M), _p(b), _p(x), _p(r))
        return r

    def transpose(self, M: Matrix):
        rp = np.zeros(M.mcol + 1, np.uint32)
        ci = np.zeros(M.nnz, np.uint32)
        va = np.zeros(M.nnz) if M.va is not None else None
        self.L.ref_transpose(*self._m(M), _p(rp), _p(ci), _p(va))
        return rp, ci, va

    def check_csrx(self, M: Matrix) -> int:
        return self.L.ref_check_csrx(*self._m(M))

    def dot(self, a, b):
        a, b = _f64(a), _f64(b)
        return self.L.ref_dot(a.shape[0], _p(a), _p(b))

    def norm2(self, a):
        a = _f64(a)
        return self.L.ref_norm2(a.shape[0], _p(a))

    def norminf(self, a):
        a = _f64(a)
        return self.L.ref_norminf(a.shape[0], _p(a))

    def axpy(self, a, x, y):
        x, y = _f64(x), _f64(y).copy()
        self.L.ref_axpy(x.shape[0], a, _p(x), _p(y))
        return y

    def xpay(self, a, x, y):
        x, y = _f64(x), _f64(y).copy()
        self.L.ref_xpay(x.shape[0], a, _p(x), _p(y))
        return y

    def axpby(self, a, b, x, y):
        x, y = _f64(x), _f64(y).copy()
        self.L.ref_axpby(x.shape[0], a, b, _p(x), _p(y))
        return y

    def waxpby(self, a, x, b, y):
        x, y = _f64(x), _f64(y)
        w = np.empty_like(x)
        self.L.ref_waxpby(x.shape[0], a, _p(x), b, _p(y), _p(w))
        return w

    def solve(self, kind, M: Matrix, b, x0, pc="identity", weight=1.0, pc_iters=1, **prm):
        P = dict(DEFAULTS)
        P.update(prm)
        params = _RefParams(P["maxIter"], P["minIter"], P["savIter"], P["restart"],
                            P["relTol"], P["absTol"])
        b = _f64(b)
        x = _f64(x0).copy()
        res = _RefResult()
        self.L.ref_solve(kind, M.nrow, M.nnz, _p(M.rp), _p(M.ci), _p(M.va), _p(M.dp),
                         1 if pc == "jacobi" else 0, pc_iters, weight, C.byref(params), _p(b),
                         _p(x), C.byref(res))
        return {"retcode": res.retcode, "numIter": res.numIter, "norm2": res.norm2,
                "normInf": res.normInf, "solve_ms": res.solve_ms, "x": x}
