"""ctypes binding of libfaspxx_hip.so (the C ABI in include/faspxx_hip.h).

The Python objects mirror the reference's classes: ``Csr`` is MAT
(src/MAT.hxx), ``Vec`` is VEC (src/VEC.hxx), ``Solver`` covers CG /
BiCGStab / GMRES / FGMRES / Jacobi (src/SOL.hxx and subclasses) and
``krylov`` is the free function Krylov() (src/Krylov.hxx:25).  Every call goes
to the HIP library; there is no CPU fallback: if the library is missing,
``load()`` raises.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import pathlib
import weakref

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libfaspxx_hip.so"

F64, F32 = 0, 1
CG, BICGSTAB, GMRES, FGMRES, JACOBI = 1, 2, 4, 5, 11

RETCODES = {
    0: "SUCCESS", -12: "ERROR_INPUT_PAR", -14: "ERROR_VEC_SIZE", -15: "ERROR_MAT_SIZE",
    -16: "ERROR_NONMATCH_SIZE", -17: "ERROR_MAT_DATA", -18: "ERROR_DIVIDE_ZERO",
    -20: "ERROR_ALLOC_MEM", -30: "ERROR_SOLVER_TYPE", -31: "ERROR_SOLVER_PCD_TYPE",
    -32: "ERROR_SOLVER_STAG", -33: "ERROR_SOLVER_SOLSTAG", -34: "ERROR_SOLVER_TOLSMALL",
    -39: "ERROR_SOLVER_MAXIT", -90: "ERROR_DEVICE",
}


class Params(C.Structure):
    """SOLParams (src/Param.hxx:57-80) + GMRES restart bounds."""

    _fields_ = [
        ("maxIter", C.c_uint32), ("minIter", C.c_uint32), ("savIter", C.c_uint32),
        ("restart", C.c_uint32), ("relTol", C.c_double), ("absTol", C.c_double),
        ("verbose", C.c_int), ("maxRestart", C.c_uint32), ("minRestart", C.c_uint32),
        ("leftPrecond", C.c_int),
    ]


class Result(C.Structure):
    _fields_ = [("retcode", C.c_int), ("numIter", C.c_uint32), ("norm2", C.c_double),
                ("normInf", C.c_double), ("solve_ms", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class FxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{RETCODES.get(code, code)}: {msg}")
        self.code = code


_lib = None
_exiting = False


@atexit.register
def _mark_exit():
    # Objects still alive at interpreter exit are not freed one by one: the
    # HIP runtime may already be shutting down, and the process is ending.
    global _exiting
    _exiting = True

_P = C.c_void_p
_SIGS = {
    "fx_last_error": (C.c_char_p, []),
    "fx_params_default": (None, [C.POINTER(Params)]),
    "fx_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "fx_ctx_create_on_stream": (C.c_int, [C.c_int, _P, C.POINTER(_P)]),
    "fx_ctx_destroy": (C.c_int, [_P]),
    "fx_ctx_sync": (C.c_int, [_P]),
    "fx_ctx_stream": (_P, [_P]),
    "fx_vec_create": (C.c_int, [_P, C.c_uint32, C.c_int, C.POINTER(_P)]),
    "fx_vec_wrap": (C.c_int, [_P, C.c_uint32, C.c_int, _P, C.POINTER(_P)]),
    "fx_vec_destroy": (C.c_int, [_P]),
    "fx_vec_size": (C.c_uint32, [_P]),
    "fx_vec_dtype": (C.c_int, [_P]),
    "fx_vec_data": (_P, [_P]),
    "fx_vec_upload": (C.c_int, [_P, _P]),
    "fx_vec_download": (C.c_int, [_P, _P]),
    "fx_vec_set": (C.c_int, [_P, C.c_double]),
    "fx_vec_copy": (C.c_int, [_P, _P]),
    "fx_axpy": (C.c_int, [_P, C.c_double, _P, _P]),
    "fx_xpay": (C.c_int, [_P, C.c_double, _P, _P]),
    "fx_axpby": (C.c_int, [_P, C.c_double, C.c_double, _P, _P]),
    "fx_waxpby": (C.c_int, [_P, C.c_double, _P, C.c_double, _P, _P]),
    "fx_scale": (C.c_int, [_P, C.c_double, _P]),
    "fx_pmul": (C.c_int, [_P, _P, _P]),
    "fx_recip": (C.c_int, [_P, _P]),
    "fx_add": (C.c_int, [_P, _P, _P]),
    "fx_shift": (C.c_int, [_P, C.c_double, _P]),
    "fx_abs": (C.c_int, [_P, _P]),
    "fx_pdiv": (C.c_int, [_P, _P, _P]),
    "fx_sub": (C.c_int, [_P, _P, _P]),
    "fx_max": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "fx_min": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "fx_dot": (C.c_int, [_P, _P, _P, C.POINTER(C.c_double)]),
    "fx_norm2": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "fx_norminf": (C.c_int, [_P, _P, C.POINTER(C.c_double)]),
    "fx_dot_dev": (C.c_int, [_P, _P, _P, _P]),
    "fx_norm2_dev": (C.c_int, [_P, _P, _P]),
    "fx_axpy_norm2": (C.c_int, [_P, C.c_double, _P, _P, C.POINTER(C.c_double)]),
    "fx_xpay_norm2": (C.c_int, [_P, C.c_double, _P, _P, C.POINTER(C.c_double)]),
    "fx_axpy_dot": (C.c_int, [_P, C.c_double, _P, _P, _P, C.POINTER(C.c_double)]),
    "fx_csr_create": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P, _P, _P,
                                C.POINTER(_P)]),
    "fx_csr_destroy": (C.c_int, [_P]),
    "fx_csr_index_bytes": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "fx_csr_layout": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "fx_spmv_timed": (C.c_int, [_P, _P, _P, _P, C.POINTER(C.c_float)]),
    "fx_csr_info": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                              C.POINTER(C.c_uint32)]),
    "fx_csr_download": (C.c_int, [_P, _P, _P, _P, _P]),
    "fx_csr_set_values_dev": (C.c_int, [_P, _P]),
    "fx_csr_values_dev": (_P, [_P]),
    "fx_csr_make_f32": (C.c_int, [_P]),
    "fx_spmv": (C.c_int, [_P, _P, _P, _P]),
    "fx_residual": (C.c_int, [_P, _P, _P, _P, _P]),
    "fx_csr_get_diag": (C.c_int, [_P, _P, _P]),
    "fx_csr_transpose": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "fx_csr_transpose_update": (C.c_int, [_P, _P, _P]),
    "fx_csr_lower_tri": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "fx_csr_upper_tri": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "fx_csr_add": (C.c_int, [_P, C.c_double, _P, C.c_double, _P, C.POINTER(_P)]),
    "fx_csr_mult": (C.c_int, [_P, _P, _P, C.POINTER(_P)]),
    "fx_csr_shift_diag": (C.c_int, [_P, _P, C.c_double]),
    "fx_op_create": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P, C.POINTER(_P)]),
    "fx_op_destroy": (C.c_int, [_P]),
    "fx_op_set_device_only": (C.c_int, [_P, C.c_int]),
    "fx_solver_setup_op": (C.c_int, [_P, _P, _P, _P]),
    "fx_csr_pattern_grad": (C.c_int, [_P, _P, _P, _P, _P]),
    "fx_pc_create_identity": (C.c_int, [_P, C.POINTER(_P)]),
    "fx_pc_create_mcgs": (C.c_int, [_P, _P, C.c_double, C.c_uint32, C.POINTER(_P)]),
    "fx_mcgs_create": (C.c_int, [_P, _P, C.c_double, C.POINTER(_P)]),
    "fx_mcgs_info": (C.c_int, [_P, C.POINTER(C.c_uint32), _P]),
    "fx_mcgs_sweep": (C.c_int, [_P, _P, _P, C.c_uint32, C.c_int]),
    "fx_mcgs_destroy": (C.c_int, [_P]),
    "fx_pc_create_jacobi": (C.c_int, [_P, _P, C.c_double, C.c_uint32, C.POINTER(_P)]),
    "fx_pc_destroy": (C.c_int, [_P]),
    "fx_pc_apply": (C.c_int, [_P, _P, _P, _P]),
    "fx_pc_apply_noalloc": (C.c_int, [_P, _P, _P, _P]),
    "fx_solver_create": (C.c_int, [_P, C.c_int, C.POINTER(_P)]),
    "fx_solver_setup": (C.c_int, [_P, _P, _P, C.POINTER(Params)]),
    "fx_solver_solve": (C.c_int, [_P, _P, _P, C.POINTER(Result)]),
    "fx_solver_history": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "fx_solver_last_path": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "fx_solver_update": (C.c_int, [_P, _P, _P, C.POINTER(Params)]),
    "fx_solver_clean": (C.c_int, [_P]),
    "fx_solver_update_op": (C.c_int, [_P, _P, _P, C.POINTER(Params)]),
    "fx_solver_destroy": (C.c_int, [_P]),
    "fx_krylov": (C.c_int, [_P, _P, _P, _P, _P, C.c_char_p, C.POINTER(Params),
                            C.POINTER(Result)]),
    "fx_solve_ir": (C.c_int, [_P, C.c_int, _P, C.c_double, C.POINTER(Params),
                              C.POINTER(Params), _P, _P, C.POINTER(Result),
                              C.POINTER(C.c_uint32)]),
    "fx_ir_create": (C.c_int, [_P, C.c_int, _P, C.c_double, C.POINTER(Params),
                               C.POINTER(Params), C.POINTER(_P)]),
    "fx_ir_solve": (C.c_int, [_P, _P, _P, C.POINTER(Result), C.POINTER(C.c_uint32)]),
    "fx_ir_destroy": (C.c_int, [_P]),
}


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} not found: build it with `python -m faspxx_amd.build` (no CPU fallback exists)")
    # PyTorch ships its own HIP runtime under the same soname as the system
    # one (libamdhip64.so.7), and whichever is loaded first serves the whole
    # process.  Loading torch first makes it the runtime the library binds
    # to, so torch tensors and the library share one runtime whatever the
    # order in which the caller initialises them.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(p))
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _chk(rc: int) -> int:
    if rc != 0:
        msg = load().fx_last_error().decode(errors="replace")
        raise FxError(rc, msg)
    return rc


def params(**kw) -> Params:
    """SOLParams defaults (src/Param.hxx:68-79), overridden by keywords."""
    p = Params()
    load().fx_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Context:
    def __init__(self, device: int = 0, stream: int | None = None):
        lib = load()
        h = C.c_void_p()
        if stream is None:
            _chk(lib.fx_ctx_create(device, C.byref(h)))
        else:
            _chk(lib.fx_ctx_create_on_stream(device, C.c_void_p(stream), C.byref(h)))
        self.h = h
        self.device = device
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    def sync(self):
        _chk(load().fx_ctx_sync(self.h))

    @property
    def stream(self) -> int:
        return load().fx_ctx_stream(self.h) or 0

    def close(self):
        """Destroy every live object created on this context, then the context
        (the C ABI requires objects to be destroyed before their context)."""
        if self.h:
            for child in list(self._children):
                child.close()
            load().fx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        if not _exiting:
            try:
                self.close()
            except Exception:
                pass


class Vec:
    """Device vector (VEC)."""

    def __init__(self, ctx: Context, n: int, dtype=F64, _wrap_ptr: int | None = None):
        lib = load()
        h = C.c_void_p()
        if _wrap_ptr is None:
            _chk(lib.fx_vec_create(ctx.h, n, dtype, C.byref(h)))
        else:
            _chk(lib.fx_vec_wrap(ctx.h, n, dtype, C.c_void_p(_wrap_ptr), C.byref(h)))
        self.h, self.ctx, self.n, self.dtype = h, ctx, n, dtype
        ctx._adopt(self)

    @classmethod
    def from_numpy(cls, ctx: Context, a: np.ndarray, dtype=None) -> "Vec":
        dt = dtype if dtype is not None else (F32 if a.dtype == np.float32 else F64)
        a = np.ascontiguousarray(a, dtype=np.float32 if dt == F32 else np.float64)
        v = cls(ctx, a.shape[0], dt)
        v.upload(a)
        return v

    @classmethod
    def wrap(cls, ctx: Context, dev_ptr: int, n: int, dtype=F64) -> "Vec":
        return cls(ctx, n, dtype, _wrap_ptr=dev_ptr)

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32 if self.dtype == F32 else np.float64)
        assert a.shape[0] == self.n
        _chk(load().fx_vec_upload(self.h, _ptr(a)))

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float32 if self.dtype == F32 else np.float64)
        _chk(load().fx_vec_download(self.h, _ptr(out)))
        return out

    @property
    def data_ptr(self) -> int:
        return load().fx_vec_data(self.h) or 0

    def set(self, v: float):
        _chk(load().fx_vec_set(self.h, v))

    def copy_to(self, dst: "Vec"):
        _chk(load().fx_vec_copy(self.h, dst.h))

    # VEC methods with the reference's argument meaning (src/VEC.hxx:99-145)
    def AXPY(self, a: float, x: "Vec"):
        _chk(load().fx_axpy(self.ctx.h, a, x.h, self.h))

    def XPAY(self, a: float, x: "Vec"):
        _chk(load().fx_xpay(self.ctx.h, a, x.h, self.h))

    def AXPBY(self, a: float, b: float, y: "Vec"):
        _chk(load().fx_axpby(self.ctx.h, a, b, y.h, self.h))

    def WAXPBY(self, a: float, x: "Vec", b: float, y: "Vec"):
        _chk(load().fx_waxpby(self.ctx.h, a, x.h, b, y.h, self.h))

    def Scale(self, a: float):
        _chk(load().fx_scale(self.ctx.h, a, self.h))

    def PointwiseMult(self, v: "Vec"):
        _chk(load().fx_pmul(self.ctx.h, v.h, self.h))

    def Reciprocal(self):
        _chk(load().fx_recip(self.ctx.h, self.h))

    def __iadd__(self, v: "Vec"):
        _chk(load().fx_add(self.ctx.h, v.h, self.h))
        return self

    def Shift(self, a: float):
        _chk(load().fx_shift(self.ctx.h, a, self.h))

    def Abs(self):
        _chk(load().fx_abs(self.ctx.h, self.h))

    def PointwiseDivide(self, v: "Vec"):
        _chk(load().fx_pdiv(self.ctx.h, v.h, self.h))

    def __isub__(self, v: "Vec"):
        _chk(load().fx_sub(self.ctx.h, v.h, self.h))
        return self

    def Max(self) -> float:
        out = C.c_double()
        _chk(load().fx_max(self.ctx.h, self.h, C.byref(out)))
        return out.value

    def Min(self) -> float:
        out = C.c_double()
        _chk(load().fx_min(self.ctx.h, self.h, C.byref(out)))
        return out.value

    def Dot(self, v: "Vec") -> float:
        out = C.c_double()
        _chk(load().fx_dot(self.ctx.h, self.h, v.h, C.byref(out)))
        return out.value

    def Norm2(self) -> float:
        out = C.c_double()
        _chk(load().fx_norm2(self.ctx.h, self.h, C.byref(out)))
        return out.value

    def NormInf(self) -> float:
        out = C.c_double()
        _chk(load().fx_norminf(self.ctx.h, self.h, C.byref(out)))
        return out.value

    # fused forms (no reference counterpart: one pass instead of two)
    def AXPY_Norm2(self, a: float, x: "Vec") -> float:
        """this += a x, then ||this||."""
        out = C.c_double()
        _chk(load().fx_axpy_norm2(self.ctx.h, a, x.h, self.h, C.byref(out)))
        return out.value

    def XPAY_Norm2(self, a: float, x: "Vec") -> float:
        """this = a this + x, then ||this||."""
        out = C.c_double()
        _chk(load().fx_xpay_norm2(self.ctx.h, a, x.h, self.h, C.byref(out)))
        return out.value

    def AXPY_Dot(self, a: float, x: "Vec", z: "Vec") -> float:
        """this += a x, then (this, z)."""
        out = C.c_double()
        _chk(load().fx_axpy_dot(self.ctx.h, a, x.h, self.h, z.h, C.byref(out)))
        return out.value

    def close(self):
        if self.h:
            load().fx_vec_destroy(self.h)
            self.h = None

    def __del__(self):
        if not _exiting:
            try:
                self.close()
            except Exception:
                pass


class Csr:
    """CSRx matrix in HBM (MAT)."""

    def __init__(self, ctx: Context, nrow: int, mcol: int, row_ptr, col_ind, values=None,
                 diag_ptr=None):
        lib = load()
        self._rp = np.ascontiguousarray(row_ptr, dtype=np.uint32)
        self._ci = np.ascontiguousarray(col_ind, dtype=np.uint32)
        self._va = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
        self._dp = None if diag_ptr is None else np.ascontiguousarray(diag_ptr, dtype=np.uint32)
        nnz = self._ci.shape[0]
        h = C.c_void_p()
        _chk(lib.fx_csr_create(ctx.h, nrow, mcol, nnz, _ptr(self._rp), _ptr(self._ci),
                               None if self._va is None else _ptr(self._va),
                               None if self._dp is None else _ptr(self._dp), C.byref(h)))
        self.h, self.ctx = h, ctx
        ctx._adopt(self)
        self.nrow, self.mcol, self.nnz = nrow, mcol, nnz
        # host copies are not needed after upload
        self._rp = self._ci = self._va = self._dp = None

    @classmethod
    def from_problem(cls, ctx: Context, A) -> "Csr":
        return cls(ctx, A.nrow, A.mcol, A.row_ptr, A.col_ind, A.values, A.diag_ptr)

    @classmethod
    def _from_handle(cls, ctx: Context, h) -> "Csr":
        obj = cls.__new__(cls)
        obj.h, obj.ctx = h, ctx
        ctx._adopt(obj)
        nr, mc, nz = C.c_uint32(), C.c_uint32(), C.c_uint32()
        _chk(load().fx_csr_info(h, C.byref(nr), C.byref(mc), C.byref(nz)))
        obj.nrow, obj.mcol, obj.nnz = nr.value, mc.value, nz.value
        return obj

    def download(self):
        rp = np.empty(self.nrow + 1, np.uint32)
        ci = np.empty(self.nnz, np.uint32)
        va = np.empty(self.nnz, np.float64)
        dp = np.empty(min(self.nrow, self.mcol), np.uint32)
        _chk(load().fx_csr_download(self.h, _ptr(rp), _ptr(ci), _ptr(va), _ptr(dp)))
        return rp, ci, va, dp

    def make_f32(self):
        _chk(load().fx_csr_make_f32(self.h))

    @property
    def index_bytes(self) -> int:
        """Bytes per column index read by the SpMV (2: 16-bit block deltas)."""
        b = C.c_int()
        _chk(load().fx_csr_index_bytes(self.h, C.byref(b)))
        return b.value

    @property
    def layout(self) -> str:
        """How the SpMV reads the matrix: 'u32', 'd16', 'pattern', 'sliced' or
        'sorted'."""
        b = C.c_int()
        _chk(load().fx_csr_layout(self.h, C.byref(b)))
        return ("u32", "d16", "pattern", "sliced", "sorted")[b.value]

    @property
    def sliced(self) -> bool:
        return self.layout == "sliced"

    def Apply(self, v: Vec, w: Vec):
        _chk(load().fx_spmv(self.ctx.h, self.h, v.h, w.h))

    def apply_timed(self, v: Vec, w: Vec) -> float:
        """Apply, timed by events attached to the kernel's dispatch (ms)."""
        ms = C.c_float()
        _chk(load().fx_spmv_timed(self.ctx.h, self.h, v.h, w.h, C.byref(ms)))
        return ms.value

    def Residual(self, b: Vec, x: Vec, r: Vec):
        _chk(load().fx_residual(self.ctx.h, self.h, b.h, x.h, r.h))

    def GetDiag(self, v: Vec):
        _chk(load().fx_csr_get_diag(self.ctx.h, self.h, v.h))

    def transpose(self) -> "Csr":
        h = C.c_void_p()
        _chk(load().fx_csr_transpose(self.ctx.h, self.h, C.byref(h)))
        return Csr._from_handle(self.ctx, h)

    def lower_tri(self) -> "Csr":
        """MAT::GetLowerTri (src/MAT.cxx:290-316)."""
        h = C.c_void_p()
        _chk(load().fx_csr_lower_tri(self.ctx.h, self.h, C.byref(h)))
        return Csr._from_handle(self.ctx, h)

    def upper_tri(self) -> "Csr":
        """MAT::GetUpperTri (src/MAT.cxx:319-345)."""
        h = C.c_void_p()
        _chk(load().fx_csr_upper_tri(self.ctx.h, self.h, C.byref(h)))
        return Csr._from_handle(self.ctx, h)

    @staticmethod
    def add(a: float, A: "Csr", b: float, B: "Csr") -> "Csr":
        """a A + b B (MAT::Add, src/MAT.cxx:652-719)."""
        h = C.c_void_p()
        _chk(load().fx_csr_add(A.ctx.h, a, A.h, b, B.h, C.byref(h)))
        return Csr._from_handle(A.ctx, h)

    @staticmethod
    def mult(L: "Csr", R: "Csr") -> "Csr":
        """L R (MAT::Mult, src/MAT.cxx:722-800)."""
        h = C.c_void_p()
        _chk(load().fx_csr_mult(L.ctx.h, L.h, R.h, C.byref(h)))
        return Csr._from_handle(L.ctx, h)

    def shift_diag(self, a: float):
        """values[diagPtr[j]] += a (MAT::Shift, src/MAT.cxx:359-364)."""
        _chk(load().fx_csr_shift_diag(self.ctx.h, self.h, a))

    def transpose_update(self, At: "Csr"):
        """Refresh At's values from this matrix (At = self.transpose())."""
        _chk(load().fx_csr_transpose_update(self.ctx.h, self.h, At.h))

    def pattern_grad(self, lam: Vec, x: Vec, out_ptr: int):
        """out_k = -lam[row(k)] * x[col(k)] (device pointer, nnz doubles)."""
        _chk(load().fx_csr_pattern_grad(self.ctx.h, self.h, lam.h, x.h, C.c_void_p(out_ptr)))

    def set_values_dev(self, dev_ptr: int):
        _chk(load().fx_csr_set_values_dev(self.h, C.c_void_p(dev_ptr)))

    @property
    def values_ptr(self) -> int:
        return load().fx_csr_values_dev(self.h) or 0

    def close(self):
        if self.h:
            load().fx_csr_destroy(self.h)
            self.h = None

    def __del__(self):
        if not _exiting:
            try:
                self.close()
            except Exception:
                pass


class Precond:
    def __init__(self, ctx: Context, kind: str = "identity", A: Csr | None = None,
                 weight: float = 1.0, iters: int = 1):
        lib = load()
        h = C.c_void_p()
        if kind == "identity":
            _chk(lib.fx_pc_create_identity(ctx.h, C.byref(h)))
        elif kind == "jacobi":
            _chk(lib.fx_pc_create_jacobi(ctx.h, A.h, weight, iters, C.byref(h)))
        elif kind == "mcgs":  # symmetric multicolor Gauss-Seidel, `iters` sweeps
            _chk(lib.fx_pc_create_mcgs(ctx.h, A.h, weight, iters, C.byref(h)))
        else:
            raise ValueError(kind)
        self.h, self.ctx, self.A = h, ctx, A
        ctx._adopt(self)

    def Solve(self, b: Vec, x: Vec):
        _chk(load().fx_pc_apply(self.ctx.h, self.h, b.h, x.h))

    def SolveNoAlloc(self, b: Vec, x: Vec):
        """Solve without device allocation (fx_pc_apply_noalloc): f64,
        identity or Jacobi, at most 4,194,304 rows."""
        _chk(load().fx_pc_apply_noalloc(self.ctx.h, self.h, b.h, x.h))

    def close(self):
        if self.h:
            load().fx_pc_destroy(self.h)
            self.h = None

    def __del__(self):
        if not _exiting:
            try:
                self.close()
            except Exception:
                pass


class MCGS:
    """Multicolor Gauss-Seidel smoother (fx_mcgs_*)."""

    def __init__(self, ctx: Context, A: Csr, weight: float = 1.0):
        h = C.c_void_p()
        _chk(load().fx_mcgs_create(ctx.h, A.h, weight, C.byref(h)))
        self.h, self.ctx, self.A = h, ctx, A
        ctx._adopt(self)

    def colors(self):
        nc = C.c_uint32()
        col = np.empty(self.A.nrow, np.uint32)
        _chk(load().fx_mcgs_info(self.h, C.byref(nc), _ptr(col)))
        return nc.value, col

    def sweep(self, b: Vec, x: Vec, sweeps: int = 1, symmetric: bool = False):
        _chk(load().fx_mcgs_sweep(self.h, b.h, x.h, sweeps, 1 if symmetric else 0))

    def close(self):
        if self.h:
            load().fx_mcgs_destroy(self.h)
            self.h = None

    def __del__(self):
        if not _exiting:
            try:
                self.close()
            except Exception:
                pass


_APPLY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)


class _View(Vec):
    """A borrowed device vector handed to an Operator's apply (not owned)."""

    def __init__(self, ctx, h, n):  # noqa: D401 -- no allocation
        self.h, self.ctx, self.n, self.dtype = C.c_void_p(h), ctx, n, F64

    def close(self):
        self.h = None

    def __del__(self):
        pass


class Operator:
    """A linear operator given by its action, LOP with a user Apply
    (src/LOP.hxx:56-59): apply(x: Vec, y: Vec) computes y = A x on the
    context's stream.  An exception in apply becomes ERROR_INPUT_PAR.

    device_only=True declares that apply only enqueues device work on the
    context's stream, the same work on every call (fx_op_set_device_only):
    the solvers then run their fused form and capture apply's work in their
    replay graphs."""

    def __init__(self, ctx: Context, n: int, apply, device_only: bool = False):
        self.ctx, self.n, self._apply = ctx, n, apply
        self.error = None

        def tramp(user, x, y):
            try:
                apply(_View(ctx, x, n), _View(ctx, y, n))
                return 0
            except Exception as e:  # reported through the return code
                self.error = e
                return -12

        self._fn = _APPLY_FN(tramp)
        h = C.c_void_p()
        _chk(load().fx_op_create(ctx.h, n, n, C.cast(self._fn, C.c_void_p), None, C.byref(h)))
        self.h = h
        if device_only:
            _chk(load().fx_op_set_device_only(h, 1))
        ctx._adopt(self)

    def close(self):
        if self.h:
            load().fx_op_destroy(self.h)
            self.h = None

    def __del__(self):
        if not _exiting:
            self.close()


class Solver:
    def __init__(self, ctx: Context, kind: int):
        h = C.c_void_p()
        _chk(load().fx_solver_create(ctx.h, kind, C.byref(h)))
        self.h, self.ctx, self.kind = h, ctx, kind
        self._keep = ()
        ctx._adopt(self)

    def Setup(self, A, pc: Precond, prm: Params):
        """A: a Csr, or an Operator (its apply runs where the reference calls
        A->Apply; the per-operation solver form)."""
        self._keep = (A, pc)
        if isinstance(A, Operator):
            _chk(load().fx_solver_setup_op(self.h, A.h, pc.h, C.byref(prm)))
        else:
            _chk(load().fx_solver_setup(self.h, A.h, pc.h, C.byref(prm)))

    def Update(self, A: "Csr", pc: Precond, prm: Params):
        """Bind A, pc and params again without a new Setup: GMRES / FGMRES
        keep their restart length (and FGMRES its Z)."""
        self._keep = (A, pc)
        _chk(load().fx_solver_update(self.h, A.h, pc.h, C.byref(prm)))

    def Clean(self):
        _chk(load().fx_solver_clean(self.h))

    def Solve(self, b: Vec, x: Vec) -> Result:
        """Returns the Result; retcode carries FaspRetCode (negative codes such
        as ERROR_SOLVER_MAXIT are results, not exceptions)."""
        res = Result()
        rc = load().fx_solver_solve(self.h, b.h, x.h, C.byref(res))
        if rc in (-90, -12, -16, -20):
            _chk(rc)
        return res

    @property
    def last_path(self) -> str:
        """'fused' (graph-replayed fused kernels) or 'ops' (per-operation
        form): the form the last Solve ran."""
        p = C.c_int()
        _chk(load().fx_solver_last_path(self.h, C.byref(p)))
        return "fused" if p.value == 1 else "ops"

    def history(self) -> np.ndarray:
        cnt = C.c_uint32()
        load().fx_solver_history(self.h, None, 0, C.byref(cnt))
        out = np.zeros(cnt.value, np.float64)
        if cnt.value:
            _chk(load().fx_solver_history(self.h, _ptr(out), cnt.value, C.byref(cnt)))
        return out

    def close(self):
        if self.h:
            load().fx_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        if not _exiting:
            try:
                self.close()
            except Exception:
                pass


def krylov(ctx: Context, A: Csr, b: Vec, x: Vec, pc: Precond, name: str, prm: Params) -> Result:
    res = Result()
    rc = load().fx_krylov(ctx.h, A.h, b.h, x.h, pc.h, name.encode(), C.byref(prm), C.byref(res))
    if rc in (-90, -12, -16, -20):
        _chk(rc)
    return res


def solve_ir(ctx: Context, A: Csr, b: Vec, x: Vec, weight: float, outer: Params,
             inner: Params, inner_kind: int = CG) -> tuple[Result, int]:
    res = Result()
    it = C.c_uint32()
    rc = load().fx_solve_ir(ctx.h, inner_kind, A.h, weight, C.byref(outer), C.byref(inner),
                            b.h, x.h, C.byref(res), C.byref(it))
    if rc in (-90, -12, -16, -20):
        _chk(rc)
    return res, it.value


class IR:
    """Reusable mixed-precision refinement solver (fx_ir_*)."""

    def __init__(self, ctx: Context, A: Csr, weight: float, outer: Params, inner: Params,
                 inner_kind: int = CG):
        h = C.c_void_p()
        _chk(load().fx_ir_create(ctx.h, inner_kind, A.h, weight, C.byref(outer), C.byref(inner),
                                 C.byref(h)))
        self.h, self.ctx, self._keep = h, ctx, (A,)
        ctx._adopt(self)

    def Solve(self, b: Vec, x: Vec) -> tuple[Result, int]:
        res = Result()
        it = C.c_uint32()
        rc = load().fx_ir_solve(self.h, b.h, x.h, C.byref(res), C.byref(it))
        if rc in (-90, -12, -16, -20):
            _chk(rc)
        return res, it.value

    def close(self):
        if self.h:
            load().fx_ir_destroy(self.h)
            self.h = None

    def __del__(self):
        if not _exiting:
            try:
                self.close()
            except Exception:
                pass
