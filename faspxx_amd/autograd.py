"""Differentiable sparse solve: x = A^{-1} b with gradients for b and for A's
stored values, on the gfx950 solvers (no reference counterpart: FASP++ has no
Python layer).

Forward: the device Krylov solver on A.  Backward, given g = dL/dx: one
adjoint solve A^T lam = g with the same solver over the CSR transpose
(formed once on the device, values refreshed from A's values), then
    dL/db = lam,        dL/dA_ij = -lam_i x_j  on A's sparsity pattern only
(one gather kernel).  The solve tolerance bounds the gradient's accuracy.

    pattern = SparsePattern(ctx, A)              # structure + transpose
    x = sparse_solve(values, b, pattern)          # values: (nnz,) f64 cuda
"""
from __future__ import annotations

import contextlib
import dataclasses

import torch

from . import fx


def _ctx_stream(ctx: fx.Context) -> torch.cuda.Stream:
    """The library context's HIP stream as a torch stream object."""
    s = ctx.stream
    if s == 0:
        return torch.cuda.default_stream(ctx.device)
    return torch.cuda.ExternalStream(s, device=torch.device("cuda", ctx.device))


@contextlib.contextmanager
def _ordered(ctx: fx.Context):
    """Order library work on the context's stream after what torch's current
    stream has queued (the inputs written there), and torch's current stream
    after the library work (the outputs it returns).  Tensors used inside
    must be allocated before entering, so that the wait covers their
    initialisation; freeing them afterwards is ordered by the second wait."""
    cur = torch.cuda.current_stream(ctx.device)
    fs = _ctx_stream(ctx)
    same = fs.cuda_stream == cur.cuda_stream
    if not same:
        fs.wait_stream(cur)
    try:
        yield
    finally:
        if not same:
            cur.wait_stream(fs)


@dataclasses.dataclass
class SolveOptions:
    """Solver for the forward and adjoint systems (FX solver type, Jacobi
    weight; weight 0 disables the preconditioner) and its SOLParams."""
    kind: int = fx.CG
    weight: float = 1.0
    maxIter: int = 20000
    relTol: float = 1e-10
    absTol: float = 0.0
    restart: int = 30


class SparsePattern:
    """A fixed sparsity pattern on the device: the CSR matrix A (values set per
    call) and its transpose, with one solver and preconditioner for each."""

    def __init__(self, ctx: fx.Context, A, opts: SolveOptions | None = None):
        self.ctx = ctx
        self.opts = opts or SolveOptions()
        self.A = fx.Csr(ctx, A.nrow, A.mcol, A.row_ptr, A.col_ind, A.values, A.diag_ptr)
        self.At = self.A.transpose()
        self.n = A.nrow
        self.nnz = A.col_ind.shape[0]
        self._solvers = {}

    def _solve(self, M: fx.Csr, b: torch.Tensor, tag: str) -> tuple[torch.Tensor, fx.Result]:
        o = self.opts
        key = tag
        if key not in self._solvers:
            if o.weight > 0:
                pc = fx.Precond(self.ctx, "jacobi", M, o.weight, 1)
            else:
                pc = fx.Precond(self.ctx, "identity")
            s = fx.Solver(self.ctx, o.kind)
            self._solvers[key] = (s, pc)
        s, pc = self._solvers[key]
        s.Setup(M, pc, fx.params(maxIter=o.maxIter, relTol=o.relTol, absTol=o.absTol,
                                 restart=o.restart))
        b = b.contiguous()
        x = torch.zeros_like(b)
        with _ordered(self.ctx):
            vb = fx.Vec.wrap(self.ctx, b.data_ptr(), self.n)
            vx = fx.Vec.wrap(self.ctx, x.data_ptr(), self.n)
            res = s.Solve(vb, vx)
            vb.close()
            vx.close()
        return x, res

    def set_values(self, values: torch.Tensor):
        values = values.contiguous()
        with _ordered(self.ctx):
            self.A.set_values_dev(values.data_ptr())
            self.A.transpose_update(self.At)
        # the Jacobi diagonals depend on the values: rebuild the solvers
        self._solvers.clear()

    def solve(self, values: torch.Tensor, b: torch.Tensor):
        self.set_values(values)
        return self._solve(self.A, b, "A")

    def adjoint(self, g: torch.Tensor):
        return self._solve(self.At, g, "At")

    def pattern_grad(self, lam: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        out = torch.empty(self.nnz, dtype=torch.float64, device=lam.device)
        with _ordered(self.ctx):
            vl = fx.Vec.wrap(self.ctx, lam.data_ptr(), self.n)
            vx = fx.Vec.wrap(self.ctx, x.data_ptr(), self.n)
            self.A.pattern_grad(vl, vx, out.data_ptr())
            vl.close()
            vx.close()
        return out


def _check(t: torch.Tensor, n: int, name: str):
    if t.dtype != torch.float64 or not t.is_cuda or t.dim() != 1 or t.shape[0] != n:
        raise ValueError(f"{name}: expected a 1-D float64 CUDA tensor of length {n}")


class SparseSolve(torch.autograd.Function):
    """x = A(values)^{-1} b; differentiable in values and b."""

    @staticmethod
    def forward(ctx, values, b, pattern: SparsePattern):
        _check(values, pattern.nnz, "values")
        _check(b, pattern.n, "b")
        x, res = pattern.solve(values.detach(), b.detach())
        if res.retcode != 0:
            raise RuntimeError(f"forward solve failed: {fx.RETCODES.get(res.retcode)}")
        ctx.pattern = pattern
        ctx.save_for_backward(values, x)
        return x

    @staticmethod
    def backward(ctx, g):
        values, x = ctx.saved_tensors
        pattern: SparsePattern = ctx.pattern
        pattern.set_values(values.detach())
        lam, res = pattern.adjoint(g.detach().contiguous())
        if res.retcode != 0:
            raise RuntimeError(f"adjoint solve failed: {fx.RETCODES.get(res.retcode)}")
        grad_values = pattern.pattern_grad(lam, x) if ctx.needs_input_grad[0] else None
        grad_b = lam if ctx.needs_input_grad[1] else None
        return grad_values, grad_b, None


def sparse_solve(values: torch.Tensor, b: torch.Tensor, pattern: SparsePattern) -> torch.Tensor:
    return SparseSolve.apply(values, b, pattern)
