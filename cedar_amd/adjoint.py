"""Differentiated 2D diffusion solves: u = A(k)^{-1} f with k = exp(theta) per cell, as a
torch.autograd.Function.

Forward: the 9-point bilinear finite-element operator is assembled on the GPU from k
(cdr_hip_q1_assemble2d_f64) and solved with the batched multigrid solver (mixed precision: an
FP32 V-cycle inside FP64 iterative refinement, or plain FP64).  Backward: A is symmetric, so the
adjoint lambda = A^{-1} dL/du reuses the same hierarchy; the per-cell gradient
dL/dk_e = -lambda_e^T K u_e (element matrix k_e K) comes from a hand-written kernel (cdr_hip_q1_gradient2d_f64),
and dL/dtheta = k dL/dk, dL/df = lambda.
"""
import torch

from . import _abi
from .cdr2 import Solver, _p, _stream


def assemble(k):
    """Stencils (nb, 5, ny+2, nx+2) from per-cell coefficients k (nb, ny+1, nx+1)."""
    k = k.contiguous()
    nb, ncy, ncx = k.shape
    so = torch.empty((nb, 5, ncy + 1, ncx + 1), dtype=torch.float64, device=k.device)
    _abi.call("cdr_hip_q1_assemble2d_f64", _p(k), _p(so), ncx - 1, ncy - 1, nb, _stream())
    return so


def cell_gradient(u, lam):
    """dL/dk per cell (nb, ny+1, nx+1) from state u and adjoint lam (nb, ny+2, nx+2)."""
    nb, jj, ii = u.shape
    g = torch.empty((nb, jj - 1, ii - 1), dtype=torch.float64, device=u.device)
    _abi.call("cdr_hip_q1_gradient2d_f64", _p(u.contiguous()), _p(lam.contiguous()), _p(g), ii - 2,
              jj - 2, nb, _stream())
    return g


def _interior(t):
    m = torch.zeros_like(t)
    m[..., 1:-1, 1:-1] = t[..., 1:-1, 1:-1]
    return m


class DiffusionSolve(torch.autograd.Function):
    """u = solve(theta, f): theta (nb, ny+1, nx+1) log-coefficients per cell, f (nb, ny+2, nx+2)."""

    @staticmethod
    def forward(ctx, theta, f, solver_kw):
        k = torch.exp(theta.detach().to(torch.float64))
        so = assemble(k)
        s = Solver(so, **solver_kw)
        u, hist = s.solve(f.detach().to(torch.float64).contiguous())
        ctx.solver, ctx.hist = s, hist
        ctx.save_for_backward(u, k)
        return u

    @staticmethod
    def backward(ctx, gu):
        u, k = ctx.saved_tensors
        lam, _ = ctx.solver.solve(_interior(gu.to(torch.float64)).contiguous())
        gk = cell_gradient(u, lam)
        return gk * k, lam, None


def solve(theta, f, **solver_kw):
    kw = dict(relaxation="line-xy", nrelax_pre=1, nrelax_post=1, maxiter=100, tol=1e-10,
              precision="mixed")
    kw.update(solver_kw)
    return DiffusionSolve.apply(theta, f, kw)
