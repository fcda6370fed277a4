"""Differentiable MLMG solve: a torch.autograd.Function over the C ABI.

The reference has no adjoint; this module adds one on top of the MLMG solve of
an ABecLaplacian (amrex_amd.mlmg.Problem):

  forward   phi_i = L(a, b)^{-1} rhs_i for each field i of a batch (MLMG.solve)
  backward  lam_i = L^{-T} dJ/dphi_i (MLMG.solveAdjoint: the same V-cycles with
            the fine residual g - L^T lam and homogeneous boundary data)
            dJ/drhs_i = lam_i
            dJ/da     = sum_i -alpha * lam_i * phi_i
            dJ/db_d   = sum_i -lam_i . d(L phi_i)/db_d   (compCoeffGradient)

With the default closure order 3 the Dirichlet boundary rows make L
non-symmetric, so the adjoint solve uses L^T explicitly rather than reusing the
forward solve.  Tensors are float64 CUDA tensors in the global (z, y, x) layout;
face coefficients of direction d have one more point along d.
"""
from __future__ import annotations

import torch


class MLMGSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rhs, a, bx, by, bz, problem, tol_rel):
        if problem.op != "abec":
            raise ValueError("MLMGSolve: the coefficient gradient needs an ABecLaplacian problem")
        problem.set_coefficients(a, bx, by, bz)
        single = rhs.dim() == 3
        R = rhs.unsqueeze(0) if single else rhs
        phis = torch.stack([problem.solve(R[i].contiguous(), tol_rel) for i in range(R.shape[0])])
        ctx.problem = problem
        ctx.tol = tol_rel
        ctx.single = single
        ctx.coef_version = problem.coef_version
        ctx.has = [t is not None for t in (a, bx, by, bz)]
        ctx.save_for_backward(phis, *[t for t in (a, bx, by, bz) if t is not None])
        return phis[0] if single else phis

    @staticmethod
    def backward(ctx, gphi):
        P = ctx.problem
        phis, *coef = ctx.saved_tensors
        it = iter(coef)
        a, bx, by, bz = [next(it) if h else None for h in ctx.has]
        if P.coef_version != ctx.coef_version:
            P.set_coefficients(a, bx, by, bz)
            ctx.coef_version = P.coef_version
        G = gphi.unsqueeze(0) if ctx.single else gphi
        lams = torch.stack([P.solve_adjoint(G[i].contiguous(), ctx.tol) for i in range(G.shape[0])])
        need = ctx.needs_input_grad
        n = P.n
        ga = torch.zeros((n[2], n[1], n[0]), dtype=torch.float64, device=phis.device) if need[1] else None
        shapes = [(n[2], n[1], n[0] + 1), (n[2], n[1] + 1, n[0]), (n[2] + 1, n[1], n[0])]
        gb = [torch.zeros(s, dtype=torch.float64, device=phis.device) if need[2 + d] else None
              for d, s in enumerate(shapes)]
        if ga is not None or any(g is not None for g in gb):
            for i in range(phis.shape[0]):
                P.coeff_gradient(lams[i], phis[i], ga, gb if any(g is not None for g in gb) else None,
                                 accumulate=i > 0)
        grad_rhs = None
        if need[0]:
            grad_rhs = lams[0] if ctx.single else lams
        return grad_rhs, ga, gb[0], gb[1], gb[2], None, None


def solve(problem, rhs, a=None, bx=None, by=None, bz=None, tol_rel=1e-10):
    """phi = L(a, b)^{-1} rhs, differentiable in rhs, a and the face coefficients."""
    return MLMGSolve.apply(rhs, a, bx, by, bz, problem, tol_rel)
