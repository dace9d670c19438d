"""Differentiable assemble-and-solve ops (torch.autograd.Function semantics).

New surface with no reference counterpart.  Forward: heterogeneous assembly
(per-element log-parameter theta_e), Dirichlet elimination and a CG solve,
all in libgeofem.  Backward (the adjoint method): with L(u(theta)),

    K_bc lambda = dL/du   (dL/du zeroed on Dirichlet DOFs, so lambda = 0 there)
    dL/dtheta_e = -lambda_e^T (dK_e/dtheta_e) u_e = -lambda_e^T K_e u_e

because K_e is linear in k_e (thermal) or E_e (elastic) and theta = log k or
log E.  The eliminated operator is SPD and symmetric (free block plus a
scaled identity on constrained rows), so the reverse solve reuses it; u_e
carries the Dirichlet values, which accounts for the theta-dependence of the
lifted right-hand side.  The loads do not depend on theta.

The reverse solve runs on dL/du scaled to unit norm, with the scale-free
breakdown test (p.Ap == 0 or a non-finite step), and scales lambda back:
CG is invariant under that scaling except for its absolute thresholds (the
FP64 path's |p.Ap| < 1e-30 and abs_tol), which would otherwise depend on the
units of the loss; abs_tol therefore acts relative to ||dL/du||.
"""
from __future__ import annotations

import copy

import torch

from .assembly import (Assembler, IsotropicElasticity, elastic_element_backward,
                       thermal_element_backward)
from .linalg import ConjugateGradient
from .runtime import default_context


def _reverse_solve(problem, K, grad_u):
    """lambda = K^-1 (dL/du with Dirichlet DOFs zeroed), solved on the
    unit-norm right-hand side."""
    g = grad_u.detach().to(dtype=torch.float64).contiguous().clone()
    g[problem.mask] = 0.0
    s = float(torch.linalg.vector_norm(g))
    if s == 0.0 or s != s:
        problem.backward_stats = None
        return g if s == 0.0 else torch.full_like(g, float("nan"))
    solver = copy.copy(problem.solver).with_relative_breakdown(True)
    lam, stats = solver.solve(K, g / s)
    problem.backward_stats = stats
    return lam * s


class ThermalProblem:
    """Steady -div(k grad T) = Q with Dirichlet data from a DofManager."""

    def __init__(self, mesh, dof_mgr, source: float, solver: ConjugateGradient | None = None,
                 ctx=None):
        self.mesh = mesh
        self.dof_mgr = dof_mgr
        self.source = float(source)
        self.solver = solver or ConjugateGradient().with_tolerance(1e-12).with_max_iterations(100000)
        self.ctx = ctx or default_context()
        self.mask = torch.as_tensor(dof_mgr.dirichlet_mask.astype(bool),
                                    device=self.ctx.torch_device())
        self.forward_stats = None
        self.backward_stats = None


class ElasticProblem:
    """Linear elasticity under a body force (gravity) with Dirichlet data."""

    def __init__(self, mesh, dof_mgr, poisson: float, density: float, gravity,
                 solver: ConjugateGradient | None = None, ctx=None):
        self.mesh = mesh
        self.dof_mgr = dof_mgr
        self.poisson = float(poisson)
        self.density = float(density)
        self.gravity = [float(g) for g in gravity]
        self.solver = solver or ConjugateGradient().with_tolerance(1e-12).with_max_iterations(100000)
        self.ctx = ctx or default_context()
        self.mask = torch.as_tensor(dof_mgr.dirichlet_mask.astype(bool),
                                    device=self.ctx.torch_device())
        self.forward_stats = None
        self.backward_stats = None


class ThermalSolve(torch.autograd.Function):
    """T = solve(log_k): per-element log-conductivity -> nodal temperature."""

    @staticmethod
    def forward(ctx, log_k, problem: ThermalProblem):
        k = torch.exp(log_k.detach().to(device=problem.ctx.torch_device(),
                                        dtype=torch.float64)).contiguous()
        K, f = Assembler.assemble_thermal_system(problem.mesh, problem.dof_mgr, 1.0,
                                                 problem.source, k_per_elem=k,
                                                 ctx=problem.ctx)
        T, stats = problem.solver.solve(K, f)
        problem.forward_stats = stats
        ctx.problem = problem
        ctx.K = K
        ctx.in_dtype = log_k.dtype
        ctx.save_for_backward(k, T)
        return T.to(log_k.dtype) if log_k.dtype != torch.float64 else T

    @staticmethod
    def backward(ctx, grad_T):
        problem = ctx.problem
        k, T = ctx.saved_tensors
        lam = _reverse_solve(problem, ctx.K, grad_T)
        grad = thermal_element_backward(problem.mesh, k, T, lam, ctx=problem.ctx)
        return grad.to(ctx.in_dtype), None


class ElasticSolve(torch.autograd.Function):
    """u = solve(log_E): per-element log Young's modulus -> displacements."""

    @staticmethod
    def forward(ctx, log_E, problem: ElasticProblem):
        E = torch.exp(log_E.detach().to(device=problem.ctx.torch_device(),
                                        dtype=torch.float64)).contiguous()
        mat = IsotropicElasticity(1.0, problem.poisson)  # E comes per element
        K, f = Assembler.assemble_elasticity_system(problem.mesh, problem.dof_mgr, mat,
                                                    problem.density, problem.gravity,
                                                    E_per_elem=E, ctx=problem.ctx)
        u, stats = problem.solver.solve(K, f)
        problem.forward_stats = stats
        ctx.problem = problem
        ctx.K = K
        ctx.in_dtype = log_E.dtype
        ctx.save_for_backward(E, u)
        return u.to(log_E.dtype) if log_E.dtype != torch.float64 else u

    @staticmethod
    def backward(ctx, grad_u):
        problem = ctx.problem
        E, u = ctx.saved_tensors
        lam = _reverse_solve(problem, ctx.K, grad_u)
        grad = elastic_element_backward(problem.mesh, E, problem.poisson, u, lam,
                                        ctx=problem.ctx)
        return grad.to(ctx.in_dtype), None


def thermal_solve(log_k, problem: ThermalProblem):
    return ThermalSolve.apply(log_k, problem)


def elastic_solve(log_E, problem: ElasticProblem):
    return ElasticSolve.apply(log_E, problem)
