"""Hypergradients by the adjoint method (reference laker/implicit_diff.py:43-151).

``v`` solves ``A v = dL/dalpha`` with the same PCG (native when the operator
and preconditioner are native); the parameter gradients are
``-d(v^T A(theta) alpha)/dtheta`` by autograd, which for the exact attention
operator runs the hand-written fused backward.
"""

from __future__ import annotations

from typing import Callable, List

import torch

from .solvers import PreconditionedConjugateGradient


def hypergradient(operator_fn: Callable[[torch.Tensor], torch.Tensor],
                  preconditioner_fn: Callable[[torch.Tensor], torch.Tensor], alpha: torch.Tensor,
                  dL_dalpha: torch.Tensor, param_list: List[torch.Tensor], pcg_tol: float = 1e-6,
                  pcg_max_iter: int = 500, verbose: bool = False) -> List[torch.Tensor]:
    pcg = PreconditionedConjugateGradient(tol=pcg_tol, max_iter=pcg_max_iter, verbose=verbose)
    with torch.no_grad():
        v = pcg.solve(operator=operator_fn, preconditioner=preconditioner_fn, rhs=dL_dalpha.detach())
    alpha_c = alpha.detach()
    v_c = v.detach().to(alpha_c.device, alpha_c.dtype)
    with torch.enable_grad():
        a_alpha = operator_fn(alpha_c.clone().requires_grad_(True))
        scalar = torch.dot(v_c.reshape(-1), a_alpha.reshape(-1).to(v_c.dtype))
    grads = []
    for p in param_list:
        if p.requires_grad:
            g = torch.autograd.grad(scalar, p, retain_graph=True, allow_unused=True)[0]
            grads.append(-g if g is not None else torch.zeros_like(p))
        else:
            grads.append(torch.zeros_like(p))
    return grads
