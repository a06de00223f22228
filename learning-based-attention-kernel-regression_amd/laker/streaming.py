"""Warm-started refits: online updates and regularisation paths (reference
laker/streaming.py:44-379).

``StreamingUpdater`` keeps the reference's three entry points and their
arguments, attributes and errors:

* ``partial_fit`` appends new observations, rebuilds the operator and the
  preconditioner, and re-solves from ``[forgetting_factor * alpha, 0]``;
* ``fit_path`` solves for a list of lambdas from the largest to the smallest,
  each solve warm-started from the previous solution, optionally reusing the
  first preconditioner;
* ``fit_continuation`` builds a geometric lambda schedule and calls
  ``fit_path``.

Every product, preconditioner build/apply and PCG loop is the library's:
a new lambda only changes the operator's diagonal shift, so the packed
embedding operands are reused (``laker_operator_set_lambda``) instead of
being re-packed for every stage.  The final relative residuals are computed
like the reference's, with the operator's product in the solver dtype, and
normed with the library's FP64 column reductions.
"""

from __future__ import annotations

import logging
from typing import Optional

import torch

from .solvers import _coldot

logger = logging.getLogger(__name__)


def _norm(v: torch.Tensor) -> float:
    return float(_coldot(v, v).sum()) ** 0.5


class StreamingUpdater:
    """Incremental updates and continuation-path fitting on a :class:`LAKERCore`."""

    def __init__(self, core) -> None:
        self.core = core

    # ------------------------------------------------------------------
    def partial_fit(self, regressor, x_new: torch.Tensor, y_new: torch.Tensor, forgetting_factor: float = 1.0,
                    rebuild_threshold: int = 100):
        """Append observations and re-solve warm-started (laker/streaming.py:75-187)."""
        if regressor.alpha is None or regressor.embeddings is None:
            raise RuntimeError("Model has not been fitted. Call fit() before partial_fit().")
        if x_new.dim() != 2:
            raise ValueError(f"x_new must be 2-D, got shape {x_new.shape}")
        if y_new.dim() != 1:
            raise ValueError(f"y_new must be 1-D, got shape {y_new.shape}")
        core = self.core
        m = x_new.shape[0]
        total_new = getattr(regressor, "partial_fit_count", 0) + m
        if total_new >= rebuild_threshold:
            regressor.partial_fit_count = 0
            raise RuntimeError("partial_fit rebuild threshold exceeded. "
                               "Please concatenate all data and call fit() for a full refit.")
        regressor.partial_fit_count = total_new

        emb_dev = regressor.embeddings.device
        with torch.no_grad():
            new_e = regressor.embedding_model(x_new.to(device=emb_dev, dtype=core.embedding_dtype))
        new_e = new_e.to(dtype=regressor.embeddings.dtype)
        old_n = regressor.embeddings.shape[0]
        regressor.embeddings = torch.cat([regressor.embeddings, new_e], dim=0)
        op = core.build_kernel_operator(regressor.embeddings)
        regressor.kernel_operator = op

        alpha_dev = regressor.alpha.device
        old_alpha = regressor.alpha * forgetting_factor
        y_old = getattr(regressor, "y_train", None)
        if y_old is None:
            y_old = torch.zeros(old_n, device=alpha_dev, dtype=regressor.alpha.dtype)
        y_ext = torch.cat([y_old.to(alpha_dev, regressor.alpha.dtype), y_new.to(alpha_dev, regressor.alpha.dtype)])
        x0 = torch.cat([old_alpha, torch.zeros(m, device=alpha_dev, dtype=regressor.alpha.dtype)])

        regressor.preconditioner = core.build_preconditioner(op.matvec, regressor.embeddings.shape[0],
                                                             diagonal=op.diagonal())
        regressor.alpha, regressor.pcg_iterations_ = core.solve_pcg(op, regressor.preconditioner, y_ext, x0=x0)
        solver = getattr(core, "last_solver", None)
        regressor.alpha_fp64 = solver.solution_fp64 if solver is not None else None
        regressor.y_train = y_ext
        if getattr(regressor, "x_train", None) is not None:
            regressor.x_train = torch.cat([regressor.x_train, x_new.to(regressor.x_train.device,
                                                                        regressor.x_train.dtype)], dim=0)
        if core.verbose:
            logger.info("partial_fit: added %d points, total=%d, PCG iters=%d", m,
                        regressor.embeddings.shape[0], regressor.pcg_iterations_)
        return regressor

    # ------------------------------------------------------------------
    def fit_path(self, regressor, x: torch.Tensor, y: torch.Tensor, lambda_reg_grid, reuse_precond: bool = True) -> dict:
        """Warm-started solves over a decreasing lambda grid (laker/streaming.py:189-304)."""
        if x.dim() != 2:
            raise ValueError(f"x must be 2-D, got shape {x.shape}")
        if y.dim() != 1:
            raise ValueError(f"y must be 1-D, got shape {y.shape}")
        if not lambda_reg_grid:
            raise ValueError("lambda_reg_grid must not be empty")
        core = self.core
        embeddings, model = core.compute_embeddings(x)
        n = embeddings.shape[0]
        lambdas = sorted(lambda_reg_grid, reverse=True)
        alphas, iters, rel_res = [], [], []
        x0 = None
        precond = None
        op = None
        for lam in lambdas:
            if op is None:
                op = core.build_kernel_operator(embeddings, lambda_reg=lam)
            else:
                op.lambda_reg = lam  # same packed embeddings, new diagonal shift
            if precond is None or not reuse_precond:
                if precond is not None and hasattr(precond, "close"):
                    precond.close()
                precond = core.build_preconditioner(op.matvec, n, gamma=core.gamma, num_probes=core.num_probes,
                                                    diagonal=op.diagonal())
            alpha, it = core.solve_pcg(op, precond, y, x0=x0)
            alphas.append(alpha)
            iters.append(it)
            yd = y.to(alpha.device, alpha.dtype)
            rel_res.append(_norm(op.matvec(alpha) - yd) / _norm(yd))
            x0 = alpha.clone()

        regressor.embeddings = embeddings
        regressor.embedding_model = model
        final = core.build_kernel_operator(embeddings, lambda_reg=lambdas[-1])
        regressor.kernel_operator = final
        regressor.preconditioner = core.build_preconditioner(final.matvec, n, gamma=core.gamma,
                                                             num_probes=core.num_probes,
                                                             diagonal=final.diagonal())
        regressor.alpha = alphas[-1]
        solver = getattr(core, "last_solver", None)
        regressor.alpha_fp64 = solver.solution_fp64 if solver is not None else None
        regressor.pcg_iterations_ = iters[-1]
        regressor.y_train = y
        path = {"lambda_reg": lambdas, "alphas": alphas, "pcg_iters": iters, "final_rel_res": rel_res}
        regressor.path_ = path
        return path

    # ------------------------------------------------------------------
    def fit_continuation(self, regressor, x: torch.Tensor, y: torch.Tensor, lambda_max: Optional[float] = None,
                         lambda_min: Optional[float] = None, n_stages: int = 5, reuse_precond: bool = True):
        """Geometric lambda schedule from lambda_max to lambda_min, then ``fit_path``
        (laker/streaming.py:306-379)."""
        core = self.core
        if lambda_max is None:
            lambda_max = 10.0 * core.lambda_reg
        if lambda_min is None:
            lambda_min = core.lambda_reg
        if n_stages < 1:
            raise ValueError(f"n_stages must be positive, got {n_stages}")
        if lambda_max <= 0 or lambda_min <= 0:
            raise ValueError("lambda_max and lambda_min must be positive")
        ratio = (lambda_min / lambda_max) ** (1.0 / max(1, n_stages - 1))
        schedule = [lambda_max * ratio ** k for k in range(n_stages)]
        schedule[-1] = lambda_min
        path = self.fit_path(regressor, x, y, lambda_reg_grid=schedule, reuse_precond=reuse_precond)
        regressor.lambda_reg = float(lambda_min)
        regressor.pcg_iterations_ = path["pcg_iters"][-1]
        return regressor
