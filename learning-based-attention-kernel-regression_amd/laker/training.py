"""Embedding training through the hand-written matvec backward.

``fit_learned_embeddings`` restates laker/training.py:93-259 (Adam on
``1/2 ||K(E(theta)) alpha - y||^2``, alpha re-solved every ``rebuild_freq``
epochs, best-snapshot restore).  Gradients reach the embedding MLP through
``AttentionKernelOperator.matvec``'s autograd function, whose backward is the
fused HIP kernel that recomputes Gram tiles (nothing n x n is stored).

``fit_validation_embeddings`` trains on the validation MSE of the PCG
solution: the direct term flows through the cross-kernel backward, the
implicit term ``-v^T (dA/dtheta) alpha`` uses one adjoint PCG solve
``A v = dL/dalpha`` (the structure of laker/implicit_diff.py:110-151).
"""

from __future__ import annotations

import logging
from typing import Optional

import torch

from .backend import compute_device
from .core import cross_kernel_matvec

logger = logging.getLogger(__name__)


class EmbeddingTrainer:
    def __init__(self, core):
        self.core = core

    def _embed(self, model, x):
        e = model(x.to(device=compute_device(), dtype=self.core.embedding_dtype))
        return e.to(dtype=self.core.dtype) if e.dtype != self.core.dtype else e

    def fit_learned_embeddings(self, regressor, x, y, lr=1e-3, epochs=50, rebuild_freq=10, patience=5):
        if x.dim() != 2:
            raise ValueError(f"x must be 2-D, got shape {x.shape}")
        if y.dim() != 1:
            raise ValueError(f"y must be 1-D, got shape {y.shape}")
        model = regressor.embedding_model
        if model is None:
            raise RuntimeError("No embedding model available.")
        trainable = [p for p in model.parameters() if p.requires_grad]
        if not trainable:
            for p in model.parameters():
                p.requires_grad = True
            trainable = list(model.parameters())
        opt = torch.optim.Adam(trainable, lr=lr)
        yd = y.to(compute_device(), self.core.dtype)
        best, bad = float("inf"), 0
        best_state = None
        alpha = regressor.alpha
        precond = regressor.preconditioner
        for epoch in range(epochs):
            opt.zero_grad()
            emb = self._embed(model, x)
            op = self.core.build_kernel_operator(emb)
            if epoch % rebuild_freq == 0 or alpha is None:
                with torch.no_grad():
                    op_d = self.core.build_kernel_operator(emb.detach())
                    precond = self.core.build_preconditioner(op_d.matvec, emb.shape[0])
                    alpha = self.core.solve_pcg(op_d, precond, yd)[0]
                regressor.preconditioner = precond
            alpha_c = alpha.detach().to(compute_device(), self.core.dtype)
            residual = op.matvec(alpha_c) - yd
            loss = 0.5 * torch.dot(residual, residual)
            loss.backward()
            opt.step()
            li = loss.item()
            if self.core.verbose and (epoch + 1) % 10 == 0:
                logger.info("Learned embeddings epoch %d/%d, loss=%.4e", epoch + 1, epochs, li)
            if li < best:
                best, bad = li, 0
                best_state = (emb.detach().clone(), alpha_c.clone(), precond)
            else:
                bad += 1
                if bad >= patience:
                    break
        emb, alpha, precond = best_state
        op = self.core.build_kernel_operator(emb)
        regressor.embeddings = emb
        regressor.kernel_operator = op
        regressor.preconditioner = precond
        regressor.alpha, regressor.pcg_iterations_ = self.core.solve_pcg(op, precond, yd)
        solver = getattr(self.core, "last_solver", None)
        regressor.alpha_fp64 = solver.solution_fp64 if solver is not None else None
        return regressor

    def validation_step(self, model, x_tr, y_tr, x_val, y_val, precond=None, rebuild: bool = True):
        """One evaluation of the validation MSE and its gradient (accumulated into ``model``).

        Returns ``(loss, precond, alpha, iterations)``.
        """
        y_tr = y_tr.to(compute_device(), self.core.dtype)
        y_val = y_val.to(compute_device(), self.core.dtype)
        emb = self._embed(model, x_tr)
        emb_q = self._embed(model, x_val)
        op = self.core.build_kernel_operator(emb)
        with torch.no_grad():
            if rebuild or precond is None:
                precond = self.core.build_preconditioner(op.matvec, emb.shape[0])
            alpha, iters = self.core.solve_pcg(op, precond, y_tr)
        alpha = alpha.detach().to(compute_device(), self.core.dtype)
        pred = cross_kernel_matvec(op, emb_q, alpha, embeddings=emb)
        r = pred - y_val
        m = y_val.shape[0]
        loss_direct = torch.dot(r, r) / m
        with torch.no_grad():
            g_alpha = op.cross_matvec_t(emb_q.detach(), (2.0 / m) * r.detach())
            v, _ = self.core.solve_pcg(op, precond, g_alpha)
            v = v.detach().to(compute_device(), self.core.dtype)
        # implicit term: d/dtheta of -v^T A(theta) alpha with v, alpha held fixed
        implicit = -torch.dot(v, op.matvec(alpha))
        total = loss_direct + implicit
        total.backward()
        return loss_direct.item(), precond, alpha, iters

    def fit_validation_embeddings(self, regressor, x_tr, y_tr, x_val, y_val, lr=1e-3, epochs=50,
                                  rebuild_freq: int = 1, history: Optional[list] = None):
        model = regressor.embedding_model if regressor.embedding_model is not None else self.core.embedding_module
        if model is None:
            raise RuntimeError("No embedding model available.")
        model = model.to(compute_device())
        for p in model.parameters():
            p.requires_grad_(True)
        opt = torch.optim.Adam(model.parameters(), lr=lr)
        precond = None
        for step in range(epochs):
            opt.zero_grad()
            loss, precond, alpha, iters = self.validation_step(model, x_tr, y_tr, x_val, y_val, precond,
                                                               rebuild=(step % rebuild_freq == 0))
            opt.step()
            if history is not None:
                history.append((loss, iters))
        regressor.embedding_model = model
        regressor.fit(x_tr, y_tr) if regressor.embedding_module is None else None
        return regressor
