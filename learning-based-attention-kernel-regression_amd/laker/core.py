"""LAKER pipeline on the HIP solver core (reference laker/core.py:73-777).

``LAKERCore`` keeps the reference's hyperparameters and method signatures:
``compute_embeddings`` -> ``build_kernel_operator`` -> ``build_preconditioner``
-> ``solve_pcg`` -> ``predict`` / ``predict_variance`` / ``predict_train``.
Extra hyperparameters: ``precision`` (``"fp32"`` plain PCG, ``"mixed"`` split-Gram
inner PCG on 16-bit MFMA + FP64 refinement), ``probe_source`` and ``probe_precision``
(CCCP probes), ``inner_tol``.
"""

from __future__ import annotations

import logging
from typing import Callable, Optional, Union

import torch
import torch.nn as nn

from . import _native as N
from .backend import compute_device, get_default_device, get_default_dtype
from .distributed_kernels import DistributedAttentionKernelOperator
from .embeddings import PositionEmbedding
from .kernels import AttentionKernelOperator, KernelOperator
from .preconditioner import AdaptivePreconditioner, CCCPPreconditioner
from .solvers import PreconditionedConjugateGradient, _coldot

logger = logging.getLogger(__name__)


def auto_chunk(n: int, chunk_size: Optional[int]) -> Optional[int]:
    """``n > 5000 -> max(1024, min(n // 10, 8192))`` (laker/core.py:340-345)."""
    if chunk_size is None and n > 5000:
        return max(1024, min(n // 10, 8192))
    return chunk_size


def native_operator(op):
    """The HIP-backed operator behind ``op`` (the inner operator of the
    single-device distributed wrapper), or ``op`` itself."""
    if isinstance(op, DistributedAttentionKernelOperator):
        return op.local_op
    return op


class _CrossFn(torch.autograd.Function):
    """``yq = exp(min(s Eq E^T, c)) alpha`` with the hand-written backward."""

    @staticmethod
    def forward(ctx, q_emb, emb, alpha2, op, policy):
        ctx.op, ctx.policy = op, policy
        q = q_emb.detach().to(op.dtype).contiguous()
        a2 = alpha2.detach().to(op.dtype).contiguous()
        ctx.save_for_backward(q, a2)
        y = torch.empty(q.shape[0], a2.shape[1], dtype=op.dtype, device=q.device)
        N.check(N.lib.laker_cross_matvec(op.handle, N.ptr(q), q.shape[0], N.ptr(a2), a2.shape[1], a2.shape[1],
                                         N.ptr(y), y.shape[1], policy, N.PREC_FP32, N.stream_ptr()))
        return y

    @staticmethod
    def backward(ctx, dy):
        q, a2 = ctx.saved_tensors
        op = ctx.op
        dy = dy.to(op.dtype).contiguous()
        dq = torch.zeros_like(q) if ctx.needs_input_grad[0] else None
        de = torch.zeros_like(op.embeddings) if ctx.needs_input_grad[1] else None
        da = torch.empty_like(a2) if ctx.needs_input_grad[2] else None
        N.check(N.lib.laker_cross_matvec_backward(op.handle, N.ptr(q), q.shape[0], N.ptr(a2), a2.shape[1],
                                                  N.ptr(dy), N.ptr(dq), N.ptr(de), N.ptr(da), ctx.policy,
                                                  N.stream_ptr()))
        return dq, de, da, None, None


def cross_kernel_matvec(op: AttentionKernelOperator, query_embeddings: torch.Tensor, alpha: torch.Tensor,
                        clamp_always: bool = False, embeddings: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Differentiable ``exp(min(s Eq E^T, c)) alpha`` through the fused kernel."""
    one_d = alpha.dim() == 1
    a2 = alpha.unsqueeze(1) if one_d else alpha
    emb = embeddings if embeddings is not None else op.embeddings
    policy = N.CLAMP_ALWAYS if clamp_always else N.CLAMP_AUTO
    y = _CrossFn.apply(query_embeddings.to(compute_device()), emb, a2.to(compute_device()), op, policy)
    return y.squeeze(1) if one_d else y


class LAKERCore:
    def __init__(self, embedding_dim: int = 10, lambda_reg: float = 1e-2, gamma: float = 1e-1,
                 num_probes: Optional[int] = None, epsilon: float = 1e-8, base_rho: float = 0.05,
                 cccp_max_iter: int = 200, cccp_tol: float = 1e-6, pcg_tol: float = 1e-6,
                 pcg_max_iter: int = 1000, chunk_size: Optional[int] = None,
                 embedding_module: Optional[nn.Module] = None, kernel_approx: Optional[str] = None,
                 num_landmarks: Optional[int] = None, num_features: Optional[int] = None,
                 k_neighbors: Optional[int] = None, grid_size: Optional[int] = None,
                 distributed: bool = False, twoscale_alpha: float = 0.5, landmark_method: str = "greedy",
                 landmark_pilot_size: int = 1000, spectral_knots: int = 5,
                 preconditioner_strategy: str = "cccp", embedding_dtype: Optional[torch.dtype] = None,
                 device: Optional[Union[str, torch.device]] = None, dtype: Optional[torch.dtype] = None,
                 verbose: bool = True, precision: str = "fp32", probe_source: str = "reference",
                 probe_precision: str = "fp32", inner_tol: float = 5e-2):
        self.embedding_dim = embedding_dim
        self.lambda_reg = lambda_reg
        self.gamma = gamma
        self.num_probes = num_probes
        self.epsilon = epsilon
        self.base_rho = base_rho
        self.cccp_max_iter = cccp_max_iter
        self.cccp_tol = cccp_tol
        self.pcg_tol = pcg_tol
        self.pcg_max_iter = pcg_max_iter
        self.chunk_size = chunk_size
        self.embedding_module = embedding_module
        self.kernel_approx = kernel_approx
        self.num_landmarks = num_landmarks
        self.num_features = num_features
        self.k_neighbors = k_neighbors
        self.grid_size = grid_size
        self.distributed = distributed
        self.twoscale_alpha = twoscale_alpha
        self.landmark_method = landmark_method
        self.landmark_pilot_size = landmark_pilot_size
        self.spectral_knots = spectral_knots
        self.preconditioner_strategy = preconditioner_strategy
        self.verbose = verbose
        self.precision = precision
        self.probe_source = probe_source
        self.probe_precision = probe_precision
        self.inner_tol = inner_tol
        if device is None:
            device = get_default_device()
        elif isinstance(device, str):
            device = torch.device(device)
        if dtype is None:
            dtype = get_default_dtype()
        self.device = device
        self.dtype = dtype
        self.embedding_dtype = embedding_dtype if embedding_dtype is not None else dtype

    # ------------------------------------------------------------------
    def compute_embeddings(self, x: torch.Tensor, embedding_model: Optional[nn.Module] = None):
        """Embeddings of the training locations (laker/core.py:247-302), on the GPU."""
        dev = compute_device()
        n, input_dim = x.shape
        if self.verbose:
            logger.info("Fitting LAKER on n=%d, dx=%d", n, input_dim)
        inp = x.to(device=dev, dtype=self.embedding_dtype)
        if self.embedding_module is not None:
            model = self.embedding_module.to(dev)
        elif embedding_model is None:
            model = PositionEmbedding(input_dim=input_dim, embedding_dim=self.embedding_dim, device=dev,
                                      dtype=self.embedding_dtype)
        else:
            model = embedding_model
        with torch.no_grad():
            emb = model(inp)
        if emb.dtype != self.dtype:
            emb = emb.to(dtype=self.dtype)
        return emb, model

    def build_kernel_operator(self, embeddings: torch.Tensor, lambda_reg: Optional[float] = None,
                              chunk_size: Optional[int] = None) -> KernelOperator:
        """Exact operator (laker/core.py:307-456); approximations are not provided here."""
        if self.kernel_approx is not None:
            raise ValueError(
                f"kernel_approx={self.kernel_approx!r} is not provided by the HIP solver core; "
                "pass kernel_approx=None (exact kernel)")
        lam = float(lambda_reg) if lambda_reg is not None else self.lambda_reg
        chunk = auto_chunk(embeddings.shape[0], chunk_size)
        if self.verbose and chunk_size is None and chunk is not None:
            logger.info("Auto-selected chunk_size=%d for n=%d", chunk, embeddings.shape[0])
        inner = "mixed" if self.precision == "mixed" else "fp32"
        if self.distributed:
            # one GPU per process: the wrapper takes the reference's single-device
            # form (laker/distributed_kernels.py:116-134, laker/core.py:348-360)
            op = DistributedAttentionKernelOperator(embeddings, lambda_reg=lam, master_device=self.device,
                                                    dtype=self.dtype, precision=inner)
            if self.verbose:
                logger.info("Using distributed kernel on %d device(s)", len(op.devices))
            return op
        return AttentionKernelOperator(embeddings, lambda_reg=lam, chunk_size=chunk, device=self.device,
                                       dtype=self.dtype, precision=inner)

    def build_preconditioner(self, matvec: Callable[[torch.Tensor], torch.Tensor], n: int,
                             gamma: Optional[float] = None, num_probes: Optional[int] = None,
                             seed: Optional[int] = None, diagonal: Optional[torch.Tensor] = None):
        """CCCP or adaptive preconditioner (laker/core.py:461-521)."""
        if self.preconditioner_strategy == "adaptive":
            pre = AdaptivePreconditioner(
                gamma=gamma if gamma is not None else self.gamma,
                num_probes=num_probes if num_probes is not None else self.num_probes, epsilon=self.epsilon,
                base_rho=self.base_rho, max_iter=self.cccp_max_iter, tol=self.cccp_tol, verbose=self.verbose,
                device=self.device, dtype=self.dtype, probe_source=self.probe_source,
                probe_precision=self.probe_precision)
            return pre.build(matvec, n, diagonal=diagonal, seed=seed)
        if self.preconditioner_strategy != "cccp":
            raise ValueError(f"preconditioner_strategy must be 'cccp' or 'adaptive', got "
                             f"{self.preconditioner_strategy!r}")
        pre = CCCPPreconditioner(num_probes=num_probes if num_probes is not None else self.num_probes,
                                 gamma=gamma if gamma is not None else self.gamma, epsilon=self.epsilon,
                                 base_rho=self.base_rho, max_iter=self.cccp_max_iter, tol=self.cccp_tol,
                                 verbose=self.verbose, device=self.device, dtype=self.dtype,
                                 probe_source=self.probe_source, probe_precision=self.probe_precision)
        pre.build(matvec, n, seed=seed)
        return pre

    def solve_pcg(self, kernel_operator: KernelOperator, preconditioner, rhs: torch.Tensor,
                  x0: Optional[torch.Tensor] = None):
        """``(alpha, iterations)`` (laker/core.py:526-575)."""
        pcg = PreconditionedConjugateGradient(tol=self.pcg_tol, max_iter=self.pcg_max_iter, verbose=self.verbose,
                                              precision=self.precision, inner_tol=self.inner_tol)
        alpha = pcg.solve(operator=kernel_operator.matvec, preconditioner=preconditioner.apply, rhs=rhs, x0=x0)
        self.last_solver = pcg
        native = native_operator(kernel_operator)
        if self.verbose and isinstance(native, AttentionKernelOperator):
            r = native.residual_fp64(rhs, pcg.solution_fp64 if pcg.solution_fp64 is not None else alpha)
            rel = float(torch.linalg.norm(r) / torch.linalg.norm(rhs.to(r.device, torch.float64)))
            logger.info("LAKER fit complete: PCG iters=%d, final rel_res=%.3e", pcg.iterations, rel)
        return alpha, pcg.iterations

    # ------------------------------------------------------------------
    def _query_embeddings(self, x: torch.Tensor, embedding_model: nn.Module) -> torch.Tensor:
        q = embedding_model(x.to(device=compute_device(), dtype=self.embedding_dtype))
        return q.to(dtype=self.dtype)

    def _predict_clamp_always(self, m: int, n: int) -> bool:
        """True where the reference takes its tiled predict branch, which always clamps."""
        chunk = self.chunk_size
        if chunk is None and max(m, n) > 5000:
            chunk = max(1024, min(max(m, n) // 10, 8192))
        if chunk is None:
            return False
        elt = 4 if self.dtype == torch.float32 else 8
        return chunk * n * elt > 64 * 1024 * 1024

    def predict(self, x: torch.Tensor, embedding_model: nn.Module, embeddings: torch.Tensor,
                kernel_operator: KernelOperator, alpha: torch.Tensor,
                residual_corrector: Optional[nn.Module] = None) -> torch.Tensor:
        """``exp(Eq E^T) alpha`` by the fused cross kernel (laker/core.py:580-661)."""
        with torch.no_grad():
            q = self._query_embeddings(x, embedding_model)
            clamp_always = self._predict_clamp_always(q.shape[0], embeddings.shape[0])
            native = native_operator(kernel_operator)
            if isinstance(native, AttentionKernelOperator):
                # an FP64 solution (the mixed solve's refined alpha) is applied in
                # FP64 on the operator's FP64 copy of E, then returned in dtype
                fp64 = alpha.dtype == torch.float64 and native.dtype == torch.float32
                out = native.cross_matvec(q, alpha, clamp_always=clamp_always,
                                          precision="fp64" if fp64 else None).to(self.dtype)
            else:
                out = kernel_operator.kernel_eval(q, embeddings) @ alpha
            if residual_corrector is not None:
                residual_corrector.eval()
                out = out + residual_corrector(x.to(out.device)).squeeze()
            return out.to(self.device)

    def predict_train(self, x: torch.Tensor, embedding_model: nn.Module, embeddings: torch.Tensor,
                      kernel_operator: KernelOperator, alpha: torch.Tensor,
                      residual_corrector: Optional[nn.Module] = None) -> torch.Tensor:
        """Differentiable predict (laker/core.py:782-849) with the hand-written cross backward."""
        q = self._query_embeddings(x, embedding_model)
        clamp_always = self._predict_clamp_always(q.shape[0], embeddings.shape[0])
        out = cross_kernel_matvec(native_operator(kernel_operator), q, alpha, clamp_always=clamp_always,
                                  embeddings=embeddings if embeddings.requires_grad else None)
        if residual_corrector is not None:
            residual_corrector.eval()
            out = out + residual_corrector(x.to(out.device)).squeeze()
        return out

    def predict_variance(self, x: torch.Tensor, embedding_model: nn.Module, embeddings: torch.Tensor,
                         kernel_operator: KernelOperator, preconditioner, alpha: torch.Tensor,
                         lambda_reg: float) -> torch.Tensor:
        """``k(q,q) - k_q^T A^{-1} k_q`` by batched PCG (laker/core.py:666-777)."""
        with torch.no_grad():
            q = self._query_embeddings(x, embedding_model)
            m, n = q.shape[0], embeddings.shape[0]
            elt = 4 if self.dtype == torch.float32 else 8
            chunk = self.chunk_size
            if chunk is None and m * n * elt > 64 * 1024 * 1024:
                chunk = max(1024, min(n // 10, 8192))
            pcg = PreconditionedConjugateGradient(tol=self.pcg_tol, max_iter=self.pcg_max_iter, verbose=False,
                                                  precision=self.precision, inner_tol=self.inner_tol)
            var = torch.empty(m, dtype=self.dtype, device=compute_device())
            step = m if (chunk is None or m <= chunk) else chunk
            for s in range(0, m, step):
                e = min(s + step, m)
                qc = q[s:e]
                k_xq = kernel_operator.kernel_eval(embeddings, qc).to(compute_device())
                v = pcg.solve(kernel_operator.matvec, preconditioner.apply, k_xq)
                k_qq = kernel_operator.kernel_eval(qc, qc).to(compute_device()).diagonal()
                if getattr(pcg, "solution_fp64", None) is not None and self.precision == "mixed":
                    # the refined solution stays FP64: var is a small difference of large terms
                    var[s:e] = (k_qq.double() - _coldot(k_xq.double(), pcg.solution_fp64)).to(self.dtype)
                else:
                    var[s:e] = (k_qq.double() - _coldot(k_xq, v.to(k_xq.dtype))).to(self.dtype)
            self.last_solver = pcg
            return var.clamp(min=0.0).to(self.device)
