"""Exact attention-kernel operator backed by the fused HIP matvec.

``AttentionKernelOperator`` keeps the reference class's attributes and methods
(laker/kernels.py:43-72 protocol, :133-354 class) and runs every product in the
HIP library: the n x n kernel is never materialised except by ``to_dense`` and
``kernel_eval``, which materialise by definition.  ``chunk_size`` is accepted
for signature compatibility; the fused kernel tiles internally.

The approximate operators of the reference (Nystrom, RFF, k-NN, SKI,
two-scale, spectral) are not part of this package; any object with the
``KernelOperator`` protocol still plugs into the solver and preconditioner.
"""

from __future__ import annotations

import ctypes
import logging
from typing import Optional, Protocol, Tuple

import torch

from . import _native as N
from .backend import compute_device

logger = logging.getLogger(__name__)

_PRECISIONS = {"fp32": N.PREC_FP32, "mixed": N.PREC_MIXED, "bf16": N.PREC_BF16, "hilo": N.PREC_HILO,
               "fp64": N.PREC_FP64}


class KernelOperator(Protocol):
    """Protocol for matrix-free kernel operators (laker/kernels.py:43-72)."""

    n: int
    embedding_dim: int
    lambda_reg: float
    dtype: torch.dtype
    device: torch.device
    shape: Tuple[int, int]

    def matvec(self, x: torch.Tensor) -> torch.Tensor: ...

    def diagonal(self) -> torch.Tensor: ...

    def to_dense(self) -> torch.Tensor: ...

    def kernel_eval(self, x: torch.Tensor, y: Optional[torch.Tensor] = None,
                    chunk_size: Optional[int] = None) -> torch.Tensor: ...


class _OperatorHandle:
    """Owns one ``laker_operator_t``."""

    def __init__(self, embeddings: torch.Tensor, lambda_reg: float, scale: float, clamp_policy: int):
        self.h = ctypes.c_void_p()
        N.check(N.lib.laker_operator_create(
            N.ptr(embeddings), embeddings.shape[0], embeddings.shape[1], N.dtype_code(embeddings.dtype),
            float(lambda_reg), float(scale), clamp_policy, N.stream_ptr(), ctypes.byref(self.h)))
        skip = ctypes.c_int32()
        mx = ctypes.c_double()
        N.check(N.lib.laker_operator_info(self.h, ctypes.byref(skip), ctypes.byref(mx)))
        self.skip_clamp = bool(skip.value)
        self.max_sq_norm = mx.value

    def close(self):
        if self.h:
            N.lib.laker_operator_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _as_2d(x: torch.Tensor) -> Tuple[torch.Tensor, bool]:
    if x.dim() == 1:
        return x.unsqueeze(1), True
    return x, False


class _MatvecFn(torch.autograd.Function):
    """``y = lambda x + exp(min(s E E^T, c)) x`` with the hand-written backward."""

    @staticmethod
    def forward(ctx, embeddings, x2, op):
        ctx.op = op
        ctx.save_for_backward(x2)
        return op._matvec_raw(x2)

    @staticmethod
    def backward(ctx, dy):
        (x2,) = ctx.saved_tensors
        op = ctx.op
        dy = dy.contiguous()
        need_e, need_x = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        dx = torch.empty_like(dy) if need_x else None
        de = torch.empty_like(op.embeddings) if need_e else None
        N.check(N.lib.laker_matvec_backward(op._h.h, N.ptr(x2), N.ptr(dy), dy.shape[1], N.ptr(dx),
                                            N.ptr(de), N.stream_ptr()))
        return de, dx, None


class AttentionKernelOperator:
    """Matrix-free ``lambda I + exp(scale * E E^T)`` (reference laker/kernels.py:133-354).

    Args:
        embeddings: ``(n, embedding_dim)`` tensor (any device; copied to the GPU).
        lambda_reg: Tikhonov regulariser added to the diagonal.
        chunk_size: accepted for compatibility (the fused kernel tiles itself).
        device: device on which results are returned (default: the input's).
        dtype: ``torch.float32`` (default for fp32 inputs) or ``torch.float64``.
        precision: matvec arithmetic for float32 operators: ``"fp32"`` (FP32
            MFMA Gram) or ``"mixed"`` (split MFMA Gram: 3-way BF16, or 2-way
            FP16 for the square operator's k=1 product when E fits FP16).
        scale: optional factor s in ``exp(s E E^T)`` (reference: 1.0).
    """

    def __init__(self, embeddings: torch.Tensor, lambda_reg: float = 1e-2, chunk_size: Optional[int] = None,
                 device: Optional[torch.device] = None, dtype: Optional[torch.dtype] = None,
                 precision: str = "fp32", scale: float = 1.0):
        if embeddings.dim() != 2:
            raise ValueError(f"embeddings must be 2-D, got shape {embeddings.shape}")
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {precision}")
        self.n = embeddings.shape[0]
        self.embedding_dim = embeddings.shape[1]
        self._lambda = float(lambda_reg)
        self.chunk_size = chunk_size
        self.device = torch.device(device) if device is not None else embeddings.device
        if dtype is None:
            dtype = embeddings.dtype if embeddings.dtype in (torch.float32, torch.float64) else torch.float32
        self.dtype = dtype
        self.precision = precision
        self.scale = float(scale)
        dev = compute_device()
        emb = embeddings.to(device=dev, dtype=dtype)
        self._grad_embeddings = emb if emb.requires_grad else None
        self.embeddings = emb.detach().contiguous()
        self.shape = (self.n, self.n)
        self._h = _OperatorHandle(self.embeddings, self._lambda, self.scale, N.CLAMP_AUTO)
        self.skip_clamp = self._h.skip_clamp

    # lambda_reg is mutable in the reference (a plain attribute)
    @property
    def lambda_reg(self) -> float:
        return self._lambda

    @lambda_reg.setter
    def lambda_reg(self, value: float) -> None:
        self._lambda = float(value)
        N.check(N.lib.laker_operator_set_lambda(self._h.h, self._lambda))

    @property
    def handle(self):
        return self._h.h

    def close(self) -> None:
        self._h.close()

    # ------------------------------------------------------------------
    def _matvec_raw(self, x2: torch.Tensor, precision: Optional[str] = None) -> torch.Tensor:
        prec = _PRECISIONS[precision or self.precision]
        if self.dtype == torch.float64:
            prec = N.PREC_FP64
        x2 = x2.to(dtype=self.dtype).contiguous()
        y = torch.empty_like(x2)
        N.check(N.lib.laker_matvec(self._h.h, N.ptr(x2), x2.shape[1], x2.shape[1], N.ptr(y), y.shape[1],
                                   prec, N.stream_ptr()))
        return y

    def matvec(self, x: torch.Tensor, precision: Optional[str] = None) -> torch.Tensor:
        """``(lambda I + G) x`` for ``x`` of shape ``(n,)`` or ``(n, k)`` (laker/kernels.py:187-208)."""
        if x.dim() not in (1, 2):
            raise ValueError(f"x must be 1-D or 2-D, got shape {x.shape}")
        if x.shape[0] != self.n:
            raise ValueError(f"x must have {self.n} rows, got shape {tuple(x.shape)}")
        out_dev = x.device
        x2, one_d = _as_2d(x.to(device=self.embeddings.device))
        grad = torch.is_grad_enabled() and (x2.requires_grad or self._grad_embeddings is not None)
        if grad and self.dtype == torch.float32:
            emb = self._grad_embeddings if self._grad_embeddings is not None else self.embeddings
            y = _MatvecFn.apply(emb, x2.to(self.dtype).contiguous(), self)
        else:
            y = self._matvec_raw(x2, precision)
        y = y.squeeze(1) if one_d else y
        return y.to(out_dev)

    __call__ = matvec

    def matvec_impl(self, x: torch.Tensor) -> torch.Tensor:
        return self.matvec(x)

    def residual_fp64(self, b: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        """``b - (lambda I + G) x`` with FP64 Gram, exp and accumulation."""
        b2, one_d = _as_2d(b.to(device=self.embeddings.device, dtype=torch.float64).contiguous())
        x2, _ = _as_2d(x.to(device=self.embeddings.device, dtype=torch.float64).contiguous())
        r = torch.empty_like(b2)
        N.check(N.lib.laker_residual_fp64(self._h.h, N.ptr(b2), N.ptr(x2), b2.shape[1], N.ptr(r),
                                          N.stream_ptr()))
        return r.squeeze(1) if one_d else r

    def diagonal(self) -> torch.Tensor:
        """``lambda + exp(s ||e_i||^2)`` (laker/kernels.py:272-283)."""
        out = torch.empty(self.n, dtype=self.dtype, device=self.embeddings.device)
        N.check(N.lib.laker_diagonal(self._h.h, N.ptr(out), N.stream_ptr()))
        return out.to(self.device)

    def to_dense(self) -> torch.Tensor:
        """Dense ``lambda I + exp(s E E^T)``, no clamp (laker/kernels.py:285-299)."""
        out = torch.empty(self.n, self.n, dtype=self.dtype, device=self.embeddings.device)
        N.check(N.lib.laker_to_dense(self._h.h, N.ptr(out), N.stream_ptr()))
        return out.to(self.device)

    def kernel_eval(self, x: torch.Tensor, y: Optional[torch.Tensor] = None,
                    chunk_size: Optional[int] = None) -> torch.Tensor:
        """``exp(min(s x y^T, c))`` materialised (laker/kernels.py:301-354)."""
        out_dev = x.device
        dev = self.embeddings.device
        xe = x.to(device=dev, dtype=self.dtype).contiguous()
        ye = self.embeddings if y is None else y.to(device=dev, dtype=self.dtype).contiguous()
        if xe.dim() != 2 or ye.dim() != 2 or xe.shape[1] != ye.shape[1]:
            raise ValueError("kernel_eval expects 2-D inputs with matching embedding_dim")
        m, p = xe.shape[0], ye.shape[0]
        out = torch.empty(m, p, dtype=self.dtype, device=dev)
        N.check(N.lib.laker_kernel_eval(N.ptr(xe), m, N.ptr(ye), p, xe.shape[1], N.dtype_code(self.dtype),
                                        self.scale, 0 if self.skip_clamp else 1, N.ptr(out), p,
                                        N.stream_ptr()))
        return out.to(out_dev)

    def cross_matvec(self, query_embeddings: torch.Tensor, x: torch.Tensor, clamp_always: bool = False,
                     precision: Optional[str] = None) -> torch.Tensor:
        """``exp(min(s Eq E^T, c)) x`` without materialising (predict, laker/core.py:580-661)."""
        dev = self.embeddings.device
        q = query_embeddings.to(device=dev, dtype=self.dtype).contiguous()
        prec = _PRECISIONS[precision or "fp32"]
        if self.dtype == torch.float64:
            prec = N.PREC_FP64
        # FP64 products take FP64 vectors (also on an FP32 operator: its FP64 copy of E)
        vdt = torch.float64 if prec == N.PREC_FP64 else self.dtype
        x2, one_d = _as_2d(x.to(device=dev, dtype=vdt).contiguous())
        y = torch.empty(q.shape[0], x2.shape[1], dtype=vdt, device=dev)
        policy = N.CLAMP_ALWAYS if clamp_always else N.CLAMP_AUTO
        N.check(N.lib.laker_cross_matvec(self._h.h, N.ptr(q), q.shape[0], N.ptr(x2), x2.shape[1],
                                         x2.shape[1], N.ptr(y), y.shape[1], policy, prec, N.stream_ptr()))
        return y.squeeze(1) if one_d else y

    def cross_matvec_t(self, query_embeddings: torch.Tensor, xq: torch.Tensor,
                       clamp_always: bool = False) -> torch.Tensor:
        """``exp(min(s E Eq^T, c)) xq`` (n x k), e.g. variance right-hand sides times weights."""
        dev = self.embeddings.device
        q = query_embeddings.to(device=dev, dtype=self.dtype).contiguous()
        x2, one_d = _as_2d(xq.to(device=dev, dtype=self.dtype).contiguous())
        y = torch.empty(self.n, x2.shape[1], dtype=self.dtype, device=dev)
        policy = N.CLAMP_ALWAYS if clamp_always else N.CLAMP_AUTO
        N.check(N.lib.laker_cross_matvec_t(self._h.h, N.ptr(q), q.shape[0], N.ptr(x2), x2.shape[1],
                                           x2.shape[1], N.ptr(y), y.shape[1], policy, N.stream_ptr()))
        return y.squeeze(1) if one_d else y


def dense_attention_matvec(embeddings: torch.Tensor, lambda_reg: float, x: torch.Tensor,
                           skip_clamp: bool = False) -> torch.Tensor:
    """``lambda x + exp(E E^T) x`` (laker/kernels.py:121-130) through the fused kernel."""
    op = AttentionKernelOperator(embeddings, lambda_reg=lambda_reg)
    try:
        return op.matvec(x)
    finally:
        op.close()
