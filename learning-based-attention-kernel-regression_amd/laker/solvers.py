"""Preconditioned conjugate gradient (reference laker/solvers.py:48-394).

``PreconditionedConjugateGradient.solve(operator, preconditioner, rhs, x0)``
keeps the reference signature, attributes (``iterations``,
``residual_norm``) and error behaviour (``ValueError`` for bad dimensions,
``RuntimeError`` on breakdown, a logged warning on non-convergence).

When ``operator`` is the bound ``matvec`` of an :class:`AttentionKernelOperator`
and ``preconditioner`` is the bound ``apply`` of a :class:`CCCPPreconditioner`
(or ``None``), the whole loop runs on device in the HIP library
(``laker_pcg_solve``): vectors resident, FP64 scalar recurrences, no host sync
per iteration.  ``precision="mixed"`` selects the split-Gram inner PCG with
FP64 iterative refinement (``laker_solve_mixed``); its stopping test is the
FP64 true residual.  Any other callables take the generic loop, which runs on
the GPU with the library's FP64 column dot products.
"""

from __future__ import annotations

import ctypes
import logging
from typing import Callable, Optional

import torch

from . import _native as N
from .backend import compute_device

logger = logging.getLogger(__name__)


def _native_operator(operator):
    from .kernels import AttentionKernelOperator

    obj = getattr(operator, "__self__", None)
    fn = getattr(operator, "__func__", None)
    if isinstance(obj, AttentionKernelOperator) and fn in (
        AttentionKernelOperator.matvec, AttentionKernelOperator.matvec_impl,
    ):
        return obj
    if isinstance(operator, AttentionKernelOperator):
        return operator
    return None


def _native_preconditioner(preconditioner):
    """``(handle owner, True)`` when ``preconditioner`` is a native CCCP/Jacobi
    apply (or an adaptive wrapper around one), ``(None, True)`` for None and
    ``(None, False)`` for any other callable."""
    from .preconditioner import AdaptivePreconditioner, CCCPPreconditioner

    if preconditioner is None:
        return None, True
    obj = getattr(preconditioner, "__self__", None)
    fn = getattr(preconditioner, "__func__", None)
    if isinstance(obj, CCCPPreconditioner) and fn in (
        CCCPPreconditioner.apply, CCCPPreconditioner.apply_1d, CCCPPreconditioner.apply_2d,
        CCCPPreconditioner.apply_impl,
    ):
        return obj, True
    if isinstance(obj, JacobiPreconditioner) and fn is JacobiPreconditioner.apply:
        return obj, True
    if isinstance(obj, AdaptivePreconditioner) and fn is AdaptivePreconditioner.apply and obj.inner is not None:
        return _native_preconditioner(obj.inner.apply)
    if isinstance(preconditioner, (CCCPPreconditioner, JacobiPreconditioner)):
        return preconditioner, True
    return None, False


class JacobiPreconditioner:
    """Diagonal preconditioner ``P = diag(lambda I + G)^{-1}`` (reference
    laker/solvers.py:505-546), held and applied by the HIP library.

    ``inv_diag = 1 / diagonal.clamp(min=eps)`` is computed on the GPU in the
    diagonal's dtype (float32 or float64; other dtypes are taken as float32);
    ``apply`` scales every row of ``(n,)`` or ``(n, k)`` input by it.  For a
    2-D input the reference's ``inv_diag * x`` broadcasts over the last axis
    (it raises unless k == n); this build applies the documented row scaling.
    Bound ``apply`` methods drive the device PCG loop like the CCCP apply.
    """

    def __init__(self, diagonal: torch.Tensor) -> None:
        from .preconditioner import _PrecondHandle

        if diagonal.dim() != 1:
            raise ValueError(f"diagonal must be 1-D, got shape {diagonal.shape}")
        dtype = diagonal.dtype if diagonal.dtype in (torch.float32, torch.float64) else torch.float32
        d = diagonal.detach().to(device=compute_device(), dtype=dtype).contiguous()
        self.n = d.shape[0]
        self.dtype = dtype
        self._h = _PrecondHandle()
        N.check(N.lib.laker_precond_create_jacobi(N.ptr(d), self.n, N.dtype_code(dtype), N.stream_ptr(),
                                                  ctypes.byref(self._h.h)))
        self._inv = None

    @property
    def handle(self):
        return self._h.h

    @property
    def inv_diag(self) -> torch.Tensor:
        """``1 / max(diag, eps)``, shape ``(n,)``."""
        if self._inv is None:
            t = torch.empty(self.n, dtype=self.dtype, device=compute_device())
            N.check(N.lib.laker_precond_copy_out(self._h.h, 4, N.ptr(t), N.stream_ptr()))
            self._inv = t
        return self._inv

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        """``P x`` for ``(n,)`` or ``(n, k)`` input."""
        if x.dim() not in (1, 2) or x.shape[0] != self.n:
            raise ValueError(f"x must be ({self.n},) or ({self.n}, k), got shape {tuple(x.shape)}")
        out_dev = x.device
        xd = x.to(device=compute_device(), dtype=self.dtype).contiguous()
        x2 = xd.unsqueeze(1) if xd.dim() == 1 else xd
        z = torch.empty_like(x2)
        N.check(N.lib.laker_precond_apply(self._h.h, N.ptr(x2), x2.shape[1], x2.shape[1], N.ptr(z), z.shape[1],
                                          N.stream_ptr()))
        z = z.squeeze(1) if x.dim() == 1 else z
        return z.to(out_dev)

    __call__ = apply

    def close(self) -> None:
        self._h.close()


def _coldot(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """FP64 column dot products of two (n, k) GPU tensors (library kernel)."""
    a2 = a if a.dim() == 2 else a.unsqueeze(1)
    b2 = b if b.dim() == 2 else b.unsqueeze(1)
    a2 = a2.contiguous()
    b2 = b2.to(a2.dtype).contiguous()
    out = torch.empty(a2.shape[1], dtype=torch.float64, device=a2.device)
    N.check(N.lib.laker_coldot(N.ptr(a2), N.ptr(b2), a2.shape[0], a2.shape[1], a2.shape[1],
                               N.dtype_code(a2.dtype), N.ptr(out), N.stream_ptr()))
    return out


class PreconditionedConjugateGradient:
    """PCG solver (reference laker/solvers.py:48-394).

    Extra arguments (not in the reference): ``precision`` (``"fp32"`` or
    ``"mixed"``), ``check_every`` (device status polling period) and, for the
    mixed path, ``inner_tol`` / ``max_outer``.
    """

    def __init__(self, tol: float = 1e-10, max_iter: Optional[int] = None, verbose: bool = True,
                 restart_freq: Optional[int] = None, breakdown_eps: Optional[float] = None,
                 autograd_safe: bool = False, precision: str = "fp32", check_every: int = 8,
                 inner_tol: float = 5e-2, max_outer: int = 20):
        if precision not in ("fp32", "mixed"):
            raise ValueError("precision must be 'fp32' or 'mixed'")
        self.tol = float(tol)
        self.max_iter = max_iter
        self.verbose = verbose
        self.restart_freq = restart_freq
        self.breakdown_eps = breakdown_eps
        self.autograd_safe = autograd_safe
        self.precision = precision
        self.check_every = int(check_every)
        self.inner_tol = float(inner_tol)
        self.max_outer = int(max_outer)
        self.iterations: int = 0
        self.residual_norm: float = float("inf")
        self.outer_iterations: int = 0
        self.converged: bool = False
        self.solution_fp64: Optional[torch.Tensor] = None

    # ------------------------------------------------------------------
    def solve(self, operator: Callable[[torch.Tensor], torch.Tensor],
              preconditioner: Optional[Callable[[torch.Tensor], torch.Tensor]], rhs: torch.Tensor,
              x0: Optional[torch.Tensor] = None) -> torch.Tensor:
        if rhs.dim() not in (1, 2):
            raise ValueError(f"rhs must be 1-D or 2-D, got shape {rhs.shape}")
        op = _native_operator(operator)
        pc, pc_ok = _native_preconditioner(preconditioner)
        if op is not None and pc is not None and (pc.dtype != op.dtype or pc.n != op.n):
            pc_ok = False  # the device loop needs a preconditioner of the operator's size and dtype
        if op is not None and pc_ok and not self.autograd_safe and not (
            torch.is_grad_enabled() and (rhs.requires_grad or (x0 is not None and x0.requires_grad))
        ):
            if self.precision == "mixed" and op.dtype == torch.float32:
                return self._solve_mixed(op, pc, rhs, x0)
            return self._solve_native(op, pc, rhs, x0)
        return self._solve_generic(operator, preconditioner, rhs, x0)

    # ------------------------------------------------------------------
    def _solve_native(self, op, pc, rhs, x0):
        out_dev = rhs.device
        n = rhs.shape[0]
        b = rhs.to(device=compute_device(), dtype=op.dtype).contiguous()
        b2 = b.unsqueeze(1) if b.dim() == 1 else b
        k = b2.shape[1]
        x = torch.empty_like(b2)
        x0c = None
        if x0 is not None:
            x0c = x0.to(device=compute_device(), dtype=op.dtype).reshape(b2.shape).contiguous()
        max_iter = self.max_iter if self.max_iter is not None else n
        res = N.PcgResult()
        eps = -1.0 if self.breakdown_eps is None else float(self.breakdown_eps)
        prec = N.PREC_MIXED if self.precision == "mixed" else N.PREC_FP32
        rc = N.lib.laker_pcg_solve(op.handle, pc.handle if pc is not None else None, N.ptr(b2), k,
                                   N.ptr(x), N.ptr(x0c), self.tol, int(max_iter), eps,
                                   int(self.restart_freq or 0), prec, self.check_every, N.stream_ptr(),
                                   ctypes.byref(res))
        self.iterations = int(res.iterations)
        self.residual_norm = float(res.residual_norm)
        self.converged = bool(res.converged)
        N.check(rc)
        if not res.zero_rhs:
            rel = self.residual_norm / res.rhs_norm if res.rhs_norm else float("inf")
            if self.converged:
                if self.verbose:
                    logger.info("PCG converged in %d iterations, rel_res=%.3e", self.iterations, rel)
            elif self.verbose:
                logger.warning("PCG did not converge in %d iterations, rel_res=%.3e", max_iter, rel)
        x = x.squeeze(1) if rhs.dim() == 1 else x
        return x.to(out_dev)

    def _solve_mixed(self, op, pc, rhs, x0):
        out_dev = rhs.device
        n = rhs.shape[0]
        b = rhs.to(device=compute_device(), dtype=torch.float64).contiguous()
        b2 = b.unsqueeze(1) if b.dim() == 1 else b
        tol = self.tol
        # ||b|| on the host only where it is used (warm start, log line): it
        # costs a device synchronisation before the solve starts
        bnorm = float(torch.linalg.norm(b2)) if (x0 is not None or self.verbose) else None
        if x0 is not None:
            # warm start: solve for the correction A d = b - A x0; the stop stays
            # relative to ||b|| (reference laker/solvers.py:148-166)
            x0d = x0.to(device=compute_device(), dtype=torch.float64).reshape(b2.shape)
            b2 = op.residual_fp64(b2, x0d)
            r0 = float(torch.linalg.norm(b2))
            if r0 > 0:
                tol = min(0.5, self.tol * bnorm / r0)
        k = b2.shape[1]
        x = torch.empty_like(b2)
        res = N.MixedResult()
        max_iter = self.max_iter if self.max_iter is not None else n
        N.check(N.lib.laker_solve_mixed(op.handle, pc.handle if pc is not None else None, N.ptr(b2), k,
                                        N.ptr(x), tol, self.max_outer, self.inner_tol, int(max_iter),
                                        N.PREC_MIXED, N.stream_ptr(), ctypes.byref(res)))
        if x0 is not None:
            x = x + x0d
        self.iterations = int(res.inner_iterations)
        self.outer_iterations = int(res.outer_iterations)
        self.residual_norm = float(res.residual_norm)
        self.converged = bool(res.converged)
        if self.verbose:
            rel = self.residual_norm / bnorm if bnorm else 0.0
            if self.converged:
                logger.info("mixed PCG: %d outer / %d inner iterations, fp64 rel_res=%.3e",
                            self.outer_iterations, self.iterations, rel)
            else:
                # as the reference's non-convergence warning (laker/solvers.py:288-294)
                logger.warning("PCG did not converge in %d iterations, rel_res=%.3e", self.iterations, rel)
        x = x.squeeze(1) if rhs.dim() == 1 else x
        self.solution_fp64 = x
        return x.to(device=out_dev, dtype=rhs.dtype if rhs.dtype.is_floating_point else torch.float32)

    # ------------------------------------------------------------------
    def _solve_generic(self, operator, preconditioner, rhs, x0):
        """The reference loop for arbitrary callables, on the GPU."""
        out_dev = rhs.device
        dev = compute_device()
        b = rhs.to(dev)
        pre = preconditioner if preconditioner is not None else (lambda v: v)
        n = b.shape[0]
        max_iter = self.max_iter if self.max_iter is not None else n
        if x0 is None:
            x = torch.zeros_like(b)
            r = b.clone()
        else:
            x = x0.to(dev, b.dtype).clone()
            r = b - operator(x)
        z = pre(r)
        p = z.clone()
        rhs_norm = float(_coldot(b, b).sum().sqrt())
        if rhs_norm == 0:
            return x.to(out_dev)
        one_d = b.dim() == 1
        eps = self.breakdown_eps if self.breakdown_eps is not None else torch.finfo(p.dtype).eps ** 0.5
        rz_old = _coldot(r, z)
        rel = float("inf")
        for it in range(max_iter):
            ap = operator(p)
            pap = _coldot(p, ap)
            pn = _coldot(p, p).sqrt()
            an = _coldot(ap, ap).sqrt()
            if bool(torch.any(pap <= -eps * pn * an)):
                raise RuntimeError(
                    "PCG breakdown: non-positive curvature detected (p^T A p <= 0). "
                    "The operator may be indefinite or the preconditioner may be unsuitable.")
            alpha = (rz_old / pap).to(b.dtype)
            alpha_b = alpha[0] if one_d else alpha.unsqueeze(0)
            if self.autograd_safe:
                x = x + alpha_b * p
                r = r - alpha_b * ap
            else:
                x.add_(alpha_b * p)
                r.sub_(alpha_b * ap)
            if self.restart_freq is not None and (it + 1) % self.restart_freq == 0:
                r = b - operator(x)
            self.residual_norm = float(_coldot(r, r).sum().sqrt())
            rel = self.residual_norm / rhs_norm
            if rel <= self.tol:
                self.iterations = it + 1
                self.converged = True
                if self.verbose:
                    logger.info("PCG converged in %d iterations, rel_res=%.3e", self.iterations, rel)
                return x.to(out_dev)
            z = pre(r)
            rz_new = _coldot(r, z)
            beta = (rz_new / rz_old).to(b.dtype)
            beta_b = beta[0] if one_d else beta.unsqueeze(0)
            if self.autograd_safe:
                p = z + beta_b * p
            else:
                p.mul_(beta_b).add_(z)
            rz_old = rz_new
        self.iterations = max_iter
        self.converged = False
        if self.verbose:
            logger.warning("PCG did not converge in %d iterations, rel_res=%.3e", max_iter, rel)
        return x.to(out_dev)
