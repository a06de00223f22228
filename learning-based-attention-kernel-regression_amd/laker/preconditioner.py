"""CCCP preconditioner ``P = Sigma^{-1/2}`` built and applied on the GPU.

Keeps the reference class's constructor, ``build(operator, n, seed)``,
``apply(x)`` and attributes (laker/preconditioner.py:71-379).  The build,
QR, FP64 CCCP loop and inverse square root run in the HIP library; the
apply is two HBM-bound passes over ``Q``.

Probes: by default (``probe_source="reference"``) the probe matrix is drawn
exactly as the reference draws it on a CPU device -- ``torch.randn(n, N_r)``
from a fresh CPU ``torch.Generator`` (seeded with ``seed``, or its default
seed 67280421310721) -- and copied to the GPU, so the preconditioner matches
the reference's CPU build.  ``probe_source="device"`` draws N(0,1) probes on
the GPU instead (faster at large n; a different but equally valid stream).

The QR factor here has a positive diagonal (Cholesky-QR), so ``q_basis`` may
differ from the reference's Householder ``Q`` by column signs; ``P`` itself
is invariant to those signs.
"""

from __future__ import annotations

import ctypes
import logging
from typing import Callable, Optional

import torch

from . import _native as N
from .backend import compute_device, get_default_device, get_default_dtype
from .utils import adaptive_shrinkage_rho, default_num_probes, eigh_stable

logger = logging.getLogger(__name__)

# probe products: "mixed" = hi/lo BF16 products of the exp tile (~2^-16, enough
# for the preconditioner), "split3" = the solver's FP32-level BF16 splitting
_PROBE_PREC = {"fp32": N.PREC_FP32, "bf16": N.PREC_BF16, "mixed": N.PREC_HILO, "split3": N.PREC_MIXED}


class _PrecondHandle:
    def __init__(self):
        self.h = ctypes.c_void_p()

    def close(self):
        if self.h:
            N.lib.laker_precond_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def preconditioner_apply_core(x, isotropic_coef, q_basis, q_eigenvalues, q_eigenvectors):
    """``a^{-1/2} x + Q V ((lam^{-1/2} - a^{-1/2}) * V^T Q^T x)`` (laker/preconditioner.py:52-68).

    Functional form for explicitly given factors, evaluated on the GPU.
    """
    inv = isotropic_coef ** (-0.5)
    dev = compute_device()
    xd = x.to(dev)
    q = q_basis.to(dev, xd.dtype)
    v = q_eigenvectors.to(dev, xd.dtype)
    scale = q_eigenvalues.to(dev, xd.dtype).rsqrt() - inv
    c = v.T @ (q.T @ xd)
    c = scale * c if xd.dim() == 1 else scale.unsqueeze(-1) * c
    return (inv * xd + q @ (v @ c)).to(x.device)


class CCCPPreconditioner:
    """Learned preconditioner (reference laker/preconditioner.py:71-379)."""

    def __init__(self, num_probes: Optional[int] = None, gamma: float = 1e-1, epsilon: float = 1e-8,
                 base_rho: float = 0.05, max_iter: int = 200, tol: float = 1e-6, verbose: bool = True,
                 device: Optional[torch.device] = None, dtype: Optional[torch.dtype] = None,
                 probe_strategy: str = "gaussian", power_iter_steps: int = 3,
                 probe_source: str = "reference", probe_precision: str = "fp32"):
        if probe_source not in ("reference", "device"):
            raise ValueError("probe_source must be 'reference' or 'device'")
        self.num_probes = num_probes
        self.gamma = float(gamma)
        self.epsilon = float(epsilon)
        self.base_rho = float(base_rho)
        self.max_iter = int(max_iter)
        self.tol = float(tol)
        self.verbose = verbose
        self.probe_strategy = probe_strategy
        self.power_iter_steps = int(power_iter_steps)
        self.probe_source = probe_source
        self.probe_precision = probe_precision
        self.device = torch.device(device) if device is not None else get_default_device()
        self.dtype = dtype if dtype is not None else get_default_dtype()
        self.n: Optional[int] = None
        self.nr: Optional[int] = None
        self.isotropic_coef: Optional[float] = None
        self.iterations: int = 0
        self.rho: Optional[float] = None
        self._h = _PrecondHandle()
        self._cache = {}

    # ------------------------------------------------------------------
    def _probes(self, n: int, nr: int, seed: Optional[int], dtype: torch.dtype) -> Optional[torch.Tensor]:
        if self.probe_source == "device":
            return None
        gen = torch.Generator(device="cpu")
        if seed is not None:
            gen.manual_seed(seed)
        return torch.randn(n, nr, dtype=dtype, generator=gen).to(compute_device())

    def _device_probes(self, n: int, nr: int, seed: Optional[int], dtype: torch.dtype) -> torch.Tensor:
        """The library's own N(0,1) probe stream (what the native build draws without probes)."""
        t = torch.empty(n, nr, dtype=dtype, device=compute_device())
        N.check(N.lib.laker_randn(N.ptr(t), n * nr, N.dtype_code(dtype),
                                  0 if seed is None else int(seed) & ((1 << 63) - 1), N.stream_ptr()))
        return t

    def _power_probes(self, operator, probes: torch.Tensor) -> torch.Tensor:
        """``probe_strategy="power_iter"`` (laker/preconditioner.py:174-184): the first
        ``max(1, int(0.25 N_r))`` probes take ``power_iter_steps`` rounds of
        ``block = orth(A block)``, the rest stay Gaussian.  The orthonormalisation
        is the library's Cholesky-QR2 (equal to the reference's Householder Q up to
        column signs, to which the build is invariant)."""
        nr = probes.shape[1]
        n_power = max(1, int(nr * 0.25))
        block = probes[:, :n_power].contiguous()
        for _ in range(self.power_iter_steps):
            block = operator(block).to(device=compute_device(), dtype=probes.dtype).contiguous()
            N.check(N.lib.laker_orthonormalize(N.ptr(block), block.shape[0], block.shape[1],
                                               N.dtype_code(block.dtype), N.stream_ptr()))
        return torch.cat([block, probes[:, n_power:]], dim=1).contiguous()

    def build(self, operator: Callable[[torch.Tensor], torch.Tensor], n: int,
              seed: Optional[int] = None) -> "CCCPPreconditioner":
        """Learn ``P`` for an ``n x n`` operator (laker/preconditioner.py:144-284)."""
        from .kernels import AttentionKernelOperator

        self.n = n
        self.nr = default_num_probes(n, self.num_probes)
        nr = self.nr
        self._h.close()
        self._cache = {}
        native = getattr(operator, "__self__", None)
        its = ctypes.c_int32()
        if isinstance(native, AttentionKernelOperator) and getattr(operator, "__func__", None) in (
            AttentionKernelOperator.matvec, AttentionKernelOperator.matvec_impl,
            AttentionKernelOperator.__call__,
        ):
            self.dtype = native.dtype
            probes = self._probes(n, nr, seed, native.dtype)
            if self.probe_strategy == "power_iter":
                if probes is None:
                    probes = self._device_probes(n, nr, seed, native.dtype)
                probes = self._power_probes(operator, probes)
            N.check(N.lib.laker_precond_build(
                native.handle, nr, self.gamma, self.epsilon, self.base_rho, self.max_iter, self.tol,
                N.ptr(probes), 0 if seed is None else int(seed) & ((1 << 63) - 1),
                _PROBE_PREC.get(self.probe_precision, N.PREC_FP32), N.stream_ptr(),
                ctypes.byref(self._h.h), ctypes.byref(its)))
        else:
            # generic callable operator: responses from the callable, rest on device
            probes = self._probes(n, nr, seed, self.dtype)
            if probes is None:
                gen = torch.Generator(device=compute_device())
                gen.manual_seed(0 if seed is None else seed)
                probes = torch.randn(n, nr, dtype=self.dtype, device=compute_device(), generator=gen)
            if self.probe_strategy == "power_iter":
                probes = self._power_probes(operator, probes)
            u = operator(probes)
            u = u.to(device=compute_device(), dtype=self.dtype).contiguous()
            N.check(N.lib.laker_precond_build_from_responses(
                N.ptr(u), n, nr, N.dtype_code(self.dtype), self.gamma, self.epsilon, self.base_rho,
                self.max_iter, self.tol, N.stream_ptr(), ctypes.byref(self._h.h), ctypes.byref(its)))
        a = ctypes.c_double()
        rho = ctypes.c_double()
        N.check(N.lib.laker_precond_info(self._h.h, ctypes.byref(a), None, ctypes.byref(rho), None, None,
                                         None, None))
        self.isotropic_coef = a.value
        self.rho = rho.value
        self.iterations = its.value
        if self.verbose:
            logger.info("CCCP preconditioner built in %d iterations (n=%d, N_r=%d)", self.iterations, n, nr)
        return self

    @property
    def handle(self):
        return self._h.h

    @property
    def built(self) -> bool:
        return bool(self._h.h)

    @classmethod
    def from_factors(cls, q_basis: torch.Tensor, qr_r: torch.Tensor, q_basis_matrix: torch.Tensor,
                     s_matrix: torch.Tensor, isotropic_coef: float, rho: float = 0.0,
                     **kwargs) -> "CCCPPreconditioner":
        """A built preconditioner from stored factors (``q_basis`` n x N_r, ``qr_r``,
        ``q_basis_matrix`` and ``s_matrix`` N_r x N_r), without the probe matvec:
        how a saved model reloads its CCCP factors."""
        pre = cls(verbose=False, dtype=q_basis.dtype, **kwargs)
        dev = compute_device()
        q = q_basis.detach().to(dev).contiguous()
        n, nr = q.shape
        if q.dtype not in (torch.float32, torch.float64):
            raise ValueError("q_basis must be float32 or float64")
        mats = [m.detach().to(dev, torch.float64).contiguous() for m in (qr_r, q_basis_matrix, s_matrix)]
        for m in mats:
            if tuple(m.shape) != (nr, nr):
                raise ValueError(f"factor matrices must be ({nr}, {nr}), got {tuple(m.shape)}")
        N.check(N.lib.laker_precond_from_factors(N.ptr(q), n, nr, N.dtype_code(q.dtype), float(isotropic_coef),
                                                 float(rho), N.ptr(mats[0]), N.ptr(mats[1]), N.ptr(mats[2]),
                                                 N.stream_ptr(), ctypes.byref(pre._h.h)))
        pre.n, pre.nr, pre.dtype = n, nr, q.dtype
        pre.isotropic_coef, pre.rho = float(isotropic_coef), float(rho)
        return pre

    # ------------------------------------------------------------------
    def _copy(self, which: int, shape, dtype) -> torch.Tensor:
        if which not in self._cache:
            if not self.built:
                return None
            t = torch.empty(shape, dtype=dtype, device=compute_device())
            N.check(N.lib.laker_precond_copy_out(self._h.h, which, N.ptr(t), N.stream_ptr()))
            self._cache[which] = t
        return self._cache[which]

    @property
    def q_basis(self) -> Optional[torch.Tensor]:
        """``(n, N_r)`` orthonormal basis of the normalised probe responses."""
        return self._copy(0, (self.n, self.nr), self.dtype) if self.built else None

    @property
    def qr_r(self) -> Optional[torch.Tensor]:
        return self._copy(1, (self.nr, self.nr), torch.float64) if self.built else None

    @property
    def q_basis_matrix(self) -> Optional[torch.Tensor]:
        return self._copy(2, (self.nr, self.nr), torch.float64) if self.built else None

    @property
    def s_matrix(self) -> Optional[torch.Tensor]:
        """``M^{-1/2} - a^{-1/2} I`` used by the apply."""
        return self._copy(3, (self.nr, self.nr), torch.float64) if self.built else None

    def _eig(self):
        if "eig" not in self._cache:
            m = self.q_basis_matrix + self.isotropic_coef * torch.eye(self.nr, dtype=torch.float64,
                                                                        device=compute_device())
            self._cache["eig"] = eigh_stable(m, eps=self.epsilon)
        return self._cache["eig"]

    @property
    def q_eigenvalues(self) -> Optional[torch.Tensor]:
        """Eigenvalues of ``M = a I + C`` (introspection; computed on first access)."""
        return self._eig()[0] if self.built else None

    @property
    def q_eigenvectors(self) -> Optional[torch.Tensor]:
        return self._eig()[1] if self.built else None

    # ------------------------------------------------------------------
    def apply(self, x: torch.Tensor) -> torch.Tensor:
        """``P x`` for ``(n,)`` or ``(n, k)`` (laker/preconditioner.py:286-352)."""
        if not self.built:
            raise RuntimeError("Preconditioner has not been built. Call build() first.")
        if x.dim() not in (1, 2):
            raise ValueError(f"x must be 1-D or 2-D, got shape {x.shape}")
        out_dev = x.device
        xd = x.to(device=compute_device(), dtype=self.dtype).contiguous()
        x2 = xd.unsqueeze(1) if xd.dim() == 1 else xd
        z = torch.empty_like(x2)
        N.check(N.lib.laker_precond_apply(self._h.h, N.ptr(x2), x2.shape[1], x2.shape[1], N.ptr(z),
                                          z.shape[1], N.stream_ptr()))
        z = z.squeeze(1) if x.dim() == 1 else z
        return z.to(out_dev)

    __call__ = apply

    def apply_1d(self, x: torch.Tensor) -> torch.Tensor:
        return self.apply(x)

    def apply_2d(self, x: torch.Tensor) -> torch.Tensor:
        return self.apply(x)

    def apply_impl(self, x: torch.Tensor) -> torch.Tensor:
        return self.apply(x)

    def to_dense(self) -> torch.Tensor:
        """Dense ``P = Sigma^{-1/2}`` (laker/preconditioner.py:354-379), by applying P to I."""
        if not self.built:
            raise RuntimeError("Preconditioner has not been built.")
        eye = torch.eye(self.n, dtype=self.dtype, device=compute_device())
        return self.apply(eye)

    def close(self) -> None:
        self._h.close()
        self._cache = {}

    @staticmethod
    def shrinkage(nr: int, n: int, gamma: float, base_rho: float) -> float:
        return adaptive_shrinkage_rho(nr, n, gamma, base_rho)


class AdaptivePreconditioner:
    """Selects Jacobi or CCCP from a cheap conditioning estimate (reference
    laker/preconditioner.py:382-599).

    ``build`` draws ``min(10, n)`` probes, applies the operator once to all of
    them, runs 5 power-iteration steps from the first response for
    ``lambda_max = v^T A v`` and takes ``lambda_min`` as the smallest probe
    Rayleigh quotient; with ``kappa = lambda_max / max(lambda_min, 1e-12)`` it
    picks ``JacobiPreconditioner(diagonal)`` if ``kappa < 1e3`` and a diagonal is
    given, CCCP if ``kappa < 1e6``, else CCCP with twice the probes.  Products
    are the operator's (native when it is the library operator); dot products
    and norms are the library's FP64 column reductions.  Probes follow
    ``probe_source`` as for :class:`CCCPPreconditioner`.
    """

    def __init__(self, gamma: float = 1e-1, num_probes: Optional[int] = None, epsilon: float = 1e-8,
                 base_rho: float = 0.05, max_iter: int = 200, tol: float = 1e-6, verbose: bool = True,
                 device: Optional[torch.device] = None, dtype: Optional[torch.dtype] = None,
                 probe_strategy: str = "gaussian", power_iter_steps: int = 3,
                 probe_source: str = "reference", probe_precision: str = "fp32"):
        if probe_source not in ("reference", "device"):
            raise ValueError("probe_source must be 'reference' or 'device'")
        self.gamma = float(gamma)
        self.num_probes = num_probes
        self.epsilon = float(epsilon)
        self.base_rho = float(base_rho)
        self.max_iter = int(max_iter)
        self.tol = float(tol)
        self.verbose = verbose
        self.probe_strategy = probe_strategy
        self.power_iter_steps = int(power_iter_steps)
        self.probe_source = probe_source
        self.probe_precision = probe_precision
        self.device = torch.device(device) if device is not None else get_default_device()
        self.dtype = dtype if dtype is not None else get_default_dtype()
        self.inner = None
        self.inner_name: Optional[str] = None
        self.condition_estimate: Optional[float] = None
        self.lambda_max: Optional[float] = None
        self.lambda_min: Optional[float] = None

    def _cccp(self, num_probes: Optional[int]) -> CCCPPreconditioner:
        return CCCPPreconditioner(num_probes=num_probes, gamma=self.gamma, epsilon=self.epsilon,
                                  base_rho=self.base_rho, max_iter=self.max_iter, tol=self.tol,
                                  verbose=self.verbose, device=self.device, dtype=self.dtype,
                                  probe_strategy=self.probe_strategy, power_iter_steps=self.power_iter_steps,
                                  probe_source=self.probe_source, probe_precision=self.probe_precision)

    def build(self, operator: Callable[[torch.Tensor], torch.Tensor], n: int,
              diagonal: Optional[torch.Tensor] = None, seed: Optional[int] = None) -> "AdaptivePreconditioner":
        """Estimate the conditioning and build the selected preconditioner
        (laker/preconditioner.py:469-574)."""
        from .kernels import AttentionKernelOperator
        from .solvers import JacobiPreconditioner, _coldot

        native = getattr(operator, "__self__", None)
        dtype = native.dtype if isinstance(native, AttentionKernelOperator) else self.dtype
        m = min(10, n)
        dev = compute_device()
        if self.probe_source == "reference":
            gen = torch.Generator(device="cpu")
            if seed is not None:
                gen.manual_seed(seed)
            probes = torch.randn(n, m, dtype=dtype, generator=gen).to(dev)
        else:
            probes = torch.empty(n, m, dtype=dtype, device=dev)
            N.check(N.lib.laker_randn(N.ptr(probes), n * m, N.dtype_code(dtype),
                                      0 if seed is None else int(seed) & ((1 << 63) - 1), N.stream_ptr()))
        op_probes = operator(probes).to(device=dev, dtype=dtype)
        v = op_probes[:, 0].contiguous()
        for _ in range(5):
            v = operator(v.unsqueeze(1)).squeeze(1).to(device=dev, dtype=dtype).contiguous()
            v = v / float(_coldot(v, v)[0]) ** 0.5
        lam_max = float(_coldot(v, operator(v).to(device=dev, dtype=dtype))[0])
        rayleigh = _coldot(probes, op_probes) / _coldot(probes, probes)
        lam_min = float(rayleigh.min())
        cond = lam_max / max(lam_min, 1e-12)
        self.lambda_max, self.lambda_min, self.condition_estimate = lam_max, lam_min, cond
        if cond < 1e3 and diagonal is not None:
            self.inner = JacobiPreconditioner(diagonal)
            self.inner_name = "jacobi"
        elif cond < 1e6:
            self.inner = self._cccp(self.num_probes).build(operator, n, seed=seed)
            self.inner_name = "cccp"
        else:
            base = self.num_probes if self.num_probes is not None else max(200, int(2 * n ** 0.5))
            self.inner = self._cccp(base * 2).build(operator, n, seed=seed)
            self.inner_name = "cccp_aggressive"
        if self.verbose:
            logger.info("AdaptivePreconditioner selected %s (kappa~%.2e)", self.inner_name, cond)
        return self

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        """Apply the selected preconditioner (laker/preconditioner.py:576-599)."""
        if self.inner is None:
            raise RuntimeError("Preconditioner has not been built. Call build() first.")
        return self.inner.apply(x)

    __call__ = apply

    @property
    def n(self):
        return getattr(self.inner, "n", None)

    def close(self) -> None:
        if self.inner is not None:
            self.inner.close()
