"""CPU oracle for the LAKER solver path -- TEST INFRASTRUCTURE ONLY.

This module is a plain torch-CPU / NumPy restatement of the reference
algorithm for the exact attention-kernel matvec, the PCG loop and the CCCP
preconditioner.  It exists to *check* the HIP product path and to time the
CPU baseline.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it; the product package never
does, and the product fails loudly when its HIP library is missing.

Every function cites the reference file:line it restates (paths relative to
the reference repository root).  The reference's own files are never
imported here.

Pinning (what checks each function independently of this restatement):

* ``exp_safe`` / ``attention_matvec`` / ``diagonal`` / ``to_dense`` /
  ``kernel_eval``: closed form ``lam*x + exp(E E^T) x`` against an FP64 dense
  product, brute-force double loops at tiny n, chunked == dense, the
  clamp threshold (tests/test_oracle.py).
* ``pcg_solve``: exact-preconditioner one-step convergence, diagonal SPD
  systems against closed forms, FP64 Cholesky agreement, breakdown on an
  indefinite diagonal, Frobenius stopping rule on synthetic columns.
* ``adaptive_shrinkage_rho``: closed-form values (n = 2000, N_r = 200, gamma = 0.1 -> 0.5).
* ``cccp_build`` / ``cccp_apply``: invariants (a > 0, tr Sigma = n, M SPD,
  P linear and symmetric, P == Sigma^{-1/2} by dense eigh at small n),
  an FP64 NumPy restatement with a different factorisation route
  (Cholesky instead of eigh) fed the same probes, and the golden vectors
  written by ``tests/golden/make_golden.py``.
* ``PositionEmbeddingOracle``: golden vectors written from the reference
  module by ``tests/golden/make_golden.py`` (exact float equality).
* ``dense_solve_fp64``: the paper's n=3 worked example
  (reference tests/test_models.py:37-88) and residual checks.
* ``jacobi_inverse`` / ``jacobi_apply``: closed forms on diagonal operators
  (the exact preconditioner of a diagonal system: one PCG step), the eps
  clamp, reference tests/test_solvers.py:59-73.
* ``adaptive_select``: diagonal operators of known spectrum (the power
  iterate tends to lambda_max, Rayleigh quotients bound lambda_min from
  above) and one operator per selection regime.
* ``power_iter_probes``: orthonormality of the power block, its span equal
  to span(A^s R_block) by projection, untouched Gaussian columns, and the
  reference's lambda_max(P A) comparison (tests/test_preconditioner.py:86-150).
* ``continuation_schedule`` / ``fit_path`` / ``partial_fit``: the closed-form
  geometric schedule, each path solution against the FP64 dense solve of its
  own lambda, and the warm-start identity (a solve started at the exact
  solution takes no step).
"""

from __future__ import annotations

import math
from typing import Callable, Optional

import numpy as np
import torch

# ---------------------------------------------------------------------------
# Kernel operator (reference laker/kernels.py)
# ---------------------------------------------------------------------------


def exp_safe(gram: torch.Tensor, skip_clamp: bool = False) -> torch.Tensor:
    """``exp(min(g, 80|700))``; clamp skipped when ``skip_clamp``.

    Restates laker/kernels.py:80-113 (out-of-place form).
    """
    if not skip_clamp:
        max_val = 80.0 if gram.dtype == torch.float32 else 700.0
        return torch.exp(gram.clamp(max=max_val))
    return torch.exp(gram)


def reference_clamp_branch_matvec(embeddings: torch.Tensor, lambda_reg: float, x: torch.Tensor) -> torch.Tensor:
    """What the reference's dense matvec returns when its clamp branch is taken.

    laker/kernels.py:128-130 calls ``exp_safe(gram, out=gram, skip_clamp=False)``
    and discards the return value; with the clamp active exp_safe clamps
    ``out`` in place and returns ``torch.exp(out)`` as a new tensor
    (laker/kernels.py:103-108), so the product uses ``min(E E^T, 80)``
    without the exponential.  Kept only to pin that behaviour against the
    reference's golden vector; the oracle's ``attention_matvec`` (and the
    product) compute the documented ``exp(min(E E^T, 80))``.
    """
    max_val = 80.0 if embeddings.dtype == torch.float32 else 700.0
    return lambda_reg * x + (embeddings @ embeddings.T).clamp(max=max_val) @ x


def skip_clamp_flag(embeddings: torch.Tensor) -> bool:
    """``max_i ||e_i||^2 < 80 (fp32) | 700 (fp64)`` -- laker/kernels.py:181-185."""
    max_sq = torch.sum(embeddings**2, dim=1).max().item()
    limit = 80.0 if embeddings.dtype == torch.float32 else 700.0
    return max_sq < limit


def attention_matvec(
    embeddings: torch.Tensor,
    lambda_reg: float,
    x: torch.Tensor,
    chunk_size: Optional[int] = None,
    skip_clamp: Optional[bool] = None,
) -> torch.Tensor:
    """``(lam I + exp(E E^T)) x`` with the reference's dense / 1-D / 2-D tiling.

    Restates laker/kernels.py:121-130 (dense) and :210-270 (chunked): 1-D row
    chunks when ``chunk*n*elt <= 64 MiB``, else 2-D tiles with a per-row-block
    accumulator.
    """
    if x.dim() not in (1, 2):
        raise ValueError(f"x must be 1-D or 2-D, got shape {x.shape}")
    if skip_clamp is None:
        skip_clamp = skip_clamp_flag(embeddings)
    n = embeddings.shape[0]
    if chunk_size is None or n <= chunk_size:
        gram = embeddings @ embeddings.T
        gram = exp_safe(gram, skip_clamp)
        return lambda_reg * x + gram @ x
    out = lambda_reg * x
    elt = 4 if embeddings.dtype == torch.float32 else 8
    if chunk_size * n * elt <= 64 * 1024 * 1024:
        for s in range(0, n, chunk_size):
            e = min(s + chunk_size, n)
            g = exp_safe(embeddings[s:e] @ embeddings.T, skip_clamp)
            out[s:e] += g @ x
        return out
    for i0 in range(0, n, chunk_size):
        i1 = min(i0 + chunk_size, n)
        acc = torch.zeros((i1 - i0,) + tuple(x.shape[1:]), dtype=x.dtype)
        for j0 in range(0, n, chunk_size):
            j1 = min(j0 + chunk_size, n)
            g = exp_safe(embeddings[i0:i1] @ embeddings[j0:j1].T, skip_clamp)
            acc += g @ x[j0:j1]
        out[i0:i1] += acc
    return out


def diagonal(embeddings: torch.Tensor, lambda_reg: float) -> torch.Tensor:
    """``lam + exp(||e_i||^2)`` (no clamp) -- laker/kernels.py:272-283."""
    return lambda_reg + torch.exp(torch.sum(embeddings**2, dim=1))


def to_dense(embeddings: torch.Tensor, lambda_reg: float) -> torch.Tensor:
    """Dense ``lam I + exp(E E^T)`` (no clamp) -- laker/kernels.py:285-299."""
    g = torch.exp(embeddings @ embeddings.T)
    g.diagonal().add_(lambda_reg)
    return g


def kernel_eval(x: torch.Tensor, y: torch.Tensor, skip_clamp: bool) -> torch.Tensor:
    """``exp(x y^T)`` with the operator's clamp flag -- laker/kernels.py:301-354."""
    return exp_safe(x @ y.T, skip_clamp)


# ---------------------------------------------------------------------------
# PCG (reference laker/solvers.py:48-394)
# ---------------------------------------------------------------------------


class PCGResult:
    def __init__(self, x, iterations, residual_norm, converged):
        self.x = x
        self.iterations = iterations
        self.residual_norm = residual_norm
        self.converged = converged


def pcg_solve(
    operator: Callable[[torch.Tensor], torch.Tensor],
    preconditioner: Callable[[torch.Tensor], torch.Tensor],
    rhs: torch.Tensor,
    x0: Optional[torch.Tensor] = None,
    tol: float = 1e-10,
    max_iter: Optional[int] = None,
    breakdown_eps: Optional[float] = None,
    restart_freq: Optional[int] = None,
) -> PCGResult:
    """PCG with host-synced scalars.

    Restates laker/solvers.py:130-187 (set-up, zero-RHS early return),
    :189-295 (1-D loop: breakdown test, in-place updates, ``iterations=it+1``
    on convergence, ``max_iter`` otherwise) and :297-394 (2-D loop: per-column
    alpha/beta, global Frobenius stopping ratio, breakdown if any column).
    """
    if rhs.dim() not in (1, 2):
        raise ValueError(f"rhs must be 1-D or 2-D, got shape {rhs.shape}")
    n = rhs.shape[0]
    max_iter = max_iter if max_iter is not None else n
    if x0 is None:
        x = torch.zeros_like(rhs)
        r = rhs.clone()
    else:
        x = x0.clone()
        r = rhs - operator(x)
    z = preconditioner(r)
    p = z.clone()
    rhs_norm = torch.linalg.norm(rhs).item()
    if rhs_norm == 0:
        return PCGResult(x, 0, float("inf"), True)
    eps = breakdown_eps if breakdown_eps is not None else torch.finfo(p.dtype).eps ** 0.5
    one_d = rhs.dim() == 1
    rz_old = torch.dot(r, z).item() if one_d else torch.sum(r * z, dim=0)
    res_norm = float("inf")
    for it in range(max_iter):
        ap = operator(p)
        if one_d:
            pap = torch.dot(p, ap).item()
            if pap <= -eps * torch.linalg.norm(p).item() * torch.linalg.norm(ap).item():
                raise RuntimeError("PCG breakdown: non-positive curvature detected (p^T A p <= 0).")
            alpha = rz_old / pap
            x.add_(p, alpha=alpha)
            r.add_(ap, alpha=-alpha)
        else:
            pap = torch.sum(p * ap, dim=0)
            if torch.any(pap <= -eps * torch.linalg.norm(p, dim=0) * torch.linalg.norm(ap, dim=0)):
                raise RuntimeError("PCG breakdown: non-positive curvature detected (p^T A p <= 0).")
            alpha = rz_old / pap
            x.add_(p * alpha.unsqueeze(0))
            r.add_(ap * (-alpha.unsqueeze(0)))
        if restart_freq is not None and (it + 1) % restart_freq == 0:
            r = rhs - operator(x)
        res_norm = torch.linalg.norm(r).item()
        if res_norm / rhs_norm <= tol:
            return PCGResult(x, it + 1, res_norm, True)
        z = preconditioner(r)
        if one_d:
            rz_new = torch.dot(r, z).item()
            p.mul_(rz_new / rz_old).add_(z)
        else:
            rz_new = torch.sum(r * z, dim=0)
            p.mul_((rz_new / rz_old).unsqueeze(0)).add_(z)
        rz_old = rz_new
    return PCGResult(x, max_iter, res_norm, False)


# ---------------------------------------------------------------------------
# CCCP preconditioner (reference laker/preconditioner.py, laker/utils.py)
# ---------------------------------------------------------------------------


def adaptive_shrinkage_rho(num_probes: int, n: int, gamma: float, base_rho: float = 0.05) -> float:
    """Shrinkage rho -- laker/utils.py:68-106."""
    if num_probes >= n:
        return base_rho
    ratio = num_probes / n
    rho = base_rho + (1.0 - base_rho) * (1.0 - ratio) * min(1.0, gamma * 10.0)
    return float(min(rho, 0.5))


def default_num_probes(n: int, num_probes: Optional[int] = None) -> int:
    """``min(num_probes or max(200, int(2 sqrt n)), n)`` -- preconditioner.py:161-163."""
    nr = num_probes if num_probes is not None else max(200, int(2 * n**0.5))
    return min(nr, n)


REFERENCE_DEFAULT_PROBE_SEED = 67280421310721


def reference_probes(n: int, nr: int, seed: Optional[int], dtype=torch.float32) -> torch.Tensor:
    """The probe matrix ``randn(n, nr)`` the reference draws on a CPU generator.

    laker/preconditioner.py:168-171: a fresh ``torch.Generator`` (default seed
    67280421310721 when ``seed`` is None, else ``manual_seed(seed)``).
    """
    gen = torch.Generator(device="cpu")
    if seed is not None:
        gen.manual_seed(seed)
    return torch.randn(n, nr, dtype=dtype, generator=gen)


def cccp_build(
    operator: Callable[[torch.Tensor], torch.Tensor],
    n: int,
    probes: torch.Tensor,
    gamma: float = 0.1,
    epsilon: float = 1e-8,
    base_rho: float = 0.05,
    max_iter: int = 200,
    tol: float = 1e-6,
) -> dict:
    """CCCP fixed point in the Q basis, in the dtype of ``probes``.

    Restates laker/preconditioner.py:144-284: one operator application to the
    probes, column normalisation clamped at ``epsilon``, reduced QR, the
    shrinkage-regularised loop (eigh of M clamped at ``epsilon``, probe
    denominators, weights ``(n/N_r)/d``, ``F = (R W R^T + gamma I)/s``, lerp
    to I by rho, trace normalisation to n, relative-change stopping test) and
    the final eigh.
    """
    dtype = probes.dtype
    nr = probes.shape[1]
    u = operator(probes)
    u = u / torch.linalg.norm(u, dim=0, keepdim=True).clamp(min=epsilon)
    q, rqr = torch.linalg.qr(u, mode="reduced")
    rho = adaptive_shrinkage_rho(nr, n, gamma, base_rho)
    s = 1.0 + gamma / n
    a = 1.0
    c = torch.zeros(nr, nr, dtype=dtype)
    eye = torch.eye(nr, dtype=dtype)
    it_done = 0
    for it in range(max_iter):
        a_prev = a
        c_prev = c.clone()
        lam, vec = torch.linalg.eigh(eye * a + c)
        lam = lam.clamp(min=epsilon)
        proj = vec.T @ rqr
        den = torch.diagonal(proj.T @ (lam.reciprocal().unsqueeze(-1) * proj)) + epsilon
        w = (n / nr) / den
        f = ((rqr * w.unsqueeze(0)) @ rqr.T + gamma * eye) * (1.0 / s)
        f_rho = torch.lerp(f, eye, rho)
        a_rho = (1.0 - rho) * gamma / s + rho
        t = n / (a_rho * (n - nr) + f_rho.diagonal().sum())
        a = float(t * a_rho)
        c = f_rho * t
        c.diagonal().sub_(a)
        iso_rel = abs(a - a_prev) / (abs(a_prev) + 1e-12)
        q_rel = torch.norm(c - c_prev, p="fro").item() / (torch.norm(c_prev, p="fro").item() + 1e-12)
        it_done = it + 1
        if max(iso_rel, q_rel) < tol:
            break
    lam, vec = torch.linalg.eigh(eye * a + c)
    lam = lam.clamp(min=epsilon)
    return {
        "q_basis": q,
        "qr_r": rqr,
        "isotropic_coef": a,
        "q_basis_matrix": c,
        "q_eigenvalues": lam,
        "q_eigenvectors": vec,
        "iterations": it_done,
        "rho": rho,
        "nr": nr,
    }


def cccp_apply(x: torch.Tensor, f: dict) -> torch.Tensor:
    """``a^{-1/2} x + Q V ((lam^{-1/2} - a^{-1/2}) * (V^T Q^T x))``.

    Restates laker/preconditioner.py:52-68.
    """
    inv = f["isotropic_coef"] ** (-0.5)
    c = f["q_eigenvectors"].T @ (f["q_basis"].T @ x)
    scale = f["q_eigenvalues"].rsqrt() - inv
    c = scale * c if x.dim() == 1 else scale.unsqueeze(-1) * c
    return inv * x + f["q_basis"] @ (f["q_eigenvectors"] @ c)


def cccp_loop_fp64_cholesky(rqr: np.ndarray, n: int, gamma: float, epsilon: float, base_rho: float,
                            max_iter: int, tol: float):
    """Independent FP64 restatement of the CCCP loop using Cholesky.

    Same fixed point as laker/preconditioner.py:207-273, but ``diag(R^T M^-1 R)``
    comes from a Cholesky factor of ``M = aI + C`` instead of an eigh; the two
    agree while the eigenvalue clamp is inactive (M = t F_rho >= t rho I > 0).
    Returns ``(a, C, iterations)``.
    """
    nr = rqr.shape[0]
    rho = adaptive_shrinkage_rho(nr, n, gamma, base_rho)
    s = 1.0 + gamma / n
    a = 1.0
    c = np.zeros((nr, nr))
    eye = np.eye(nr)
    it_done = 0
    for it in range(max_iter):
        a_prev, c_prev = a, c.copy()
        l_fac = np.linalg.cholesky(a * eye + c)
        z = np.linalg.solve(l_fac, rqr)
        den = np.sum(z * z, axis=0) + epsilon
        w = (n / nr) / den
        f = ((rqr * w[None, :]) @ rqr.T + gamma * eye) / s
        f_rho = (1 - rho) * f + rho * eye
        a_rho = (1 - rho) * gamma / s + rho
        t = n / (a_rho * (n - nr) + np.trace(f_rho))
        a = t * a_rho
        c = t * f_rho - a * eye
        iso_rel = abs(a - a_prev) / (abs(a_prev) + 1e-12)
        q_rel = np.linalg.norm(c - c_prev) / (np.linalg.norm(c_prev) + 1e-12)
        it_done = it + 1
        if max(iso_rel, q_rel) < tol:
            break
    return a, c, it_done


# ---------------------------------------------------------------------------
# Embedding (reference laker/embeddings.py:94-242)
# ---------------------------------------------------------------------------


class PositionEmbeddingOracle:
    """Fourier features + 2-layer tanh MLP drawn from one CPU generator.

    Restates laker/embeddings.py:94-218 (draw order: freq=randn(in, F)/sigma,
    phase=rand(F)*2pi, then per Linear weight rand*2b-b and bias rand*2b-b with
    b=sqrt(1/fan_in)) and :220-242 (forward).
    """

    def __init__(self, input_dim: int, embedding_dim: int, num_fourier: Optional[int] = None,
                 sigma: float = 10.0, seed: int = 42, dtype=torch.float32):
        nf = num_fourier if num_fourier is not None else embedding_dim * 2
        hidden = max(embedding_dim, nf // 2)
        gen = torch.Generator(device="cpu").manual_seed(seed)
        self.freq = torch.randn(input_dim, nf, generator=gen, dtype=dtype) / sigma
        self.phase = torch.rand(nf, generator=gen, dtype=dtype) * 2.0 * math.pi
        self.layers = []
        for fan_in, fan_out in ((nf, hidden), (hidden, embedding_dim)):
            b = math.sqrt(1.0 / fan_in)
            w = torch.rand((fan_out, fan_in), generator=gen, dtype=dtype) * (2 * b) - b
            bias = torch.rand((fan_out,), generator=gen, dtype=dtype) * (2 * b) - b
            self.layers.append((w, bias))

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        if x.dim() == 1:
            x = x.unsqueeze(0)
        h = torch.cos(2.0 * math.pi * (x @ self.freq) + self.phase)
        (w1, b1), (w2, b2) = self.layers
        h = torch.tanh(torch.nn.functional.linear(h, w1, b1))
        return torch.nn.functional.linear(h, w2, b2)


# ---------------------------------------------------------------------------
# Pipeline (reference laker/core.py, laker/models.py)
# ---------------------------------------------------------------------------


def auto_chunk(n: int, chunk_size: Optional[int] = None) -> Optional[int]:
    """``n > 5000 -> max(1024, min(n // 10, 8192))`` -- laker/core.py:340-345."""
    if chunk_size is None and n > 5000:
        return max(1024, min(n // 10, 8192))
    return chunk_size


def fit(x: torch.Tensor, y: torch.Tensor, embedding_dim: int = 10, lambda_reg: float = 1e-2,
        gamma: float = 0.1, num_probes: Optional[int] = None, epsilon: float = 1e-8,
        base_rho: float = 0.05, cccp_max_iter: int = 200, cccp_tol: float = 1e-6,
        pcg_tol: float = 1e-6, pcg_max_iter: int = 1000, chunk_size: Optional[int] = None,
        seed: Optional[int] = None, embeddings: Optional[torch.Tensor] = None,
        dtype=torch.float32) -> dict:
    """``LAKERRegressor.fit`` on the CPU in ``dtype``.

    Restates laker/models.py:392-455 -> laker/core.py:247-302 (embeddings),
    :307-367 (exact operator + auto chunk), :461-521 (CCCP), :526-575 (PCG).
    """
    x = x.to(dtype)
    y = y.to(dtype)
    if embeddings is None:
        emb = PositionEmbeddingOracle(x.shape[1], embedding_dim, dtype=dtype)
        with torch.no_grad():
            e = emb(x)
    else:
        emb = None
        e = embeddings.to(dtype)
    n = e.shape[0]
    chunk = auto_chunk(n, chunk_size)
    skip = skip_clamp_flag(e)

    def op(v):
        return attention_matvec(e, lambda_reg, v, chunk, skip)

    nr = default_num_probes(n, num_probes)
    probes = reference_probes(n, nr, seed, dtype)
    fac = cccp_build(op, n, probes, gamma, epsilon, base_rho, cccp_max_iter, cccp_tol)
    res = pcg_solve(op, lambda v: cccp_apply(v, fac), y, tol=pcg_tol, max_iter=pcg_max_iter)
    return {"embeddings": e, "alpha": res.x, "iterations": res.iterations, "factors": fac,
            "embedding_model": emb, "skip_clamp": skip, "chunk": chunk,
            "residual_norm": res.residual_norm}


def predict(query_embeddings: torch.Tensor, embeddings: torch.Tensor, alpha: torch.Tensor,
            skip_clamp: bool, clamp_always: bool = False) -> torch.Tensor:
    """``exp(Eq E^T) alpha`` -- laker/core.py:580-661 (the tiled branch always clamps)."""
    return exp_safe(query_embeddings @ embeddings.T, skip_clamp and not clamp_always) @ alpha


def predict_variance(query_embeddings, embeddings, lambda_reg, op, pre, skip_clamp,
                     pcg_tol=1e-6, pcg_max_iter=1000):
    """``diag K_qq - colsum(K_xq * A^-1 K_xq)`` clamped >= 0 -- laker/core.py:666-777."""
    k_xq = kernel_eval(embeddings, query_embeddings, skip_clamp)
    res = pcg_solve(op, pre, k_xq, tol=pcg_tol, max_iter=pcg_max_iter)
    k_qq = kernel_eval(query_embeddings, query_embeddings, skip_clamp).diagonal()
    return (k_qq - torch.sum(k_xq * res.x, dim=0)).clamp(min=0.0), res


# ---------------------------------------------------------------------------
# FP64 dense oracle
# ---------------------------------------------------------------------------


def dense_matrix_fp64(embeddings, lambda_reg: float, scale: float = 1.0) -> np.ndarray:
    """``lam I + exp(scale E E^T)`` in FP64 (NumPy)."""
    e = np.asarray(embeddings, dtype=np.float64)
    a = np.exp(scale * (e @ e.T))
    a[np.diag_indices_from(a)] += lambda_reg
    return a


def dense_solve_fp64(embeddings, lambda_reg: float, rhs) -> np.ndarray:
    """Exact solution of ``(lam I + exp(E E^T)) x = rhs`` by FP64 Cholesky."""
    import scipy.linalg as sla

    a = dense_matrix_fp64(embeddings, lambda_reg)
    c = sla.cho_factor(a, lower=True, check_finite=False)
    return sla.cho_solve(c, np.asarray(rhs, dtype=np.float64), check_finite=False)


def matvec_fp64(embeddings, lambda_reg: float, x, scale: float = 1.0, clamp: Optional[float] = None,
                block: int = 2048) -> np.ndarray:
    """Row-blocked FP64 ``(lam I + exp(min(scale E E^T, clamp))) x`` (NumPy)."""
    e = np.asarray(embeddings, dtype=np.float64)
    xv = np.asarray(x, dtype=np.float64)
    out = lambda_reg * xv.copy()
    for i0 in range(0, e.shape[0], block):
        g = scale * (e[i0:i0 + block] @ e.T)
        if clamp is not None:
            np.minimum(g, clamp, out=g)
        out[i0:i0 + block] += np.exp(g) @ xv
    return out


def cross_matvec_fp64(query, embeddings, x, scale: float = 1.0, clamp: Optional[float] = None,
                      block: int = 2048) -> np.ndarray:
    """FP64 ``exp(min(scale Q E^T, clamp)) x`` (NumPy)."""
    q = np.asarray(query, dtype=np.float64)
    e = np.asarray(embeddings, dtype=np.float64)
    xv = np.asarray(x, dtype=np.float64)
    out = np.zeros((q.shape[0],) + xv.shape[1:])
    for i0 in range(0, q.shape[0], block):
        g = scale * (q[i0:i0 + block] @ e.T)
        if clamp is not None:
            np.minimum(g, clamp, out=g)
        out[i0:i0 + block] = np.exp(g) @ xv
    return out


# ---------------------------------------------------------------------------
# Jacobi / adaptive preconditioners, power-iteration probes, warm-started refits
# (reference laker/solvers.py:505-546, laker/preconditioner.py:174-184,382-599,
# laker/streaming.py:75-379)
# ---------------------------------------------------------------------------


def jacobi_inverse(diagonal: torch.Tensor) -> torch.Tensor:
    """``1 / diagonal.clamp(min=eps(dtype))`` -- laker/solvers.py:521-532."""
    return 1.0 / diagonal.clamp(min=torch.finfo(diagonal.dtype).eps)


def jacobi_apply(inv_diag: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """``inv_diag * x`` -- laker/solvers.py:534-546 for (n,) input.  For (n, k)
    input the rows are scaled (the documented intent; the reference's
    broadcast runs over the last axis and raises unless k == n)."""
    return inv_diag * x if x.dim() == 1 else inv_diag.unsqueeze(1) * x


def adaptive_select(operator: Callable[[torch.Tensor], torch.Tensor], probes: torch.Tensor,
                    have_diagonal: bool) -> dict:
    """Conditioning estimate and choice of AdaptivePreconditioner.build
    (laker/preconditioner.py:469-574) on the given ``min(10, n)`` probes: one
    operator application to all probes, 5 normalised power steps from the first
    response, ``lambda_max = v^T A v``, ``lambda_min`` = smallest probe Rayleigh
    quotient, ``kappa = lambda_max / max(lambda_min, 1e-12)``; Jacobi below 1e3
    (when a diagonal is given), CCCP below 1e6, else CCCP with doubled probes."""
    op_probes = operator(probes)
    v = op_probes[:, 0].clone()
    for _ in range(5):
        v = operator(v.unsqueeze(1)).squeeze(1)
        v = v / torch.linalg.norm(v)
    lam_max = torch.dot(v, operator(v)).item()
    rayleigh = torch.sum(probes * op_probes, dim=0) / torch.sum(probes ** 2, dim=0)
    lam_min = rayleigh.min().item()
    cond = lam_max / max(lam_min, 1e-12)
    if cond < 1e3 and have_diagonal:
        choice = "jacobi"
    elif cond < 1e6:
        choice = "cccp"
    else:
        choice = "cccp_aggressive"
    return {"lambda_max": lam_max, "lambda_min": lam_min, "cond": cond, "choice": choice}


def power_iter_probes(operator: Callable[[torch.Tensor], torch.Tensor], probes: torch.Tensor,
                      steps: int = 3) -> torch.Tensor:
    """The "power_iter" probe strategy (laker/preconditioner.py:174-184): the
    first ``max(1, int(0.25 N_r))`` columns take ``steps`` rounds of
    ``Q(A block)`` (reduced Householder QR), the rest stay as drawn."""
    nr = probes.shape[1]
    n_power = max(1, int(nr * 0.25))
    block = probes[:, :n_power]
    for _ in range(steps):
        block = operator(block)
        block = torch.linalg.qr(block, mode="reduced")[0]
    return torch.cat([block, probes[:, n_power:]], dim=1)


def continuation_schedule(lambda_max: float, lambda_min: float, n_stages: int) -> list:
    """``lambda_k = lambda_max r^k``, ``r = (lambda_min/lambda_max)^(1/(n-1))``,
    last entry set to lambda_min -- laker/streaming.py:351-373."""
    if n_stages < 1:
        raise ValueError(f"n_stages must be positive, got {n_stages}")
    if lambda_max <= 0 or lambda_min <= 0:
        raise ValueError("lambda_max and lambda_min must be positive")
    ratio = (lambda_min / lambda_max) ** (1.0 / max(1, n_stages - 1))
    sched = [lambda_max * ratio ** k for k in range(n_stages)]
    sched[-1] = lambda_min
    return sched


def fit_path(embeddings: torch.Tensor, y: torch.Tensor, lambda_grid, probes: torch.Tensor,
             reuse_precond: bool = True, gamma: float = 0.1, epsilon: float = 1e-8, base_rho: float = 0.05,
             cccp_max_iter: int = 200, cccp_tol: float = 1e-6, pcg_tol: float = 1e-6,
             pcg_max_iter: int = 1000) -> dict:
    """StreamingUpdater.fit_path on fixed embeddings (laker/streaming.py:189-304):
    lambdas sorted descending, each PCG warm-started from the previous alpha,
    the CCCP factors of the first stage reused when ``reuse_precond``, final
    relative residuals ``||A alpha - y|| / ||y||`` in the solver dtype.  The
    same ``probes`` feed every CCCP build (the reference draws them from the
    same default-seeded generator each time)."""
    lambdas = sorted(lambda_grid, reverse=True)
    n = embeddings.shape[0]
    skip = skip_clamp_flag(embeddings)
    alphas, iters, rel = [], [], []
    x0 = None
    fac = None
    for lam in lambdas:
        def op(v, lam=lam):
            return attention_matvec(embeddings, lam, v, None, skip)

        if fac is None or not reuse_precond:
            fac = cccp_build(op, n, probes, gamma, epsilon, base_rho, cccp_max_iter, cccp_tol)
        res = pcg_solve(op, lambda v, f=fac: cccp_apply(v, f), y, x0=x0, tol=pcg_tol, max_iter=pcg_max_iter)
        alphas.append(res.x)
        iters.append(res.iterations)
        rel.append((torch.linalg.norm(op(res.x) - y) / torch.linalg.norm(y)).item())
        x0 = res.x.clone()
    return {"lambda_reg": lambdas, "alphas": alphas, "pcg_iters": iters, "final_rel_res": rel}


def partial_fit(embeddings: torch.Tensor, new_embeddings: torch.Tensor, alpha: torch.Tensor, y_old: torch.Tensor,
                y_new: torch.Tensor, lambda_reg: float, probes: torch.Tensor, forgetting_factor: float = 1.0,
                gamma: float = 0.1, epsilon: float = 1e-8, base_rho: float = 0.05, cccp_max_iter: int = 200,
                cccp_tol: float = 1e-6, pcg_tol: float = 1e-6, pcg_max_iter: int = 1000) -> dict:
    """StreamingUpdater.partial_fit (laker/streaming.py:75-187) on given
    embeddings: append, rebuild the operator and the CCCP factors (on
    ``probes``), and re-solve from ``[forgetting_factor * alpha, 0]``."""
    e = torch.cat([embeddings, new_embeddings], dim=0)
    n = e.shape[0]
    skip = skip_clamp_flag(e)

    def op(v):
        return attention_matvec(e, lambda_reg, v, None, skip)

    y = torch.cat([y_old, y_new])
    x0 = torch.cat([alpha * forgetting_factor, torch.zeros(new_embeddings.shape[0], dtype=alpha.dtype)])
    fac = cccp_build(op, n, probes, gamma, epsilon, base_rho, cccp_max_iter, cccp_tol)
    res = pcg_solve(op, lambda v: cccp_apply(v, fac), y, x0=x0, tol=pcg_tol, max_iter=pcg_max_iter)
    return {"embeddings": e, "alpha": res.x, "iterations": res.iterations, "y": y}
