"""Host-side numerical helpers of the CCCP build (reference laker/utils.py:68-134)."""

from __future__ import annotations

import torch


def adaptive_shrinkage_rho(num_probes: int, problem_size: int, gamma: float, base_rho: float = 0.05) -> float:
    """Shrinkage rho from the undersampling ratio (laker/utils.py:68-106).

    ``base_rho`` when ``N_r >= n``; else
    ``min(0.5, base_rho + (1 - base_rho)(1 - N_r/n) min(1, 10 gamma))``.
    The device build computes the same value in C++.
    """
    if num_probes >= problem_size:
        return base_rho
    ratio = num_probes / problem_size
    rho = base_rho + (1.0 - base_rho) * (1.0 - ratio) * min(1.0, gamma * 10.0)
    return float(min(rho, 0.5))


def eigh_stable(mat: torch.Tensor, eps: float = 1e-10):
    """``eigh`` with eigenvalues clamped to ``>= eps`` (laker/utils.py:109-134).

    Used only to expose ``q_eigenvalues`` / ``q_eigenvectors`` for
    introspection; the device build never needs an eigendecomposition.
    """
    w, v = torch.linalg.eigh(mat)
    return w.clamp(min=eps), v


def default_num_probes(n: int, num_probes=None) -> int:
    """``min(num_probes or max(200, int(2 sqrt n)), n)`` (laker/preconditioner.py:161-163)."""
    nr = num_probes if num_probes is not None else max(200, int(2 * n**0.5))
    return min(nr, n)
