"""Scikit-learn style estimator (reference laker/models.py:47-884).

``LAKERRegressor`` keeps the reference constructor (validation, defaults),
hyperparameter delegation to the core (``HYPERPARAMS``), ``fit``,
``predict``, ``predict_variance``, ``score`` (negative RMSE), ``score_r2``,
``get_params`` / ``set_params``, ``save`` / ``load``,
``fit_learned_embeddings`` and the warm-started ``partial_fit`` /
``fit_path`` / ``fit_continuation``.  Added keyword arguments: ``precision``,
``probe_source``, ``probe_precision``, ``inner_tol``.
"""

from __future__ import annotations

import logging
from typing import Any, Optional, Union

import torch
import torch.nn as nn

from .backend import to_tensor
from .core import LAKERCore
from .streaming import StreamingUpdater

logger = logging.getLogger(__name__)

_OUT_OF_SCOPE = (
    "is not provided by the MI355X solver core (only the exact-kernel CCCP/PCG path is)"
)


class LAKERRegressor:
    HYPERPARAMS = (
        "embedding_dim", "lambda_reg", "gamma", "num_probes", "epsilon", "base_rho", "cccp_max_iter",
        "cccp_tol", "pcg_tol", "pcg_max_iter", "chunk_size", "kernel_approx", "num_landmarks",
        "num_features", "k_neighbors", "grid_size", "distributed", "twoscale_alpha", "landmark_method",
        "landmark_pilot_size", "spectral_knots", "preconditioner_strategy", "embedding_dtype", "device",
        "dtype", "verbose", "embedding_module", "precision", "probe_source", "probe_precision", "inner_tol",
    )

    def __init__(self, embedding_dim: int = 10, lambda_reg: float = 1e-2, gamma: float = 1e-1,
                 num_probes: Optional[int] = None, epsilon: float = 1e-8, base_rho: float = 0.05,
                 cccp_max_iter: int = 200, cccp_tol: float = 1e-6, pcg_tol: float = 1e-6,
                 pcg_max_iter: int = 1000, chunk_size: Optional[int] = None,
                 embedding_module: Optional[nn.Module] = None, kernel_approx: Optional[str] = None,
                 num_landmarks: Optional[int] = None, num_features: Optional[int] = None,
                 k_neighbors: Optional[int] = None, grid_size: Optional[int] = None,
                 distributed: bool = False, twoscale_alpha: float = 0.5, landmark_method: str = "greedy",
                 landmark_pilot_size: int = 1000, spectral_knots: int = 5, preconditioner: str = "cccp",
                 embedding_dtype: Optional[torch.dtype] = None,
                 device: Optional[Union[str, torch.device]] = None, dtype: Optional[torch.dtype] = None,
                 verbose: bool = True, residual_corrector: Optional[nn.Module] = None,
                 precision: str = "fp32", probe_source: str = "reference", probe_precision: str = "fp32",
                 inner_tol: float = 5e-2):
        # validation (laker/models.py:241-288)
        if embedding_dim <= 0:
            raise ValueError(f"embedding_dim must be positive, got {embedding_dim}")
        if lambda_reg <= 0:
            raise ValueError(f"lambda_reg must be positive, got {lambda_reg}")
        if gamma < 0:
            raise ValueError(f"gamma must be non-negative, got {gamma}")
        if epsilon <= 0:
            raise ValueError(f"epsilon must be positive, got {epsilon}")
        if not 0 <= base_rho <= 1:
            raise ValueError(f"base_rho must be in [0, 1], got {base_rho}")
        if cccp_max_iter <= 0:
            raise ValueError(f"cccp_max_iter must be positive, got {cccp_max_iter}")
        if cccp_tol <= 0:
            raise ValueError(f"cccp_tol must be positive, got {cccp_tol}")
        if pcg_tol <= 0:
            raise ValueError(f"pcg_tol must be positive, got {pcg_tol}")
        if pcg_max_iter <= 0:
            raise ValueError(f"pcg_max_iter must be positive, got {pcg_max_iter}")
        if kernel_approx not in (None, "nystrom", "rff", "knn", "ski", "twoscale", "spectral"):
            raise ValueError(
                "kernel_approx must be None, 'nystrom', 'rff', 'knn', 'ski', "
                f"'twoscale', or 'spectral', got {kernel_approx}")
        if k_neighbors is not None and k_neighbors <= 0:
            raise ValueError(f"k_neighbors must be positive, got {k_neighbors}")
        if grid_size is not None and grid_size < 2:
            raise ValueError(f"grid_size must be at least 2, got {grid_size}")
        if not 0.0 <= twoscale_alpha <= 1.0:
            raise ValueError(f"twoscale_alpha must be in [0, 1], got {twoscale_alpha}")
        if landmark_method not in ("greedy", "leverage"):
            raise ValueError(f"landmark_method must be 'greedy' or 'leverage', got {landmark_method}")
        if landmark_pilot_size <= 0:
            raise ValueError(f"landmark_pilot_size must be positive, got {landmark_pilot_size}")
        if preconditioner not in ("cccp", "adaptive"):
            raise ValueError(f"preconditioner must be 'cccp' or 'adaptive', got {preconditioner}")
        if spectral_knots <= 0:
            raise ValueError(f"spectral_knots must be positive, got {spectral_knots}")
        if precision not in ("fp32", "mixed"):
            raise ValueError(f"precision must be 'fp32' or 'mixed', got {precision}")
        object.__setattr__(self, "core", LAKERCore(
            embedding_dim=embedding_dim, lambda_reg=lambda_reg, gamma=gamma, num_probes=num_probes,
            epsilon=epsilon, base_rho=base_rho, cccp_max_iter=cccp_max_iter, cccp_tol=cccp_tol,
            pcg_tol=pcg_tol, pcg_max_iter=pcg_max_iter, chunk_size=chunk_size,
            embedding_module=embedding_module, kernel_approx=kernel_approx, num_landmarks=num_landmarks,
            num_features=num_features, k_neighbors=k_neighbors, grid_size=grid_size,
            distributed=distributed, twoscale_alpha=twoscale_alpha, landmark_method=landmark_method,
            landmark_pilot_size=landmark_pilot_size, spectral_knots=spectral_knots,
            preconditioner_strategy=preconditioner, embedding_dtype=embedding_dtype, device=device,
            dtype=dtype, verbose=verbose, precision=precision, probe_source=probe_source,
            probe_precision=probe_precision, inner_tol=inner_tol))
        self.embeddings: Optional[torch.Tensor] = None
        self.alpha: Optional[torch.Tensor] = None
        self.alpha_fp64: Optional[torch.Tensor] = None
        self.kernel_operator = None
        self.preconditioner = None
        self.embedding_model: Optional[nn.Module] = None
        self.residual_corrector = residual_corrector
        self.x_train: Optional[torch.Tensor] = None
        self.y_train: Optional[torch.Tensor] = None
        self.partial_fit_count: int = 0
        self.pcg_iterations_: int = 0
        self.streaming = StreamingUpdater(self.core)

    # hyperparameter delegation (laker/models.py:343-387)
    def __getattr__(self, name: str) -> Any:
        if name in LAKERRegressor.HYPERPARAMS and "core" in self.__dict__:
            return getattr(self.__dict__["core"], name)
        raise AttributeError(f"'{type(self).__name__}' has no attribute '{name}'")

    def __setattr__(self, name: str, value: Any) -> None:
        if name in self.HYPERPARAMS and "core" in self.__dict__:
            if name == "device" and isinstance(value, str):
                value = torch.device(value)
            if name in ("dtype", "embedding_dtype") and isinstance(value, str):
                value = torch.float32 if "float32" in value else torch.float64
            setattr(self.core, name, value)
        else:
            super().__setattr__(name, value)

    # ------------------------------------------------------------------
    def fit(self, x, y, x0: Optional[torch.Tensor] = None, seed: Optional[int] = None) -> "LAKERRegressor":
        """Embeddings -> operator -> CCCP -> PCG (laker/models.py:392-455)."""
        x = to_tensor(x, device=self.device, dtype=self.dtype)
        y = to_tensor(y, device=self.device, dtype=self.dtype).squeeze()
        if x.dim() != 2:
            raise ValueError(f"x must be 2-D, got shape {x.shape}")
        if y.dim() != 1:
            raise ValueError(f"y must be 1-D, got shape {y.shape}")
        self.x_train = x
        self.y_train = y
        self.embeddings, self.embedding_model = self.core.compute_embeddings(x)
        op = self.core.build_kernel_operator(self.embeddings)
        self.kernel_operator = op
        self.preconditioner = self.core.build_preconditioner(op.matvec, self.embeddings.shape[0], seed=seed,
                                                             diagonal=op.diagonal())
        self.alpha, self.pcg_iterations_ = self.core.solve_pcg(op, self.preconditioner, y, x0=x0)
        solver = getattr(self.core, "last_solver", None)
        self.alpha_fp64 = solver.solution_fp64 if solver is not None else None
        return self

    def _check_fitted(self) -> None:
        if self.alpha is None or self.embeddings is None:
            raise RuntimeError("Model has not been fitted. Call fit() first.")

    def predict(self, x) -> torch.Tensor:
        """``exp(e(x) E^T) alpha`` (laker/models.py:457-495)."""
        self._check_fitted()
        x = to_tensor(x, device=self.device, dtype=self.dtype)
        if x.dim() != 2:
            raise ValueError(f"x must be 2-D, got shape {x.shape}")
        # the mixed solve's FP64 solution, where there is one, keeps its accuracy in the predictions
        alpha = self.alpha_fp64 if self.alpha_fp64 is not None else self.alpha
        return self.core.predict(x, self.embedding_model, self.embeddings, self.kernel_operator, alpha,
                                 self.residual_corrector)

    def predict_variance(self, x) -> torch.Tensor:
        """Posterior variance, clamped at 0 (laker/models.py:497-537)."""
        self._check_fitted()
        x = to_tensor(x, device=self.device, dtype=self.dtype)
        if x.dim() != 2:
            raise ValueError(f"x must be 2-D, got shape {x.shape}")
        if self.preconditioner is None:  # loaded from a file without stored factors
            op = self.kernel_operator
            self.preconditioner = self.core.build_preconditioner(op.matvec, self.embeddings.shape[0],
                                                                 diagonal=op.diagonal())
        return self.core.predict_variance(x, self.embedding_model, self.embeddings, self.kernel_operator,
                                          self.preconditioner, self.alpha, self.lambda_reg)

    def score(self, x, y) -> float:
        """Negative RMSE (laker/models.py:715-736)."""
        y_pred = self.predict(x)
        y_true = to_tensor(y, device=y_pred.device, dtype=self.dtype).squeeze()
        return -torch.sqrt(torch.mean((y_pred - y_true) ** 2)).item()

    def score_r2(self, x, y) -> float:
        """Coefficient of determination (laker/models.py:738-770)."""
        y_pred = self.predict(x)
        y_true = to_tensor(y, device=y_pred.device, dtype=self.dtype).squeeze()
        ss_res = torch.sum((y_true - y_pred) ** 2).item()
        ss_tot = torch.sum((y_true - torch.mean(y_true)) ** 2).item()
        if ss_tot == 0:
            return 1.0 if ss_res == 0 else 0.0
        return 1.0 - ss_res / ss_tot

    # ------------------------------------------------------------------
    def fit_learned_embeddings(self, x, y, lr: float = 1e-3, epochs: int = 50, rebuild_freq: int = 10,
                               patience: int = 5) -> "LAKERRegressor":
        """End-to-end embedding training through the hand-written backward (laker/models.py:638-650)."""
        from .training import EmbeddingTrainer

        x = to_tensor(x, device=self.device, dtype=self.dtype)
        y = to_tensor(y, device=self.device, dtype=self.dtype).squeeze()
        return EmbeddingTrainer(self.core).fit_learned_embeddings(self, x, y, lr, epochs, rebuild_freq, patience)

    def fit_validation_embeddings(self, x_train, y_train, x_val, y_val, lr: float = 1e-3, epochs: int = 50):
        """Train the embedding on validation MSE through the implicit adjoint (bilevel)."""
        from .training import EmbeddingTrainer

        return EmbeddingTrainer(self.core).fit_validation_embeddings(
            self, to_tensor(x_train, device=self.device, dtype=self.dtype),
            to_tensor(y_train, device=self.device, dtype=self.dtype).squeeze(),
            to_tensor(x_val, device=self.device, dtype=self.dtype),
            to_tensor(y_val, device=self.device, dtype=self.dtype).squeeze(), lr, epochs)

    def _unavailable(self, name):
        raise NotImplementedError(f"{name} {_OUT_OF_SCOPE}")

    def fit_with_search(self, *a, **k):
        self._unavailable("fit_with_search")

    def fit_with_bo(self, *a, **k):
        self._unavailable("fit_with_bo")

    # streaming (laker/models.py:595-636)
    def partial_fit(self, x_new, y_new, forgetting_factor: float = 1.0,
                    rebuild_threshold: int = 100) -> "LAKERRegressor":
        """Append observations and re-solve warm-started (laker/streaming.py:75-187)."""
        x_new = to_tensor(x_new, device=self.device, dtype=self.dtype)
        y_new = to_tensor(y_new, device=self.device, dtype=self.dtype).squeeze()
        return self.streaming.partial_fit(self, x_new, y_new, forgetting_factor, rebuild_threshold)

    def fit_path(self, x, y, lambda_reg_grid, reuse_precond: bool = True) -> dict:
        """Warm-started solves over a lambda grid (laker/streaming.py:189-304)."""
        x = to_tensor(x, device=self.device, dtype=self.dtype)
        y = to_tensor(y, device=self.device, dtype=self.dtype).squeeze()
        return self.streaming.fit_path(self, x, y, lambda_reg_grid, reuse_precond)

    def fit_continuation(self, x, y, lambda_max: Optional[float] = None, lambda_min: Optional[float] = None,
                         n_stages: int = 5, reuse_precond: bool = True) -> "LAKERRegressor":
        """Geometric lambda continuation (laker/streaming.py:306-379)."""
        x = to_tensor(x, device=self.device, dtype=self.dtype)
        y = to_tensor(y, device=self.device, dtype=self.dtype).squeeze()
        return self.streaming.fit_continuation(self, x, y, lambda_max, lambda_min, n_stages, reuse_precond)

    def fit_residual_corrector(self, *a, **k):
        self._unavailable("fit_residual_corrector")

    def fit_bilevel(self, *a, **k):
        self._unavailable("fit_bilevel")

    def fit_uncertainty_aware(self, *a, **k):
        self._unavailable("fit_uncertainty_aware")

    def condition_number(self) -> float:
        self._unavailable("condition_number")

    # ------------------------------------------------------------------
    def get_params(self, deep: bool = True) -> dict:
        """Constructor parameters (laker/models.py:778-826)."""
        return {
            "embedding_dim": self.embedding_dim, "lambda_reg": self.lambda_reg, "gamma": self.gamma,
            "num_probes": self.num_probes, "epsilon": self.epsilon, "base_rho": self.base_rho,
            "cccp_max_iter": self.cccp_max_iter, "cccp_tol": self.cccp_tol, "pcg_tol": self.pcg_tol,
            "pcg_max_iter": self.pcg_max_iter, "chunk_size": self.chunk_size,
            "embedding_module": self.embedding_module, "kernel_approx": self.kernel_approx,
            "num_landmarks": self.num_landmarks, "num_features": self.num_features,
            "k_neighbors": self.k_neighbors, "grid_size": self.grid_size, "distributed": self.distributed,
            "twoscale_alpha": self.twoscale_alpha, "landmark_method": self.landmark_method,
            "landmark_pilot_size": self.landmark_pilot_size, "preconditioner": self.preconditioner_strategy,
            "embedding_dtype": (str(self.embedding_dtype).replace("torch.", "") if self.embedding_dtype else None),
            "device": str(self.device), "dtype": str(self.dtype).replace("torch.", ""),
            "verbose": self.verbose, "residual_corrector": self.residual_corrector,
            "precision": self.precision, "probe_source": self.probe_source,
            "probe_precision": self.probe_precision, "inner_tol": self.inner_tol,
        }

    def set_params(self, **params) -> "LAKERRegressor":
        """Update hyperparameters by name (laker/models.py:828-846)."""
        for key, value in params.items():
            if key == "preconditioner":
                key = "preconditioner_strategy"
            if key not in self.HYPERPARAMS and key != "residual_corrector":
                raise ValueError(f"Invalid parameter {key!r} for LAKERRegressor")
            setattr(self, key, value)
        return self

    # ------------------------------------------------------------------
    def save(self, path: str) -> None:
        from .persistence import save_model

        save_model(self, path)

    @classmethod
    def load(cls, path: str) -> "LAKERRegressor":
        from .persistence import load_model

        return load_model(path)
