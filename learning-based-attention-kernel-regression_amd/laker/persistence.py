"""Model save / load in the reference's ``torch.save`` dict format (laker/persistence.py:74-328).

The dict carries the same hyperparameter keys, ``embeddings``, ``alpha`` and
the embedding-module state dict, so files written by either implementation
hold the same fields.  It is extended compatibly with the CCCP factors
(``preconditioner_*`` keys: Q, the QR factor R, C, S = M^{-1/2} - a^{-1/2} I,
isotropic_coef and rho), which the reference's format lacks: a model saved
here reloads with its preconditioner handle rebuilt from those factors, with
no probe matvec or CCCP iteration, so a served system is ready to solve at
once.  Files without those keys (the reference's) load as before and build
the preconditioner on demand.  Loading uses ``torch.load(weights_only=True)``
(no code from the file runs) and rebuilds the exact operator.
"""

from __future__ import annotations

import torch

_HP = ("embedding_dim", "lambda_reg", "gamma", "num_probes", "epsilon", "base_rho", "cccp_max_iter",
       "cccp_tol", "pcg_tol", "pcg_max_iter", "chunk_size", "kernel_approx", "num_landmarks",
       "num_features", "k_neighbors", "grid_size", "distributed")


def save_model(reg, path: str) -> None:
    if reg.alpha is None:
        raise RuntimeError("Model has not been fitted. Call fit() before saving.")
    state = {k: getattr(reg, k) for k in _HP}
    state.update({
        "device": str(reg.device), "dtype": str(reg.dtype),
        "embedding_dtype": str(reg.embedding_dtype) if reg.embedding_dtype else None,
        "verbose": reg.verbose, "embeddings": reg.embeddings.detach().cpu(), "alpha": reg.alpha.detach().cpu(),
        "precision": reg.precision,
    })
    if reg.alpha_fp64 is not None:
        state["alpha_fp64"] = reg.alpha_fp64.detach().cpu()
    pre = _cccp_of(reg.preconditioner)
    if pre is not None:
        state.update({
            "preconditioner_q_basis": pre.q_basis.detach().cpu(),
            "preconditioner_qr_r": pre.qr_r.detach().cpu(),
            "preconditioner_q_basis_matrix": pre.q_basis_matrix.detach().cpu(),
            "preconditioner_s_matrix": pre.s_matrix.detach().cpu(),
            "preconditioner_isotropic_coef": float(pre.isotropic_coef),
            "preconditioner_rho": float(pre.rho if pre.rho is not None else 0.0),
        })
    if reg.embedding_model is not None:
        state["embedding_model_state"] = {k: v.detach().cpu() for k, v in reg.embedding_model.state_dict().items()}
        state["embedding_model_class"] = reg.embedding_model.__class__.__name__
        state["embedding_model_module"] = reg.embedding_model.__class__.__module__
        if hasattr(reg.embedding_model, "input_dim"):
            state["input_dim"] = reg.embedding_model.input_dim
    torch.save(state, path)


def load_model(path: str):
    from .embeddings import PositionEmbedding
    from .models import LAKERRegressor

    state = torch.load(path, weights_only=True, map_location="cpu")
    dtype = torch.float32 if "float32" in state["dtype"] else torch.float64
    ed = state.get("embedding_dtype")
    emb_dtype = None if not ed else (torch.float32 if "float32" in ed else torch.float64)
    kwargs = {k: state.get(k) for k in _HP if k in state}
    reg = LAKERRegressor(dtype=dtype, embedding_dtype=emb_dtype, verbose=state.get("verbose", False),
                         precision=state.get("precision", "fp32"), **kwargs)
    reg.embeddings = state["embeddings"].to(reg.core.device if reg.core.device.type != "cpu" else "cuda")
    reg.alpha = state["alpha"].to(reg.embeddings.device)
    if "alpha_fp64" in state:
        reg.alpha_fp64 = state["alpha_fp64"].to(reg.embeddings.device)
    if "embedding_model_state" in state:
        if state.get("embedding_model_class") == "PositionEmbedding" and "input_dim" in state:
            model = PositionEmbedding(input_dim=state["input_dim"], embedding_dim=state["embedding_dim"],
                                      device=reg.embeddings.device, dtype=emb_dtype or dtype)
            model.load_state_dict(state["embedding_model_state"])
            reg.embedding_model = model
    reg.kernel_operator = reg.core.build_kernel_operator(reg.embeddings)
    if "preconditioner_q_basis" in state:
        from .preconditioner import CCCPPreconditioner

        reg.preconditioner = CCCPPreconditioner.from_factors(
            state["preconditioner_q_basis"], state["preconditioner_qr_r"], state["preconditioner_q_basis_matrix"],
            state["preconditioner_s_matrix"], state["preconditioner_isotropic_coef"],
            state.get("preconditioner_rho", 0.0), gamma=reg.gamma, epsilon=reg.epsilon, base_rho=reg.base_rho,
            max_iter=reg.cccp_max_iter, tol=reg.cccp_tol)
    return reg


def _cccp_of(pre):
    """The built CCCP preconditioner behind ``pre`` (itself or an adaptive wrapper's), else None."""
    from .preconditioner import AdaptivePreconditioner, CCCPPreconditioner

    if isinstance(pre, AdaptivePreconditioner):
        pre = pre.inner
    if isinstance(pre, CCCPPreconditioner) and pre.built:
        return pre
    return None
