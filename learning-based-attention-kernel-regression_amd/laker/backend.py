"""Device / dtype defaults (reference laker/backend.py:41-244).

The solver always computes on the ROCm GPU.  ``device`` settings keep the
reference meaning of "where user-visible tensors live": CPU inputs are moved to
the GPU for the computation and results are returned on the requested device.
``LAKER_DEVICE`` / ``LAKER_DTYPE`` are honoured as in the reference
(laker/backend.py:49-87).
"""

from __future__ import annotations

import logging
import os
from typing import Optional, Union

import torch

logger = logging.getLogger(__name__)

DEFAULT_DEVICE: torch.device = torch.device("cuda")
DEFAULT_DTYPE: torch.dtype = torch.float32
_ENV_INIT_DONE = False


def _init_from_env() -> None:
    global _ENV_INIT_DONE
    if _ENV_INIT_DONE:
        return
    _ENV_INIT_DONE = True
    env_device = os.environ.get("LAKER_DEVICE")
    if env_device:
        set_default_device(env_device)
    env_dtype = os.environ.get("LAKER_DTYPE")
    if env_dtype == "float32":
        set_default_dtype(torch.float32)
    elif env_dtype == "float64":
        set_default_dtype(torch.float64)


def get_default_device() -> torch.device:
    _init_from_env()
    return DEFAULT_DEVICE


def set_default_device(device: Optional[Union[str, torch.device]] = None) -> torch.device:
    """Set the device on which results are returned (computation is on the GPU)."""
    global DEFAULT_DEVICE
    device = torch.device("cuda") if device is None else torch.device(device)
    DEFAULT_DEVICE = device
    logger.info("Default device set to %s", device)
    return device


def get_default_dtype() -> torch.dtype:
    _init_from_env()
    return DEFAULT_DTYPE


def set_default_dtype(dtype: torch.dtype) -> None:
    global DEFAULT_DTYPE
    DEFAULT_DTYPE = dtype
    logger.info("Default dtype set to %s", dtype)


def compute_device() -> torch.device:
    """The GPU every HIP kernel runs on."""
    if not torch.cuda.is_available():
        raise RuntimeError("LAKER's HIP backend requires a ROCm GPU; none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def to_compute(t: torch.Tensor, dtype: Optional[torch.dtype] = None) -> torch.Tensor:
    """Contiguous copy (or view) of ``t`` on the compute GPU in ``dtype``."""
    dev = compute_device()
    return t.to(device=dev, dtype=dtype if dtype is not None else t.dtype).contiguous()


def to_tensor(data, device: Optional[torch.device] = None, dtype: Optional[torch.dtype] = None) -> torch.Tensor:
    """Array-like to tensor on ``device``/``dtype`` (reference laker/backend.py:211-244)."""
    if device is None:
        device = get_default_device()
    if dtype is None:
        dtype = get_default_dtype()
    if isinstance(data, torch.Tensor):
        return data.to(device=device, dtype=dtype)
    return torch.as_tensor(data, device=device, dtype=dtype)
