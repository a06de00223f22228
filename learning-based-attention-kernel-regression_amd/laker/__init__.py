"""LAKER on MI355X: learning-based attention-kernel regression with a HIP solver core.

Drop-in for the reference package's exact-kernel path: ``LAKERRegressor``,
``AttentionKernelOperator``, ``CCCPPreconditioner``,
``PreconditionedConjugateGradient``, ``PositionEmbedding`` and the data
helpers keep the reference names, signatures and error behaviour
(laker/__init__.py:82-122).  Every matvec, preconditioner build/apply and
PCG loop runs in ``liblaker_hip.so`` (hand-written gfx950 kernels); importing
the package fails if the library is missing.
"""

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

from . import _native  # noqa: F401  (raises ImportError if liblaker_hip.so is missing)
from .backend import get_default_device, set_default_device
from .data import generate_grid, generate_radio_field, pathloss_problem
from .embeddings import PositionEmbedding
from .kernels import AttentionKernelOperator, KernelOperator
from .models import LAKERRegressor
from .preconditioner import AdaptivePreconditioner, CCCPPreconditioner
from .solvers import JacobiPreconditioner, PreconditionedConjugateGradient
from .streaming import StreamingUpdater

__version__ = "0.4.0"

__all__ = [
    "get_default_device",
    "set_default_device",
    "generate_grid",
    "generate_radio_field",
    "pathloss_problem",
    "PositionEmbedding",
    "AttentionKernelOperator",
    "KernelOperator",
    "CCCPPreconditioner",
    "AdaptivePreconditioner",
    "JacobiPreconditioner",
    "PreconditionedConjugateGradient",
    "StreamingUpdater",
    "LAKERRegressor",
]
