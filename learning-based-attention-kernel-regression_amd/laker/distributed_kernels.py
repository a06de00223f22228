"""``DistributedAttentionKernelOperator`` on one GPU per process.

The reference wrapper (laker/distributed_kernels.py:42-339) shards the
embeddings over every visible device when there is more than one, and
otherwise collapses to a single inner ``AttentionKernelOperator``
(:116-134) to which ``matvec``, ``diagonal``, ``to_dense`` and
``kernel_eval`` delegate.  This build runs one process per GPU, so the
wrapper always takes that single-device form: the inner operator is the
HIP-backed ``AttentionKernelOperator`` on the process's compute device, and
``single_device`` is ``True``.  ``matvec`` is the inner operator's own bound
method, so the CCCP build and the PCG loop recognise it and run on device.
Sharding by output rows across processes is a design, not code here.
"""

from __future__ import annotations

import logging
from typing import Optional

import torch

from .backend import compute_device
from .kernels import AttentionKernelOperator

logger = logging.getLogger(__name__)


class DistributedAttentionKernelOperator:
    """Single-device form of the reference's distributed operator.

    Args:
        embeddings: ``(n, embedding_dim)`` embeddings.
        lambda_reg: Tikhonov regulariser.
        master_device: device on which results are returned (default: the
            embeddings' device), as in the reference.
        dtype: ``torch.float32`` or ``torch.float64`` (default: the
            embeddings' dtype).
        precision: matvec arithmetic of the inner operator (``"fp32"`` or
            ``"mixed"``).
    """

    def __init__(self, embeddings: torch.Tensor, lambda_reg: float = 1e-2,
                 master_device: Optional[torch.device] = None, dtype: Optional[torch.dtype] = None,
                 precision: str = "fp32") -> None:
        if dtype is None:
            dtype = embeddings.dtype
        if master_device is None:
            master_device = embeddings.device
        self.n = embeddings.shape[0]
        self.embedding_dim = embeddings.shape[1]
        self.lambda_reg = float(lambda_reg)
        self.master_device = torch.device(master_device)
        self.device = self.master_device
        self.dtype = dtype
        self.shape = (self.n, self.n)
        # recorded for inspection, as in the reference (laker/distributed_kernels.py:105-111)
        self.skip_clamp = True
        # one GPU per process: the wrapper's device list is this process's compute device
        self.devices = [compute_device()]
        visible = torch.cuda.device_count() if torch.cuda.is_available() else 0
        if visible > 1:
            logger.info("DistributedAttentionKernelOperator: %d GPUs visible; this build runs one GPU per "
                        "process and uses %s only", visible, self.devices[0])
        self.single_device = True
        self.local_op = AttentionKernelOperator(embeddings.to(dtype=dtype), lambda_reg=lambda_reg, chunk_size=None,
                                                device=self.master_device, dtype=dtype, precision=precision)
        # bound methods of the inner operator: the native solver paths see them as such
        self.matvec = self.local_op.matvec
        self.matvec_impl = self.local_op.matvec

    @property
    def embeddings(self) -> torch.Tensor:
        return self.local_op.embeddings

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return self.local_op.matvec(x)

    def diagonal(self) -> torch.Tensor:
        """``lambda + exp(||e_i||^2)`` (laker/distributed_kernels.py:262-278)."""
        return self.local_op.diagonal()

    def to_dense(self) -> torch.Tensor:
        """Dense ``lambda I + exp(E E^T)`` (laker/distributed_kernels.py:280-301)."""
        return self.local_op.to_dense()

    def kernel_eval(self, x: torch.Tensor, y: Optional[torch.Tensor] = None,
                    chunk_size: Optional[int] = None) -> torch.Tensor:
        """``exp(x y^T)`` (laker/distributed_kernels.py:303-339)."""
        return self.local_op.kernel_eval(x, y, chunk_size=chunk_size)

    def close(self) -> None:
        self.local_op.close()
