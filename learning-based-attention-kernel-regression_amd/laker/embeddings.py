"""Random-Fourier-feature position embedding (reference laker/embeddings.py:42-258).

The parameters are drawn from a dedicated CPU ``torch.Generator`` in the
reference's order -- ``freq = randn(in, F) / sigma``, ``phase = rand(F) 2 pi``,
then for each Linear ``weight = rand * 2b - b`` and ``bias = rand * 2b - b``
with ``b = sqrt(1 / fan_in)`` -- and then moved to the requested device, so the
module equals the reference module built on a CPU device for any target
device.  This module is the input producer of the solver (it runs once per
fit); it is a plain ``nn.Module`` so it stays trainable.
"""

from __future__ import annotations

import math
from typing import Optional

import torch
import torch.nn as nn

from .backend import get_default_device, get_default_dtype


class PositionEmbedding(nn.Module):
    def __init__(self, input_dim: int, embedding_dim: int, num_fourier: Optional[int] = None,
                 sigma: float = 10.0, seed: int = 42, device: Optional[torch.device] = None,
                 dtype: Optional[torch.dtype] = None):
        super().__init__()
        self.input_dim = input_dim
        self.embedding_dim = embedding_dim
        self.sigma = sigma
        if num_fourier is None:
            num_fourier = embedding_dim * 2
        self.num_fourier = num_fourier
        if device is None:
            device = get_default_device()
        if dtype is None:
            dtype = get_default_dtype()
        gen = torch.Generator(device="cpu").manual_seed(seed)
        freq = torch.randn(input_dim, num_fourier, generator=gen, dtype=dtype) / sigma
        phase = torch.rand(num_fourier, generator=gen, dtype=dtype) * 2.0 * math.pi
        self.register_buffer("freq", freq)
        self.register_buffer("phase", phase)
        hidden = max(embedding_dim, num_fourier // 2)
        self.mlp = nn.Sequential(
            nn.Linear(num_fourier, hidden, dtype=dtype),
            nn.Tanh(),
            nn.Linear(hidden, embedding_dim, dtype=dtype),
        )
        with torch.no_grad():
            for layer in self.mlp:
                if isinstance(layer, nn.Linear):
                    bound = math.sqrt(1.0 / layer.weight.shape[1])
                    layer.weight.copy_(torch.rand(layer.weight.shape, generator=gen, dtype=dtype) * (2 * bound) - bound)
                    layer.bias.copy_(torch.rand(layer.bias.shape, generator=gen, dtype=dtype) * (2 * bound) - bound)
        self.to(device=device, dtype=dtype)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """``mlp(cos(2 pi x @ freq + phase))`` (laker/embeddings.py:220-242)."""
        if x.dim() == 1:
            x = x.unsqueeze(0)
        x = x.to(device=self.freq.device, dtype=self.freq.dtype)
        features = torch.cos(2.0 * math.pi * (x @ self.freq) + self.phase)
        return self.mlp(features)

    def extra_repr(self) -> str:
        return (f"input_dim={self.input_dim}, embedding_dim={self.embedding_dim}, "
                f"num_fourier={self.num_fourier}, sigma={self.sigma}")
