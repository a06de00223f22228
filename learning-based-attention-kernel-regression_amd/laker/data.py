"""Synthetic radio-map inputs (reference laker/data.py), host side.

``pathloss_problem`` is the seeded generator of the benchmark workloads: n
locations uniform in [0,100]^2, three log-distance sources with per-source
exponent in [2,4] (the formula of laker/data.py:161-169 with d0 = 1 and 0 dBm
powers), N(0,1) noise, standardised targets.  Draw order on one CPU
generator: locations, source positions, exponents, noise.
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch


def pathloss_problem(n: int, seed: int = 0, num_sources: int = 3, k: int = 1,
                     dtype: torch.dtype = torch.float32) -> Tuple[torch.Tensor, torch.Tensor]:
    """``(x, y)`` with ``x`` (n, 2) and ``y`` (n,) -- or (n, k) with k independent noise draws."""
    gen = torch.Generator(device="cpu").manual_seed(int(seed))
    x = torch.rand(n, 2, generator=gen, dtype=dtype) * 100.0
    tx = torch.rand(num_sources, 2, generator=gen, dtype=dtype) * 100.0
    eta = torch.rand(num_sources, generator=gen, dtype=dtype) * 2.0 + 2.0
    clean = torch.zeros(n, dtype=dtype)
    for j in range(num_sources):
        d = torch.norm(x - tx[j], dim=1).clamp(min=1.0)
        clean = clean - 10.0 * eta[j] * torch.log10(d)
    cols = []
    for _ in range(k):
        y = clean + torch.randn(n, generator=gen, dtype=dtype)
        cols.append((y - y.mean()) / y.std())
    y = cols[0] if k == 1 else torch.stack(cols, dim=1)
    return x, y


def generate_radio_field(locations: torch.Tensor, transmitters: torch.Tensor, powers: torch.Tensor,
                         path_loss_exponent: float = 2.0, reference_distance: float = 1.0,
                         shadow_sigma: float = 1.5, seed: Optional[int] = None):
    """``(rss_clean, rss_noisy)``: sum_j P_j - 10 eta log10(max(d_j, d0)/d0), plus shadowing.

    Restates laker/data.py:84-181 / :184-219 (shape checks included).
    """
    if locations.dim() != 2 or transmitters.dim() != 2 or powers.dim() != 1:
        raise ValueError("locations/transmitters must be 2-D and powers 1-D")
    if transmitters.shape[0] != powers.shape[0]:
        raise ValueError("number of transmitters and powers differ")
    if locations.shape[1] != transmitters.shape[1]:
        raise ValueError("spatial dimensions of locations and transmitters differ")
    gen = None
    if seed is not None:
        gen = torch.Generator(device=locations.device).manual_seed(seed)
    n = locations.shape[0]
    clean = torch.zeros(n, device=locations.device, dtype=locations.dtype)
    for tx, p in zip(transmitters, powers):
        d = torch.norm(locations - tx, dim=1).clamp(min=reference_distance)
        clean += p - 10.0 * path_loss_exponent * torch.log10(d / reference_distance)
    noise = torch.randn(n, device=locations.device, dtype=locations.dtype, generator=gen)
    return clean, clean + shadow_sigma * noise


def generate_grid(bounds, grid_size: int, device=None, dtype=None) -> torch.Tensor:
    """Row-major ``grid_size^2 x 2`` lattice over ``bounds`` (laker/data.py:222-257)."""
    x_min, x_max, y_min, y_max = bounds
    x = torch.linspace(x_min, x_max, grid_size, device=device, dtype=dtype)
    y = torch.linspace(y_min, y_max, grid_size, device=device, dtype=dtype)
    gx, gy = torch.meshgrid(x, y, indexing="ij")
    return torch.stack([gx.reshape(-1), gy.reshape(-1)], dim=1)
