"""Cholesky-QR2 of tall blocks on the GPU (laker_orthonormalize): the probe QR
of the CCCP build and the power-iteration orthonormalisation.  Widths below,
at and across the 64-column blocks of the blocked Cholesky (lower-triangle
Gram, triangular V L^-T, in-place panels), checked against NumPy's
Householder QR in FP64: the orthonormal factors agree up to column signs and
span the input."""

import numpy as np
import pytest
import torch

from laker import _native as N

pytestmark = pytest.mark.gpu


def orthonormalize(x):
    y = x.clone().contiguous()
    N.check(N.lib.laker_orthonormalize(N.ptr(y), y.shape[0], y.shape[1], N.dtype_code(y.dtype), N.stream_ptr()))
    torch.cuda.synchronize()
    return y


@pytest.mark.parametrize("n,k", [(70, 1), (500, 63), (500, 64), (3000, 65), (3000, 200), (20000, 333)])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_orthonormalize_matches_householder(cuda, n, k, dtype):
    g = torch.Generator().manual_seed(n + k)
    # columns of very different scale and some correlation: the QR has real work to do
    x = torch.randn(n, k, dtype=torch.float64, generator=g)
    x = x @ (torch.eye(k, dtype=torch.float64) + 0.3 * torch.randn(k, k, dtype=torch.float64, generator=g))
    x = x * torch.logspace(0, 3, k, dtype=torch.float64)
    xd = x.to(dtype)
    print(f"orthonormalize n={n} k={k} {dtype}", flush=True)
    q = orthonormalize(xd.to(cuda)).double().cpu().numpy()
    qh, _ = np.linalg.qr(xd.double().numpy())
    # FP64: rounding of order kappa(x) u (kappa ~ 1e4 here); FP32: the output rounds an FP64 basis
    tol = 1e-10 if dtype == torch.float64 else 2e-5
    # orthonormal columns (Cholesky-QR2: ~ n u in FP64)
    assert np.max(np.abs(q.T @ q - np.eye(k))) < (1e-11 if dtype == torch.float64 else 5e-6 * np.sqrt(n))
    # the same basis as Householder QR, column by column up to sign (both triangularise the input)
    s = np.sign(np.sum(q * qh, axis=0))
    assert np.all(s != 0)
    assert np.max(np.abs(q - qh * s)) < tol * np.sqrt(k), np.max(np.abs(q - qh * s))
