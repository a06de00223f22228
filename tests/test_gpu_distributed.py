"""``DistributedAttentionKernelOperator`` in its single-device form, through the
C ABI (reference tests/test_distributed.py:9-88), and ``distributed=True`` fits."""

import numpy as np
import pytest
import torch

import laker
import laker_oracle as O
from laker.distributed_kernels import DistributedAttentionKernelOperator

pytestmark = pytest.mark.gpu


def _inputs(n, seed=42):
    g = torch.Generator().manual_seed(seed)
    e = torch.randn(n, 8, generator=g, dtype=torch.float64)
    x = torch.randn(n, generator=g, dtype=torch.float64)
    return e, x


def test_single_device_fallback(cuda):
    e, x = _inputs(50)
    dist_op = DistributedAttentionKernelOperator(e, lambda_reg=1e-2, dtype=torch.float64)
    assert dist_op.single_device
    assert dist_op.shape == (50, 50) and dist_op.n == 50 and dist_op.embedding_dim == 8
    y = dist_op.matvec(x)
    d = dist_op.diagonal()
    assert y.shape == (50,) and d.shape == (50,)
    # results come back on the master device (the embeddings' device here: CPU)
    assert y.device == e.device
    # the inner operator is the HIP one: its matvec is what the solvers receive
    assert isinstance(dist_op.local_op, laker.AttentionKernelOperator)
    assert dist_op.matvec.__self__ is dist_op.local_op


def test_matvec_matches_single_operator_and_fp64(cuda):
    e, x = _inputs(50)
    dist_op = DistributedAttentionKernelOperator(e, lambda_reg=1e-2, dtype=torch.float64)
    single = laker.AttentionKernelOperator(e, lambda_reg=1e-2, dtype=torch.float64)
    y_dist = dist_op.matvec(x)
    assert torch.allclose(y_dist, single.matvec(x), atol=1e-6)
    # reference tests compare at 1e-6; the FP64 product holds 1e-12 (the clamp is active: max ||e||^2 > 80?)
    clamp = None if O.skip_clamp_flag(e) else 700.0
    ref = O.matvec_fp64(e.numpy(), 1e-2, x.numpy(), clamp=clamp)
    np.testing.assert_allclose(y_dist.numpy(), ref, rtol=1e-12, atol=1e-12)
    x2 = torch.randn(50, 3, dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    np.testing.assert_allclose(dist_op.matvec(x2).numpy(), O.matvec_fp64(e.numpy(), 1e-2, x2.numpy(), clamp=clamp),
                               rtol=1e-12, atol=1e-12)


def test_diagonal_matches(cuda):
    e, _ = _inputs(50)
    dist_op = DistributedAttentionKernelOperator(e, lambda_reg=1e-2, dtype=torch.float64)
    single = laker.AttentionKernelOperator(e, lambda_reg=1e-2, dtype=torch.float64)
    assert torch.allclose(dist_op.diagonal(), single.diagonal(), atol=1e-6)
    np.testing.assert_allclose(dist_op.diagonal().numpy(), O.diagonal(e, 1e-2).numpy(), rtol=1e-14)


def test_to_dense_matches(cuda):
    e, _ = _inputs(30)
    dist_op = DistributedAttentionKernelOperator(e, lambda_reg=1e-2, dtype=torch.float64)
    single = laker.AttentionKernelOperator(e, lambda_reg=1e-2, dtype=torch.float64)
    assert torch.allclose(dist_op.to_dense(), single.to_dense(), atol=1e-6)
    np.testing.assert_allclose(dist_op.to_dense().numpy(), O.to_dense(e, 1e-2).numpy(), rtol=1e-13)


def test_kernel_eval_matches(cuda):
    e, _ = _inputs(50)
    xq = torch.randn(10, 8, dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    dist_op = DistributedAttentionKernelOperator(e, lambda_reg=1e-2, dtype=torch.float64)
    single = laker.AttentionKernelOperator(e, lambda_reg=1e-2, dtype=torch.float64)
    k = dist_op.kernel_eval(xq)
    assert k.shape == (10, 50)
    assert torch.allclose(k, single.kernel_eval(xq), atol=1e-6)


def test_on_device_master(cuda):
    """reference test_distributed_multi_device_matvec: vectors on the GPU stay there."""
    g = torch.Generator().manual_seed(42)
    e = torch.randn(100, 8, generator=g, dtype=torch.float64).to(cuda)
    x = torch.randn(100, generator=g, dtype=torch.float64).to(cuda)
    dist_op = DistributedAttentionKernelOperator(e, lambda_reg=1e-2, master_device=torch.device("cuda"))
    y = dist_op.matvec(x)
    assert y.shape == (100,) and y.device.type == "cuda"


def test_distributed_fit_matches_plain_fit(cuda, golden):
    """distributed=True on one GPU: the same fit as distributed=False (n=2000 configuration)."""
    g = golden("fit_n2000.npz")
    x, y = torch.from_numpy(g["x"]), torch.from_numpy(g["y"])
    kw = dict(embedding_dim=10, lambda_reg=1e-2, gamma=0.1, pcg_tol=1e-6, pcg_max_iter=1000, verbose=False)
    plain = laker.LAKERRegressor(**kw).fit(x, y)
    dist = laker.LAKERRegressor(distributed=True, **kw).fit(x, y)
    assert isinstance(dist.kernel_operator, DistributedAttentionKernelOperator)
    assert dist.kernel_operator.single_device
    # same device path: bit-identical solution and iteration count
    assert dist.pcg_iterations_ == plain.pcg_iterations_
    assert torch.equal(dist.alpha, plain.alpha)
    xt = torch.from_numpy(g["xt"])
    assert torch.equal(dist.predict(xt), plain.predict(xt))
    ref_it = int(g["iterations"])
    assert abs(dist.pcg_iterations_ - ref_it) <= 0.15 * ref_it
