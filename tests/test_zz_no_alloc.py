"""laker_matvec_noalloc through the C ABI against laker_matvec (GPU).

Two calls in a row on one GPU, the second on a smaller operator, so the
shared scratch array holds the first call's partial sums when the second
starts; each result must match laker_matvec's and the FP64 product."""

import ctypes

import numpy as np
import pytest
import torch

import laker
import laker._native as N
import laker_oracle as O

pytestmark = pytest.mark.gpu


def _noalloc(op, x2, precision):
    y = torch.empty_like(x2)
    rc = N.lib.laker_matvec_noalloc(op.handle, N.ptr(x2), x2.shape[1], x2.shape[1], N.ptr(y), y.shape[1],
                                    precision, N.stream_ptr())
    return rc, y


def _case(cuda, n, k, d, precision, seed):
    g = torch.Generator().manual_seed(seed)
    e = (torch.randn(n, d, generator=g) * 0.4).to(cuda)
    x = torch.randn(n, k, generator=g).to(cuda).contiguous()
    op = laker.AttentionKernelOperator(e, lambda_reg=0.05, precision="mixed" if precision == N.PREC_MIXED else "fp32")
    rc, y = _noalloc(op, x, precision)
    assert rc == N.LAKER_OK, N.last_error()
    y_ref = torch.empty_like(x)
    N.check(N.lib.laker_matvec(op.handle, N.ptr(x), k, k, N.ptr(y_ref), k, precision, N.stream_ptr()))
    torch.cuda.synchronize()
    e64, x64 = e.double().cpu().numpy(), x.double().cpu().numpy()
    ref64 = O.matvec_fp64(e64, 0.05, x64)
    scale = 0.05 * np.abs(x64) + O.matvec_fp64(e64, 0.0, np.abs(x64))
    err = float(np.max(np.abs(y.double().cpu().numpy() - ref64) / scale))
    err_entry = float(np.max(np.abs(y.double().cpu().numpy() - y_ref.double().cpu().numpy()) / scale))
    print(f"n={n} k={k} d={d} precision={precision}: vs FP64 {err:.2e}, vs laker_matvec {err_entry:.2e}", flush=True)
    assert err < 2e-5
    assert err_entry < 4e-5
    return op


@pytest.mark.parametrize("precision", [N.PREC_FP32, N.PREC_MIXED])
def test_noalloc_matches_matvec_twice(cuda, precision):
    op1 = _case(cuda, 20000, 3, 10, precision, 1)   # first call: larger operator, k = 3
    op2 = _case(cuda, 5000, 1, 16, precision, 2)    # second call: smaller operator, k = 1
    op3 = _case(cuda, 3001, 2, 10, precision, 3)    # third call: ragged n, k = 2
    for op in (op1, op2, op3):
        op.close()


def test_noalloc_rejects_beyond_maximum(cuda):
    # n = 131,073 with k = 3 needs more than the 16 MiB scratch: rejected on the host
    e = torch.zeros(131073, 2, device=cuda)
    op = laker.AttentionKernelOperator(e, lambda_reg=1.0)
    x = torch.zeros(131073, 3, device=cuda)
    rc, _ = _noalloc(op, x, N.PREC_FP32)
    assert rc == N.LAKER_EINVAL and "maximum" in N.last_error()
    rc, _ = _noalloc(op, x[:, :2].contiguous(), N.PREC_FP64)
    assert rc == N.LAKER_EINVAL
    x5 = torch.zeros(131073, 5, device=cuda)
    rc, _ = _noalloc(op, x5, N.PREC_FP32)
    assert rc == N.LAKER_EINVAL
    # k = 2 at this n fits (n <= 262144 for k <= 2)
    rc, y = _noalloc(op, x[:, :2].contiguous(), N.PREC_FP32)
    assert rc == N.LAKER_OK, N.last_error()
    torch.cuda.synchronize()
    # E = 0: y = lambda x + (all-ones kernel) x = 0 for x = 0
    assert float(y.abs().max()) == 0.0
    op.close()
