"""Batched predictive variance at full size: n=50,000 training points,
embedding_dim=16, lambda_reg=1e-3, 512 query locations solved as 512
right-hand sides of one batched PCG (reference laker/core.py:666-777), in
plain fp32 and in the mixed refined form, against the FP64 variance of the
first 64 queries from a dense FP64 Cholesky factor on the GPU."""

import time

import pytest
import torch

import laker

pytestmark = pytest.mark.gpu

N_TRAIN, D, LAM, M, M64 = 50_000, 16, 1e-3, 512, 64


def fp64_variance(e, q, lam):
    """k(q,q) - k_q^T A^{-1} k_q with A = lam I + exp(E E^T), all FP64."""
    a = torch.exp(e @ e.T)
    a.diagonal().add_(lam)
    l_fac = torch.linalg.cholesky(a)
    del a
    kxq = torch.exp(e @ q.T)
    w = torch.linalg.solve_triangular(l_fac, kxq, upper=False)
    del l_fac
    return torch.exp((q * q).sum(1)) - (w * w).sum(0)


def test_variance_c2_fp32_and_mixed(cuda):
    x, y = laker.pathloss_problem(N_TRAIN, seed=0)
    xq = torch.rand(M, 2, generator=torch.Generator().manual_seed(1)) * 100.0
    out = {}
    for precision in ("mixed", "fp32"):
        t0 = time.perf_counter()
        model = laker.LAKERRegressor(embedding_dim=D, lambda_reg=LAM, pcg_tol=1e-6, pcg_max_iter=1000,
                                     verbose=False, precision=precision, probe_source="device",
                                     probe_precision="mixed" if precision == "mixed" else "fp32")
        model.fit(x, y)
        var = model.predict_variance(xq)
        torch.cuda.synchronize()
        solver = model.core.last_solver
        print(f"variance {precision}: {time.perf_counter() - t0:.1f} s, batched PCG iterations "
              f"{solver.iterations}, outer {getattr(solver, 'outer_iterations', 0)}", flush=True)
        out[precision] = var.double()
        emb = model.embedding_model
        e = model.embeddings.double()
    with torch.no_grad():
        q = emb(xq[:M64].to(cuda)).double()
    ref = fp64_variance(e, q, LAM)
    scale = torch.exp((q * q).sum(1))
    errs = {}
    for precision in out:
        errs[precision] = float(((out[precision][:M64] - ref.clamp(min=0.0)).abs() / scale).max())
        print(f"variance {precision}: max |var - var64| / k(q,q) over {M64} queries = {errs[precision]:.2e}",
              flush=True)
    # At lambda = 1e-3 and n = 50,000 the FP32-level operator error exceeds
    # lambda, so refinement cannot contract to 1e-6: the mixed solve keeps its
    # best FP64 iterate.  Both legs are measured at about 1e-6 of k(q,q).
    assert bool((out["mixed"] >= 0).all()) and out["mixed"].shape == (M,)
    assert errs["mixed"] < 1e-5, errs
    assert errs["fp32"] < 1e-5, errs
