"""CPU checks that pin the oracle functions the other oracle tests leave open.

Each check compares an oracle function with something independent of its own
formula: a brute-force double loop, a reference golden vector, a closed form,
or another branch of the reference's tiling that must give the same numbers.
A transposed operand, a dropped scale or clamp, or a wrong branch condition
fails at least one of them.
"""

import math

import numpy as np
import pytest
import torch

import laker_oracle as O


def brute_cross(q, e, x, scale=1.0, clamp=None):
    """out[i, c] = sum_j exp(min(scale * sum_t q[i,t] e[j,t], clamp)) x[j, c], by loops."""
    q = np.asarray(q, np.float64)
    e = np.asarray(e, np.float64)
    x = np.asarray(x, np.float64)
    x2 = x.reshape(x.shape[0], -1)
    out = np.zeros((q.shape[0], x2.shape[1]))
    for i in range(q.shape[0]):
        for j in range(e.shape[0]):
            g = scale * sum(q[i, t] * e[j, t] for t in range(q.shape[1]))
            if clamp is not None:
                g = min(g, clamp)
            out[i] += math.exp(g) * x2[j]
    return out.reshape((q.shape[0],) + x.shape[1:])


def test_cross_matvec_fp64_brute_force():
    rng = np.random.default_rng(3)
    q = rng.normal(size=(5, 3))
    e = rng.normal(size=(7, 3))
    x1 = rng.normal(size=7)
    x2 = rng.normal(size=(7, 4))
    # block=2 walks the query rows in several blocks
    for x in (x1, x2):
        np.testing.assert_allclose(O.cross_matvec_fp64(q, e, x, block=2), brute_cross(q, e, x), rtol=1e-13)
        np.testing.assert_allclose(O.cross_matvec_fp64(q, e, x, scale=0.7, clamp=0.5, block=3),
                                   brute_cross(q, e, x, 0.7, 0.5), rtol=1e-13)
    # the clamp is active in that case: without it the result differs
    assert not np.allclose(brute_cross(q, e, x1, 0.7, 0.5), brute_cross(q, e, x1, 0.7))


def test_cross_matvec_fp64_against_reference_kernel_eval(golden):
    """The reference's kernel_eval(q) (exp(q E^T), laker/kernels.py:301-354) times x."""
    g = golden("operator.npz")
    rng = np.random.default_rng(4)
    e, q, keval = g["f64_e"], g["f64_q"], g["f64_keval"]
    x = rng.normal(size=(e.shape[0], 3))
    np.testing.assert_allclose(O.cross_matvec_fp64(q, e, x), keval @ x, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(O.cross_matvec_fp64(q, e, x[:, 0]), keval @ x[:, 0], rtol=1e-12, atol=1e-12)
    # and the FP32 reference output to FP32 accuracy
    e32, q32, k32 = g["f32_e"], g["f32_q"], g["f32_keval"]
    x32 = rng.normal(size=e32.shape[0])
    np.testing.assert_allclose(O.cross_matvec_fp64(q32, e32, x32), k32.astype(np.float64) @ x32, rtol=2e-5,
                               atol=1e-5)


def test_matvec_fp64_brute_force():
    rng = np.random.default_rng(5)
    e = rng.normal(size=(9, 3))
    x = rng.normal(size=(9, 2))
    lam = 0.3
    ref = lam * x + brute_cross(e, e, x, 1.3, 0.9)
    np.testing.assert_allclose(O.matvec_fp64(e, lam, x, scale=1.3, clamp=0.9, block=4), ref, rtol=1e-13)
    ref1 = lam * x[:, 0] + brute_cross(e, e, x[:, 0])
    np.testing.assert_allclose(O.matvec_fp64(e, lam, x[:, 0], block=2), ref1, rtol=1e-13)


def test_exp_safe_clamp_thresholds():
    """exp(min(g, 80)) in fp32, exp(min(g, 700)) in fp64, no clamp when skipped (laker/kernels.py:80-113)."""
    g32 = torch.tensor([1.0, 79.0, 95.0], dtype=torch.float32)
    out = O.exp_safe(g32)
    assert out[2].item() == pytest.approx(math.exp(80.0), rel=1e-6)
    assert out[1].item() == pytest.approx(math.exp(79.0), rel=1e-6)
    assert torch.isinf(O.exp_safe(g32, skip_clamp=True)[2])
    g64 = torch.tensor([699.0, 720.0], dtype=torch.float64)
    out64 = O.exp_safe(g64)
    assert out64[1].item() == pytest.approx(math.exp(700.0), rel=1e-14)
    assert out64[0].item() == pytest.approx(math.exp(699.0), rel=1e-14)


def test_to_dense_brute_force_and_no_clamp():
    """lam I + exp(E E^T) without the clamp (laker/kernels.py:285-299)."""
    rng = np.random.default_rng(6)
    e = rng.normal(size=(6, 2))
    lam = 0.25
    ref = brute_cross(e, e, np.eye(6)) + lam * np.eye(6)
    np.testing.assert_allclose(O.to_dense(torch.from_numpy(e), lam).numpy(), ref, rtol=1e-13)
    # norms far above the fp64 clamp point: entries keep exp(g) for g > 700
    big = torch.tensor([[26.6], [26.5]], dtype=torch.float64)  # 26.6^2 = 707.56
    d = O.to_dense(big, 0.0).numpy()
    assert d[0, 0] == pytest.approx(math.exp(26.6 * 26.6), rel=1e-12)
    assert d[0, 0] > math.exp(700.0) * 1e3


def test_attention_matvec_2d_tile_branch():
    """The reference's 2-D tiling (chunk * n * elt > 64 MiB, laker/kernels.py:246-270)."""
    n, chunk = 9000, 2000
    assert chunk * n * 4 > 64 * 1024 * 1024  # fp32 takes the 2-D branch
    g = torch.Generator().manual_seed(8)
    e64 = torch.randn(n, 4, generator=g, dtype=torch.float64) * 0.4
    x64 = torch.randn(n, 3, generator=g, dtype=torch.float64)
    # fp64: 2-D tiles against the dense branch (same arithmetic up to summation order)
    y2d = O.attention_matvec(e64, 0.05, x64, chunk_size=chunk)
    yd = O.attention_matvec(e64, 0.05, x64, chunk_size=None)
    np.testing.assert_allclose(y2d.numpy(), yd.numpy(), rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(y2d.numpy(), O.matvec_fp64(e64.numpy(), 0.05, x64.numpy()), rtol=1e-11,
                               atol=1e-11)
    # fp32, one right-hand side: against the FP64 product to FP32 accuracy
    e32, x32 = e64.float(), x64[:, 0].float()
    y32 = O.attention_matvec(e32, 0.05, x32, chunk_size=chunk).double().numpy()
    y_ref = O.matvec_fp64(e32.double().numpy(), 0.05, x32.double().numpy())
    scale = 0.05 * np.abs(x32.double().numpy()) + O.matvec_fp64(e32.double().numpy(), 0.0,
                                                               np.abs(x32.double().numpy()))
    assert np.max(np.abs(y32 - y_ref) / scale) < 2e-5


@pytest.mark.parametrize("n,expected", [(100, None), (5000, None), (5001, 1024), (20000, 2000),
                                        (50000, 5000), (81920, 8192), (100000, 8192), (300000, 8192)])
def test_auto_chunk_closed_form(n, expected):
    """n > 5000 -> max(1024, min(n // 10, 8192)) (laker/core.py:340-345); a given chunk is kept."""
    assert O.auto_chunk(n) == expected
    assert O.auto_chunk(n, 333) == 333


def test_auto_chunk_matches_package():
    import laker.core as core

    for n in (10, 5000, 5001, 12345, 65536, 100000):
        assert core.auto_chunk(n, None) == O.auto_chunk(n)


def test_predict_variance_exact_n3():
    """k(q,q) - k_q^T A^{-1} k_q at n = 3 against the explicit inverse
    (reference tests/test_math_correctness.py:503-551 checks 1e-3; this checks 1e-10)."""
    g = torch.Generator().manual_seed(42)
    e = torch.randn(3, 2, generator=g, dtype=torch.float64)
    eq = torch.randn(2, 2, generator=g, dtype=torch.float64)
    lam = 0.1
    a = O.dense_matrix_fp64(e.numpy(), lam)
    kq = np.exp(eq.numpy() @ e.numpy().T)  # (2, 3)
    exact = np.exp(np.sum(eq.numpy() ** 2, axis=1)) - np.einsum("ij,ji->i", kq, np.linalg.inv(a) @ kq.T)

    def op(v):
        return O.attention_matvec(e, lam, v)

    var, res = O.predict_variance(eq, e, lam, op, lambda v: v, True, pcg_tol=1e-14, pcg_max_iter=50)
    np.testing.assert_allclose(var.numpy(), np.maximum(exact, 0.0), rtol=1e-10, atol=1e-12)


def test_predict_variance_cccp_against_dense():
    """Batched PCG variance (2-D solve, Frobenius stop) with a CCCP preconditioner vs the dense formula."""
    g = torch.Generator().manual_seed(9)
    n, m = 300, 7
    e = torch.randn(n, 5, generator=g, dtype=torch.float64) * 0.5
    eq = torch.randn(m, 5, generator=g, dtype=torch.float64) * 0.5
    lam = 1e-2

    def op(v):
        return O.attention_matvec(e, lam, v)

    probes = O.reference_probes(n, 60, 5, torch.float64)
    fac = O.cccp_build(op, n, probes, 0.1, 1e-8, 0.05, 200, 1e-6)
    var, res = O.predict_variance(eq, e, lam, op, lambda v: O.cccp_apply(v, fac), True, pcg_tol=1e-12,
                                  pcg_max_iter=500)
    a = O.dense_matrix_fp64(e.numpy(), lam)
    kxq = np.exp(e.numpy() @ eq.numpy().T)
    exact = np.exp(np.sum(eq.numpy() ** 2, axis=1)) - np.sum(kxq * np.linalg.solve(a, kxq), axis=0)
    kqq = np.exp(np.sum(eq.numpy() ** 2, axis=1))
    assert np.max(np.abs(var.numpy() - np.maximum(exact, 0.0)) / kqq) < 1e-7
    assert res.iterations < 500


def test_pcg_result_fields():
    r = O.pcg_solve(lambda v: 2.0 * v, lambda v: 0.5 * v, torch.ones(4, dtype=torch.float64), tol=1e-12,
                    max_iter=5)
    assert isinstance(r, O.PCGResult)
    assert r.iterations == 1 and r.converged
    np.testing.assert_allclose(r.x.numpy(), 0.5 * np.ones(4), rtol=1e-15)
    assert r.residual_norm == pytest.approx(0.0, abs=1e-15)
