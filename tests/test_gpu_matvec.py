"""Parity of the fused HIP matvec family with the FP64 oracle (GPU).

Error metric: |y - y_ref| / (lambda |x| + |G| |x|) row-wise (scale of the
summands), so cancellation in the sum does not inflate the tolerance.
Tolerances: FP32 Gram 2e-5, split Gram ("mixed": FP16 2-way for the square
k=1 product, BF16 3-way otherwise) 2e-5, hi/lo right-hand sides 2e-4, BF16
right-hand sides 1e-2, FP64 1e-12 or 1e-13.
"""

import numpy as np
import pytest
import torch

import laker
import laker_oracle as O

pytestmark = pytest.mark.gpu

TOL = {"fp32": 2e-5, "mixed": 2e-5, "hilo": 2e-4, "bf16": 1e-2}


def scale_rows(e, lam, x, clamp=None):
    e = np.asarray(e, np.float64)
    g = e @ e.T
    if clamp is not None:
        g = np.minimum(g, clamp)
    ax = np.abs(np.asarray(x, np.float64))
    return lam * ax + np.exp(g) @ ax


def norm_err(y, yref, scale):
    y = np.asarray(y, np.float64)
    return float(np.max(np.abs(y - yref) / scale))


def make(n, d, k, seed, amp=0.6, dtype=np.float32):
    rng = np.random.default_rng(seed)
    e = (rng.standard_normal((n, d)) * amp / np.sqrt(max(d, 1)) * 2).astype(dtype)
    x = rng.standard_normal((n, k) if k > 1 else n).astype(dtype)
    return e, x


CASES = [(1, 1, 1), (5, 3, 1), (31, 10, 1), (64, 10, 2), (100, 6, 3), (257, 16, 4), (300, 10, 5),
         (513, 16, 31), (700, 10, 64), (1000, 32, 70), (2000, 10, 1), (2000, 16, 1), (5000, 10, 1),
         (9000, 16, 1), (600, 4, 97), (800, 16, 130), (1500, 10, 200), (400, 5, 1), (777, 20, 1),
         (1000, 32, 1)]


@pytest.mark.parametrize("n,d,k", CASES)
@pytest.mark.parametrize("precision", ["fp32", "mixed", "hilo", "bf16"])
def test_matvec_matches_fp64(cuda, n, d, k, precision):
    e, x = make(n, d, k, seed=n + d + k)
    print(f"matvec n={n} d={d} k={k} precision={precision}", flush=True)
    op = laker.AttentionKernelOperator(torch.from_numpy(e).to(cuda), lambda_reg=0.05, precision=precision)
    y = op.matvec(torch.from_numpy(x).to(cuda)).cpu().numpy()
    yref = O.matvec_fp64(e, 0.05, x)
    s = scale_rows(e, 0.05, x)
    assert y.shape == x.shape
    # d > 16 with k > 4 has no packed 3-way kernel instance: hi/lo products
    tol = TOL["hilo"] if (precision == "mixed" and d > 16 and k > 4) else TOL[precision]
    assert norm_err(y, yref, s) < tol


def test_matvec_fp64_operator(cuda):
    e, x = make(400, 10, 3, seed=3, dtype=np.float64)
    op = laker.AttentionKernelOperator(torch.from_numpy(e).to(cuda), lambda_reg=0.05)
    y = op.matvec(torch.from_numpy(x).to(cuda)).cpu().numpy()
    assert norm_err(y, O.matvec_fp64(e, 0.05, x), scale_rows(e, 0.05, x)) < 1e-12


def test_matvec_reference_golden(cuda, golden):
    g = golden("operator.npz")
    for tag, tol in (("f32", 2e-5), ("f64", 1e-12)):
        e = g[f"{tag}_e"]
        op = laker.AttentionKernelOperator(torch.from_numpy(e).to(cuda), lambda_reg=0.05)
        for xk, yk in (("x1", "y1"), ("x2", "y2")):
            x = g[f"{tag}_{xk}"]
            y = op.matvec(torch.from_numpy(x).to(cuda)).cpu().numpy()
            assert norm_err(y, g[f"{tag}_{yk}"], scale_rows(e, 0.05, x)) < tol
        np.testing.assert_allclose(op.diagonal().cpu().numpy(), g[f"{tag}_diag"], rtol=1e-6)
        k = op.kernel_eval(torch.from_numpy(g[f"{tag}_q"]).to(cuda)).cpu().numpy()
        np.testing.assert_allclose(k, g[f"{tag}_keval"], rtol=2e-6)


def test_clamp_regime(cuda, golden):
    g = golden("operator.npz")
    e, x = g["clamp_e"], g["clamp_x1"]
    op = laker.AttentionKernelOperator(torch.from_numpy(e).to(cuda), lambda_reg=0.05)
    assert op.skip_clamp is False
    y = op.matvec(torch.from_numpy(x).to(cuda)).cpu().numpy()
    yref = O.matvec_fp64(e, 0.05, x, clamp=80.0)
    assert norm_err(y, yref, scale_rows(e, 0.05, x, clamp=80.0)) < 2e-5


@pytest.mark.parametrize("big", [2000.0, 5.0e4])
def test_square_product_split_range(cuda, big):
    """The mixed square k=1 product takes its FP16 2-way split only while the
    scaled embedding fits FP16 (max |c e| < 2^14 with c = scale log2(e)): one row
    e_0 = (big, 0, ..., 0) sets the range, every other row has a zero first
    coordinate, so G[0, j] = 0 exactly and G[0, 0] = big^2 is clamped.  At
    big = 5e4, c e_0 overflows FP16: a wrong choice of split shows up as NaN."""
    e, x = make(600, 10, 1, seed=31)
    e[:, 0] = 0.0
    e[0] = 0.0
    e[0, 0] = big
    op = laker.AttentionKernelOperator(torch.from_numpy(e).to(cuda), lambda_reg=0.05, precision="mixed")
    assert op.skip_clamp is False
    y = op.matvec(torch.from_numpy(x).to(cuda)).cpu().numpy()
    assert np.isfinite(y).all()
    yref = O.matvec_fp64(e, 0.05, x, clamp=80.0)
    assert norm_err(y, yref, scale_rows(e, 0.05, x, clamp=80.0)) < TOL["mixed"]


def test_cpu_inputs_roundtrip_device(cuda):
    e, x = make(50, 4, 1, seed=9)
    op = laker.AttentionKernelOperator(torch.from_numpy(e), lambda_reg=0.1)
    y = op.matvec(torch.from_numpy(x))
    assert y.device.type == "cpu" and y.shape == (50,)
    with pytest.raises(ValueError):
        op.matvec(torch.zeros(50, 2, 2))


@pytest.mark.parametrize("m,n,k", [(1, 50, 1), (17, 300, 1), (300, 1000, 2), (129, 2000, 5)])
def test_cross_matvec(cuda, m, n, k):
    e, x = make(n, 10, k, seed=m)
    q, _ = make(m, 10, 1, seed=m + 1)
    op = laker.AttentionKernelOperator(torch.from_numpy(e).to(cuda), lambda_reg=0.05)
    y = op.cross_matvec(torch.from_numpy(q).to(cuda), torch.from_numpy(x).to(cuda)).cpu().numpy()
    yref = O.cross_matvec_fp64(q, e, x)
    s = np.exp(q.astype(np.float64) @ e.astype(np.float64).T) @ np.abs(x.astype(np.float64))
    assert norm_err(y, yref, s) < 2e-5
    # transposed: exp(E Q^T) xq
    xq = np.random.default_rng(m).standard_normal((m, k) if k > 1 else m).astype(np.float32)
    yt = op.cross_matvec_t(torch.from_numpy(q).to(cuda), torch.from_numpy(xq).to(cuda)).cpu().numpy()
    yt_ref = O.cross_matvec_fp64(e, q, xq)
    st = np.exp(e.astype(np.float64) @ q.astype(np.float64).T) @ np.abs(xq.astype(np.float64))
    assert norm_err(yt, yt_ref, st) < 2e-5


@pytest.mark.parametrize("n,d", [(300, 10), (2049, 10), (7000, 16)])
@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_symmetric_product_matches_general(cuda, n, d, precision):
    """The square operator's k=1 product (symmetric kernel: each exp once per
    pair) against the general rectangular kernel on the same embeddings."""
    e, x = make(n, d, 1, seed=n + 7)
    print(f"symmetric vs general n={n} d={d} precision={precision}", flush=True)
    et, xt = torch.from_numpy(e).to(cuda), torch.from_numpy(x).to(cuda)
    op = laker.AttentionKernelOperator(et, lambda_reg=0.0, precision=precision)
    y_sym = op.matvec(xt).double().cpu().numpy()
    y_gen = op.cross_matvec(et, xt).double().cpu().numpy()
    s = scale_rows(e, 0.0, x)
    assert norm_err(y_sym, y_gen, s) < 2 * TOL[precision]
    assert norm_err(y_sym, O.matvec_fp64(e, 0.0, x), s) < TOL[precision]
    # bit-identical on repeat (fixed-order reductions)
    assert np.array_equal(op.matvec(xt).double().cpu().numpy(), y_sym)


@pytest.mark.parametrize("n,amp", [(3000, 0.6), (1500, 9.0), (1, 0.6), (17, 0.6), (257, 0.6), (1025, 0.6),
                                   (2049, 9.0)])
def test_residual_fp64_single_rhs(cuda, n, amp):
    """k=1 FP64 residual (symmetric FP64 MFMA Gram + table exp2, each exp once
    per pair) against NumPy FP64, with and without the clamp, on sizes that
    leave partial 16-row tiles and partial 256-row blocks."""
    e, _ = make(n, 10, 1, seed=n, amp=amp)
    rng = np.random.default_rng(7)
    x = rng.standard_normal(n) * 50
    b = rng.standard_normal(n)
    op = laker.AttentionKernelOperator(torch.from_numpy(e).to(cuda), lambda_reg=0.01)
    r = op.residual_fp64(torch.from_numpy(b).to(cuda), torch.from_numpy(x).to(cuda)).cpu().numpy()
    e64 = e.astype(np.float64)
    g = e64 @ e64.T
    if not op.skip_clamp:
        g = np.minimum(g, 80.0)
    kx = np.exp(g)
    rref = b - (0.01 * x + kx @ x)
    s = np.abs(b) + 0.01 * np.abs(x) + kx @ np.abs(x)
    assert (amp > 5) == (not op.skip_clamp)
    assert np.max(np.abs(r - rref) / s) < 1e-13


def test_to_dense_diagonal_kernel_eval(cuda):
    e, _ = make(70, 5, 1, seed=2)
    op = laker.AttentionKernelOperator(torch.from_numpy(e).to(cuda), lambda_reg=0.2)
    a = op.to_dense().cpu().numpy()
    ref = O.dense_matrix_fp64(e, 0.2)
    np.testing.assert_allclose(a, ref, rtol=2e-6)
    np.testing.assert_allclose(op.diagonal().cpu().numpy(), np.diag(ref), rtol=2e-6)
    q, _ = make(9, 5, 1, seed=4)
    k = op.kernel_eval(torch.from_numpy(q).to(cuda)).cpu().numpy()
    np.testing.assert_allclose(k, np.exp(q.astype(np.float64) @ e.astype(np.float64).T), rtol=2e-6)


def test_residual_fp64(cuda):
    e, _ = make(900, 10, 1, seed=5)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((900, 2)) * 50
    b = rng.standard_normal((900, 2))
    op = laker.AttentionKernelOperator(torch.from_numpy(e).to(cuda), lambda_reg=0.01)
    r = op.residual_fp64(torch.from_numpy(b).to(cuda), torch.from_numpy(x).to(cuda)).cpu().numpy()
    rref = b - O.matvec_fp64(e, 0.01, x)
    s = np.abs(b) + scale_rows(e, 0.01, x)
    assert np.max(np.abs(r - rref) / s) < 1e-13


@pytest.mark.parametrize("n", [20000, 100000])
def test_full_size_rows_subset(cuda, n):
    """Full-size operators (the n=100k workload): sampled rows against FP64."""
    x_loc, y = laker.pathloss_problem(n, seed=0)
    emb = laker.PositionEmbedding(2, 10, device=cuda)
    with torch.no_grad():
        e = emb(x_loc.to(cuda)).contiguous()
    rows = np.random.default_rng(0).choice(n, 48, replace=False)
    e64 = e.double().cpu().numpy()
    yv = y.numpy().astype(np.float64)
    g = np.exp(e64[rows] @ e64.T)
    ref = 0.01 * yv[rows] + g @ yv
    s = 0.01 * np.abs(yv[rows]) + g @ np.abs(yv)
    for precision, tol in (("fp32", 2e-5), ("mixed", 2e-5)):
        op = laker.AttentionKernelOperator(e, lambda_reg=0.01, precision=precision)
        out = op.matvec(y.to(cuda)).double().cpu().numpy()[rows]
        err = float(np.max(np.abs(out - ref) / s))
        print(f"n={n} {precision}: max row error {err:.2e}", flush=True)
        assert err < tol, precision
        op.close()
