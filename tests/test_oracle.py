"""Pin the CPU oracle against golden vectors, closed forms and brute force (no GPU)."""

import math

import numpy as np
import pytest
import torch

import laker_oracle as O


# ---------------------------------------------------------------- kernel operator
def brute_matvec(e, lam, x, clamp=None):
    n = e.shape[0]
    y = lam * np.array(x, dtype=np.float64)
    for i in range(n):
        for j in range(n):
            g = float(np.dot(e[i], e[j]))
            if clamp is not None:
                g = min(g, clamp)
            y[i] += math.exp(g) * x[j]
    return y


def test_matvec_brute_force_tiny():
    rng = np.random.default_rng(0)
    e = rng.standard_normal((9, 3))
    x = rng.standard_normal(9)
    y = O.attention_matvec(torch.from_numpy(e), 0.3, torch.from_numpy(x)).numpy()
    np.testing.assert_allclose(y, brute_matvec(e, 0.3, x), rtol=1e-12)
    np.testing.assert_allclose(O.matvec_fp64(e, 0.3, x), brute_matvec(e, 0.3, x), rtol=1e-12)


def test_matvec_chunked_equals_dense():
    rng = np.random.default_rng(1)
    e = torch.from_numpy(rng.standard_normal((100, 6)) * 0.5)
    x = torch.from_numpy(rng.standard_normal((100, 3)))
    full = O.attention_matvec(e, 0.1, x)
    for chunk in (16, 33, 99):
        np.testing.assert_allclose(O.attention_matvec(e, 0.1, x, chunk_size=chunk).numpy(), full.numpy(),
                                   rtol=1e-12)


def test_clamp_semantics():
    e = torch.tensor([[10.0, 0.0], [9.5, 0.0]], dtype=torch.float64)
    assert not O.skip_clamp_flag(e.float())
    assert O.skip_clamp_flag(e)  # 100 < 700 for fp64
    x = torch.tensor([1.0, 0.0], dtype=torch.float64)
    y = O.attention_matvec(e.float(), 0.0, x.float(), skip_clamp=False).double()
    assert abs(y[0].item() - math.exp(80.0)) / math.exp(80.0) < 1e-6
    assert abs(y[1].item() - math.exp(80.0)) / math.exp(80.0) < 1e-6


def test_operator_golden(golden):
    g = golden("operator.npz")
    for tag in ("f32", "f64"):
        e = torch.from_numpy(g[f"{tag}_e"])
        for xk, yk in (("x1", "y1"), ("x2", "y2")):
            x = torch.from_numpy(g[f"{tag}_{xk}"])
            y = O.attention_matvec(e, 0.05, x).numpy()
            tol = 1e-5 if tag == "f32" else 1e-12
            np.testing.assert_allclose(y, g[f"{tag}_{yk}"], rtol=tol, atol=tol)
            yc = O.attention_matvec(e, 0.05, x, chunk_size=16).numpy()
            np.testing.assert_allclose(yc, g[f"{tag}_{yk}c"], rtol=tol, atol=tol)
        np.testing.assert_allclose(O.diagonal(e, 0.05).numpy(), g[f"{tag}_diag"], rtol=1e-6)
        k = O.kernel_eval(torch.from_numpy(g[f"{tag}_q"]), e, O.skip_clamp_flag(e)).numpy()
        np.testing.assert_allclose(k, g[f"{tag}_keval"], rtol=1e-6)
    # clamp regime (max ||e||^2 >= 80): the reference drops the exponential there
    # (exp_safe's return value is discarded); the oracle keeps the documented
    # exp(min(g, 80)) and restates the reference's actual output separately.
    e = torch.from_numpy(g["clamp_e"])
    x1 = torch.from_numpy(g["clamp_x1"])
    assert bool(g["clamp_skip"]) == O.skip_clamp_flag(e) is False
    np.testing.assert_allclose(O.reference_clamp_branch_matvec(e, 0.05, x1).numpy(), g["clamp_y1"],
                               rtol=1e-5, atol=1e-3)
    y = O.attention_matvec(e.double(), 0.05, x1.double(), skip_clamp=False).numpy()
    yb = brute_matvec(g["clamp_e"].astype(np.float64), 0.05, g["clamp_x1"].astype(np.float64), clamp=700.0)
    np.testing.assert_allclose(y, yb, rtol=1e-10)


# ---------------------------------------------------------------- PCG
def test_pcg_exact_preconditioner_one_step():
    n = 20
    a = torch.diag(torch.arange(1, n + 1, dtype=torch.float64))
    b = torch.randn(n, dtype=torch.float64, generator=torch.Generator().manual_seed(0))
    res = O.pcg_solve(lambda v: a @ v, lambda v: torch.linalg.solve(a, v), b, tol=1e-12, max_iter=n)
    assert res.iterations == 1
    np.testing.assert_allclose(res.x.numpy(), (b / torch.arange(1, n + 1)).numpy(), rtol=1e-12)


def test_pcg_diagonal_closed_form_and_cg_bound():
    # CG on a diagonal SPD matrix with m distinct eigenvalues converges in <= m steps
    d = torch.tensor([1.0, 1.0, 2.0, 2.0, 5.0, 5.0, 5.0], dtype=torch.float64)
    b = torch.arange(1, 8, dtype=torch.float64)
    res = O.pcg_solve(lambda v: d * v, lambda v: v, b, tol=1e-12, max_iter=50)
    assert res.iterations <= 3
    np.testing.assert_allclose(res.x.numpy(), (b / d).numpy(), rtol=1e-10)


def test_pcg_breakdown_and_zero_rhs():
    d = torch.tensor([1.0, -2.0, 3.0], dtype=torch.float64)
    with pytest.raises(RuntimeError, match="breakdown"):
        O.pcg_solve(lambda v: d * v, lambda v: v, torch.tensor([0.0, 1.0, 0.0], dtype=torch.float64))
    res = O.pcg_solve(lambda v: d * v, lambda v: v, torch.zeros(3, dtype=torch.float64))
    assert res.iterations == 0 and float(res.x.abs().sum()) == 0.0


def test_pcg_frobenius_stopping_rule():
    # two columns with very different scales: the global ratio stops both together
    d = torch.linspace(1.0, 50.0, 40, dtype=torch.float64)
    b = torch.stack([torch.ones(40, dtype=torch.float64), 1e-6 * torch.linspace(-1, 1, 40, dtype=torch.float64)], 1)
    res = O.pcg_solve(lambda v: d.unsqueeze(1) * v, lambda v: v, b, tol=1e-8, max_iter=200)
    r = b - d.unsqueeze(1) * res.x
    assert torch.linalg.norm(r) / torch.linalg.norm(b) <= 1e-8
    # the small column alone is not resolved to 1e-8 relative: global rule, not per column
    res1 = O.pcg_solve(lambda v: d * v, lambda v: v, b[:, 1], tol=1e-8, max_iter=200)
    assert res1.iterations >= 1


def test_pcg_matches_fp64_cholesky(golden):
    g = golden("cccp_pcg_f64.npz")
    e = torch.from_numpy(g["e"])
    b = torch.from_numpy(g["b"])
    fac = O.cccp_build(lambda v: O.attention_matvec(e, 1e-2, v), 300, torch.from_numpy(g["probes"]),
                       0.1, 1e-8, 0.05, 200, 1e-6)
    res = O.pcg_solve(lambda v: O.attention_matvec(e, 1e-2, v), lambda v: O.cccp_apply(v, fac), b,
                      tol=1e-10, max_iter=300)
    x64 = O.dense_solve_fp64(g["e"], 1e-2, g["b"])
    assert np.linalg.norm(res.x.numpy() - x64) / np.linalg.norm(x64) < 1e-6
    # the reference's own trajectory, bit for bit
    assert res.iterations == int(g["iterations"])
    np.testing.assert_allclose(res.x.numpy(), g["x"], rtol=1e-9, atol=1e-9)
    res2 = O.pcg_solve(lambda v: O.attention_matvec(e, 1e-2, v), lambda v: O.cccp_apply(v, fac),
                       torch.from_numpy(g["b2"]), tol=1e-8, max_iter=300)
    assert res2.iterations == int(g["iterations2"])
    np.testing.assert_allclose(res2.x.numpy(), g["x2"], rtol=1e-9, atol=1e-9)


# ---------------------------------------------------------------- CCCP
def test_shrinkage_closed_form():
    assert O.adaptive_shrinkage_rho(200, 2000, 0.1, 0.05) == 0.5
    assert O.adaptive_shrinkage_rho(2000, 2000, 0.1, 0.05) == 0.05
    assert abs(O.adaptive_shrinkage_rho(50, 100, 0.01, 0.05) - (0.05 + 0.95 * 0.5 * 0.1)) < 1e-15
    assert O.default_num_probes(100000) == 632 and O.default_num_probes(2000) == 200
    assert O.default_num_probes(50000) == 447 and O.default_num_probes(20000) == 282


def test_cccp_golden_and_invariants(golden):
    g = golden("cccp_pcg_f64.npz")
    e = torch.from_numpy(g["e"])
    op = lambda v: O.attention_matvec(e, 1e-2, v)  # noqa: E731
    fac = O.cccp_build(op, 300, torch.from_numpy(g["probes"]), 0.1, 1e-8, 0.05, 200, 1e-6)
    assert abs(fac["isotropic_coef"] - float(g["isotropic_coef"])) < 1e-12
    np.testing.assert_allclose(fac["q_basis_matrix"].numpy(), g["q_basis_matrix"], rtol=1e-9, atol=1e-12)
    pv = O.cccp_apply(torch.from_numpy(g["v"]), fac).numpy()
    np.testing.assert_allclose(pv, g["pv"], rtol=1e-9, atol=1e-12)
    # invariants: a > 0, tr(Sigma) = n, M SPD, P symmetric and == Sigma^{-1/2}
    a, q, c = fac["isotropic_coef"], fac["q_basis"], fac["q_basis_matrix"]
    assert a > 0
    sigma = a * torch.eye(300, dtype=torch.float64) + q @ c @ q.T
    assert abs(torch.trace(sigma).item() - 300) < 1e-8
    assert torch.linalg.eigvalsh(a * torch.eye(60, dtype=torch.float64) + c).min() > 0
    p = O.cccp_apply(torch.eye(300, dtype=torch.float64), fac)
    np.testing.assert_allclose(p.numpy(), p.T.numpy(), atol=1e-10)
    w, v = torch.linalg.eigh(sigma)
    np.testing.assert_allclose(p.numpy(), (v @ torch.diag(w.rsqrt()) @ v.T).numpy(), atol=1e-8)


def test_cccp_cholesky_route_agrees_with_eigh(golden):
    g = golden("cccp_pcg_f64.npz")
    a, c, its = O.cccp_loop_fp64_cholesky(g["qr_r"], 300, 0.1, 1e-8, 0.05, 200, 1e-6)
    assert abs(a - float(g["isotropic_coef"])) < 1e-12
    np.testing.assert_allclose(c, g["q_basis_matrix"], rtol=1e-8, atol=1e-11)


def test_cccp_linearity():
    rng = np.random.default_rng(4)
    e = torch.from_numpy(rng.standard_normal((40, 4)) * 0.5)
    op = lambda v: O.attention_matvec(e, 1e-2, v)  # noqa: E731
    fac = O.cccp_build(op, 40, O.reference_probes(40, 30, 0, torch.float64), 0.1, 1e-8, 0.05, 15, 1e-5)
    u, v = torch.from_numpy(rng.standard_normal(40)), torch.from_numpy(rng.standard_normal(40))
    lhs = O.cccp_apply(2.5 * u - 1.3 * v, fac)
    rhs = 2.5 * O.cccp_apply(u, fac) - 1.3 * O.cccp_apply(v, fac)
    np.testing.assert_allclose(lhs.numpy(), rhs.numpy(), rtol=1e-10, atol=1e-12)


# ---------------------------------------------------------------- embeddings and pipeline
def test_position_embedding_golden(golden):
    g = golden("position_embedding.npz")
    x = torch.from_numpy(g["x"])
    for d in (10, 16):
        emb = O.PositionEmbeddingOracle(2, d)
        np.testing.assert_array_equal(emb.freq.numpy(), g[f"freq{d}"])
        np.testing.assert_array_equal(emb.phase.numpy(), g[f"phase{d}"])
        np.testing.assert_array_equal(emb.layers[0][0].numpy(), g[f"w1_{d}"])
        np.testing.assert_array_equal(emb.layers[1][1].numpy(), g[f"b2_{d}"])
        np.testing.assert_array_equal(emb(x).numpy(), g[f"e{d}"])


def test_fit_n2000_matches_reference_bit_for_bit(golden):
    g = golden("fit_n2000.npz")
    r = O.fit(torch.from_numpy(g["x"]), torch.from_numpy(g["y"]))
    assert r["iterations"] == int(g["iterations"])
    np.testing.assert_array_equal(r["embeddings"].numpy(), g["embeddings"])
    np.testing.assert_array_equal(r["alpha"].numpy(), g["alpha"])
    assert r["factors"]["isotropic_coef"] == pytest.approx(float(g["isotropic_coef"]), rel=1e-7)
    q = r["embedding_model"](torch.from_numpy(g["xt"]))
    pred = O.predict(q, r["embeddings"], r["alpha"], r["skip_clamp"])
    np.testing.assert_allclose(pred.numpy(), g["pred"], rtol=1e-6, atol=1e-6)


def test_paper_n3_example(golden):
    g = golden("paper_n3.npz")
    a = O.dense_matrix_fp64(g["e"], 0.1)
    alpha = O.dense_solve_fp64(g["e"], 0.1, g["y"])
    np.testing.assert_allclose(a @ alpha, g["y"], rtol=1e-12)
    res = O.pcg_solve(lambda v: O.attention_matvec(torch.from_numpy(g["e"]), 0.1, v), lambda v: v,
                      torch.from_numpy(g["y"]), tol=1e-12, max_iter=10)
    np.testing.assert_allclose(res.x.numpy(), alpha, rtol=1e-8)
    pred = O.cross_matvec_fp64(g["e_star"], g["e"], alpha)
    assert np.isfinite(pred).all()


def test_pathloss_problem_matches_golden(golden):
    import laker

    g = golden("fit_n2000.npz")
    x, y = laker.pathloss_problem(2000, seed=0)
    np.testing.assert_array_equal(x.numpy(), g["x"])
    np.testing.assert_array_equal(y.numpy(), g["y"])
