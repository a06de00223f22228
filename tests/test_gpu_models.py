"""LAKERRegressor end to end on the GPU against the reference's golden vectors,
the paper's n=3 example and the FP64 oracle."""

import os

import numpy as np
import pytest
import torch

import laker
import laker_oracle as O

pytestmark = pytest.mark.gpu


def test_fit_predict_score_n2000(cuda, golden):
    g = golden("fit_n2000.npz")
    x, y = torch.from_numpy(g["x"]), torch.from_numpy(g["y"])
    model = laker.LAKERRegressor(embedding_dim=10, lambda_reg=1e-2, gamma=0.1, pcg_tol=1e-6,
                                 pcg_max_iter=1000, verbose=False)
    model.fit(x, y)
    e = model.embeddings.cpu().numpy()
    assert np.max(np.abs(e - g["embeddings"])) < 2e-5
    ref_it = int(g["iterations"])
    assert abs(model.pcg_iterations_ - ref_it) <= 0.15 * ref_it
    e64 = g["embeddings"].astype(np.float64)
    a64 = O.dense_solve_fp64(e64, 1e-2, g["y"].astype(np.float64))
    q = O.PositionEmbeddingOracle(2, 10)(torch.from_numpy(g["xt"])).double().numpy()
    pred64 = O.cross_matvec_fp64(q, e64, a64)
    pred = model.predict(torch.from_numpy(g["xt"])).double().cpu().numpy()
    err_pred = np.linalg.norm(pred - pred64) / np.linalg.norm(pred64)
    err_ref = np.linalg.norm(g["pred"] - pred64) / np.linalg.norm(pred64)
    print(f"prediction error vs FP64: HIP {err_pred:.2e}, reference {err_ref:.2e}", flush=True)
    # both FP32 solves sit ~1e-2 from the FP64 predictions (kappa ~ 2e5), so they
    # are compared through their distance to FP64, not to each other
    assert err_pred < max(3 * err_ref, 1e-4)
    assert np.linalg.norm(pred - g["pred"]) / np.linalg.norm(pred64) < 2 * (err_pred + err_ref)
    assert model.score(x, y) == pytest.approx(float(g["score"]), abs=2e-3)
    assert model.score_r2(x, y) <= 1.0


def test_fit_mixed_precision_n2000(cuda, golden):
    g = golden("fit_n2000.npz")
    x, y = torch.from_numpy(g["x"]), torch.from_numpy(g["y"])
    model = laker.LAKERRegressor(lambda_reg=1e-2, pcg_tol=1e-6, verbose=False, precision="mixed")
    model.fit(x, y)
    e64 = model.embeddings.double().cpu().numpy()
    a64 = O.dense_solve_fp64(e64, 1e-2, g["y"].astype(np.float64))
    am = model.alpha_fp64.cpu().numpy()
    assert np.linalg.norm(am - a64) / np.linalg.norm(a64) < 2e-5
    a = O.dense_matrix_fp64(e64, 1e-2)
    assert np.linalg.norm(a @ am - g["y"]) / np.linalg.norm(g["y"]) <= 1.01e-6


def test_paper_example(cuda, golden):
    g = golden("paper_n3.npz")
    e = torch.from_numpy(g["e"])
    y = torch.from_numpy(g["y"])

    class FixedEmbedding(torch.nn.Module):
        def forward(self, x):
            return e.to(x.device)

    model = laker.LAKERRegressor(embedding_dim=2, lambda_reg=0.1, gamma=0.0, num_probes=3, cccp_max_iter=10,
                                 cccp_tol=1e-6, pcg_tol=1e-12, pcg_max_iter=10, embedding_module=FixedEmbedding(),
                                 verbose=False, dtype=torch.float64)
    model.fit(torch.zeros(3, 2, dtype=torch.float64), y)
    alpha = O.dense_solve_fp64(g["e"], 0.1, g["y"])
    np.testing.assert_allclose(model.alpha.cpu().numpy(), alpha, rtol=1e-4, atol=1e-3)
    es = torch.from_numpy(g["e_star"])

    class QueryEmbedding(torch.nn.Module):
        def forward(self, x):
            return es.to(x.device)

    model.embedding_model = QueryEmbedding()
    pred = model.predict(torch.zeros(1, 2, dtype=torch.float64)).item()
    assert abs(pred - O.cross_matvec_fp64(g["e_star"], g["e"], alpha)[0]) < 1e-3


def test_variance_tiny_exact(cuda):
    gen = torch.Generator().manual_seed(42)
    e = torch.randn(3, 2, dtype=torch.float64, generator=gen)
    y = torch.randn(3, dtype=torch.float64, generator=gen)
    eq = torch.randn(1, 2, dtype=torch.float64, generator=gen)
    a = O.dense_matrix_fp64(e.numpy(), 0.1)
    kq = np.exp(eq.numpy() @ e.numpy().T)
    exact = float((np.exp(eq.numpy() @ eq.numpy().T)[0, 0] - kq @ np.linalg.solve(a, kq.T)).item())

    class Fixed(torch.nn.Module):
        def __init__(self, t):
            super().__init__()
            self.t = t

        def forward(self, x):
            return self.t.to(x.device)

    model = laker.LAKERRegressor(embedding_dim=2, lambda_reg=0.1, gamma=0.0, num_probes=3, cccp_max_iter=5,
                                 pcg_tol=1e-12, pcg_max_iter=10, dtype=torch.float64, verbose=False,
                                 embedding_module=Fixed(e))
    model.fit(torch.zeros(3, 2, dtype=torch.float64), y)
    model.embedding_model = Fixed(eq)
    var = model.predict_variance(torch.zeros(1, 2, dtype=torch.float64)).item()
    assert abs(var - max(exact, 0.0)) < 1e-3


def test_variance_batched_matches_fp64(cuda):
    x, y = laker.pathloss_problem(1500, seed=2)
    model = laker.LAKERRegressor(embedding_dim=16, lambda_reg=1e-3, pcg_tol=1e-8, verbose=False)
    model.fit(x, y)
    xq = torch.rand(40, 2, generator=torch.Generator().manual_seed(1)) * 100
    var = model.predict_variance(xq).double().cpu().numpy()
    e = model.embeddings.double().cpu().numpy()
    with torch.no_grad():
        q = model.embedding_model(xq.to(cuda)).double().cpu().numpy()
    a = O.dense_matrix_fp64(e, 1e-3)
    kxq = np.exp(e @ q.T)
    exact = np.maximum(np.exp(np.sum(q * q, 1)) - np.sum(kxq * np.linalg.solve(a, kxq), 0), 0)
    scale = np.exp(np.sum(q * q, 1))
    assert np.max(np.abs(var - exact) / scale) < 1e-4


def test_chunked_predict_and_save_load(cuda, tmp_path):
    x, y = laker.pathloss_problem(300, seed=4)
    model = laker.LAKERRegressor(embedding_dim=6, num_probes=40, cccp_max_iter=20, pcg_tol=1e-6,
                                 chunk_size=16, verbose=False)
    model.fit(x, y)
    xt = torch.rand(50, 2, generator=torch.Generator().manual_seed(0)) * 100
    p1 = model.predict(xt)
    model.chunk_size = None
    p2 = model.predict(xt)
    assert torch.allclose(p1, p2, rtol=1e-5, atol=1e-6)
    path = os.path.join(tmp_path, "m.pt")
    model.save(path)
    loaded = laker.LAKERRegressor.load(path)
    assert torch.allclose(loaded.predict(xt).to(p2.device), p2, rtol=1e-5, atol=1e-6)


def test_bf16_embedding_dtype(cuda):
    x, y = laker.pathloss_problem(100, seed=42)
    model = laker.LAKERRegressor(embedding_dim=10, num_probes=50, cccp_max_iter=20, cccp_tol=1e-4,
                                 pcg_tol=1e-6, pcg_max_iter=500, embedding_dtype=torch.bfloat16,
                                 dtype=torch.float32, verbose=False)
    model.fit(x, y)
    assert model.embeddings.dtype == torch.float32 and model.alpha is not None
