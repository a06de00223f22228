"""Hand-written matvec / cross-matvec backward against FP64 autograd and central
finite differences (GPU).  Gradients are FP32 (tiles recomputed); tolerance
1e-4 relative to the FP64 reference gradient's norm."""

import numpy as np
import pytest
import torch

import laker
import laker_oracle as O
from laker.core import cross_kernel_matvec

pytestmark = pytest.mark.gpu


def dense_fp64(e, lam, x):
    return lam * x + torch.exp(e @ e.T) @ x


@pytest.mark.parametrize("n,d,k", [(64, 4, 1), (150, 10, 3), (300, 16, 1)])
def test_matvec_backward_vs_fp64_autograd(cuda, n, d, k):
    gen = torch.Generator().manual_seed(n)
    e = (torch.randn(n, d, generator=gen, dtype=torch.float64) * 0.4)
    x = torch.randn(n, k, generator=gen, dtype=torch.float64)
    w = torch.randn(n, k, generator=gen, dtype=torch.float64)
    # FP64 CPU reference gradients
    e_r = e.clone().requires_grad_(True)
    x_r = x.clone().requires_grad_(True)
    (w * dense_fp64(e_r, 0.05, x_r)).sum().backward()
    # HIP backward
    e_g = e.float().to(cuda).requires_grad_(True)
    x_g = x.float().to(cuda).requires_grad_(True)
    op = laker.AttentionKernelOperator(e_g, lambda_reg=0.05)
    y = op.matvec(x_g)
    (w.float().to(cuda) * y).sum().backward()
    de = e_g.grad.double().cpu()
    dx = x_g.grad.double().cpu()
    assert float(torch.linalg.norm(de - e_r.grad) / torch.linalg.norm(e_r.grad)) < 1e-4
    assert float(torch.linalg.norm(dx - x_r.grad) / torch.linalg.norm(x_r.grad)) < 1e-4


def test_cross_backward_vs_fp64_autograd(cuda):
    gen = torch.Generator().manual_seed(5)
    e = torch.randn(200, 8, generator=gen, dtype=torch.float64) * 0.4
    q = torch.randn(37, 8, generator=gen, dtype=torch.float64) * 0.4
    a = torch.randn(200, generator=gen, dtype=torch.float64)
    w = torch.randn(37, generator=gen, dtype=torch.float64)
    er, qr, ar = (t.clone().requires_grad_(True) for t in (e, q, a))
    (w * (torch.exp(qr @ er.T) @ ar)).sum().backward()
    eg, qg, ag = (t.float().to(cuda).requires_grad_(True) for t in (e, q, a))
    op = laker.AttentionKernelOperator(eg, lambda_reg=0.1)
    yq = cross_kernel_matvec(op, qg, ag, embeddings=eg)
    (w.float().to(cuda) * yq).sum().backward()
    for got, ref in ((eg.grad, er.grad), (qg.grad, qr.grad), (ag.grad, ar.grad)):
        assert float(torch.linalg.norm(got.double().cpu() - ref) / torch.linalg.norm(ref)) < 1e-4


def test_backward_central_differences(cuda):
    gen = torch.Generator().manual_seed(9)
    e = torch.randn(120, 6, generator=gen, dtype=torch.float64) * 0.4
    x = torch.randn(120, generator=gen, dtype=torch.float64)
    w = torch.randn(120, generator=gen, dtype=torch.float64)
    direction = torch.randn(120, 6, generator=gen, dtype=torch.float64)
    h = 1e-3

    def loss(ee):
        return float((w * dense_fp64(ee, 0.05, x)).sum())

    fd = (loss(e + h * direction) - loss(e - h * direction)) / (2 * h)
    eg = e.float().to(cuda).requires_grad_(True)
    op = laker.AttentionKernelOperator(eg, lambda_reg=0.05)
    (w.float().to(cuda) * op.matvec(x.float().to(cuda))).sum().backward()
    an = float((eg.grad.double().cpu() * direction).sum())
    assert abs(an - fd) / abs(fd) < 1e-3


def test_validation_gradient_vs_finite_differences(cuda):
    """Validation-MSE gradient (direct + implicit adjoint term) vs FP64 central differences."""
    x_tr, y_tr = laker.pathloss_problem(300, seed=0)
    x_va, y_va = laker.pathloss_problem(80, seed=1)
    torch.manual_seed(0)
    mlp = torch.nn.Sequential(torch.nn.Linear(2, 16), torch.nn.Tanh(), torch.nn.Linear(16, 4))
    scale_in = 1.0 / 100.0

    class Emb(torch.nn.Module):
        def __init__(self, m):
            super().__init__()
            self.m = m

        def forward(self, x):
            return self.m(x * scale_in)

    emb = Emb(mlp).to(cuda)
    core = laker.core.LAKERCore(lambda_reg=1e-2, pcg_tol=1e-10, pcg_max_iter=2000, num_probes=60, verbose=False)
    from laker.training import EmbeddingTrainer

    trainer = EmbeddingTrainer(core)
    emb.zero_grad()
    trainer.validation_step(emb, x_tr, y_tr, x_va, y_va)
    params = list(emb.parameters())
    direction = [torch.randn_like(p) for p in params]
    an = sum(float((p.grad * dvec).sum()) for p, dvec in zip(params, direction))

    def loss64(shift):
        m64 = Emb(torch.nn.Sequential(torch.nn.Linear(2, 16), torch.nn.Tanh(), torch.nn.Linear(16, 4))).double()
        with torch.no_grad():
            for p64, p, dvec in zip(m64.parameters(), params, direction):
                p64.copy_(p.detach().double().cpu() + shift * dvec.double().cpu())
            et = m64(x_tr.double()).numpy()
            ev = m64(x_va.double()).numpy()
        alpha = O.dense_solve_fp64(et, 1e-2, y_tr.double().numpy())
        r = np.exp(ev @ et.T) @ alpha - y_va.double().numpy()
        return float(np.mean(r * r))

    h = 1e-3
    fd = (loss64(h) - loss64(-h)) / (2 * h)
    print(f"validation gradient: analytic {an:.6e} fd {fd:.6e}", flush=True)
    assert abs(an - fd) / abs(fd) < 2e-2


def test_fit_learned_embeddings_reduces_residual(cuda):
    x, y = laker.pathloss_problem(200, seed=3)
    model = laker.LAKERRegressor(embedding_dim=6, num_probes=40, cccp_max_iter=20, pcg_tol=1e-6, verbose=False)
    model.fit(x, y)
    res0 = float(torch.linalg.norm(model.kernel_operator.matvec(model.alpha) - y.to(cuda)))
    model.fit_learned_embeddings(x, y, lr=1e-3, epochs=5, rebuild_freq=2)
    res1 = float(torch.linalg.norm(model.kernel_operator.matvec(model.alpha) - y.to(cuda)))
    assert res1 <= 3 * res0 + 1e-6
