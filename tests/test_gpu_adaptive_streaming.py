"""Jacobi and adaptive preconditioners, power-iteration probes, stored CCCP
factors and the warm-started refits on the GPU, against the oracle."""

import numpy as np
import pytest
import torch

import laker
import laker_oracle as O
from laker.solvers import _native_preconditioner

pytestmark = pytest.mark.gpu


# ---------------------------------------------------------------- Jacobi
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_jacobi_matches_oracle_bitwise(cuda, dtype):
    e = torch.randn(500, 6, dtype=dtype, generator=torch.Generator().manual_seed(0)) * 0.5
    op = laker.AttentionKernelOperator(e.to(cuda), lambda_reg=0.3)
    d = op.diagonal()
    d[3] = 0.0  # exercises the eps clamp
    jac = laker.JacobiPreconditioner(d)
    inv_ref = O.jacobi_inverse(d.cpu())
    print(f"jacobi {dtype}: n=500", flush=True)
    assert torch.equal(jac.inv_diag.cpu(), inv_ref)
    x1 = torch.randn(500, dtype=dtype, device=cuda)
    x3 = torch.randn(500, 3, dtype=dtype, device=cuda)
    assert torch.equal(jac.apply(x1).cpu(), O.jacobi_apply(inv_ref, x1.cpu()))
    assert torch.equal(jac.apply(x3).cpu(), O.jacobi_apply(inv_ref, x3.cpu()))
    with pytest.raises(ValueError):
        jac.apply(torch.zeros(7, device=cuda, dtype=dtype))


def test_jacobi_pcg_native_matches_oracle(cuda):
    # reference tests/test_solvers.py:59-73 on the device loop
    e = torch.randn(30, 4, generator=torch.Generator().manual_seed(3))
    b = torch.randn(30, generator=torch.Generator().manual_seed(4))
    op = laker.AttentionKernelOperator(e.to(cuda), lambda_reg=1.0)
    jac = laker.JacobiPreconditioner(op.diagonal())
    assert _native_preconditioner(jac.apply)[0] is jac
    pcg = laker.PreconditionedConjugateGradient(tol=1e-8, max_iter=200, verbose=False)
    x = pcg.solve(op.matvec, jac.apply, b.to(cuda))
    res = torch.linalg.norm(op.matvec(x) - b.to(cuda)) / torch.linalg.norm(b)
    assert res.item() < 1e-5
    inv = O.jacobi_inverse(O.diagonal(e, 1.0))
    ref = O.pcg_solve(lambda v: O.attention_matvec(e, 1.0, v), lambda v: O.jacobi_apply(inv, v), b, tol=1e-8,
                      max_iter=200)
    print(f"jacobi PCG iterations {pcg.iterations} (oracle {ref.iterations})", flush=True)
    assert abs(pcg.iterations - ref.iterations) <= 1
    # the mixed solve drives the same handle and reaches the FP64 target
    mixed = laker.PreconditionedConjugateGradient(tol=1e-6, max_iter=200, verbose=False, precision="mixed")
    mixed.solve(op.matvec, jac.apply, b.to(cuda))
    a = O.dense_matrix_fp64(e.double().numpy(), 1.0)
    rel = np.linalg.norm(a @ mixed.solution_fp64.cpu().numpy() - b.double().numpy()) / np.linalg.norm(b.numpy())
    assert mixed.converged and rel < 1.01e-6


# ---------------------------------------------------------------- adaptive
@pytest.mark.parametrize("n,lam,with_diag,expect", [(300, 1.0, True, "jacobi"), (300, 1.0, False, "cccp"),
                                                    (1500, 1e-2, True, "cccp")])
def test_adaptive_selection_matches_oracle(cuda, n, lam, with_diag, expect):
    e = torch.randn(n, 4, dtype=torch.float64, generator=torch.Generator().manual_seed(0)) * 0.3
    op = laker.AttentionKernelOperator(e.to(cuda), lambda_reg=lam)
    pre = laker.AdaptivePreconditioner(num_probes=40, verbose=False, dtype=torch.float64)
    pre.build(op.matvec, n, diagonal=op.diagonal() if with_diag else None, seed=5)
    probes = O.reference_probes(n, 10, 5, torch.float64)
    ref = O.adaptive_select(lambda v: O.attention_matvec(e, lam, v), probes, with_diag)
    print(f"adaptive lam={lam}: kappa {pre.condition_estimate:.6e} (oracle {ref['cond']:.6e}) -> "
          f"{pre.inner_name}", flush=True)
    assert pre.inner_name == ref["choice"] == expect
    assert pre.condition_estimate == pytest.approx(ref["cond"], rel=1e-9)
    assert pre.lambda_max == pytest.approx(ref["lambda_max"], rel=1e-9)
    # the device PCG runs on the selected handle
    inner, ok = _native_preconditioner(pre.apply)
    assert ok and inner is pre.inner
    b = torch.randn(n, dtype=torch.float64, device=cuda)
    x = laker.PreconditionedConjugateGradient(tol=1e-10, max_iter=1000, verbose=False).solve(op.matvec, pre.apply, b)
    assert (torch.linalg.norm(op.matvec(x) - b) / torch.linalg.norm(b)).item() < 1e-9


def test_adaptive_aggressive_branch_generic_operator(cuda):
    # a non-positive probe Rayleigh quotient clamps lambda_min: CCCP with twice the probes
    n = 240
    d = torch.cat([torch.tensor([50.0]), torch.full((n - 1,), -1.0)]).double().to(cuda)

    def op(v):
        return d * v if v.dim() == 1 else d.unsqueeze(1) * v

    pre = laker.AdaptivePreconditioner(num_probes=60, verbose=False, dtype=torch.float64)
    pre.build(op, n, diagonal=d, seed=2)
    ref = O.adaptive_select(lambda v: (d.cpu() * v) if v.dim() == 1 else d.cpu().unsqueeze(1) * v,
                            O.reference_probes(n, 10, 2, torch.float64), True)
    assert pre.inner_name == ref["choice"] == "cccp_aggressive"
    assert pre.inner.nr == min(2 * 60, n)
    with pytest.raises(RuntimeError, match="not been built"):
        laker.AdaptivePreconditioner(verbose=False).apply(torch.zeros(3, device=cuda))


def test_regressor_adaptive_strategy(cuda):
    x, y = laker.pathloss_problem(2000, seed=0)
    m = laker.LAKERRegressor(verbose=False, preconditioner="adaptive").fit(x, y)
    assert m.preconditioner.inner_name in ("cccp", "cccp_aggressive")
    ref = laker.LAKERRegressor(verbose=False).fit(x, y)
    # the n = 2000 problem's conditioning selects CCCP on the same default probes: the same factors and solve
    assert m.preconditioner.inner.isotropic_coef == pytest.approx(ref.preconditioner.isotropic_coef, rel=1e-12)
    assert torch.equal(m.alpha.cpu(), ref.alpha.cpu())


# ---------------------------------------------------------------- power-iteration probes
def _gap_embedding():
    # reference tests/test_preconditioner.py:86-150: n=500, d_e=8, spectrum gap by SVD
    g = torch.Generator().manual_seed(0)
    u = torch.linalg.qr(torch.randn(500, 8, dtype=torch.float64, generator=g))[0]
    v = torch.linalg.qr(torch.randn(8, 8, dtype=torch.float64, generator=g))[0]
    s = torch.tensor([2.0, 1.8, 0.6, 0.5, 0.4, 0.3, 0.2, 0.1], dtype=torch.float64)
    return u @ torch.diag(s) @ v


def test_power_iter_probes_match_oracle(cuda):
    e = _gap_embedding()
    op = laker.AttentionKernelOperator(e.to(cuda), lambda_reg=1e-3)
    kw = dict(num_probes=50, gamma=1e-1, epsilon=1e-8, base_rho=0.05, max_iter=80, tol=1e-5, verbose=False,
              dtype=torch.float64, power_iter_steps=3)
    pre_p = laker.CCCPPreconditioner(probe_strategy="power_iter", **kw).build(op.matvec, 500, seed=42)
    pre_g = laker.CCCPPreconditioner(probe_strategy="gaussian", **kw).build(op.matvec, 500, seed=42)

    def opc(v):
        return O.attention_matvec(e, 1e-3, v)

    probes = O.power_iter_probes(opc, O.reference_probes(500, 50, 42, torch.float64), 3)
    fac = O.cccp_build(opc, 500, probes, 1e-1, 1e-8, 0.05, 80, 1e-5)
    print(f"power_iter: a {pre_p.isotropic_coef:.12e} (oracle {fac['isotropic_coef']:.12e})", flush=True)
    assert pre_p.isotropic_coef == pytest.approx(fac["isotropic_coef"], rel=1e-8)
    v = torch.randn(500, dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    np.testing.assert_allclose(pre_p.apply(v.to(cuda)).cpu().numpy(), O.cccp_apply(v, fac).numpy(),
                               rtol=1e-7, atol=1e-9)

    def lam_max(mv, steps=40):
        w = torch.randn(500, dtype=torch.float64, generator=torch.Generator().manual_seed(4)).to(cuda)
        w = w / torch.linalg.norm(w)
        for _ in range(steps):
            w = mv(w)
            w = w / torch.linalg.norm(w)
        return torch.dot(w, mv(w)).item()

    l0 = lam_max(op.matvec)
    lg = lam_max(lambda w: pre_g.apply(op.matvec(w)))
    lp = lam_max(lambda w: pre_p.apply(op.matvec(w)))
    print(f"lambda_max: A {l0:.4e}, P_gauss A {lg:.4e}, P_power A {lp:.4e}", flush=True)
    assert lg < l0 and lp < l0 and lp < 2.0 * lg


# ---------------------------------------------------------------- stored factors
def test_saved_factors_reload_without_rebuild(cuda, tmp_path):
    x, y = laker.pathloss_problem(2000, seed=0)
    m = laker.LAKERRegressor(verbose=False).fit(x, y)
    path = str(tmp_path / "model.pt")
    m.save(path)
    m2 = laker.LAKERRegressor.load(path)
    assert m2.preconditioner is not None and m2.preconditioner.isotropic_coef == m.preconditioner.isotropic_coef
    v = torch.randn(2000, 3, device=cuda)
    assert torch.equal(m2.preconditioner.apply(v), m.preconditioner.apply(v))
    xq = torch.rand(64, 2) * 100.0
    assert torch.equal(m2.predict(xq), m.predict(xq))
    var2 = m2.predict_variance(xq[:8])
    var1 = m.predict_variance(xq[:8])
    assert torch.allclose(var2, var1, rtol=1e-6, atol=0)
    # a file without the factor keys (the reference's format) still loads and builds on demand
    state = torch.load(path, weights_only=True)
    for k in [k for k in state if k.startswith("preconditioner_")]:
        del state[k]
    torch.save(state, path)
    m3 = laker.LAKERRegressor.load(path)
    assert m3.preconditioner is None
    assert m3.predict_variance(xq[:8]).shape == (8,) and m3.preconditioner is not None


# ---------------------------------------------------------------- warm-started refits
def _small_problem(n, seed=42):
    torch.manual_seed(seed)
    x = torch.rand(n, 2, dtype=torch.float64) * 100.0
    y = torch.randn(n, dtype=torch.float64)
    return x, y


_KW = dict(embedding_dim=10, lambda_reg=1e-2, gamma=1e-1, num_probes=50, cccp_max_iter=20, cccp_tol=1e-4,
           pcg_tol=1e-6, pcg_max_iter=500, dtype=torch.float64, verbose=False)


def test_partial_fit_matches_oracle(cuda):
    # reference tests/test_advanced_features.py:298-324
    x, y = _small_problem(100)
    m = laker.LAKERRegressor(**_KW).fit(x[:80], y[:80])
    old_n = m.embeddings.shape[0]
    e_old, a_old = m.embeddings.cpu(), m.alpha.cpu()
    m.partial_fit(x[80:90], y[80:90])
    assert m.embeddings.shape[0] == old_n + 10 and m.alpha.shape[0] == old_n + 10
    assert m.predict(torch.rand(10, 2, dtype=torch.float64) * 100.0).shape == (10,)
    e_new = m.embeddings[80:].cpu()
    ref = O.partial_fit(e_old, e_new, a_old, y[:80], y[80:90], 1e-2, O.reference_probes(90, 50, None, torch.float64),
                        gamma=1e-1, cccp_max_iter=20, cccp_tol=1e-4, pcg_tol=1e-6, pcg_max_iter=500)
    err = (torch.linalg.norm(m.alpha.cpu() - ref["alpha"]) / torch.linalg.norm(ref["alpha"])).item()
    print(f"partial_fit: iterations {m.pcg_iterations_} (oracle {ref['iterations']}), alpha rel diff {err:.2e}",
          flush=True)
    # the FP64 iterates agree closely (the factors differ at rounding level, amplified by the
    # conditioning); near the 1e-6 stopping threshold the counts may differ by a few steps
    assert abs(m.pcg_iterations_ - ref["iterations"]) <= 3
    assert err < (1e-6 if m.pcg_iterations_ == ref["iterations"] else 1e-3)
    with pytest.raises(RuntimeError, match="rebuild threshold"):
        m.partial_fit(x[90:100], y[90:100], rebuild_threshold=15)
    with pytest.raises(RuntimeError, match="not been fitted"):
        laker.LAKERRegressor(**_KW).partial_fit(x[:2], y[:2])


def test_fit_path_matches_oracle(cuda):
    # reference tests/test_advanced_features.py:133-159
    x, y = _small_problem(100)
    m = laker.LAKERRegressor(**_KW)
    path = m.fit_path(x, y, lambda_reg_grid=[1e-3, 1e-1, 1e-2])
    assert path["lambda_reg"] == [1e-1, 1e-2, 1e-3]
    assert len(path["alphas"]) == len(path["pcg_iters"]) == len(path["final_rel_res"]) == 3
    ref = O.fit_path(m.embeddings.cpu(), y, [1e-3, 1e-1, 1e-2], O.reference_probes(100, 50, None, torch.float64),
                     gamma=1e-1, cccp_max_iter=20, cccp_tol=1e-4, pcg_tol=1e-6, pcg_max_iter=500)
    for a, ra, it, rit, rr, rrr in zip(path["alphas"], ref["alphas"], path["pcg_iters"], ref["pcg_iters"],
                                      path["final_rel_res"], ref["final_rel_res"]):
        err = (torch.linalg.norm(a.cpu() - ra) / torch.linalg.norm(ra)).item()
        print(f"fit_path stage: iterations {it} (oracle {rit}), alpha rel diff {err:.2e}, rel_res {rr:.3e} "
              f"(oracle {rrr:.3e})", flush=True)
        # the last stage (lambda = 1e-3 on the first stage's factors) is ill-conditioned:
        # its count moves by a few percent with rounding-level differences
        assert abs(it - rit) <= max(3, 0.05 * rit) and err < (1e-6 if it == rit else 1e-3) and rr < 2e-6
    assert torch.equal(m.alpha, path["alphas"][-1])
    with pytest.raises(ValueError):
        m.fit_path(x, y, lambda_reg_grid=[])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fit_continuation(cuda, precision):
    # reference tests/test_math_correctness.py:175-225 and test_advanced_features.py:475-500
    x, y = _small_problem(60)
    kw = dict(_KW, embedding_dim=4, num_probes=30, cccp_max_iter=10, pcg_max_iter=200)
    if precision == "mixed":
        kw.update(dtype=torch.float32, precision="mixed")
    m = laker.LAKERRegressor(**kw)
    m.fit_continuation(x, y, lambda_max=1.0, lambda_min=1e-2, n_stages=5, reuse_precond=True)
    lambdas = m.path_["lambda_reg"]
    assert all(a > b for a, b in zip(lambdas, lambdas[1:])) and abs(lambdas[-1] - 1e-2) < 1e-10
    assert m.lambda_reg == 1e-2 and m.path_["pcg_iters"][0] < 50
    # the final stage meets its residual target in FP64 (mixed: its own stopping test)
    a = O.dense_matrix_fp64(m.embeddings.double().cpu().numpy(), 1e-2)
    sol = (m.alpha_fp64 if precision == "mixed" else m.alpha).double().cpu().numpy()
    rel = np.linalg.norm(a @ sol - y.numpy()) / np.linalg.norm(y.numpy())
    print(f"continuation {precision}: iterations {m.path_['pcg_iters']}, FP64 relative residual {rel:.2e}",
          flush=True)
    assert rel < (1.01e-6 if precision == "mixed" else 2e-6)
