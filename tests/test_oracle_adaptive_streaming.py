"""Pin the oracle's Jacobi / adaptive / power-iteration / warm-start restatements
with closed forms and known spectra (no GPU)."""

import math

import numpy as np
import pytest
import torch

import laker_oracle as O


def _diag_op(d):
    return lambda v: d * v if v.dim() == 1 else d.unsqueeze(1) * v


# ---------------------------------------------------------------- Jacobi
def test_jacobi_exact_on_diagonal_system():
    d = torch.linspace(0.5, 40.0, 50, dtype=torch.float64)
    inv = O.jacobi_inverse(d)
    b = torch.randn(50, dtype=torch.float64, generator=torch.Generator().manual_seed(0))
    res = O.pcg_solve(_diag_op(d), lambda v: O.jacobi_apply(inv, v), b, tol=1e-12, max_iter=50)
    # the exact preconditioner: one step, x = b / d
    assert res.iterations == 1
    np.testing.assert_allclose(res.x.numpy(), (b / d).numpy(), rtol=1e-14)
    # (n, k) input scales rows
    x2 = torch.randn(50, 3, dtype=torch.float64)
    np.testing.assert_allclose(O.jacobi_apply(inv, x2).numpy(), (x2 / d.unsqueeze(1)).numpy(), rtol=1e-15)


def test_jacobi_eps_clamp():
    for dt in (torch.float32, torch.float64):
        d = torch.tensor([0.0, -3.0, 2.0, 1e-30], dtype=dt)
        inv = O.jacobi_inverse(d)
        eps = torch.finfo(dt).eps
        assert inv[0].item() == pytest.approx(1.0 / eps) and inv[1].item() == pytest.approx(1.0 / eps)
        assert inv[2].item() == 0.5 and inv[3].item() == pytest.approx(1.0 / eps)


def test_jacobi_pcg_reference_case():
    # reference tests/test_solvers.py:59-73 (n=30, d_e=4, lambda=1: diagonally dominant)
    e = torch.randn(30, 4, generator=torch.Generator().manual_seed(3))
    b = torch.randn(30, generator=torch.Generator().manual_seed(4))
    inv = O.jacobi_inverse(O.diagonal(e, 1.0))
    op = lambda v: O.attention_matvec(e, 1.0, v)  # noqa: E731
    res = O.pcg_solve(op, lambda v: O.jacobi_apply(inv, v), b, tol=1e-8, max_iter=200)
    assert (torch.linalg.norm(op(res.x) - b) / torch.linalg.norm(b)).item() < 1e-5


# ---------------------------------------------------------------- adaptive selection
def test_adaptive_estimates_on_known_spectrum():
    n = 200
    d = torch.cat([torch.tensor([1000.0, 400.0], dtype=torch.float64), torch.linspace(1.0, 100.0, n - 2,
                                                                                       dtype=torch.float64)])
    probes = torch.randn(n, 10, dtype=torch.float64, generator=torch.Generator().manual_seed(5))
    r = O.adaptive_select(_diag_op(d), probes, have_diagonal=True)
    # 5 power steps from A r: the error in lambda_max decays like (400/1000)^12
    assert r["lambda_max"] == pytest.approx(1000.0, rel=1e-4)
    # every Rayleigh quotient lies inside the spectrum, near the mean eigenvalue
    assert d.min().item() <= r["lambda_min"] <= d.max().item()
    ray = torch.sum(probes * (d.unsqueeze(1) * probes), 0) / torch.sum(probes ** 2, 0)
    assert r["lambda_min"] == pytest.approx(ray.min().item(), rel=1e-14)
    assert r["cond"] == pytest.approx(r["lambda_max"] / r["lambda_min"], rel=1e-14)


def test_adaptive_selection_regimes():
    n = 120
    probes = torch.randn(n, 10, dtype=torch.float64, generator=torch.Generator().manual_seed(6))
    well = torch.linspace(1.0, 50.0, n, dtype=torch.float64)
    assert O.adaptive_select(_diag_op(well), probes, True)["choice"] == "jacobi"
    # no diagonal: Jacobi is never chosen
    assert O.adaptive_select(_diag_op(well), probes, False)["choice"] == "cccp"
    # kappa between 1e3 and 1e6: the lambda_min estimate is a Rayleigh quotient,
    # so the ratio is lambda_max over roughly the mean eigenvalue
    mid = torch.cat([torch.tensor([2e5], dtype=torch.float64), torch.full((n - 1,), 1.0, dtype=torch.float64)])
    r = O.adaptive_select(_diag_op(mid), probes, True)
    assert 1e3 <= r["cond"] < 1e6 and r["choice"] == "cccp"
    # a non-positive Rayleigh quotient clamps lambda_min at 1e-12: aggressive CCCP
    ind = torch.cat([torch.tensor([50.0], dtype=torch.float64), torch.full((n - 1,), -1.0, dtype=torch.float64)])
    r = O.adaptive_select(_diag_op(ind), probes, True)
    assert r["lambda_min"] < 0 and r["cond"] >= 1e6 and r["choice"] == "cccp_aggressive"


# ---------------------------------------------------------------- power-iteration probes
def test_power_iter_probes_span_and_untouched_columns():
    e = torch.randn(200, 5, dtype=torch.float64, generator=torch.Generator().manual_seed(7)) * 0.4
    op = lambda v: O.attention_matvec(e, 1e-2, v)  # noqa: E731
    probes = torch.randn(200, 40, dtype=torch.float64, generator=torch.Generator().manual_seed(8))
    out = O.power_iter_probes(op, probes, steps=3)
    k = max(1, int(40 * 0.25))
    q = out[:, :k]
    np.testing.assert_allclose((q.T @ q).numpy(), np.eye(k), atol=1e-12)
    # span(Q) = span(A^3 R_block): the projector onto Q reproduces A^3 R_block
    a3r = op(op(op(probes[:, :k])))
    a3r = a3r / torch.linalg.norm(a3r, dim=0, keepdim=True)
    resid = a3r - q @ (q.T @ a3r)
    assert torch.linalg.norm(resid).item() < 1e-8
    assert torch.equal(out[:, k:], probes[:, k:])
    # one probe column: n_power = max(1, int(0.25)) = 1
    assert O.power_iter_probes(op, probes[:, :1], steps=1).shape == (200, 1)


# ---------------------------------------------------------------- warm-started refits
def test_continuation_schedule_closed_form():
    s = O.continuation_schedule(1.0, 1e-2, 5)
    np.testing.assert_allclose(s, [1.0, 10 ** -0.5, 0.1, 10 ** -1.5, 1e-2], rtol=1e-14)
    assert s[-1] == 1e-2 and all(a > b for a, b in zip(s, s[1:]))
    assert O.continuation_schedule(1.0, 1e-2, 1) == [1e-2]
    with pytest.raises(ValueError):
        O.continuation_schedule(1.0, 1e-2, 0)
    with pytest.raises(ValueError):
        O.continuation_schedule(-1.0, 1e-2, 3)


def test_fit_path_each_stage_solves_its_system():
    e = torch.randn(120, 4, dtype=torch.float64, generator=torch.Generator().manual_seed(9)) * 0.5
    y = torch.randn(120, dtype=torch.float64, generator=torch.Generator().manual_seed(10))
    probes = O.reference_probes(120, 60, None, torch.float64)
    path = O.fit_path(e, y, [1e-3, 1e-1, 1e-2], probes, reuse_precond=True, pcg_tol=1e-10, cccp_max_iter=20)
    assert path["lambda_reg"] == [1e-1, 1e-2, 1e-3]
    for lam, a, rr in zip(path["lambda_reg"], path["alphas"], path["final_rel_res"]):
        ref = O.dense_solve_fp64(e.numpy(), lam, y.numpy())
        assert np.linalg.norm(a.numpy() - ref) / np.linalg.norm(ref) < 1e-6
        assert rr < 1e-9


def test_partial_fit_solves_extended_system_and_warm_start_identity():
    e = torch.randn(90, 4, dtype=torch.float64, generator=torch.Generator().manual_seed(11)) * 0.5
    y = torch.randn(90, dtype=torch.float64, generator=torch.Generator().manual_seed(12))
    lam = 1e-2
    alpha0 = torch.from_numpy(O.dense_solve_fp64(e[:80].numpy(), lam, y[:80].numpy()))
    probes = O.reference_probes(90, 50, None, torch.float64)
    out = O.partial_fit(e[:80], e[80:], alpha0, y[:80], y[80:], lam, probes, pcg_tol=1e-10, cccp_max_iter=20)
    ref = O.dense_solve_fp64(e.numpy(), lam, y.numpy())
    assert np.linalg.norm(out["alpha"].numpy() - ref) / np.linalg.norm(ref) < 1e-6
    # nothing appended: the warm start is the solution already (one step at most)
    none = O.partial_fit(e[:80], e[80:80], alpha0, y[:80], y[80:80], lam, O.reference_probes(80, 50, None,
                                                                                              torch.float64),
                         pcg_tol=1e-6, cccp_max_iter=20)
    assert none["iterations"] <= 1
    assert math.isfinite(float(none["alpha"].sum()))
