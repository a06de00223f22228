"""Differentiable embedding training at full size (GPU).

n = 20,000 training + 5,000 validation locations drawn as in the n = 2000
path-loss problem (seed 0), a PositionEmbedding-style MLP 2 -> 64 -> 64 -> 16
(tanh, Xavier-uniform weights, zero biases, torch seed 0) evaluated in bf16,
an fp32 solver (exact kernel, CCCP, PCG tol 1e-6), and Adam steps at lr 1e-3
on the validation MSE of the PCG solution: direct term through the fused
cross product's backward, implicit term through the adjoint PCG solve and the
hand-written matvec backward (laker/training.py:93-259,
laker/implicit_diff.py:110-151 restated).  Gradient values against central
finite differences and FP64 autograd are pinned at reduced size in
test_gpu_backward.py; this test pins that the full-size loop runs on the
device path, stays finite and lowers the validation loss.

Where the loop stops: at lr 1e-3 the training itself grows the embedding
(max ||e||^2 from 1.1 to about 10.5 in 13 Adam steps, whatever lambda in
[1e-2, 5]); the kernel's largest entries grow as exp(max ||e||^2), u_fp32
||K|| passes lambda and the FP32 operator turns numerically indefinite.  At
lambda = 1e-2 the fp32 PCG raises the reference's breakdown error at step 9
(max ||e||^2 = 6.2); at lambda = 1.0 at step 13 (max ||e||^2 = 10.6), where a
solve that does not detect the breakdown instead runs out of iterations with
a meaningless solution (validation MSE of order 1e5).  The mixed solver does
not rescue the run there: its inner FP32 operator has the same error, its
refinement cannot contract, and the gradients of its non-converged solves
drive the embedding further out (max ||e||^2 = 17.8 by step 49, validation
MSE stuck at 1.0).  So the full-size test runs the fp32 solver at lambda =
1.0 up to 50 steps, stops at the first breakdown, and requires it to come
only once max ||e||^2 > 9.
"""

import math

import pytest
import torch

import laker
from laker.training import EmbeddingTrainer

pytestmark = pytest.mark.gpu

N_TRAIN, N_VAL, STEPS = 20_000, 5_000, 50


class _Mlp(torch.nn.Module):
    def __init__(self):
        super().__init__()
        self.net = torch.nn.Sequential(torch.nn.Linear(2, 64), torch.nn.Tanh(), torch.nn.Linear(64, 64),
                                       torch.nn.Tanh(), torch.nn.Linear(64, 16))
        for m in self.net:
            if isinstance(m, torch.nn.Linear):
                torch.nn.init.xavier_uniform_(m.weight)
                torch.nn.init.zeros_(m.bias)

    def forward(self, x):
        return self.net(x * (1.0 / 100.0))  # locations live in [0, 100]^2


def test_validation_training_full_size(cuda):
    x, y = laker.pathloss_problem(N_TRAIN + N_VAL, seed=0)
    x_tr, y_tr, x_va, y_va = x[:N_TRAIN], y[:N_TRAIN], x[N_TRAIN:], y[N_TRAIN:]
    torch.manual_seed(0)
    model = _Mlp().to(device=cuda, dtype=torch.bfloat16)
    core = laker.core.LAKERCore(lambda_reg=1.0, gamma=0.1, pcg_tol=1e-6, pcg_max_iter=1000, verbose=False,
                                embedding_dtype=torch.bfloat16)
    trainer = EmbeddingTrainer(core)
    opt = torch.optim.Adam(model.parameters(), lr=1e-3)
    losses, iters, norms = [], [], []
    precond, broke = None, None
    for step in range(STEPS):
        opt.zero_grad()
        with torch.no_grad():
            e = model(x_tr.to(cuda, torch.bfloat16)).float()
            norms.append(float((e * e).sum(1).max()))
        try:
            loss, precond, _, it = trainer.validation_step(model, x_tr, y_tr, x_va, y_va, precond, rebuild=True)
        except RuntimeError as exc:
            assert "PCG breakdown" in str(exc), exc
            broke = step
            break
        grads = [p.grad for p in model.parameters()]
        assert all(g is not None and bool(torch.isfinite(g).all()) for g in grads), f"step {step}: bad gradient"
        assert math.isfinite(loss), f"step {step}: loss {loss}"
        assert it < 1000, f"step {step}: PCG did not converge (max ||e||^2 {norms[-1]:.2f})"
        opt.step()
        losses.append(loss)
        iters.append(it)
    print(f"{len(losses)} steps, breakdown at step {broke}; validation MSE {losses[0]:.4f} -> {losses[-1]:.4f}; "
          f"PCG iterations {iters}; max ||e||^2 {[round(v, 2) for v in norms]}", flush=True)
    assert len(losses) >= 12
    if broke is not None:
        assert norms[broke] > 9.0, norms
    assert losses[-1] < 0.5 * losses[0]


def test_lambda_1e2_fp32_breakdown_and_mixed_solves(cuda):
    """The same training at the n = 2000 configuration's lambda = 1e-2.

    The embedding drifts (max ||e||^2 from 1.1 to above 5 in under ten Adam
    steps), u_fp32 ||K|| passes lambda and the plain fp32 PCG -- the
    reference's arithmetic and breakdown test (laker/solvers.py:243-246) --
    raises its breakdown error (measured: step 9, max ||e||^2 = 6.2).  Run by
    the mixed solver, the same steps' solves of A alpha = y reach an FP64
    relative residual <= 1e-6 while max ||e||^2 <= 5; beyond that (step 8,
    max ||e||^2 = 5.5, u_fp32 ||K|| ~ 6e-2 > lambda) the refinement cannot
    contract either, and the solver says so (converged False, residual kept
    at its best FP64 iterate).  Fifty steps at lambda = 1e-2 therefore need
    an FP64 inner operator."""
    x, y = laker.pathloss_problem(N_TRAIN + N_VAL, seed=0)
    x_tr, y_tr, x_va, y_va = x[:N_TRAIN], y[:N_TRAIN], x[N_TRAIN:], y[N_TRAIN:]
    y64 = y_tr.to(cuda, torch.float64)
    results = {}
    for precision in ("fp32", "mixed"):
        torch.manual_seed(0)
        model = _Mlp().to(device=cuda, dtype=torch.bfloat16)
        core = laker.core.LAKERCore(lambda_reg=1e-2, gamma=0.1, pcg_tol=1e-6, pcg_max_iter=1000, verbose=False,
                                    embedding_dtype=torch.bfloat16, precision=precision)
        trainer = EmbeddingTrainer(core)
        opt = torch.optim.Adam(model.parameters(), lr=1e-3)
        steps, broke, norms, relres, conv = 0, None, [], [], []
        stop = results.get("fp32", {}).get("broke", 15)
        for step in range(15 if precision == "fp32" else stop):
            opt.zero_grad()
            with torch.no_grad():
                e = model(x_tr.to(cuda, torch.bfloat16)).float()
                norms.append(float((e * e).sum(1).max()))
            try:
                loss, precond, alpha, it = trainer.validation_step(model, x_tr, y_tr, x_va, y_va, None,
                                                                   rebuild=True)
            except RuntimeError as exc:
                assert "PCG breakdown" in str(exc), exc
                broke = step
                break
            if precision == "mixed":
                # the solve of A alpha = y at this step, checked in FP64
                with torch.no_grad():
                    op = core.build_kernel_operator(e.contiguous())
                    pre = core.build_preconditioner(op.matvec, e.shape[0])
                    a, _ = core.solve_pcg(op, pre, y_tr.to(cuda))
                    sol = core.last_solver.solution_fp64
                    conv.append(core.last_solver.converged)
                    r = op.residual_fp64(y64, sol)
                    relres.append(float(torch.linalg.norm(r) / torch.linalg.norm(y64)))
                    op.close()
                    pre.close()
            opt.step()
            steps += 1
        results[precision] = {"broke": broke, "steps": steps, "max_sq_norms": norms, "relres": relres,
                              "converged": conv}
        print(f"lambda=1e-2 {precision}: {steps} steps, breakdown at step {broke}; max ||e||^2 "
              f"{[round(v, 2) for v in norms]}; FP64 residuals of the solves {['%.1e' % v for v in relres]}, "
              f"converged {conv}", flush=True)
    f = results["fp32"]
    assert f["broke"] is not None and f["max_sq_norms"][f["broke"]] > 5.0
    m = results["mixed"]
    assert m["broke"] is None and m["steps"] == f["broke"]
    for nrm, rel, ok in zip(m["max_sq_norms"], m["relres"], m["converged"]):
        # converged means the FP64 target was met; otherwise the solver reports it
        assert (ok and rel <= 1e-6) or not ok, (nrm, rel, ok)
        if nrm <= 5.0:
            assert ok and rel <= 1e-6, (nrm, rel)
