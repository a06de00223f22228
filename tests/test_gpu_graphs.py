"""HIP-graph replay of the device PCG iterations: identical results to eager
launches (fp32 and mixed, one and several right-hand sides), with many
replays per solve and eager work (status polls, FP64 residuals, new
refinement steps) between them."""

import numpy as np
import pytest
import torch

import laker
from laker import _native as N

pytestmark = pytest.mark.gpu


def _solve(e, y, precision, graphs):
    N.check(N.lib.laker_set_graphs(1 if graphs else 0))
    try:
        op = laker.AttentionKernelOperator(e, lambda_reg=1e-2, precision=precision)
        pre = laker.CCCPPreconditioner(gamma=0.1, verbose=False, probe_source="device",
                                       probe_precision="mixed" if precision == "mixed" else "fp32")
        pre.build(op.matvec, e.shape[0], seed=3)
        pcg = laker.PreconditionedConjugateGradient(tol=1e-6, max_iter=1000, verbose=False, precision=precision,
                                                    check_every=8)
        x = pcg.solve(op.matvec, pre.apply, y)
        sol = pcg.solution_fp64 if precision == "mixed" else x
        out = (sol.detach().cpu().clone(), pcg.iterations, pcg.outer_iterations)
        op.close()
        pre.close()
        return out
    finally:
        N.check(N.lib.laker_set_graphs(1))


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
@pytest.mark.parametrize("k", [1, 3, 6])
def test_graph_replay_matches_eager(cuda, precision, k):
    x, y = laker.pathloss_problem(3000, seed=11, k=k)
    emb = laker.PositionEmbedding(2, 10, device=cuda)
    with torch.no_grad():
        e = emb(x.to(cuda)).contiguous()
    y = y.to(cuda)
    print(f"graphs vs eager: precision={precision} k={k}", flush=True)
    xe, ite, oute = _solve(e, y, precision, graphs=False)
    xg, itg, outg = _solve(e, y, precision, graphs=True)
    assert ite == itg and oute == outg, (ite, itg, oute, outg)
    assert ite > 16  # several replays of the 8-iteration graph
    assert torch.equal(xe, xg)
    assert np.isfinite(xg.numpy()).all()
