"""Batched mixed-precision solves (laker_batch_*) through the C ABI (GPU).

Systems of uneven n, k and embedding width share the batch's launches.  Each
system's FP64 solution must meet the FP64 residual target, in about the
iterations laker_solve_mixed takes for it (the batched product splits its
columns fewer ways, so the rounding differs), and must be bit-identical
whether the system is solved alone or in a mix that systems join and leave."""

import numpy as np
import pytest
import torch

import laker
import laker_oracle as O
from laker.batching import MixedBatch

pytestmark = pytest.mark.gpu

# (n, k, d, seed): n from the serving stream's small range, k across the three
# product column groups, both embedding widths
CASES = [(1291, 8, 10, 1), (2801, 11, 16, 9), (4873, 2, 10, 15), (1153, 25, 10, 3), (7356, 17, 16, 4),
         (3217, 32, 10, 7)]


def _system(cuda, n, k, d, seed):
    x, y = laker.pathloss_problem(n, seed=seed, k=k)
    emb = laker.PositionEmbedding(2, d, device=cuda)
    with torch.no_grad():
        e = emb(x.to(cuda)).contiguous()
    op = laker.AttentionKernelOperator(e, lambda_reg=1e-2, precision="mixed")
    pre = laker.CCCPPreconditioner(gamma=0.1, verbose=False, probe_source="device", probe_precision="mixed")
    pre.build(op.matvec, n, seed=seed)
    return op, pre, y.to(cuda).contiguous()


def _relres(op, y, sol):
    r = op.residual_fp64(y.double(), sol)
    return float(torch.linalg.norm(r) / torch.linalg.norm(y.double()))


def test_batch_matches_unbatched_solves(cuda):
    systems = [_system(cuda, *c) for c in CASES]
    ref = []
    for (op, pre, y), (n, k, d, _) in zip(systems, CASES):
        pcg = laker.PreconditionedConjugateGradient(tol=1e-6, max_iter=1000, verbose=False, precision="mixed")
        pcg.solve(op.matvec, pre.apply, y)
        ref.append((pcg.solution_fp64.clone(), pcg.iterations, pcg.outer_iterations))
    batch = MixedBatch(max_systems=8)
    for i, (op, pre, y) in enumerate(systems):
        batch.add(op, pre, y, key=i)
    done = {e.key: e for e in batch.run()}
    torch.cuda.synchronize()
    assert sorted(done) == list(range(len(CASES)))
    for i, ((op, pre, y), (n, k, d, _)) in enumerate(zip(systems, CASES)):
        e = done[i]
        assert e.error is None, e.error
        rel = _relres(op, y, e.x)
        print(f"n={n} k={k} d={d}: batched {e.result.inner_iterations} iterations / {e.result.outer_iterations} "
              f"steps, FP64 residual {rel:.2e}; alone {ref[i][1]} / {ref[i][2]}", flush=True)
        assert e.result.converged == 1 and rel <= 1e-6
        assert abs(e.result.inner_iterations - ref[i][1]) <= 0.1 * ref[i][1] + 5
        # same solution as the unbatched solve to the accuracy both reached
        a = O.dense_matrix_fp64(op.embeddings.double().cpu().numpy(), 1e-2)
        xb, xr = e.x.cpu().numpy(), ref[i][0].cpu().numpy()
        assert np.linalg.norm(a @ (xb - xr)) / np.linalg.norm(y.double().cpu().numpy()) <= 2e-6
    batch.close()


def test_batch_alone_equals_mix_with_staggered_joins(cuda):
    systems = [_system(cuda, *c) for c in CASES]
    alone = []
    for op, pre, y in systems:
        b = MixedBatch(max_systems=1)
        b.add(op, pre, y)
        alone.append(b.run()[0].x.clone())
        b.close()
    # mixed, joining at different times while others are mid-solve
    batch = MixedBatch(max_systems=4)
    pending = list(range(len(systems)))
    out = {}
    steps = 0
    while pending or batch.entries:
        while pending and batch.free > 0 and (steps % 3 == 0 or not batch.entries):
            i = pending.pop(0)
            op, pre, y = systems[i]
            batch.add(op, pre, y, key=i)
        for e in batch.step():
            out[e.key] = e
        steps += 1
    torch.cuda.synchronize()
    for i in range(len(systems)):
        assert out[i].error is None
        assert torch.equal(out[i].x, alone[i]), f"case {i} differs alone vs in the mix"
    batch.close()
