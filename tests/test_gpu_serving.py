"""Continuous-batching serving on the GPU: every request's solution is the same
alone and in the concurrent mix (bit-identical: fixed-order reductions, probes
seeded by the request id), and every solution meets the FP64 residual."""

import pytest
import torch

import laker
from laker.serving import SolveServer, make_request, poisson_arrivals

pytestmark = pytest.mark.gpu


def test_alone_equals_mix(cuda):
    arr = poisson_arrivals(6, rate=50.0, seed=1)
    reqs = [make_request(r, n_min=1000, n_max=6000, k_max=6, arrival=a) for r, a in enumerate(arr)]
    print("serving requests", [(r.n, r.k, r.embedding_dim) for r in reqs], flush=True)
    mix = SolveServer(slots=4, precision="mixed")
    res_mix = mix.run(reqs, realtime=True)
    alone = SolveServer(slots=1, precision="mixed")
    res_alone = alone.run(reqs, realtime=False)
    assert mix.max_inflight >= 2
    for a, b, r in zip(res_alone, res_mix, reqs):
        assert a.error is None and b.error is None, (a.error, b.error)
        assert torch.equal(a.alpha, b.alpha), f"request {r.rid} differs alone vs mixed"
        op = laker.AttentionKernelOperator(mix._embedding(r.embedding_dim)(r.x.to(cuda)).detach(), lambda_reg=1e-2)
        res = op.residual_fp64(r.y.to(cuda).double(), b.alpha)
        rel = float(torch.linalg.norm(res) / torch.linalg.norm(r.y.double()))
        assert rel <= 1.01e-6, (r.rid, rel)
    s = SolveServer.summary(res_mix, mix.elapsed_s)
    assert s["solves"] == len(reqs) and s["p99_ms"] >= s["p50_ms"] > 0


def test_hbm_admission_forced_cap(cuda):
    reqs = [make_request(r, n_min=1500, n_max=1500, k_max=2) for r in range(4)]
    from laker.serving import footprint_bytes

    cap = footprint_bytes(1500, 2, 16) + 1  # room for one request at a time
    srv = SolveServer(slots=4, hbm_budget_bytes=cap, precision="fp32")
    res = srv.run(reqs, realtime=False)
    assert srv.max_inflight == 1 and all(r.error is None for r in res)


def test_alone_equals_mix_stream_shape(cuda):
    """Requests at the serving stream's own shape: ids 0-9 and 29 of the stream
    (n 1,153-60,390 log-uniform, d in {10, 16}, 1-29 right-hand sides)
    arriving as a Poisson stream (4 requests/s, seed 0)."""
    rids = list(range(10)) + [29]
    arr = poisson_arrivals(len(rids), rate=4.0, seed=0)
    reqs = [make_request(r, arrival=a) for r, a in zip(rids, arr)]
    print("serving requests (rid, n, k, d):", [(r.rid, r.n, r.k, r.embedding_dim) for r in reqs], flush=True)
    assert max(r.n for r in reqs) > 50000 and max(r.k for r in reqs) >= 25 and {r.embedding_dim for r in reqs} == {10, 16}
    mix = SolveServer(slots=4, precision="mixed")
    res_mix = mix.run(reqs, realtime=True)
    alone = SolveServer(slots=1, precision="mixed")
    res_alone = alone.run(reqs, realtime=False)
    print("mix:", SolveServer.summary(res_mix, mix.elapsed_s), flush=True)
    assert mix.max_inflight >= 2
    for a, b, r in zip(res_alone, res_mix, reqs):
        assert a.rid == b.rid == r.rid
        assert a.error is None and b.error is None, (a.error, b.error)
        assert torch.equal(a.alpha, b.alpha), f"request {r.rid} differs alone vs mixed"
        op = laker.AttentionKernelOperator(mix._embedding(r.embedding_dim)(r.x.to(cuda)).detach(), lambda_reg=1e-2)
        res = op.residual_fp64(r.y.to(cuda).double(), b.alpha)
        rel = float(torch.linalg.norm(res) / torch.linalg.norm(r.y.double()))
        print(f"request {r.rid}: n={r.n} k={r.k} d={r.embedding_dim} FP64 residual {rel:.2e} "
              f"iterations {b.iterations} latency {b.latency_s * 1e3:.0f} ms", flush=True)
        assert rel <= 1e-6, (r.rid, rel)
        op.close()
