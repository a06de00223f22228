"""Serving scheduler logic on the host (no GPU): slot admission, HBM-footprint
admission control, arrival ordering, failure isolation, latency statistics."""

import threading
import time

import pytest
import torch

from laker.serving import SolveRequest, SolveResult, SolveServer, footprint_bytes, make_request, poisson_arrivals


def fake_request(rid, n=100, k=1, arrival=0.0):
    return SolveRequest(rid=rid, x=torch.zeros(n, 2), y=torch.zeros(n, k), arrival=arrival)


def make_server(slots, budget, delay=0.02, fail_ids=()):
    live = {"now": 0, "max": 0}
    lock = threading.Lock()

    def solve(req, slot):
        with lock:
            live["now"] += 1
            live["max"] = max(live["max"], live["now"])
        time.sleep(delay)
        with lock:
            live["now"] -= 1
        if req.rid in fail_ids:
            raise RuntimeError("forced failure")
        return SolveResult(rid=req.rid, alpha=None, n=req.n, k=req.k, latency_s=0, queue_s=0, service_s=0)

    srv = SolveServer(slots=slots, hbm_budget_bytes=budget, solve_fn=solve,
                      footprint_fn=lambda n, k, d: n * 1000)
    return srv, live


def test_slots_bound_concurrency():
    srv, live = make_server(slots=3, budget=10**12)
    res = srv.run([fake_request(i) for i in range(12)], realtime=False)
    assert [r.rid for r in res] == list(range(12))
    assert live["max"] == 3 and srv.max_inflight == 3


def test_hbm_budget_limits_admission():
    # each request costs 100k "bytes"; the budget admits two at a time
    srv, live = make_server(slots=8, budget=250_000)
    srv.run([fake_request(i, n=100) for i in range(10)], realtime=False)
    assert live["max"] == 2 and srv.max_bytes_inflight <= 250_000


def test_oversized_request_is_refused():
    srv, _ = make_server(slots=2, budget=50_000)
    with pytest.raises(ValueError, match="HBM budget"):
        srv.run([fake_request(0, n=100)], realtime=False)


def test_failure_is_isolated_and_arrivals_are_honoured():
    srv, _ = make_server(slots=2, budget=10**12, delay=0.01, fail_ids=(3,))
    arr = [0.0, 0.0, 0.05, 0.05, 0.1]
    res = srv.run([fake_request(i, arrival=a) for i, a in enumerate(arr)], realtime=True)
    assert res[3].error is not None and all(r.error is None for i, r in enumerate(res) if i != 3)
    assert all(r.queue_s >= -1e-3 for r in res)
    assert res[4].latency_s >= res[4].service_s - 1e-6
    s = SolveServer.summary(res, srv.elapsed_s)
    assert s["solves"] == 4 and s["errors"] == 1 and s["p99_ms"] >= s["p50_ms"]


def test_request_stream_definition():
    r = make_request(5)
    r2 = make_request(5)
    assert torch.equal(r.x, r2.x) and torch.equal(r.y, r2.y)
    assert 1000 <= r.n <= 65536 and r.embedding_dim in (10, 16) and 1 <= r.k <= 32
    arr = poisson_arrivals(100, rate=10.0)
    assert all(b >= a for a, b in zip(arr, arr[1:]))
    assert footprint_bytes(65536, 32, 16) < 2 << 30
