"""Continuous-batching solve server for streams of LAKER systems on one GPU.

Requests (locations, targets with k right-hand sides, embedding_dim) arrive
over time.  The server keeps ``slots`` build lanes, each a host worker bound
to its own HIP stream; a dispatcher admits a waiting request as soon as a lane
is free *and* the request's HBM footprint fits the remaining budget.  A lane
computes the request's embedding, packs its operator and builds its CCCP
preconditioner.  Then:

* small systems (n <= ``batch_max_n``, 2..32 right-hand sides, mixed
  precision) are handed to the batch: one device batch (``MixedBatch``, one
  HIP stream) runs the inner PCG iterations of every system in it as shared
  kernel launches -- a launch per kernel of the iteration for all of them --
  and takes each system's FP64 refinement steps between polls; systems join
  as soon as their build is done and a batch slot is free, and leave when
  they have converged, so the batch is continuously refilled;
* other systems solve on their lane (mixed- or FP32-precision PCG, HIP-graph
  replay of the iterations), overlapping with the batch on the device.

E, the factors and the Krylov state stay resident until the request
completes; its HBM footprint stays reserved until then.  Results are
deterministic per request (fixed-order reductions, probes seeded by the
request id, and a batched system's iterates do not depend on the other
systems in the batch), so a request's solution is bit-identical whether it is
solved alone or in the mix.  Latency = completion - arrival (queueing
included); the server reports solves/sec and p50/p99.

No reference counterpart: the reference solves one system per call
(laker/models.py:392-455).
"""

from __future__ import annotations

import collections
import dataclasses
import math
import threading
import time
from typing import Callable, List, Optional

import torch


@dataclasses.dataclass
class SolveRequest:
    rid: int
    x: torch.Tensor            # (n, 2) locations (host or device)
    y: torch.Tensor            # (n,) or (n, k) targets
    embedding_dim: int = 10
    arrival: float = 0.0       # seconds since the start of the stream
    lambda_reg: float = 1e-2

    @property
    def n(self) -> int:
        return self.x.shape[0]

    @property
    def k(self) -> int:
        return 1 if self.y.dim() == 1 else self.y.shape[1]


@dataclasses.dataclass
class SolveResult:
    rid: int
    alpha: Optional[torch.Tensor]
    n: int
    k: int
    latency_s: float
    queue_s: float
    service_s: float
    iterations: int = 0
    outer: int = 0
    relres: float = float("nan")
    slot: int = -1
    error: Optional[str] = None
    batched: bool = False


def footprint_bytes(n: int, k: int, d: int, nr: Optional[int] = None) -> int:
    """Peak device bytes of one request (build peak dominates)."""
    nr = nr if nr is not None else min(n, max(200, int(2 * math.sqrt(n))))
    ka32, ka16, ka16h = 2 * ((d + 1) // 2), 16 * ((6 * d + 15) // 16), 16 * ((3 * d + 15) // 16)
    npad = 64 * ((n + 63) // 64)
    packs = npad * (2 * ka32 * 4 + 2 * ka16 * 2 + 2 * ka16h * 2) + n * d * 12
    # Q, probe responses and FP64 Cholesky-QR work (n x nr each), the FP16 copy
    # of Q of the mixed k = 1 applies, and the nr x nr factors
    build = n * nr * (4 + 4 + 8 + 8 + 2) + nr * nr * 8 * 14
    solve = n * k * (6 * 4 + 3 * 8) + 64 * n * max(k, 4) * 4
    # partial sums of the symmetric k=1 product (column partials ~n^2/512
    # doubles, row partials per 1024-column segment) and the packed BF16
    # right-hand sides of the multi-RHS product (3 parts x 512 16-byte vectors
    # per 32-row tile)
    sym = (n * n // 512 + (n // 1024 + 1) * n) * 8
    xpack = ((n + 31) // 32) * 3 * 512 * 16
    return int(packs + build + solve + sym + xpack + (64 << 20))


def make_request(rid: int, seed_base: int = 0, n_min: int = 1000, n_max: int = 65536,
                 k_max: int = 32, arrival: float = 0.0) -> SolveRequest:
    """The serving-stream request ``rid``: n log-uniform in [n_min, n_max],
    d in {10, 16}, k in [1, k_max], inputs as in the n=2000 configuration,
    everything seeded by the request id."""
    from .data import pathloss_problem

    g = torch.Generator().manual_seed(seed_base * 1_000_003 + rid)
    u = torch.rand(3, generator=g, dtype=torch.float64)
    n = int(round(math.exp(math.log(n_min) + float(u[0]) * (math.log(n_max) - math.log(n_min)))))
    d = 10 if float(u[1]) < 0.5 else 16
    k = 1 + int(float(u[2]) * k_max)
    k = min(k, k_max)
    x, y = pathloss_problem(n, seed=rid, k=k)
    return SolveRequest(rid=rid, x=x, y=y, embedding_dim=d, arrival=arrival)


def poisson_arrivals(count: int, rate: float, seed: int = 0) -> List[float]:
    g = torch.Generator().manual_seed(seed)
    gaps = -torch.log(torch.rand(count, generator=g, dtype=torch.float64)) / rate
    return torch.cumsum(gaps, 0).tolist()


class SolveServer:
    """Continuous-batching server (see module docstring).

    Args:
        slots: build lanes (HIP streams) -- and solve lanes for systems
            outside the batch.
        hbm_budget_bytes: admission cap on the summed request footprints
            (default: 70% of the GPU's memory).
        precision: ``"mixed"`` or ``"fp32"``.
        tol / inner_tol: PCG tolerances.
        batch_max_n: largest n solved in the device batch (0 disables the
            batch; mixed precision only).
        batch_systems: systems the batch solves at once.
        solve_fn: injectable per-request work ``(req, slot) -> SolveResult``
            (host-side tests of the scheduler use a stand-in; no batch then).
    """

    def __init__(self, slots: int = 4, hbm_budget_bytes: Optional[int] = None, precision: str = "mixed",
                 tol: float = 1e-6, inner_tol: float = 5e-2, gamma: float = 0.1,
                 solve_fn: Optional[Callable] = None, footprint_fn: Callable = footprint_bytes,
                 batch_max_n: int = 8192, batch_systems: int = 64):
        self.slots = int(slots)
        self.precision = precision
        self.tol = tol
        self.inner_tol = inner_tol
        self.gamma = gamma
        self._solve_fn = solve_fn
        self._footprint = footprint_fn
        if hbm_budget_bytes is None:
            hbm_budget_bytes = int(0.7 * torch.cuda.get_device_properties(0).total_memory) \
                if torch.cuda.is_available() else 1 << 40
        self.budget = int(hbm_budget_bytes)
        self.batch_max_n = int(batch_max_n) if precision == "mixed" and solve_fn is None else 0
        self.batch_systems = int(batch_systems)
        self._emb = {}
        self._streams = None
        self._batch = None
        self.max_inflight = 0
        self.max_bytes_inflight = 0
        self.max_batched = 0

    # ------------------------------------------------------------------
    def _embedding(self, d: int):
        from .backend import compute_device
        from .embeddings import PositionEmbedding

        if d not in self._emb:
            self._emb[d] = PositionEmbedding(2, d, device=compute_device())
        return self._emb[d]

    def batchable(self, req: SolveRequest) -> bool:
        return self.batch_max_n > 0 and req.n <= self.batch_max_n and 2 <= req.k <= 32

    def build_one(self, req: SolveRequest):
        """Embedding, operator and CCCP preconditioner of a request, on the current stream."""
        from .backend import compute_device
        from .kernels import AttentionKernelOperator
        from .preconditioner import CCCPPreconditioner

        dev = compute_device()
        with torch.no_grad():
            e = self._embedding(req.embedding_dim)(req.x.to(dev)).contiguous()
        y = req.y.to(dev, torch.float32).contiguous()
        op = AttentionKernelOperator(e, lambda_reg=req.lambda_reg,
                                     precision="mixed" if self.precision == "mixed" else "fp32")
        pre = CCCPPreconditioner(gamma=self.gamma, verbose=False, probe_source="device",
                                 probe_precision="mixed" if self.precision == "mixed" else "fp32")
        pre.build(op.matvec, req.n, seed=req.rid)
        return op, pre, y

    def solve_one(self, req: SolveRequest, slot: int) -> SolveResult:
        """Device path of one request on the current stream (outside the batch)."""
        from .solvers import PreconditionedConjugateGradient

        op, pre, y = self.build_one(req)
        pcg = PreconditionedConjugateGradient(tol=self.tol, max_iter=1000, verbose=False,
                                              precision=self.precision, inner_tol=self.inner_tol)
        a = pcg.solve(op.matvec, pre.apply, y)
        alpha = pcg.solution_fp64 if pcg.solution_fp64 is not None else a
        relres = pcg.residual_norm / float(torch.linalg.norm(y.double())) if pcg.residual_norm < math.inf else 0.0
        op.close()
        pre.close()
        return SolveResult(rid=req.rid, alpha=alpha, n=req.n, k=req.k, latency_s=0.0, queue_s=0.0, service_s=0.0,
                           iterations=pcg.iterations, outer=pcg.outer_iterations, relres=relres, slot=slot)

    # ------------------------------------------------------------------
    def run(self, requests: List[SolveRequest], realtime: bool = True) -> List[SolveResult]:
        """Serve ``requests`` (sorted by arrival).  With ``realtime`` arrivals are
        honoured against the wall clock; otherwise every request is available
        at t=0 (a saturating stream)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        foot = {r.rid: self._footprint(r.n, r.k, r.embedding_dim) for r in reqs}
        for r in reqs:
            if foot[r.rid] > self.budget:
                raise ValueError(f"request {r.rid} needs {foot[r.rid]} bytes > HBM budget {self.budget}")
        use_gpu = self._solve_fn is None
        if use_gpu and self._streams is None:
            self._streams = [torch.cuda.Stream() for _ in range(self.slots)]
        use_batch = use_gpu and self.batch_max_n > 0 and any(self.batchable(r) for r in reqs)
        if use_batch and self._batch is None:
            from .batching import MixedBatch

            self._batch = MixedBatch(max_systems=self.batch_systems, tol=self.tol, inner_tol=self.inner_tol)
        lock = threading.Condition()
        free_slots = list(range(self.slots))
        inflight_bytes = [0]
        inflight = [0]
        results: List[SolveResult] = []
        handoff = collections.deque()   # built systems waiting for a batch slot
        remaining = [len(reqs)]
        t0 = time.perf_counter()

        def finish(res: SolveResult, req: SolveRequest, t_admit: float, lane: Optional[int] = None):
            t_done = time.perf_counter() - t0
            arrival = req.arrival if realtime else 0.0
            res.latency_s = t_done - arrival
            res.queue_s = t_admit - arrival
            res.service_s = t_done - t_admit
            with lock:
                if lane is not None:
                    free_slots.append(lane)
                results.append(res)
                inflight_bytes[0] -= foot[req.rid]
                inflight[0] -= 1
                remaining[0] -= 1
                lock.notify_all()

        def error_result(req, slot, exc):
            return SolveResult(rid=req.rid, alpha=None, n=req.n, k=req.k, latency_s=0, queue_s=0, service_s=0,
                               slot=slot, error=repr(exc))

        def worker(req: SolveRequest, slot: int, t_admit: float):
            res = None
            try:
                if use_gpu:
                    with torch.cuda.stream(self._streams[slot]):
                        if use_batch and self.batchable(req):
                            built = self.build_one(req)
                            torch.cuda.current_stream().synchronize()
                            with lock:
                                handoff.append((req, slot, t_admit, built))
                                free_slots.append(slot)
                                lock.notify_all()
                            return
                        res = self.solve_one(req, slot)
                        torch.cuda.current_stream().synchronize()
                else:
                    res = self._solve_fn(req, slot)
            except Exception as exc:  # reported per request, the server keeps going
                res = error_result(req, slot, exc)
            res.slot = slot
            finish(res, req, t_admit, lane=slot)

        def batch_loop():
            batch = self._batch
            waiting = {}
            while True:
                with lock:
                    while not handoff and not batch.entries and remaining[0] > 0:
                        lock.wait(timeout=0.05)
                    if remaining[0] == 0 and not batch.entries and not handoff:
                        return
                    joins = []
                    room = batch.free
                    while handoff and room > 0:
                        joins.append(handoff.popleft())
                        room -= 1
                for req, slot, t_admit, (op, pre, y) in joins:
                    try:
                        e = batch.add(op, pre, y if y.dim() == 2 else y.unsqueeze(1), key=req.rid)
                        waiting[req.rid] = (req, slot, t_admit, op, pre, e)
                    except Exception as exc:
                        op.close()
                        pre.close()
                        finish(error_result(req, slot, exc), req, t_admit)
                self.max_batched = max(self.max_batched, len(batch.entries))
                if not batch.entries:
                    continue
                try:
                    done = batch.step()
                except Exception as exc:  # a failed poll ends every system in the batch
                    done = list(batch.entries.values())
                    batch.entries.clear()
                    for e in done:
                        e.error = repr(exc)
                for e in done:
                    req, slot, t_admit, op, pre, _ = waiting.pop(e.key)
                    if e.error is not None:
                        res = error_result(req, slot, e.error)
                    else:
                        r = e.result
                        alpha = e.x if req.y.dim() == 2 else e.x.squeeze(1)
                        res = SolveResult(rid=req.rid, alpha=alpha, n=req.n, k=req.k, latency_s=0.0, queue_s=0.0,
                                          service_s=0.0, iterations=r.inner_iterations, outer=r.outer_iterations,
                                          relres=r.residual_norm / r.rhs_norm if r.rhs_norm else 0.0, slot=slot,
                                          batched=True)
                    op.close()
                    pre.close()
                    finish(res, req, t_admit)

        threads = []
        if use_batch:
            bth = threading.Thread(target=batch_loop, daemon=True)
            bth.start()
            threads.append(bth)
        i = 0
        with lock:
            while i < len(reqs):
                now = time.perf_counter() - t0
                r = reqs[i]
                if realtime and r.arrival > now:
                    lock.wait(timeout=r.arrival - now)
                    continue
                if free_slots and inflight_bytes[0] + foot[r.rid] <= self.budget:
                    slot = free_slots.pop(0)
                    inflight_bytes[0] += foot[r.rid]
                    inflight[0] += 1
                    self.max_inflight = max(self.max_inflight, inflight[0])
                    self.max_bytes_inflight = max(self.max_bytes_inflight, inflight_bytes[0])
                    th = threading.Thread(target=worker, args=(r, slot, now), daemon=True)
                    th.start()
                    threads.append(th)
                    i += 1
                else:
                    lock.wait(timeout=0.05)
        for th in threads:
            th.join()
        self.elapsed_s = time.perf_counter() - t0
        return sorted(results, key=lambda r: r.rid)

    @staticmethod
    def summary(results: List[SolveResult], elapsed_s: float) -> dict:
        lat = sorted(r.latency_s for r in results if r.error is None)
        if not lat:
            return {"solves": 0}

        def pct(p):
            idx = min(len(lat) - 1, max(0, int(math.ceil(p / 100.0 * len(lat))) - 1))
            return lat[idx]

        return {
            "solves": len(lat),
            "errors": sum(r.error is not None for r in results),
            "solves_per_sec": len(lat) / elapsed_s,
            "rhs_per_sec": sum(r.k for r in results if r.error is None) / elapsed_s,
            "p50_ms": 1000 * pct(50),
            "p99_ms": 1000 * pct(99),
            "elapsed_s": elapsed_s,
        }
