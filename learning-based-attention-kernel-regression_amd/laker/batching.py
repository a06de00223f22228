"""Batched mixed-precision solves of many small systems (``laker_batch_*``).

``MixedBatch`` holds a device batch: systems (an operator, its CCCP
preconditioner and 2..32 right-hand sides each, uneven n) join with ``add``,
``step`` advances every running system by a few inner PCG iterations in
shared kernel launches and takes the FP64 refinement steps of the systems
whose step ended, and finished systems come back from ``step`` with their
FP64 solutions.  A system's iterates do not depend on which other systems
share the batch, so its solution is the same solved alone or in a mix.
No reference counterpart: the reference solves one system per call
(laker/models.py:392-455).
"""

from __future__ import annotations

import ctypes
import dataclasses
from typing import Any, Dict, List, Optional

import torch

from . import _native as N


@dataclasses.dataclass
class BatchEntry:
    slot: int
    key: Any
    op: Any
    pre: Any
    b: torch.Tensor       # (n, k) FP64 right-hand sides, device
    x: torch.Tensor       # (n, k) FP64 solution, device
    result: Optional[N.MixedResult] = None
    error: Optional[str] = None


class MixedBatch:
    """A device batch of mixed-precision solves on one HIP stream.

    Args:
        max_systems: systems that can solve at once.
        stream: torch CUDA stream the batch's work is ordered on (default: a new one).
        tol / inner_tol / max_outer / inner_max_iter: as ``laker_solve_mixed``.
    """

    def __init__(self, max_systems: int = 32, stream: Optional[torch.cuda.Stream] = None, tol: float = 1e-6,
                 inner_tol: float = 5e-2, max_outer: int = 20, inner_max_iter: int = 1000,
                 iterations_per_step: int = 8):
        self.stream = stream if stream is not None else torch.cuda.Stream()
        self.tol, self.inner_tol = float(tol), float(inner_tol)
        self.max_outer, self.inner_max_iter = int(max_outer), int(inner_max_iter)
        self.iterations_per_step = int(iterations_per_step)
        self.max_systems = int(max_systems)
        self.h = ctypes.c_void_p()
        N.check(N.lib.laker_batch_create(self.max_systems, ctypes.c_void_p(self.stream.cuda_stream),
                                         ctypes.byref(self.h)))
        self.entries: Dict[int, BatchEntry] = {}
        self._finished = (ctypes.c_int32 * self.max_systems)()

    @staticmethod
    def eligible(op, pre, k: int) -> bool:
        """Systems the batch takes: FP32 operator, FP32 CCCP handle with N_r % 4 == 0, 2 <= k <= 32."""
        return (op.dtype == torch.float32 and 2 <= k <= 32 and getattr(pre, "nr", 0) % 4 == 0
                and getattr(pre, "dtype", None) == torch.float32)

    @property
    def free(self) -> int:
        return self.max_systems - len(self.entries)

    def add(self, op, pre, y: torch.Tensor, key: Any = None) -> BatchEntry:
        """Joins the system (lambda I + K) x = y; its set-up runs on the batch stream now."""
        with torch.cuda.stream(self.stream):
            b = y.to(device=op.embeddings.device, dtype=torch.float64).contiguous()
            if b.dim() == 1:
                b = b.unsqueeze(1)
            x = torch.empty_like(b)
        slot = ctypes.c_int32()
        N.check(N.lib.laker_batch_add(self.h, op.handle, pre.handle, N.ptr(b), b.shape[1], N.ptr(x), self.tol,
                                      self.max_outer, self.inner_tol, self.inner_max_iter, ctypes.byref(slot)))
        e = BatchEntry(slot.value, key, op, pre, b, x)
        self.entries[e.slot] = e
        return e

    def step(self) -> List[BatchEntry]:
        """One poll; returns the entries that finished (result or error set)."""
        nf, na = ctypes.c_int32(), ctypes.c_int32()
        N.check(N.lib.laker_batch_step(self.h, self.iterations_per_step, self._finished, ctypes.byref(nf),
                                       ctypes.byref(na)))
        done = []
        for i in range(nf.value):
            slot = self._finished[i]
            e = self.entries.pop(slot)
            res = N.MixedResult()
            rc = N.lib.laker_batch_result(self.h, slot, ctypes.byref(res))
            e.result = res
            if rc != N.LAKER_OK:
                e.error = N.last_error()
            done.append(e)
        return done

    def run(self) -> List[BatchEntry]:
        """Steps until every system has finished."""
        out = []
        while self.entries:
            out.extend(self.step())
        return out

    def close(self) -> None:
        if self.h:
            N.lib.laker_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
