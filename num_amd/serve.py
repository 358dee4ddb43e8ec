"""Serving CG solves with continuous batching on one GPU.

``CGServer`` admits solve requests into a running ``CGEngine`` batch:

* requests wait in a FIFO queue from arrival until their device memory
  (matrix, FP32 copy, b, x and workspace) fits the admission budget and a
  batch slot is free;
* admitted requests join the running batch at the next engine step, so their
  iterations overlap with everything already on the device; matrices and
  Krylov vectors stay resident until the solve finishes;
* a finished request leaves the batch, returns its memory to the budget and
  records its latency (completion - arrival).

Each request's arithmetic is independent of what else is in the batch (the
engine reduces every system over its own partials in its own order), so a
result is bit-identical alone and in the mix.  ``run_open_loop`` replays a
timed arrival trace (open loop: arrivals never wait for completions) and
reports p50 / p99 latency and solves/s.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import kokkos_spmv as ks
from .batched import CGEngine


@dataclass
class Request:
    rid: int
    t_arrival: float          # seconds from the start of the trace
    dims: tuple
    shift: float
    seed: int
    tol: float = 1e-10
    t_admit: float = math.nan
    t_done: float = math.nan
    n: int = 0
    nnz: int = 0
    result: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


def request_bytes(n: int, nnz: int, mixed: bool) -> int:
    """Device bytes a request holds while admitted: CSR (FP64 values, int32
    columns and rowptr, plus the FP32 copy in mixed mode), b, x, the engine
    workspace and the row-tile schedule."""
    csr = 12 * nnz + 4 * (n + 1) + (4 * nnz if mixed else 0)
    vecs = 16 * n
    work = CGEngine.work_bytes(n, mixed)
    sched = 16 * (nnz // 1024 + 2)
    return csr + vecs + work + sched


class CGServer:
    def __init__(self, budget_bytes: float = 200e9, max_systems: int = 4096,
                 mixed: bool = True, iters_per_step: int = 8, max_iter: int = 20000,
                 delta: float = 0.1, nt64: bool = True, nt32: bool = False):
        self.budget = float(budget_bytes)
        self.in_use = 0.0
        self.peak_in_use = 0.0
        self.mixed = bool(mixed)
        self.iters_per_step = int(iters_per_step)
        self.max_iter = int(max_iter)
        self.delta = float(delta)
        self.engine = CGEngine(max_systems=max_systems, nt64=nt64, nt32=nt32)
        self.max_systems = int(max_systems)
        self.queue: collections.deque = collections.deque()
        self.running: dict = {}    # slot -> (request, bytes, keep-alive tensors)
        self.done: list = []
        self.clock0 = None

    def now(self) -> float:
        return time.perf_counter() - self.clock0

    def submit(self, req: Request) -> None:
        self.queue.append(req)

    def _admit(self) -> int:
        admitted = 0
        while self.queue and len(self.running) < self.max_systems:
            req = self.queue[0]
            n = int(np.prod(req.dims))
            nnz = n * (2 * len(req.dims) + 1)  # upper bound before assembly
            need = request_bytes(n, nnz, self.mixed)
            if self.in_use + need > self.budget and self.running:
                break  # wait for memory to free up
            self.queue.popleft()
            A = ks.upload_stencil(req.dims, req.shift)
            gen = torch.Generator(device="cuda")
            gen.manual_seed(int(req.seed))
            b = torch.rand(A.nrows, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
            x = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
            slot = self.engine.add(A, b, x, tol=req.tol, max_iter=self.max_iter,
                                   mixed=self.mixed, delta=self.delta)
            req.n, req.nnz = A.nrows, A.nnz
            req.t_admit = self.now()
            self.in_use += need
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            self.running[slot] = (req, need, (A, b, x))
            admitted += 1
        return admitted

    def _collect(self) -> int:
        finished = 0
        for slot in self.engine.finished():
            req, need, keep = self.running[slot]
            res = self.engine.result(slot)
            req.result = res
            req.t_done = self.now()
            req.solution = keep[2]
            self.engine.remove(slot)
            del self.running[slot]
            self.in_use -= need
            self.done.append(req)
            finished += 1
        return finished

    def pump(self) -> None:
        """Admit what fits, run one engine step, retire finished solves."""
        self._admit()
        if self.running:
            self.engine.step(self.iters_per_step)
            self._collect()

    def run_open_loop(self, requests, keep_solutions: bool = False,
                      max_wall: float | None = None) -> dict:
        """Replay ``requests`` (sorted by arrival time) open-loop; returns the
        latency / throughput summary.  Requests not finished when
        ``max_wall`` seconds have passed are reported as unfinished."""
        pending = collections.deque(sorted(requests, key=lambda r: r.t_arrival))
        self.clock0 = time.perf_counter()
        while pending or self.queue or self.running:
            t = self.now()
            if max_wall is not None and t > max_wall:
                break
            while pending and pending[0].t_arrival <= t:
                self.submit(pending.popleft())
            if not self.queue and not self.running:
                if pending:
                    time.sleep(max(0.0, min(pending[0].t_arrival - t, 0.01)))
                continue
            self.pump()
        wall = self.now()
        if not keep_solutions:
            for r in self.done:
                r.__dict__.pop("solution", None)
        return summarize(self.done, wall, len(requests))


def summarize(done, wall: float, submitted: int) -> dict:
    lat = np.array([r.latency for r in done]) if done else np.zeros(0)
    ok = sum(1 for r in done if r.result.get("converged"))
    t_first = min((r.t_arrival for r in done), default=0.0)
    t_last = max((r.t_done for r in done), default=0.0)
    return {
        "submitted": submitted, "completed": len(done), "converged": ok,
        "wall_s": wall,
        "solves_per_s": len(done) / max(t_last - t_first, 1e-9) if done else 0.0,
        "p50_ms": float(np.percentile(lat, 50) * 1e3) if done else None,
        "p99_ms": float(np.percentile(lat, 99) * 1e3) if done else None,
        "max_residual": max((r.result.get("residual", 0.0) for r in done), default=None),
        "mean_iters": float(np.mean([r.result.get("iters", 0) for r in done])) if done else None,
    }


def requests_from_trace(trace, tol: float = 1e-10):
    """generators.served_requests(...) dicts -> Request objects."""
    return [Request(i, q["t"], tuple(q["dims"]), q["shift"], q["seed"], tol)
            for i, q in enumerate(trace)]


def one_at_a_time(requests, mixed: bool = False) -> dict:
    """Baseline: each request built and solved by itself with plain FP64
    cg_solve_local (or cg_solve_mixed), one after another."""
    ks.init(0)
    t0 = time.perf_counter()
    iters = []
    for req in requests:
        A = ks.upload_stencil(req.dims, req.shift)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(int(req.seed))
        b = torch.rand(A.nrows, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1
        x = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
        if mixed:
            r = ks.cg_solve_mixed(A, b, x, tol=req.tol)
        else:
            r = ks.cg_solve_local(A, b, x, tol=req.tol, max_iter=20000)
        iters.append(r["iters"])
    el = time.perf_counter() - t0
    return {"requests": len(requests), "seconds": el, "solves_per_s": len(requests) / el,
            "mean_iters": float(np.mean(iters)) if iters else None}
