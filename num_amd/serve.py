"""Serving CG solves with continuous batching on one GPU.

``CGServer`` admits solve requests into a running ``CGEngine`` batch:

* requests wait in a FIFO queue from arrival until their device memory
  (matrix, FP32 copy, b, x and workspace, one buffer) fits the admission
  budget and a batch slot is free;
* an admitted request's system is assembled on the device (no host work
  proportional to its size, no synchronisation) and joins the running batch
  at the next engine step, so its iterations overlap with everything already
  on the device; matrices and Krylov vectors stay resident until the solve
  finishes;
* steps are pipelined: the host admits and retires requests while the
  device runs the previous step;
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
import sys
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import kokkos_spmv as ks
from .batched import CGEngine
from .generators import stencil_nnz


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


def request_bytes(dims, mixed: bool) -> int:
    """Device bytes a request holds while admitted: its assembled stencil
    (rowptr, int32 columns, FP64 values, the FP32 copy in mixed mode), b, x
    and the engine workspace -- one buffer."""
    return CGEngine.stencil_bytes(dims, mixed)


class CGServer:
    """Continuous-batching CG server on one GPU (see the module docstring).

    ``timeout_s``: a request still running that long after admission is
    cancelled (its slot and memory are freed) and reported with
    ``result["timeout"] = True``.  A request larger than the whole budget is
    rejected at once (``result["rejected"] = True``).  ``delta``: refine once
    the FP32 residual has dropped by this factor (0.003 measured best for
    the served stream: 3.0k solves/s at lambda = 10k against 2.7k with 0.1,
    as each refinement costs an FP64 iteration of its system); ``nt32``:
    stream the FP32 matrix copies non-temporally (measured slower on the
    served stream: 2.68k against 2.74k solves/s).  ``ka_chunk``,
    ``iters_per_step`` and ``nt32`` are what autotune.tune_engine picks."""

    def __init__(self, budget_bytes: float = 200e9, max_systems: int = 4096,
                 mixed: bool = True, iters_per_step: int = 8, max_iter: int = 20000,
                 delta: float = 0.003, nt64: bool = True, nt32: bool = False,
                 timeout_s: float | None = None, admit_per_pump: int = 32,
                 ka_chunk: int = 0):
        self.budget = float(budget_bytes)
        self.in_use = 0.0
        self.peak_in_use = 0.0
        self.mixed = bool(mixed)
        self.iters_per_step = int(iters_per_step)
        self.max_iter = int(max_iter)
        self.delta = float(delta)
        self.timeout_s = timeout_s
        # admissions between two launches: bounded so a burst of arrivals
        # never leaves the device without queued work for long
        self.admit_per_pump = int(admit_per_pump)
        self.engine = CGEngine(max_systems=max_systems, nt64=nt64, nt32=nt32)
        self.engine.set_schedule(ka_chunk)
        self.max_systems = int(max_systems)
        self.queue: collections.deque = collections.deque()
        self.running: dict = {}    # slot -> (request, bytes)
        self.done: list = []
        self.keep_solutions = False
        self.clock0 = time.perf_counter()

    def now(self) -> float:
        return time.perf_counter() - self.clock0

    def submit(self, req: Request) -> None:
        self.queue.append(req)

    def _finish(self, req: Request, result: dict) -> None:
        req.result = result
        req.t_done = self.now()
        self.done.append(req)

    def _admit(self) -> int:
        admitted = 0
        while (self.queue and len(self.running) < self.max_systems
               and admitted < self.admit_per_pump):
            req = self.queue[0]
            need = request_bytes(req.dims, self.mixed)
            if need > self.budget:
                self.queue.popleft()
                self._finish(req, {"rejected": True, "converged": False})
                continue
            if self.in_use + need > self.budget:
                break  # wait for memory to free up (FIFO: no overtaking)
            self.queue.popleft()
            slot = self.engine.add_stencil(req.dims, req.shift, req.seed, tol=req.tol,
                                           max_iter=self.max_iter, mixed=self.mixed,
                                           delta=self.delta, nbytes=need)
            req.n = math.prod(req.dims)
            req.nnz = stencil_nnz(req.dims)
            req.t_admit = self.now()
            self.in_use += need
            self.peak_in_use = max(self.peak_in_use, self.in_use)
            self.running[slot] = (req, need)
            admitted += 1
        return admitted

    def _release(self, slot: int) -> Request:
        req, need = self.running.pop(slot)
        if self.keep_solutions:
            req.solution = self.engine.solution(slot)
        self.engine.remove(slot)
        self.in_use -= need
        return req

    def _collect(self) -> int:
        finished = 0
        for slot in self.engine.finished():
            res = self.engine.result(slot)
            req = self._release(slot)
            self._finish(req, res)
            finished += 1
        if self.timeout_s is not None:
            t = self.now()
            for slot in [s for s, (r, _) in self.running.items()
                         if t - r.t_admit > self.timeout_s]:
                req = self._release(slot)
                self._finish(req, {"timeout": True, "converged": False})
                finished += 1
        return finished

    def pump(self) -> None:
        """Admit what fits, queue one engine step (the previous one keeps the
        device busy meanwhile), retire finished solves."""
        self._admit()
        if self.running:
            self.engine.launch(self.iters_per_step)
            self._collect()

    def run_open_loop(self, requests, keep_solutions: bool = False,
                      max_wall: float | None = None,
                      progress_every: float | None = None) -> dict:
        """Replay ``requests`` (sorted by arrival time) open-loop; returns the
        latency / throughput summary.  Requests not finished when
        ``max_wall`` seconds have passed are reported as unfinished.  With
        ``progress_every`` a one-line status goes to stderr that often."""
        self.keep_solutions = keep_solutions
        pending = collections.deque(sorted(requests, key=lambda r: r.t_arrival))
        self.clock0 = time.perf_counter()
        next_report = progress_every
        while pending or self.queue or self.running:
            t = self.now()
            if max_wall is not None and t > max_wall:
                break
            if next_report is not None and t >= next_report:
                next_report = t + progress_every
                print(f"serve: t={t:.0f}s done={len(self.done)} queued={len(self.queue)} "
                      f"running={len(self.running)} pending={len(pending)}",
                      file=sys.stderr, flush=True)
            while pending and pending[0].t_arrival <= t:
                self.submit(pending.popleft())
            if not self.queue and not self.running:
                if pending:
                    time.sleep(max(0.0, min(pending[0].t_arrival - t, 0.01)))
                continue
            self.pump()
        self.engine.poll(wait=True)
        wall = self.now()
        return summarize(self.done, wall, len(requests))


def summarize(done, wall: float, submitted: int) -> dict:
    lat = np.array([r.latency for r in done]) if done else np.zeros(0)
    ok = sum(1 for r in done if r.result.get("converged"))
    t_first = min((r.t_arrival for r in done), default=0.0)
    t_last = max((r.t_done for r in done), default=0.0)
    return {
        "submitted": submitted, "completed": len(done), "converged": ok,
        "rejected": sum(1 for r in done if r.result.get("rejected")),
        "timed_out": sum(1 for r in done if r.result.get("timeout")),
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


def one_at_a_time(requests, mixed: bool = False, max_iter: int = 20000) -> dict:
    """Baseline: each request built and solved by itself, one after another,
    with plain FP64 cg_solve_local (or cg_solve_mixed).  The set-up is built
    on the device as the engine builds it (the stencil assembled by
    upload_stencil, b by the same counter-based draws, x = 0).
    ``solves_per_s`` counts the solve calls only (set-up excluded, which
    favours the baseline); ``wall_solves_per_s`` includes the set-up."""
    ks.init(0)
    t0 = time.perf_counter()
    t_solve = 0.0
    iters = []
    for req in requests:
        A = ks.upload_stencil(req.dims, req.shift)
        b = ks.hash_uniform(A.nrows, req.seed)
        x = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        if mixed:
            r = ks.cg_solve_mixed(A, b, x, tol=req.tol, max_iter=max_iter)
        else:
            r = ks.cg_solve_local(A, b, x, tol=req.tol, max_iter=max_iter)
        t_solve += time.perf_counter() - t1
        iters.append(r["iters"])
    el = time.perf_counter() - t0
    return {"requests": len(requests), "seconds": el, "solve_seconds": t_solve,
            "solves_per_s": len(requests) / t_solve if t_solve > 0 else None,
            "wall_solves_per_s": len(requests) / el,
            "mean_iters": float(np.mean(iters)) if iters else None}
