"""Batched CG: many independent SPD systems advancing together on one GPU.

``CGEngine`` wraps the native engine (include/num_amd.h, num_engine_*): each
admitted system keeps its CSR handle, right-hand side, iterate and workspace
resident in HBM, and every ``step(k)`` runs k CG iterations of all live
systems with two launches per iteration (row tiles of every system in one
SpMV launch; row chunks of every system for the vector updates, whose last
chunk per system also finishes that system's scalars).  A system that converges leaves the launch lists at the
next step; a system admitted between steps joins the running batch.  Each
system's arithmetic depends only on its own data, so its result is
bit-identical whether it runs alone or with any other systems.

Modes: FP64 (the reference's CG sequence, src/cg_solver.hpp:133-196) or mixed
(FP32 CG steps with FP64-accumulated dots inside FP64 reliable-update
refinement; convergence on the FP64 true residual).  The mixed form refines
once the FP32 residual has dropped by ``delta`` (default 0.003: on the
4096-system batch 80.8k solves/s against 73.1k with 0.1, since every
refinement costs its system an FP64 iteration).

Systems come either as a CrsMatrix with caller-owned b / x (``add``) or as a
Dirichlet stencil assembled on the device in one buffer (``add_stencil``,
the serving path: one allocation and two kernel launches per request, no
host synchronisation).  ``launch`` / ``poll`` pipeline the steps: the host
admits and retires systems while the device runs the previous step.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import call
from .kokkos_spmv import CrsMatrix, _dev_tensor, _ptr, _stream

ENGINE_MAX_NNZ_PER_TILE = 7 * 256  # the engine's tile kernel holds 7 entries per thread


def _al256(b: int) -> int:
    return (b + 255) & ~255


def stencil_layout(dims, mixed: bool) -> dict:
    """Byte offsets inside an assembled stencil system's buffer (the layout
    num_engine_add_stencil carves): x, b, rowptr, colind, values[, values32],
    rpid (a row's pattern id: its neighbour bits, uint8), rpat (64 patterns
    of 8 int16 column - row offsets, then their lengths, 4 bits each),
    workspace."""
    from .generators import stencil_nnz
    n = 1
    for d in dims:
        n *= int(d)
    nnz = stencil_nnz(dims)
    off, o = {}, 0
    for name, size in (("x", 8 * n), ("b", 8 * n), ("rowptr", 4 * (n + 1)),
                       ("colind", 4 * nnz), ("values", 8 * nnz)) + \
            ((("values32", 4 * nnz),) if mixed else ()) + (("rpid", n), ("rpat", 64 * 16 + 32)):
        off[name] = o
        o += _al256(size)
    off["work"] = o
    off["n"], off["nnz"] = n, nnz
    return off


def _grid(dims):
    dims = tuple(int(d) for d in dims)
    if not 1 <= len(dims) <= 3:
        raise ValueError("dims must have 1, 2 or 3 entries")
    return (len(dims),) + dims + (1,) * (3 - len(dims))


class CGEngine:
    """A pool of up to ``max_systems`` concurrent CG solves.  ``graph``:
    replay each step's kernels from a captured hipGraph."""

    def __init__(self, max_systems: int = 4096, nt64: bool = True, nt32: bool = True,
                 graph: bool = False):
        h = C.c_void_p()
        call("num_engine_create", int(max_systems), int(bool(nt64)), int(bool(nt32)),
             C.byref(h))
        self._h = h
        if graph:
            call("num_engine_set_graph", h, 1)
        self._sys = {}
        self.max_systems = int(max_systems)
        self.ka_chunk = 0

    def set_schedule(self, ka_chunk: int = 0) -> None:
        """Tiles per XCD turn in the engine's SpMV launch (0 = by tile count);
        the autotuner's knob.  Never changes a result."""
        call("num_engine_set_schedule", self._h, int(ka_chunk))
        self.ka_chunk = int(ka_chunk)

    @staticmethod
    def work_bytes(n: int, mixed: bool = False) -> int:
        return int(_lib.load().num_engine_work_bytes(int(n), 1 if mixed else 0))

    def add(self, A: CrsMatrix, b, x, tol: float = 1e-10, max_iter: int = 1000,
            mixed: bool = False, delta: float = 0.003, work=None, tune: bool = False) -> int:
        """Admit A x = b (x holds x0, receives the solution).  Returns the slot.
        ``tune``: first let the autotuner pick A's row-tile schedule among
        those the engine's tile kernel runs (<= 7 entries per thread),
        remembered per matrix structure."""
        b = _dev_tensor(b, "b", 1)
        x = _dev_tensor(x, "x", 1)
        if A.nrows != A.ncols or b.shape[0] != A.nrows or x.shape[0] != A.nrows:
            raise ValueError("engine.add: A must be square and b, x of length A.nrows")
        if tune:
            from .autotune import tune_spmv
            tune_spmv(A, max_npt=ENGINE_MAX_NNZ_PER_TILE // 256)
        info = A.schedule_info()
        if info["nnz_per_block"] > ENGINE_MAX_NNZ_PER_TILE:
            # only the tile capacity changes: rows per tile, XCD chunk and
            # load kinds stay what the caller (or the autotuner) chose
            A.set_schedule(rows_per_block=info["rows_per_block"],
                           nnz_per_block=ENGINE_MAX_NNZ_PER_TILE, chunk=info["chunk"],
                           nt64=info["nt64"], nt32=info["nt32"])
        if mixed:
            A.ensure_f32()
        nbytes = self.work_bytes(A.nrows, mixed)
        if work is None:
            work = torch.empty(nbytes, dtype=torch.uint8, device=x.device)
        slot = C.c_int32()
        call("num_engine_add", self._h, A.handle, _ptr(b), _ptr(x), float(tol), int(max_iter),
             1 if mixed else 0, float(delta), _ptr(work), nbytes, _stream(), C.byref(slot))
        self._sys[slot.value] = (A, b, x, work, bool(mixed))
        return slot.value

    @staticmethod
    def stencil_bytes(dims, mixed: bool = False) -> int:
        b = int(_lib.load().num_engine_stencil_bytes(*_grid(dims), 1 if mixed else 0))
        if b < 0:
            raise ValueError(f"invalid stencil grid {tuple(dims)}")
        return b

    def add_stencil(self, dims, shift: float, seed: int, tol: float = 1e-10,
                    max_iter: int = 1000, mixed: bool = False, delta: float = 0.003,
                    buf=None, nbytes: int | None = None) -> int:
        """Admit the Dirichlet stencil system on grid ``dims`` (x fastest) with
        diagonal shift ``shift`` and b = generators.hash_uniform(n, seed),
        x0 = 0, assembled on the device inside one buffer (allocated here
        unless given).  Nothing waits for the device.  Returns the slot."""
        g = _grid(dims)
        if nbytes is None:
            nbytes = self.stencil_bytes(dims, mixed)
        if buf is None:
            buf = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        elif not (buf.is_cuda and buf.dtype == torch.uint8 and buf.is_contiguous()
                  and buf.numel() >= nbytes):
            raise ValueError(f"add_stencil: buf must be a contiguous CUDA uint8 tensor of "
                             f">= {nbytes} bytes")
        slot = C.c_int32()
        call("num_engine_add_stencil", self._h, *g, float(shift), int(seed) % 2**64, float(tol),
             int(max_iter), 1 if mixed else 0, float(delta), _ptr(buf), buf.numel(), _stream(),
             C.byref(slot))
        n = g[1] * g[2] * g[3]
        self._sys[slot.value] = (buf, buf[: 8 * n].view(torch.float64))
        return slot.value

    def solution(self, slot: int) -> torch.Tensor:
        """The solution vector of a system admitted with add_stencil (a view
        of its buffer) or the x given to add."""
        ent = self._sys[int(slot)]
        return ent[1] if len(ent) == 2 else ent[2]

    def launch(self, iters: int = 8) -> None:
        """Queue ``iters`` iterations of every live system and return (at most
        two steps stay queued; the older one is waited for first)."""
        call("num_engine_launch", self._h, int(iters), _stream())

    def poll(self, wait: bool = False) -> tuple:
        """Retire completed steps (all, waiting, if ``wait``); returns
        (systems not known finished, steps still queued)."""
        a, b = C.c_int32(), C.c_int32()
        call("num_engine_poll", self._h, 1 if wait else 0, C.byref(a), C.byref(b))
        return a.value, b.value

    def step(self, iters: int = 8) -> int:
        """Run ``iters`` iterations of every live system; returns how many
        are still running (the call returns after the device work)."""
        n = C.c_int32()
        call("num_engine_step", self._h, int(iters), C.byref(n), _stream())
        return n.value

    def result(self, slot: int) -> dict:
        res = _lib.CgResult()
        nref = C.c_int32()
        call("num_engine_result", self._h, int(slot), C.byref(res), C.byref(nref), _stream())
        return {"iters": int(res.iters), "residual": float(res.residual),
                "converged": bool(res.converged), "breakdown": bool(res.breakdown),
                "refinements": int(nref.value)}

    def finished(self) -> list:
        """Slots whose system has finished and not been removed."""
        cnt = C.c_int32()
        call("num_engine_finished", self._h, None, 0, C.byref(cnt))
        if cnt.value == 0:
            return []
        buf = (C.c_int32 * cnt.value)()
        call("num_engine_finished", self._h, buf, cnt.value, C.byref(cnt))
        return list(buf)

    def remove(self, slot: int) -> None:
        call("num_engine_remove", self._h, int(slot))
        self._sys.pop(int(slot), None)

    def info(self) -> dict:
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        call("num_engine_info", self._h, C.byref(a), C.byref(b), C.byref(c))
        return {"live": a.value, "tiles": b.value, "chunks": c.value}

    def run(self, iters_per_step: int = 16) -> None:
        """Step until every admitted system has finished."""
        while self.step(iters_per_step) > 0:
            pass

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().num_engine_destroy(h)
            except Exception:
                pass
            self._h = None


def solve_batch(systems, tol: float = 1e-10, max_iter: int = 1000, mixed: bool = False,
                delta: float = 0.003, iters_per_step: int = 16, engine: CGEngine | None = None):
    """Solve every (A, b, x) of ``systems`` together; x is overwritten.
    Returns the per-system result dicts in input order."""
    eng = engine or CGEngine(max_systems=max(1, len(systems)))
    slots = [eng.add(A, b, x, tol=tol, max_iter=max_iter, mixed=mixed, delta=delta)
             for A, b, x in systems]
    eng.run(iters_per_step)
    out = [eng.result(s) for s in slots]
    for s in slots:
        eng.remove(s)
    return out
