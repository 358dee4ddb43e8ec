"""Drop-in replacement for the reference's ``kokkos_spmv`` module on MI355X.

Same names, arguments, defaults, return values and exceptions as the nanobind
module of src/bindings.cpp:38-290, with PyTorch-ROCm tensors (``device='cuda'``,
float64, contiguous) in place of the CUDA ndarrays:

    init, finalize, rank, size, barrier            bindings.cpp:53-76
    CrsMatrix (nrows, ncols, nnz, repr)            bindings.cpp:79-88
    DistCrsMatrix (global_nrows, local_row_start,  bindings.cpp:90-94
                   local_row_end)
    upload_csr, distribute_csr                     bindings.cpp:97-148
    spmv(A, x, y, kokkoskernels=True)              bindings.cpp:152-183
    batch_spmv(A, X, Y)                            bindings.cpp:186-215
    dist_spmv(A_dist, x_local, y_local)            bindings.cpp:218-236
    cg_solve(A_dist, b, x, tol, max_iter)          bindings.cpp:239-270
    cg_solve_local(A, b, x, tol, max_iter)         bindings.cpp:273-290

plus the vector ops of src/vec_ops.hpp (axpy, axpy2, dot, norm2, copy, fill,
scale), which the reference keeps C++-only.  Every call runs hand-written HIP
kernels from libnum_amd.so and, like every reference call (each ends in a
Kokkos::fence), returns after the device work is complete; pass
``sync=False`` to leave the work queued on the current stream.

Errors: wrong shapes raise ValueError (the reference's std::invalid_argument),
wrong dtype/device/rank raise TypeError (nanobind's ndarray signature
mismatch), numerical breakdown raises RuntimeError.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import call

_state = {"initialized": False, "device": 0}


# ---------------------------------------------------------------- runtime --

def init(device_id: int = 0):
    """Select the GPU; returns (rank, size).  Idempotent (bindings.cpp:53-66)."""
    lib = _lib.load()
    if not torch.cuda.is_available():
        raise RuntimeError("num_amd: no ROCm GPU is visible (torch.cuda.is_available() "
                           "is False); there is no CPU fallback")
    torch.cuda.set_device(int(device_id))
    _lib.check(lib.num_init(int(device_id)), "init")
    _state["initialized"] = True
    _state["device"] = int(device_id)
    return rank(), size()


def finalize() -> None:
    """bindings.cpp:68-70; frees nothing the handles still own."""
    _lib.check(_lib.load().num_finalize(), "finalize")


def _dist():
    import torch.distributed as dist
    return dist if dist.is_available() and dist.is_initialized() else None


def rank() -> int:
    d = _dist()
    return d.get_rank() if d else 0


def size() -> int:
    d = _dist()
    return d.get_world_size() if d else 1


def barrier() -> None:
    d = _dist()
    if d:
        d.barrier()


def _ensure_init():
    if not _state["initialized"]:
        init(torch.cuda.current_device() if torch.cuda.is_available() else 0)


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _sync():
    torch.cuda.current_stream().synchronize()


# ------------------------------------------------------------ validation --

def _dev_tensor(t, name: str, ndim: int, dtype=torch.float64) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch.Tensor on the GPU, got {type(t).__name__}")
    if t.device.type != "cuda":
        raise TypeError(f"{name}: expected a GPU tensor, got device {t.device}")
    if t.dtype != dtype:
        raise TypeError(f"{name}: expected dtype {dtype}, got {t.dtype}")
    if t.dim() != ndim:
        raise TypeError(f"{name}: expected {ndim}-D tensor, got {t.dim()}-D")
    if not t.is_contiguous():
        raise ValueError(f"{name}: tensor must be contiguous (row-major)")
    return t


def _host_array(a, name: str, dtype) -> np.ndarray:
    if isinstance(a, torch.Tensor):
        if a.device.type != "cpu":
            raise TypeError(f"{name}: expected a host array")
        a = a.numpy()
    a = np.asarray(a)
    if a.ndim != 1:
        raise TypeError(f"{name}: expected a 1-D array")
    if a.dtype != dtype:
        raise TypeError(f"{name}: expected dtype {np.dtype(dtype).name}, got {a.dtype}")
    return np.ascontiguousarray(a)


def _ptr(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


# ------------------------------------------------------------- CrsMatrix --

class CrsMatrix:
    """Device CSR matrix (src/crs_matrix.hpp:37-73).  Owns its device arrays
    (torch tensors) and the native handle with its row-block schedule."""

    __slots__ = ("_h", "_rowptr", "_colind", "_values", "_values32", "_rel16", "_d16", "_xcd",
                 "_nrows", "_ncols", "_nnz", "_h_rowptr", "__weakref__")

    def __init__(self):
        raise TypeError("CrsMatrix objects are created by upload_csr()")

    @classmethod
    def _from_device(cls, h_rowptr: np.ndarray, rowptr: torch.Tensor, colind: torch.Tensor,
                     values: torch.Tensor, nrows: int, ncols: int,
                     index: str = "auto") -> "CrsMatrix":
        self = object.__new__(cls)
        self._h = None
        self._rowptr, self._colind, self._values = rowptr, colind, values
        self._values32 = None
        self._rel16 = None
        self._d16 = None
        self._xcd = None
        self._h_rowptr = h_rowptr
        h = C.c_void_p()
        call("num_csr_create", h_rowptr.ctypes.data_as(C.c_void_p), _ptr(rowptr), _ptr(colind),
             _ptr(values), int(nrows), int(ncols), C.byref(h))
        self._h = h
        r, c, z = C.c_int32(), C.c_int32(), C.c_int64()
        call("num_csr_shape", h, C.byref(r), C.byref(c), C.byref(z))
        self._nrows, self._ncols, self._nnz = r.value, c.value, z.value
        if index == "auto" and self._nnz > 0 and int(np.diff(h_rowptr).max()) <= 32:
            # delta-coded column ids for the FP64 kernels where they fit
            # nnz + 8 codes (8 of padding), then one row-length byte per row
            codes = torch.empty(self._nnz + 8 + (self._nrows + 1) // 2, dtype=torch.int16,
                                device=values.device)
            ok = C.c_int32()
            call("num_csr_attach_d16", h, _ptr(codes), _stream(), C.byref(ok))
            self._d16 = codes if ok.value else None
        return self

    @property
    def layout(self) -> str:
        """'xcd' when spmv() runs the column-partitioned layout (see
        upload_csr), else 'rows'."""
        return "xcd" if self._xcd is not None else "rows"

    @property
    def index_format(self) -> str:
        """What the FP64 SpMV / CG kernels read for the columns.
        'windowpat': the windowed sliced layout with row patterns (every
        tile reads x from a few column ranges fetched into LDS; each row a
        4-bit id into its tile's table of slot patterns, tables shared by
        the tiles that have equal ones; rows <= 8 entries).
        'window16': the same layout with a 16-bit slot per entry (a tile with
        more than 16 distinct row patterns, or set_windows(patterns=False)).
        'delta16': the delta-coded column ids
        (every row <= 32 entries in ascending column order, codes within 16
        bits) when the schedule streams the FP64 matrix from HBM (nt64; a
        matrix that stays in the Infinity Cache reads its int32 columns
        faster).  Else 'int32'."""
        w = self._window_info()
        if w:
            return "windowpat" if w == 2 else "window16"
        if self._d16 is None or not self.schedule_info()["nt64"]:
            return "int32"
        return "delta16"

    @property
    def nrows(self) -> int:
        return self._nrows

    @property
    def ncols(self) -> int:
        return self._ncols

    @property
    def nnz(self) -> int:
        return self._nnz

    @property
    def handle(self):
        return self._h

    def __repr__(self) -> str:
        return f"<CrsMatrix nrows={self._nrows} ncols={self._ncols} nnz={self._nnz}>"

    def ensure_f32(self) -> None:
        """Attach an FP32 copy of the values (mixed-precision inner solve)."""
        if self._values32 is None:
            self._values32 = self._values.to(torch.float32)
            call("num_csr_attach_f32", self._h, _ptr(self._values32))
            # int16 column offsets for the FP32 walk where they fit
            rel = torch.empty(max(1, self.nnz), dtype=torch.int16, device=self._values.device)
            ok = C.c_int32()
            call("num_csr_attach_rel16", self._h, _ptr(rel), _stream(), C.byref(ok))
            self._rel16 = rel if ok.value else None

    def set_schedule(self, rows_per_block: int = 0, nnz_per_block: int = 0, chunk: int = -1,
                     nt64: int = -1, nt32: int = -1) -> None:
        """Rebuild the row-tile schedule (0 / -1 = default for that knob)."""
        call("num_csr_set_schedule", self._h, int(rows_per_block), int(nnz_per_block),
             int(chunk), int(nt64), int(nt32))
        if self._d16 is not None:  # re-coded for the new tiles, or detached
            on = C.c_int32()
            call("num_csr_index_info", self._h, C.byref(on))
            if not on.value:
                self._d16 = None

    def set_windows(self, enable: bool | None = True, patterns: bool = True) -> None:
        """Build (True) or drop (False) the windowed sliced layout
        (num_csr_set_windows): a banded or stencil matrix whose tiles each
        read x from a few contiguous column ranges fetches them once per
        tile into LDS and streams its rows in sliced order.  None restores
        the default (built for matrices of at least 65,536 rows).
        ``patterns=False`` keeps a 16-bit slot per entry instead of the row
        patterns.  Same results either way (bit-identical)."""
        mode = -1 if enable is None else int(bool(enable))
        if mode == 1 and not patterns:
            mode = 2
        call("num_csr_set_windows", self._h, mode, _stream())

    def set_mixed_path(self, mode: int = -1) -> None:
        """Kernel of cg_solve_mixed's FP32 steps (num_csr_set_mixed_path):
        1 the windowed sliced step, 0 the staged pair / row-tile kernels (the
        pair kernel with the int16 offsets' row patterns where built), 2 the
        staged kernels with 2 offset bytes per entry, -1 (default) windowed
        when the FP32 copy streams from HBM and the matrix's long-range
        column offset is at least 16,384 rows, else as 0."""
        call("num_csr_set_mixed_path", self._h, int(mode))

    @property
    def patterns(self) -> dict:
        """Which row-pattern forms are built (num_csr_pattern_info): 'windows'
        for the windowed layout, 'offsets' for the FP32 pair kernel's int16
        column offsets."""
        w, o = C.c_int32(), C.c_int32()
        call("num_csr_pattern_info", self._h, C.byref(w), C.byref(o))
        return {"windows": bool(w.value), "offsets": bool(o.value)}

    def _window_info(self) -> int:
        on = C.c_int32()
        call("num_csr_window_info", self._h, C.byref(on))
        return on.value

    @property
    def windowed(self) -> bool:
        """True when the FP64 kernels read x through column windows."""
        return self._window_info() != 0

    def schedule_info(self) -> dict:
        v = [C.c_int32() for _ in range(8)]
        call("num_csr_schedule_info", self._h, *[C.byref(x) for x in v])
        keys = ("nblocks", "rows_per_block", "nnz_per_block", "long_rows", "max_row_nnz",
                "chunk", "nt64", "nt32")
        return {k: x.value for k, x in zip(keys, v)}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().num_csr_destroy(h)
            except Exception:
                pass
            self._h = None


# x blocks of at most 4 MiB (one XCD's L2), at least one per XCD: on the Zipf
# 4M matrix 4 / 2 / 1 MiB blocks gave 289 / 315 / 359 us (fewer segments)
_XCD_BLOCK_BYTES = 4 << 20
_NUM_XCD = 8


def _wants_xcd(rp: np.ndarray, ci: np.ndarray, nrows: int, ncols: int) -> bool:
    """The column-partitioned layout pays when x is far larger than one
    XCD's 4 MB L2 and most gathers land far from their row (random
    columns): then nearly every gather fills a 128-byte L2 line."""
    nnz = int(rp[-1])
    if ncols * 8 <= 2 * _XCD_BLOCK_BYTES or nnz < (1 << 20):
        return False
    # a strided sample of at most 1M entries decides
    idx = np.arange(0, nnz, max(1, nnz // (1 << 20)), dtype=np.int64)
    rows = np.searchsorted(rp, idx, side="right") - 1
    far = np.abs(ci[idx].astype(np.int64) - rows) > (_XCD_BLOCK_BYTES // 8)
    return bool(far.mean() > 0.5)


def _xcd_plan(rp: np.ndarray, ci: np.ndarray, vv: np.ndarray, nrows: int, ncols: int):
    """Segments of the column-partitioned layout: columns cut into nb equal
    blocks (nb a multiple of 8, each block's x at most 4 MiB); a segment is
    one row's entries in one block, in their original order; segments are
    block-major, rows ascending inside a block."""
    per = _NUM_XCD * _XCD_BLOCK_BYTES // 8
    nb = _NUM_XCD * max(1, -(-ncols // per))
    bs = -(-ncols // nb)
    lens = np.diff(rp)
    rows = np.repeat(np.arange(nrows, dtype=np.int64), lens)
    key = (ci.astype(np.int64) // bs) * nrows + rows
    order = np.argsort(key, kind="stable")
    kk = key[order]
    first = np.ones(kk.shape[0], dtype=bool)
    first[1:] = kk[1:] != kk[:-1]
    seg_start = np.flatnonzero(first)
    seg_key = kk[seg_start]
    seg_blk = seg_key // nrows
    seg_row = seg_key - seg_blk * nrows
    nseg = seg_start.shape[0]
    seg_ptr = np.empty(nseg + 1, dtype=np.int32)
    seg_ptr[:-1] = seg_start
    seg_ptr[-1] = kk.shape[0]
    block_start = np.searchsorted(seg_blk, np.arange(nb + 1)).astype(np.int32)
    by_row = np.lexsort((seg_blk, seg_row)).astype(np.int32)  # row, then block
    row_seg_ptr = np.zeros(nrows + 1, dtype=np.int32)
    np.cumsum(np.bincount(seg_row, minlength=nrows), out=row_seg_ptr[1:])
    return (nb, seg_ptr, ci[order].astype(np.int32), vv[order], block_start, row_seg_ptr,
            by_row)


def _attach_xcd(A: CrsMatrix, rp, ci, vv) -> None:
    nb, seg_ptr, cols, vals, block_start, row_seg_ptr, row_seg_idx = _xcd_plan(
        rp, ci, vv, A.nrows, A.ncols)
    dev = A._values.device
    S = CrsMatrix._from_device(seg_ptr, torch.from_numpy(seg_ptr).to(dev),
                               torch.from_numpy(cols).to(dev), torch.from_numpy(vals).to(dev),
                               seg_ptr.shape[0] - 1, A.ncols, index="int32")
    d_rsp = torch.from_numpy(row_seg_ptr).to(dev)
    d_rsi = torch.from_numpy(row_seg_idx).to(dev) if row_seg_idx.size else \
        torch.zeros(1, dtype=torch.int32, device=dev)
    call("num_csr_attach_xcd", A._h, S._h, int(nb), block_start.ctypes.data_as(C.c_void_p),
         _ptr(d_rsp), _ptr(d_rsi))
    A._xcd = (S, d_rsp, d_rsi, int(seg_ptr.shape[0] - 1))


def upload_csr(rowptr, colind, values, nrows: int, ncols: int, *,
               index: str = "auto", layout: str = "auto") -> CrsMatrix:
    """Copy a host CSR matrix to the GPU (bindings.cpp:97-123).

    rowptr int32[nrows+1], colind int32[nnz], values float64[nnz] (host).
    Raises ValueError for non-positive dimensions (crs_matrix.hpp:52-53) and,
    beyond the reference, for inconsistent lengths or out-of-range columns.
    index="auto" also builds the delta-coded column ids (2 bytes per entry
    more on the device) that the FP64 SpMV and CG read instead of colind
    where every row holds <= 32 entries in ascending column order;
    index="int32" keeps the int32 columns only.  Results are the same
    either way.
    layout="auto" also builds, for a matrix whose x (more than 8 MiB) is
    gathered mostly far from each row (random columns), a column-partitioned
    copy that spmv() runs instead (each XCD gathers from its own x blocks of
    up to 4 MiB; 12 bytes per entry and about 12 per row-block pair more on the
    device); layout="xcd" forces it, layout="rows" never builds it.  Rows
    whose entries lie in one block give the same bits either way, the others
    agree to rounding.
    """
    if index not in ("auto", "int32"):
        raise ValueError("index must be 'auto' or 'int32'")
    if layout not in ("auto", "rows", "xcd"):
        raise ValueError("layout must be 'auto', 'rows' or 'xcd'")
    _ensure_init()
    nrows, ncols = int(nrows), int(ncols)
    rp = _host_array(rowptr, "rowptr", np.int32)
    ci = _host_array(colind, "colind", np.int32)
    vv = _host_array(values, "values", np.float64)
    if nrows <= 0 or ncols <= 0:
        raise ValueError("CrsMatrixHandle: non-positive dimensions")
    if rp.shape[0] < nrows + 1:
        raise ValueError(f"rowptr has {rp.shape[0]} entries, needs nrows+1 = {nrows + 1}")
    rp = rp[: nrows + 1]
    nnz = int(rp[-1])
    if rp[0] != 0 or np.any(np.diff(rp) < 0):
        raise ValueError("rowptr must start at 0 and be non-decreasing")
    if ci.shape[0] < nnz or vv.shape[0] < nnz:
        raise ValueError(f"colind/values shorter than nnz = rowptr[nrows] = {nnz}")
    ci, vv = ci[:nnz], vv[:nnz]
    if nnz and (int(ci.min()) < 0 or int(ci.max()) >= ncols):
        raise ValueError(f"column index outside [0, {ncols})")
    dev = torch.device("cuda", torch.cuda.current_device())
    d_rp = torch.from_numpy(rp).to(dev)
    d_ci = torch.from_numpy(ci).to(dev) if nnz else torch.zeros(1, dtype=torch.int32, device=dev)
    d_v = torch.from_numpy(vv).to(dev) if nnz else torch.zeros(1, dtype=torch.float64, device=dev)
    A = CrsMatrix._from_device(rp, d_rp, d_ci, d_v, nrows, ncols, index=index)
    if nnz and (layout == "xcd" or (layout == "auto" and _wants_xcd(rp, ci, nrows, ncols))):
        _attach_xcd(A, rp, ci, vv)
    return A


def upload_stencil(dims, shift: float = 0.0) -> CrsMatrix:
    """The 1-/2-/3-D Dirichlet Laplacian (5-point in 2-D, 7-point in 3-D) on
    ``dims`` (x fastest) plus ``shift`` on the diagonal, assembled directly on
    the GPU (the host computes only rowptr).  Same matrix as
    ``generators.stencil(dims, shift)``."""
    from .generators import stencil_rowptr
    _ensure_init()
    dims = tuple(int(d) for d in dims)
    if not 1 <= len(dims) <= 3 or min(dims) < 1:
        raise ValueError("dims must hold 1 to 3 positive sizes")
    rp = stencil_rowptr(dims)
    n = int(np.prod(dims))
    nnz = int(rp[-1])
    dev = torch.device("cuda", torch.cuda.current_device())
    d_rp = torch.from_numpy(rp).to(dev)
    d_ci = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)
    d_v = torch.empty(max(nnz, 1), dtype=torch.float64, device=dev)
    full = dims + (1,) * (3 - len(dims))
    call("num_stencil_fill", len(dims), full[0], full[1], full[2], float(shift), _ptr(d_rp),
         _ptr(d_ci), _ptr(d_v), _stream())
    return CrsMatrix._from_device(rp, d_rp, d_ci, d_v, n, n)


def hash_uniform(n: int, seed: int) -> torch.Tensor:
    """generators.hash_uniform(n, seed) computed on the GPU (the right-hand
    side the engine assembles for a served request)."""
    _ensure_init()
    out = torch.empty(max(1, int(n)), dtype=torch.float64,
                      device=torch.device("cuda", torch.cuda.current_device()))[: int(n)]
    call("num_hash_uniform", int(n), int(seed) % 2**64, _ptr(out), _stream())
    return out


# ------------------------------------------------------------------ SpMV --

def spmv(A: CrsMatrix, x, y, kokkoskernels: bool = True, *, sync: bool = True) -> None:
    """y = A @ x in place (bindings.cpp:152-183).  ``kokkoskernels`` is
    accepted for compatibility; both settings run the same row-binned kernel."""
    x = _dev_tensor(x, "x", 1)
    y = _dev_tensor(y, "y", 1)
    if x.shape[0] != A.ncols:
        raise ValueError("spmv: x length != A.ncols")
    if y.shape[0] != A.nrows:
        raise ValueError("spmv: y length != A.nrows")
    if A._xcd is not None:
        # per-call partials from torch's stream-ordered allocator (no hipMalloc
        # after the first call), so calls on different streams never share them
        work = torch.empty(max(1, A._xcd[3]), dtype=torch.float64, device=y.device)
        call("num_spmv_xcd", A._h, _ptr(x), _ptr(y), _ptr(work), _stream())
    else:
        call("num_spmv", A._h, _ptr(x), _ptr(y), _stream())
    if sync:
        _sync()


def batch_spmv(A: CrsMatrix, X, Y, *, sync: bool = True) -> None:
    """Y = A @ X with X (ncols, k), Y (nrows, k) row-major (bindings.cpp:186-215)."""
    X = _dev_tensor(X, "X", 2)
    Y = _dev_tensor(Y, "Y", 2)
    k = X.shape[1]
    if X.shape[0] != A.ncols:
        raise ValueError("batch_spmv: X.shape[0] != A.ncols")
    if Y.shape[0] != A.nrows or Y.shape[1] != k:
        raise ValueError("batch_spmv: Y must be (A.nrows, k)")
    call("num_batch_spmv", A._h, _ptr(X), _ptr(Y), int(k), _stream())
    if sync:
        _sync()


# ---------------------------------------------------------------- vec ops --

def _same_len(*ts):
    n = ts[0].shape[0]
    for t in ts[1:]:
        if t.shape[0] != n:
            raise ValueError("vector lengths differ")
    return n


def axpy(alpha: float, x, y, z, *, sync: bool = True) -> None:
    """z = alpha*x + y (vec_ops.hpp:16-26); z may alias x or y."""
    x, y, z = (_dev_tensor(t, nm, 1) for t, nm in ((x, "x"), (y, "y"), (z, "z")))
    n = _same_len(x, y, z)
    call("num_axpy", n, float(alpha), _ptr(x), _ptr(y), _ptr(z), _stream())
    if sync:
        _sync()


def axpy2(alpha: float, x, beta: float, y, z, *, sync: bool = True) -> None:
    """z = alpha*x + beta*y (vec_ops.hpp:31-42)."""
    x, y, z = (_dev_tensor(t, nm, 1) for t, nm in ((x, "x"), (y, "y"), (z, "z")))
    n = _same_len(x, y, z)
    call("num_axpy2", n, float(alpha), _ptr(x), float(beta), _ptr(y), _ptr(z), _stream())
    if sync:
        _sync()


def _dot_call(x, y, sqrt: bool) -> float:
    n = x.shape[0]
    ws = torch.empty(int(_lib.load().num_dot_workspace_bytes(n)) // 8 + 1,
                     dtype=torch.float64, device=x.device)
    out = ws[-1:]
    if sqrt:
        call("num_norm2", n, _ptr(x), _ptr(out), _ptr(ws), _stream())
    else:
        call("num_dot", n, _ptr(x), _ptr(y), _ptr(out), _ptr(ws), _stream())
    return float(out.item())


def dot(x, y) -> float:
    """sum_i x_i y_i (vec_ops.hpp:47-59), deterministic two-stage reduction."""
    x = _dev_tensor(x, "x", 1)
    y = _dev_tensor(y, "y", 1)
    _same_len(x, y)
    return _dot_call(x, y, False)


def norm2(x) -> float:
    """||x||_2 (vec_ops.hpp:64-69)."""
    x = _dev_tensor(x, "x", 1)
    return _dot_call(x, x, True)


def copy(src, dst, *, sync: bool = True) -> None:
    """dst = src (vec_ops.hpp:74-79)."""
    src = _dev_tensor(src, "src", 1)
    dst = _dev_tensor(dst, "dst", 1)
    n = _same_len(src, dst)
    call("num_copy", n, _ptr(src), _ptr(dst), _stream())
    if sync:
        _sync()


def fill(x, val: float, *, sync: bool = True) -> None:
    """x[i] = val (vec_ops.hpp:84-88)."""
    x = _dev_tensor(x, "x", 1)
    call("num_fill", x.shape[0], float(val), _ptr(x), _stream())
    if sync:
        _sync()


def scale(x, alpha: float, *, sync: bool = True) -> None:
    """x *= alpha (vec_ops.hpp:93-100)."""
    x = _dev_tensor(x, "x", 1)
    call("num_scale", x.shape[0], float(alpha), _ptr(x), _stream())
    if sync:
        _sync()


# -------------------------------------------------------------------- CG --

def _cg_result(res) -> dict:
    return {"iters": int(res.iters), "residual": float(res.residual),
            "converged": bool(res.converged)}


def cg_solve_local(A: CrsMatrix, b, x, tol: float = 1e-10, max_iter: int = 1000) -> dict:
    """Single-GPU CG (bindings.cpp:273-290, cg_solver.hpp:133-196).

    x is the initial guess and is overwritten with the solution.  Returns
    {'iters', 'residual', 'converged'}.  Raises RuntimeError when p'Ap
    underflows 1e-300 (the guard of cg_solver.hpp:84-85, applied here too).
    """
    b = _dev_tensor(b, "b", 1)
    x = _dev_tensor(x, "x", 1)
    if A.nrows != A.ncols:
        raise ValueError("cg_solve_local: A must be square")
    if b.shape[0] != A.nrows or x.shape[0] != A.nrows:
        raise ValueError("cg_solve_local: b and x must have length A.nrows")
    lib = _lib.load()
    nbytes = int(lib.num_cg_workspace_bytes(A._h))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=x.device)
    res = _lib.CgResult()
    rc = lib.num_cg_solve(A._h, _ptr(b), _ptr(x), float(tol), int(max_iter), _ptr(ws), nbytes,
                          C.byref(res), _stream())
    _sync()
    _lib.check(rc, "cg_solve_local")
    return _cg_result(res)


def cg_solve_mixed(A: CrsMatrix, b, x, tol: float = 1e-10, max_iter: int = 100000,
                   delta: float = 0.01) -> dict:
    """FP32-inner / FP64-refined CG (no reference counterpart).

    CG steps run in FP32 (FP32 matrix copy and vectors, FP64-accumulated
    dots).  Each time the FP32 residual has dropped by ``delta`` since the last
    refinement, the GPU refines in FP64: x += x32, r = b - A x (FP64 true
    residual, also the convergence test, strict < tol), the FP32 residual is
    re-seeded from r and the search direction kept.  Returns {'outer' (FP64
    refinements), 'inner_total' / 'iters' (FP32 steps), 'residual' (FP64 true
    residual), 'converged'}.
    """
    b = _dev_tensor(b, "b", 1)
    x = _dev_tensor(x, "x", 1)
    if A.nrows != A.ncols:
        raise ValueError("cg_solve_mixed: A must be square")
    if b.shape[0] != A.nrows or x.shape[0] != A.nrows:
        raise ValueError("cg_solve_mixed: b and x must have length A.nrows")
    A.ensure_f32()
    lib = _lib.load()
    nbytes = int(lib.num_cg_mixed_workspace_bytes(A._h))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=x.device)
    res = _lib.MixedResult()
    rc = lib.num_cg_solve_mixed(A._h, _ptr(b), _ptr(x), float(tol), int(max_iter),
                                float(delta), _ptr(ws), nbytes, C.byref(res), _stream())
    _sync()
    _lib.check(rc, "cg_solve_mixed")
    return {"outer": int(res.outer), "inner_total": int(res.inner_total),
            "iters": int(res.inner_total), "residual": float(res.true_residual),
            "converged": bool(res.converged)}


# ------------------------------------------------------ one rank's share --

class DistCrsMatrix:
    """One rank's share of the 1-D row partition (crs_matrix.hpp:101-114).

    ``local`` is the remapped local matrix ([own | ghost] columns,
    dist_spmv.hpp:62-82); ``ghost_ids`` the global id of every ghost slot in
    slot order; ``send_ids`` / ``send_idx`` the rows this rank sends each
    neighbour (halo_exchange.hpp:154-191); ``row_runs`` the interior and
    boundary row ranges of the split SpMV.
    """

    def __init__(self):
        raise TypeError("DistCrsMatrix objects are created by distribute_csr()")

    @property
    def global_nrows(self) -> int:
        return self._global_nrows

    @property
    def local_row_start(self) -> int:
        return self._start

    @property
    def local_row_end(self) -> int:
        return self._end

    @property
    def n_ghost(self) -> int:
        return int(self.ghost_ids.shape[0])

    def ghost_values(self, x_global) -> torch.Tensor:
        """The ghost block a halo exchange would deliver: x_global[ghost_ids]."""
        if isinstance(x_global, torch.Tensor):
            return x_global.index_select(0, self._ghost_idx_dev)
        return torch.from_numpy(np.ascontiguousarray(np.asarray(x_global)[self.ghost_ids])).to(
            self.local._values.device)

    def __repr__(self) -> str:
        return (f"<DistCrsMatrix rank={self.rank}/{self.size} rows=[{self._start}, {self._end}) "
                f"ghosts={self.n_ghost}>")


def _runs(mask: np.ndarray):
    """Maximal runs of True in mask -> [(begin, end)]."""
    m = np.concatenate([[False], mask, [False]]).astype(np.int8)
    d = np.diff(m)
    return list(zip(np.nonzero(d == 1)[0].tolist(), np.nonzero(d == -1)[0].tolist()))


def distribute_csr(rowptr, colind, values, global_nrows: int, global_ncols: int, *,
                   rank: int | None = None, size: int | None = None) -> DistCrsMatrix:
    """Build this rank's local matrix exactly as distribute_csr does
    (dist_spmv.hpp:27-100; ghost order of halo_exchange.hpp:62-75,106-128).

    ``rank``/``size`` default to this process's rank and world size; passing
    them builds any rank's share on this one GPU (one GPU's share of an
    N-slab run at its real shape).
    """
    from .generators import slab_share
    _ensure_init()
    rank = globals()["rank"]() if rank is None else int(rank)
    size = globals()["size"]() if size is None else int(size)
    if not (0 <= rank < size):
        raise ValueError(f"rank {rank} outside [0, {size})")
    rp = _host_array(rowptr, "rowptr", np.int32)[: int(global_nrows) + 1]
    ci = _host_array(colind, "colind", np.int32)
    vv = _host_array(values, "values", np.float64)
    sh = slab_share(rp, ci, vv, int(global_nrows), int(global_ncols), rank, size)
    self = object.__new__(DistCrsMatrix)
    self.rank, self.size = rank, size
    self._global_nrows = int(global_nrows)
    self._start, self._end = sh.local_row_start, sh.local_row_end
    self.local = upload_csr(sh.local.rowptr, sh.local.colind, sh.local.values, sh.local.nrows,
                            sh.local.ncols)
    self.ghost_ids = sh.ghost_ids
    self.recv = sh.recv
    dev = self.local._values.device
    self._ghost_idx_dev = torch.from_numpy(sh.ghost_ids).to(dev)
    n_local = sh.local.nrows
    boundary = np.zeros(n_local, dtype=bool)
    boundary[sh.boundary_rows] = True
    self.interior_runs = _runs(~boundary)
    self.boundary_runs = _runs(boundary)
    self.interior_rows = sh.interior_rows
    self.boundary_rows = sh.boundary_rows
    # rows every neighbour needs from us, by rank ascending, ids ascending
    send = []
    j0, j1 = int(rp[self._start]), int(rp[self._end])
    cols_all = ci[: int(rp[-1])]
    owner_rows = np.repeat(np.arange(int(global_nrows), dtype=np.int64), np.diff(rp))
    mine = (cols_all >= self._start) & (cols_all < self._end)
    mine[j0:j1] = False
    if mine.any():
        from .generators import row_split
        starts = np.array([row_split(int(global_nrows), r, size)[0] for r in range(size)])
        req_rank = np.searchsorted(starts, owner_rows[mine], side="right") - 1
        req_col = cols_all[mine].astype(np.int64)
        for r in np.unique(req_rank):
            ids = np.unique(req_col[req_rank == r])
            send.append((int(r), ids))
    self.send = send
    idx = (np.concatenate([ids for _, ids in send]) - self._start).astype(np.int32) \
        if send else np.zeros(0, np.int32)
    self.send_idx = torch.from_numpy(idx).to(dev)
    return self


def dist_spmv(A_dist: DistCrsMatrix, x_local, y_local, x_ghost=None, *,
              overlap: bool = False, sync: bool = True) -> None:
    """y_local = (A x)[own rows] (bindings.cpp:218-236, dist_spmv.hpp:118-163).

    Reads [x_local | x_ghost] without building x_ext.  With more than one
    rank the ghost block comes from the halo exchange, which is not run on one
    GPU: pass ``x_ghost`` (e.g. ``A_dist.ghost_values(x_global)``).  The ghost
    block is complete on entry, so the share runs as one launch over all its
    rows; ``overlap=True`` runs the interior rows (no ghost columns) on
    x_local first and the boundary rows after them, the order an exchange in
    flight would need (3 launches for a middle slab: 20.9 us against about
    15 us for one launch on the 192^3 rank-3 share).  Same results either way.
    """
    A = A_dist.local
    x_local = _dev_tensor(x_local, "x_local", 1)
    y_local = _dev_tensor(y_local, "y_local", 1)
    if x_local.shape[0] != A.nrows or y_local.shape[0] != A.nrows:
        raise ValueError("dist_spmv: x_local and y_local must have the local row count")
    ng = A.ncols - A.nrows
    if ng > 0:
        if x_ghost is None:
            raise RuntimeError("dist_spmv: this share has ghost columns; the halo exchange "
                               "needs the other ranks - pass x_ghost")
        x_ghost = _dev_tensor(x_ghost, "x_ghost", 1)
        if x_ghost.shape[0] != ng:
            raise ValueError(f"dist_spmv: x_ghost must have {ng} entries")
    st = _stream()
    gp = _ptr(x_ghost) if ng > 0 else C.c_void_p(0)
    if overlap:
        for b, e in A_dist.interior_runs:
            call("num_spmv_rows", A._h, int(b), int(e), _ptr(x_local), _ptr(y_local), st)
        for b, e in A_dist.boundary_runs:
            call("num_spmv_split", A._h, int(b), int(e), _ptr(x_local), gp, _ptr(y_local), st)
    elif A.nrows > 0:
        call("num_spmv_split", A._h, 0, int(A.nrows), _ptr(x_local), gp, _ptr(y_local), st)
    if sync:
        _sync()


def halo_pack(A_dist: DistCrsMatrix, x_local, send_buf=None, *, sync: bool = True):
    """send_buf[i] = x_local[send_idx[i]] (halo_exchange.hpp:222-235)."""
    x_local = _dev_tensor(x_local, "x_local", 1)
    n = int(A_dist.send_idx.shape[0])
    if send_buf is None:
        send_buf = torch.empty(n, dtype=torch.float64, device=x_local.device)
    send_buf = _dev_tensor(send_buf, "send_buf", 1)
    if send_buf.shape[0] < n:
        raise ValueError("halo_pack: send_buf too small")
    call("num_halo_pack", _ptr(A_dist.send_idx), n, _ptr(x_local), _ptr(send_buf), _stream())
    if sync:
        _sync()
    return send_buf


def send_ranges(A_dist: DistCrsMatrix):
    """[(neighbour rank, offset, count)] of each neighbour's part of the packed
    send buffer, neighbours ascending (halo_exchange.hpp:154-191)."""
    out, off = [], 0
    for r, ids in A_dist.send:
        out.append((int(r), off, int(ids.shape[0])))
        off += int(ids.shape[0])
    return out


def halo_scatter(A_dist: DistCrsMatrix, recv_buf, x_ghost, src_rank: int, *,
                 sync: bool = True) -> None:
    """x_ghost[offset + i] = recv_buf[i] for the ghost slots owned by
    ``src_rank`` (halo_scatter, halo_exchange.hpp:287-300): the values that
    rank sent, in its send order, which is this rank's ghost order for that
    owner (ascending global id)."""
    recv_buf = _dev_tensor(recv_buf, "recv_buf", 1)
    x_ghost = _dev_tensor(x_ghost, "x_ghost", 1)
    if x_ghost.shape[0] != A_dist.n_ghost:
        raise ValueError(f"halo_scatter: x_ghost must have {A_dist.n_ghost} entries")
    for r, off, cnt in A_dist.recv:
        if int(r) == int(src_rank):
            break
    else:
        raise ValueError(f"halo_scatter: rank {src_rank} owns no ghost of this share")
    if recv_buf.shape[0] < cnt:
        raise ValueError(f"halo_scatter: recv_buf holds {recv_buf.shape[0]} values, "
                         f"rank {src_rank} sends {cnt}")
    call("num_halo_scatter", _ptr(recv_buf), int(cnt), _ptr(x_ghost), int(off), _stream())
    if sync:
        _sync()


# ------------------------------------------------- host round-trip helpers --

def spmv_host(rowptr, colind, values, nrows: int, ncols: int, x) -> np.ndarray:
    """Upload, y = A x, download (spmv_host_roundtrip, spmv_kernel.hpp:186-217):
    host arrays in, host array out, through num_spmv_host."""
    _ensure_init()
    rp = _host_array(rowptr, "rowptr", np.int32)
    ci = _host_array(colind, "colind", np.int32)
    vv = _host_array(values, "values", np.float64)
    xx = _host_array(x, "x", np.float64)
    if rp.shape[0] != int(nrows) + 1 or xx.shape[0] != int(ncols):
        raise ValueError("spmv_host: rowptr must hold nrows+1 entries and x ncols")
    y = np.empty(int(nrows), np.float64)
    call("num_spmv_host", rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
         vv.ctypes.data_as(C.c_void_p), int(nrows), int(ncols), xx.ctypes.data_as(C.c_void_p),
         y.ctypes.data_as(C.c_void_p))
    return y


def batch_spmv_host(rowptr, colind, values, nrows: int, ncols: int, X) -> np.ndarray:
    """batch_spmv_host_roundtrip (batch_spmv_kernel.hpp:179-209): X (ncols, k)
    row-major on the host in, Y (nrows, k) out, through num_batch_spmv_host."""
    _ensure_init()
    rp = _host_array(rowptr, "rowptr", np.int32)
    ci = _host_array(colind, "colind", np.int32)
    vv = _host_array(values, "values", np.float64)
    X = np.ascontiguousarray(X, np.float64)
    if X.ndim != 2 or X.shape[0] != int(ncols) or rp.shape[0] != int(nrows) + 1:
        raise ValueError("batch_spmv_host: X must be (ncols, k), rowptr nrows+1 entries")
    k = X.shape[1]
    Y = np.empty((int(nrows), k), np.float64)
    call("num_batch_spmv_host", rp.ctypes.data_as(C.c_void_p), ci.ctypes.data_as(C.c_void_p),
         vv.ctypes.data_as(C.c_void_p), int(nrows), int(ncols), int(k),
         X.ctypes.data_as(C.c_void_p), Y.ctypes.data_as(C.c_void_p))
    return Y


def cg_solve(A_dist: DistCrsMatrix, b, x, tol: float = 1e-10, max_iter: int = 1000) -> dict:
    """Distributed CG (bindings.cpp:239-270).  On one process (size 1) this is
    cg_solve_local on the whole matrix; a share of a multi-rank partition needs
    the other ranks' halo and allreduce and raises RuntimeError."""
    if A_dist.size != 1:
        raise RuntimeError("cg_solve: the share of a multi-rank partition needs the halo "
                           "exchange and allreduce of the other ranks")
    return cg_solve_local(A_dist.local, b, x, tol, max_iter)
