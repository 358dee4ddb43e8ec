"""Deterministic CSR generators for the workloads this package is measured on.

All matrices are int32-indexed CSR with float64 values and, unless stated,
columns sorted within each row.  Random draws use ``numpy.random.default_rng``
(the reference's convention, python/matrix_generators.py:98).

* ``laplacian_2d(L)``: 5-point Dirichlet Laplacian, diagonal 4, off-diagonals
  -1, row-major (x fastest) -- the closed form of
  python/matrix_generators.py:16-37 (scipy's ``diags`` sum stores no zeros).
* ``laplacian_3d(L)``: 7-point Dirichlet Laplacian, diagonal 6, off-diagonals
  -1 -- python/matrix_generators.py:42-80.
* ``stencil(dims, shift)``: the same stencils on a rectangular grid with a
  diagonal shift (batched / served workloads).
* ``zipf_matrix``: irregular power-law rows for the row-binned SpMV.
* ``diffusion_3d``: variable-coefficient 7-point diffusion, SPD.
* ``slab_share``: one rank's share of the 1-D row partition of
  src/dist_spmv.hpp:27-100 with the ghost ordering of
  src/halo_exchange.hpp:62-75,106-128.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


@dataclass
class CSR:
    rowptr: np.ndarray  # int32 [nrows+1]
    colind: np.ndarray  # int32 [nnz]
    values: np.ndarray  # float64 [nnz]
    nrows: int
    ncols: int
    meta: dict = field(default_factory=dict)

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.values, self.colind, self.rowptr),
                             shape=(self.nrows, self.ncols))


def _grid_stencil(dims, diag_of_row, off_val_fn=None):
    """Generic Dirichlet stencil on a grid ``dims`` = (nx, ny[, nz]), x fastest.

    Neighbour order per row is by column index: -z, -y, -x, self, +x, +y, +z.
    ``diag_of_row``: float64 array [n] with the diagonal; ``off_val_fn(k, rows)``
    returns the off-diagonal values for offset k (default -1).
    """
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    strides = [1]
    for d in dims[:-1]:
        strides.append(strides[-1] * d)
    idx = np.arange(n, dtype=np.int64)
    coords = []
    rem = idx
    for d in dims:
        coords.append(rem % d)
        rem = rem // d
    # offsets sorted by column: negative high strides first
    offs = []
    for a in reversed(range(len(dims))):
        offs.append((a, -1))
    offs.append((None, 0))
    for a in range(len(dims)):
        offs.append((a, +1))
    masks = []
    for a, s in offs:
        if a is None:
            masks.append(np.ones(n, dtype=bool))
        elif s < 0:
            masks.append(coords[a] > 0)
        else:
            masks.append(coords[a] < dims[a] - 1)
    counts = np.zeros(n, dtype=np.int64)
    for m in masks:
        counts += m
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    nnz = int(rowptr[-1])
    colind = np.empty(nnz, dtype=np.int32)
    values = np.empty(nnz, dtype=np.float64)
    before = np.zeros(n, dtype=np.int64)
    for k, ((a, s), m) in enumerate(zip(offs, masks)):
        rows = idx[m]
        pos = rowptr[:-1][m] + before[m]
        if a is None:
            colind[pos] = rows
            values[pos] = diag_of_row[m]
        else:
            colind[pos] = rows + s * strides[a]
            values[pos] = -1.0 if off_val_fn is None else off_val_fn(k, a, s, m)
        before += m
    del before
    return CSR(rowptr.astype(np.int32), colind, values, n, n)


def stencil_rowptr(dims) -> np.ndarray:
    """rowptr of the Dirichlet stencil on ``dims`` (x fastest): row degree is
    1 + sum over axes of (has lower neighbour) + (has upper neighbour)."""
    dims = tuple(int(d) for d in dims)
    deg = np.ones((1,) * len(dims), dtype=np.int32)
    for a, d in enumerate(dims):
        c = np.arange(d)
        da = (c > 0).astype(np.int32) + (c < d - 1).astype(np.int32)
        shape = [1] * len(dims)
        shape[len(dims) - 1 - a] = d
        deg = deg + da.reshape(shape)
    rowptr = np.zeros(int(np.prod(dims)) + 1, dtype=np.int64)
    np.cumsum(deg.reshape(-1), out=rowptr[1:])
    return rowptr.astype(np.int32)


def laplacian_2d(L: int) -> CSR:
    """5-point Dirichlet Laplacian on an L x L grid: n = L^2, nnz = 5n - 4L."""
    A = _grid_stencil((L, L), np.full(L * L, 4.0))
    A.meta = {"kind": "laplacian_2d", "L": L}
    return A


def laplacian_3d(L: int) -> CSR:
    """7-point Dirichlet Laplacian on an L^3 grid: n = L^3, nnz = 7n - 6L^2."""
    A = _grid_stencil((L, L, L), np.full(L ** 3, 6.0))
    A.meta = {"kind": "laplacian_3d", "L": L}
    return A


def stencil(dims, shift: float = 0.0) -> CSR:
    """2-D 5-point / 3-D 7-point Dirichlet Laplacian on ``dims`` plus ``shift``*I."""
    nd = len(dims)
    base = 4.0 if nd == 2 else 6.0
    if nd == 1:
        base = 2.0
    n = int(np.prod(dims))
    A = _grid_stencil(dims, np.full(n, base + float(shift)))
    A.meta = {"kind": f"stencil_{nd}d", "dims": tuple(int(d) for d in dims), "shift": float(shift)}
    return A


def laplacian_1d(N: int) -> CSR:
    """Tridiagonal [-1, 2, -1] (tests/test_cg.cpp:76-92)."""
    A = _grid_stencil((N,), np.full(N, 2.0))
    A.meta = {"kind": "laplacian_1d", "N": N}
    return A


def diffusion_3d(L: int, sigma: float = 1.0, seed: int = 3) -> CSR:
    """Variable-coefficient 7-point diffusion on an L^3 grid, Dirichlet walls.

    Face coefficients are lognormal(0, sigma), drawn in this order from
    ``default_rng(seed)``: x-faces shape (L, L, L+1) [z, y, face-x], then
    y-faces (L, L+1, L), then z-faces (L+1, L, L).  Row i: A_ii = sum of its six
    face coefficients (wall faces included), A_ij = -c(face i|j).  Symmetric,
    strictly diagonally dominant in rows touching a wall, irreducible -> SPD.
    """
    rng = np.random.default_rng(seed)
    cx = rng.lognormal(0.0, sigma, size=(L, L, L + 1))
    cy = rng.lognormal(0.0, sigma, size=(L, L + 1, L))
    cz = rng.lognormal(0.0, sigma, size=(L + 1, L, L))
    # face coefficients on the low / high side of each cell, flattened (z,y,x)
    xlo = cx[:, :, :L].reshape(-1)
    xhi = cx[:, :, 1:].reshape(-1)
    ylo = cy[:, :L, :].reshape(-1)
    yhi = cy[:, 1:, :].reshape(-1)
    zlo = cz[:L, :, :].reshape(-1)
    zhi = cz[1:, :, :].reshape(-1)
    diag = xlo + xhi + ylo + yhi + zlo + zhi
    face = {(2, -1): zlo, (1, -1): ylo, (0, -1): xlo, (0, 1): xhi, (1, 1): yhi, (2, 1): zhi}

    def offval(k, a, s, m):
        return -face[(a, s)][m]

    A = _grid_stencil((L, L, L), diag, offval)
    A.meta = {"kind": "diffusion_3d", "L": L, "sigma": sigma, "seed": seed}
    return A


def zipf_matrix(n: int = 4_000_000, a: float = 2.0, kmax: int = 4096, seed: int = 2) -> CSR:
    """Irregular matrix: row nnz = clip(zipf(a), 1, kmax), columns uniform
    without replacement and sorted, values uniform(-1, 1).

    Draw order from ``default_rng(seed)``: row lengths (n), then one int64
    column draw per non-zero; duplicate columns inside a row are redrawn (in
    row-major order of the duplicates) until none remain; then the values.
    """
    rng = np.random.default_rng(seed)
    k = np.clip(rng.zipf(a, size=n), 1, kmax).astype(np.int64)
    k = np.minimum(k, n)
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(k, out=rowptr[1:])
    nnz = int(rowptr[-1])
    rows = np.repeat(np.arange(n, dtype=np.int64), k)
    cols = rng.integers(0, n, size=nnz, dtype=np.int64)
    while True:
        key = rows * n + cols
        order = np.argsort(key, kind="stable")
        ks = key[order]
        dup = np.zeros(nnz, dtype=bool)
        dup[order[1:][ks[1:] == ks[:-1]]] = True
        nd = int(dup.sum())
        if nd == 0:
            cols = cols[order]
            break
        cols[dup] = rng.integers(0, n, size=nd, dtype=np.int64)
    del rows
    vals = rng.uniform(-1.0, 1.0, size=nnz)
    A = CSR(rowptr.astype(np.int32), cols.astype(np.int32), vals, n, n)
    A.meta = {"kind": "zipf", "n": n, "a": a, "kmax": kmax, "seed": seed}
    return A


def lcg_matrix(N: int = 10_000, nnz_per_row: int = 7, seed: int = 42):
    """The LCG matrix and x of tests/test_spmv_kernel.cpp:163-193 (unsorted,
    possibly repeated columns).  Returns (CSR, x)."""
    s = seed

    def lcg():
        nonlocal s
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        return s

    x = np.array([(lcg() % 1000) * 0.001 - 0.5 for _ in range(N)], dtype=np.float64)
    rowptr = [0]
    colind = []
    vals = []
    for i in range(N):
        colind.append(i)
        vals.append(float(nnz_per_row))
        for _ in range(1, nnz_per_row):
            j = lcg() % N
            if j != i:
                colind.append(j)
                vals.append(-1.0)
        rowptr.append(len(colind))
    A = CSR(np.array(rowptr, np.int32), np.array(colind, np.int32), np.array(vals, np.float64),
            N, N)
    A.meta = {"kind": "lcg", "N": N, "seed": seed}
    return A, x


# ---------------------------------------------------------------------------
# One rank's share of the 1-D row partition (src/dist_spmv.hpp:27-100)
# ---------------------------------------------------------------------------

def row_split(global_nrows: int, rank: int, size: int):
    """Even split of src/dist_spmv.hpp:38-43: first ``extra`` ranks get one more."""
    base, extra = divmod(global_nrows, size)
    start = rank * base + min(rank, extra)
    end = start + base + (1 if rank < extra else 0)
    return start, end


@dataclass
class SlabShare:
    local: CSR                 # remapped local matrix, ncols = n_local + n_ghost
    global_nrows: int
    local_row_start: int
    local_row_end: int
    ghost_ids: np.ndarray      # global ids of ghost slots, in slot order
    ghost_owner: np.ndarray    # owning rank of each ghost slot
    recv: list                 # [(rank, offset, count)] in slot order
    interior_rows: np.ndarray  # local rows touching no ghost column
    boundary_rows: np.ndarray  # local rows touching at least one ghost column


def slab_share(rowptr, colind, values, global_nrows: int, global_ncols: int, rank: int,
               size: int) -> SlabShare:
    """Build rank ``rank``'s local matrix exactly as distribute_csr does.

    * rows [start, end) by the even split (dist_spmv.hpp:38-43);
    * ghost columns grouped by owner rank ascending, each group in ascending
      global id (std::map/std::set order, halo_exchange.hpp:62-75,106-128);
    * own columns -> col - start, ghost columns -> n_local + slot
      (dist_spmv.hpp:62-82); column order inside a row is NOT re-sorted.
    Raises RuntimeError for a column outside every rank's rows
    (halo_exchange.hpp:57-59).
    """
    rowptr = np.asarray(rowptr)
    colind = np.asarray(colind)
    values = np.asarray(values)
    start, end = row_split(global_nrows, rank, size)
    n_local = end - start
    j0, j1 = int(rowptr[start]), int(rowptr[end])
    lc = colind[j0:j1].astype(np.int64)
    lv = np.array(values[j0:j1], dtype=np.float64, copy=True)
    lr = (rowptr[start:end + 1].astype(np.int64) - j0)
    if lc.size and (lc.min() < 0 or lc.max() >= global_nrows):
        bad = int(lc[(lc < 0) | (lc >= global_nrows)][0])
        raise RuntimeError(f"build_ghost_map: column index {bad} is out of the global row range")
    is_ghost = (lc < start) | (lc >= end)
    ghost_ids = np.unique(lc[is_ghost])  # ascending global id
    # owner of each ghost (vectorised even-split inverse)
    starts = np.array([row_split(global_nrows, r, size)[0] for r in range(size)] + [global_nrows])
    owner = np.searchsorted(starts, ghost_ids, side="right") - 1
    # group by owner ascending, ids ascending inside a group: ids sorted and
    # owners monotone in id, so ascending id order already is that order.
    slot = np.searchsorted(ghost_ids, lc[is_ghost])
    new_c = np.empty_like(lc)
    new_c[~is_ghost] = lc[~is_ghost] - start
    new_c[is_ghost] = n_local + slot
    recv = []
    off = 0
    for r in np.unique(owner):
        cnt = int((owner == r).sum())
        recv.append((int(r), off, cnt))
        off += cnt
    row_of = np.repeat(np.arange(n_local), np.diff(lr))
    touches = np.zeros(n_local, dtype=bool)
    touches[row_of[is_ghost]] = True
    local = CSR(lr.astype(np.int32), new_c.astype(np.int32), lv, n_local,
                n_local + len(ghost_ids))
    local.meta = {"kind": "slab", "rank": rank, "size": size}
    return SlabShare(local, global_nrows, start, end, ghost_ids.astype(np.int64),
                     owner.astype(np.int32), recv,
                     np.nonzero(~touches)[0].astype(np.int32),
                     np.nonzero(touches)[0].astype(np.int32))


# ---------------------------------------------------------------------------
# Batched and served workloads
# ---------------------------------------------------------------------------

def rect_dims(n_target: float, aspect: float):
    nx = max(1, int(round(math.sqrt(n_target * aspect))))
    ny = max(1, int(round(n_target / nx)))
    return nx, ny


def batched_systems(count: int = 4096, nmin: int = 64, nmax: int = 65536, seed: int = 4):
    """4096 independent SPD systems: 2-D 5-point Laplacians on random rectangles
    (aspect uniform [1,4]) with n log-uniform in [nmin, nmax], plus a diagonal
    shift uniform[0.01, 1.0]; b uniform(-1, 1).  Draw order from
    ``default_rng(seed)``: log n (count), aspect (count), shift (count), then
    each system's b in system order.  Returns a list of (dims, shift, b)."""
    rng = np.random.default_rng(seed)
    logn = rng.uniform(math.log(nmin), math.log(nmax), size=count)
    aspect = rng.uniform(1.0, 4.0, size=count)
    shift = rng.uniform(0.01, 1.0, size=count)
    out = []
    for i in range(count):
        dims = rect_dims(math.exp(logn[i]), aspect[i])
        n = dims[0] * dims[1]
        b = rng.uniform(-1.0, 1.0, size=n)
        out.append((dims, float(shift[i]), b))
    return out


# ---- the reference's other matrix families (python/matrix_generators.py) --
# Restated for the benchmark harness and breadth tests of the row binning.
# lattice_gauge draws its phases in the reference's order, so with the same
# seed it builds the same matrix; random_sparse and power_law repeat the
# reference's draw sequence call for call.

def _to_csr(A) -> CSR:
    A = A.tocsr()
    A.sum_duplicates()
    A.sort_indices()
    return CSR(A.indptr.astype(np.int32), A.indices.astype(np.int32),
               A.data.astype(np.float64), A.shape[0], A.shape[1])


def lattice_gauge(L: int, mass: float = 0.1, seed: int = 42) -> CSR:
    """Scalar 2-D lattice 'Dirac' operator on an L x L torus with random U(1)
    link phases, symmetrised (python/matrix_generators.py:85-139): diagonal
    4 + mass, hopping -0.5 Re(exp(+-i phase)) to the four periodic
    neighbours, then A = (A + A^T) / 2.  Draws: phi_x (L, L), then phi_y."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    phi_x = rng.uniform(0, 2 * np.pi, (L, L))
    phi_y = rng.uniform(0, 2 * np.pi, (L, L))
    N = L * L
    y, x = np.divmod(np.arange(N), L)

    def idx(xx, yy):
        return (yy % L) * L + (xx % L)
    i = np.arange(N)
    # per site, in the reference's append order: self, +x, -x, +y, -y
    rows = np.stack([i, i, i, i, i], axis=1).reshape(-1)
    cols = np.stack([i, idx(x + 1, y), idx(x - 1, y), idx(x, y + 1), idx(x, y - 1)],
                    axis=1).reshape(-1)
    vals = np.stack([
        np.full(N, 4.0 + mass),
        -0.5 * np.exp(1j * phi_x[y, x]).real,
        -0.5 * np.exp(-1j * phi_x[y, (x - 1) % L]).real,
        -0.5 * np.exp(1j * phi_y[y, x]).real,
        -0.5 * np.exp(-1j * phi_y[(y - 1) % L, x]).real,
    ], axis=1).reshape(-1)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N))
    return _to_csr(0.5 * (A + A.T))


def random_sparse(N: int, nnz_per_row: int = 10, seed: int = 0, symmetric: bool = True) -> CSR:
    """Random sparse SPD matrix (python/matrix_generators.py:144-170): row i
    draws min(nnz_per_row - 1, i) distinct columns j < i, each with value
    -U(0.1, 1) (mirrored when symmetric); the diagonal is the absolute row
    sum + 1 (strict diagonal dominance)."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for i in range(N):
        n_off = min(nnz_per_row - 1, i)
        if n_off > 0:
            js = rng.choice(i, size=n_off, replace=False)
            for j in js:
                v = rng.uniform(0.1, 1.0)
                rows.append(i); cols.append(int(j)); vals.append(-v)
                if symmetric:
                    rows.append(int(j)); cols.append(i); vals.append(-v)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(N, N), dtype=np.float64)
    diag = np.asarray(abs(A).sum(axis=1)).ravel() + 1.0
    return _to_csr(A + sp.diags(diag))


def power_law(N: int, m: int = 3, seed: int = 1) -> CSR:
    """The reference's 'graph Laplacian' of a Barabasi-Albert preferential-
    attachment graph (python/matrix_generators.py:175-221): a clique of m + 1
    nodes, then each new node links to m distinct earlier nodes drawn with
    probability proportional to degree.  The reference builds the adjacency
    with -1 entries and returns diag(degree) - that, i.e. D + Adj (the
    signless Laplacian), which is reproduced here: symmetric, weakly
    diagonally dominant, positive semi-definite."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    edges = set()
    degrees = np.zeros(N, dtype=np.int64)
    first = min(m + 1, N)
    for i in range(first):
        for j in range(i):
            edges.add((i, j))
            degrees[i] += 1
            degrees[j] += 1
    for node in range(first, N):
        deg_sum = degrees[:node].sum()
        if deg_sum == 0:
            targets = rng.choice(node, size=min(m, node), replace=False)
        else:
            targets = rng.choice(node, size=min(m, node), replace=False,
                                 p=degrees[:node] / deg_sum)
        for t in targets:
            edges.add((node, int(t)))
            degrees[node] += 1
            degrees[t] += 1
    e = np.array(sorted(edges), dtype=np.int64).reshape(-1, 2)
    rows = np.concatenate([e[:, 0], e[:, 1]])
    cols = np.concatenate([e[:, 1], e[:, 0]])
    A = sp.csr_matrix((-np.ones(rows.size), (rows, cols)), shape=(N, N), dtype=np.float64)
    d = np.asarray(abs(A).sum(axis=1)).ravel()
    return _to_csr(sp.diags(d) - A)


MATRIX_FAMILIES = {
    "laplacian_2d": lambda N: laplacian_2d(int(N ** 0.5)),
    "laplacian_3d": lambda N: laplacian_3d(int(round(N ** (1 / 3)))),
    "lattice_gauge": lambda N: lattice_gauge(int(N ** 0.5)),
    "random_sparse": lambda N: random_sparse(N),
    "power_law": lambda N: power_law(N),
}


def get_matrix(name: str, N: int) -> CSR:
    """The reference's get_matrix(name, N) (python/matrix_generators.py:226-258)."""
    if name not in MATRIX_FAMILIES:
        raise ValueError(f"Unknown matrix '{name}'. Choose from: {list(MATRIX_FAMILIES)}")
    return MATRIX_FAMILIES[name](N)


def batched_specs(count: int = 4096, nmin: int = 64, nmax: int = 65536, seed: int = 4):
    """The batched workload as request descriptions: the same draws as
    ``batched_systems`` for the sizes, aspects and shifts, then one seed per
    system (b = ``hash_uniform(n, seed_i)``, assembled on the device).
    Returns a list of (dims, shift, seed_i)."""
    rng = np.random.default_rng(seed)
    logn = rng.uniform(math.log(nmin), math.log(nmax), size=count)
    aspect = rng.uniform(1.0, 4.0, size=count)
    shift = rng.uniform(0.01, 1.0, size=count)
    seeds = rng.integers(0, 2**31 - 1, size=count)
    return [(rect_dims(math.exp(logn[i]), aspect[i]), float(shift[i]), int(seeds[i]))
            for i in range(count)]


def hash_uniform(n: int, seed: int, start: int = 0) -> np.ndarray:
    """Counter-based uniform draws in [-1, 1): element i is 2u - 1 with u the
    top 53 bits of SplitMix64 (Steele, Lea & Flood, OOPSLA 2014) applied to
    seed * 0xD1B54A32D192ED03 + (i + 1) * 0x9E3779B97F4A7C15 (mod 2^64).  The
    engine assembles a served request's right-hand side with the same
    function on the device, so host and device draws agree bit for bit."""
    i = np.arange(start, start + int(n), dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(int(seed) % 2**64) * np.uint64(0xD1B54A32D192ED03) + \
            (i + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * 2.0**-53
    return 2.0 * u - 1.0


def stencil_nnz(dims) -> int:
    """Non-zeros of the Dirichlet stencil on ``dims``: n + 2 (L_a - 1) n / L_a
    per axis."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    return n + sum(2 * (L - 1) * (n // L) for L in dims)


def served_requests(rate: float, duration: float, seed: int = 5, nmin: int = 1000,
                    nmax: int = 2_000_000, p2d: float = 0.7, max_requests: int | None = None):
    """Open-loop Poisson arrivals at ``rate`` req/s for ``duration`` s.  Each
    request is a 2-D (p=0.7) or 3-D Laplacian with shift uniform[0.001, 1] and
    n log-uniform in [nmin, nmax]; its b is ``hash_uniform(n, seed)`` of the
    request's own seed.  Returns a list of dicts {t, dims, shift, seed}."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        two_d = rng.uniform() < p2d
        n = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
        shift = rng.uniform(0.001, 1.0)
        if two_d:
            L = max(2, int(round(math.sqrt(n))))
            dims = (L, L)
        else:
            L = max(2, int(round(n ** (1.0 / 3.0))))
            dims = (L, L, L)
        reqs.append({"t": t, "dims": dims, "shift": float(shift),
                     "seed": int(rng.integers(0, 2**31 - 1))})
        if max_requests is not None and len(reqs) >= max_requests:
            break
    return reqs
