"""Shared checks for the GPU tests.

``stencil_residual`` computes ||b - A x|| / ||b|| for the shifted Dirichlet
stencil systems of the batched and served workloads directly from the
stencil's closed form (diagonal 2*ndim + shift, -1 to each grid neighbour
inside the box, x fastest), with torch tensor arithmetic in float64.  It
never touches the library's matrices or kernels, so it is an independent
check of a solution the engine returns.
"""
from __future__ import annotations

import numpy as np
import torch

from num_amd import generators as g


def stencil_apply(x: torch.Tensor, dims, shift: float) -> torch.Tensor:
    """A x for the Dirichlet stencil on ``dims`` (x fastest), float64."""
    dims = tuple(int(d) for d in dims)
    v = x.reshape(tuple(reversed(dims)))  # (nz, ny, nx) / (ny, nx) / (nx,)
    y = (2.0 * len(dims) + shift) * v
    for ax in range(v.dim()):
        n = v.shape[ax]
        if n < 2:
            continue
        lo = v.narrow(ax, 0, n - 1)
        hi = v.narrow(ax, 1, n - 1)
        y.narrow(ax, 1, n - 1).sub_(lo)   # left / lower neighbour
        y.narrow(ax, 0, n - 1).sub_(hi)   # right / upper neighbour
    return y.reshape(-1)


def stencil_residual(x: torch.Tensor, dims, shift: float, seed: int) -> float:
    """FP64 relative residual of x for the system with b = hash_uniform(n, seed)."""
    n = int(np.prod(dims))
    b = torch.from_numpy(g.hash_uniform(n, seed)).to(x.device)
    r = b - stencil_apply(x.to(torch.float64), dims, shift)
    return float(torch.linalg.norm(r) / torch.linalg.norm(b))


def shuffled_stencil_rows(nx, ny, seed):
    """The 2-D 5-point stencil (diagonal 4.5, -1 to each neighbour) with
    each grid line's rows stored in one of two shuffled entry orders (by the
    line's parity) and the diagonal split into two duplicate entries on every
    third grid column: unsorted and duplicate columns, which the CSR contract
    allows (the LCG known-answer matrix has both), in few enough distinct
    rows that a tile keeps at most 16 row patterns.  Returns (rowptr,
    colind, values, n)."""
    rng = np.random.default_rng(seed)
    n = nx * ny
    orders = [rng.permutation(5) for _ in range(2)]
    rp, ci, va = [0], [], []
    for r in range(n):
        ix, iy = r % nx, r // nx
        ent = [(r - nx, -1.0) if iy > 0 else None, (r - 1, -1.0) if ix > 0 else None,
               (r, 4.5), (r + 1, -1.0) if ix < nx - 1 else None,
               (r + nx, -1.0) if iy < ny - 1 else None]
        row = []
        for k in orders[iy % 2]:
            if ent[k] is None:
                continue
            c, v = ent[k]
            if c == r and ix % 3 == 0:
                row += [(c, 0.5 * v), (c, 0.5 * v)]
            else:
                row.append((c, v))
        ci += [c for c, _ in row]
        va += [v for _, v in row]
        rp.append(len(ci))
    return (np.array(rp, np.int32), np.array(ci, np.int32), np.array(va, np.float64), n)
