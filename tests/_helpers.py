"""Shared checks for the GPU tests.

``stencil_residual`` computes ||b - A x|| / ||b|| for the shifted Dirichlet
stencil systems of the batched and served workloads directly from the
stencil's closed form (diagonal 2*ndim + shift, -1 to each grid neighbour
inside the box, x fastest), with torch tensor arithmetic in float64.  It
never touches the library's matrices or kernels, so it is an independent
check of a solution the engine returns.

``shares_spmv`` / ``shares_cg`` run the reference's distributed SpMV and CG
over every rank's share in turn on one GPU (halo pack, scatter, share SpMV;
global dots as sums of the shares' dots), for the counterparts of the
reference's multi-rank tests.
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


def shares_spmv(ks, D, v, out, xg):
    """out[q] = share q's rows of A v on one GPU: every share packs its halo
    (halo_pack), each neighbour's part is scattered into share q's ghost
    block (halo_scatter), then share q's SpMV (dist_spmv); the transport
    between pack and scatter is the part a second GPU would carry."""
    size = len(D)
    bufs = [ks.halo_pack(D[s], v[s]) for s in range(size)]
    for q in range(size):
        for src, _, _ in D[q].recv:
            off, cnt = next((o, c) for dst, o, c in ks.send_ranges(D[src]) if dst == q)
            ks.halo_scatter(D[q], bufs[src][off:off + cnt].contiguous(), xg[q], src_rank=src)
        ks.dist_spmv(D[q], v[q], out[q], xg[q])


def shares_cg(ks, D, b, tol, max_iter):
    """The distributed CG of cg_solver.hpp:38-125 over every share of D in
    turn on one GPU: r = b - A x0 (x0 = 0), p = r, ||b|| (0 -> 1), then per
    iteration A p by shares_spmv, alpha = (r,r) / (p,Ap) with the guard
    |p.Ap| < 1e-300, x += alpha p, r -= alpha Ap, the strict test
    ||r|| / ||b|| < tol, beta, p = r + beta p; every global dot is the sum of
    the shares' dots.  Returns (gathered x, iterations, converged)."""
    size = len(D)
    nl = [d.local_row_end - d.local_row_start for d in D]
    vec = lambda n: torch.zeros(n, dtype=torch.float64, device="cuda")
    bl = [torch.from_numpy(np.ascontiguousarray(b[d.local_row_start:d.local_row_end])).cuda()
          for d in D]
    x, r, p, Ap = ([vec(n) for n in nl] for _ in range(4))
    xg = [vec(d.n_ghost) for d in D]
    gdot = lambda u, v: sum(ks.dot(u[s], v[s]) for s in range(size))
    shares_spmv(ks, D, x, Ap, xg)
    for s in range(size):
        ks.axpy(-1.0, Ap[s], bl[s], r[s])
        ks.copy(r[s], p[s])
    bnorm = float(np.sqrt(gdot(bl, bl))) or 1.0
    rr = gdot(r, r)
    gather = lambda: np.concatenate([v.cpu().numpy() for v in x])
    if np.sqrt(rr) / bnorm < tol:
        return gather(), 0, True
    for it in range(max_iter):
        shares_spmv(ks, D, p, Ap, xg)
        pAp = gdot(p, Ap)
        if abs(pAp) < 1e-300:
            raise RuntimeError("CG: p'Ap is near zero")
        alpha = rr / pAp
        for s in range(size):
            ks.axpy(alpha, p[s], x[s], x[s])
            ks.axpy(-alpha, Ap[s], r[s], r[s])
        rr_new = gdot(r, r)
        if np.sqrt(rr_new) / bnorm < tol:
            return gather(), it + 1, True
        beta, rr = rr_new / rr, rr_new
        for s in range(size):
            ks.axpy2(1.0, r[s], beta, p[s], p[s])
    return gather(), max_iter, False
