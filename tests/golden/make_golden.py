"""Write the small golden fixtures under tests/golden/ from the pinned oracle.

The reference stores no golden vectors (its tests generate every input at run
time from closed forms, seeded RNGs or its LCG).  These fixtures freeze the
oracle's outputs on the same generated inputs so both the oracle and the HIP
path are checked against fixed data:

  tridiag4.npz      4x4 tridiagonal KAT (tests/test_spmv_kernel.cpp:65-89),
                    y = A x and Y = A I4 (test [4])
  lcg10k.npz        LCG N=1e4 matrix (tests/test_spmv_kernel.cpp:163-193):
                    x, y = A x, and a digest of the CSR arrays
  lap2d64.npz       2-D 64^2 Laplacian: x ~ U(-1,1) seed 1 and y = A x;
                    CG with b ~ U(-1,1) seed 0, tol 1e-8: iters, residual, x
  slab16.npz        3-D 16^3 Laplacian, rank 1 of 4: ghost ids, send ids,
                    interior/boundary counts, y_local for x ~ U(-1,1) seed 5

Run:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from num_amd import generators as g  # noqa: E402


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def main():
    oracle.set_threads(1)
    rp = np.array([0, 2, 5, 8, 10], np.int32)
    ci = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3], np.int32)
    v = np.array([2, -1, -1, 2, -1, -1, 2, -1, -1, 2], np.float64)
    x = np.array([1.0, 2, 3, 4])
    np.savez_compressed(os.path.join(HERE, "tridiag4.npz"), rowptr=rp, colind=ci, values=v,
                        x=x, y=oracle.spmv(rp, ci, v, x),
                        Y=oracle.batch_spmv(rp, ci, v, np.eye(4)))

    A, x = g.lcg_matrix()
    np.savez_compressed(os.path.join(HERE, "lcg10k.npz"), x=x,
                        y=oracle.spmv(A.rowptr, A.colind, A.values, x),
                        digest=np.array(digest(A.rowptr, A.colind, A.values)))

    A = g.laplacian_2d(64)
    x = np.random.default_rng(1).uniform(-1, 1, A.ncols)
    b = np.random.default_rng(0).uniform(-1, 1, A.nrows)
    xs, r = oracle.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-8, max_iter=2000)
    np.savez_compressed(os.path.join(HERE, "lap2d64.npz"), x=x,
                        y=oracle.spmv(A.rowptr, A.colind, A.values, x), cg_x=xs,
                        cg_iters=np.array(r["iters"]), cg_residual=np.array(r["residual"]),
                        digest=np.array(digest(A.rowptr, A.colind, A.values)))

    A = g.laplacian_3d(16)
    sh = g.slab_share(A.rowptr, A.colind, A.values, A.nrows, A.ncols, 1, 4)
    xg = np.random.default_rng(5).uniform(-1, 1, A.ncols)
    xl = xg[sh.local_row_start:sh.local_row_end]
    x_ext = np.concatenate([xl, xg[sh.ghost_ids]])
    L = sh.local
    y_local = oracle.spmv(L.rowptr, L.colind, L.values, x_ext)
    np.savez_compressed(os.path.join(HERE, "slab16.npz"), ghost_ids=sh.ghost_ids,
                        n_interior=np.array(len(sh.interior_rows)),
                        n_boundary=np.array(len(sh.boundary_rows)),
                        start=np.array(sh.local_row_start), end=np.array(sh.local_row_end),
                        x_global=xg, y_local=y_local,
                        digest=np.array(digest(L.rowptr, L.colind, L.values)))
    oracle.set_threads(0)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
