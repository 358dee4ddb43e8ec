"""The reference's Python accuracy suite (tests/test_accuracy.py) on the HIP
path, at its sizes and with its matrix families, inputs and tolerances, plus
the oracle as a second, tighter reference:

1. single-GPU SpMV, N = 50,000, laplacian_2d / laplacian_3d / random_sparse /
   lattice_gauge, x standard normal from default_rng(0) drawn in family
   order, both kokkoskernels settings: rel. error < 1e-10 against SciPy
   (test_accuracy.py:85-119); against the oracle bit-identical on rows of
   <= 32 entries, else within 1e-12 of sum |a_ij x_j|;
2. batch SpMV, N = 20,000, k = 64, three families, X standard normal from
   default_rng(1): < 1e-10 against SciPy (:124-155), bit-identical to the
   oracle;
3. cg_solve_local, N = 10,000, laplacian_3d / random_sparse, b = 1,
   tol 1e-10: solution within 1e-6 of SciPy CG at rtol 1e-12 (:160-195),
   iterations within +-2 % of the oracle's;
4. distributed SpMV, N = 40,000, laplacian_2d / laplacian_3d, x standard
   normal from default_rng(2), every rank's share of 2- and 4-rank
   partitions run on this GPU and gathered: < 1e-10 against SciPy
   (:200-250), bit-identical to the oracle.
The distributed CG (:255-310) needs the other ranks' halo exchange and
allreduce inside every iteration, which one GPU cannot run.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla
import torch

from num_amd import generators as g

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _rel_err(a, b):
    d = float(np.linalg.norm(b))
    return float(np.linalg.norm(a)) if d == 0.0 else float(np.linalg.norm(a - b) / d)


def _up(ks, A):
    return ks.upload_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols)


@pytest.fixture(scope="module")
def spmv_cases():
    # x drawn from one generator in the reference's family order (:92-99)
    rng = np.random.default_rng(0)
    out = []
    for name in ("laplacian_2d", "laplacian_3d", "random_sparse", "lattice_gauge"):
        A = g.get_matrix(name, 50_000)
        out.append((name, A, rng.standard_normal(A.ncols)))
    return out


def test_spmv_single_gpu_50k(ks, orc, spmv_cases):
    for name, A, x in spmv_cases:
        print(f"accuracy: spmv {name} n={A.nrows} nnz={A.nnz}", flush=True)
        y_sp = A.to_scipy() @ x
        y_or = orc.spmv(A.rowptr, A.colind, A.values, x)
        y_abs = orc.spmv(A.rowptr, A.colind, np.abs(A.values), np.abs(x))
        short = np.diff(A.rowptr) <= 32
        dA = _up(ks, A)
        for kk in (True, False):
            y = torch.full((A.nrows,), float("nan"), dtype=torch.float64, device="cuda")
            ks.spmv(dA, _dev(x), y, kokkoskernels=kk)
            got = y.cpu().numpy()
            assert _rel_err(got, y_sp) < 1e-10, (name, kk)
            assert np.array_equal(got[short], y_or[short]), (name, kk)
            assert np.all(np.abs(got - y_or) <= 1e-12 * y_abs), (name, kk)


@pytest.mark.parametrize("name", ["laplacian_2d", "laplacian_3d", "random_sparse"])
def test_batch_spmv_20k_k64(ks, orc, name):
    A = g.get_matrix(name, 20_000)
    X = np.random.default_rng(1).standard_normal((A.ncols, 64))
    print(f"accuracy: batch_spmv k=64 {name} n={A.nrows}", flush=True)
    Y = torch.full((A.nrows, 64), float("nan"), dtype=torch.float64, device="cuda")
    ks.batch_spmv(_up(ks, A), _dev(X), Y)
    got = Y.cpu().numpy()
    assert _rel_err(got, np.asarray(A.to_scipy() @ X)) < 1e-10
    assert np.array_equal(got, orc.batch_spmv(A.rowptr, A.colind, A.values, X))


@pytest.mark.parametrize("name", ["laplacian_3d", "random_sparse"])
def test_cg_local_10k_vs_scipy(ks, orc, name):
    A = g.get_matrix(name, 10_000)
    b = np.ones(A.nrows)
    x_sp, info = spla.cg(A.to_scipy(), b, rtol=1e-12, maxiter=5000)
    assert info == 0
    print(f"accuracy: cg_solve_local {name} n={A.nrows}", flush=True)
    x = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
    res = ks.cg_solve_local(_up(ks, A), _dev(b), x, tol=1e-10, max_iter=5000)
    assert res["converged"]
    assert _rel_err(x.cpu().numpy(), x_sp) < 1e-6
    _, ro = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-10, max_iter=5000)
    assert abs(res["iters"] - ro["iters"]) <= max(1, 0.02 * ro["iters"]), (res, ro)


@pytest.mark.parametrize("name", ["laplacian_2d", "laplacian_3d"])
@pytest.mark.parametrize("size", [2, 4])
def test_distributed_spmv_40k_gathered(ks, orc, name, size):
    A = g.get_matrix(name, 40_000)
    x = np.random.default_rng(2).standard_normal(A.ncols)
    y_or = orc.spmv(A.rowptr, A.colind, A.values, x)
    dx = _dev(x)
    parts = []
    for r in range(size):
        D = ks.distribute_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols, rank=r,
                              size=size)
        s, e = D.local_row_start, D.local_row_end
        print(f"accuracy: dist_spmv {name} rank {r} of {size}", flush=True)
        yl = torch.full((e - s,), float("nan"), dtype=torch.float64, device="cuda")
        ks.dist_spmv(D, dx[s:e].contiguous(), yl, D.ghost_values(dx) if D.n_ghost else None)
        parts.append(yl.cpu().numpy())
    y = np.concatenate(parts)
    assert _rel_err(y, A.to_scipy() @ x) < 1e-10
    assert np.array_equal(y, y_or)


@pytest.mark.parametrize("size", [2, 4])
def test_distributed_cg_20k_vs_scipy(ks, size):
    """tests/test_accuracy.py:265-322 on one GPU: the distributed CG over
    every share in turn (laplacian_3d, N = 20k, b = 1, tol 1e-10, 5000
    iterations) against SciPy's CG at rtol 1e-12, solution within 1e-6."""
    import scipy.sparse.linalg as spla
    from _helpers import shares_cg
    A = g.get_matrix("laplacian_3d", 20_000)
    b = np.ones(A.nrows)
    x_ref, info = spla.cg(A.to_scipy(), b, rtol=1e-12, maxiter=5000)
    assert info == 0
    D = [ks.distribute_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols, rank=r, size=size)
         for r in range(size)]
    x, iters, conv = shares_cg(ks, D, b, tol=1e-10, max_iter=5000)
    assert conv and iters > 0
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-6
