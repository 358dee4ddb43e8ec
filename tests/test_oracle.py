"""Pin the CPU oracle (oracle/oracle.c) before trusting it as the parity bar.

Each oracle function is checked against what fixes it independently:
  * the reference's known-answer tests (tests/test_spmv_kernel.cpp [1]-[5],
    tests/test_cg.cpp [1]-[7], [9]), reproduced from their source text;
  * library routines (scipy.sparse A @ x, scipy.sparse.linalg.cg);
  * closed forms (Laplacian * const = missing-neighbour count, sine modes are
    eigenvectors), exact integer sums, brute-force dense products;
  * and, for each check family, a planted mistake in the oracle (mutant
    switch in oracle.c) that the same check must catch.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from num_amd import generators as g


@pytest.fixture(autouse=True)
def _faithful():
    oracle.lib().orc_set_mutant(0)
    yield
    oracle.lib().orc_set_mutant(0)


def rel(a, b):
    d = float(np.linalg.norm(b))
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b))) / (d if d else 1.0)


TRI4 = (np.array([0, 2, 5, 8, 10], np.int32), np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3], np.int32),
        np.array([2, -1, -1, 2, -1, -1, 2, -1, -1, 2], np.float64))


def dense_of(rp, ci, v, n, m):
    D = np.zeros((n, m))
    for i in range(n):
        for j in range(rp[i], rp[i + 1]):
            D[i, ci[j]] += v[j]
    return D


# ---------------------------------------------------------------- SpMV -----

def kat_tridiag():
    y = oracle.spmv(*TRI4, np.array([1.0, 2, 3, 4]))
    return y.tolist() == [0.0, 0.0, 0.0, 5.0]


def kat_lcg():
    A, x = g.lcg_matrix()
    y = oracle.spmv(A.rowptr, A.colind, A.values, x)
    # dense_spmv of tests/test_spmv_kernel.cpp:44-56, written out with scipy's
    # duplicate-summing COO -> dense path
    D = sp.csr_matrix((A.values, A.colind, A.rowptr), shape=(A.nrows, A.ncols)).toarray()
    ref = D @ x
    return np.max(np.abs(y - ref)) / np.max(np.abs(ref)) < 1e-12


def closed_form_laplacian():
    L = 9
    A = g.laplacian_3d(L)
    y = oracle.spmv(A.rowptr, A.colind, A.values, np.ones(A.ncols))
    idx = np.arange(A.nrows)
    c = [idx % L, (idx // L) % L, idx // (L * L)]
    deg = sum((ci > 0).astype(int) + (ci < L - 1).astype(int) for ci in c)
    return np.array_equal(y, (6 - deg).astype(float))


def closed_form_sine():
    L = 11
    A = g.laplacian_2d(L)
    idx = np.arange(A.nrows)
    i, j = idx % L, idx // L
    mode = np.sin(np.pi * 2 * (i + 1) / (L + 1)) * np.sin(np.pi * 3 * (j + 1) / (L + 1))
    lam = (2 - 2 * np.cos(2 * np.pi / (L + 1))) + (2 - 2 * np.cos(3 * np.pi / (L + 1)))
    return rel(oracle.spmv(A.rowptr, A.colind, A.values, mode), lam * mode) < 1e-13


def scipy_rect():
    rng = np.random.default_rng(3)
    M = sp.random(300, 170, density=0.05, random_state=4, format="csr")
    x = rng.standard_normal(170)
    return rel(oracle.spmv(M.indptr, M.indices, M.data, x, nrows=300), M @ x) < 1e-14


SPMV_CHECKS = [kat_tridiag, kat_lcg, closed_form_laplacian, closed_form_sine, scipy_rect]


@pytest.mark.parametrize("check", SPMV_CHECKS, ids=lambda f: f.__name__)
def test_spmv_pinned(check):
    assert check()


def test_identity_kat():
    N = 100
    x = np.arange(N) * 0.1
    y = oracle.spmv(np.arange(N + 1), np.arange(N), np.ones(N), x)
    assert np.array_equal(y, x)


def test_spmv_brute_force():
    rng = np.random.default_rng(0)
    for n, m, d in ((1, 1, 1.0), (7, 5, 0.6), (64, 300, 0.05), (2048, 2048, 0.003)):
        M = sp.random(n, m, density=d, random_state=int(rng.integers(1 << 30)), format="csr")
        x = rng.standard_normal(m)
        assert rel(oracle.spmv(M.indptr, M.indices, M.data, x, nrows=n), M.toarray() @ x) < 1e-13


def test_spmv_threads_invariant():
    A = g.zipf_matrix(n=20000, kmax=500, seed=1)
    x = np.random.default_rng(2).uniform(-1, 1, A.ncols)
    oracle.set_threads(1)
    y1 = oracle.spmv(A.rowptr, A.colind, A.values, x)
    oracle.set_threads(8)
    y8 = oracle.spmv(A.rowptr, A.colind, A.values, x)
    oracle.set_threads(0)
    assert np.array_equal(y1, y8)


@pytest.mark.parametrize("m", [1, 2])
def test_spmv_mutants_caught(m):
    with oracle.mutant(m):
        assert not all(c() for c in SPMV_CHECKS)


def test_spmv_f32_matches_sequential_float32():
    A = g.diffusion_3d(6, seed=3)
    v32 = A.values.astype(np.float32)
    x = np.random.default_rng(1).uniform(-1, 1, A.ncols).astype(np.float32)
    y = oracle.spmv_f32(A.rowptr, A.colind, v32, x)
    ref = np.zeros(A.nrows, np.float32)
    for i in range(A.nrows):
        acc = np.float32(0.0)
        for j in range(A.rowptr[i], A.rowptr[i + 1]):
            acc = np.float32(acc + np.float32(v32[j] * x[A.colind[j]]))
        ref[i] = acc
    assert np.array_equal(y, ref)


# ---------------------------------------------------------- batch SpMV -----

def batch_kat_identity():
    Y = oracle.batch_spmv(*TRI4, np.eye(4))
    return np.array_equal(Y, dense_of(*TRI4, 4, 4))


def batch_scipy():
    A = g.laplacian_3d(6)
    X = np.random.default_rng(5).standard_normal((A.ncols, 5))
    return rel(oracle.batch_spmv(A.rowptr, A.colind, A.values, X), A.to_scipy() @ X) < 1e-14


def batch_k1_equals_spmv():
    A, x = g.lcg_matrix(N=500)
    Y = oracle.batch_spmv(A.rowptr, A.colind, A.values, x[:, None])
    return np.array_equal(Y[:, 0], oracle.spmv(A.rowptr, A.colind, A.values, x))


BATCH_CHECKS = [batch_kat_identity, batch_scipy, batch_k1_equals_spmv]


@pytest.mark.parametrize("check", BATCH_CHECKS, ids=lambda f: f.__name__)
def test_batch_pinned(check):
    assert check()


def test_batch_mutant_caught():
    with oracle.mutant(7):
        assert not all(c() for c in BATCH_CHECKS)


# ------------------------------------------------------------- BLAS-1 ------

def blas_exact_integer_dot():
    n = 100_003
    x = (np.arange(n) % 13 - 6).astype(np.float64)
    y = (np.arange(n) % 7 - 3).astype(np.float64)
    return oracle.dot(x, y) == float(np.dot(x.astype(np.int64), y.astype(np.int64)))


def blas_fsum_dot():
    x = np.random.default_rng(1).uniform(-1, 1, 54321)
    y = np.random.default_rng(2).uniform(-1, 1, 54321)
    return abs(oracle.dot(x, y) - math.fsum((x * y).tolist())) < 1e-12


def blas_axpy2_elementwise():
    x = np.random.default_rng(3).uniform(-1, 1, 1000)
    y = np.random.default_rng(4).uniform(-1, 1, 1000)
    return np.array_equal(oracle.axpy2(2.0, x, -0.5, y), 2.0 * x + (-0.5) * y)


def blas_fill_every_element():
    x = np.random.default_rng(5).uniform(-1, 1, 1001)
    return np.array_equal(oracle.fill(x, -2.5), np.full(1001, -2.5))


def blas_copy_every_element():
    src = np.random.default_rng(6).uniform(-1, 1, 1001)
    dst = np.full(1001, np.nan)
    return np.array_equal(oracle.copy(src, dst), src)


BLAS_CHECKS = [blas_exact_integer_dot, blas_fsum_dot, blas_axpy2_elementwise,
               blas_fill_every_element, blas_copy_every_element]


@pytest.mark.parametrize("check", BLAS_CHECKS, ids=lambda f: f.__name__)
def test_blas_pinned(check):
    assert check()


@pytest.mark.parametrize("m", [5, 6, 8, 9])
def test_blas_mutants_caught(m):
    with oracle.mutant(m):
        assert not all(c() for c in BLAS_CHECKS)


def test_axpy_norm_scale():
    x = np.random.default_rng(3).uniform(-1, 1, 777)
    y = np.random.default_rng(4).uniform(-1, 1, 777)
    assert np.array_equal(oracle.axpy(0.3, x, y), 0.3 * x + y)
    assert np.array_equal(oracle.scale(x, -2.0), -2.0 * x)
    assert abs(oracle.norm2(x) - math.sqrt(math.fsum((x * x).tolist()))) < 1e-13


# ----------------------------------------------------------------- CG -------

def _csr(rows):
    rp, ci, v = [0], [], []
    for row in rows:
        for c, val in row:
            ci.append(c)
            v.append(val)
        rp.append(len(ci))
    return np.array(rp, np.int32), np.array(ci, np.int32), np.array(v, np.float64)


def cg_identity():  # [1]
    N = 50
    b = np.arange(1, N + 1, dtype=float)
    x, r = oracle.cg_solve_local(*_csr([[(i, 1.0)] for i in range(N)]), b, tol=1e-12,
                                 max_iter=200)
    return rel(x, b) < 1e-10 and r["converged"] and r["iters"] <= 2


def cg_2x2():  # [2]
    x, r = oracle.cg_solve_local(*_csr([[(0, 4.0), (1, 1.0)], [(0, 1.0), (1, 3.0)]]),
                                 np.array([1.0, 2.0]), tol=1e-14, max_iter=100)
    return rel(x, [1 / 11, 7 / 11]) < 1e-10 and r["converged"] and r["iters"] <= 2


def cg_diagonal():  # [3]
    d = np.arange(1, 101, dtype=float)
    x, r = oracle.cg_solve_local(*_csr([[(i, d[i])] for i in range(100)]), np.ones(100),
                                 tol=1e-12, max_iter=5000)
    return rel(x, 1 / d) < 1e-8 and r["converged"]


def cg_laplacian_1d():  # [4] + finite termination on 25 active eigenvalues
    A = g.laplacian_1d(50)
    x, r = oracle.cg_solve_local(A.rowptr, A.colind, A.values, np.ones(50), tol=1e-10,
                                 max_iter=500)
    return r["converged"] and r["residual"] < 1e-10 and 0 < r["iters"] <= 27


def cg_sine():  # [6]
    N = 500
    A = g.laplacian_1d(N)
    xs = np.sin(np.pi * (np.arange(N) + 1.0) / (N + 1.0))
    b = A.to_scipy() @ xs
    x, r = oracle.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-10, max_iter=2000)
    return r["converged"] and rel(x, xs) < 1e-8


def cg_vs_scipy():
    A = g.laplacian_2d(24)
    b = np.random.default_rng(0).uniform(-1, 1, A.nrows)
    x_ref, info = spla.cg(A.to_scipy(), b, rtol=1e-13, maxiter=5000)
    x, r = oracle.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-11, max_iter=5000)
    return info == 0 and r["converged"] and rel(x, x_ref) < 1e-8


def cg_warm_start():
    """x0 != 0 exercises r = b - A x0 (with x0 = 0 the sign of A x0 is invisible)."""
    A = g.laplacian_2d(16)
    rng = np.random.default_rng(7)
    b = rng.uniform(-1, 1, A.nrows)
    x0 = rng.uniform(-1, 1, A.nrows)
    x, r = oracle.cg_solve_local(A.rowptr, A.colind, A.values, b, x0=x0, tol=1e-12,
                                 max_iter=2000)
    x_ref = spla.spsolve(A.to_scipy().tocsc(), b)
    return r["converged"] and rel(x, x_ref) < 1e-9


CG_CHECKS = [cg_identity, cg_2x2, cg_diagonal, cg_laplacian_1d, cg_sine, cg_vs_scipy,
             cg_warm_start]


@pytest.mark.parametrize("check", CG_CHECKS, ids=lambda f: f.__name__)
def test_cg_pinned(check):
    assert check()


@pytest.mark.parametrize("m", [3, 4])
def test_cg_mutants_caught(m):
    with oracle.mutant(m):
        assert not all(c() for c in CG_CHECKS)


def test_cg_zero_rhs():  # [5]
    A = g.laplacian_1d(20)
    x, r = oracle.cg_solve_local(A.rowptr, A.colind, A.values, np.zeros(20), tol=1e-10,
                                 max_iter=100)
    assert r == {"iters": 0, "residual": 0.0, "converged": True}
    assert np.all(x == 0)


def test_cg_max_iter():  # [7]
    A = g.laplacian_1d(50)
    _, r = oracle.cg_solve_local(A.rowptr, A.colind, A.values, np.ones(50), tol=1e-14,
                                 max_iter=5)
    assert r["iters"] == 5 and not r["converged"]


def test_cg_repeat_bit_identical():  # [9]
    A = g.laplacian_1d(30)
    xs = [oracle.cg_solve_local(A.rowptr, A.colind, A.values, np.ones(30), tol=1e-12,
                                max_iter=500)[0] for _ in range(3)]
    assert np.array_equal(xs[0], xs[1]) and np.array_equal(xs[0], xs[2])


def test_cg_config_cpu_path():
    """2-D 256^2, b = U(-1,1) seed 0, tol 1e-8, 2000 max: converges, true
    residual matches the recursive one, scipy agrees."""
    A = g.laplacian_2d(256)
    b = np.random.default_rng(0).uniform(-1, 1, A.nrows)
    x, r = oracle.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-8, max_iter=2000)
    assert r["converged"] and r["residual"] < 1e-8
    true = np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b)
    assert true < 2e-8


def test_mixed_reaches_fp64_residual():
    A = g.diffusion_3d(10, seed=3)
    b = np.random.default_rng(0).uniform(-1, 1, A.nrows)
    x, r = oracle.cg_mixed(A.rowptr, A.colind, A.values, b, tol=1e-10)
    assert r["converged"]
    true = np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b)
    assert true < 1e-10 and r["true_residual"] < 1e-10
