"""GPU parity for cg_solve_local and the mixed-precision refined solver.

Known-answer tests follow tests/test_cg.cpp [1]-[7], [9] of the reference;
parity against the oracle (the reference's CG sequence restated in C) is:
same convergence flag, iteration count within +-2% (the parity bar for CG), both
recursive residuals below tol, and solutions within what the conditioning
allows.
"""
import numpy as np
import pytest
import torch

from num_amd import generators as g

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _up(ks, A):
    return ks.upload_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols)


def _solve(ks, A, b, tol, max_iter, x0=None):
    dA = _up(ks, A)
    x = _dev(np.zeros(A.nrows) if x0 is None else x0)
    r = ks.cg_solve_local(dA, _dev(b), x, tol=tol, max_iter=max_iter)
    return x.cpu().numpy(), r


def rel(a, b):
    d = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (d if d else 1.0)


def _csr(rows):
    rp, ci, v = [0], [], []
    for row in rows:
        for c, val in row:
            ci.append(c)
            v.append(val)
        rp.append(len(ci))
    n = len(rows)
    return g.CSR(np.array(rp, np.int32), np.array(ci, np.int32), np.array(v, np.float64), n, n)


def test_identity_cg(ks):  # [1]
    N = 50
    A = _csr([[(i, 1.0)] for i in range(N)])
    b = np.arange(1, N + 1, dtype=np.float64)
    x, r = _solve(ks, A, b, 1e-12, 200)
    assert rel(x, b) < 1e-10 and r["converged"] and r["iters"] <= 2


def test_2x2_spd(ks):  # [2]
    A = _csr([[(0, 4.0), (1, 1.0)], [(0, 1.0), (1, 3.0)]])
    x, r = _solve(ks, A, np.array([1.0, 2.0]), 1e-14, 100)
    assert rel(x, np.array([1 / 11, 7 / 11])) < 1e-10 and r["converged"] and r["iters"] <= 2


def test_diagonal_spd(ks):  # [3]
    N = 100
    d = np.arange(1, N + 1, dtype=np.float64)
    A = _csr([[(i, d[i])] for i in range(N)])
    x, r = _solve(ks, A, np.ones(N), 1e-12, 5000)
    assert rel(x, 1 / d) < 1e-8 and r["converged"]
    # n distinct eigenvalues -> CG terminates in <= n steps in exact arithmetic
    assert r["iters"] <= N


def test_laplacian_1d_convergence(ks, orc):  # [4]
    A = g.laplacian_1d(50)
    x, r = _solve(ks, A, np.ones(50), 1e-10, 500)
    assert r["converged"] and r["residual"] < 1e-10 and r["iters"] > 0
    # symmetric b on a symmetric tridiagonal: 25 distinct active eigenvalues
    assert r["iters"] <= 25 + 2
    _, ro = orc.cg_solve_local(A.rowptr, A.colind, A.values, np.ones(50), tol=1e-10, max_iter=500)
    assert r["iters"] == ro["iters"]


def test_zero_rhs(ks):  # [5]
    A = g.laplacian_1d(20)
    x, r = _solve(ks, A, np.zeros(20), 1e-10, 100)
    assert r["converged"] and r["iters"] == 0 and r["residual"] == 0.0
    assert np.all(x == 0.0)


def test_manufactured_sine(ks):  # [6]
    N = 500
    A = g.laplacian_1d(N)
    xs = np.sin(np.pi * (np.arange(N) + 1.0) / (N + 1.0))
    b = 2.0 * xs
    b[1:] -= xs[:-1]
    b[:-1] -= xs[1:]
    x, r = _solve(ks, A, b, 1e-10, 2000)
    assert r["converged"] and rel(x, xs) < 1e-8


def test_max_iter_respected(ks):  # [7]
    A = g.laplacian_1d(50)
    x, r = _solve(ks, A, np.ones(50), 1e-14, 5)
    assert r["iters"] <= 5 and not r["converged"]
    assert r["iters"] == 5


def test_repeated_solves_bit_identical(ks):  # [9]
    A = g.laplacian_1d(30)
    sols = [_solve(ks, A, np.ones(30), 1e-12, 500)[0] for _ in range(3)]
    assert np.array_equal(sols[0], sols[1]) and np.array_equal(sols[0], sols[2])


def test_breakdown_guard(ks):
    # singular 2x2 with b in the null space direction: p'Ap == 0 -> RuntimeError
    A = _csr([[(0, 1.0), (1, -1.0)], [(0, -1.0), (1, 1.0)]])
    dA = _up(ks, A)
    with pytest.raises(RuntimeError):
        ks.cg_solve_local(dA, _dev(np.array([1.0, 1.0])), _dev(np.zeros(2)), 1e-10, 10)


def test_nan_runs_to_max_iter_like_reference(ks, orc):
    # cg_solve_local has no breakdown guard (cg_solver.hpp:165-167) and the
    # distributed one, |pAp| < 1e-300, is false for NaN: a NaN right-hand side
    # runs max_iter iterations and returns a non-converged NaN residual
    A = g.laplacian_1d(8)
    b = np.ones(8)
    b[3] = np.nan
    dA = _up(ks, A)
    res = ks.cg_solve_local(dA, _dev(b), _dev(np.zeros(8)), 1e-10, 5)
    _, ro = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, np.zeros(8), 1e-10, 5)
    assert res["iters"] == ro["iters"] == 5
    assert not res["converged"] and not ro["converged"]
    assert np.isnan(res["residual"]) and np.isnan(ro["residual"])


def test_warm_start(ks, orc):
    A = g.laplacian_2d(32)
    b = np.random.default_rng(0).uniform(-1, 1, A.nrows)
    x0 = np.random.default_rng(1).uniform(-1, 1, A.nrows)
    x, r = _solve(ks, A, b, 1e-10, 2000, x0=x0)
    xo, ro = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, x0=x0, tol=1e-10, max_iter=2000)
    assert r["converged"] and ro["converged"]
    assert abs(r["iters"] - ro["iters"]) <= max(1, 0.02 * ro["iters"])


def test_config_cpu_path_parity(ks, orc):
    """2-D 256^2 Laplacian, b = U(-1,1) seed 0, x0 = 0, tol 1e-8, 2000 max."""
    A = g.laplacian_2d(256)
    b = np.random.default_rng(0).uniform(-1, 1, A.nrows)
    x, r = _solve(ks, A, b, 1e-8, 2000)
    xo, ro = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-8, max_iter=2000)
    assert r["converged"] and ro["converged"]
    assert abs(r["iters"] - ro["iters"]) <= max(1, int(0.02 * ro["iters"]))
    assert r["residual"] < 1e-8
    true = np.linalg.norm(b - A.to_scipy() @ x) / np.linalg.norm(b)
    assert true < 2e-8
    assert rel(x, xo) < 1e-6


def test_vs_scipy_cg(ks):
    import scipy.sparse.linalg as spla
    A = g.laplacian_3d(20)
    b = np.ones(A.nrows)
    x_ref, info = spla.cg(A.to_scipy(), b, rtol=1e-12, maxiter=5000)
    assert info == 0
    x, r = _solve(ks, A, b, 1e-10, 5000)
    assert r["converged"] and rel(x, x_ref) < 1e-6  # tests/test_accuracy.py:193


def test_mixed_refined(ks, orc):
    A = g.diffusion_3d(24, sigma=1.0, seed=3)
    b = np.random.default_rng(0).uniform(-1, 1, A.nrows)
    dA = _up(ks, A)
    x = _dev(np.zeros(A.nrows))
    r = ks.cg_solve_mixed(dA, _dev(b), x, tol=1e-10)
    assert r["converged"] and r["residual"] < 1e-10
    xh = x.cpu().numpy()
    true = np.linalg.norm(b - A.to_scipy() @ xh) / np.linalg.norm(b)
    assert true < 1e-10
    x64, r64 = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-10, max_iter=20000)
    assert rel(xh, x64) < 1e-7


def _hub_spd(L=40, hub_deg=120, seed=1):
    """2-D Laplacian plus one 'hub' row/column joined to hub_deg nodes:
    symmetric, strictly diagonally dominant (SPD), with one row far longer
    than 32 entries, so the FP32 path keeps int32 columns."""
    import scipy.sparse as sp
    A = g.laplacian_2d(L).to_scipy().tolil()
    n = A.shape[0]
    rng = np.random.default_rng(seed)
    nbrs = rng.choice(np.arange(1, n), hub_deg, replace=False)
    for j in nbrs:
        A[0, j] = -0.5
        A[j, 0] = -0.5
        A[j, j] = A[j, j] + 0.5
    A[0, 0] = 0.5 * hub_deg + 1.0
    A = sp.csr_matrix(A)
    A.sort_indices()
    return g.CSR(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64),
                 n, n)


@pytest.mark.parametrize("case", ["diffusion", "lap2d_odd", "hub"])
def test_mixed_paths(ks, orc, case):
    """The paired int16-offset FP32 path (stencils) and the int32 path (a row
    longer than 32 entries) both reach the FP64 true residual, and agree
    with the FP64 oracle solution."""
    if case == "diffusion":
        A = g.diffusion_3d(30, sigma=1.0, seed=5)
    elif case == "lap2d_odd":
        A = g.stencil((123, 77), 0.01)   # odd tile count: the last pair is a single tile
    else:
        A = _hub_spd()
    b = np.random.default_rng(2).uniform(-1, 1, A.nrows)
    dA = _up(ks, A)
    dA.ensure_f32()
    assert (dA._rel16 is not None) == (case != "hub")
    x = _dev(np.zeros(A.nrows))
    r = ks.cg_solve_mixed(dA, _dev(b), x, tol=1e-10)
    assert r["converged"] and r["residual"] < 1e-10
    xh = x.cpu().numpy()
    true = np.linalg.norm(b - A.to_scipy() @ xh) / np.linalg.norm(b)
    assert true < 1e-10
    x64, r64 = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-10, max_iter=20000)
    assert rel(xh, x64) < 1e-7
    # the FP32 steps stay close to the FP64 iteration count
    assert r["inner_total"] <= 1.5 * r64["iters"] + 10
    if case != "hub":
        # the pair kernel read the offsets' row patterns; with 2 offset bytes
        # per entry (mixed path 2) it must give the same bits
        assert dA.patterns["offsets"]
        dA.set_mixed_path(2)
        x2 = _dev(np.zeros(A.nrows))
        r2 = ks.cg_solve_mixed(dA, _dev(b), x2, tol=1e-10)
        assert r2 == r and np.array_equal(x2.cpu().numpy(), xh)


def test_mixed_offsets_fall_back_past_16_patterns(ks, orc):
    """Rows coupled at random offsets in +-1..+-12: a pair of tiles holds far
    more than 16 distinct offset lists, so the matrix keeps the int16 offset
    per entry (no pattern table), and the mixed solve still reaches the FP64
    true residual and the oracle's solution."""
    import scipy.sparse as sp
    rng = np.random.default_rng(11)
    n = 3000
    offs = np.concatenate([-np.arange(1, 13), np.arange(1, 13)])
    r = np.arange(n)
    c = np.clip(r + rng.choice(offs, n), 0, n - 1)
    keep = c != r
    M = sp.coo_matrix((np.full(int(keep.sum()), -0.3), (r[keep], c[keep])), shape=(n, n))
    M = (M + M.T).tocsr()
    M = M + sp.diags(np.asarray(abs(M).sum(axis=1)).ravel() + 0.5)
    M = sp.csr_matrix(M)
    M.sort_indices()
    A = g.CSR(M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.astype(np.float64),
              n, n)
    assert np.diff(A.rowptr).max() <= 8
    b = rng.uniform(-1, 1, n)
    dA = _up(ks, A)
    dA.ensure_f32()
    assert dA._rel16 is not None and not dA.patterns["offsets"]
    x = _dev(np.zeros(n))
    res = ks.cg_solve_mixed(dA, _dev(b), x, tol=1e-10)
    assert res["converged"]
    xh = x.cpu().numpy()
    assert np.linalg.norm(b - M @ xh) / np.linalg.norm(b) < 1e-10
    x64, _ = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-10, max_iter=20000)
    assert rel(xh, x64) < 1e-7


def test_mixed_offsets_unsorted_duplicate_columns(ks, orc):
    """A 2-D stencil stored with rows in one of four entry orders and the
    diagonal split into two duplicate entries on every third row (the CSR
    contract allows both): the offset patterns carry each row's stored
    order and its length (a duplicated diagonal is two zero offsets), and
    the mixed solve reaches the FP64 true residual and the oracle's
    solution, with the same bits as the int16 offsets per entry."""
    from _helpers import shuffled_stencil_rows
    import scipy.sparse as sp
    rp, ci, va, n = shuffled_stencil_rows(90, 60, seed=6)
    A = g.CSR(rp, ci, va, n, n)
    M = sp.csr_matrix((va.copy(), ci.copy(), rp.copy()), shape=(n, n))
    b = np.random.default_rng(8).uniform(-1, 1, n)
    dA = _up(ks, A)
    dA.ensure_f32()
    assert dA.patterns["offsets"]
    x = _dev(np.zeros(n))
    res = ks.cg_solve_mixed(dA, _dev(b), x, tol=1e-10)
    assert res["converged"]
    xh = x.cpu().numpy()
    assert np.linalg.norm(b - M @ xh) / np.linalg.norm(b) < 1e-10
    x64, _ = orc.cg_solve_local(rp, ci, va, b, tol=1e-10, max_iter=20000)
    assert rel(xh, x64) < 1e-7
    dA.set_mixed_path(2)
    x2 = _dev(np.zeros(n))
    res2 = ks.cg_solve_mixed(dA, _dev(b), x2, tol=1e-10)
    assert res2 == res and np.array_equal(x2.cpu().numpy(), xh)


@pytest.mark.parametrize("mixed", [False, True])
def test_full_width_update_groups(ks, orc, mixed):
    """n >= 2^20 rows: the update kernel runs its full 256 workgroups, so the
    FP64 solver's last-workgroup fold draws 256 tickets per iteration and
    every workgroup of the mixed SpMV kernel folds 256 partials.  Same
    iteration count as the oracle (+-2%), the true residual, and bit-identical
    repeats (the fold's order does not depend on which workgroup is last)."""
    A = g.stencil((1024, 1024), 0.05)
    b = np.random.default_rng(3).uniform(-1, 1, A.nrows)
    dA = _up(ks, A)
    if mixed:
        dA.ensure_f32()
    xs = []
    for _ in range(2):
        x = _dev(np.zeros(A.nrows))
        if mixed:
            r = ks.cg_solve_mixed(dA, _dev(b), x, tol=1e-10)
        else:
            r = ks.cg_solve_local(dA, _dev(b), x, tol=1e-10, max_iter=5000)
        assert r["converged"] and r["residual"] < 1e-10
        xs.append(x.cpu().numpy())
    assert np.array_equal(xs[0], xs[1])
    true = np.linalg.norm(b - A.to_scipy() @ xs[0]) / np.linalg.norm(b)
    assert true < (1e-10 if mixed else 2e-10)
    xo, ro = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-10, max_iter=5000)
    assert rel(xs[0], xo) < 1e-7
    if not mixed:
        assert abs(r["iters"] - ro["iters"]) <= max(1, int(0.02 * ro["iters"]))


@pytest.mark.parametrize("max_iter", [0, 1, 2, 3, 4, 5, 6, 7])
def test_iterate_after_each_iteration_count(ks, orc, max_iter):
    # the x update runs every other iteration (two at once); a solve that
    # stops after an odd or an even number of iterations must return the
    # iterate of exactly that many iterations (the oracle's x after max_iter)
    A = g.laplacian_2d(24)
    b = np.random.default_rng(7).uniform(-1, 1, A.nrows)
    x, r = _solve(ks, A, b, 1e-14, max_iter)
    xo, ro = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, np.zeros(A.nrows), 1e-14,
                                max_iter)
    assert r["iters"] == ro["iters"] == max_iter and not r["converged"]
    assert rel(x, xo) < 1e-13, rel(x, xo)


@pytest.mark.parametrize("L", [5, 6, 9])
def test_converged_iterate_odd_and_even_counts(ks, orc, L):
    # n distinct eigenvalues: CG converges in a few iterations, odd or even
    A = g.laplacian_1d(L)
    b = np.random.default_rng(L).uniform(-1, 1, L)
    x, r = _solve(ks, A, b, 1e-12, 100)
    xo, ro = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, np.zeros(L), 1e-12, 100)
    assert r["converged"] and r["iters"] == ro["iters"]
    assert rel(x, xo) < 1e-12
    dense = np.zeros((L, L))
    for i in range(L):
        dense[i, A.colind[A.rowptr[i]:A.rowptr[i + 1]]] = A.values[A.rowptr[i]:A.rowptr[i + 1]]
    assert np.linalg.norm(dense @ x - b) <= 1e-11 * np.linalg.norm(b)


@pytest.mark.parametrize("mixed", [False, True])
def test_breakdown_keeps_last_iterate(ks, mixed):
    # A = diag(1, 0), b = [1, 1]: the first step gives x = [2, 2]; the second
    # direction [0, 2] has p.Ap = 0 (breakdown).  The returned iterate must
    # hold the first step's update: the FP64 solver runs its x update every
    # other iteration, so that update is still pending when the solve stops.
    A = _csr([[(0, 1.0)], [(1, 0.0)]])
    dA = _up(ks, A)
    x = _dev(np.zeros(2))
    with pytest.raises(RuntimeError):
        if mixed:
            ks.cg_solve_mixed(dA, _dev(np.array([1.0, 1.0])), x, tol=1e-10, max_iter=10)
        else:
            ks.cg_solve_local(dA, _dev(np.array([1.0, 1.0])), x, 1e-10, 10)
    assert x.cpu().numpy().tolist() == [2.0, 2.0]


@pytest.mark.parametrize("max_iter", [0, 1, 2, 3, 4, 5, 8])
def test_mixed_stop_after_each_step_count(ks, max_iter):
    # the mixed solver ends with an FP64 refinement whose true residual it
    # reports: the returned x must be the iterate that residual belongs to,
    # whatever the parity of the last step (a pending x32 update is folded in)
    A = g.laplacian_2d(20)
    b = np.random.default_rng(3).uniform(-1, 1, A.nrows)
    dA = _up(ks, A)
    x = _dev(np.zeros(A.nrows))
    r = ks.cg_solve_mixed(dA, _dev(b), x, tol=1e-14, max_iter=max_iter)
    assert not r["converged"] and r["inner_total"] == max_iter
    xh = x.cpu().numpy()
    true = np.linalg.norm(b - A.to_scipy() @ xh) / np.linalg.norm(b)
    assert abs(true - r["residual"]) <= 1e-12 * r["residual"], (true, r["residual"])
