"""GPU parity: SpMV, FP32 SpMV, multi-RHS SpMV and the vector ops through the
C ABI, against the CPU oracle on the same seeded inputs.

Bar: the row-tile kernel sums every row of <= 32 entries in index order
from 0.0 with one thread, so those rows must be bit-identical to the
oracle's sequential row loop (src/spmv_kernel.hpp:90-97); longer rows (a
wave or a workgroup per row) are checked to 1e-12 relative (the target of
the reference's LCG test, tests/test_spmv_kernel.cpp:205).
"""
import math

import numpy as np
import pytest
import torch

from num_amd import generators as g

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _up(ks, A):
    return ks.upload_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols)


def _spmv(ks, dA, x):
    y = torch.full((dA.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    ks.spmv(dA, _dev(x), y)
    return y.cpu().numpy()


def rel(a, b):
    d = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / (d if d else 1.0)


# -- known-answer tests of tests/test_spmv_kernel.cpp ------------------------

def test_tridiag4_kat(ks):
    # [1] and [2]: A = tridiag(-1, 2, -1) 4x4, x = [1,2,3,4] -> [0,0,0,5]
    rp = np.array([0, 2, 5, 8, 10], np.int32)
    ci = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3], np.int32)
    v = np.array([2, -1, -1, 2, -1, -1, 2, -1, -1, 2], np.float64)
    dA = ks.upload_csr(rp, ci, v, 4, 4)
    for kk in (True, False):
        y = torch.zeros(4, dtype=torch.float64, device="cuda")
        ks.spmv(dA, _dev(np.array([1.0, 2, 3, 4])), y, kokkoskernels=kk)
        assert y.cpu().numpy().tolist() == [0.0, 0.0, 0.0, 5.0]
    assert repr(dA) == "<CrsMatrix nrows=4 ncols=4 nnz=10>"


def test_identity_kat(ks):
    # [3]: identity N=100, x_i = 0.1 i -> y == x exactly
    N = 100
    rp = np.arange(N + 1, dtype=np.int32)
    dA = ks.upload_csr(rp, np.arange(N, dtype=np.int32), np.ones(N), N, N)
    x = np.arange(N) * 0.1
    assert np.array_equal(_spmv(ks, dA, x), x)


def test_lcg_kat_bitexact(ks, orc):
    # [5]: LCG N=1e4 matrix, unsorted and repeated columns
    A, x = g.lcg_matrix()
    y = _spmv(ks, _up(ks, A), x)
    y_ref = orc.spmv(A.rowptr, A.colind, A.values, x)
    assert np.array_equal(y, y_ref)
    assert np.max(np.abs(y - y_ref)) / np.max(np.abs(y_ref)) < 1e-12


@pytest.mark.parametrize("L", [1, 2, 7, 64, 256])
def test_laplacian_2d_bitexact(ks, orc, L):
    A = g.laplacian_2d(L)
    x = np.random.default_rng(1).uniform(-1, 1, A.ncols)
    assert np.array_equal(_spmv(ks, _up(ks, A), x), orc.spmv(A.rowptr, A.colind, A.values, x))


def test_laplacian_3d_closed_forms(ks):
    # L * ones = number of missing neighbours; L * sine mode = lambda * mode
    L = 24
    A = g.laplacian_3d(L)
    dA = _up(ks, A)
    y = _spmv(ks, dA, np.ones(A.ncols))
    idx = np.arange(A.nrows)
    c = [idx % L, (idx // L) % L, idx // (L * L)]
    deg = sum((ci > 0).astype(int) + (ci < L - 1).astype(int) for ci in c)
    assert np.array_equal(y, (6 - deg).astype(np.float64))
    k = (1, 2, 3)
    s = [np.sin(np.pi * kk * (ci + 1) / (L + 1)) for kk, ci in zip(k, c)]
    mode = s[0] * s[1] * s[2]
    lam = sum(2 - 2 * np.cos(np.pi * kk / (L + 1)) for kk in k)
    assert rel(_spmv(ks, dA, mode), lam * mode) < 1e-13


def test_zipf_long_rows(ks, orc):
    A = g.zipf_matrix(n=30000, kmax=4096, seed=7)
    info = ks.upload_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols).schedule_info()
    assert info["max_row_nnz"] > 2048 and info["long_rows"] > 0
    x = np.random.default_rng(3).uniform(-1, 1, A.ncols)
    y = _spmv(ks, _up(ks, A), x)
    y_ref = orc.spmv(A.rowptr, A.colind, A.values, x)
    assert rel(y, y_ref) < 1e-12
    # rows of <= 32 entries are walked by one thread in index order even in
    # tiles that also hold longer rows: bit-identical to the row loop
    short = np.diff(A.rowptr) <= 32
    assert np.array_equal(y[short], y_ref[short])


def test_arrow_matrix(ks, orc):
    # An arrow matrix: row 0 holds every column (a single row of 200,000
    # entries, far past a tile's capacity: a workgroup of its own), column 0
    # is read by every row, and the diagonal.  Row 0 within 1e-12 of the
    # oracle, every other row (2 entries) bit-identical; the multi-RHS
    # kernels on the same matrix (k = 3: lane groups, k = 128: a wave per
    # row) bit-identical on the short rows too.
    n = 200_000
    rng = np.random.default_rng(21)
    rp = np.empty(n + 1, np.int64)
    rp[0], rp[1] = 0, n
    rp[2:] = n + 2 * np.arange(1, n)
    ci = np.empty(rp[-1], np.int32)
    ci[:n] = np.arange(n)
    ci[n::2] = 0
    ci[n + 1::2] = np.arange(1, n)
    va = rng.uniform(-1, 1, rp[-1])
    rp = rp.astype(np.int32)
    dA = ks.upload_csr(rp, ci, va, n, n)
    assert dA.schedule_info()["long_rows"] == 1
    x = rng.uniform(-1, 1, n)
    y = _spmv(ks, dA, x)
    y_ref = orc.spmv(rp, ci, va, x)
    assert abs(y[0] - y_ref[0]) <= 1e-12 * np.abs(va[:n] * x).sum()
    assert np.array_equal(y[1:], y_ref[1:])
    for k in (3, 128):
        X = rng.standard_normal((n, k))
        Y = torch.zeros((n, k), dtype=torch.float64, device="cuda")
        ks.batch_spmv(dA, _dev(X), Y)
        Yr = orc.batch_spmv(rp, ci, va, X)
        Yg = Y.cpu().numpy()
        assert np.array_equal(Yg[1:], Yr[1:])
        assert np.all(np.abs(Yg[0] - Yr[0]) <= 1e-12 * (np.abs(va[:n, None] * X).sum(axis=0)))


@pytest.mark.parametrize("lo,hi,frac_short", [(33, 300, 0.0), (1, 1800, 0.5), (30, 40, 0.3)])
def test_medium_rows(ks, orc, lo, hi, frac_short):
    """Tiles with many rows longer than 32 entries (each goes to a wave, the
    m-th to wave m % 4) mixed with short rows."""
    rng = np.random.default_rng(lo + hi)
    n = 20000
    k = rng.integers(lo, hi + 1, n)
    short_rows = rng.random(n) < frac_short
    k[short_rows] = rng.integers(0, 6, int(short_rows.sum()))
    rp = np.concatenate([[0], np.cumsum(k)]).astype(np.int32)
    ci = np.sort(rng.integers(0, n, int(rp[-1])).astype(np.int32))
    v = rng.uniform(-1, 1, int(rp[-1]))
    dA = ks.upload_csr(rp, ci, v, n, n)
    x = rng.uniform(-1, 1, n)
    y = _spmv(ks, dA, x)
    y_ref = orc.spmv(rp, ci, v, x)
    assert rel(y, y_ref) < 1e-12
    s = k <= 32
    assert np.array_equal(y[s], y_ref[s])


@pytest.mark.parametrize("sched", [(256, 2048), (64, 512), (16, 4096), (256, 1024), (1, 1)])
def test_schedules_agree(ks, orc, sched):
    A = g.zipf_matrix(n=5000, kmax=600, seed=11)
    dA = _up(ks, A)
    dA.set_schedule(*sched)
    x = np.random.default_rng(4).uniform(-1, 1, A.ncols)
    assert rel(_spmv(ks, dA, x), orc.spmv(A.rowptr, A.colind, A.values, x)) < 1e-12


def test_plane_aligned_chunks(ks, orc):
    # 80^3: the z-plane offset is 6400 rows, so the automatic XCD chunk is
    # 6400 / 8 = 800 rows = 3 tiles of 256 rows.  64^3 (4096 / 8 = 512 rows
    # = 2 tiles) and a 2-D 200^2 grid (offset 200 rows) are under the 3-tile
    # minimum and keep the size-based chunk (1 tile at these sizes).  The
    # order of the workgroups never changes a result.
    assert _up(ks, g.laplacian_3d(64)).schedule_info()["chunk"] == 1
    A = g.laplacian_3d(80)
    dA = _up(ks, A)
    assert dA.schedule_info()["chunk"] == 3
    x = np.random.default_rng(6).uniform(-1, 1, A.ncols)
    y = _spmv(ks, dA, x)
    assert np.array_equal(y, orc.spmv(A.rowptr, A.colind, A.values, x))
    for ch in (1, 64):
        dA.set_schedule(chunk=ch)
        assert dA.schedule_info()["chunk"] == ch
        assert np.array_equal(_spmv(ks, dA, x), y)
    B = g.laplacian_2d(200)
    assert _up(ks, B).schedule_info()["chunk"] == 1


def test_empty_rows_and_rect(ks, orc):
    rng = np.random.default_rng(5)
    nrows, ncols = 3000, 777
    k = rng.integers(0, 5, nrows)
    k[::7] = 0
    rp = np.concatenate([[0], np.cumsum(k)]).astype(np.int32)
    ci = rng.integers(0, ncols, int(rp[-1])).astype(np.int32)
    v = rng.uniform(-1, 1, int(rp[-1]))
    dA = ks.upload_csr(rp, ci, v, nrows, ncols)
    x = rng.uniform(-1, 1, ncols)
    y = _spmv(ks, dA, x)
    assert np.array_equal(y, orc.spmv(rp, ci, v, x))
    assert np.all(y[::7] == 0.0)


def test_spmv_f32(ks, orc):
    A = g.diffusion_3d(20, seed=3)
    dA = _up(ks, A)
    dA.ensure_f32()
    x = np.random.default_rng(2).uniform(-1, 1, A.ncols).astype(np.float32)
    y = torch.zeros(A.nrows, dtype=torch.float32, device="cuda")
    from num_amd import _lib
    import ctypes as C
    _lib.call("num_spmv_f32", dA.handle, C.c_void_p(_dev(x).data_ptr()),
              C.c_void_p(y.data_ptr()), None)
    torch.cuda.synchronize()
    y_ref = orc.spmv_f32(A.rowptr, A.colind, A.values.astype(np.float32), x)
    assert np.array_equal(y.cpu().numpy(), y_ref)


def test_shape_errors(ks):
    A = g.laplacian_2d(4)
    dA = _up(ks, A)
    with pytest.raises(ValueError):
        ks.spmv(dA, torch.zeros(15, dtype=torch.float64, device="cuda"),
                torch.zeros(16, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        ks.spmv(dA, torch.zeros(16, dtype=torch.float64, device="cuda"),
                torch.zeros(17, dtype=torch.float64, device="cuda"))
    with pytest.raises(TypeError):
        ks.spmv(dA, torch.zeros(16, dtype=torch.float32, device="cuda"),
                torch.zeros(16, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        ks.upload_csr(A.rowptr, A.colind, A.values, 0, 16)


# -- multi-RHS (src/batch_spmv_kernel.hpp) -----------------------------------

def test_batch_identity_kat(ks):
    # [4]: tridiag 4x4, X = I4 -> Y[:, b] = A[:, b]
    rp = np.array([0, 2, 5, 8, 10], np.int32)
    ci = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3], np.int32)
    v = np.array([2, -1, -1, 2, -1, -1, 2, -1, -1, 2], np.float64)
    dA = ks.upload_csr(rp, ci, v, 4, 4)
    Y = torch.zeros((4, 4), dtype=torch.float64, device="cuda")
    ks.batch_spmv(dA, torch.eye(4, dtype=torch.float64, device="cuda"), Y)
    dense = np.zeros((4, 4))
    for i in range(4):
        for j in range(rp[i], rp[i + 1]):
            dense[i, ci[j]] = v[j]
    assert np.array_equal(Y.cpu().numpy(), dense)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 7, 16, 33, 64, 100])
def test_batch_spmv(ks, orc, k):
    A = g.laplacian_3d(12)
    dA = _up(ks, A)
    X = np.random.default_rng(k).standard_normal((A.ncols, k))
    Y = torch.zeros((A.nrows, k), dtype=torch.float64, device="cuda")
    ks.batch_spmv(dA, _dev(X), Y)
    Y_ref = orc.batch_spmv(A.rowptr, A.colind, A.values, X)
    assert np.array_equal(Y.cpu().numpy(), Y_ref)
    y1 = orc.spmv(A.rowptr, A.colind, A.values, X[:, 0])
    assert np.array_equal(Y.cpu().numpy()[:, 0], y1)


@pytest.mark.parametrize("k", [2, 5, 8, 64, 128])
def test_batch_spmv_irregular(ks, orc, k):
    # Zipf rows of 1..300 entries: rows longer than one 8-entry batch, and
    # lane groups of one wave walking rows of different lengths.
    A = g.zipf_matrix(4000, 2.0, 300, seed=11)
    dA = _up(ks, A)
    X = np.random.default_rng(100 + k).standard_normal((A.ncols, k))
    Y = torch.zeros((A.nrows, k), dtype=torch.float64, device="cuda")
    ks.batch_spmv(dA, _dev(X), Y)
    assert np.array_equal(Y.cpu().numpy(), orc.batch_spmv(A.rowptr, A.colind, A.values, X))


@pytest.mark.parametrize("k", [2, 16, 37, 128])
def test_batch_spmv_xcd_chunks(ks, orc, k):
    # 40^3 = 64,000 rows: enough row blocks for the chunked XCD order at
    # k >= 16, and a partial last chunk.
    A = g.laplacian_3d(40)
    dA = _up(ks, A)
    X = np.random.default_rng(200 + k).uniform(-1, 1, (A.ncols, k))
    Y = torch.zeros((A.nrows, k), dtype=torch.float64, device="cuda")
    ks.batch_spmv(dA, _dev(X), Y)
    assert np.array_equal(Y.cpu().numpy(), orc.batch_spmv(A.rowptr, A.colind, A.values, X))


@pytest.mark.parametrize("k", [64, 65, 128, 192, 256])
def test_batch_spmv_wide(ks, orc, k):
    # A wave per row: 37^2 = 1369 rows leaves a last wave with fewer rows
    # than the others; k = 192 and 256 take two column passes (the first with
    # idle lanes in its second pass, the second without); k = 65 the 8-byte
    # form.  The appended Zipf rows (up to 40 entries) take the direct walk.
    A2 = g.laplacian_2d(37)
    Z = g.zipf_matrix(A2.nrows, 2.0, 40, seed=12)
    rp = np.concatenate([A2.rowptr, A2.rowptr[-1] + Z.rowptr[1:]]).astype(np.int32)
    ci = np.concatenate([A2.colind, Z.colind]).astype(np.int32)
    va = np.concatenate([A2.values, Z.values])
    n = 2 * A2.nrows - 1
    rp = rp[: n + 1]
    ci, va = ci[: rp[-1]], va[: rp[-1]]
    dA = ks.upload_csr(rp, ci, va, n, A2.ncols)
    X = np.random.default_rng(300 + k).standard_normal((A2.ncols, k))
    Y = torch.zeros((n, k), dtype=torch.float64, device="cuda")
    ks.batch_spmv(dA, _dev(X), Y)
    assert np.array_equal(Y.cpu().numpy(), orc.batch_spmv(rp, ci, va, X))


@pytest.mark.parametrize("k", [6, 40])
def test_batch_spmv_8byte_aligned(ks, orc, k):
    # X and Y one element into their buffers: the 8-byte-load form for even k
    # (k = 40: the wave-per-row kernel, which 8-byte loads take above k = 32).
    A = g.laplacian_2d(30)
    dA = _up(ks, A)
    X = np.random.default_rng(5).standard_normal((A.ncols, k))
    xb = torch.zeros(A.ncols * k + 1, dtype=torch.float64, device="cuda")
    yb = torch.zeros(A.nrows * k + 1, dtype=torch.float64, device="cuda")
    Xd = xb[1:].view(A.ncols, k)
    Xd.copy_(_dev(X))
    Yd = yb[1:].view(A.nrows, k)
    assert Xd.data_ptr() % 16 == 8 and Yd.data_ptr() % 16 == 8
    ks.batch_spmv(dA, Xd, Yd)
    assert np.array_equal(Yd.cpu().numpy(), orc.batch_spmv(A.rowptr, A.colind, A.values, X))
    assert yb[0].item() == 0.0


def test_batch_shape_errors(ks):
    A = g.laplacian_2d(4)
    dA = _up(ks, A)
    with pytest.raises(ValueError):
        ks.batch_spmv(dA, torch.zeros((15, 3), dtype=torch.float64, device="cuda"),
                      torch.zeros((16, 3), dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        ks.batch_spmv(dA, torch.zeros((16, 3), dtype=torch.float64, device="cuda"),
                      torch.zeros((16, 2), dtype=torch.float64, device="cuda"))


# -- vector ops (src/vec_ops.hpp) ---------------------------------------------

@pytest.mark.parametrize("n", [0, 1, 63, 64, 1000, 123457, 3_000_001])
def test_vec_ops(ks, orc, n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n)
    y = rng.uniform(-1, 1, n)
    dx, dy = _dev(x), _dev(y)
    z = torch.empty(n, dtype=torch.float64, device="cuda")
    ks.axpy(0.37, dx, dy, z)
    assert np.array_equal(z.cpu().numpy(), orc.axpy(0.37, x, y))
    ks.axpy2(-1.5, dx, 0.25, dy, z)
    assert np.array_equal(z.cpu().numpy(), orc.axpy2(-1.5, x, 0.25, y))
    w = dx.clone()
    ks.scale(w, 3.0)
    assert np.array_equal(w.cpu().numpy(), orc.scale(x, 3.0))
    ks.fill(w, 2.5)
    assert np.all(w.cpu().numpy() == 2.5)
    ks.copy(dy, w)
    assert np.array_equal(w.cpu().numpy(), y)
    d = ks.dot(dx, dy)
    exact = math.fsum((x * y).tolist()) if n else 0.0
    assert abs(d - exact) <= 1e-14 * max(1.0, float(np.abs(x * y).sum()))
    nn = ks.norm2(dx)
    assert abs(nn - math.sqrt(math.fsum((x * x).tolist()))) <= 1e-14 * max(1.0, nn)


def test_dot_integer_exact(ks):
    n = 1_000_003
    x = (np.arange(n) % 17 - 8).astype(np.float64)
    y = (np.arange(n) % 5 - 2).astype(np.float64)
    assert ks.dot(_dev(x), _dev(y)) == float(np.dot(x.astype(np.int64), y.astype(np.int64)))


def test_axpy_aliasing(ks, orc):
    x = np.random.default_rng(9).uniform(-1, 1, 5000)
    y = np.random.default_rng(10).uniform(-1, 1, 5000)
    dx, dy = _dev(x), _dev(y)
    ks.axpy(2.0, dx, dy, dy)  # z aliases y (vec_ops.hpp:14)
    assert np.array_equal(dy.cpu().numpy(), orc.axpy(2.0, x, y))


def test_autotuner_keeps_results(ks, orc, tmp_path):
    """The tuner times candidate schedules on the matrix and applies the
    fastest; the result is unchanged (bit-exact rows of <= 32 entries) and
    the choice is cached by structure."""
    from num_amd.autotune import tune_spmv
    cache = str(tmp_path / "sched.json")
    for A in (g.laplacian_3d(40), g.zipf_matrix(n=20000, kmax=3000, seed=5)):
        dA = _up(ks, A)
        x = np.random.default_rng(9).uniform(-1, 1, A.ncols)
        y_ref = orc.spmv(A.rowptr, A.colind, A.values, x)
        t = tune_spmv(dA, reps=3, cache_file=cache)
        assert not t["cached"] and t["timed"] > 4 and t["us"] <= t["default_us"]
        y = _spmv(ks, dA, x)
        short = np.diff(A.rowptr) <= 32
        assert np.array_equal(y[short], y_ref[short]) and rel(y, y_ref) < 1e-12
        again = tune_spmv(_up(ks, A), cache_file=cache)
        assert again["cached"] and again["schedule"] == t["schedule"]


# -- host round-trip helpers (spmv_host_roundtrip / batch_spmv_host_roundtrip)

def test_spmv_host_roundtrip_kats(ks, orc):
    # spmv_kernel.hpp:186-217 through num_spmv_host: the 4x4 tridiagonal KAT
    rp = np.array([0, 2, 5, 8, 10], np.int32)
    ci = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3], np.int32)
    v = np.array([2, -1, -1, 2, -1, -1, 2, -1, -1, 2], np.float64)
    y = ks.spmv_host(rp, ci, v, 4, 4, np.array([1.0, 2, 3, 4]))
    assert y.tolist() == [0.0, 0.0, 0.0, 5.0]
    # the LCG KAT (unsorted, repeated columns) bit-exact against the oracle
    A, x = g.lcg_matrix()
    y = ks.spmv_host(A.rowptr, A.colind, A.values, A.nrows, A.ncols, x)
    assert np.array_equal(y, orc.spmv(A.rowptr, A.colind, A.values, x))


def test_batch_spmv_host_roundtrip_kats(ks, orc):
    # batch_spmv_kernel.hpp:179-209 through num_batch_spmv_host: X = I -> Y = A
    rp = np.array([0, 2, 5, 8, 10], np.int32)
    ci = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3], np.int32)
    v = np.array([2, -1, -1, 2, -1, -1, 2, -1, -1, 2], np.float64)
    Y = ks.batch_spmv_host(rp, ci, v, 4, 4, np.eye(4))
    dense = np.zeros((4, 4))
    for r in range(4):
        dense[r, ci[rp[r]:rp[r + 1]]] = v[rp[r]:rp[r + 1]]
    assert np.array_equal(Y, dense)
    A, _ = g.lcg_matrix()
    X = np.random.default_rng(4).uniform(-1, 1, (A.ncols, 5))
    Y = ks.batch_spmv_host(A.rowptr, A.colind, A.values, A.nrows, A.ncols, X)
    assert np.array_equal(Y, orc.batch_spmv(A.rowptr, A.colind, A.values, X))


def test_host_roundtrip_errors(ks):
    rp = np.array([0, 1, 2], np.int32)
    with pytest.raises(ValueError):  # column out of range, refused before upload
        ks.spmv_host(rp, np.array([0, 5], np.int32), np.ones(2), 2, 2, np.ones(2))
    with pytest.raises(ValueError):
        ks.spmv_host(rp, np.array([0, 1], np.int32), np.ones(2), 2, 2, np.ones(3))
    with pytest.raises(ValueError):
        ks.batch_spmv_host(rp, np.array([0, 1], np.int32), np.ones(2), 2, 2, np.ones((3, 2)))


def test_device_create_rejects_bad_columns(ks):
    # num_csr_create over device arrays checks column ids on the device
    import ctypes as C
    from num_amd._lib import call
    rp = np.array([0, 2, 3], np.int32)
    for bad in ([0, 2, 1], [0, -1, 1]):
        d_rp = torch.from_numpy(rp).cuda()
        d_ci = torch.tensor(bad, dtype=torch.int32, device="cuda")
        d_v = torch.ones(3, dtype=torch.float64, device="cuda")
        h = C.c_void_p()
        with pytest.raises(ValueError, match="column index"):
            call("num_csr_create", rp.ctypes.data_as(C.c_void_p), C.c_void_p(d_rp.data_ptr()),
                 C.c_void_p(d_ci.data_ptr()), C.c_void_p(d_v.data_ptr()), 2, 2, C.byref(h))
