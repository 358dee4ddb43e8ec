"""CPU tests of the host-side logic: input generators, the row partition and
ghost layout of one rank's share (the semantics tests/test_halo_exchange.py of
the reference checks with several MPI ranks, run here as single-process
simulations of every rank), the golden fixtures, and the C ABI library
(loads and exports every symbol include/num_amd.h declares -- no compute
calls, there is no GPU here)."""
import os
import re

import numpy as np
import pytest

import oracle
from num_amd import generators as g

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


# ------------------------------------------------------------ generators ---

@pytest.mark.parametrize("L", [1, 2, 5, 64])
def test_laplacian_2d_shape(L):
    A = g.laplacian_2d(L)
    n = L * L
    assert A.nrows == n and A.nnz == 5 * n - 4 * L
    S = A.to_scipy()
    assert abs(S - S.T).max() == 0 and S.has_sorted_indices
    assert np.all(S.diagonal() == 4.0)


@pytest.mark.parametrize("L", [1, 3, 16])
def test_laplacian_3d_shape(L):
    A = g.laplacian_3d(L)
    n = L ** 3
    assert A.nrows == n and A.nnz == 7 * n - 6 * L * L
    S = A.to_scipy()
    assert abs(S - S.T).max() == 0 and S.has_sorted_indices


def test_stencil_shift_and_rect():
    A = g.stencil((7, 3), shift=0.25)
    S = A.to_scipy()
    assert A.nrows == 21 and np.all(S.diagonal() == 4.25)
    # row sums: shift + number of missing neighbours
    ones = S @ np.ones(21)
    assert ones.min() >= 0.25


def test_diffusion_spd():
    A = g.diffusion_3d(8, sigma=1.0, seed=3)
    S = A.to_scipy().toarray()
    assert np.array_equal(S, S.T)
    assert np.linalg.eigvalsh(S).min() > 0
    assert A.nnz == 7 * 512 - 6 * 64


def test_zipf_properties():
    A = g.zipf_matrix(n=50_000, kmax=4096, seed=2)
    k = np.diff(A.rowptr)
    assert k.min() >= 1 and k.max() <= 4096
    assert np.all(A.values > -1) and np.all(A.values < 1)
    for i in np.random.default_rng(0).integers(0, A.nrows, 200):
        c = A.colind[A.rowptr[i]:A.rowptr[i + 1]]
        assert np.all(np.diff(c) > 0)  # sorted, no repeats
    B = g.zipf_matrix(n=50_000, kmax=4096, seed=2)
    assert np.array_equal(A.colind, B.colind) and np.array_equal(A.values, B.values)


def test_lcg_matrix_matches_reference_text():
    A, x = g.lcg_matrix()
    assert A.nrows == 10_000
    # first LCG draws from seed 42: seed*1664525 + 1013904223 mod 2^32
    s = (42 * 1664525 + 1013904223) & 0xFFFFFFFF
    assert x[0] == (s % 1000) * 0.001 - 0.5
    assert np.all(A.values[A.rowptr[:-1]] == 7.0) and np.all(A.colind[A.rowptr[:-1]] ==
                                                               np.arange(A.nrows))


def test_batched_systems():
    sysl = g.batched_systems(count=256, seed=4)
    ns = np.array([d[0] * d[1] for d, _, _ in sysl])
    assert len(sysl) == 256 and ns.min() >= 32 and ns.max() <= 70000
    assert all(0.01 <= s <= 1.0 for _, s, _ in sysl)


def test_served_requests():
    r = g.served_requests(rate=100.0, duration=10.0, seed=5)
    assert 800 < len(r) < 1200
    t = np.array([q["t"] for q in r])
    assert np.all(np.diff(t) > 0) and t[-1] < 10.0


# ---------------------------------------------- partition and ghost layout ---

def test_hash_uniform_is_splitmix64():
    """b of a served request: SplitMix64 outputs for state 0 are the published
    e220a8397b1dcdaf, 6e789e6aa1b965f4, 06c45d188009454f (seed 0 draws
    i = 0, 1, 2), mapped to 2 * (z >> 11) * 2^-53 - 1."""
    known = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    want = [2.0 * ((z >> 11) * 2.0**-53) - 1.0 for z in known]
    assert g.hash_uniform(3, 0).tolist() == want
    # counter-based: any window of indices is a slice of the full draw
    full = g.hash_uniform(1000, 77)
    assert np.array_equal(g.hash_uniform(10, 77, start=500), full[500:510])
    assert full.min() >= -1.0 and full.max() < 1.0
    assert not np.array_equal(g.hash_uniform(1000, 78), full)
    big = g.hash_uniform(200_000, 5)
    assert abs(big.mean()) < 0.01 and abs(big.var() - 1.0 / 3.0) < 0.01


@pytest.mark.parametrize("dims", [(1,), (9,), (1, 1), (13, 7), (2, 2, 2), (6, 5, 4), (33, 1, 2)])
def test_stencil_nnz_and_layout(dims):
    from num_amd.batched import stencil_layout
    A = g.stencil(dims, 0.0)
    assert g.stencil_nnz(dims) == A.nnz
    lay = stencil_layout(dims, True)
    assert lay["n"] == A.nrows and lay["nnz"] == A.nnz
    offs = [lay[k] for k in ("x", "b", "rowptr", "colind", "values", "values32", "work")]
    assert offs == sorted(offs) and all(o % 256 == 0 for o in offs)
    assert lay["b"] >= 8 * A.nrows and lay["work"] - lay["values32"] >= 4 * A.nnz


@pytest.mark.parametrize("name", ["laplacian_2d", "laplacian_3d", "lattice_gauge",
                                  "random_sparse", "power_law"])
def test_matrix_families(name):
    """The reference's families (python/matrix_generators.py): square,
    symmetric, sorted int32 CSR; SPD by strict diagonal dominance (the
    Laplacians and power_law are weakly dominant: PSD / SPD by structure)."""
    A = g.get_matrix(name, 3000)
    S = A.to_scipy()
    assert A.nrows == A.ncols and A.rowptr.dtype == np.int32 and A.colind.dtype == np.int32
    assert abs(S - S.T).max() == 0.0
    for i in range(0, A.nrows, 97):
        c = A.colind[A.rowptr[i]:A.rowptr[i + 1]]
        assert np.all(np.diff(c) > 0)
    d = S.diagonal()
    off = np.asarray(abs(S).sum(axis=1)).ravel() - np.abs(d)
    if name in ("lattice_gauge", "random_sparse"):
        assert np.all(d > off)
    else:
        assert np.all(d >= off - 1e-12)
    if name == "power_law":
        # the reference subtracts an adjacency built with -1 entries, so the
        # matrix is D + Adj (signless Laplacian): row sums are 2 * degree
        assert np.array_equal(S @ np.ones(A.nrows), 2.0 * d)
        assert np.diff(A.rowptr).max() > 10 * np.median(np.diff(A.rowptr))  # heavy tail
    with pytest.raises(ValueError):
        g.get_matrix("nope", 10)


def test_lattice_gauge_small_torus_closed_form():
    """L = 3: every site has 4 distinct neighbours; A + A^T halves each
    hopping pair, so A[i, j] = -0.25 (Re e^{i phi} + Re e^{-i phi}) for the
    shared link."""
    A = g.lattice_gauge(3, mass=0.1, seed=42).to_scipy().toarray()
    rng = np.random.default_rng(42)
    phi_x = rng.uniform(0, 2 * np.pi, (3, 3))
    # site (x=0, y=0) -> (x=1, y=0) uses phi_x[0, 0] both ways
    assert A[0, 0] == 4.1
    assert np.isclose(A[0, 1], -0.5 * np.cos(phi_x[0, 0]))


def test_row_partition_coverage_and_balance():
    # tests/test_halo_exchange.py:89-141 semantics at N = 127
    for size in range(1, 9):
        spans = [g.row_split(127, r, size) for r in range(size)]
        assert spans[0][0] == 0 and spans[-1][1] == 127
        assert all(spans[i][1] == spans[i + 1][0] for i in range(size - 1))
        lens = [e - s for s, e in spans]
        assert max(lens) - min(lens) <= 1


def test_ghosts_1d():
    # tests/test_halo_exchange.py:146-200: 1-D Laplacian N=200, interior ranks
    # need exactly start-1 and end
    A = g.laplacian_1d(200)
    for size in (2, 4, 5):
        for r in range(size):
            sh = g.slab_share(A.rowptr, A.colind, A.values, 200, 200, r, size)
            want = []
            if r > 0:
                want.append(sh.local_row_start - 1)
            if r < size - 1:
                want.append(sh.local_row_end)
            assert sh.ghost_ids.tolist() == want


def test_ghosts_2d_order():
    # tests/test_halo_exchange.py:205-258: 2-D L=16; ghosts grouped by owner
    # rank ascending, ids ascending (halo_exchange.hpp:62-75,106-128)
    A = g.laplacian_2d(16)
    sh = g.slab_share(A.rowptr, A.colind, A.values, 256, 256, 1, 3)
    assert np.all(np.diff(sh.ghost_owner) >= 0)
    for rank, off, cnt in sh.recv:
        ids = sh.ghost_ids[off:off + cnt]
        assert np.all(np.diff(ids) > 0)
        assert np.all(sh.ghost_owner[off:off + cnt] == rank)


def test_slab_192_share_sizes():
    """One GPU's share of the 8-slab 192^3 partition: rank 3 owns planes
    72-95 and reads ghost planes 71 and 96 (closed form)."""
    L = 192
    A = g.laplacian_3d(L)
    sh = g.slab_share(A.rowptr, A.colind, A.values, A.nrows, A.ncols, 3, 8)
    P = L * L
    assert (sh.local_row_start, sh.local_row_end) == (72 * P, 96 * P)
    assert sh.local.nrows == 884_736 and sh.local.ncols - sh.local.nrows == 73_728
    assert sh.local.nnz == 6_174_720
    assert len(sh.interior_rows) == 811_008 and len(sh.boundary_rows) == 73_728
    assert np.array_equal(sh.ghost_ids[:P], np.arange(71 * P, 72 * P))
    assert np.array_equal(sh.ghost_ids[P:], np.arange(96 * P, 97 * P))
    assert sh.recv == [(2, 0, P), (4, P, P)]


def test_slab_rows_equal_global_rows():
    A = g.laplacian_3d(20)
    x = np.random.default_rng(3).uniform(-1, 1, A.ncols)
    y = oracle.spmv(A.rowptr, A.colind, A.values, x)
    for size in (2, 3, 8):
        for r in range(size):
            sh = g.slab_share(A.rowptr, A.colind, A.values, A.nrows, A.ncols, r, size)
            x_ext = np.concatenate([x[sh.local_row_start:sh.local_row_end], x[sh.ghost_ids]])
            L = sh.local
            yl = oracle.spmv(L.rowptr, L.colind, L.values, x_ext)
            assert np.array_equal(yl, y[sh.local_row_start:sh.local_row_end])
            # interior rows touch no ghost column
            for i in sh.interior_rows[:50]:
                assert np.all(L.colind[L.rowptr[i]:L.rowptr[i + 1]] < L.nrows)


def test_slab_zero_and_basis():
    # x = 0 -> y = 0 exactly; basis vector at a partition boundary column
    # (tests/test_halo_exchange.py:314-415)
    A = g.laplacian_2d(12)
    sh = g.slab_share(A.rowptr, A.colind, A.values, A.nrows, A.ncols, 1, 3)
    L = sh.local
    assert np.all(oracle.spmv(L.rowptr, L.colind, L.values, np.zeros(L.ncols)) == 0.0)
    for col in (sh.local_row_start - 1, sh.local_row_end):
        e = np.zeros(A.ncols)
        e[col] = 1.0
        x_ext = np.concatenate([e[sh.local_row_start:sh.local_row_end], e[sh.ghost_ids]])
        yl = oracle.spmv(L.rowptr, L.colind, L.values, x_ext)
        dense_col = A.to_scipy()[:, col].toarray().ravel()
        assert np.array_equal(yl, dense_col[sh.local_row_start:sh.local_row_end])


def test_slab_bad_column_raises():
    rp = np.array([0, 1, 2], np.int32)
    ci = np.array([0, 5], np.int32)
    with pytest.raises(RuntimeError):
        g.slab_share(rp, ci, np.ones(2), 2, 2, 1, 2)


# ------------------------------------------------------------- fixtures ---

def test_golden_fixtures_reproduce():
    import hashlib

    def digest(*arrays):
        h = hashlib.sha256()
        for a in arrays:
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()

    f = np.load(os.path.join(GOLD, "tridiag4.npz"))
    assert f["y"].tolist() == [0.0, 0.0, 0.0, 5.0]
    assert np.array_equal(oracle.batch_spmv(f["rowptr"], f["colind"], f["values"], np.eye(4)),
                          f["Y"])
    A, x = g.lcg_matrix()
    f = np.load(os.path.join(GOLD, "lcg10k.npz"))
    assert str(f["digest"]) == digest(A.rowptr, A.colind, A.values)
    assert np.array_equal(f["x"], x)
    assert np.array_equal(oracle.spmv(A.rowptr, A.colind, A.values, x), f["y"])
    A = g.laplacian_2d(64)
    f = np.load(os.path.join(GOLD, "lap2d64.npz"))
    assert str(f["digest"]) == digest(A.rowptr, A.colind, A.values)
    assert np.array_equal(oracle.spmv(A.rowptr, A.colind, A.values, f["x"]), f["y"])
    oracle.set_threads(1)
    b = np.random.default_rng(0).uniform(-1, 1, A.nrows)
    xs, r = oracle.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-8, max_iter=2000)
    oracle.set_threads(0)
    assert r["iters"] == int(f["cg_iters"]) and np.array_equal(xs, f["cg_x"])
    A = g.laplacian_3d(16)
    f = np.load(os.path.join(GOLD, "slab16.npz"))
    sh = g.slab_share(A.rowptr, A.colind, A.values, A.nrows, A.ncols, 1, 4)
    assert np.array_equal(sh.ghost_ids, f["ghost_ids"])
    assert str(f["digest"]) == digest(sh.local.rowptr, sh.local.colind, sh.local.values)


# ------------------------------------------------------------ C ABI lib ---

def _declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "num_amd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return sorted(set(re.findall(r"\b(num_[a-z0-9_]+)\s*\(", hdr)))


def test_library_exports_every_declared_symbol():
    from num_amd import _lib
    L = _lib.load()
    syms = _declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), f"{s} declared in include/num_amd.h but not exported"
    assert set(syms) == set(_lib.SIGNATURES), "ctypes table out of sync with the header"
    assert L.num_abi_version() == 1


def test_library_rejects_bad_arguments_without_gpu():
    """Argument checks that run before any device call."""
    import ctypes as C
    from num_amd import _lib
    L = _lib.load()
    assert L.num_spmv(None, None, None, None) == _lib.NUM_EINVAL
    assert "NULL" in _lib.last_error()
    assert L.num_axpy(-1, 1.0, None, None, None, None) == _lib.NUM_EINVAL
    assert L.num_halo_pack(None, -1, None, None, None) == _lib.NUM_EINVAL
    out = C.c_void_p()
    rp = np.array([0, 1], np.int32)
    assert L.num_csr_create_host(rp.ctypes.data_as(C.c_void_p), None, None, 0, 1,
                                 C.byref(out)) == _lib.NUM_EINVAL
    flag = C.c_int32(7)
    # the derived-layout entry points check their arguments before any device call
    assert L.num_csr_attach_d16(None, None, None, C.byref(flag)) == _lib.NUM_EINVAL
    assert L.num_csr_index_info(None, C.byref(flag)) == _lib.NUM_EINVAL
    assert L.num_csr_attach_xcd(None, None, 8, None, None, None) == _lib.NUM_EINVAL
    assert L.num_spmv_xcd(None, None, None, None, None) == _lib.NUM_EINVAL
    assert "NULL" in _lib.last_error()
    with pytest.raises(ValueError):
        _lib.check(_lib.NUM_EINVAL, "x")
    with pytest.raises(RuntimeError):
        _lib.check(_lib.NUM_ERUNTIME, "x")


def test_harness_cpu_leg_and_csv_schema(tmp_path, monkeypatch):
    """The harness's CPU leg and its CSV schema (python/benchmark_spmv.py:
    370-377: backend, N, nnz, elapsed_us, bw_GB_s, gflops)."""
    import csv as _csv
    from num_amd import harness
    monkeypatch.chdir(tmp_path)
    assert harness.main(["spmv", "--N", "4096", "--matrix", "laplacian_2d", "--repeat", "3",
                         "--backend", "scipy"]) == 0
    files = list((tmp_path / "benchmark" / "results").glob("spmv_*.csv"))
    assert len(files) == 1
    rows = list(_csv.DictReader(open(files[0])))
    assert list(rows[0].keys()) == ["backend", "N", "nnz", "elapsed_us", "bw_GB_s", "gflops"]
    assert rows[0]["backend"] == "scipy_cpu" and int(rows[0]["nnz"]) == 5 * 4096 - 4 * 64
    m = harness.spmv_metrics(1e-3, 1000, 5000, k=2)
    assert m["bw_GB_s"] == (5000 * 12 + 2 * 1000 * 8 + 1000 * 8) / 1e-3 / 1e9
    assert m["gflops"] == 2 * 5000 * 2 / 1e-3 / 1e9


def test_serving_host_logic():
    """Request sizing and the latency summary, without a GPU: the stencil
    buffer size is pure host arithmetic in the library."""
    from num_amd.batched import CGEngine, stencil_layout
    from num_amd.serve import Request, summarize, requests_from_trace
    for dims, mixed in (((30, 20), False), ((30, 20), True), ((9, 8, 7), True), ((5,), False)):
        lay = stencil_layout(dims, mixed)
        b = CGEngine.stencil_bytes(dims, mixed)
        assert b == lay["work"] + CGEngine.work_bytes(lay["n"], mixed)
    with pytest.raises(ValueError):
        CGEngine.stencil_bytes((40000, 2), True)   # grid line above 32767 points
    with pytest.raises(ValueError):
        CGEngine.stencil_bytes((2, 2, 2, 2), True)
    tr = g.served_requests(rate=50.0, duration=2.0, seed=5, nmax=5000)
    reqs = requests_from_trace(tr)
    assert [r.rid for r in reqs] == list(range(len(reqs)))
    assert all(r.tol == 1e-10 for r in reqs)
    done = []
    for i, (lat, conv) in enumerate([(0.001, True), (0.004, True), (0.010, False), (0.002, True)]):
        r = Request(i, t_arrival=0.1 * i, dims=(4, 4), shift=0.1, seed=i)
        r.t_done = r.t_arrival + lat
        r.result = {"converged": conv, "iters": 10 + i, "residual": 1e-11 * (i + 1)}
        done.append(r)
    s = summarize(done, wall=1.0, submitted=5)
    assert s["completed"] == 4 and s["converged"] == 3 and s["submitted"] == 5
    assert s["p50_ms"] == pytest.approx(3.0) and s["p99_ms"] <= 10.0
    assert s["solves_per_s"] == pytest.approx(4 / (0.3 + 0.002))
    assert s["mean_iters"] == 11.5 and s["max_residual"] == pytest.approx(4e-11)
