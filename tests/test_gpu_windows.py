"""GPU parity of the windowed sliced layout (num_csr_set_windows): where every
tile of the FP64 schedule reads x from at most 4 contiguous column ranges of
at most 1152 columns in all and every row holds <= 8 entries, the FP64 SpMV
and CG walk a sliced copy of the tile (entry k of row i at slice k, lane i)
and read x from those ranges, fetched once per tile into LDS.  The entries
and their summation order are unchanged, so every result must be
bit-identical to the oracle's row loop (src/spmv_kernel.hpp:90-97) and to
the delta-coded and int32-column kernels; matrices outside the layout must
keep the other index forms.  The layout is built by default from 65,536
rows; the smaller test matrices ask for it (_up).
"""
import numpy as np
import pytest
import torch

from num_amd import generators as g
from _helpers import shuffled_stencil_rows

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _up(ks, rowptr, colind, values, nrows, ncols):
    dA = ks.upload_csr(rowptr, colind, values, nrows, ncols)
    dA.set_windows(True)
    return dA


def _spmv(ks, dA, x):
    y = torch.full((dA.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    ks.spmv(dA, _dev(x), y)
    return y.cpu().numpy()


def _with_empty_rows(A, every=7):
    lens = np.diff(A.rowptr).copy()
    keep = np.ones(A.nnz, bool)
    for r in range(0, A.nrows, every):
        keep[A.rowptr[r]:A.rowptr[r + 1]] = False
    lens[::every] = 0
    rp = np.zeros(A.nrows + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    return rp, A.colind[keep].astype(np.int32), A.values[keep]


def _banded(n, m, w, seed):
    """n x m, row r holds w consecutive columns from min(r, m - w)."""
    rp = np.arange(0, w * (n + 1), w, dtype=np.int32)
    ci = np.concatenate([min(r, m - w) + np.arange(w) for r in range(n)]).astype(np.int32)
    va = np.random.default_rng(seed).uniform(-1, 1, ci.size)
    return rp, ci, va


CASES = {
    "lap3d_20": lambda: g.laplacian_3d(20),
    "lap3d_33": lambda: g.laplacian_3d(33),     # last tile partial, mixed widths
    "lap2d_70": lambda: g.laplacian_2d(70),
    "diffusion_16": lambda: g.diffusion_3d(16, sigma=1.0, seed=3),
    "stencil_96x96x5": lambda: g.stencil((96, 96, 5), 0.25),  # plane 9216: 3 windows
    # periodic 600 x 600 torus: the tiles of the first and last grid rows
    # read the far side's columns (a fourth window), the others three
    "lattice_600": lambda: g.lattice_gauge(600),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("nt64", [1, 0])
def test_spmv_windows_bit_identical(ks, orc, name, nt64):
    # row patterns (every case holds at most 16 per tile), a slot per entry,
    # no windows: the same bits as the oracle's row loop
    A = CASES[name]()
    dA = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols)
    dA.set_schedule(nt64=nt64)
    assert dA.index_format == "windowpat" and dA.patterns["windows"]
    x = np.random.default_rng(7).uniform(-1, 1, A.ncols)
    y = _spmv(ks, dA, x)
    assert np.array_equal(y, orc.spmv(A.rowptr, A.colind, A.values, x))
    dA.set_windows(True, patterns=False)
    assert dA.index_format == "window16"
    assert np.array_equal(_spmv(ks, dA, x), y)
    dA.set_windows(False)
    assert not dA.windowed
    assert np.array_equal(_spmv(ks, dA, x), y)
    dA.set_windows(True)
    assert dA.index_format == "windowpat"
    assert np.array_equal(_spmv(ks, dA, x), y)


def test_row_patterns_fall_back_to_slots(ks, orc):
    # 300 rows: the diagonal and one neighbour at a random offset in
    # +-1..+-12 (25 column offsets: the windows still fit), so a tile holds
    # up to 24 distinct row patterns, more than 16: the windows keep a slot
    # per entry; still bit-identical
    rng = np.random.default_rng(5)
    n = 300
    offs = np.concatenate([-np.arange(1, 13), np.arange(1, 13)])
    rows = []
    for r in range(n):
        d = int(rng.choice(offs))
        c = min(max(r + d, 0), n - 1)
        rows.append(sorted({r, c}))
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(c) for c in rows])
    ci = np.concatenate(rows).astype(np.int32)
    va = rng.uniform(-1, 1, ci.size)
    dA = _up(ks, rp, ci, va, n, n)
    assert dA.index_format == "window16"
    x = rng.uniform(-1, 1, n)
    assert np.array_equal(_spmv(ks, dA, x), orc.spmv(rp, ci, va, x))


def test_windows_unsorted_duplicate_columns(ks, orc):
    # the row patterns keep each row's entries in their stored order, so the
    # sums stay the row loop's: bit-identical with patterns, slots and none
    rp, ci, va, n = shuffled_stencil_rows(90, 60, seed=4)
    x = np.random.default_rng(13).uniform(-1, 1, n)
    ref = orc.spmv(rp, ci, va, x)
    for on, pat, fmt in ((True, True, "windowpat"), (True, False, "window16"),
                         (False, True, None)):
        dA = _up(ks, rp, ci, va, n, n)
        dA.set_windows(on, patterns=pat)
        if fmt:
            assert dA.index_format == fmt
        assert np.array_equal(_spmv(ks, dA, x), ref), (on, pat)


@pytest.mark.parametrize("w", range(1, 9))
def test_windows_every_width(ks, orc, w):
    # tiles of width w: the packed slot groups (4 / 2 / 1 per load) and the
    # FP32 pairs take every shape; rows of w - 1 entries mixed in
    n = 256 * 6 + 37
    rp, ci, va = _banded(n, n, w, seed=w)
    lens = np.full(n, w)
    lens[::3] = max(w - 1, 0)
    keep = np.concatenate([np.arange(rp[r], rp[r] + lens[r]) for r in range(n)]).astype(np.int64)
    rp2 = np.zeros(n + 1, np.int32)
    np.cumsum(lens, out=rp2[1:])
    ci2, va2 = ci[keep], va[keep]
    dA = _up(ks, rp2, ci2, va2, n, n)
    assert dA.windowed
    x = np.random.default_rng(w).uniform(-1, 1, n)
    assert np.array_equal(_spmv(ks, dA, x), orc.spmv(rp2, ci2, va2, x))


def test_windows_empty_rows_and_specials(ks, orc):
    A = g.laplacian_3d(14)
    rp, ci, va = _with_empty_rows(A)
    dA = _up(ks, rp, ci, va, A.nrows, A.ncols)
    assert dA.windowed
    x = np.random.default_rng(8).standard_normal(A.ncols)
    x[5] = np.inf
    x[300] = np.nan
    y = _spmv(ks, dA, x)
    ref = orc.spmv(rp, ci, va, x)
    assert np.array_equal(np.isnan(y), np.isnan(ref))
    ok = ~np.isnan(ref)
    assert np.array_equal(y[ok], ref[ok])
    assert np.all(y[::7] == 0.0)


def _banded_wrap(n, m, w, seed):
    """n x m, row r holds w consecutive columns from r mod (m - w): a band
    that wraps to column 0 past the last column (a few column offsets per
    tile, as the window build needs)."""
    rp = np.arange(0, w * (n + 1), w, dtype=np.int32)
    ci = np.concatenate([r % (m - w) + np.arange(w) for r in range(n)]).astype(np.int32)
    va = np.random.default_rng(seed).uniform(-1, 1, ci.size)
    return rp, ci, va


@pytest.mark.parametrize("n,m", [(3000, 2000), (2000, 3000)])
def test_windows_rectangular(ks, orc, n, m):
    rp, ci, va = _banded_wrap(n, m, 6, seed=n)
    dA = _up(ks, rp, ci, va, n, m)
    assert dA.windowed
    x = np.random.default_rng(9).uniform(-1, 1, m)
    assert np.array_equal(_spmv(ks, dA, x), orc.spmv(rp, ci, va, x))


def test_windows_refused_outside_the_layout(ks, orc):
    x_of = lambda m: np.random.default_rng(10).uniform(-1, 1, m)
    # scattered columns: more than 32 column offsets per tile
    L, _ = g.lcg_matrix(5000, 5, seed=3)
    dL = _up(ks, L.rowptr, L.colind, L.values, L.nrows, L.ncols)
    assert not dL.windowed
    assert np.array_equal(_spmv(ks, dL, x_of(L.ncols)),
                          orc.spmv(L.rowptr, L.colind, L.values, x_of(L.ncols)))
    # rows of 9 entries: wider than the 8 slices
    rp, ci, va = _banded(2000, 2000, 9, seed=1)
    dB = _up(ks, rp, ci, va, 2000, 2000)
    assert not dB.windowed and dB.index_format in ("delta16", "int32")
    assert np.array_equal(_spmv(ks, dB, x_of(2000)), orc.spmv(rp, ci, va, x_of(2000)))
    # 3-D stencils: rows 600 apart make 5 ranges per tile (over 4); rows 250
    # apart merge into a middle range of 256 + 2*250, which with the two
    # plane ranges needs 1268 > 1152 slots
    for dims in ((600, 4, 3), (250, 4, 3)):
        S = g.stencil(dims, 0.25)
        dS = _up(ks, S.rowptr, S.colind, S.values, S.nrows, S.ncols)
        assert not dS.windowed, dims
        assert np.array_equal(_spmv(ks, dS, x_of(S.ncols)),
                              orc.spmv(S.rowptr, S.colind, S.values, x_of(S.ncols)))
    # switched off on request
    A = g.laplacian_3d(12)
    dA = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols)
    dA.set_windows(False)
    assert not dA.windowed


def test_windows_survive_reschedule(ks, orc):
    A = g.laplacian_3d(24)
    dA = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols)
    x = np.random.default_rng(11).uniform(-1, 1, A.ncols)
    ref = orc.spmv(A.rowptr, A.colind, A.values, x)
    for rpb, npb, nt in ((64, 0, 1), (200, 900, 0), (0, 0, -1)):
        dA.set_schedule(rows_per_block=rpb, nnz_per_block=npb, nt64=nt)
        assert dA.windowed, (rpb, npb, nt)
        assert np.array_equal(_spmv(ks, dA, x), ref)
    dA.set_windows(False)
    dA.set_schedule(rows_per_block=128)
    assert not dA.windowed  # the switch outlives the new tiles
    assert np.array_equal(_spmv(ks, dA, x), ref)


@pytest.mark.parametrize("name", ["lap3d_20", "diffusion_16", "lap3d_33", "lattice_600"])
@pytest.mark.parametrize("nt64", [1, 0])
def test_cg_windows_bit_identical(ks, name, nt64):
    # the windowed CG kernel reads p_new of its own rows from the windows:
    # same bits as the delta-coded / int32 kernels, which compute it again,
    # with row patterns or a slot per entry
    A = CASES[name]()
    b = _dev(np.random.default_rng(12).uniform(-1, 1, A.nrows))
    out = []
    for on, pat in ((True, True), (True, False), (False, True)):
        dA = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols)
        dA.set_windows(on, patterns=pat)
        dA.set_schedule(nt64=nt64)
        assert dA.windowed == on
        if on:
            assert dA.index_format == ("windowpat" if pat else "window16")
        x = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
        res = ks.cg_solve_local(dA, b, x, tol=1e-10, max_iter=2000)
        out.append((res, x.cpu().numpy()))
    (r0, x0) = out[0]
    assert r0["converged"]
    for r1, x1 in out[1:]:
        assert r0["iters"] == r1["iters"] and r0["residual"] == r1["residual"]
        assert np.array_equal(x0, x1)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", ["diffusion_16", "lap3d_33", "stencil_96x96x5"])
def test_mixed_windows_streamed(ks, orc, name):
    # Mixed path 1 sends the mixed solver's FP32 steps and refinements
    # through the windowed sliced kernel (FP32 copy streamed: nt32 = 1, the
    # policy of an FP32 matrix too large for the Infinity Cache); path 0,
    # with the FP32 copy cached (nt32 = 0), through the staged pair kernel.
    # Both must reach the FP64 true residual and the oracle's FP64 solution;
    # their per-tile partials are grouped differently (one per tile against
    # one per pair), so the step counts may differ by rounding.
    A = CASES[name]()
    b = np.random.default_rng(21).uniform(-1, 1, A.nrows)
    x64, _ = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-12, max_iter=20000)
    steps = {}
    for nt32 in (1, 0):
        dA = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols)
        dA.ensure_f32()
        dA.set_schedule(nt32=nt32)
        dA.set_mixed_path(nt32)
        assert dA.windowed
        x = _dev(np.zeros(A.nrows))
        r = ks.cg_solve_mixed(dA, _dev(b), x, tol=1e-10)
        assert r["converged"] and r["residual"] < 1e-10, (nt32, r)
        xh = x.cpu().numpy()
        true = np.linalg.norm(b - A.to_scipy() @ xh) / np.linalg.norm(b)
        assert true < 1e-10 and abs(true - r["residual"]) <= 1e-12 * r["residual"] + 1e-16
        assert _rel(xh, x64) < 1e-7
        steps[nt32] = r["inner_total"]
        if nt32:
            # the windowed step over the row patterns and over a slot per
            # entry: the same bits
            assert dA.patterns["windows"]
            dA.set_windows(True, patterns=False)
            dA.set_schedule(nt32=1)
            assert not dA.patterns["windows"]
            x2 = _dev(np.zeros(A.nrows))
            r2 = ks.cg_solve_mixed(dA, _dev(b), x2, tol=1e-10)
            assert r2 == r and np.array_equal(x2.cpu().numpy(), xh)
    assert abs(steps[1] - steps[0]) <= max(2, 0.02 * steps[0]), steps


@pytest.mark.parametrize("rpb", [100, 256])
def test_mixed_windows_odd_tiles(ks, orc, rpb):
    # the windowed FP32 step takes two tiles per workgroup: tiles of 100
    # rows (the pattern ids of lanes past a tile's rows unused) and a tile
    # count that leaves the last pair one tile, with patterns and with slots
    A = g.stencil((31, 29, 23), 0.01)
    b = np.random.default_rng(23).uniform(-1, 1, A.nrows)
    x64, _ = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-12, max_iter=20000)
    out = []
    for pat in (True, False):
        dA = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols)
        dA.ensure_f32()
        dA.set_windows(True, patterns=pat)
        dA.set_schedule(rows_per_block=rpb, nt32=1)
        dA.set_mixed_path(1)
        nb = dA.schedule_info()["nblocks"]
        assert dA.windowed and dA.patterns["windows"] == pat and nb % 2 == 1, nb
        x = _dev(np.zeros(A.nrows))
        r = ks.cg_solve_mixed(dA, _dev(b), x, tol=1e-10)
        xh = x.cpu().numpy()
        assert r["converged"]
        assert np.linalg.norm(b - A.to_scipy() @ xh) / np.linalg.norm(b) < 1e-10
        assert _rel(xh, x64) < 1e-7
        out.append((r, xh))
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])


def test_windows_default_from_65536_rows(ks):
    # built by default for matrices of at least 65,536 rows only
    small = g.laplacian_3d(40)   # 64,000 rows
    large = g.laplacian_3d(41)   # 68,921 rows
    ds = ks.upload_csr(small.rowptr, small.colind, small.values, small.nrows, small.ncols)
    dl = ks.upload_csr(large.rowptr, large.colind, large.values, large.nrows, large.ncols)
    assert not ds.windowed and dl.windowed
    ds.set_windows(True)
    dl.set_windows(False)
    assert ds.windowed and not dl.windowed
    ds.set_windows(None)
    dl.set_windows(None)
    assert not ds.windowed and dl.windowed
    from num_amd._lib import call
    with pytest.raises(ValueError):
        call("num_csr_set_windows", dl.handle, 3, None)
    with pytest.raises(ValueError):
        dl.set_mixed_path(3)
