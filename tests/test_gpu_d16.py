"""GPU parity of the delta-coded column ids (num_csr_attach_d16): the FP64
SpMV and CG read uint16 column gaps instead of int32 columns where every row
holds <= 32 entries in ascending column order and the matrix streams from
HBM.  The columns and their order
are unchanged, so every result must be bit-identical to the int32 path
(index="int32") and to the oracle's row loop (src/spmv_kernel.hpp:90-97);
matrices outside the format must stay on int32 columns.
"""
import numpy as np
import pytest
import torch

from num_amd import generators as g

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _up(ks, rp, ci, va, n, m, index="auto"):
    # The codes are read only when the schedule streams the matrix (nt64);
    # these test matrices would stay in the Infinity Cache, so stream them.
    # The windowed layout, which would replace the codes for these stencils,
    # is switched off (tests/test_gpu_windows.py covers it).
    dA = ks.upload_csr(rp, ci, va, n, m, index=index)
    dA.set_windows(False)
    dA.set_schedule(nt64=1)
    return dA


def _spmv(ks, dA, x):
    y = torch.full((dA.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    ks.spmv(dA, _dev(x), y)
    return y.cpu().numpy()


def _plane_stencil(px, py, nz):
    """3-D 7-point stencil on px x py x nz: column gaps of px*py - px (the
    +-z neighbour over the +-y one), the largest code a 3-D grid needs."""
    return g.stencil((px, py, nz), 0.25)


def _with_empty_rows(A, every=7):
    """A copy of A with every `every`-th row emptied (the tile bases then
    come from the remaining rows)."""
    lens = np.diff(A.rowptr)
    keep = np.ones(A.nnz, bool)
    for r in range(0, A.nrows, every):
        keep[A.rowptr[r]:A.rowptr[r + 1]] = False
    lens = lens.copy()
    lens[::every] = 0
    rp = np.zeros(A.nrows + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    return rp, A.colind[keep].astype(np.int32), A.values[keep]


CASES = {
    "lap3d_20": lambda: g.laplacian_3d(20),
    "lap2d_70": lambda: g.laplacian_2d(70),
    "diffusion_16": lambda: g.diffusion_3d(16, sigma=1.0, seed=3),
    "plane_60025": lambda: _plane_stencil(245, 245, 3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_spmv_d16_bit_identical(ks, orc, name):
    A = CASES[name]()
    dA = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols)
    dI = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols, index="int32")
    assert dA.index_format == "delta16"
    assert dI.index_format == "int32"
    x = np.random.default_rng(7).uniform(-1, 1, A.ncols)
    y = _spmv(ks, dA, x)
    assert np.array_equal(y, _spmv(ks, dI, x))
    assert np.array_equal(y, orc.spmv(A.rowptr, A.colind, A.values, x))


def test_spmv_d16_empty_rows(ks, orc):
    A = g.laplacian_3d(14)
    rp, ci, va = _with_empty_rows(A)
    dA = _up(ks, rp, ci, va, A.nrows, A.ncols)
    assert dA.index_format == "delta16"
    x = np.random.default_rng(8).standard_normal(A.ncols)
    y = _spmv(ks, dA, x)
    assert np.array_equal(y, orc.spmv(rp, ci, va, x))
    assert np.all(y[::7] == 0.0)


def test_d16_refused_outside_the_format(ks, orc):
    # gaps of 280*250 - 280 = 69,720 > 65,535
    A = _plane_stencil(280, 250, 3)
    dA = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols)
    assert dA.index_format == "int32"
    x = np.random.default_rng(9).uniform(-1, 1, A.ncols)
    assert np.array_equal(_spmv(ks, dA, x), orc.spmv(A.rowptr, A.colind, A.values, x))
    # a row out of column order
    B = g.laplacian_2d(20)
    ci = B.colind.copy()
    va = B.values.copy()
    r = 45
    s, e = B.rowptr[r], B.rowptr[r + 1]
    ci[s:e] = ci[s:e][::-1]
    va[s:e] = va[s:e][::-1]
    dB = _up(ks, B.rowptr, ci, va, B.nrows, B.ncols)
    assert dB.index_format == "int32"
    x = np.random.default_rng(10).uniform(-1, 1, B.ncols)
    assert np.array_equal(_spmv(ks, dB, x), orc.spmv(B.rowptr, ci, va, x))
    # rows longer than 32 entries
    Z = g.zipf_matrix(3000, 2.0, 100, seed=3)
    assert _up(ks, Z.rowptr, Z.colind, Z.values, Z.nrows, Z.ncols).index_format == "int32"


def test_d16_survives_reschedule(ks, orc):
    A = g.laplacian_3d(24)
    dA = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols)
    x = np.random.default_rng(11).uniform(-1, 1, A.ncols)
    ref = orc.spmv(A.rowptr, A.colind, A.values, x)
    for rpb, npb in ((64, 0), (200, 900), (0, 0)):
        dA.set_schedule(rows_per_block=rpb, nnz_per_block=npb, nt64=1)
        assert dA.index_format == "delta16"
        assert np.array_equal(_spmv(ks, dA, x), ref)
    dA.set_schedule(nt64=0)  # cache-resident policy: int32 columns, same result
    assert dA.index_format == "int32"
    assert np.array_equal(_spmv(ks, dA, x), ref)


@pytest.mark.parametrize("name", ["lap3d_20", "diffusion_16"])
def test_cg_d16_bit_identical(ks, name):
    A = CASES[name]()
    b = _dev(np.random.default_rng(12).uniform(-1, 1, A.nrows))
    out = []
    for index in ("auto", "int32"):
        dA = _up(ks, A.rowptr, A.colind, A.values, A.nrows, A.ncols, index=index)
        x = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
        res = ks.cg_solve_local(dA, b, x, tol=1e-10, max_iter=2000)
        out.append((res, x.cpu().numpy()))
    (r0, x0), (r1, x1) = out
    assert r0["converged"] and r0["iters"] == r1["iters"]
    assert r0["residual"] == r1["residual"]
    assert np.array_equal(x0, x1)


def _banded(n, w, lead, seed):
    """n rows of exactly w ascending columns (min(r, n - w) + k), except row 0
    with `lead` entries: every later tile then holds exactly its entry
    capacity and starts at a column-id offset that moves with `lead`."""
    lens = np.full(n, w, np.int64)
    lens[0] = lead
    rp = np.zeros(n + 1, np.int32)
    np.cumsum(lens, out=rp[1:])
    ci = np.concatenate([min(r, n - w) + np.arange(lens[r]) for r in range(n)]).astype(np.int32)
    va = np.random.default_rng(seed).uniform(-1, 1, ci.size)
    return rp, ci, va


@pytest.mark.parametrize("w,npb", [(5, 1280), (7, 0), (12, 3072), (32, 4096)])
def test_d16_full_tiles_every_offset(ks, orc, w, npb):
    # Tiles filled to the staging capacity (256 threads x 5/7/12/16 entries)
    # at start offsets that include the worst ones, an odd value offset and a
    # code offset of 7 mod 8 (all 8 code offsets for w = 7, 12, 32): the LDS
    # stage is sized to exactly what such a tile needs.
    n = 256 * 12
    x = np.random.default_rng(13).uniform(-1, 1, n)
    for lead in range(8):
        rp, ci, va = _banded(n, w, lead, seed=lead)
        dA = ks.upload_csr(rp, ci, va, n, n)
        dA.set_windows(False)
        dA.set_schedule(rows_per_block=0, nnz_per_block=npb, nt64=1)
        assert dA.index_format == "delta16"
        assert np.array_equal(_spmv(ks, dA, x), orc.spmv(rp, ci, va, x)), lead
