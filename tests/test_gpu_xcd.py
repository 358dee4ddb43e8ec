"""GPU parity of the column-partitioned SpMV (num_spmv_xcd): each XCD
gathers from its own x blocks, then each row's per-block partial sums are
added in block order.  Rows whose entries lie in one block (and hold <= 32
entries) must be bit-identical to the oracle's row loop
(src/spmv_kernel.hpp:90-97); the whole product within 1e-12 relative (the
reference's bar, tests/test_spmv_kernel.cpp:205)."""
import numpy as np
import pytest
import torch

from num_amd import generators as g

pytestmark = pytest.mark.gpu


def _check(ks, orc, rp, ci, vv, n, m, seed):
    from num_amd import kokkos_spmv as kmod
    dA = ks.upload_csr(rp, ci, vv, n, m, layout="xcd")
    assert dA.layout == "xcd"
    x = np.random.default_rng(seed).uniform(-1, 1, m)
    y = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    ks.spmv(dA, torch.from_numpy(x).cuda(), y)
    got = y.cpu().numpy()
    ref = orc.spmv(rp, ci, vv, x)
    assert np.linalg.norm(got - ref) <= 1e-12 * max(np.linalg.norm(ref), 1e-300)
    nb = kmod._xcd_plan(rp, ci, vv, n, m)[0]
    bs = -(-m // nb)
    lens = np.diff(rp)
    one = np.array([len(np.unique(ci[rp[r]:rp[r + 1]] // bs)) <= 1 for r in range(n)])
    exact = one & (lens <= 32)
    assert np.array_equal(got[exact], ref[exact])
    # repeat: the same bits
    y2 = torch.empty_like(y)
    ks.spmv(dA, torch.from_numpy(x).cuda(), y2)
    assert torch.equal(y, y2)
    # the row layout of the same matrix agrees to rounding, exactly on short rows
    dR = ks.upload_csr(rp, ci, vv, n, m, layout="rows")
    assert dR.layout == "rows"
    yr = torch.empty_like(y)
    ks.spmv(dR, torch.from_numpy(x).cuda(), yr)
    yr = yr.cpu().numpy()
    assert np.array_equal(yr[exact], got[exact])
    return dA


def test_xcd_zipf(ks, orc):
    Z = g.zipf_matrix(200_000, 2.0, 600, seed=13)
    _check(ks, orc, Z.rowptr, Z.colind, Z.values, Z.nrows, Z.ncols, 1)


def test_xcd_wide_x_many_blocks(ks, orc):
    # 5M columns: 16 blocks (two per XCD), rows of 1..40 random columns
    rng = np.random.default_rng(5)
    n, m = 60_000, 5_000_000
    k = np.clip(rng.zipf(1.8, n), 1, 40)
    rp = np.concatenate([[0], np.cumsum(k)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(m, size=int(c), replace=False)) for c in k]
                        ).astype(np.int32)
    vv = rng.uniform(-1, 1, int(rp[-1]))
    from num_amd import kokkos_spmv as kmod
    assert kmod._xcd_plan(rp, ci, vv, n, m)[0] == 16
    _check(ks, orc, rp, ci, vv, n, m, 2)


def test_xcd_empty_rows_rect_and_tiny(ks, orc):
    rng = np.random.default_rng(6)
    n, m = 3000, 777
    k = rng.integers(0, 5, n)
    k[::7] = 0
    rp = np.concatenate([[0], np.cumsum(k)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(m, size=int(c), replace=False)) for c in k]
                        ).astype(np.int32)
    vv = rng.uniform(-1, 1, int(rp[-1]))
    _check(ks, orc, rp, ci, vv, n, m, 3)
    # a few entries only: most XCD groups empty
    rp2 = np.array([0, 1, 1, 3], np.int32)
    ci2 = np.array([5, 0, 9], np.int32)
    vv2 = np.array([2.0, -1.0, 0.5])
    _check(ks, orc, rp2, ci2, vv2, 3, 10, 4)
