"""Host-side checks of the column-partitioned layout's plan
(kokkos_spmv._xcd_plan / _wants_xcd): the segments cover every entry once,
in block-major / row-ascending order, and the per-segment partial sums added
per row in block order give A x (checked against scipy)."""
import numpy as np
import scipy.sparse as sp

from num_amd import generators as g
from num_amd import kokkos_spmv as ks


def _plan_spmv(rp, ci, vv, n, m, x):
    nb, seg_ptr, cols, vals, bstart, rsp, rsi = ks._xcd_plan(rp, ci, vv, n, m)
    prod = vals * x[cols]
    part = np.array([prod[seg_ptr[s]:seg_ptr[s + 1]].sum() for s in range(len(seg_ptr) - 1)])
    y = np.zeros(n)
    for r in range(n):
        for i in range(rsp[r], rsp[r + 1]):
            y[r] += part[rsi[i]]
    return y, (nb, seg_ptr, cols, bstart, rsp, rsi)


def test_plan_reproduces_spmv_and_orders():
    A = g.zipf_matrix(3000, 2.0, 200, seed=21)
    x = np.random.default_rng(2).uniform(-1, 1, A.ncols)
    y, (nb, seg_ptr, cols, bstart, rsp, rsi) = _plan_spmv(A.rowptr, A.colind, A.values,
                                                          A.nrows, A.ncols, x)
    ref = sp.csr_matrix((A.values, A.colind, A.rowptr), shape=(A.nrows, A.ncols)) @ x
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)
    assert nb % 8 == 0 and bstart[0] == 0 and bstart[-1] == len(seg_ptr) - 1
    assert seg_ptr[-1] == A.nnz and np.all(np.diff(seg_ptr) > 0)
    bs = -(-A.ncols // nb)
    for j in range(nb):  # every segment of block j holds only block-j columns
        lo, hi = seg_ptr[bstart[j]], seg_ptr[bstart[j + 1]]
        assert np.all(cols[lo:hi] // bs == j)
    # each row's segments: as many as the blocks it touches, in block order
    for r in range(0, A.nrows, 97):
        blocks = np.unique(A.colind[A.rowptr[r]:A.rowptr[r + 1]] // bs)
        segs = rsi[rsp[r]:rsp[r + 1]]
        assert len(segs) == len(blocks)
        got = [int(np.searchsorted(bstart, s, side="right") - 1) for s in segs]
        assert got == list(blocks)


def test_plan_empty_rows_rect():
    rng = np.random.default_rng(3)
    n, m = 500, 1300
    k = rng.integers(0, 6, n)
    k[::5] = 0
    rp = np.concatenate([[0], np.cumsum(k)]).astype(np.int32)
    ci = np.concatenate([np.sort(rng.choice(m, size=int(c), replace=False)) for c in k]
                        ).astype(np.int32)
    vv = rng.uniform(-1, 1, int(rp[-1]))
    x = rng.uniform(-1, 1, m)
    y, (_, _, _, _, rsp, _) = _plan_spmv(rp, ci, vv, n, m, x)
    ref = sp.csr_matrix((vv, ci, rp), shape=(n, m)) @ x
    assert np.allclose(y, ref, rtol=0, atol=1e-14)
    assert np.all(np.diff(rsp)[::5] == 0)


def test_wants_xcd():
    # random columns over 4M: far gathers, x of 32 MB -> the partitioned layout
    n = 4_000_000
    rng = np.random.default_rng(4)
    rp = np.arange(0, 2 * n + 1, 2, dtype=np.int32)
    ci = rng.integers(0, n, 2 * n).astype(np.int32)
    assert ks._wants_xcd(rp, ci, n, n)
    # a banded pattern of the same size: near gathers -> rows
    rows = np.repeat(np.arange(n), 2)
    ci2 = np.clip(rows + rng.integers(-1000, 1000, 2 * n), 0, n - 1).astype(np.int32)
    assert not ks._wants_xcd(rp, ci2, n, n)
    # small x -> rows
    assert not ks._wants_xcd(rp[:1001], ci[:2000] % 1000, 1000, 1000)
