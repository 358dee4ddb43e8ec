"""The closed-form stencil residual used by the GPU tests, pinned on the CPU
against the assembled CSR stencil (scipy) and the oracle's SpMV."""
import numpy as np
import pytest
import torch

from num_amd import generators as g
from _helpers import stencil_apply, stencil_residual


@pytest.mark.parametrize("dims,shift", [((7,), 0.5), ((5, 9), 0.01), ((4, 3, 6), 0.3),
                                        ((1, 5), 1.0), ((3, 1, 2), 0.2)])
def test_stencil_apply_equals_csr(orc, dims, shift):
    A = g.stencil(dims, shift)
    x = np.random.default_rng(3).uniform(-1, 1, A.nrows)
    got = stencil_apply(torch.from_numpy(x), dims, shift).numpy()
    ref = orc.spmv(A.rowptr, A.colind, A.values, x)
    assert np.allclose(got, ref, rtol=0, atol=1e-14)
    assert np.allclose(got, A.to_scipy() @ x, rtol=0, atol=1e-14)


def test_stencil_residual_of_exact_solve():
    import scipy.sparse.linalg as sla
    dims, shift, seed = (6, 5), 0.2, 77
    A = g.stencil(dims, shift)
    b = g.hash_uniform(A.nrows, seed)
    x = sla.spsolve(A.to_scipy().tocsc(), b)
    assert stencil_residual(torch.from_numpy(x), dims, shift, seed) < 1e-14
    assert stencil_residual(torch.zeros(A.nrows, dtype=torch.float64), dims, shift, seed) == 1.0


def test_shuffled_stencil_rows_is_the_stencil():
    # the reordered matrix with duplicates summed is the 2-D stencil itself,
    # and it really holds unsorted rows and duplicate columns
    from _helpers import shuffled_stencil_rows
    import scipy.sparse as sp
    rp, ci, va, n = shuffled_stencil_rows(30, 20, seed=4)
    M = sp.csr_matrix((va.copy(), ci.copy(), rp.copy()), shape=(n, n))  # scipy sorts in place
    M.sum_duplicates()
    S = g.stencil((30, 20), 0.5).to_scipy()
    assert abs(M - S).max() == 0.0
    rows = [ci[rp[i]:rp[i + 1]] for i in range(n)]
    assert sum(np.any(np.diff(c) < 0) for c in rows) > n // 2
    assert sum(len(set(c)) < len(c) for c in rows) >= n // 4
