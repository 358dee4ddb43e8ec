"""num_amd.generators' restatements of the reference's matrix families
(python/matrix_generators.py:16-258) against the reference's own outputs,
frozen in tests/golden/generators.npz by tests/golden/make_generator_golden.py:
same seed, same matrix, bit for bit (both in scipy's canonical CSR form:
duplicates summed, columns sorted)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from num_amd import generators as g

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "generators.npz"))

OURS = {
    "lattice_gauge_L2": lambda: g.lattice_gauge(2),
    "lattice_gauge_L3": lambda: g.lattice_gauge(3),
    "lattice_gauge_L12": lambda: g.lattice_gauge(12),
    "lattice_gauge_L9_m05": lambda: g.lattice_gauge(9, mass=0.5),
    "random_sparse_N1": lambda: g.random_sparse(1),
    "random_sparse_N5": lambda: g.random_sparse(5),
    "random_sparse_N200": lambda: g.random_sparse(200),
    "random_sparse_N60_k4_nonsym": lambda: g.random_sparse(60, nnz_per_row=4, symmetric=False),
    "power_law_N1": lambda: g.power_law(1),
    "power_law_N4": lambda: g.power_law(4),
    "power_law_N200": lambda: g.power_law(200),
    "power_law_N80_m2": lambda: g.power_law(80, m=2),
}
for _name in ("laplacian_2d", "laplacian_3d", "lattice_gauge", "random_sparse", "power_law"):
    OURS[f"get_matrix_{_name}_N256"] = (lambda nm: lambda: g.get_matrix(nm, 256))(_name)


def _golden(key):
    shape = tuple(int(v) for v in GOLD[key + "__shape"])
    A = sp.csr_matrix((GOLD[key + "__data"], GOLD[key + "__indices"], GOLD[key + "__indptr"]),
                      shape=shape)
    A.sum_duplicates()
    A.sort_indices()
    return A


def test_fixture_covers_every_case():
    keys = {k.split("__")[0] for k in GOLD.files}
    assert keys == set(OURS)


@pytest.mark.parametrize("key", sorted(OURS))
def test_generator_equals_reference(key):
    ref = _golden(key)
    A = OURS[key]()
    assert (A.nrows, A.ncols) == ref.shape
    assert np.array_equal(A.rowptr, ref.indptr)
    assert np.array_equal(A.colind, ref.indices)
    assert np.array_equal(A.values, ref.data), np.max(np.abs(A.values - ref.data))
