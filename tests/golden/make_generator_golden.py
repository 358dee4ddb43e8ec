"""Write tests/golden/generators.npz: the reference's own matrix generators'
outputs at small sizes, so that num_amd.generators' restatements of them are
pinned exactly (same seed -> same matrix).

The families lattice_gauge, random_sparse and power_law
(python/matrix_generators.py:85-221) depend on the RNG call sequence and on
set iteration order; no closed form or library routine fixes their output,
so this script imports the reference module (it runs only where the
reference checkout is present, never on the GPU box) and stores, per case,
the CSR arrays it returns (indptr, indices, data) and whether they are in
scipy's canonical form.  Only those arrays are committed.

Run:  python tests/golden/make_generator_golden.py
"""
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/python/matrix_generators.py"

# (key, function name, args, kwargs)
CASES = [
    ("lattice_gauge_L2", "make_lattice_gauge", (2,), {}),
    ("lattice_gauge_L3", "make_lattice_gauge", (3,), {}),
    ("lattice_gauge_L12", "make_lattice_gauge", (12,), {}),
    ("lattice_gauge_L9_m05", "make_lattice_gauge", (9,), {"mass": 0.5}),
    ("random_sparse_N1", "make_random_sparse", (1,), {}),
    ("random_sparse_N5", "make_random_sparse", (5,), {}),
    ("random_sparse_N200", "make_random_sparse", (200,), {}),
    ("random_sparse_N60_k4_nonsym", "make_random_sparse", (60,),
     {"nnz_per_row": 4, "symmetric": False}),
    ("power_law_N1", "make_power_law", (1,), {}),
    ("power_law_N4", "make_power_law", (4,), {}),
    ("power_law_N200", "make_power_law", (200,), {}),
    ("power_law_N80_m2", "make_power_law", (80,), {"m": 2}),
]
# the registry factory (get_matrix) at one size per family
REGISTRY_N = 256
REGISTRY = ["laplacian_2d", "laplacian_3d", "lattice_gauge", "random_sparse", "power_law"]


def main():
    spec = importlib.util.spec_from_file_location("ref_matrix_generators", REF)
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    out = {}

    def put(key, A):
        A = A.tocsr()
        out[key + "__indptr"] = A.indptr.astype(np.int64)
        out[key + "__indices"] = A.indices.astype(np.int64)
        out[key + "__data"] = A.data.astype(np.float64)
        out[key + "__shape"] = np.array(A.shape, np.int64)
        out[key + "__canonical"] = np.array(bool(A.has_canonical_format))
    for key, fn, args, kw in CASES:
        put(key, getattr(mg, fn)(*args, **kw))
    for name in REGISTRY:
        put(f"get_matrix_{name}_N{REGISTRY_N}", mg.get_matrix(name, REGISTRY_N))
    path = os.path.join(HERE, "generators.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {len(CASES) + len(REGISTRY)} matrices, "
          f"{os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
