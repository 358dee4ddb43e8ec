"""GPU legs of the benchmark harness (the reference's benchmark_spmv.py /
benchmark_cg.py command lines) on small matrices of every family."""
import csv

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("matrix", ["laplacian_3d", "lattice_gauge", "random_sparse",
                                    "power_law"])
def test_harness_spmv(ks, tmp_path, monkeypatch, matrix):
    from num_amd import harness
    monkeypatch.chdir(tmp_path)
    assert harness.main(["spmv", "--N", "3000", "--matrix", matrix, "--repeat", "5",
                         "--batch", "4", "--backend", "num_amd", "num_amd_custom"]) == 0
    (f,) = list((tmp_path / "benchmark" / "results").glob("spmv_*.csv"))
    rows = {r["backend"]: r for r in csv.DictReader(open(f))}
    assert set(rows) == {"num_amd_1gpu", "num_amd_custom_1gpu", "num_amd_batch_k4"}
    assert all(float(r["elapsed_us"]) > 0 for r in rows.values())
    assert rows["num_amd_batch_k4"]["k"] == "4"


@pytest.mark.parametrize("matrix", ["laplacian_3d", "lattice_gauge", "random_sparse"])
def test_harness_cg(ks, tmp_path, monkeypatch, matrix):
    from num_amd import harness
    monkeypatch.chdir(tmp_path)
    assert harness.main(["cg", "--N", "4000", "--matrix", matrix, "--mixed"]) == 0
    (f,) = list((tmp_path / "benchmark" / "results").glob("cg_*.csv"))
    rows = {r["backend"]: r for r in csv.DictReader(open(f))}
    assert {"scipy_cg", "num_amd_cg_1gpu", "num_amd_cg_mixed_1gpu", "spmv_only_1gpu"} <= set(rows)
    for tag in ("scipy_cg", "num_amd_cg_1gpu", "num_amd_cg_mixed_1gpu"):
        assert rows[tag]["converged"] == "True" and float(rows[tag]["residual"]) < 1e-8
    # same CG sequence as scipy's unpreconditioned CG: iteration counts close
    assert abs(int(rows["num_amd_cg_1gpu"]["iters"]) - int(rows["scipy_cg"]["iters"])) <= \
        max(2, 0.05 * int(rows["scipy_cg"]["iters"]))


def test_harness_verify(ks, capsys):
    # the check groups of python/verify_accuracy.py (single SpMV, batch SpMV,
    # local CG against SciPy, the distributed SpMV over 3 shares run in turn)
    from num_amd import harness
    assert harness.main(["verify", "--ranks", "3"]) == 0
    out = capsys.readouterr().out
    assert "ALL 4 test groups PASSED" in out and "FAIL" not in out
    assert out.count("PASS") == 4 + 2 + 2 + 2 + 1
