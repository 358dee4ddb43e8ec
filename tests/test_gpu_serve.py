"""GPU tests of the serving loop: device-assembled stencils equal the host
generator, a request's result is bit-identical alone and inside the
continuous batch, and admission respects the memory budget."""
import numpy as np
import pytest
import torch

from num_amd import generators as g

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,shift", [((9,), 0.5), ((13, 7), 0.01), ((6, 5, 4), 0.2),
                                        ((1, 1), 1.0), ((200, 150), 0.001)])
def test_upload_stencil_equals_generator(ks, dims, shift):
    A = g.stencil(dims, shift)
    d = ks.upload_stencil(dims, shift)
    assert (d.nrows, d.nnz) == (A.nrows, A.nnz)
    assert np.array_equal(d._rowptr.cpu().numpy(), A.rowptr)
    assert np.array_equal(d._colind.cpu().numpy(), A.colind)
    assert np.array_equal(d._values.cpu().numpy(), A.values)


def _trace(count, seed, nmax=60_000):
    tr = g.served_requests(rate=2000.0, duration=10.0, seed=seed, nmin=1000, nmax=nmax,
                           max_requests=count)
    from num_amd.serve import requests_from_trace
    return requests_from_trace(tr)


@pytest.mark.parametrize("mixed", [False, True])
def test_served_results_equal_alone(ks, mixed):
    from num_amd.serve import CGServer
    reqs = _trace(24, seed=5)
    srv = CGServer(budget_bytes=50e9, max_systems=64, mixed=mixed, iters_per_step=4)
    summ = srv.run_open_loop(reqs, keep_solutions=True)
    assert summ["completed"] == len(reqs) and summ["converged"] == len(reqs)
    assert summ["p99_ms"] >= summ["p50_ms"] > 0
    by_id = {r.rid: r for r in srv.done}
    # several requests must really have overlapped on the device
    overlap = sum(1 for r in srv.done for q in srv.done
                  if r.rid < q.rid and r.t_admit < q.t_done and q.t_admit < r.t_done)
    assert overlap > 0
    for rid in (0, 5, 11, 23):
        r = by_id[rid]
        alone = CGServer(budget_bytes=50e9, max_systems=1, mixed=mixed, iters_per_step=3)
        from num_amd.serve import Request
        one = Request(0, 0.0, r.dims, r.shift, r.seed, r.tol)
        alone.run_open_loop([one], keep_solutions=True)
        a = alone.done[0]
        assert a.result["iters"] == r.result["iters"]
        assert torch.equal(a.solution, r.solution)
        # true residual in FP64
        A = g.stencil(r.dims, r.shift)
        gen = torch.Generator(device="cuda")
        gen.manual_seed(int(r.seed))
        b = (torch.rand(A.nrows, dtype=torch.float64, device="cuda", generator=gen) * 2 - 1)
        b = b.cpu().numpy()
        true = np.linalg.norm(b - A.to_scipy() @ r.solution.cpu().numpy()) / np.linalg.norm(b)
        assert true < (1e-10 if mixed else 2e-10)


def test_budget_limits_admission(ks):
    from num_amd.serve import CGServer, request_bytes
    reqs = _trace(6, seed=7, nmax=20_000)
    biggest = max(request_bytes(int(np.prod(r.dims)), int(np.prod(r.dims)) * 7, True)
                  for r in reqs)
    budget = biggest * 1.01
    srv = CGServer(budget_bytes=budget, max_systems=16, mixed=True, iters_per_step=4)
    summ = srv.run_open_loop(reqs)
    assert summ["completed"] == len(reqs) and summ["converged"] == len(reqs)
    assert 0 < srv.peak_in_use <= budget
    assert srv.in_use == 0.0
