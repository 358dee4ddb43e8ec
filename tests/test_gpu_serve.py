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


@pytest.mark.parametrize("dims", [(9,), (13, 7), (6, 5, 4), (1, 1), (1,), (200, 150),
                                  (33, 17, 9)])
@pytest.mark.parametrize("mixed", [False, True])
def test_engine_stencil_assembly(ks, dims, mixed):
    """The engine's one-kernel request assembly (closed-form rowptr, columns,
    values, FP32 copy, b = hash_uniform, x = 0) equals the host generators."""
    from num_amd.batched import CGEngine, stencil_layout
    eng = CGEngine(max_systems=2)
    shift, seed = 0.25, 1234 + len(dims)
    s = eng.add_stencil(dims, shift, seed, mixed=mixed)
    torch.cuda.synchronize()
    buf = eng._sys[s][0].cpu().numpy()
    lay = stencil_layout(dims, mixed)
    n, nnz = lay["n"], lay["nnz"]
    A = g.stencil(dims, shift)
    assert (A.nrows, A.nnz) == (n, nnz)

    def arr(name, dt, count):
        o = lay[name]
        return buf[o:o + count * np.dtype(dt).itemsize].view(dt)
    assert np.array_equal(arr("rowptr", np.int32, n + 1), A.rowptr)
    assert np.array_equal(arr("colind", np.int32, nnz), A.colind)
    assert np.array_equal(arr("values", np.float64, nnz), A.values)
    if mixed:
        assert np.array_equal(arr("values32", np.float32, nnz), A.values.astype(np.float32))
    # each row's pattern, decoded from its id byte and the table, lists
    # exactly the row's column - row offsets, zero past the row
    pid = arr("rpid", np.uint8, n).astype(np.int64)
    tab = arr("rpat", np.int16, 64 * 8).reshape(64, 8).astype(np.int64)
    lens = np.diff(A.rowptr)
    assert lens.max() <= 8
    rows = np.repeat(np.arange(n), lens)
    slot = np.arange(nnz) - np.repeat(A.rowptr[:-1], lens)
    assert np.array_equal(tab[pid[rows], slot], A.colind - rows)
    mask = np.arange(8)[None, :] >= lens[:, None]
    assert not tab[pid][mask].any()
    # and its length, a 4-bit field of the length words after the table
    lw = arr("rpat", np.uint64, 64 * 2 + 4)[128:]
    nib = np.array([(int(lw[f >> 4]) >> (4 * (f & 15))) & 15 for f in range(64)])
    assert np.array_equal(nib[pid], lens)
    assert np.array_equal(arr("b", np.float64, n), g.hash_uniform(n, seed))
    assert np.array_equal(arr("x", np.float64, n), np.zeros(n))
    assert torch.equal(eng.solution(s).cpu(), torch.zeros(n, dtype=torch.float64))
    assert CGEngine.stencil_bytes(dims, mixed) >= lay["work"] + CGEngine.work_bytes(n, mixed)


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
        b = g.hash_uniform(A.nrows, r.seed)
        true = np.linalg.norm(b - A.to_scipy() @ r.solution.cpu().numpy()) / np.linalg.norm(b)
        assert true < (1e-10 if mixed else 2e-10)
        if not mixed:
            # the same system through the one-at-a-time FP64 solver: the same
            # CG sequence, iterations within the +-2% bar
            dA = ks.upload_stencil(r.dims, r.shift)
            x1 = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
            one = ks.cg_solve_local(dA, torch.from_numpy(b).cuda(), x1, tol=r.tol,
                                    max_iter=20000)
            assert abs(one["iters"] - r.result["iters"]) <= max(1, 0.02 * one["iters"])


def test_budget_limits_admission(ks):
    from num_amd.serve import CGServer, request_bytes
    reqs = _trace(6, seed=7, nmax=20_000)
    biggest = max(request_bytes(r.dims, True) for r in reqs)
    budget = biggest * 1.01
    srv = CGServer(budget_bytes=budget, max_systems=16, mixed=True, iters_per_step=4)
    summ = srv.run_open_loop(reqs)
    assert summ["completed"] == len(reqs) and summ["converged"] == len(reqs)
    assert 0 < srv.peak_in_use <= budget
    assert srv.in_use == 0.0


def test_rejection_and_timeout(ks):
    from num_amd.serve import CGServer, Request, request_bytes
    small = Request(0, 0.0, (20, 20), 0.5, 1)
    huge = Request(1, 0.0, (400, 400), 0.5, 2)
    srv = CGServer(budget_bytes=request_bytes(small.dims, True) * 2, max_systems=4)
    summ = srv.run_open_loop([small, huge])
    assert summ["completed"] == 2 and summ["rejected"] == 1 and summ["converged"] == 1
    assert srv.in_use == 0.0
    # timeout 0: every request still running at its first collection is
    # cancelled, its slot and memory returned
    reqs = _trace(8, seed=3, nmax=200_000)
    srv = CGServer(budget_bytes=50e9, max_systems=8, timeout_s=0.0)
    summ = srv.run_open_loop(reqs)
    assert summ["completed"] == len(reqs)
    assert summ["timed_out"] + summ["converged"] == len(reqs) and summ["timed_out"] > 0
    assert srv.in_use == 0.0 and srv.engine.info()["live"] == 0


@pytest.mark.parametrize("mixed", [False, True])
def test_pipelined_launch_equals_step(ks, mixed):
    """launch/poll (steps queued while the host works) and the synchronous
    step give bit-identical solutions and iteration counts."""
    from num_amd.batched import CGEngine
    grids = [((30, 20), 0.01), ((12, 11, 10), 0.3), ((500,), 0.001), ((7, 64), 1.0),
             ((1, 1), 0.5), ((2,), 0.1), ((3, 1), 0.2)]  # tiny: fewer than 4 entries
    out = []
    for pipelined in (False, True):
        eng = CGEngine(max_systems=8)
        slots = [eng.add_stencil(d, sh, 10 + i, tol=1e-10, max_iter=5000, mixed=mixed)
                 for i, (d, sh) in enumerate(grids)]
        if pipelined:
            while True:
                eng.launch(5)
                live, inflight = eng.poll()
                if live == 0:
                    break
            eng.poll(wait=True)
        else:
            eng.run(5)
        out.append([(eng.result(s)["iters"], eng.solution(s).clone()) for s in slots])
    for (i0, x0), (i1, x1) in zip(*out):
        assert i0 == i1 and torch.equal(x0, x1)
    # every solution, tiny grids included (1-D and single-row systems read
    # the neighbour-pattern table's short patterns), against the stencil's
    # closed form, independent of the library
    from _helpers import stencil_residual
    for i, ((d, sh), (_, x)) in enumerate(zip(grids, out[0])):
        assert stencil_residual(x, d, sh, 10 + i) < 1e-10, d


@pytest.mark.parametrize("n,seed", [(0, 1), (1, 7), (1000, 12345), (300_001, 2**40 + 3)])
def test_device_hash_uniform_equals_host(ks, n, seed):
    # the device draws of num_hash_uniform are the host generator's, bit for bit
    got = ks.hash_uniform(n, seed).cpu().numpy()
    assert np.array_equal(got, g.hash_uniform(n, seed))
