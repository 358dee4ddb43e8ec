"""GPU tests of the batched CG engine: every system of a batch must match the
single-system solver / oracle, a result must be bit-identical alone and in
any mix, systems admitted into a running batch must behave the same, and the
edge cases of tests/test_cg.cpp (zero RHS, max_iter, breakdown) must hold per
system without disturbing the others."""
import numpy as np
import pytest
import torch

from num_amd import generators as g

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _systems(count, seed=4, nmax=5000):
    return g.batched_systems(count=count, nmin=16, nmax=nmax, seed=seed)


def _upload(ks, dims, shift):
    A = g.stencil(dims, shift)
    return A, ks.upload_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols)


def test_batch_matches_single_and_oracle(ks, orc):
    from num_amd.batched import solve_batch
    sysl = _systems(40)
    mats, entries = [], []
    for dims, shift, b in sysl:
        A, dA = _upload(ks, dims, shift)
        mats.append((A, b))
        entries.append((dA, _dev(b), _dev(np.zeros(A.nrows))))
    res = solve_batch(entries, tol=1e-10, max_iter=5000)
    for (A, b), (dA, db, x), r in zip(mats, entries, res):
        assert r["converged"]
        xo, ro = orc.cg_solve_local(A.rowptr, A.colind, A.values, b, tol=1e-10, max_iter=5000)
        assert abs(r["iters"] - ro["iters"]) <= max(1, int(0.02 * ro["iters"]))
        xs = x.cpu().numpy()
        assert np.linalg.norm(xs - xo) / np.linalg.norm(xo) < 1e-8
        true = np.linalg.norm(b - A.to_scipy() @ xs) / np.linalg.norm(b)
        assert true < 2e-10


def test_alone_equals_in_mix_bitwise(ks):
    from num_amd.batched import CGEngine
    sysl = _systems(12, seed=9)
    ups = [_upload(ks, d, s) for d, s, _ in sysl]
    # all together
    eng = CGEngine(max_systems=16)
    xs_mix = []
    slots = []
    for (A, dA), (_, _, b) in zip(ups, sysl):
        x = _dev(np.zeros(A.nrows))
        xs_mix.append(x)
        slots.append(eng.add(dA, _dev(b), x, tol=1e-10, max_iter=5000))
    eng.run(7)
    mix = [eng.result(s) for s in slots]
    # each alone, in a fresh engine
    for k, ((A, dA), (_, _, b)) in enumerate(zip(ups, sysl)):
        e1 = CGEngine(max_systems=1)
        x = _dev(np.zeros(A.nrows))
        s = e1.add(dA, _dev(b), x, tol=1e-10, max_iter=5000)
        e1.run(5)
        r = e1.result(s)
        assert r["iters"] == mix[k]["iters"] and r["residual"] == mix[k]["residual"]
        assert torch.equal(x, xs_mix[k])


def test_admission_into_running_batch(ks):
    from num_amd.batched import CGEngine
    sysl = _systems(10, seed=5)
    ups = [_upload(ks, d, s) for d, s, _ in sysl]
    eng = CGEngine(max_systems=16)
    xs, slots = [], []
    for i, ((A, dA), (_, _, b)) in enumerate(zip(ups, sysl)):
        x = _dev(np.zeros(A.nrows))
        xs.append(x)
        slots.append(eng.add(dA, _dev(b), x, tol=1e-10, max_iter=5000))
        eng.step(3)  # the earlier systems are mid-solve when the next arrives
    eng.run(8)
    for i, ((A, dA), (_, _, b)) in enumerate(zip(ups, sysl)):
        e1 = CGEngine(max_systems=1)
        x = _dev(np.zeros(A.nrows))
        s = e1.add(dA, _dev(b), x, tol=1e-10, max_iter=5000)
        e1.run(16)
        assert torch.equal(x, xs[i])
        assert e1.result(s)["iters"] == eng.result(slots[i])["iters"]


def test_mixed_batch_true_residual(ks):
    from num_amd.batched import solve_batch
    sysl = _systems(20, seed=6)
    mats, entries = [], []
    for dims, shift, b in sysl:
        A, dA = _upload(ks, dims, shift)
        mats.append((A, b))
        entries.append((dA, _dev(b), _dev(np.zeros(A.nrows))))
    res = solve_batch(entries, tol=1e-10, max_iter=20000, mixed=True)
    for (A, b), (dA, db, x), r in zip(mats, entries, res):
        assert r["converged"] and r["residual"] < 1e-10 and r["refinements"] >= 1
        xs = x.cpu().numpy()
        true = np.linalg.norm(b - A.to_scipy() @ xs) / np.linalg.norm(b)
        assert true < 1e-10


def test_edge_cases_per_system(ks):
    from num_amd.batched import CGEngine
    eng = CGEngine(max_systems=8)
    A, dA = _upload(ks, (20,), 0.0)
    # zero RHS -> 0 iterations, residual 0 (tests/test_cg.cpp [5])
    x0 = _dev(np.zeros(20))
    s0 = eng.add(dA, _dev(np.zeros(20)), x0, tol=1e-10, max_iter=100)
    # max_iter = 5 (tests/test_cg.cpp [7])
    A2, dA2 = _upload(ks, (50,), 0.0)
    x1 = _dev(np.zeros(50))
    s1 = eng.add(dA2, _dev(np.ones(50)), x1, tol=1e-14, max_iter=5)
    # breakdown: singular 2x2 with p'Ap = 0
    rp = np.array([0, 2, 4], np.int32)
    ci = np.array([0, 1, 0, 1], np.int32)
    v = np.array([1.0, -1.0, -1.0, 1.0])
    dS = ks.upload_csr(rp, ci, v, 2, 2)
    s2 = eng.add(dS, _dev(np.ones(2)), _dev(np.zeros(2)), tol=1e-10, max_iter=10)
    # a NaN in b: that system stops unconverged, nothing else is touched
    A4, dA4 = _upload(ks, (40,), 0.5)
    b4 = np.ones(40)
    b4[17] = np.nan
    s4 = eng.add(dA4, _dev(b4), _dev(np.zeros(40)), tol=1e-10, max_iter=50)
    # a healthy system beside them
    A3, dA3 = _upload(ks, (30, 20), 0.1)
    b3 = np.random.default_rng(1).uniform(-1, 1, A3.nrows)
    x3 = _dev(np.zeros(A3.nrows))
    s3 = eng.add(dA3, _dev(b3), x3, tol=1e-10, max_iter=2000)
    eng.run(4)
    r0, r1, r2, r3, r4 = (eng.result(s) for s in (s0, s1, s2, s3, s4))
    assert r0["iters"] == 0 and r0["converged"] and r0["residual"] == 0.0
    assert r1["iters"] == 5 and not r1["converged"]
    assert r2["breakdown"] and not r2["converged"]
    assert not r4["converged"] and r4["iters"] <= 50
    assert r3["converged"]
    true = np.linalg.norm(b3 - A3.to_scipy() @ x3.cpu().numpy()) / np.linalg.norm(b3)
    assert true < 2e-10


@pytest.mark.parametrize("mixed", [False, True])
def test_graph_replay_equals_eager(ks, mixed):
    """Steps replayed from a captured hipGraph give the eager engine's
    iterates after EVERY step (FP64 x is the live iterate), with eager torch
    work and synchronisations between replays, admissions into the running
    batch (re-capture) and systems finishing (re-capture)."""
    from num_amd.batched import CGEngine
    grids = [((40, 30), 0.01), ((9, 8, 7), 0.2), ((300,), 0.05), ((64, 16), 0.5)]
    late = [((25, 25), 0.02), ((5, 6, 7), 0.3)]
    engines = [CGEngine(max_systems=8), CGEngine(max_systems=8, graph=True)]
    slots = [[e.add_stencil(d, sh, 20 + i, max_iter=4000, mixed=mixed)
              for i, (d, sh) in enumerate(grids)] for e in engines]
    for step in range(400):
        live = [e.step(6) for e in engines]
        assert live[0] == live[1]
        junk = torch.randn(1 << 16, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        for s0, s1 in zip(slots[0], slots[1]):
            assert torch.equal(engines[0].solution(s0), engines[1].solution(s1)), (step, s0)
        if step == 3:
            for k, e in enumerate(engines):
                slots[k] += [e.add_stencil(d, sh, 90 + i, max_iter=4000, mixed=mixed)
                             for i, (d, sh) in enumerate(late)]
        if live[0] == 0 and step > 3:
            break
    assert live == [0, 0]
    for s0, s1 in zip(slots[0], slots[1]):
        r0, r1 = engines[0].result(s0), engines[1].result(s1)
        assert r0 == r1 and r0["converged"]
    del junk


@pytest.mark.parametrize("mixed", [False, True])
def test_large_system_wide_chunks(ks, mixed):
    """A system of 819,200 rows gets update chunks wider than 2048 rows (at
    most 256 chunks per system): its result is converged, meets the FP64
    true residual, and is bit-identical alone and next to a small system."""
    from num_amd.batched import CGEngine
    big, small = ((1024, 800), 0.05, 11), ((40, 30), 0.3, 12)

    def run(specs):
        eng = CGEngine(max_systems=4)
        slots = [eng.add_stencil(d, sh, sd, tol=1e-10, max_iter=5000, mixed=mixed)
                 for d, sh, sd in specs]
        eng.run(8)
        return [(eng.result(s), eng.solution(s).clone()) for s in slots]

    (r_alone, x_alone), = run([big])
    (r_small, _), (r_mix, x_mix) = run([small, big])
    assert r_alone["converged"] and r_mix["converged"] and r_small["converged"]
    assert r_alone["iters"] == r_mix["iters"]
    assert torch.equal(x_alone, x_mix)
    A = g.stencil(big[0], big[1])
    b = g.hash_uniform(A.nrows, big[2])
    true = np.linalg.norm(b - A.to_scipy() @ x_alone.cpu().numpy()) / np.linalg.norm(b)
    assert true < (1e-10 if mixed else 2e-10)


def test_schedule_held_while_in_engine(ks):
    """A live engine system expands its launch lists from its matrix's tiles:
    rescheduling that matrix is refused until the system is removed, and the
    solve is unaffected by the refused attempt."""
    from num_amd.batched import CGEngine
    A, dA = _upload(ks, (40, 30), 0.05)
    b = np.random.default_rng(2).uniform(-1, 1, A.nrows)
    x_alone = _dev(np.zeros(A.nrows))
    ref = CGEngine(max_systems=1)
    s0 = ref.add(dA, _dev(b), x_alone, tol=1e-10, max_iter=2000)
    ref.run(4)
    r_alone = ref.result(s0)
    ref.remove(s0)
    eng = CGEngine(max_systems=2)
    x = _dev(np.zeros(A.nrows))
    s = eng.add(dA, _dev(b), x, tol=1e-10, max_iter=2000)
    eng.step(3)  # mid-solve
    before = dA.schedule_info()
    with pytest.raises(RuntimeError, match="engine"):
        dA.set_schedule(rows_per_block=64)
    with pytest.raises(RuntimeError, match="engine"):
        import ctypes as C
        from num_amd._lib import call
        call("num_csr_destroy", dA.handle)
    assert dA.schedule_info() == before
    eng.run(4)
    r = eng.result(s)
    assert r["converged"] and r["iters"] == r_alone["iters"]
    assert torch.equal(x, x_alone)
    eng.remove(s)
    dA.set_schedule(rows_per_block=64)  # released
    assert dA.schedule_info()["rows_per_block"] == 64


def test_engine_autotune_keeps_results(ks):
    """The engine tuner picks a schedule among its candidates, remembers it per
    workload class, and no candidate changes a result: the tuned batch equals
    the default-schedule batch bit for bit."""
    from num_amd import autotune
    from num_amd.batched import CGEngine
    specs = g.batched_specs(count=48, nmin=64, nmax=4000, seed=9)
    for mixed in (False, True):
        t = autotune.tune_engine(specs, mixed=mixed, sample=16)
        sch = t["schedule"]
        assert sch["ka_chunk"] in autotune.ENGINE_KA_CHUNKS
        assert sch["iters_per_step"] in autotune.ENGINE_ITERS
        assert t["timed"] > 1 and not t["cached"]
        assert autotune.tune_engine(specs, mixed=mixed, sample=16)["cached"]
        sols = []
        for ka, it, nt in ((0, 16, True), (sch["ka_chunk"], sch["iters_per_step"], sch["nt32"]),
                           (1, 8, False)):
            eng = CGEngine(max_systems=len(specs), nt32=nt)
            eng.set_schedule(ka)
            slots = [eng.add_stencil(d, sh, sd, tol=1e-10, max_iter=5000, mixed=mixed)
                     for d, sh, sd in specs]
            eng.run(it)
            assert all(eng.result(s)["converged"] for s in slots)
            sols.append([eng.solution(s).clone() for s in slots])
        for other in sols[1:]:
            assert all(torch.equal(a, b) for a, b in zip(sols[0], other))


def test_engine_add_tune_bounded_depth(ks):
    """add(..., tune=True) lets the SpMV tuner pick the matrix schedule within
    the engine's 7 entries per thread; the solve matches the untuned one."""
    from num_amd.batched import CGEngine
    A, dA = _upload(ks, (70, 50), 0.02)
    b = np.random.default_rng(5).uniform(-1, 1, A.nrows)
    xs = []
    for tune in (False, True):
        eng = CGEngine(max_systems=1)
        x = _dev(np.zeros(A.nrows))
        s = eng.add(dA, _dev(b), x, tol=1e-10, max_iter=3000, tune=tune)
        assert dA.schedule_info()["nnz_per_block"] <= 7 * 256
        eng.run(8)
        assert eng.result(s)["converged"]
        eng.remove(s)
        xs.append(x.cpu().numpy())
    assert np.linalg.norm(xs[0] - xs[1]) <= 1e-8 * np.linalg.norm(xs[0])
