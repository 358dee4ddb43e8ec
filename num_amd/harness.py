"""Benchmark harness with the reference's command lines and CSV schemas.

    python -m num_amd.harness spmv [--N 100000] [--matrix laplacian_3d]
                                   [--batch K] [--repeat 100]
                                   [--backend scipy num_amd num_amd_custom] [--no-csv]
    python -m num_amd.harness cg   [--N 100000] [--matrix laplacian_3d]
                                   [--tol 1e-8] [--maxiter 2000] [--no-csv] [--mixed]
    python -m num_amd.harness verify [--ranks 4]

They mirror python/benchmark_spmv.py:294-377 and python/benchmark_cg.py:
278-363: the same matrix families (generators.get_matrix), the same x
(standard normal, seed 42) and b (ones), the same metrics (elapsed_us,
bw_GB_s = (12 nnz + 8 k N + 8 N) / t, gflops = 2 nnz k / t; elapsed_s,
iters, residual, converged) and CSV files benchmark/results/{spmv,cg}_<ts>.csv
with the same columns, so the reference's plot_results.py reads them.
GPU times are the median of HIP-event pairs on the stream the kernels run on
(20 warm-ups, as python/benchmark_spmv.py:89-102).  Backend tags:
scipy_cpu / scipy_cg (CPU), num_amd_1gpu and num_amd_custom_1gpu (spmv with
kokkoskernels True / False: the same kernel), num_amd_batch_k<K>,
num_amd_cg_1gpu, num_amd_cg_mixed_1gpu, spmv_only_1gpu.

``verify`` runs the four check groups of python/verify_accuracy.py:67-216
against SciPy with the same matrices, sizes, seeds and bars: single SpMV
(N = 50k, four families, < 1e-10), batch SpMV (N = 20k, k = 64, < 1e-10),
local CG (N = 10k, b = 1, tol 1e-10, 2000 iterations, solution within 1e-6
of SciPy's CG at rtol 1e-12) and the distributed SpMV (N = 40k, < 1e-10).
The reference runs the last group under MPI with two or more ranks; on one
GPU every rank's share runs here in turn -- its halo packed by its
neighbours' shares and scattered into its ghost block -- and the gathered
rows are checked.  Exit status 0 when every group passes.
"""
from __future__ import annotations

import argparse
import csv
import sys
import time
from datetime import datetime
from pathlib import Path

import numpy as np

RESULTS_DIR = Path("benchmark") / "results"


def spmv_metrics(elapsed_s: float, N: int, nnz: int, k: int = 1) -> dict:
    """python/benchmark_spmv.py:107-113 (its bytes model omits rowptr)."""
    mem_bytes = nnz * 12 + k * N * 8 + N * 8
    return {"elapsed_us": elapsed_s * 1e6, "bw_GB_s": mem_bytes / elapsed_s / 1e9,
            "gflops": 2 * nnz * k / elapsed_s / 1e9}


def _median_events(fn, warmup: int = 20, repeat: int = 100) -> float:
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    times = []
    for _ in range(repeat):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        fn()
        e1.record(st)
        e1.synchronize()
        times.append(e0.elapsed_time(e1) * 1e-3)
    return float(np.median(times))


def _median_cpu(fn, warmup: int = 5, repeat: int = 100) -> float:
    for _ in range(warmup):
        fn()
    times = []
    for _ in range(repeat):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def _write_csv(prefix: str, rows: list) -> Path:
    RESULTS_DIR.mkdir(parents=True, exist_ok=True)
    path = RESULTS_DIR / f"{prefix}_{datetime.now().strftime('%Y%m%d_%H%M%S')}.csv"
    keys, seen = [], set()
    for r in rows:
        for k in r:
            if k not in seen:
                keys.append(k)
                seen.add(k)
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=keys, extrasaction="ignore")
        w.writeheader()
        w.writerows(rows)
    return path


def _matrix(name: str, N: int):
    """get_matrix, plus 'zipf': the irregular power-law rows of the Zipf(2.0)
    workload (n = N, rows clipped to [1, 4096], seed 2)."""
    from . import generators as g
    if name == "zipf":
        return g.zipf_matrix(N, 2.0, 4096, seed=2)
    return g.get_matrix(name, N)


def run_spmv(args) -> list:
    A = _matrix(args.matrix, args.N)
    N, nnz = A.nrows, A.nnz
    x = np.random.default_rng(42).standard_normal(A.ncols)
    print(f"\nSpMV Benchmark - {args.matrix}  N={N:,}  nnz={nnz:,}  nnz/row={nnz / N:.1f}")
    print("-" * 78)
    print(f"{'Backend':<32} {'Time(us)':>10} {'BW(GB/s)':>10} {'GFLOPS':>8}")
    print("-" * 78)
    rows = []

    def report(tag, m):
        print(f"  {tag:<30} {m['elapsed_us']:>10.1f} {m['bw_GB_s']:>10.2f} {m['gflops']:>8.2f}")
        rows.append({"backend": tag, "N": N, "nnz": nnz, **m})

    if "scipy" in args.backend:
        S = A.to_scipy()
        report("scipy_cpu", spmv_metrics(_median_cpu(lambda: S @ x, repeat=args.repeat), N, nnz))
    gpu = [b for b in args.backend if b.startswith("num_amd")]
    if gpu or args.batch > 0:
        import torch
        from . import kokkos_spmv as ks
        ks.init(0)
        dA = ks.upload_csr(A.rowptr, A.colind, A.values, N, A.ncols)
        dx = torch.from_numpy(x).cuda()
        dy = torch.empty(N, dtype=torch.float64, device="cuda")
        for tag, kk in (("num_amd", True), ("num_amd_custom", False)):
            if tag in args.backend:
                t = _median_events(lambda: ks.spmv(dA, dx, dy, kk, sync=False),
                                   repeat=args.repeat)
                report(f"{tag}_1gpu", spmv_metrics(t, N, nnz))
        if args.batch > 0:
            k = args.batch
            print(f"\n  --- Batch SpMV (k={k}) ---")
            X = torch.from_numpy(np.random.default_rng(7).standard_normal((A.ncols, k))).cuda()
            Y = torch.empty((N, k), dtype=torch.float64, device="cuda")
            t = _median_events(lambda: ks.batch_spmv(dA, X, Y, sync=False), repeat=args.repeat)
            m = spmv_metrics(t, N, nnz, k=k)
            m["k"] = k
            report(f"num_amd_batch_k{k}", m)
    return rows


def run_cg(args) -> list:
    import scipy.sparse.linalg as spla
    from . import generators as g
    A = g.get_matrix(args.matrix, args.N)
    N, nnz = A.nrows, A.nnz
    b = np.ones(N, dtype=np.float64)
    print(f"\nCG Benchmark - {args.matrix}  N={N:,}  nnz={nnz:,}  tol={args.tol:.0e}  np=1")
    print("-" * 78)
    print(f"  {'Backend':<30} {'Time(s)':>9} {'Iters':>7} {'Residual':>12} Converged")
    print("-" * 78)
    rows = []

    def report(m):
        conv = "YES" if m.get("converged") else "NO "
        print(f"  {m['backend']:<30} {m['elapsed_s']:>9.3f} {m.get('iters', '-'):>7} "
              f"{m.get('residual', float('nan')):>12.3e} {conv}")
        rows.append({"N": N, "nnz": nnz, **m})

    S = A.to_scipy()
    if not args.no_scipy:
        it = [0]
        t0 = time.perf_counter()
        xs, info = spla.cg(S, b, rtol=args.tol, maxiter=args.maxiter,
                           callback=lambda _: it.__setitem__(0, it[0] + 1))
        el = time.perf_counter() - t0
        report({"backend": "scipy_cg", "elapsed_s": el, "iters": it[0],
                "residual": float(np.linalg.norm(b - S @ xs) / np.linalg.norm(b)),
                "converged": info == 0})
    import torch
    from . import kokkos_spmv as ks
    ks.init(0)
    dA = ks.upload_csr(A.rowptr, A.colind, A.values, N, A.ncols)
    db = torch.from_numpy(b).cuda()

    def solve(mixed):
        x = torch.zeros(N, dtype=torch.float64, device="cuda")
        if mixed:
            return ks.cg_solve_mixed(dA, db, x, tol=args.tol, max_iter=args.maxiter)
        return ks.cg_solve_local(dA, db, x, tol=args.tol, max_iter=args.maxiter)

    for mixed, tag in ((False, "num_amd_cg_1gpu"), (True, "num_amd_cg_mixed_1gpu")):
        if mixed and not args.mixed:
            continue
        if mixed:
            dA.ensure_f32()
        solve(mixed)  # warm-up, as python/benchmark_cg.py:165
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = solve(mixed)
        torch.cuda.synchronize()
        report({"backend": tag, "elapsed_s": time.perf_counter() - t0, "iters": r["iters"],
                "residual": r["residual"], "converged": r["converged"], "n_ranks": 1})
    x = torch.from_numpy(np.random.default_rng(42).standard_normal(N)).cuda()
    y = torch.empty(N, dtype=torch.float64, device="cuda")
    t = _median_events(lambda: ks.spmv(dA, x, y, sync=False), repeat=50)
    m = spmv_metrics(t, N, nnz)
    rows.append({"N": N, "nnz": nnz, "backend": "spmv_only_1gpu", "elapsed_s": t,
                 "n_ranks": 1, **m})
    print(f"  {'spmv_only_1gpu':<30} {t:>9.6f}       -            -")
    return rows


def _rel_err(a, b) -> float:
    """||a - b|| / ||b|| (||a|| when b is zero), python/verify_accuracy.py:45-50."""
    d = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) if d == 0 else float(np.linalg.norm(a - b) / d)


def _verdict(label: str, err: float, tol: float = 1e-10) -> bool:
    ok = err < tol
    print(f"  {'PASS' if ok else 'FAIL'}  {label:56s}  rel_err = {err:.3e}"
          + ("" if ok else f"  (tol={tol:.0e})"))
    return ok


def run_verify(args) -> bool:
    """The check groups of python/verify_accuracy.py on the HIP path."""
    import scipy.sparse.linalg as spla
    import torch
    from . import generators as g
    from . import kokkos_spmv as ks
    ks.init(0)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    groups = []

    print("\n[1] Single-GPU SpMV")
    ok = True
    for name in ("laplacian_2d", "laplacian_3d", "random_sparse", "lattice_gauge"):
        A = g.get_matrix(name, 50_000)
        x = np.random.default_rng(0).standard_normal(A.ncols)
        dA = ks.upload_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols)
        y = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
        ks.spmv(dA, dev(x), y)
        ok &= _verdict(f"spmv  {name:20s} N={A.nrows:>7,}",
                       _rel_err(y.cpu().numpy(), A.to_scipy() @ x))
    groups.append(ok)

    print("\n[2] Batch SpMV")
    ok = True
    for name in ("laplacian_2d", "random_sparse"):
        A = g.get_matrix(name, 20_000)
        X = np.random.default_rng(1).standard_normal((A.ncols, 64))
        dA = ks.upload_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols)
        Y = torch.zeros((A.nrows, 64), dtype=torch.float64, device="cuda")
        ks.batch_spmv(dA, dev(X), Y)
        ok &= _verdict(f"batch_spmv k= 64  {name:20s} N={A.nrows:>7,}",
                       _rel_err(Y.cpu().numpy(), A.to_scipy() @ X))
    groups.append(ok)

    print("\n[3] Local CG solver")
    ok = True
    for name in ("laplacian_3d", "random_sparse"):
        A = g.get_matrix(name, 10_000)
        b = np.ones(A.nrows)
        x_ref, info = spla.cg(A.to_scipy(), b, rtol=1e-12, maxiter=2000)
        if info != 0:
            print(f"  [WARN] SciPy CG failed (info={info}) for {name}")
        dA = ks.upload_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols)
        x = torch.zeros(A.nrows, dtype=torch.float64, device="cuda")
        r = ks.cg_solve_local(dA, dev(b), x, tol=1e-10, max_iter=2000)
        ok &= _verdict(f"cg_local  {name:20s} N={A.nrows:>7,}  iters={r['iters']:4d}  "
                       f"converged={r['converged']}", _rel_err(x.cpu().numpy(), x_ref), 1e-6)
    groups.append(ok)

    print("\n[4] Distributed SpMV")
    ok = True
    size = int(args.ranks)
    for name in ("laplacian_2d", "laplacian_3d"):
        A = g.get_matrix(name, 40_000)
        x = np.random.default_rng(2).standard_normal(A.ncols)
        D = [ks.distribute_csr(A.rowptr, A.colind, A.values, A.nrows, A.ncols, rank=r, size=size)
             for r in range(size)]
        bufs = [ks.halo_pack(D[r], dev(x[D[r].local_row_start:D[r].local_row_end]))
                for r in range(size)]
        y = np.empty(A.nrows)
        for q in range(size):
            xg = torch.zeros(D[q].n_ghost, dtype=torch.float64, device="cuda")
            for src, _, _ in D[q].recv:
                off, cnt = next((o, c) for dst, o, c in ks.send_ranges(D[src]) if dst == q)
                ks.halo_scatter(D[q], bufs[src][off:off + cnt].contiguous(), xg, src_rank=src)
            s, e = D[q].local_row_start, D[q].local_row_end
            yl = torch.empty(e - s, dtype=torch.float64, device="cuda")
            ks.dist_spmv(D[q], dev(x[s:e]), yl, xg)
            y[s:e] = yl.cpu().numpy()
        ok &= _verdict(f"dist_spmv  {name:20s} N={A.nrows:>7,}  np={size}",
                       _rel_err(y, A.to_scipy() @ x))
    groups.append(ok)

    n_pass = sum(groups)
    print("\n" + "=" * 72)
    print(f"  ALL {len(groups)} test groups PASSED" if all(groups)
          else f"  {n_pass}/{len(groups)} test groups passed - see failures above")
    print("=" * 72)
    return all(groups)


def main(argv=None) -> int:
    from .generators import MATRIX_FAMILIES
    ap = argparse.ArgumentParser(description="num_amd benchmark harness")
    sub = ap.add_subparsers(dest="cmd", required=True)
    sp_ = sub.add_parser("spmv", help="SpMV benchmark (python/benchmark_spmv.py)")
    sp_.add_argument("--N", type=int, default=100_000)
    sp_.add_argument("--matrix", default="laplacian_3d", choices=list(MATRIX_FAMILIES) + ["zipf"])
    sp_.add_argument("--batch", type=int, default=0)
    sp_.add_argument("--repeat", type=int, default=100)
    sp_.add_argument("--backend", nargs="+", default=["scipy", "num_amd", "num_amd_custom"])
    sp_.add_argument("--no-csv", action="store_true")
    cg_ = sub.add_parser("cg", help="CG benchmark (python/benchmark_cg.py)")
    cg_.add_argument("--N", type=int, default=100_000)
    cg_.add_argument("--matrix", default="laplacian_3d",
                     choices=["laplacian_2d", "laplacian_3d", "lattice_gauge", "random_sparse"])
    cg_.add_argument("--tol", type=float, default=1e-8)
    cg_.add_argument("--maxiter", type=int, default=2000)
    cg_.add_argument("--mixed", action="store_true", help="also run the mixed refined solver")
    cg_.add_argument("--no-scipy", action="store_true")
    cg_.add_argument("--no-csv", action="store_true")
    vf_ = sub.add_parser("verify", help="accuracy checks (python/verify_accuracy.py)")
    vf_.add_argument("--ranks", type=int, default=4,
                     help="shares of the distributed SpMV check, run in turn on one GPU")
    args = ap.parse_args(argv)
    if args.cmd == "verify":
        return 0 if run_verify(args) else 1
    rows = run_spmv(args) if args.cmd == "spmv" else run_cg(args)
    if not args.no_csv and rows:
        print(f"\n  Results saved to {_write_csv(args.cmd, rows)}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
