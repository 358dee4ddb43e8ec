"""Benchmark harness with the reference's command lines and CSV schemas.

    python -m num_amd.harness spmv [--N 100000] [--matrix laplacian_3d]
                                   [--batch K] [--repeat 100]
                                   [--backend scipy num_amd num_amd_custom] [--no-csv]
    python -m num_amd.harness cg   [--N 100000] [--matrix laplacian_3d]
                                   [--tol 1e-8] [--maxiter 2000] [--no-csv] [--mixed]

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
    args = ap.parse_args(argv)
    rows = run_spmv(args) if args.cmd == "spmv" else run_cg(args)
    if not args.no_csv and rows:
        print(f"\n  Results saved to {_write_csv(args.cmd, rows)}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
