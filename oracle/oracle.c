/*
 * oracle.c -- CPU restatement of the reference's Kokkos CPU path for CSR SpMV,
 * batched (multi-RHS) SpMV, the BLAS-1 vector ops and unpreconditioned CG.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: it may be loaded
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and there only as the checker or the timed CPU baseline.  The product path
 * (num_amd) never links, loads or calls it.
 *
 * Built with:  gcc -O3 -fopenmp -ffp-contract=off -shared -fPIC
 * (-ffp-contract=off keeps every a*b+c as a rounded multiply followed by a
 *  rounded add, which is what the reference's scalar loops mean.)
 *
 * Reference files followed (paths relative to the reference checkout):
 *   src/spmv_kernel.hpp:79-98        SpMVFunctorFlat: acc=0; acc += v*x[c]; y=acc
 *   src/batch_spmv_kernel.hpp:94-102 BatchSpMVFunctorMDRange row loop per rhs b
 *   src/vec_ops.hpp:16-100           axpy / axpy2 / local_dot / norm2 / copy / fill / scale
 *   src/cg_solver.hpp:133-196        cg_solve_local, exact operation order
 *
 * Parity pinning (each function is checked by tests/test_oracle.py against the
 * reference's known-answer tests, scipy, closed forms and brute force; see the
 * per-function notes below).  A mutant switch (orc_set_mutant) lets the tests
 * plant a plausible mistake in the oracle and show that the checks catch it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

/* ------------------------------------------------------------------------ */
/* Test-only mutant switch: 0 = faithful.  Non-zero values plant one mistake.  */
/*   1 spmv drops the last term of every row                                 */
/*   2 spmv multiplies by the transpose (y[c] += v*x[row])                    */
/*   3 cg initial residual with the wrong sign (r = Ap + b)                   */
/*   4 cg beta with swapped indices (beta = rr_old / rr_new)                  */
/*   5 dot skips the last element                                            */
/*   6 axpy2 swaps its two coefficients                                      */
/*   7 batch spmv reads X column-major instead of row-major                   */
/*   8 fill leaves the last element untouched                                 */
/*   9 copy copies all but the first element                                  */
/* ------------------------------------------------------------------------ */
static int g_mutant = 0;
static int g_threads = 0; /* 0 = OpenMP default */

void orc_set_mutant(int m) { g_mutant = m; }
int orc_get_mutant(void) { return g_mutant; }
void orc_set_threads(int t) { g_threads = t; }
int orc_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

static int nthreads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

/* ---- SpMV: src/spmv_kernel.hpp:79-98 ------------------------------------ */
/* Pinned by: test_spmv_kernel.cpp [1] (4x4 tridiagonal -> [0,0,0,5]),
 * [3] identity, [5] LCG N=1e4 (unsorted, repeated columns) vs dense loop;
 * scipy A@x; dense brute force; Laplacian * const = boundary degree. */
void orc_spmv(int32_t nrows, const int32_t* rowptr, const int32_t* colind,
              const double* vals, const double* x, double* y)
{
    if (g_mutant == 2) {
        /* planted mistake: transpose */
        memset(y, 0, sizeof(double) * (size_t)nrows);
        for (int32_t i = 0; i < nrows; ++i)
            for (int32_t j = rowptr[i]; j < rowptr[i + 1]; ++j)
                y[colind[j]] += vals[j] * x[i];
        return;
    }
    const int drop = (g_mutant == 1);
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int32_t i = 0; i < nrows; ++i) {
        double acc = 0.0;
        const int32_t end = rowptr[i + 1] - (drop && rowptr[i + 1] > rowptr[i] ? 1 : 0);
        for (int32_t j = rowptr[i]; j < end; ++j)
            acc += vals[j] * x[colind[j]];
        y[i] = acc;
    }
}

/* FP32 form of the same row loop (values, x and the accumulator in fp32).
 * Used for the mixed-precision inner solve.  Pinned by: identical structure to
 * orc_spmv, checked against numpy float32 row sums in sequential order. */
void orc_spmv_f32(int32_t nrows, const int32_t* rowptr, const int32_t* colind,
                  const float* vals, const float* x, float* y)
{
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int32_t i = 0; i < nrows; ++i) {
        float acc = 0.0f;
        for (int32_t j = rowptr[i]; j < rowptr[i + 1]; ++j)
            acc += vals[j] * x[colind[j]];
        y[i] = acc;
    }
}

/* ---- Batch SpMV: src/batch_spmv_kernel.hpp:94-102 ----------------------- */
/* X is (ncols, k) row-major, Y is (nrows, k) row-major (LayoutRight views,
 * src/batch_spmv_kernel.hpp:84-87).  Pinned by test_spmv_kernel.cpp [4]
 * (X = I4 -> columns of A), k=1 == orc_spmv, scipy A@X. */
void orc_batch_spmv(int32_t nrows, int32_t ncols, int32_t k, const int32_t* rowptr,
                    const int32_t* colind, const double* vals, const double* X, double* Y)
{
    const int colmajor = (g_mutant == 7);
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int32_t i = 0; i < nrows; ++i) {
        for (int32_t b = 0; b < k; ++b) {
            double acc = 0.0;
            for (int32_t j = rowptr[i]; j < rowptr[i + 1]; ++j) {
                const int64_t c = colind[j];
                const double xv = colmajor ? X[(int64_t)b * ncols + c] : X[c * k + b];
                acc += vals[j] * xv;
            }
            Y[(int64_t)i * k + b] = acc;
        }
    }
}

/* ---- BLAS-1: src/vec_ops.hpp ------------------------------------------- */
/* axpy z = alpha*x + y  (vec_ops.hpp:16-26), aliasing allowed. */
void orc_axpy(int64_t n, double alpha, const double* x, const double* y, double* z)
{
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int64_t i = 0; i < n; ++i) z[i] = alpha * x[i] + y[i];
}

/* axpy2 z = alpha*x + beta*y  (vec_ops.hpp:31-42). */
void orc_axpy2(int64_t n, double alpha, const double* x, double beta, const double* y,
               double* z)
{
    if (g_mutant == 6) { double t = alpha; alpha = beta; beta = t; }
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int64_t i = 0; i < n; ++i) z[i] = alpha * x[i] + beta * y[i];
}

/* dot = sum x_i*y_i  (vec_ops.hpp:47-59).  Kokkos' OpenMP parallel_reduce
 * gives each thread a contiguous chunk, sums it in index order and combines
 * the thread partials in thread order; this does the same with T threads.
 * With T = 1 it is the plain sequential sum.  Pinned by exact sums of
 * integer-valued vectors, math.fsum to 1e-15, and the 2x2 CG KAT. */
double orc_dot(int64_t n, const double* x, const double* y)
{
    if (g_mutant == 5 && n > 0) n -= 1;
    const int T = nthreads();
    double part[1024];
    const int TT = T > 1024 ? 1024 : T;
#pragma omp parallel num_threads(TT)
    {
        const int t = omp_get_thread_num();
        const int nt = omp_get_num_threads();
        const int64_t chunk = (n + nt - 1) / nt;
        const int64_t lo = (int64_t)t * chunk;
        const int64_t hi = lo + chunk < n ? lo + chunk : n;
        double acc = 0.0;
        for (int64_t i = lo; i < hi; ++i) acc += x[i] * y[i];
        part[t] = acc;
        if (t == 0) {
            for (int u = nt; u < TT; ++u) part[u] = 0.0;
        }
    }
    double s = 0.0;
    for (int t = 0; t < TT; ++t) s += part[t];
    return s;
}

/* norm2 = sqrt(dot(x,x))  (vec_ops.hpp:64-69). */
double orc_norm2(int64_t n, const double* x) { return sqrt(orc_dot(n, x, x)); }

/* copy / fill / scale  (vec_ops.hpp:74-100).  Pinned by elementwise
 * identities over every element (a copy equal to its source, a fill of a
 * value equal to np.full, a scale equal to the product), including the first
 * and last element; mutants 8 and 9 are caught. */
void orc_copy(int64_t n, const double* src, double* dst)
{
    if (g_mutant == 9 && n > 0) {
        memcpy(dst + 1, src + 1, sizeof(double) * (size_t)(n - 1));
        return;
    }
    memcpy(dst, src, sizeof(double) * (size_t)n);
}
void orc_fill(int64_t n, double* x, double val)
{
    const int64_t m = (g_mutant == 8 && n > 0) ? n - 1 : n;
    for (int64_t i = 0; i < m; ++i) x[i] = val;
}
void orc_scale(int64_t n, double* x, double alpha)
{
#pragma omp parallel for schedule(static) num_threads(nthreads())
    for (int64_t i = 0; i < n; ++i) x[i] *= alpha;
}

/* ---- CG: src/cg_solver.hpp:133-196 -------------------------------------- */
/* Exact operation order of cg_solve_local:
 *   Ap = A x0; r = -1*Ap + b; p = r; bn = ||b|| (0 -> 1); rr = r.r;
 *   if sqrt(rr)/bn < tol: return {0, sqrt(rr)/bn, true}
 *   loop iter: Ap = A p; pAp = p.Ap; alpha = rr/pAp; x = alpha*p + x;
 *              r = -alpha*Ap + r; rrn = r.r; rel = sqrt(rrn)/bn;
 *              if rel < tol: return {iter+1, rel, true}
 *              beta = rrn/rr; rr = rrn; p = 1*r + beta*p
 *   return {max_iter, sqrt(r.r)/bn, false}
 * Returns 0.  Pinned by test_cg.cpp [1]-[7],[9], scipy cg, the exact
 * solutions (identity, 2x2 -> [1/11, 7/11], diagonal -> 1/d, sine mode) and
 * the finite-termination property (n distinct eigenvalues -> <= n steps). */
int orc_cg_solve_local(int32_t n, const int32_t* rowptr, const int32_t* colind,
                       const double* vals, const double* b, double* x, double tol,
                       int32_t max_iter, int32_t* iters, double* final_res,
                       int32_t* converged)
{
    double* r = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* p = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double* Ap = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!r || !p || !Ap) { free(r); free(p); free(Ap); return -1; }

    orc_spmv(n, rowptr, colind, vals, x, Ap);
    if (g_mutant == 3) orc_axpy(n, 1.0, Ap, b, r);
    else orc_axpy(n, -1.0, Ap, b, r);
    orc_copy(n, r, p);

    double bn = orc_norm2(n, b);
    if (bn == 0.0) bn = 1.0;
    double rr = orc_dot(n, r, r);
    if (sqrt(rr) / bn < tol) {
        *iters = 0; *final_res = sqrt(rr) / bn; *converged = 1;
        free(r); free(p); free(Ap);
        return 0;
    }
    for (int32_t it = 0; it < max_iter; ++it) {
        orc_spmv(n, rowptr, colind, vals, p, Ap);
        const double pAp = orc_dot(n, p, Ap);
        const double alpha = rr / pAp;
        orc_axpy(n, alpha, p, x, x);
        orc_axpy(n, -alpha, Ap, r, r);
        const double rrn = orc_dot(n, r, r);
        const double rel = sqrt(rrn) / bn;
        if (rel < tol) {
            *iters = it + 1; *final_res = rel; *converged = 1;
            free(r); free(p); free(Ap);
            return 0;
        }
        const double beta = (g_mutant == 4) ? rr / rrn : rrn / rr;
        rr = rrn;
        orc_axpy2(n, 1.0, r, beta, p, p);
    }
    *iters = max_iter;
    *final_res = sqrt(orc_dot(n, r, r)) / bn;
    *converged = 0;
    free(r); free(p); free(Ap);
    return 0;
}

/* ---- Mixed-precision refinement (no reference counterpart) --------------- */
/* The oracle for the FP32-inner / FP64-refined form with reliable updates:
 * CG steps in FP32 storage (FP32 matrix copy, FP32 vectors, dot products
 * accumulated in FP64, scalars in FP64), in the reference's CG order.  When
 * the FP32 residual has dropped by `delta` since the last refinement, or
 * looks converged, the iterate is refined in FP64: x64 += x32, x32 = 0,
 * r = b - A x64 (orc_spmv / orc_axpy order), the FP32 residual is re-seeded
 * from r, beta = r32.r32 / rr, and the search direction is kept.  The FP64
 * true residual is the convergence test (strict <).  Pinned by that true
 * residual ||b - A x|| / ||b|| < tol, a property independent of rounding
 * details, and by the solution agreeing with orc_cg_solve_local's. */
static double dot_f32_acc64(int64_t n, const float* a, const float* b)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
    return s;
}

/* x64 += x32, x32 = 0, r = b - A x64 (FP64), r32 = (float) r.
 * Returns ||r||^2 in *rr64 and r32.r32 in *rr32. */
static void refine(int32_t n, const int32_t* rowptr, const int32_t* colind, const double* vals,
                   const double* b, double* x, float* x32, double* r, float* r32,
                   double* rr64, double* rr32)
{
    for (int32_t i = 0; i < n; ++i) { x[i] = x[i] + (double)x32[i]; x32[i] = 0.0f; }
    orc_spmv(n, rowptr, colind, vals, x, r);
    orc_axpy(n, -1.0, r, b, r);
    double s64 = 0.0, s32 = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        r32[i] = (float)r[i];
        s64 += r[i] * r[i];
        s32 += (double)r32[i] * (double)r32[i];
    }
    *rr64 = s64;
    *rr32 = s32;
}

int orc_cg_mixed(int32_t n, const int32_t* rowptr, const int32_t* colind, const double* vals,
                 const double* b, double* x, double tol, int32_t max_iter, double delta,
                 int32_t* nref, int32_t* steps, double* true_res, int32_t* converged)
{
    const int32_t nnz = rowptr[n];
    float* v32 = (float*)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    float* r32 = (float*)malloc(sizeof(float) * (size_t)n);
    float* x32 = (float*)calloc((size_t)n, sizeof(float));
    float* pold = (float*)calloc((size_t)n, sizeof(float));
    float* pnew = (float*)malloc(sizeof(float) * (size_t)n);
    float* q = (float*)malloc(sizeof(float) * (size_t)n);
    if (!v32 || !r || !r32 || !x32 || !pold || !pnew || !q) return -1;
    for (int32_t j = 0; j < nnz; ++j) v32[j] = (float)vals[j];
    double bn = sqrt(orc_dot(n, b, b));
    if (bn == 0.0) bn = 1.0;
    double rr64, rr32;
    refine(n, rowptr, colind, vals, b, x, x32, r, r32, &rr64, &rr32);
    int32_t refs = 1, it = 0, rc = 0;
    double rel = sqrt(rr64) / bn;
    *converged = 0;
    if (rel < tol) { *converged = 1; goto out; }
    if (max_iter <= 0) goto out;
    double rr = rr32, rr_ref = rr32, beta = 0.0;
    for (;;) {
        const float bf = (float)beta;
        for (int32_t i = 0; i < n; ++i) pnew[i] = r32[i] + bf * pold[i];
        orc_spmv_f32(n, rowptr, colind, v32, pnew, q);
        const double pq = dot_f32_acc64(n, pnew, q);
        if (!(fabs(pq) >= 1e-300)) { rc = 1; break; }
        const float al = (float)(rr / pq), nal = -al;
        double rrn = 0.0;
        for (int32_t i = 0; i < n; ++i) {
            x32[i] = al * pnew[i] + x32[i];
            r32[i] = nal * q[i] + r32[i];
            rrn += (double)r32[i] * (double)r32[i];
        }
        ++it;
        float* t = pold; pold = pnew; pnew = t;
        const double est = sqrt(rrn) / bn;
        const int final_ = it >= max_iter;
        if (final_ || est < tol || sqrt(rrn) < delta * sqrt(rr_ref) || !(rrn == rrn)) {
            refine(n, rowptr, colind, vals, b, x, x32, r, r32, &rr64, &rr32);
            ++refs;
            rel = sqrt(rr64) / bn;
            if (rel < tol) { *converged = 1; break; }
            if (final_ || !(rel == rel)) break;
            beta = rr32 / rr;
            rr = rr32;
            rr_ref = rr32;
        } else {
            beta = rrn / rr;
            rr = rrn;
        }
    }
    for (int32_t i = 0; i < n; ++i) x[i] = x[i] + (double)x32[i];
out:
    *nref = refs;
    *steps = it;
    *true_res = rel;
    free(v32); free(r); free(r32); free(x32); free(pold); free(pnew); free(q);
    return rc;
}
