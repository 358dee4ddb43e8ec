/*
 * oracle_kats.c -- every function of the CPU oracle (oracle/oracle.c) run
 * under AddressSanitizer + UndefinedBehaviorSanitizer on the reference's
 * known-answer tests: tests/test_spmv_kernel.cpp [1] 4x4 tridiagonal ->
 * [0,0,0,5], [3] identity, [4] batch with X = I, and tests/test_cg.cpp [2]
 * 2x2 -> [1/11, 7/11], [5] zero right-hand side -> 0 iterations, plus the
 * BLAS-1 identities and the mixed solver on a 1-D Laplacian.  Test
 * infrastructure, like the oracle itself.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void orc_spmv(int32_t, const int32_t*, const int32_t*, const double*, const double*, double*);
void orc_spmv_f32(int32_t, const int32_t*, const int32_t*, const float*, const float*, float*);
void orc_batch_spmv(int32_t, int32_t, int32_t, const int32_t*, const int32_t*, const double*,
                    const double*, double*);
void orc_axpy(int64_t, double, const double*, const double*, double*);
void orc_axpy2(int64_t, double, const double*, double, const double*, double*);
double orc_dot(int64_t, const double*, const double*);
double orc_norm2(int64_t, const double*);
void orc_copy(int64_t, const double*, double*);
void orc_fill(int64_t, double*, double);
void orc_scale(int64_t, double*, double);
int orc_cg_solve_local(int32_t, const int32_t*, const int32_t*, const double*, const double*,
                       double*, double, int32_t, int32_t*, double*, int32_t*);
int orc_cg_mixed(int32_t, const int32_t*, const int32_t*, const double*, const double*, double*,
                 double, int32_t, double, int32_t*, int32_t*, double*, int32_t*);

static int g_fail = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            fprintf(stderr, "FAIL line %d: %s\n", __LINE__, #c);  \
            ++g_fail;                                              \
        }                                                          \
    } while (0)

int main(void)
{
    /* [1] tridiag(-1, 2, -1) 4x4, x = [1,2,3,4] -> [0,0,0,5] */
    const int32_t rp[5] = {0, 2, 5, 8, 10};
    const int32_t ci[10] = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3};
    const double v[10] = {2, -1, -1, 2, -1, -1, 2, -1, -1, 2};
    const double x[4] = {1, 2, 3, 4};
    double y[4];
    orc_spmv(4, rp, ci, v, x, y);
    CHECK(y[0] == 0 && y[1] == 0 && y[2] == 0 && y[3] == 5);
    {
        const float v32[10] = {2, -1, -1, 2, -1, -1, 2, -1, -1, 2};
        const float x32[4] = {1, 2, 3, 4};
        float y32[4];
        orc_spmv_f32(4, rp, ci, v32, x32, y32);
        CHECK(y32[0] == 0 && y32[3] == 5);
    }
    /* [4] batch with X = I (4x4, row-major) -> Y = A */
    {
        double X[16] = {0}, Y[16];
        for (int i = 0; i < 4; ++i) X[i * 4 + i] = 1.0;
        orc_batch_spmv(4, 4, 4, rp, ci, v, X, Y);
        for (int r = 0; r < 4; ++r)
            for (int j = rp[r]; j < rp[r + 1]; ++j) CHECK(Y[r * 4 + ci[j]] == v[j]);
    }
    /* [3] identity N = 100 */
    {
        enum { N = 100 };
        int32_t irp[N + 1], ici[N];
        double iv[N], ix[N], iy[N];
        for (int i = 0; i < N; ++i) {
            irp[i] = i;
            ici[i] = i;
            iv[i] = 1.0;
            ix[i] = 0.1 * i;
        }
        irp[N] = N;
        orc_spmv(N, irp, ici, iv, ix, iy);
        for (int i = 0; i < N; ++i) CHECK(iy[i] == ix[i]);
    }
    /* BLAS-1 (vec_ops.hpp:16-100) on integer-valued vectors: exact */
    {
        enum { N = 1001 };
        double a[N], b[N], c[N];
        for (int i = 0; i < N; ++i) {
            a[i] = (double)(i % 13 - 6);
            b[i] = (double)(i % 7 - 3);
        }
        long long ref = 0;
        for (int i = 0; i < N; ++i) ref += (long long)(i % 13 - 6) * (i % 7 - 3);
        CHECK(orc_dot(N, a, b) == (double)ref);
        orc_axpy(N, 2.0, a, b, c);
        for (int i = 0; i < N; ++i) CHECK(c[i] == 2.0 * a[i] + b[i]);
        orc_axpy2(N, 2.0, a, -3.0, b, c);
        for (int i = 0; i < N; ++i) CHECK(c[i] == 2.0 * a[i] - 3.0 * b[i]);
        orc_copy(N, a, c);
        for (int i = 0; i < N; ++i) CHECK(c[i] == a[i]);
        orc_fill(N, c, -2.5);
        for (int i = 0; i < N; ++i) CHECK(c[i] == -2.5);
        orc_scale(N, c, -2.0);
        for (int i = 0; i < N; ++i) CHECK(c[i] == 5.0);
        CHECK(fabs(orc_norm2(N, c) - 5.0 * sqrt((double)N)) < 1e-12);
    }
    /* test_cg.cpp [2]: [[4,1],[1,3]] x = [1,2] -> [1/11, 7/11] */
    {
        const int32_t crp[3] = {0, 2, 4};
        const int32_t cci[4] = {0, 1, 0, 1};
        const double cv[4] = {4, 1, 1, 3};
        const double cb[2] = {1, 2};
        double cx[2] = {0, 0}, res = -1;
        int32_t it = -1, conv = -1;
        CHECK(orc_cg_solve_local(2, crp, cci, cv, cb, cx, 1e-12, 100, &it, &res, &conv) == 0);
        CHECK(conv == 1 && it <= 2);
        CHECK(fabs(cx[0] - 1.0 / 11) < 1e-12 && fabs(cx[1] - 7.0 / 11) < 1e-12);
        /* [5] zero right-hand side: 0 iterations, residual exactly 0 */
        const double zb[2] = {0, 0};
        double zx[2] = {0, 0};
        CHECK(orc_cg_solve_local(2, crp, cci, cv, zb, zx, 1e-10, 100, &it, &res, &conv) == 0);
        CHECK(it == 0 && res == 0.0 && conv == 1);
    }
    /* mixed refinement on the 1-D Laplacian N = 64, b = 1: FP64 true residual */
    {
        enum { N = 64 };
        int32_t lrp[N + 1], lci[3 * N];
        double lv[3 * N], lb[N], lx[N], ly[N];
        int32_t p = 0;
        for (int i = 0; i < N; ++i) {
            lrp[i] = p;
            if (i > 0) { lci[p] = i - 1; lv[p++] = -1.0; }
            lci[p] = i; lv[p++] = 2.0;
            if (i < N - 1) { lci[p] = i + 1; lv[p++] = -1.0; }
            lb[i] = 1.0;
            lx[i] = 0.0;
        }
        lrp[N] = p;
        int32_t outer = 0, inner = 0, conv = 0;
        double tres = 1.0;
        const int rc = orc_cg_mixed(N, lrp, lci, lv, lb, lx, 1e-10, 10000, 0.01, &outer, &inner,
                                    &tres, &conv);
        CHECK(rc == 0 && conv == 1 && tres < 1e-10);
        orc_spmv(N, lrp, lci, lv, lx, ly);
        double r2 = 0.0;
        for (int i = 0; i < N; ++i) r2 += (lb[i] - ly[i]) * (lb[i] - ly[i]);
        CHECK(sqrt(r2) / sqrt((double)N) < 1e-10);
    }
    if (g_fail) {
        fprintf(stderr, "%d oracle checks failed\n", g_fail);
        return 1;
    }
    printf("oracle_kats: all checks passed\n");
    return 0;
}
