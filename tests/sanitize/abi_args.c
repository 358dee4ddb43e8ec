/*
 * abi_args.c -- the argument checks of the C ABI (include/num_amd.h), called
 * directly from C under AddressSanitizer + UndefinedBehaviorSanitizer.
 *
 * Every call below must be refused (or answered) by the library's host code
 * before it touches the GPU, so this runs on a machine without one: wrong
 * sizes, NULL pointers, non-monotone rowptr, out-of-range columns, invalid
 * stencil shapes and the workspace-size queries.  Each refusal must carry
 * the documented code and a non-empty, NUL-terminated message.
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "num_amd.h"

static int g_fail = 0;

static void expect(int got, int want, const char* what)
{
    const char* msg = num_last_error();
    if (got != want) {
        fprintf(stderr, "FAIL %s: code %d, expected %d (%s)\n", what, got, want, msg);
        ++g_fail;
        return;
    }
    if (want != NUM_OK && (!msg || !msg[0] || strlen(msg) >= 1024)) {
        fprintf(stderr, "FAIL %s: no error message\n", what);
        ++g_fail;
    }
}

static void expect64(int64_t got, int64_t want, const char* what)
{
    if (got != want) {
        fprintf(stderr, "FAIL %s: %lld, expected %lld\n", what, (long long)got, (long long)want);
        ++g_fail;
    }
}

int main(void)
{
    const int32_t rp_ok[5] = {0, 2, 5, 8, 10};
    const int32_t ci_ok[10] = {0, 1, 0, 1, 2, 1, 2, 3, 2, 3};
    const double v_ok[10] = {2, -1, -1, 2, -1, -1, 2, -1, -1, 2};
    const int32_t rp_bad0[5] = {1, 2, 5, 8, 10};
    const int32_t rp_dec[5] = {0, 2, 1, 8, 10};
    const int32_t ci_hi[10] = {0, 1, 0, 1, 2, 1, 2, 4, 2, 3};
    const int32_t ci_neg[10] = {0, 1, -1, 1, 2, 1, 2, 3, 2, 3};
    num_csr* A = (num_csr*)0x1;
    double x[4] = {1, 2, 3, 4}, y[4];
    double X[8] = {0}, Y[8];

    expect64(num_abi_version(), 1, "abi version");

    /* CSR handle from host arrays (crs_matrix.hpp:45-72 plus range checks) */
    expect(num_csr_create_host(rp_ok, ci_ok, v_ok, 4, 4, NULL), NUM_EINVAL, "create: out NULL");
    expect(num_csr_create_host(rp_ok, ci_ok, v_ok, 0, 4, &A), NUM_EINVAL, "create: nrows 0");
    if (A != NULL) {
        fprintf(stderr, "FAIL create: *out not cleared on error\n");
        ++g_fail;
    }
    expect(num_csr_create_host(rp_ok, ci_ok, v_ok, 4, -3, &A), NUM_EINVAL, "create: ncols < 0");
    expect(num_csr_create_host(NULL, ci_ok, v_ok, 4, 4, &A), NUM_EINVAL, "create: rowptr NULL");
    expect(num_csr_create_host(rp_bad0, ci_ok, v_ok, 4, 4, &A), NUM_EINVAL, "create: rowptr[0]");
    expect(num_csr_create_host(rp_dec, ci_ok, v_ok, 4, 4, &A), NUM_EINVAL, "create: decreasing");
    expect(num_csr_create_host(rp_ok, ci_hi, v_ok, 4, 4, &A), NUM_EINVAL, "create: col >= ncols");
    expect(num_csr_create_host(rp_ok, ci_neg, v_ok, 4, 4, &A), NUM_EINVAL, "create: col < 0");
    expect(num_csr_create_host(rp_ok, NULL, v_ok, 4, 4, &A), NUM_EINVAL, "create: colind NULL");
    expect(num_csr_create_host(rp_ok, ci_ok, NULL, 4, 4, &A), NUM_EINVAL, "create: values NULL");
    /* device-pointer create: the same rowptr checks, then NULL device arrays */
    expect(num_csr_create(rp_ok, NULL, NULL, NULL, 4, 4, &A), NUM_EINVAL, "create dev: NULL");
    expect(num_csr_create(rp_dec, (const int32_t*)8, (const int32_t*)8, (const double*)8, 4, 4,
                          &A),
           NUM_EINVAL, "create dev: decreasing");
    expect(num_csr_create(rp_ok, rp_ok, ci_ok, v_ok, 4, 4, NULL), NUM_EINVAL, "create dev: out");

    /* NULL handles */
    expect(num_csr_shape(NULL, NULL, NULL, NULL), NUM_EINVAL, "shape: NULL");
    expect(num_csr_destroy(NULL), NUM_OK, "destroy: NULL");
    expect(num_csr_set_schedule(NULL, 0, 0, -1, -1, -1), NUM_EINVAL, "set_schedule: NULL");
    expect(num_csr_attach_f32(NULL, NULL), NUM_EINVAL, "attach_f32: NULL");
    expect(num_csr_set_windows(NULL, 1, NULL), NUM_EINVAL, "set_windows: NULL");
    expect(num_csr_window_info(NULL, NULL), NUM_EINVAL, "window_info: NULL");
    expect(num_csr_set_mixed_path(NULL, 0), NUM_EINVAL, "set_mixed_path: NULL");
    expect(num_csr_pattern_info(NULL, NULL, NULL), NUM_EINVAL, "pattern_info: NULL");
    expect(num_spmv(NULL, x, y, NULL), NUM_EINVAL, "spmv: NULL matrix");
    expect(num_batch_spmv(NULL, X, Y, 2, NULL), NUM_EINVAL, "batch_spmv: NULL matrix");
    expect64(num_cg_workspace_bytes(NULL), -1, "cg workspace: NULL");
    expect64(num_cg_mixed_workspace_bytes(NULL), -1, "mixed workspace: NULL");
    {
        num_cg_result r;
        expect(num_cg_solve(NULL, x, y, 1e-10, 10, NULL, 0, &r, NULL), NUM_EINVAL, "cg: NULL");
    }

    /* host round trips refuse before any upload */
    expect(num_spmv_host(rp_ok, ci_hi, v_ok, 4, 4, x, y), NUM_EINVAL, "spmv_host: column");
    expect(num_spmv_host(rp_ok, ci_ok, v_ok, 4, 4, NULL, y), NUM_EINVAL, "spmv_host: x NULL");
    expect(num_batch_spmv_host(rp_ok, ci_ok, v_ok, 4, 4, 0, X, Y), NUM_EINVAL, "batch_host: k");
    expect(num_batch_spmv_host(rp_ok, ci_ok, v_ok, 4, 4, 2, X, NULL), NUM_EINVAL,
           "batch_host: Y NULL");

    /* vector ops (vec_ops.hpp:16-100): lengths and pointers; n = 0 is a no-op */
    expect(num_axpy(-1, 1.0, x, x, y, NULL), NUM_EINVAL, "axpy: n < 0");
    expect(num_axpy(4, 1.0, NULL, x, y, NULL), NUM_EINVAL, "axpy: NULL");
    expect(num_axpy(0, 1.0, NULL, NULL, NULL, NULL), NUM_OK, "axpy: n = 0");
    expect(num_axpy2(-1, 1.0, x, 2.0, x, y, NULL), NUM_EINVAL, "axpy2: n < 0");
    expect(num_axpy2(4, 1.0, x, 2.0, NULL, y, NULL), NUM_EINVAL, "axpy2: NULL");
    expect(num_scale(-5, 2.0, x, NULL), NUM_EINVAL, "scale: n < 0");
    expect(num_scale(4, 2.0, NULL, NULL), NUM_EINVAL, "scale: NULL");
    expect(num_fill(-1, 0.0, x, NULL), NUM_EINVAL, "fill: n < 0");
    expect(num_fill(3, 0.0, NULL, NULL), NUM_EINVAL, "fill: NULL");
    expect(num_copy(-1, x, y, NULL), NUM_EINVAL, "copy: n < 0");
    expect(num_copy(4, x, x, NULL), NUM_OK, "copy: same buffer");
    expect(num_dot(-1, x, x, y, X, NULL), NUM_EINVAL, "dot: n < 0");
    expect(num_dot(4, x, x, y, NULL, NULL), NUM_EINVAL, "dot: workspace NULL");
    expect(num_dot(4, NULL, x, y, X, NULL), NUM_EINVAL, "dot: x NULL");
    expect(num_dot_noalloc(-1, x, x, y, NULL), NUM_EINVAL, "dot_noalloc: n < 0");
    expect(num_dot_noalloc(4, x, x, NULL, NULL), NUM_EINVAL, "dot_noalloc: out NULL");
    expect(num_norm2(4, x, NULL, X, NULL), NUM_EINVAL, "norm2: out NULL");
    expect64(num_dot_workspace_bytes(0), 8, "dot workspace: n = 0");
    expect64(num_dot_workspace_bytes(-7), 8, "dot workspace: n < 0");
    expect64(num_dot_workspace_bytes((int64_t)1 << 40), 2048 * 8, "dot workspace: capped");
    expect(num_hash_uniform(-1, 1, y, NULL), NUM_EINVAL, "hash_uniform: n < 0");
    expect(num_hash_uniform(4, 1, NULL, NULL), NUM_EINVAL, "hash_uniform: NULL");
    expect(num_hash_uniform(0, 1, NULL, NULL), NUM_OK, "hash_uniform: n = 0");

    /* halo pack / scatter (halo_exchange.hpp:222-300) */
    expect(num_halo_pack(NULL, -1, x, y, NULL), NUM_EINVAL, "pack: count < 0");
    expect(num_halo_pack(NULL, 3, x, y, NULL), NUM_EINVAL, "pack: idx NULL");
    expect(num_halo_pack(NULL, 0, NULL, NULL, NULL), NUM_OK, "pack: count 0");
    expect(num_halo_scatter(x, 2, y, -1, NULL), NUM_EINVAL, "scatter: offset < 0");
    expect(num_halo_scatter(NULL, 2, y, 0, NULL), NUM_EINVAL, "scatter: buf NULL");

    /* stencils and the engine's size queries */
    expect(num_stencil_fill(4, 3, 1, 1, 0.0, NULL, NULL, NULL, NULL), NUM_EINVAL, "stencil: ndim");
    expect(num_stencil_fill(2, 3, 0, 1, 0.0, NULL, NULL, NULL, NULL), NUM_EINVAL, "stencil: ny");
    expect(num_stencil_fill(1, 3, 2, 1, 0.0, NULL, NULL, NULL, NULL), NUM_EINVAL,
           "stencil: unused dim");
    expect(num_stencil_fill(2, 3, 3, 1, 0.0, NULL, NULL, NULL, NULL), NUM_EINVAL,
           "stencil: NULL arrays");
    expect64(num_engine_stencil_bytes(2, 4, 4, 1, 2), -1, "stencil bytes: mode");
    expect64(num_engine_stencil_bytes(3, 200, 200, 200, 0), -1, "stencil bytes: plane");
    expect64(num_engine_stencil_bytes(0, 4, 4, 1, 0), -1, "stencil bytes: ndim");
    if (num_engine_stencil_bytes(2, 4, 4, 1, 1) <= num_engine_stencil_bytes(2, 4, 4, 1, 0)) {
        fprintf(stderr, "FAIL stencil bytes: mixed must need more than FP64\n");
        ++g_fail;
    }
    expect64(num_engine_work_bytes(-1, 0), -1, "engine work: n < 0");
    expect64(num_engine_work_bytes(10, 3), -1, "engine work: mode");
    expect64(num_engine_work_bytes(1000, 0), 4 * 8192, "engine work: fp64");
    {
        num_engine* E = (num_engine*)0x1;
        expect(num_engine_create(0, 1, 1, &E), NUM_EINVAL, "engine: 0 systems");
        expect(num_engine_create(4, 1, 1, NULL), NUM_EINVAL, "engine: out NULL");
        expect(num_engine_destroy(NULL), NUM_OK, "engine destroy: NULL");
        expect(num_engine_remove(NULL, 0), NUM_EINVAL, "engine remove: NULL");
        expect(num_engine_launch(NULL, 1, NULL), NUM_EINVAL, "engine launch: NULL");
        expect(num_engine_info(NULL, NULL, NULL, NULL), NUM_EINVAL, "engine info: NULL");
    }

    if (g_fail) {
        fprintf(stderr, "%d argument checks failed\n", g_fail);
        return 1;
    }
    printf("abi_args: all argument checks passed\n");
    return 0;
}
