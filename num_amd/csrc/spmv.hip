// spmv.hip -- y = A x (FP64 / FP32), row-range SpMV, multi-RHS SpMV, halo
// pack/scatter and the host round-trip helpers.
#include <vector>

#include "spmv_launch.hpp"

namespace numamd {

template <typename T, class Gather>
static int launch_spmv_g(const num_csr* A, const Schedule& s, const T* vals, const Gather& x,
                         T* y, hipStream_t stream)
{
    launch_tiles<T>(A, s, vals, x, EpiY<T>{y}, stream);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

template <typename T>
int launch_spmv(const num_csr* A, const Schedule& s, const T* vals, const T* x, T* y,
                hipStream_t stream)
{
    return launch_spmv_g<T>(A, s, vals, GatherX<T>{x}, y, stream);
}

template int launch_spmv<double>(const num_csr*, const Schedule&, const double*, const double*,
                                 double*, hipStream_t);
template int launch_spmv<float>(const num_csr*, const Schedule&, const float*, const float*,
                                float*, hipStream_t);

// ---------------------------------------------------------------------------
// Multi-RHS: Y(i, b) = sum_j A_ij X(c_j, b), row-major X (ncols, k), Y (nrows, k)
// (src/batch_spmv_kernel.hpp:94-102 order: sequential over the row's
// non-zeros from 0.0).  KP lanes per row, lane b covers rhs b, b+KP, ...; the
// KP lanes of a row read the same value/column (one broadcast load) and
// consecutive X entries (one coalesced load per KP lanes).
// ---------------------------------------------------------------------------
template <int KP>
__global__ __launch_bounds__(256) void batch_spmv_kernel(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const double* __restrict__ vals, const double* __restrict__ X, double* __restrict__ Y,
    int nrows, int k)
{
    constexpr int RPB = 256 / KP;
    const int row = blockIdx.x * RPB + threadIdx.x / KP;
    const int lane = threadIdx.x % KP;
    if (row >= nrows) return;
    const int js = rowptr[row], je = rowptr[row + 1];
    for (int b = lane; b < k; b += KP) {
        double acc = 0.0;
        for (int j = js; j < je; ++j) acc += vals[j] * X[(int64_t)colind[j] * k + b];
        Y[(int64_t)row * k + b] = acc;
    }
}

template <int KP>
static void launch_batch(const num_csr* A, const double* X, double* Y, int k,
                         hipStream_t stream)
{
    constexpr int RPB = 256 / KP;
    const int grid = (A->nrows + RPB - 1) / RPB;
    batch_spmv_kernel<KP><<<grid, 256, 0, stream>>>(A->rowptr, A->colind, A->vals, X, Y,
                                                    A->nrows, k);
}

// One row of the Dirichlet stencil on an nx x ny x nz grid (x fastest):
// entries in column order -z, -y, -x, self, +x, +y, +z, off-diagonals -1,
// diagonal 2*ndim + shift.
__global__ void stencil_fill_kernel(int ndim, int nx, int ny, int nz, double shift,
                                    const int32_t* __restrict__ rowptr,
                                    int32_t* __restrict__ colind, double* __restrict__ vals)
{
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ix = (int)(i % nx);
    const int iy = (int)((i / nx) % ny);
    const int iz = (int)(i / ((int64_t)nx * ny));
    const int sy = nx, sz = nx * ny;
    int p = rowptr[i];
    const int row = (int)i;
    if (ndim >= 3 && iz > 0) { colind[p] = row - sz; vals[p++] = -1.0; }
    if (ndim >= 2 && iy > 0) { colind[p] = row - sy; vals[p++] = -1.0; }
    if (ix > 0) { colind[p] = row - 1; vals[p++] = -1.0; }
    colind[p] = row;
    vals[p++] = 2.0 * ndim + shift;
    if (ix < nx - 1) { colind[p] = row + 1; vals[p++] = -1.0; }
    if (ndim >= 2 && iy < ny - 1) { colind[p] = row + sy; vals[p++] = -1.0; }
    if (ndim >= 3 && iz < nz - 1) { colind[p] = row + sz; vals[p++] = -1.0; }
}

__global__ void halo_pack_kernel(const int32_t* __restrict__ idx, int count,
                                 const double* __restrict__ x, double* __restrict__ buf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) buf[i] = x[idx[i]];
}

__global__ void halo_scatter_kernel(const double* __restrict__ buf, int count,
                                    double* __restrict__ xg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) xg[i] = buf[i];
}

}  // namespace numamd

using namespace numamd;

extern "C" {

int num_spmv(const num_csr* A, const double* d_x, double* d_y, void* stream)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (!d_x || !d_y) return fail(NUM_EINVAL, "x or y is NULL");
    int rc = check_device(A);
    if (rc) return rc;
    return launch_spmv<double>(A, A->sched, A->vals, d_x, d_y, (hipStream_t)stream);
}

int num_spmv_rows(const num_csr* A, int32_t row_begin, int32_t row_end, const double* d_x,
                  double* d_y, void* stream)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (!d_x || !d_y) return fail(NUM_EINVAL, "x or y is NULL");
    int rc = check_device(A);
    if (rc) return rc;
    const Schedule* s = nullptr;
    rc = range_schedule(A, row_begin, row_end, &s);
    if (rc) return rc;
    return launch_spmv<double>(A, *s, A->vals, d_x, d_y, (hipStream_t)stream);
}

int num_spmv_split(const num_csr* A, int32_t row_begin, int32_t row_end, const double* d_x_own,
                   const double* d_x_ghost, double* d_y, void* stream)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (!d_x_own || !d_y) return fail(NUM_EINVAL, "x_own or y is NULL");
    if (A->ncols > A->nrows && !d_x_ghost)
        return fail(NUM_EINVAL, "matrix has %d ghost columns but x_ghost is NULL",
                    A->ncols - A->nrows);
    int rc = check_device(A);
    if (rc) return rc;
    const Schedule* s = nullptr;
    rc = range_schedule(A, row_begin, row_end, &s);
    if (rc) return rc;
    return launch_spmv_g<double>(A, *s, A->vals,
                                 GatherSplit<double>{d_x_own, d_x_ghost, A->nrows}, d_y,
                                 (hipStream_t)stream);
}

int num_spmv_f32(const num_csr* A, const float* d_x, float* d_y, void* stream)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (!A->vals32) return fail(NUM_EINVAL, "no FP32 values attached (num_csr_attach_f32)");
    if (!d_x || !d_y) return fail(NUM_EINVAL, "x or y is NULL");
    int rc = check_device(A);
    if (rc) return rc;
    return launch_spmv<float>(A, A->sched, A->vals32, d_x, d_y, (hipStream_t)stream);
}

int num_batch_spmv(const num_csr* A, const double* d_X, double* d_Y, int32_t k, void* stream)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (k < 0) return fail(NUM_EINVAL, "k must be >= 0");
    if (k == 0) return NUM_OK;  // src/batch_spmv_kernel.hpp:130
    if (!d_X || !d_Y) return fail(NUM_EINVAL, "X or Y is NULL");
    int rc = check_device(A);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    if (k == 1) return launch_spmv<double>(A, A->sched, A->vals, d_X, d_Y, st);
    if (k <= 2) launch_batch<2>(A, d_X, d_Y, k, st);
    else if (k <= 4) launch_batch<4>(A, d_X, d_Y, k, st);
    else if (k <= 8) launch_batch<8>(A, d_X, d_Y, k, st);
    else if (k <= 16) launch_batch<16>(A, d_X, d_Y, k, st);
    else if (k <= 32) launch_batch<32>(A, d_X, d_Y, k, st);
    else launch_batch<64>(A, d_X, d_Y, k, st);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

int num_stencil_fill(int32_t ndim, int32_t nx, int32_t ny, int32_t nz, double shift,
                     const int32_t* d_rowptr, int32_t* d_colind, double* d_vals, void* stream)
{
    if (ndim < 1 || ndim > 3) return fail(NUM_EINVAL, "ndim must be 1, 2 or 3");
    if (nx < 1 || ny < 1 || nz < 1) return fail(NUM_EINVAL, "grid sizes must be >= 1");
    if ((ndim < 3 && nz != 1) || (ndim < 2 && ny != 1))
        return fail(NUM_EINVAL, "unused grid dimensions must be 1");
    const int64_t n = (int64_t)nx * ny * nz;
    if (n * (2 * ndim + 1) >= (int64_t)1 << 31) return fail(NUM_EINVAL, "grid too large");
    if (!d_rowptr || !d_colind || !d_vals) return fail(NUM_EINVAL, "NULL pointer");
    stencil_fill_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        ndim, nx, ny, nz, shift, d_rowptr, d_colind, d_vals);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

int num_halo_pack(const int32_t* d_idx, int32_t count, const double* d_x, double* d_buf,
                  void* stream)
{
    if (count < 0) return fail(NUM_EINVAL, "count must be >= 0");
    if (count == 0) return NUM_OK;
    if (!d_idx || !d_x || !d_buf) return fail(NUM_EINVAL, "NULL pointer");
    halo_pack_kernel<<<(count + 255) / 256, 256, 0, (hipStream_t)stream>>>(d_idx, count, d_x,
                                                                            d_buf);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

int num_halo_scatter(const double* d_buf, int32_t count, double* d_x_ghost, int32_t offset,
                     void* stream)
{
    if (count < 0 || offset < 0) return fail(NUM_EINVAL, "count and offset must be >= 0");
    if (count == 0) return NUM_OK;
    if (!d_buf || !d_x_ghost) return fail(NUM_EINVAL, "NULL pointer");
    halo_scatter_kernel<<<(count + 255) / 256, 256, 0, (hipStream_t)stream>>>(
        d_buf, count, d_x_ghost + offset);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

static int roundtrip(const int32_t* h_rowptr, const int32_t* h_colind, const double* h_vals,
                     int32_t nrows, int32_t ncols, int32_t k, const double* h_x, double* h_y)
{
    num_csr* A = nullptr;
    int rc = num_csr_create_host(h_rowptr, h_colind, h_vals, nrows, ncols, &A);
    if (rc) return rc;
    double* dx = nullptr;
    double* dy = nullptr;
    const size_t kx = (size_t)ncols * (size_t)(k > 0 ? k : 1);
    const size_t ky = (size_t)nrows * (size_t)(k > 0 ? k : 1);
    hipError_t e = hipMalloc(&dx, kx * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dy, ky * sizeof(double));
    if (e == hipSuccess) e = hipMemcpy(dx, h_x, kx * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        rc = (k == 0) ? num_spmv(A, dx, dy, nullptr) : num_batch_spmv(A, dx, dy, k, nullptr);
        if (rc == NUM_OK) {
            e = hipDeviceSynchronize();
            if (e == hipSuccess)
                e = hipMemcpy(h_y, dy, ky * sizeof(double), hipMemcpyDeviceToHost);
        }
    }
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    num_csr_destroy(A);
    if (rc) return rc;
    if (e != hipSuccess) return fail(NUM_EHIP, "round trip failed: %s", hipGetErrorString(e));
    return NUM_OK;
}

int num_spmv_host(const int32_t* h_rowptr, const int32_t* h_colind, const double* h_vals,
                  int32_t nrows, int32_t ncols, const double* h_x, double* h_y)
{
    if (!h_x || !h_y) return fail(NUM_EINVAL, "x or y is NULL");
    return roundtrip(h_rowptr, h_colind, h_vals, nrows, ncols, 0, h_x, h_y);
}

int num_batch_spmv_host(const int32_t* h_rowptr, const int32_t* h_colind, const double* h_vals,
                        int32_t nrows, int32_t ncols, int32_t k, const double* h_X, double* h_Y)
{
    if (k <= 0) return fail(NUM_EINVAL, "k must be >= 1");
    if (!h_X || !h_Y) return fail(NUM_EINVAL, "X or Y is NULL");
    return roundtrip(h_rowptr, h_colind, h_vals, nrows, ncols, k, h_X, h_Y);
}

}  // extern "C"
