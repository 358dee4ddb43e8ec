// blas1.hip -- vector ops of src/vec_ops.hpp on the device: axpy, axpy2,
// scale, fill, copy, and a deterministic two-stage dot / norm2.
#include <mutex>

#include "blas1.hpp"

namespace numamd {

// Scratch of num_dot_noalloc: the stage-1 partial sums of the two-stage dot,
// one double per stage-1 workgroup (dot_grid caps them at kMaxRedBlocks, so
// any n fits: 16 KiB).  Every GPU the library runs on holds its own copy
// until the process ends.  In host code `scratch` only names the symbol:
// kernels get the per-device address resolved by noalloc_setup.
constexpr int kNoAllocPartials = kMaxRedBlocks;
static __device__ __align__(256) double scratch[kNoAllocPartials];

constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxScratchDevices] = {};

int noalloc_setup()
{
    int dev = -1;
    NUM_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices)
        return fail(NUM_EINVAL, "device %d outside the %d devices the static scratch serves", dev,
                    kMaxScratchDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return NUM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(NUM_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                    hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
    return NUM_OK;
}

// The current device's scratch address (resolved on its first use there).
static int noalloc_scratch(double** out)
{
    int dev = -1;
    NUM_HIP(hipGetDevice(&dev));
    if (dev >= 0 && dev < kMaxScratchDevices) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (g_scratch_ptr[dev]) {
            *out = g_scratch_ptr[dev];
            return NUM_OK;
        }
    }
    const int rc = noalloc_setup();
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    *out = g_scratch_ptr[dev];
    return NUM_OK;
}

__global__ __launch_bounds__(kRedBlock) void axpy_kernel(int64_t n, double a,
                                                         const double* __restrict__ x,
                                                         const double* y, double* z)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        z[i] = a * x[i] + y[i];
}

__global__ __launch_bounds__(kRedBlock) void axpy2_kernel(int64_t n, double a, const double* x,
                                                          double b, const double* y, double* z)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        z[i] = a * x[i] + b * y[i];
}

__global__ __launch_bounds__(kRedBlock) void scale_kernel(int64_t n, double a, double* x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        x[i] *= a;
}

__global__ __launch_bounds__(kRedBlock) void fill_kernel(int64_t n, double v, double* x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        x[i] = v;
}

// Stage 1: block b sums x_i*y_i over i = b*BS + t + k*grid*BS in order of k
// per thread, then a fixed butterfly/wave-order block sum -> partial[b].
__global__ __launch_bounds__(kRedBlock) void dot_partial_kernel(int64_t n,
                                                                const double* __restrict__ x,
                                                                const double* __restrict__ y,
                                                                double* __restrict__ partial)
{
    __shared__ double red[kRedBlock / kWave];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        acc += x[i] * y[i];
    const double s = block_sum<kRedBlock>(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Stage 2: one block sums `np` partials in a fixed order; optional sqrt.
__global__ __launch_bounds__(kRedBlock) void sum_partials_kernel(const double* __restrict__ partial,
                                                                 int np, double* out, int do_sqrt)
{
    __shared__ double red[kRedBlock / kWave];
    double acc = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedBlock) acc += partial[i];
    const double s = block_sum<kRedBlock>(acc, red);
    if (threadIdx.x == 0) out[0] = do_sqrt ? sqrt(s) : s;
}

int elementwise_grid(int64_t n)
{
    int64_t g = (n + kRedBlock * 4 - 1) / (kRedBlock * 4);
    if (g < 1) g = 1;
    if (g > 8192) g = 8192;
    return (int)g;
}

int dot_grid(int64_t n)
{
    int64_t g = (n + kRedBlock * 8 - 1) / (kRedBlock * 8);
    if (g < 1) g = 1;
    if (g > kMaxRedBlocks) g = kMaxRedBlocks;
    return (int)g;
}

int launch_dot(int64_t n, const double* x, const double* y, double* out, double* partial,
               int do_sqrt, hipStream_t stream)
{
    const int g = dot_grid(n);
    dot_partial_kernel<<<g, kRedBlock, 0, stream>>>(n, x, y, partial);
    sum_partials_kernel<<<1, kRedBlock, 0, stream>>>(partial, g, out, do_sqrt);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

}  // namespace numamd

using namespace numamd;

extern "C" {

int num_axpy(int64_t n, double alpha, const double* d_x, const double* d_y, double* d_z,
             void* stream)
{
    if (n < 0) return fail(NUM_EINVAL, "n must be >= 0");
    if (n == 0) return NUM_OK;
    if (!d_x || !d_y || !d_z) return fail(NUM_EINVAL, "NULL vector");
    axpy_kernel<<<elementwise_grid(n), kRedBlock, 0, (hipStream_t)stream>>>(n, alpha, d_x, d_y,
                                                                            d_z);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

int num_axpy2(int64_t n, double alpha, const double* d_x, double beta, const double* d_y,
              double* d_z, void* stream)
{
    if (n < 0) return fail(NUM_EINVAL, "n must be >= 0");
    if (n == 0) return NUM_OK;
    if (!d_x || !d_y || !d_z) return fail(NUM_EINVAL, "NULL vector");
    axpy2_kernel<<<elementwise_grid(n), kRedBlock, 0, (hipStream_t)stream>>>(n, alpha, d_x, beta,
                                                                             d_y, d_z);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

int num_scale(int64_t n, double alpha, double* d_x, void* stream)
{
    if (n < 0) return fail(NUM_EINVAL, "n must be >= 0");
    if (n == 0) return NUM_OK;
    if (!d_x) return fail(NUM_EINVAL, "NULL vector");
    scale_kernel<<<elementwise_grid(n), kRedBlock, 0, (hipStream_t)stream>>>(n, alpha, d_x);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

int num_fill(int64_t n, double val, double* d_x, void* stream)
{
    if (n < 0) return fail(NUM_EINVAL, "n must be >= 0");
    if (n == 0) return NUM_OK;
    if (!d_x) return fail(NUM_EINVAL, "NULL vector");
    fill_kernel<<<elementwise_grid(n), kRedBlock, 0, (hipStream_t)stream>>>(n, val, d_x);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

int num_copy(int64_t n, const double* d_src, double* d_dst, void* stream)
{
    if (n < 0) return fail(NUM_EINVAL, "n must be >= 0");
    if (n == 0 || d_src == d_dst) return NUM_OK;
    if (!d_src || !d_dst) return fail(NUM_EINVAL, "NULL vector");
    NUM_HIP(hipMemcpyAsync(d_dst, d_src, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice,
                           (hipStream_t)stream));
    return NUM_OK;
}

int64_t num_dot_workspace_bytes(int64_t n)
{
    return (int64_t)dot_grid(n < 0 ? 0 : n) * (int64_t)sizeof(double);
}

int num_dot(int64_t n, const double* d_x, const double* d_y, double* d_out, void* d_work,
            void* stream)
{
    if (n < 0) return fail(NUM_EINVAL, "n must be >= 0");
    if (!d_out || !d_work) return fail(NUM_EINVAL, "out or workspace is NULL");
    if (n > 0 && (!d_x || !d_y)) return fail(NUM_EINVAL, "NULL vector");
    return launch_dot(n, d_x, d_y, d_out, (double*)d_work, 0, (hipStream_t)stream);
}

int num_dot_noalloc(int64_t n, const double* d_x, const double* d_y, double* d_out,
                    void* stream)
{
    if (n < 0) return fail(NUM_EINVAL, "n must be >= 0");
    if (!d_out) return fail(NUM_EINVAL, "out is NULL");
    if (n > 0 && (!d_x || !d_y)) return fail(NUM_EINVAL, "NULL vector");
    const int g = dot_grid(n);
    if (g > kNoAllocPartials)
        return fail(NUM_EINVAL, "dot of length %lld needs %d partials; the static scratch "
                    "holds %d", (long long)n, g, kNoAllocPartials);
    double* partial = nullptr;
    const int rc = noalloc_scratch(&partial);
    if (rc) return rc;
    // the same two kernels as num_dot: stage 1 writes partial[0, g) before
    // stage 2 reads them, in stream order, so nothing needs clearing
    return launch_dot(n, d_x, d_y, d_out, partial, 0, (hipStream_t)stream);
}

int num_norm2(int64_t n, const double* d_x, double* d_out, void* d_work, void* stream)
{
    if (n < 0) return fail(NUM_EINVAL, "n must be >= 0");
    if (!d_out || !d_work) return fail(NUM_EINVAL, "out or workspace is NULL");
    if (n > 0 && !d_x) return fail(NUM_EINVAL, "NULL vector");
    return launch_dot(n, d_x, d_x, d_out, (double*)d_work, 1, (hipStream_t)stream);
}

}  // extern "C"
