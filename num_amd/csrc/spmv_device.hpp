// spmv_device.hpp -- row-block CSR SpMV kernel for gfx950, shared by the plain
// SpMV entry points and the fused CG kernels.
//
// Work unit: one workgroup of BS threads owns a block of consecutive rows
// [blk_row[2t], blk_row[2t+2]) whose non-zeros fit the block's LDS budget
// (BS*NPT entries).  Three phases:
//   1. stream: every thread loads NPT non-zeros at stride BS, so each wave
//      load instruction reads 64 consecutive values / column indices (fully
//      coalesced whatever the row lengths), gathers x[col], and stores the
//      rounded product v*x into LDS;
//   2. reduce: G lanes per row (blk_row[2t+1]; G = 1 when no row of the block
//      has more than 32 entries) sum the row's products from LDS; with G = 1
//      the sum runs in index order from
//      0.0, the exact arithmetic of the reference's row loop
//      (src/spmv_kernel.hpp:90-97) -- bit-identical results;
//   3. epilogue: a functor consumes (row, value) -- store y, or the fused CG
//      updates -- and can contribute to a deterministic per-block partial sum.
// A row longer than the LDS budget gets a workgroup of its own ("long" block)
// that accumulates at stride BS and reduces with a fixed butterfly.
// Blocks are mapped XCD-contiguously (common.hpp: xcd_tile).
#pragma once
#include "common.hpp"

namespace numamd {

template <typename T>
__device__ __forceinline__ T shfl_xor_w(T v, int off, int width)
{
    return __shfl_xor(v, off, width);
}

// x accessors: one vector, or the [own | ghost] pair of a rank's slab where
// columns >= n_own address the ghost block (no x_ext copy is ever built).
template <typename T>
struct GatherX {
    const T* __restrict__ x;
    __device__ __forceinline__ T operator()(int c) const { return x[c]; }
};

template <typename T>
struct GatherSplit {
    const T* __restrict__ x_own;
    const T* __restrict__ x_ghost;
    int n_own;
    __device__ __forceinline__ T operator()(int c) const
    {
        return c < n_own ? x_own[c] : x_ghost[c - n_own];
    }
};

__device__ __forceinline__ int pow2_floor(int v)
{
    return v <= 1 ? 1 : (1 << (31 - __clz(v)));
}

// Kernel body; `tile` is the logical row block.  Returns the two epilogue
// contributions of this thread (zero for non-leader lanes).
template <typename T, int BS, int NPT, class Epi, class Gather>
__device__ __forceinline__ void spmv_block_body(
    int tile, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const T* __restrict__ vals, const Gather& x, const int32_t* __restrict__ blk_row,
    const Epi& epi, T* __restrict__ prod, double& ca, double& cb)
{
    constexpr int NNZ_MAX = BS * NPT;
    const int t = threadIdx.x;
    // schedule entries are [first row, lanes per row] pairs, plus a sentinel
    const int r0 = blk_row[2 * tile];
    const int G = blk_row[2 * tile + 1];
    const int r1 = blk_row[2 * tile + 2];
    const int j0 = rowptr[r0];
    const int j1 = rowptr[r1];
    const int nz = j1 - j0;
    ca = 0.0;
    cb = 0.0;
    if (nz > NNZ_MAX) {
        // long row: one row, whole workgroup, stride-BS partial sums then a
        // fixed-shape tree (wave butterfly, then waves in order).
        T acc = T(0);
        for (int j = j0 + t; j < j1; j += BS) acc += vals[j] * x(colind[j]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += shfl_xor_w(acc, off, 64);
        if ((t & (kWave - 1)) == 0) prod[t / kWave] = acc;
        __syncthreads();
        if (t == 0) {
            T s = T(0);
#pragma unroll
            for (int w = 0; w < BS / kWave; ++w) s += prod[w];
            epi(r0, s, ca, cb);
        }
        return;
    }
    // 1. stream phase
#pragma unroll
    for (int s = 0; s < NPT; ++s) {
        const int j = j0 + t + s * BS;
        if (j < j1) prod[j - j0] = vals[j] * x(colind[j]);
    }
    __syncthreads();
    // 2. reduce phase: G lanes per row, chosen by the scheduler from the
    //    block's longest row (G = 1 whenever every row has <= 32 entries)
    const int nr = r1 - r0;
    const int lr = t / G;
    const int lane = t & (G - 1);
    T acc = T(0);
    const bool active = lr < nr;
    if (active) {
        const int rs = rowptr[r0 + lr] - j0;
        const int re = rowptr[r0 + lr + 1] - j0;
        for (int k = rs + lane; k < re; k += G) acc += prod[k];
    }
    if (G > 1) {
        for (int off = G >> 1; off > 0; off >>= 1) acc += shfl_xor_w(acc, off, G);
    }
    if (active && lane == 0) epi(r0 + lr, acc, ca, cb);
}

// Generic kernel: per-block partial sums (if Epi::kRed) go to
// epi.partial_a / partial_b[tile].
template <typename T, int BS, int NPT, class Epi, class Gather = GatherX<T>>
__global__ __launch_bounds__(BS) void spmv_rowblock_kernel(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const T* __restrict__ vals, Gather x, const int32_t* __restrict__ blk_row,
    int ntiles, int per, Epi epi)
{
    __shared__ T prod[BS * NPT];
    __shared__ double red[2 * BS / kWave];
    const int tile = xcd_tile(blockIdx.x, ntiles, per);
    if (tile < 0) return;
    if (epi.skip()) return;
    double ca, cb;
    spmv_block_body<T, BS, NPT, Epi, Gather>(tile, rowptr, colind, vals, x, blk_row, epi, prod, ca, cb);
    if constexpr (Epi::kRed == 1) {
        const double s = block_sum<BS>(ca, red);
        if (threadIdx.x == 0) epi.partial_a[tile] = s;
    } else if constexpr (Epi::kRed == 2) {
        block_sum2<BS>(ca, cb, red);
        if (threadIdx.x == 0) {
            epi.partial_a[tile] = ca;
            epi.partial_b[tile] = cb;
        }
    }
}

// Plain y = A x epilogue with no reduction and no skip flag.
template <typename T>
struct EpiY {
    static constexpr int kRed = 0;
    T* __restrict__ y;
    double* partial_a = nullptr;
    double* partial_b = nullptr;
    __device__ __forceinline__ bool skip() const { return false; }
    __device__ __forceinline__ void operator()(int row, T v, double&, double&) const
    {
        y[row] = v;
    }
};

}  // namespace numamd
