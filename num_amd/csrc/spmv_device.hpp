// spmv_device.hpp -- row-tile CSR SpMV kernel for gfx950, shared by the plain
// SpMV entry points and the fused CG kernels.
//
// Work unit: one workgroup of BS threads owns a tile of consecutive rows
// [r0, r1) whose non-zeros [j0, j1) fit the tile's LDS budget (BS*NPT
// entries).  Phases:
//   1. stage: every thread issues NPT non-temporal loads of values and column
//      indices at stride BS (each wave instruction reads 64 consecutive
//      entries: fully coalesced whatever the row lengths) and parks them in
//      LDS.  All loads are issued before any is used (clamped indices, no
//      per-element branches), so each thread keeps 2*NPT loads in flight.
//      The row bounds rowptr[r], rowptr[r+1] are fetched in the same batch.
//   2. walk: G lanes per row (G = 1 unless the tile holds a row with more
//      than 32 entries) walk the row's entries in LDS and gather x[col].  For
//      stencil rows consecutive lanes own consecutive rows, so the gathers for
//      one neighbour offset are contiguous across the wave.  With G = 1 the
//      row is summed in index order from 0.0 with separately rounded products,
//      the arithmetic of the reference's row loop (src/spmv_kernel.hpp:90-97):
//      bit-identical results.
//   3. epilogue: a functor consumes (row, value) -- store y, or the fused CG
//      updates -- and can contribute to a deterministic per-tile partial sum.
// A row longer than the LDS budget gets a workgroup of its own ("long" tile)
// that accumulates straight from global memory at stride BS and reduces with
// a fixed butterfly.
// Tiles are dealt to the 8 XCDs in chunks of consecutive tiles (common.hpp:
// chunk_tile), so each XCD's L2 sees neighbouring x windows.
#pragma once
#include "common.hpp"

namespace numamd {

// Matrix streams: non-temporal (NT) loads keep a matrix that is read once
// per call from evicting x out of L2 / the Infinity Cache; a matrix small
// enough to stay resident in the 256 MB Infinity Cache across calls (the
// FP32 copy inside a CG solve) is loaded normally so it does stay.
template <bool NT, typename T>
__device__ __forceinline__ T ld_mat(const T* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// x accessors.  prepare() runs once per workgroup before any gather (it may
// load scalars the accessor needs).  GatherX: one vector.  GatherSplit: the
// [own | ghost] pair of a rank's slab, columns >= n_own address the ghost
// block (no x_ext copy is ever built).
template <typename T>
struct GatherX {
    const T* __restrict__ x;
    __device__ __forceinline__ void prepare() {}
    __device__ __forceinline__ T operator()(int c) const { return x[c]; }
};

template <typename T>
struct GatherSplit {
    const T* __restrict__ x_own;
    const T* __restrict__ x_ghost;
    int n_own;
    __device__ __forceinline__ void prepare() {}
    __device__ __forceinline__ T operator()(int c) const
    {
        return c < n_own ? x_own[c] : x_ghost[c - n_own];
    }
};

// Body for one tile.  Leaves this thread's two epilogue contributions in
// ca/cb (zero for lanes that finish no row).
template <typename T, int BS, int NPT, bool NT, class Epi, class Gather>
__device__ __forceinline__ void spmv_tile_body(
    int tile, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const T* __restrict__ vals, const Gather& x, const TileEntry* __restrict__ sched,
    const Epi& epi, T* __restrict__ lv, int32_t* __restrict__ lc, double& ca, double& cb)
{
    constexpr int NNZ_MAX = BS * NPT;
    const int t = threadIdx.x;
    const TileEntry e0 = sched[tile];
    const TileEntry e1 = sched[tile + 1];
    const int r0 = e0.row, G = e0.lanes, j0 = e0.nz0;
    const int r1 = e1.row, j1 = e1.nz0;
    const int nz = j1 - j0;
    ca = 0.0;
    cb = 0.0;
    if (nz > NNZ_MAX) {
        // long row: one row, whole workgroup; 4 independent loads in flight
        // per thread, then a fixed-shape tree (wave butterfly, waves in order).
        T acc = T(0);
        int j = j0 + t;
        for (; j + 3 * BS < j1; j += 4 * BS) {
            T v[4];
            int c[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v[u] = ld_mat<NT>(vals + j + u * BS);
                c[u] = ld_mat<NT>(colind + j + u * BS);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) acc += v[u] * x(c[u]);
        }
        for (; j < j1; j += BS) acc += vals[j] * x(colind[j]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if ((t & (kWave - 1)) == 0) lv[t / kWave] = acc;
        __syncthreads();
        if (t == 0) {
            T s = T(0);
#pragma unroll
            for (int w = 0; w < BS / kWave; ++w) s += lv[w];
            epi(r0, s, ca, cb, x);
        }
        return;
    }
    // row bounds for the walk phase, fetched together with the stage loads
    const int nr = r1 - r0;
    const int lr = t / G;
    const int lane = t & (G - 1);
    const bool active = lr < nr;
    int rb = 0, re = 0;
    if (active) {
        rb = rowptr[r0 + lr] - j0;
        re = rowptr[r0 + lr + 1] - j0;
    }
    // 1. stage
    if (nz > 0) {
        const int last = j1 - 1;
        T v[NPT];
        int c[NPT];
#pragma unroll
        for (int s = 0; s < NPT; ++s) {
            const int j = min(j0 + t + s * BS, last);
            v[s] = ld_mat<NT>(vals + j);
            c[s] = ld_mat<NT>(colind + j);
        }
#pragma unroll
        for (int s = 0; s < NPT; ++s) {
            const int k = t + s * BS;
            if (k < nz) {
                lv[k] = v[s];
                lc[k] = c[s];
            }
        }
    }
    __syncthreads();
    // 2. walk: lane visits k = rb + lane, rb + lane + G, ...; 8 gathers in
    //    flight before the (ordered) sums
    T acc = T(0);
    if (active) {
        int k = rb + lane;
        for (; k + 7 * G < re; k += 8 * G) {
            T vv[8], xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                vv[u] = lv[k + u * G];
                xv[u] = x(lc[k + u * G]);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += vv[u] * xv[u];
        }
        T vv[7], xv[7];
#pragma unroll
        for (int u = 0; u < 7; ++u) {
            const int kk = k + u * G;
            const bool ok = kk < re;
            const int ks = ok ? kk : 0;  // slot 0 is always inside the LDS array
            vv[u] = lv[ks];
            const int c = lc[ks];
            xv[u] = x(ok ? c : 0);       // column 0 always exists
        }
#pragma unroll
        for (int u = 0; u < 7; ++u)
            if (k + u * G < re) acc += vv[u] * xv[u];
    }
    if (G > 1) {
        for (int off = G >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
    }
    if (active && lane == 0) epi(r0 + lr, acc, ca, cb, x);
}

// Generic kernel: per-tile partial sums (if Epi::kRed) go to
// epi.partial_a / partial_b[tile].
template <typename T, int BS, int NPT, bool NT, class Epi, class Gather = GatherX<T>>
__global__ __launch_bounds__(BS) void spmv_tile_kernel(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const T* __restrict__ vals, Gather x, const TileEntry* __restrict__ sched, int ntiles,
    int chunk, Epi epi)
{
    __shared__ T lv[BS * NPT];
    __shared__ int32_t lc[BS * NPT];
    __shared__ double red[2 * BS / kWave];
    const int tile = chunk_tile(blockIdx.x, ntiles, chunk);
    if (tile < 0) return;
    if (epi.skip()) return;
    x.prepare();
    double ca, cb;
    spmv_tile_body<T, BS, NPT, NT, Epi, Gather>(tile, rowptr, colind, vals, x, sched, epi, lv, lc,
                                            ca, cb);
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
    template <class Gather>
    __device__ __forceinline__ void operator()(int row, T v, double&, double&, const Gather&) const
    {
        y[row] = v;
    }
};

}  // namespace numamd
