// spmv.hip -- y = A x (FP64 / FP32), row-range SpMV, multi-RHS SpMV, halo
// pack/scatter and the host round-trip helpers.
#include <algorithm>
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
// non-zeros from 0.0, one rounded product and one rounded sum per entry).
//
// Two kernels, both memory-bound on X and Y (8k bytes each per row):
// * k < 64, batch_spmv_kernel: a group of G lanes owns a row; lane l covers
//   the V consecutive right-hand sides b = l*V + m*G*V (m = 0, 1, ...), so one
//   X row is read as G*V*8 contiguous bytes (V = 2: 16-byte loads).  The
//   row's entries are taken U at a time: U column/value loads (the same
//   address for the G lanes of a group, one request), then U independent X
//   loads, all issued before the first sum.
// * k >= 64, batch_wave_kernel: a wave per row, R rows per wave, the rows'
//   entries held in registers and the X loads of row r + 1 issued before row
//   r is summed (below).
// In both the sums run in entry order.
//
// X rows are re-read by every row that references them (7 times for a 3-D
// stencil), and at k = 128 one X row is 1 KB, so the reuse must come from the
// XCD's own L2: row blocks are dealt to the 8 XCDs in chunks of consecutive
// blocks (chunk_tile, batch_chunk), so the rows a block shares X rows with
// (the neighbouring rows, the next grid line and, with plane-aligned chunks,
// the next plane's rows 8 chunks on) mostly run on the same XCD: an XCD
// then holds its own eighth of three planes' X rows (3.75 MB at k = 128 on
// 100^3, within its 4 MB L2), and fetches 1.27x the algorithmic read bytes
// instead of 2.99x with 2048-row chunks.  Y is written once and never read
// here: non-temporal stores.
// ---------------------------------------------------------------------------
constexpr int kBatchBS = 256;
constexpr int kBatchU = 8;
constexpr int kBatchChunkRows = 2048;

typedef double d2v __attribute__((ext_vector_type(2)));

template <int V> struct BatchVec;
template <> struct BatchVec<1> {
    typedef double T;
    static __device__ __forceinline__ T load(const double* p) { return *p; }
    static __device__ __forceinline__ double get(const T& v, int) { return v; }
    static __device__ __forceinline__ void store_nt(double* p, const double* a)
    {
        __builtin_nontemporal_store(a[0], p);
    }
};
template <> struct BatchVec<2> {
    typedef d2v T;
    static __device__ __forceinline__ T load(const double* p) { return *(const d2v*)p; }
    static __device__ __forceinline__ double get(const T& v, int i) { return v[i]; }
    static __device__ __forceinline__ void store_nt(double* p, const double* a)
    {
        const d2v o = {a[0], a[1]};
        __builtin_nontemporal_store(o, (d2v*)p);
    }
};

// One row's Y(row, b..) for the right-hand sides b = lane*V + m*G*V, entries
// read straight from colind / vals, kBatchU at a time.
template <int G, int V>
__device__ __forceinline__ void batch_row_direct(const int32_t* __restrict__ colind,
                                                 const double* __restrict__ vals,
                                                 const double* __restrict__ X,
                                                 double* __restrict__ Y, int row, int js, int je,
                                                 int k, int lane)
{
    typedef BatchVec<V> Vec;
    for (int b = lane * V; b < k; b += G * V) {
        double acc[V];
#pragma unroll
        for (int i = 0; i < V; ++i) acc[i] = 0.0;
        for (int j = js; j < je; j += kBatchU) {
            int c[kBatchU];
            double v[kBatchU];
            typename Vec::T xv[kBatchU];
#pragma unroll
            for (int u = 0; u < kBatchU; ++u) {
                if (j + u < je) {
                    c[u] = colind[j + u];
                    v[u] = vals[j + u];
                }
            }
#pragma unroll
            for (int u = 0; u < kBatchU; ++u)
                if (j + u < je) xv[u] = Vec::load(X + (int64_t)c[u] * k + b);
#pragma unroll
            for (int u = 0; u < kBatchU; ++u) {
                if (j + u < je) {
#pragma unroll
                    for (int i = 0; i < V; ++i) acc[i] = acc[i] + v[u] * Vec::get(xv[u], i);
                }
            }
        }
        Vec::store_nt(Y + (int64_t)row * k + b, acc);
    }
}

template <int G, int V>
__global__ __launch_bounds__(kBatchBS) void batch_spmv_kernel(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const double* __restrict__ vals, const double* __restrict__ X, double* __restrict__ Y,
    int nrows, int k, int nblocks, int chunk)
{
    constexpr int RPB = kBatchBS / G;  // rows per workgroup
    const int t = chunk_tile(blockIdx.x, nblocks, chunk);
    if (t < 0) return;
    const int row = t * RPB + (int)threadIdx.x / G;
    if (row >= nrows) return;
    batch_row_direct<G, V>(colind, vals, X, Y, row, rowptr[row], rowptr[row + 1], k,
                           (int)threadIdx.x % G);
}

__device__ __forceinline__ int readlane_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double readlane_d(double v, int l)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The pipelined walk of batch_wave_kernel.  FULL: all R rows exist and every
// lane has a right-hand side (k a multiple of 64 * V), so no store is guarded;
// a guarded store lets the compiler sink the loads feeding it into the guard,
// behind the next row's loads, which undoes the pipelining.
template <int V, int R, bool FULL>
__device__ __forceinline__ void batch_wave_rows(const double* __restrict__ X,
                                                double* __restrict__ Y, int rpl, int e0,
                                                int nr, int cl, double vl, int r0, int k,
                                                int lane)
{
    typedef BatchVec<V> Vec;
    for (int cb = 0; cb < k; cb += kWave * V) {
        const int b = cb + lane * V;
        const bool act = FULL || b < k;
        const int bx = act ? b : 0;  // idle lanes load an in-bounds column
        typename Vec::T xb[2][kBatchU];
#pragma unroll
        for (int r = 0; r <= R; ++r) {
            if (r < R) {  // issue row r (clamped to the last row) into buffer r & 1
                const int rr = FULL ? r : min(r, nr - 1);
                const int s = readlane_i(rpl, rr) - e0;
                const int len = readlane_i(rpl, rr + 1) - e0 - s;
#pragma unroll
                for (int u = 0; u < kBatchU; ++u) {
                    const int e = u < len ? s + u : max(s + len - 1, 0);
                    const int c = readlane_i(cl, e);
                    xb[r & 1][u] = Vec::load(X + (int64_t)c * k + bx);
                }
            }
            if (r > 0 && (FULL || r - 1 < nr)) {  // sum row r - 1
                const int q = r - 1;
                const int s = readlane_i(rpl, q) - e0;
                const int len = readlane_i(rpl, q + 1) - e0 - s;
                double acc[V];
#pragma unroll
                for (int i = 0; i < V; ++i) acc[i] = 0.0;
                // Branch-free: every slot's sum is formed and kept only for
                // u < len, so no load is sunk into a branch and waited for.
#pragma unroll
                for (int u = 0; u < kBatchU; ++u) {
                    const double v = readlane_d(vl, min(s + u, kWave - 1));
#pragma unroll
                    for (int i = 0; i < V; ++i) {
                        const double nxt = acc[i] + v * Vec::get(xb[q & 1][u], i);
                        acc[i] = u < len ? nxt : acc[i];
                    }
                }
                if (act) Vec::store_nt(Y + (int64_t)(r0 + q) * k + b, acc);
            }
        }
    }
}

// Wide k, one wave (64 lanes) per row: each wave takes R consecutive rows.
// Their bounds are one load (lane i holds rowptr[r0 + i]); when the R rows
// hold at most 64 entries and none more than kBatchU, their columns and
// values are one more load (lane e holds entry e) and every X load after that
// takes its column from a register (readlane), with no index load in between.
// The X loads of row r + 1 are issued before row r is summed (two register
// buffers), so each wave keeps a row's X rows in flight while it adds.  Every
// load slot is filled (a short row repeats its last column, the last row of
// the matrix is loaded again), so the compiler's wait counts stay exact.
// Rows outside that shape take batch_row_direct.
template <int V, int R>
__global__ __launch_bounds__(kBatchBS) void batch_wave_kernel(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const double* __restrict__ vals, const double* __restrict__ X, double* __restrict__ Y,
    int nrows, int k, int nblocks, int chunk)
{
    static_assert(R * kBatchU <= kWave, "entries of R rows must fit one wave");
    const int t = chunk_tile(blockIdx.x, nblocks, chunk);
    if (t < 0) return;
    const int lane = (int)threadIdx.x & (kWave - 1);
    const int r0 = (t * (kBatchBS / kWave) + (int)threadIdx.x / kWave) * R;
    if (r0 >= nrows) return;
    const int nr = min(R, nrows - r0);
    const int rpl = rowptr[r0 + min(lane, nr)];
    const int e0 = readlane_i(rpl, 0);
    int maxlen = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (r < nr) maxlen = max(maxlen, readlane_i(rpl, r + 1) - readlane_i(rpl, r));
    if (maxlen > kBatchU) {
        for (int r = 0; r < nr; ++r)
            batch_row_direct<kWave, V>(colind, vals, X, Y, r0 + r, readlane_i(rpl, r),
                                       readlane_i(rpl, r + 1), k, lane);
        return;
    }
    // nr <= R rows of <= kBatchU entries: at most R * kBatchU <= 64 entries.
    const int ne = readlane_i(rpl, nr) - e0;
    int cl = 0;
    double vl = 0.0;
    if (lane < ne) {
        cl = colind[e0 + lane];
        vl = vals[e0 + lane];
    }
    if (nr == R && k % (kWave * V) == 0)
        batch_wave_rows<V, R, true>(X, Y, rpl, e0, nr, cl, vl, r0, k, lane);
    else
        batch_wave_rows<V, R, false>(X, Y, rpl, e0, nr, cl, vl, r0, k, lane);
}

// XCD chunk (in row blocks of rpb rows) of the multi-RHS kernels: aligned to
// the matrix's long-range offset as the SpMV's (csr.hpp: plane_chunk), up to
// 16384 rows; else about kBatchChunkRows rows.  100^3: k = 8 54.6 -> 48.6
// us, k = 128 691 -> 646 us against 2048-row chunks.
static int batch_chunk(const num_csr* A, int rpb, int nblocks)
{
    const int pc = plane_chunk(A, rpb, 16384 / rpb);
    if (pc > 0) return pc;
    return nblocks >= 8 * 64 ? std::max(1, kBatchChunkRows / rpb) : 1;
}

template <int G, int V>
static void launch_batch(const num_csr* A, const double* X, double* Y, int k,
                         hipStream_t stream)
{
    constexpr int RPB = kBatchBS / G;
    const int nblocks = (A->nrows + RPB - 1) / RPB;
    const int chunk = batch_chunk(A, RPB, nblocks);
    const int grid = chunk_grid(nblocks, chunk);
    batch_spmv_kernel<G, V><<<grid, kBatchBS, 0, stream>>>(A->rowptr, A->colind, A->vals, X,
                                                           Y, A->nrows, k, nblocks, chunk);
}

template <int V, int R>
static void launch_batch_wave(const num_csr* A, const double* X, double* Y, int k,
                              hipStream_t stream)
{
    constexpr int RPB = (kBatchBS / kWave) * R;
    const int nblocks = (A->nrows + RPB - 1) / RPB;
    const int chunk = batch_chunk(A, RPB, nblocks);
    const int grid = chunk_grid(nblocks, chunk);
    batch_wave_kernel<V, R><<<grid, kBatchBS, 0, stream>>>(A->rowptr, A->colind, A->vals, X,
                                                           Y, A->nrows, k, nblocks, chunk);
}

// k >= kBatchWaveK, or k > kBatchWaveK1 with 8-byte X loads (odd k or
// unaligned X, Y): a wave per row (16-byte X loads when k > 64 is even and
// X, Y are 16-byte aligned).  Smaller k: G lanes per row, G the smallest
// power of two with G*V >= k (V = 2 when k is even and aligned).  100^3:
// k = 33 589 -> 358 us and k = 63 657 -> 418 us on the wave kernel, while
// even k = 40-56 stay 3-4 % faster on 32 lanes of 16-byte loads.
constexpr int kBatchWaveK = 64;
constexpr int kBatchWaveK1 = 32;
// Rows per wave: 2 when the wave's X loads cover whole 128-column passes
// (k a multiple of 128 with 16-byte loads: 100^3, k = 128 647 -> 617 us,
// k = 256 1349 -> 1313 us), else 4 (k = 64 403 against 414 us with 2, k = 96
// 498 against 529 us).
constexpr int kBatchWaveRows = 4;
constexpr int kBatchWaveRowsFull = 2;
static void dispatch_batch(const num_csr* A, const double* X, double* Y, int k,
                           hipStream_t st)
{
    const bool v2 = (k % 2 == 0) && ((uintptr_t)X % 16 == 0) && ((uintptr_t)Y % 16 == 0);
    if (k >= kBatchWaveK || (!v2 && k > kBatchWaveK1)) {
        if (v2 && k > kWave && k % (2 * kWave) == 0)
            launch_batch_wave<2, kBatchWaveRowsFull>(A, X, Y, k, st);
        else if (v2 && k > kWave) launch_batch_wave<2, kBatchWaveRows>(A, X, Y, k, st);
        else launch_batch_wave<1, kBatchWaveRows>(A, X, Y, k, st);
        return;
    }
    const int w = v2 ? k / 2 : k;
    if (v2) {
        if (w <= 1) launch_batch<1, 2>(A, X, Y, k, st);
        else if (w <= 2) launch_batch<2, 2>(A, X, Y, k, st);
        else if (w <= 4) launch_batch<4, 2>(A, X, Y, k, st);
        else if (w <= 8) launch_batch<8, 2>(A, X, Y, k, st);
        else if (w <= 16) launch_batch<16, 2>(A, X, Y, k, st);
        else launch_batch<32, 2>(A, X, Y, k, st);
    } else {
        if (w <= 2) launch_batch<2, 1>(A, X, Y, k, st);
        else if (w <= 4) launch_batch<4, 1>(A, X, Y, k, st);
        else if (w <= 8) launch_batch<8, 1>(A, X, Y, k, st);
        else if (w <= 16) launch_batch<16, 1>(A, X, Y, k, st);
        else launch_batch<32, 1>(A, X, Y, k, st);  // w = k <= kBatchWaveK1
    }
}

// ---------------------------------------------------------------------------
// Column-partitioned SpMV (num_csr_attach_xcd), for matrices whose x gathers
// are spread far wider than an XCD's 4 MB L2 (random columns: the Zipf
// power-law rows).  x is cut into column blocks of up to 4 MB, and block j is
// served by XCD j % 8 only, so each XCD's L2 holds the one x block it gathers
// from instead of filling a 128-byte line per random 8-byte gather.  Pass 1
// (xcd_seg_kernel) walks the segment matrix (a row per (block, row) pair,
// block-major) in tiles of up to kSegRows segments and kSegNnz entries taken
// from its own XCD's group (workgroup b runs on XCD b % 8), writing one
// partial sum per segment.  Its gathers are dense: thread t gathers x for the
// entries it stages, whatever the segment lengths (3 entries on average, 62 %
// single; a row-per-thread walk left most lanes idle), and the segments are
// summed from the products in LDS.  Pass 2 sums each row's partials in block
// order.  A row whose entries all fall in one block keeps the row loop's sum
// bit for bit (0.0 + s = s); other rows are the sum of their blocks' partials.
// Zipf n = 4M: 290 us with the row-tile kernel as pass 1, 225 us now (pass 1
// 186 us, bound by L2 hit latency on ~24M random gathers: the L1 waits on L2
// data 72 % of its cycles; pass 2 39 us).
// ---------------------------------------------------------------------------
struct XcdGroups {
    int32_t off[kNumXcd + 1];
};

// Pass 1 over one tile of segments [s0, s1) holding entries [j0, j1):
// thread t stages entries j0 + t + u*BS (u < kSegNpt) and gathers their x at
// once (every lane busy, however short the segments), products to LDS; then
// thread t sums segments t + r*BS (r < kSegRpt) of <= kShortRow entries in
// index order from 0.0 (bit-identical to the row loop), and a longer segment
// of the tile goes to a wave, whose 64 lanes stride over its products (4
// groups of 64 in order, then the tail) before a fixed butterfly.  A
// segment longer than a tile holds (kSegNnz entries) is a tile of its own,
// summed by the whole workgroup straight from memory.
template <bool NT>
__global__ __launch_bounds__(kSpmvBS) void xcd_seg_kernel(
    const int32_t* __restrict__ segptr, const int32_t* __restrict__ colind,
    const double* __restrict__ vals, const double* __restrict__ x,
    const TileEntry* __restrict__ sched, const int32_t* __restrict__ xcd_tiles, XcdGroups g,
    double* __restrict__ part)
{
    constexpr int BS = kSpmvBS, NPT = kSegNpt, RPT = kSegRpt, NW = BS / kWave;
    __shared__ double lp[BS * NPT];
    __shared__ unsigned long long wmask[RPT * NW];
    const int k = blockIdx.x % kNumXcd;
    const int q = blockIdx.x / kNumXcd;
    if (q >= g.off[k + 1] - g.off[k]) return;
    const int tile = xcd_tiles[g.off[k] + q];
    const TileEntry e0 = sched[tile], e1 = sched[tile + 1];
    const int t = threadIdx.x, w = t / kWave, l = t & (kWave - 1);
    const int s0 = e0.row, j0 = e0.nz0;
    const int ns = e1.row - s0, nz = e1.nz0 - j0;
    if (e0.wide == kWalkLong) {
        // one long segment: 8 loads in flight per thread, fixed tree
        double acc = 0.0;
        const int j1 = j0 + nz, last = j1 - 1;
        for (int j = j0 + t; j < j1; j += 8 * BS) {
            double v[8], xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int jj = min(j + u * BS, last);
                v[u] = ld_mat<NT>(vals + jj);
                xv[u] = x[ld_mat<NT>(colind + jj)];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j + u * BS < j1) acc += v[u] * xv[u];
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (l == 0) lp[w] = acc;
        __syncthreads();
        if (t == 0) {
            double sum = 0.0;
#pragma unroll
            for (int i = 0; i < NW; ++i) sum += lp[i];
            __builtin_nontemporal_store(sum, part + s0);
        }
        return;
    }
    // segment bounds (branch-free: a thread past the tile reads the end twice)
    int sb[RPT], se[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int i = t + r * BS;
        const int li = min(i, ns);
        sb[r] = segptr[s0 + li];
        se[r] = segptr[s0 + li + (i < ns ? 1 : 0)];
    }
    if (nz > 0) {
        const int last = j0 + nz - 1;
        double v[NPT], xv[NPT];
        int c[NPT];
#pragma unroll
        for (int u = 0; u < NPT; ++u) {
            const int j = min(j0 + t + u * BS, last);
            v[u] = ld_mat<NT>(vals + j);
            c[u] = ld_mat<NT>(colind + j);
        }
        // clamped slots repeat the tile's last entry: one line, coalesced
#pragma unroll
        for (int u = 0; u < NPT; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < NPT; ++u) {
            const int kk = t + u * BS;
            if (kk < nz) lp[kk] = v[u] * xv[u];
        }
    }
    __syncthreads();
    const bool wide = e0.wide != kWalkShort;  // uniform over the workgroup
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int i = t + r * BS;
        const int b = sb[r] - j0, e = se[r] - j0;
        if (i < ns && e - b <= kShortRow) {
            double acc = 0.0;
            for (int kk = b; kk < e; ++kk) acc += lp[kk];
            __builtin_nontemporal_store(acc, part + s0 + i);
        }
        if (wide) {
            const unsigned long long bal = __ballot(i < ns && e - b > kShortRow);
            if (l == 0) wmask[r * NW + w] = bal;
        }
    }
    if (!wide) return;
    __syncthreads();
    // the m-th longer segment (in segment order) goes to wave m % NW
    int m = 0;
    for (int i = 0; i < RPT * NW; ++i) {
        unsigned long long bits = wmask[i];
        while (bits) {
            const int bpos = __builtin_ctzll(bits);
            bits &= bits - 1;
            if ((m++ % NW) != w) continue;
            // wmask[r * NW + v] covers segments r*BS + v*64 + bit
            const int seg = (i / NW) * BS + (i % NW) * kWave + bpos;
            const int mb = segptr[s0 + seg] - j0, me = segptr[s0 + seg + 1] - j0;
            double acc = 0.0;
            int kk = mb + l;
            for (; kk + 3 * kWave < me; kk += 4 * kWave) {
#pragma unroll
                for (int u = 0; u < 4; ++u) acc += lp[kk + u * kWave];
            }
            for (; kk < me; kk += kWave) acc += lp[kk];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
            if (l == 0) __builtin_nontemporal_store(acc, part + s0 + seg);
        }
    }
}

// Pass 2: y[r] = the row's segment partials summed in block order from 0.0.
// A row has at most one segment per column block; up to 8 segment ids, then
// their partials, are loaded at once (lanes past the row's segments masked
// off) instead of one dependent pair per segment.
__global__ void xcd_reduce_kernel(int32_t n, const int32_t* __restrict__ row_seg_ptr,
                                  const int32_t* __restrict__ row_seg_idx,
                                  const double* __restrict__ part, double* __restrict__ y)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int b = row_seg_ptr[r], e = row_seg_ptr[r + 1];
    double acc = 0.0;
    for (int i = b; i < e; i += 8) {
        int idx[8];
        double pv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i + u < e) idx[u] = row_seg_idx[i + u];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i + u < e) pv[u] = part[idx[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (i + u < e) acc += pv[u];
    }
    __builtin_nontemporal_store(acc, y + r);
}

static void launch_xcd(const num_csr* A, const double* x, double* part, hipStream_t st)
{
    const num_csr* S = A->xcd_seg;
    XcdGroups g;
    int32_t most = 0;
    for (int k = 0; k <= kNumXcd; ++k) g.off[k] = A->xcd_off[k];
    for (int k = 0; k < kNumXcd; ++k) most = std::max(most, g.off[k + 1] - g.off[k]);
    if (most == 0) return;
    const TileEntry* tiles = static_cast<const TileEntry*>(A->xcd_sched.d_tiles);
    xcd_seg_kernel<true><<<most * kNumXcd, kSpmvBS, 0, st>>>(S->rowptr, S->colind, S->vals, x,
                                                             tiles, A->xcd_tiles, g, part);
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

int num_spmv_xcd(const num_csr* A, const double* d_x, double* d_y, double* d_work, void* stream)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (!A->xcd_seg) return fail(NUM_EINVAL, "no column-partitioned layout attached");
    if (!d_x || !d_y || (A->xcd_seg->nrows > 0 && !d_work))
        return fail(NUM_EINVAL, "x, y or work is NULL");
    int rc = check_device(A);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    launch_xcd(A, d_x, d_work, st);
    NUM_LAUNCH_CHECK();
    if (A->nrows > 0) {
        xcd_reduce_kernel<<<(A->nrows + 255) / 256, 256, 0, st>>>(
            A->nrows, A->xcd_row_seg_ptr, A->xcd_row_seg_idx, d_work, d_y);
        NUM_LAUNCH_CHECK();
    }
    return NUM_OK;
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
    dispatch_batch(A, d_X, d_Y, k, st);
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
