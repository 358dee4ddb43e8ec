// common.hpp -- error plumbing, launch constants and wave/block reductions
// shared by every translation unit of libnum_amd.so (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/num_amd.h"

namespace numamd {

// ---------------------------------------------------------------- errors ---
// Thread-local last-error text (num_last_error()).  Return codes follow
// include/num_amd.h: NUM_EINVAL maps to ValueError in Python, the others to
// RuntimeError.
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);
const char* last_error();

#define NUM_HIP(expr)                                                               \
    do {                                                                            \
        hipError_t e__ = (expr);                                                    \
        if (e__ != hipSuccess)                                                      \
            return ::numamd::fail(NUM_EHIP, "%s failed: %s (%s:%d)", #expr,         \
                                  hipGetErrorString(e__), __FILE__, __LINE__);      \
    } while (0)

#define NUM_LAUNCH_CHECK()                                                          \
    do {                                                                            \
        hipError_t e__ = hipGetLastError();                                         \
        if (e__ != hipSuccess)                                                      \
            return ::numamd::fail(NUM_EHIP, "kernel launch failed: %s (%s:%d)",     \
                                  hipGetErrorString(e__), __FILE__, __LINE__);      \
    } while (0)

// ------------------------------------------------------------ constants ---
constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kNumXcd = 8;         // MI355X: 8 XCDs, workgroups dealt round-robin
constexpr int kRedBlock = 256;     // block size of elementwise / reduction kernels
constexpr int kMaxRedBlocks = 2048;  // cap on per-launch partial sums

// Workgroup index -> tile.  Workgroup b runs on XCD b % 8 (round-robin
// dispatch); tiles are cut into chunks of `chunk` consecutive tiles and chunk
// ch goes to XCD ch % 8, so every XCD (with its private L2) walks runs of
// neighbouring rows while all XCDs stay close together in the matrix.
// chunk = 1 is the plain hardware order.  Returns -1 for padding workgroups.
__device__ __forceinline__ int chunk_tile(int b, int ntiles, int chunk)
{
    const int k = b % kNumXcd;
    const int q = b / kNumXcd;  // q-th workgroup this XCD receives
    const int t = ((q / chunk) * kNumXcd + k) * chunk + (q % chunk);
    return t < ntiles ? t : -1;
}
// Chunk size used for a launch of `ntiles` tiles and the grid covering it.
inline int tile_chunk(int ntiles) { return ntiles >= 16384 ? 64 : ntiles >= 4096 ? 16 : 1; }
inline int chunk_grid(int ntiles, int chunk)
{
    const int nchunks = (ntiles + chunk - 1) / chunk;
    return ((nchunks + kNumXcd - 1) / kNumXcd) * kNumXcd * chunk;
}

// Row lengths (one byte per row) stored after the nnz delta codes and their
// 8 padding codes in the buffer of num_csr_attach_d16.
constexpr int kD16Pad = 8;
__host__ __device__ inline uint8_t* d16_lens(uint16_t* codes, int64_t nnz)
{
    return reinterpret_cast<uint8_t*>(codes + nnz + kD16Pad);
}
__host__ __device__ inline const uint8_t* d16_lens(const uint16_t* codes, int64_t nnz)
{
    return reinterpret_cast<const uint8_t*>(codes + nnz + kD16Pad);
}

// Windowed sliced layout (spmv_sell_run).  Column windows of a tile: at
// most kXwWin contiguous column ranges holding every column the tile's
// entries read, kXwSlots columns in all.  Per tile the record holds the
// windows' first columns (lo[kXwWin]), their cumulative ends in LDS slots
// (end[kXwWin]; an unused window has lo 0 and the previous end), then the
// tile's base in the sliced arrays (in 256-entry slices) and its width
// (longest row, 1 to kSellW entries).  A tile of width w takes w slices of
// each array, with no padding between tiles:
//   FP64 values: entry k of row i at (base + k) * 256 + i;
//   slots (uint16) and FP32 values: entry k of row i in packed groups,
//     slots four per 8-byte load (then two per 4-byte, then one), FP32
//     values two per 8-byte load (then one) -- sell_pos gives the position.
constexpr int kXwWin = 4;
constexpr int kXwSlots = 1152;
constexpr int kSellW = 8;
constexpr int kXwBase = 2 * kXwWin;
constexpr int kXwWidth = 2 * kXwWin + 1;
constexpr int kXwRec = 2 * kXwWin + 2;
// Row patterns of the windowed layout: a row's window slots minus its lane
// in the tile (plus kXwPatBias, as uint16; 0 past the row's length) are the
// same for every row of a tile that reads the same column offsets (a
// stencil's interior rows), so a tile keeps a table of at most kXwPat
// distinct patterns, two uint64 of four 16-bit slots each (tiles with equal
// tables share one), and a 4-bit id per row replaces the row's 2-byte slots
// and its length byte: the length sits in bits 12-15 of the first slot
// (biased slots stay below 4096).
constexpr int kXwPat = 16;
constexpr int kXwPatBias = 256;
static_assert(kXwSlots + kXwPatBias <= 4096 && kSellW < 16, "length field of a row pattern");
static_assert(kXwPat <= 16, "row pattern ids are 4 bits");
// Row patterns of the int16 column offsets (num_csr_attach_rel16): a row's
// offsets column - row (at most 8, zeros past its length, as uint16) and its
// length per pair of schedule tiles, at most kRelPat distinct patterns per
// pair (the table: 2 * kRelPat pattern words, then one word of 4-bit
// lengths), one id byte per row (the FP32 pair kernel reads them instead of
// 2 bytes per entry and rowptr).
constexpr int kRelPat = 16;
constexpr int kRelPatW = 8;
// Words of a row-pattern table of NPW pattern words (two per pattern): the
// patterns, then their lengths, 4 bits each, 16 to a word.
template <int NPW>
constexpr int relpat_len_words() { return (NPW / 2 + 15) / 16; }
template <int NPW>
constexpr int relpat_words() { return NPW + relpat_len_words<NPW>(); }

// Position of entry k of lane i (of bs lanes) in a tile's packed region of
// w slices, elements of which `pack` (4 for slots, 2 for FP32 values) share
// one load: full groups of `pack`, then groups of halves down to single
// elements.  Positions count elements from the tile's region start.
__host__ __device__ inline int64_t sell_pos(int k, int i, int w, int pack, int bs)
{
    int64_t base = 0;
    for (int g = pack; g >= 1; g >>= 1) {
        const int n = (w / g) * g;  // slots taken by groups of size g
        if (k < n) return base + (int64_t)(k / g) * bs * g + (int64_t)i * g + (k % g);
        base += (int64_t)n * bs;
        k -= n;
        w -= n;
    }
    return base;  // not reached for k < w
}

// Schedule entry of one SpMV row tile; the entry after the last tile holds
// {r_end, 0, rowptr[r_end], 0}.
struct TileEntry {
    int32_t row;    // first row
    int32_t wide;   // walk mode: kWalkShort, kWalkMedium or kWalkLong
    int32_t nz0;    // rowptr[row]
    int32_t base;   // delta-coded index only: the tile's base column, else 0
};
// Row classes.  A row of up to kShortRow entries is walked by one thread; a
// longer row of a staged tile by a whole wave; a row of more than kLongRow
// entries gets a tile (workgroup) of its own and is read straight from
// global memory.
constexpr int kShortRow = 32;
constexpr int kLongRow = 256;
enum : int32_t { kWalkShort = 0, kWalkMedium = 1, kWalkLong = 2 };

// ------------------------------------------------------------ reductions ---
// Deterministic butterfly sum over a whole wave64: every lane ends with the
// same value, combined in a fixed pattern independent of timing.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Deterministic block sum; result valid in thread 0.  `red` must hold
// blockDim.x / 64 doubles.  Ends with a barrier so `red` can be reused.
template <int BS>
__device__ __forceinline__ double block_sum(double v, double* red)
{
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < BS / kWave; ++i) s += red[i];
    }
    __syncthreads();
    return s;
}

// Sum of n per-tile partials by one workgroup of BS threads in a fixed order:
// thread t adds p[t], p[t + BS], ... with U independent loads in flight (a
// batch; predicated past n, so up to U * BS partials take one memory round
// trip), then the block tree.  Result valid in thread 0.  `red` holds BS / 64
// doubles.  The per-thread order does not depend on U.
template <int BS, int U = 8>
__device__ __forceinline__ double sum_partials_block(const double* __restrict__ p, int n,
                                                     double* red)
{
    double acc = 0.0;
    for (int base = threadIdx.x; base < n; base += U * BS) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = base + u * BS;
            v[u] = i < n ? p[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (base + u * BS < n) acc += v[u];
    }
    return block_sum<BS>(acc, red);
}

// Block sum returned to every thread (each adds the wave totals in the same
// order, so all hold the same bits; thread 0 matches block_sum).  No trailing
// barrier: `red` must not be written again in this launch.
template <int BS>
__device__ __forceinline__ double block_total_once(double v, double* red)
{
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < BS / kWave; ++i) s += red[i];
    return s;
}

// ---- finalize by the last workgroup of a launch ---------------------------
// Each workgroup publishes its partial sum with an agent-scope store (written
// through to the level all XCDs share, so none of the workgroup's other
// output has to be written back early), waits for that store to complete and
// draws a ticket with an agent-scope atomic.  The workgroup that draws the
// last ticket re-arms the counter for the next launch, reads every partial
// with agent-scope loads and finalizes.  No workgroup ever waits for another,
// so nothing spins.  Every workgroup must read whatever the finalizer writes
// before it draws its ticket.
__device__ __forceinline__ void publish(double* slot, double v)
{
    __hip_atomic_store(slot, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by every thread after thread 0 has published (if it publishes);
// returns, to every thread, whether this workgroup drew the last of `total`
// tickets (default: one per workgroup of the launch).  `flag` is a
// __shared__ int.
__device__ __forceinline__ bool take_ticket(uint32_t* counter, int* flag, uint32_t total = 0)
{
    if (threadIdx.x == 0) {
        __builtin_amdgcn_s_waitcnt(0);  // the published store has completed
        const uint32_t t =
            __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int last = (t == (total ? total : gridDim.x) - 1) ? 1 : 0;
        if (last) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *flag = last;
    }
    __syncthreads();
    return *flag != 0;
}

// sum_partials_block over published partials (agent-scope loads), same order.
template <int BS>
__device__ __forceinline__ double sum_published_block(const double* p, int n, double* red)
{
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += BS)
        acc += __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return block_sum<BS>(acc, red);
}

// Two independent block sums in one pass (r.r and b.b of the CG set-up).
template <int BS>
__device__ __forceinline__ void block_sum2(double& a, double& b, double* red)
{
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) { red[w] = a; red[BS / kWave + w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0;
#pragma unroll
        for (int i = 0; i < BS / kWave; ++i) { sa += red[i]; sb += red[BS / kWave + i]; }
        a = sa;
        b = sb;
    }
    __syncthreads();
}

}  // namespace numamd
