// csr.hpp -- the CSR handle behind num_csr* and its row-block schedule.
#pragma once
#include <atomic>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"

namespace numamd {

constexpr int kSpmvBS = 256;  // threads per row-block workgroup

// LDS entries per thread the tile kernel is instantiated for.
constexpr int kNptChoices[] = {5, 7, 12, 16};
constexpr int kNptMax = 16;

// Pass-1 tiles of the column-partitioned SpMV: up to kSegRpt segments and
// kSegNpt entries per thread.
constexpr int kSegNpt = 8;
constexpr int kSegRpt = 4;
constexpr int kSegRows = kSegRpt * kSpmvBS;
constexpr int kSegNnz = kSegNpt * kSpmvBS;

// Row-tile schedule over rows [row_begin, row_end): consecutive rows packed
// greedily into tiles of at most rows_per_block rows and nnz_per_block
// non-zeros; a row with more than nnz_per_block non-zeros is a tile alone.
struct Schedule {
    int32_t row_begin = 0, row_end = 0;
    int32_t rows_per_block = 0, nnz_per_block = 0;
    int32_t npt = 0;             // LDS entries per thread (kNptChoices)
    int32_t nblocks = 0;         // tiles
    int32_t long_rows = 0;       // single-row tiles longer than nnz_per_block
    int32_t chunk = 1;           // XCD chunk size for chunk_tile
    int32_t nt64 = 1, nt32 = 1;  // non-temporal matrix loads (FP64 / FP32 values)
    void* d_tiles = nullptr;     // device TileEntry[nblocks + 1]
};

}  // namespace numamd

struct num_csr {
    int device = 0;
    int32_t nrows = 0, ncols = 0;
    int64_t nnz = 0;
    const int32_t* rowptr = nullptr;
    const int32_t* colind = nullptr;
    const double* vals = nullptr;
    const float* vals32 = nullptr;
    const int16_t* rel16 = nullptr;     // column - row as int16 (FP32 path), if attached
    // Row-pattern form of rel16 per pair of schedule tiles (kRelPat), owned:
    // relpat_words<2 * kRelPat>() uint64 per pair (patterns, then lengths),
    // then one pattern id per row.
    uint64_t* rel_pat = nullptr;
    uint8_t* rel_pid = nullptr;
    // Delta-coded column ids (FP64 path on the whole-matrix schedule), if
    // attached: see num_csr_attach_d16.  d16_buf is the caller's buffer
    // (nnz + 8 codes, then one length byte per row: d16_lens), kept so a
    // new schedule can re-code into it.
    const uint16_t* d16 = nullptr;
    uint16_t* d16_buf = nullptr;
    // Windowed sliced layout (num_csr_set_windows), owned, built over
    // A->sched when every tile fits: per-tile records (kXwRec int32: window
    // first columns, cumulative slot ends, slice base, width), the FP64
    // values and window slots in sliced order, one length byte per row
    // (sell_codes and sell_lens point into the sell_vals allocation).
    int32_t* xw = nullptr;
    double* sell_vals = nullptr;
    float* sell_vals32 = nullptr;       // when vals32 is attached
    uint16_t* sell_codes = nullptr;
    uint8_t* sell_lens = nullptr;
    int32_t sell_w = 0;                 // slices per tile (tile b starts at b * sell_w)
    // Row-pattern form of the window slots (when every tile holds at most
    // kXwPat distinct patterns), owned, one allocation: the distinct tables
    // (2 * kXwPat uint64 each; tiles with equal tables share one), each
    // tile's table index (sell_tix), and each row's 4-bit pattern id,
    // kSpmvBS / 2 bytes per tile (sell_pid: lane i of tile b in byte
    // b * kSpmvBS / 2 + i / 2, bits 4 * (i & 1)).
    uint64_t* sell_pat = nullptr;
    int32_t* sell_tix = nullptr;
    uint8_t* sell_pid = nullptr;
    int32_t pat_mode = -1;              // row patterns: 0 off (the slots), else when they fit
    int32_t xw_mode = -1;               // 1 on, 0 off, -1 from kXwAutoRows rows
    int32_t mix_path = -1;              // mixed FP32 step: 1 windowed, 0 staged, -1 auto
    bool owns = false;                  // rowptr/colind/vals allocated here
    std::vector<int32_t> h_rowptr;      // host copy, drives scheduling
    int32_t max_row_nnz = 0;
    // Dominant long-range column offset |col - row| (the largest offset of
    // the highest power-of-two band that holds >= 10 % of the entries: the
    // z-plane offset nx*ny of a 3-D stencil), 0 if none; see plane_chunk.
    int32_t offset_d = 0;
    // Column-partitioned layout (num_csr_attach_xcd), if attached: the
    // segment matrix (one row per (column block, row) pair holding entries,
    // block-major), each row's segments in block order, and the segment
    // matrix's tiles grouped by the XCD that serves their column block.
    const num_csr* xcd_seg = nullptr;            // the caller's; outlives A
    const int32_t* xcd_row_seg_ptr = nullptr;    // [nrows + 1]
    const int32_t* xcd_row_seg_idx = nullptr;    // [nseg]
    numamd::Schedule xcd_sched;                  // owned: the segment tiles of pass 1
    int32_t* xcd_tiles = nullptr;                // owned: [tiles], grouped by XCD
    int32_t xcd_off[numamd::kNumXcd + 1] = {};   // group bounds in xcd_tiles
    numamd::Schedule sched;             // whole-matrix schedule
    // Engine systems admitted with this matrix and not yet removed: their
    // launch lists are expanded from sched.d_tiles, so the schedule must not
    // be rebuilt (and the handle not destroyed) while this is non-zero.
    mutable std::atomic<int32_t> engine_holds{0};
    mutable std::mutex mu;              // guards range_sched
    mutable std::map<std::pair<int32_t, int32_t>, numamd::Schedule> range_sched;
};

namespace numamd {

// Builds `s` for rows [r0, r1) of A and uploads it.  Returns NUM_OK or error.
// rows_per_block may reach max_rows (the row-tile kernels walk one row per
// thread: kSpmvBS; the column-partitioned pass 1 walks kSegRows); a row
// longer than long_row entries gets a tile of its own.
int build_schedule(const num_csr* A, int32_t r0, int32_t r1, int32_t rows_per_block,
                   int32_t nnz_per_block, int32_t chunk, int32_t nt64, int32_t nt32,
                   Schedule& s, int32_t max_rows = kSpmvBS, int32_t long_row = kLongRow);
// Default tile capacity for A: the smallest kNptChoices entry covering the
// mean row length, times the block size.
int32_t default_nnz_per_block(const num_csr* A);
void free_schedule(Schedule& s);
// Frees the windowed sliced layout of A (if any).
void free_windows(num_csr* A);
// XCD chunk (in units of rows_per_unit rows) that deals the rows offset_d
// apart to the same XCD: offset_d / 8 rows, so rows offset_d apart sit
// exactly 8 chunks apart and the x entries one chunk gathers across the
// long-range offset were fetched into that XCD's L2 by its previous chunk.
// 0 when there is no such offset or the chunk would fall outside
// [1, max_units].
int32_t plane_chunk(const num_csr* A, double rows_per_unit, int32_t max_units = 64);
// Schedule for a row range (cached per range); whole range -> A->sched.
int range_schedule(const num_csr* A, int32_t r0, int32_t r1, const Schedule** out);
int check_device(const num_csr* A);

// Launch y = A x over a schedule (T = double or float).
template <typename T>
int launch_spmv(const num_csr* A, const Schedule& s, const T* vals, const T* x, T* y,
                hipStream_t stream);

}  // namespace numamd
