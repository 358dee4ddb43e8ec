// runtime.hip -- errors, device selection, CSR handle life cycle and the
// row-block scheduler.
#include <algorithm>
#include <climits>
#include <cstring>
#include <new>

#include "blas1.hpp"
#include "csr.hpp"

namespace numamd {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

const char* last_error() { return g_err; }

int check_device(const num_csr* A)
{
    int dev = -1;
    NUM_HIP(hipGetDevice(&dev));
    if (dev != A->device)
        return fail(NUM_EINVAL, "matrix lives on device %d but device %d is current", A->device,
                    dev);
    return NUM_OK;
}

static int pick_npt(int32_t nnz_per_block)
{
    for (int c : kNptChoices)
        if (c * kSpmvBS >= nnz_per_block) return c;
    return kNptMax;
}

int32_t default_nnz_per_block(const num_csr* A)
{
    const double avg = A->nrows > 0 ? (double)A->nnz / (double)A->nrows : 1.0;
    for (int c : kNptChoices)
        if ((double)c >= avg) return c * kSpmvBS;
    return kNptMax * kSpmvBS;
}

// Walk mode of a staged tile: kWalkShort when every row has <= kShortRow
// entries (one thread per row, summed in index order: bit-identical to the
// row loop), kWalkMedium when some row is longer (those rows get a wave each).
static int32_t walk_mode(int32_t maxlen) { return maxlen > kShortRow ? kWalkMedium : kWalkShort; }

void free_schedule(Schedule& s)
{
    if (s.d_tiles) (void)hipFree(s.d_tiles);
    s.d_tiles = nullptr;
    s.nblocks = 0;
}

// Infinity Cache (MALL) capacity below which a matrix copy is loaded with
// normal (caching) loads so repeated SpMVs inside a solve find it resident.
constexpr double kMallResident = 256.0 * 1024 * 1024;

static int32_t auto_nt(const num_csr* A, int value_bytes)
{
    const double bytes = (double)A->nnz * (value_bytes + 4) + 4.0 * (A->nrows + 1);
    return bytes > kMallResident ? 1 : 0;
}

// Smallest plane-aligned XCD chunk (in tiles) the schedule takes.
constexpr int32_t kMinPlaneChunk = 3;

int build_schedule(const num_csr* A, int32_t r0, int32_t r1, int32_t rows_per_block,
                   int32_t nnz_per_block, int32_t chunk, int32_t nt64, int32_t nt32,
                   Schedule& s, int32_t max_rows, int32_t long_row)
{
    if (rows_per_block < 1 || rows_per_block > max_rows)
        return fail(NUM_EINVAL, "rows_per_block must be in [1, %d], got %d", max_rows,
                    rows_per_block);
    if (nnz_per_block < 1 || nnz_per_block > kNptMax * kSpmvBS)
        return fail(NUM_EINVAL, "nnz_per_block must be in [1, %d], got %d", kNptMax * kSpmvBS,
                    nnz_per_block);
    if (r0 < 0 || r1 > A->nrows || r0 > r1)
        return fail(NUM_EINVAL, "row range [%d, %d) outside [0, %d)", r0, r1, A->nrows);
    const int32_t* rp = A->h_rowptr.data();
    std::vector<TileEntry> tiles;
    tiles.reserve((size_t)(r1 - r0) / (size_t)std::max(1, rows_per_block / 2) + 8);
    int32_t longs = 0;
    int32_t r = r0;
    while (r < r1) {
        const int32_t start = r;
        const int32_t len = rp[r + 1] - rp[r];
        if (len > nnz_per_block || len > long_row) {
            ++r;
            ++longs;
            tiles.push_back({start, kWalkLong, rp[start], 0});
            continue;
        }
        int64_t nz = 0;
        int32_t maxlen = 0;
        while (r < r1 && r - start < rows_per_block) {
            const int32_t l = rp[r + 1] - rp[r];
            if (nz + l > nnz_per_block || l > long_row) break;
            nz += l;
            maxlen = std::max(maxlen, l);
            ++r;
        }
        tiles.push_back({start, walk_mode(maxlen), rp[start], 0});
    }
    tiles.push_back({r1, 0, rp[r1], 0});
    Schedule ns;
    ns.row_begin = r0;
    ns.row_end = r1;
    ns.rows_per_block = rows_per_block;
    ns.nnz_per_block = nnz_per_block;
    ns.npt = pick_npt(nnz_per_block);
    ns.nblocks = (int32_t)tiles.size() - 1;
    ns.long_rows = longs;
    if (chunk > 0) {
        ns.chunk = chunk;
    } else {
        const int32_t pc =
            ns.nblocks > 0 ? plane_chunk(A, (double)(r1 - r0) / ns.nblocks) : 0;
        // Plane-aligned chunks of at least kMinPlaneChunk tiles: 192^3
        // Laplacian 18 tiles 103.3 us against 106.3 us for 64, 128^3 8 tiles
        // 30.4 against 34.8 us, 80^3 3 tiles 9.9 against 10.1 us.  Below that
        // the size-based chunk is faster: 2-D 3000^2 (1 tile) 112.0 against
        // 101.6 us, 40x40x4000 (1 tile) 106 against 93 us, 64x64x2000
        // (2 tiles) 123.5 against 119.7 us.
        ns.chunk = pc >= kMinPlaneChunk && ns.nblocks >= 8 * pc ? pc : tile_chunk(ns.nblocks);
    }
    ns.nt64 = nt64 >= 0 ? (nt64 != 0) : auto_nt(A, 8);
    ns.nt32 = nt32 >= 0 ? (nt32 != 0) : auto_nt(A, 4);
    const size_t bytes = tiles.size() * sizeof(TileEntry);
    NUM_HIP(hipMalloc(&ns.d_tiles, bytes));
    hipError_t e = hipMemcpy(ns.d_tiles, tiles.data(), bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(ns.d_tiles);
        return fail(NUM_EHIP, "schedule upload failed: %s", hipGetErrorString(e));
    }
    free_schedule(s);
    s = ns;
    return NUM_OK;
}

int range_schedule(const num_csr* A, int32_t r0, int32_t r1, const Schedule** out)
{
    if (r0 == 0 && r1 == A->nrows) {
        *out = &A->sched;
        return NUM_OK;
    }
    std::lock_guard<std::mutex> lk(A->mu);
    auto key = std::make_pair(r0, r1);
    auto it = A->range_sched.find(key);
    if (it == A->range_sched.end()) {
        Schedule s;
        int rc = build_schedule(A, r0, r1, A->sched.rows_per_block, A->sched.nnz_per_block,
                                -1, A->sched.nt64, A->sched.nt32, s);
        if (rc) return rc;
        it = A->range_sched.emplace(key, std::move(s)).first;
    }
    *out = &it->second;
    return NUM_OK;
}

static int validate_rowptr(const int32_t* h_rowptr, int32_t nrows, int32_t ncols)
{
    if (nrows <= 0 || ncols <= 0)
        return fail(NUM_EINVAL, "CrsMatrixHandle: non-positive dimensions");
    if (!h_rowptr) return fail(NUM_EINVAL, "rowptr is NULL");
    if (h_rowptr[0] != 0) return fail(NUM_EINVAL, "rowptr[0] must be 0, got %d", h_rowptr[0]);
    for (int32_t i = 0; i < nrows; ++i)
        if (h_rowptr[i + 1] < h_rowptr[i])
            return fail(NUM_EINVAL, "rowptr decreases at row %d", i);
    return NUM_OK;
}

// |col - row| over all entries by power-of-two band (band b: offsets in
// [2^b, 2^(b+1)); zero offsets skipped), with each band's largest offset.
// Workgroup-local counts in LDS, one global atomic per band per workgroup.
// Also flags (*bad = 1) a column index outside [0, ncols): every kernel
// gathers x[col] unchecked, so such a matrix is refused at create.
__global__ void offset_hist_kernel(int32_t nrows, int32_t ncols,
                                   const int32_t* __restrict__ rowptr,
                                   const int32_t* __restrict__ colind,
                                   unsigned long long* __restrict__ count,
                                   int32_t* __restrict__ maxo, int32_t* __restrict__ bad)
{
    __shared__ unsigned int lcount[32];
    __shared__ int lmax[32];
    if (threadIdx.x < 32) {
        lcount[threadIdx.x] = 0;
        lmax[threadIdx.x] = 0;
    }
    __syncthreads();
    int out_of_range = 0;
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nrows; r += gridDim.x * blockDim.x) {
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            const int c = colind[j];
            if (c < 0 || c >= ncols) {
                out_of_range = 1;
                continue;
            }
            const int d = c - r;
            const int o = d < 0 ? -d : d;
            if (o > 0) {
                const int b = 31 - __clz(o);
                atomicAdd(&lcount[b], 1u);
                atomicMax(&lmax[b], o);
            }
        }
    }
    if (out_of_range) atomicOr(bad, 1);
    __syncthreads();
    if (threadIdx.x < 32 && lcount[threadIdx.x]) {
        atomicAdd(&count[threadIdx.x], (unsigned long long)lcount[threadIdx.x]);
        atomicMax(&maxo[threadIdx.x], lmax[threadIdx.x]);
    }
}

// A->offset_d from the band histogram (synchronous); NUM_EINVAL if a
// column index lies outside [0, ncols).
static int find_offset(num_csr* A)
{
    A->offset_d = 0;
    if (A->nnz == 0) return NUM_OK;
    // the arrays may have been filled on any stream (set-up only, once)
    NUM_HIP(hipDeviceSynchronize());
    constexpr size_t kBytes = 32 * sizeof(unsigned long long) + 33 * sizeof(int32_t);
    void* d = nullptr;
    NUM_HIP(hipMalloc(&d, kBytes));
    auto* d_count = static_cast<unsigned long long*>(d);
    auto* d_max = reinterpret_cast<int32_t*>(d_count + 32);
    int32_t* d_bad = d_max + 32;
    unsigned long long h_count[32];
    int32_t h_max[33];
    hipError_t e = hipMemset(d, 0, kBytes);
    if (e == hipSuccess) {
        const int grid = std::min(2048, (A->nrows + 255) / 256);
        offset_hist_kernel<<<grid, 256>>>(A->nrows, A->ncols, A->rowptr, A->colind, d_count,
                                          d_max, d_bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(h_count, d_count, sizeof(h_count), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(h_max, d_max, sizeof(h_max), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(NUM_EHIP, "offset histogram failed: %s", hipGetErrorString(e));
    if (h_max[32] != 0)
        return fail(NUM_EINVAL, "column index outside [0, %d)", A->ncols);
    unsigned long long cum = 0;
    for (int b = 31; b >= 0; --b) {
        cum += h_count[b];
        if (cum * 10 >= (unsigned long long)A->nnz) {
            A->offset_d = h_max[b];
            break;
        }
    }
    return NUM_OK;
}

int32_t plane_chunk(const num_csr* A, double rows_per_unit, int32_t max_units)
{
    if (A->offset_d <= 0 || rows_per_unit <= 0) return 0;
    const double c = A->offset_d / (double)kNumXcd / rows_per_unit;
    const int32_t r = (int32_t)(c + 0.5);
    return r >= 1 && r <= max_units ? r : 0;
}

static int finish_create(num_csr* A)
{
    int32_t mx = 0;
    for (int32_t i = 0; i < A->nrows; ++i)
        mx = std::max(mx, A->h_rowptr[i + 1] - A->h_rowptr[i]);
    A->max_row_nnz = mx;
    int rc = find_offset(A);
    if (rc) return rc;
    return build_schedule(A, 0, A->nrows, kSpmvBS, default_nnz_per_block(A), -1, -1, -1,
                          A->sched);
}

// rel[j] = colind[j] - row(j) for every entry; *bad = 1 if one does not fit
// in int16.
__global__ void rel16_kernel(int32_t nrows, const int32_t* __restrict__ rowptr,
                             const int32_t* __restrict__ colind, int16_t* __restrict__ rel,
                             int32_t* __restrict__ bad)
{
    const int32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    int flag = 0;
    for (int32_t j = rowptr[r]; j < rowptr[r + 1]; ++j) {
        const int32_t d = colind[j] - r;
        if (d < -32768 || d > 32767) flag = 1;
        rel[j] = (int16_t)d;
    }
    if (flag) atomicOr(bad, 1);
}

// Row patterns of the int16 offsets, one workgroup per pair of schedule
// tiles (at most 2 * 256 rows, two per thread): each row's offsets column -
// row (at most kRelPatW, zeros past its length) packed in two uint64, with
// its length (a zero offset is the diagonal, so the length is part of the
// pattern); the pair's distinct patterns numbered in order of their first
// row, in rounds (the first row without an id leads, every row with its
// pattern takes the next id); the rows' id bytes and the pair's table
// (relpat_words: patterns, then 4-bit lengths) written.  *bad = 1 if a row
// holds more than kRelPatW entries or the pair more than kRelPat patterns.
__global__ __launch_bounds__(256) void relpat_kernel(const int32_t* __restrict__ rowptr,
                                                     const int16_t* __restrict__ rel,
                                                     const TileEntry* __restrict__ tiles,
                                                     int32_t ntiles, uint64_t* __restrict__ pat,
                                                     uint8_t* __restrict__ pid,
                                                     int32_t* __restrict__ bad)
{
    __shared__ unsigned long long s_q[4 * 256];
    __shared__ int s_l[2 * 256];
    __shared__ unsigned long long s_tab[2 * kRelPat];
    __shared__ int s_tl[kRelPat];
    __shared__ int s_lead;
    const int p = blockIdx.x, t = threadIdx.x;
    const int r0 = tiles[2 * p].row, r1 = tiles[min(2 * p + 2, ntiles)].row;
    uint64_t q[2][2] = {{0, 0}, {0, 0}};
    int myid[2] = {-1, -1}, ln[2] = {0, 0};
    bool act[2];
    int over = 0;
    for (int h = 0; h < 2; ++h) {
        const int lr = t + h * 256, r = r0 + lr;
        act[h] = r < r1;
        if (act[h]) {
            const int js = rowptr[r], len = rowptr[r + 1] - js;
            ln[h] = len;
            if (len > kRelPatW) over = 1;
            for (int k = 0; k < min(len, kRelPatW); ++k) {
                const uint64_t d = (uint64_t)(uint16_t)rel[js + k];
                if (k < 4) q[h][0] |= d << (16 * k);
                else q[h][1] |= d << (16 * (k - 4));
            }
        }
        s_q[2 * lr] = q[h][0];
        s_q[2 * lr + 1] = q[h][1];
        s_l[lr] = ln[h];
    }
    if (over) atomicOr(bad, 1);  // the host drops the table; the ids are not used
    int np = 0;
    for (;;) {
        if (t == 0) s_lead = INT_MAX;
        __syncthreads();
        for (int h = 0; h < 2; ++h)
            if (act[h] && myid[h] < 0) atomicMin(&s_lead, t + h * 256);
        __syncthreads();
        const int lead = s_lead;
        if (lead == INT_MAX) break;
        if (np == kRelPat) {
            if (t == 0) atomicOr(bad, 1);
            return;  // uniform: every thread saw the same lead and np
        }
        const unsigned long long a = s_q[2 * lead], c = s_q[2 * lead + 1];
        const int l = s_l[lead];
        for (int h = 0; h < 2; ++h)
            if (act[h] && myid[h] < 0 && q[h][0] == a && q[h][1] == c && ln[h] == l)
                myid[h] = np;
        if (t == 0) {
            s_tab[2 * np] = a;
            s_tab[2 * np + 1] = c;
            s_tl[np] = l;
        }
        ++np;
        __syncthreads();  // s_lead is reset only after every thread read it
    }
    for (int h = 0; h < 2; ++h)
        if (act[h]) pid[r0 + t + h * 256] = (uint8_t)myid[h];
    uint64_t* out = pat + (size_t)p * relpat_words<2 * kRelPat>();
    if (t < 2 * kRelPat) out[t] = t < 2 * np ? s_tab[t] : 0ull;
    if (t == 0) {
        uint64_t lw = 0;
        for (int i = 0; i < np; ++i) lw |= (uint64_t)s_tl[i] << (4 * i);
        out[2 * kRelPat] = lw;
    }
    static_assert(relpat_len_words<2 * kRelPat>() == 1, "one length word per pair");
}

// Delta codes, one workgroup per tile (rows_per_block <= blockDim): the
// tile's base column is the smallest first column of its rows; each row's
// first entry is coded against the base, every later entry against the
// previous entry.  *bad = 1 if a code is negative (columns out of order) or
// exceeds 16 bits.
__global__ void d16_build_kernel(const int32_t* __restrict__ rowptr,
                                 const int32_t* __restrict__ colind, TileEntry* __restrict__ tiles,
                                 uint16_t* __restrict__ code, uint8_t* __restrict__ lens,
                                 int32_t* __restrict__ bad)
{
    __shared__ int base;
    const int r0 = tiles[blockIdx.x].row, r1 = tiles[blockIdx.x + 1].row;
    if (threadIdx.x == 0) base = INT32_MAX;
    __syncthreads();
    for (int r = r0 + (int)threadIdx.x; r < r1; r += blockDim.x) {
        const int js = rowptr[r], je = rowptr[r + 1];
        if (je > js) atomicMin(&base, colind[js]);
    }
    __syncthreads();
    const int b = base == INT32_MAX ? 0 : base;
    int flag = 0;
    for (int r = r0 + (int)threadIdx.x; r < r1; r += blockDim.x) {
        int prev = b;
        lens[r] = (uint8_t)(rowptr[r + 1] - rowptr[r]);  // <= kShortRow
        for (int j = rowptr[r]; j < rowptr[r + 1]; ++j) {
            const int c = colind[j];
            const int d = c - prev;
            if (d < 0 || d > 65535) flag = 1;
            code[j] = (uint16_t)d;
            prev = c;
        }
    }
    if (flag) atomicOr(bad, 1);
    if (threadIdx.x == 0) tiles[blockIdx.x].base = b;
}

// (Re)code A->d16_buf against A->sched; attaches on success, detaches
// otherwise.  Synchronises `st`.
static int build_d16(num_csr* A, hipStream_t st)
{
    A->d16 = nullptr;
    if (!A->d16_buf || A->nnz == 0 || A->max_row_nnz > kShortRow || A->sched.long_rows != 0 ||
        A->sched.nblocks == 0 || A->sched.rows_per_block > kSpmvBS)
        return NUM_OK;
    int32_t* d_bad = nullptr;
    NUM_HIP(hipMalloc(&d_bad, sizeof(int32_t)));
    int32_t h_bad = 0;
    hipError_t e = hipMemcpyAsync(d_bad, &h_bad, sizeof(int32_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        d16_build_kernel<<<A->sched.nblocks, kSpmvBS, 0, st>>>(
            A->rowptr, A->colind, static_cast<TileEntry*>(A->sched.d_tiles), A->d16_buf,
            d16_lens(A->d16_buf, A->nnz), d_bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(&h_bad, d_bad, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_bad);
    if (e != hipSuccess) return fail(NUM_EHIP, "delta-code build failed: %s", hipGetErrorString(e));
    if (h_bad == 0) A->d16 = A->d16_buf;
    return NUM_OK;
}

static void free_relpat(num_csr* A)
{
    // rel_pid lives in the rel_pat allocation
    if (A->rel_pat) (void)hipFree(A->rel_pat);
    A->rel_pat = nullptr;
    A->rel_pid = nullptr;
}

// (Re)builds the row patterns of A->rel16 over the pairs of A->sched's tiles
// (relpat_kernel); nothing is kept when a row is longer than kRelPatW, a
// pair holds more than kRelPat patterns, or a tile has more than 256 rows.
// Synchronises `st`.
static int build_relpat(num_csr* A, hipStream_t st)
{
    free_relpat(A);
    const Schedule& s = A->sched;
    if (!A->rel16 || A->max_row_nnz > kRelPatW || s.nblocks == 0 ||
        s.rows_per_block > 256 || s.long_rows != 0)
        return NUM_OK;
    const int32_t npairs = (s.nblocks + 1) / 2;
    const size_t npat = (size_t)npairs * relpat_words<2 * kRelPat>();
    int32_t* d_bad = nullptr;
    hipError_t e = hipMalloc(&d_bad, sizeof(int32_t));
    if (e == hipSuccess) e = hipMemsetAsync(d_bad, 0, sizeof(int32_t), st);
    if (e == hipSuccess) e = hipMalloc(&A->rel_pat, npat * sizeof(uint64_t) + A->nrows);
    if (e == hipSuccess) {
        A->rel_pid = reinterpret_cast<uint8_t*>(A->rel_pat + npat);
        relpat_kernel<<<npairs, 256, 0, st>>>(A->rowptr, A->rel16,
                                              static_cast<const TileEntry*>(s.d_tiles),
                                              s.nblocks, A->rel_pat, A->rel_pid, d_bad);
        e = hipGetLastError();
    }
    int32_t h_bad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h_bad, d_bad, sizeof(int32_t),
                                            hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d_bad) (void)hipFree(d_bad);
    if (e != hipSuccess || h_bad) {
        free_relpat(A);
        if (e != hipSuccess)
            return fail(NUM_EHIP, "offset pattern build failed: %s", hipGetErrorString(e));
    }
    return NUM_OK;
}

// Windows of one tile (rows [r0, r1)), one workgroup per tile: the columns
// its entries read plus its own rows as columns (so the CG kernel finds
// p_new of its own rows there).  Each entry's column offset (column - row)
// goes into an LDS hash set of kXwOffsets offsets, which keeps each
// offset's smallest and largest column; thread 0 then sorts those ranges
// by first column and merges the ones that overlap or lie at most kXwGap
// columns apart (the columns in between lie inside [0, ncols) too).  Writes
// the tile's record (windows, first slice b * stride, width) and sets *fail
// when the tile has more offsets than the set holds, more than
// kXwWin windows or more than kXwSlots slots.
constexpr int kXwOffsets = 32;
constexpr int kXwGap = 16;

__global__ __launch_bounds__(kSpmvBS) void xw_tile_kernel(const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colind,
                                                        const TileEntry* __restrict__ tiles,
                                                        int32_t ncols, int32_t stride,
                                                        int32_t* __restrict__ rec,
                                                        int32_t* __restrict__ fail)
{
    __shared__ int s_off[kXwOffsets], s_lo[kXwOffsets], s_hi[kXwOffsets];
    __shared__ int s_fail, s_width;
    const int b = blockIdx.x, t = threadIdx.x;
    const int r0 = tiles[b].row, r1 = tiles[b + 1].row;
    if (t < kXwOffsets) {
        s_off[t] = INT_MIN;  // empty
        s_lo[t] = INT_MAX;
        s_hi[t] = INT_MIN;
    }
    if (t == 0) {
        s_fail = 0;
        s_width = 1;  // at least one slice per tile
    }
    __syncthreads();
    auto insert = [&](int d, int lo, int hi) {
        int h = (int)(((uint32_t)d * 2654435761u) >> 27) & (kXwOffsets - 1);
        for (int probe = 0; probe < kXwOffsets; ++probe) {
            const int old = atomicCAS(&s_off[h], INT_MIN, d);
            if (old == INT_MIN || old == d) {
                atomicMin(&s_lo[h], lo);
                atomicMax(&s_hi[h], hi);
                return;
            }
            h = (h + 1) & (kXwOffsets - 1);
        }
        s_fail = 1;
    };
    if (t == 0 && r0 < min(r1, ncols)) insert(0, r0, min(r1, ncols));  // the own rows
    const int r = r0 + t;
    if (r < r1) {
        const int js = rowptr[r], je = rowptr[r + 1];
        atomicMax(&s_width, je - js);
        for (int j = js; j < je; ++j) {
            const int c = colind[j];
            insert(c - r, c, c + 1);
        }
    }
    __syncthreads();
    if (t != 0) return;
    int lo[kXwOffsets], hi[kXwOffsets], n = 0;
    for (int q = 0; q < kXwOffsets; ++q)
        if (s_off[q] != INT_MIN) {
            int k = n++;  // insertion by first column
            while (k > 0 && lo[k - 1] > s_lo[q]) {
                lo[k] = lo[k - 1];
                hi[k] = hi[k - 1];
                --k;
            }
            lo[k] = s_lo[q];
            hi[k] = s_hi[q];
        }
    int nw = 0, total = 0, bad = s_fail;
    int wl[kXwWin], wh[kXwWin];
    for (int a = 0; a < n && !bad; ++a) {
        if (nw > 0 && lo[a] <= wh[nw - 1] + kXwGap) {
            wh[nw - 1] = max(wh[nw - 1], hi[a]);
            continue;
        }
        if (nw == kXwWin) {
            bad = 1;
            break;
        }
        wl[nw] = lo[a];
        wh[nw] = hi[a];
        ++nw;
    }
    int32_t* R = rec + (size_t)b * kXwRec;
    for (int w = 0; w < kXwWin; ++w) {
        if (w < nw) total += wh[w] - wl[w];
        R[w] = w < nw ? wl[w] : 0;
        R[kXwWin + w] = total;
    }
    R[kXwBase] = b * stride;
    R[kXwWidth] = s_width;
    if (bad || total > kXwSlots) atomicOr(fail, 1);
}

// Each entry's slot in its tile's windows at its packed position
// (sell_pos, four per group), zeros in the padding, and one length byte
// per row; one workgroup per tile, one thread per row slot.  With `pat`:
// also the row-pattern form (kXwPat): each row's slots minus its lane plus
// kXwPatBias (0 past its length), its length in bits 12-15 of the first
// (a biased slot is < kXwSlots + kXwPatBias < 4096), are compared in
// rounds, the tile's distinct patterns numbered in order of their first
// row, then the rows' 4-bit ids (two lanes to a byte, kSpmvBS / 2 bytes per
// tile) and the tile's table are written; a tile with more than kXwPat
// patterns sets *pfail.
__global__ __launch_bounds__(kSpmvBS) void xw_code_kernel(const int32_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ colind,
                                                        const TileEntry* __restrict__ tiles,
                                                        const int32_t* __restrict__ rec,
                                                        uint16_t* __restrict__ codes,
                                                        uint8_t* __restrict__ lens,
                                                        uint64_t* __restrict__ pat,
                                                        uint8_t* __restrict__ pid,
                                                        int32_t* __restrict__ pfail)
{
    __shared__ unsigned long long s_q[2 * kSpmvBS];
    __shared__ unsigned long long s_tab[2 * kXwPat];
    __shared__ int s_lead;
    const int b = blockIdx.x, i = threadIdx.x;
    const int r0 = tiles[b].row, r1 = tiles[b + 1].row;
    const int32_t* R = rec + (size_t)b * kXwRec;
    const int width = R[kXwWidth];
    uint16_t* __restrict__ out = codes + (int64_t)R[kXwBase] * kSpmvBS;
    const int r = r0 + i;
    const int js = r < r1 ? rowptr[r] : 0;
    const int len = r < r1 ? rowptr[r + 1] - js : 0;
    if (r < r1) lens[r] = (uint8_t)len;
    uint64_t q0 = 0, q1 = 0;
    for (int k = 0; k < width; ++k) {
        int s = 0;
        if (k < len) {
            const int c = colind[js + k];
            int w = 0;
            while (w + 1 < kXwWin && R[kXwWin + w] < R[kXwWin + w + 1] && c >= R[w + 1]) ++w;
            s = (w == 0 ? 0 : R[kXwWin + w - 1]) + c - R[w];
            const uint64_t rel = (uint64_t)(uint16_t)(s - i + kXwPatBias);
            if (k < 4) q0 |= rel << (16 * k);
            else q1 |= rel << (16 * (k - 4));
        }
        out[sell_pos(k, i, width, 4, kSpmvBS)] = (uint16_t)s;
    }
    q0 |= (uint64_t)len << 12;
    if (!pat) return;
    // number the distinct patterns in order of first row: each round the
    // first row without an id leads, every row with its pattern takes the
    // next id (at most kXwPat + 1 rounds, all threads in step)
    const bool active = r < r1;
    s_q[2 * i] = q0;
    s_q[2 * i + 1] = q1;
    int myid = -1, np = 0, over = 0;
    for (;;) {
        if (i == 0) s_lead = INT_MAX;
        __syncthreads();
        if (active && myid < 0) atomicMin(&s_lead, i);
        __syncthreads();
        const int lead = s_lead;
        if (lead == INT_MAX) break;
        if (np == kXwPat) {
            over = 1;
            break;
        }
        const unsigned long long a = s_q[2 * lead], c = s_q[2 * lead + 1];
        if (active && myid < 0 && q0 == a && q1 == c) myid = np;
        if (i == 0) {
            s_tab[2 * np] = a;
            s_tab[2 * np + 1] = c;
        }
        ++np;
        __syncthreads();  // s_lead is reset only after every thread read it
    }
    if (over) {
        if (i == 0) atomicOr(pfail, 1);
        return;
    }
    // lanes i and i + 1 share a wave: the even lane writes both ids
    const int idv = active ? myid : 0;
    const int hi = __shfl_down(idv, 1, kWave);
    if ((i & 1) == 0) pid[(size_t)b * (kSpmvBS / 2) + (i >> 1)] = (uint8_t)(idv | (hi << 4));
    if (i < 2 * kXwPat) pat[(size_t)b * 2 * kXwPat + i] = i < 2 * np ? s_tab[i] : 0ull;
}

// Scatter of the values into the sliced layout, one workgroup per tile and
// one thread per row slot, zeros in the padding (rows past the tile's last,
// entries past a row's length, up to the tile's width): FP64 values one
// slice per entry index, FP32 values packed in pairs (sell_pos).
template <typename T, int Pack>
__global__ void sell_scatter_kernel(const int32_t* __restrict__ rowptr,
                                    const T* __restrict__ vals,
                                    const TileEntry* __restrict__ tiles,
                                    const int32_t* __restrict__ rec, T* __restrict__ svals)
{
    const int b = blockIdx.x, i = threadIdx.x;
    const int r0 = tiles[b].row, r1 = tiles[b + 1].row;
    const int32_t* R = rec + (size_t)b * kXwRec;
    T* __restrict__ out = svals + (int64_t)R[kXwBase] * kSpmvBS;
    const int width = R[kXwWidth];
    const int js = r0 + i < r1 ? rowptr[r0 + i] : 0;
    const int len = r0 + i < r1 ? rowptr[r0 + i + 1] - js : 0;
    for (int k = 0; k < width; ++k)
        out[sell_pos(k, i, width, Pack, kSpmvBS)] = k < len ? vals[js + k] : T(0);
}

// The FP32 values of A in the sliced layout (A->xw built, vals32 attached).
static int build_sell32(num_csr* A, hipStream_t st)
{
    if (A->sell_vals32) (void)hipFree(A->sell_vals32);
    A->sell_vals32 = nullptr;
    if (!A->xw || !A->vals32) return NUM_OK;
    const Schedule& s = A->sched;
    const size_t nval = (size_t)s.nblocks * A->sell_w * kSpmvBS;
    NUM_HIP(hipMalloc(&A->sell_vals32, nval * sizeof(float)));
    sell_scatter_kernel<float, 2><<<s.nblocks, kSpmvBS, 0, st>>>(
        A->rowptr, A->vals32, static_cast<const TileEntry*>(s.d_tiles), A->xw, A->sell_vals32);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)hipFree(A->sell_vals32);
        A->sell_vals32 = nullptr;
        return fail(NUM_EHIP, "FP32 sliced values failed: %s", hipGetErrorString(e));
    }
    return NUM_OK;
}

static void free_patterns(num_csr* A)
{
    // sell_tix and sell_pid live in the sell_pat allocation
    if (A->sell_pat) (void)hipFree(A->sell_pat);
    A->sell_pat = nullptr;
    A->sell_tix = nullptr;
    A->sell_pid = nullptr;
}

// Keeps one copy of each distinct tile table: `tabs` holds every tile's
// table (2 * kXwPat words), `pid` the 4-bit ids; the final allocation is
// the distinct tables, each tile's index into them, and the ids (moved
// over).  Interior tiles of a stencil share a few dozen tables, which then
// stay in the L2 instead of streaming 256 bytes per tile.  Synchronises.
static int share_tables(num_csr* A, const uint64_t* tabs, const uint8_t* pid, hipStream_t st)
{
    const int32_t nb = A->sched.nblocks;
    constexpr int TW = 2 * kXwPat;
    std::vector<uint64_t> h((size_t)nb * TW);
    NUM_HIP(hipMemcpyAsync(h.data(), tabs, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           st));
    NUM_HIP(hipStreamSynchronize(st));
    std::map<std::vector<uint64_t>, int32_t> seen;
    std::vector<uint64_t> uniq;
    std::vector<int32_t> tix(nb);
    for (int32_t b = 0; b < nb; ++b) {
        std::vector<uint64_t> key(h.begin() + (size_t)b * TW, h.begin() + (size_t)(b + 1) * TW);
        auto it = seen.find(key);
        if (it == seen.end()) {
            it = seen.emplace(std::move(key), (int32_t)(uniq.size() / TW)).first;
            uniq.insert(uniq.end(), h.begin() + (size_t)b * TW, h.begin() + (size_t)(b + 1) * TW);
        }
        tix[b] = it->second;
    }
    const size_t tab_bytes = (uniq.size() * sizeof(uint64_t) + 255) & ~size_t(255);
    const size_t tix_bytes = ((size_t)nb * sizeof(int32_t) + 255) & ~size_t(255);
    const size_t pid_bytes = (size_t)nb * (kSpmvBS / 2);
    char* buf = nullptr;
    NUM_HIP(hipMalloc(&buf, tab_bytes + tix_bytes + pid_bytes));
    hipError_t e = hipMemcpyAsync(buf, uniq.data(), uniq.size() * sizeof(uint64_t),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(buf + tab_bytes, tix.data(), (size_t)nb * sizeof(int32_t),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(buf + tab_bytes + tix_bytes, pid, pid_bytes, hipMemcpyDeviceToDevice,
                           st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
        (void)hipFree(buf);
        return fail(NUM_EHIP, "row pattern tables failed: %s", hipGetErrorString(e));
    }
    free_patterns(A);
    A->sell_pat = reinterpret_cast<uint64_t*>(buf);
    A->sell_tix = reinterpret_cast<int32_t*>(buf + tab_bytes);
    A->sell_pid = reinterpret_cast<uint8_t*>(buf + tab_bytes + tix_bytes);
    return NUM_OK;
}

void free_windows(num_csr* A)
{
    // sell_codes and sell_lens live in the sell_vals allocation
    if (A->xw) (void)hipFree(A->xw);
    if (A->sell_vals) (void)hipFree(A->sell_vals);
    if (A->sell_vals32) (void)hipFree(A->sell_vals32);
    free_patterns(A);
    A->xw = nullptr;
    A->sell_vals = nullptr;
    A->sell_vals32 = nullptr;
    A->sell_codes = nullptr;
    A->sell_lens = nullptr;
    A->sell_w = 0;
}

// Smallest matrix (rows) for which the default mode builds the layout.
constexpr int32_t kXwAutoRows = 65536;

// (Re)builds the windowed layout of A->sched (spmv_sell_run) on the device,
// three kernels and one read of the failure flag: every tile's windows and
// width (xw_tile_kernel), each entry's packed slot and the row lengths
// (xw_code_kernel), the FP64 values in sliced order (sell_scatter_kernel);
// the FP32 values too when attached.  Tile b's slices start at b * W, W =
// the longest row (a stencil's tiles nearly all have that width; a tile
// loads only its own width).  Nothing is kept -- the kernels keep the
// delta-coded or int32 index -- when windows are switched off (or, in the
// default mode, the matrix is small), the schedule has long or wide rows,
// or any tile does not fit.  Synchronises `st`.
static int build_xw(num_csr* A, hipStream_t st)
{
    free_windows(A);
    const Schedule& s = A->sched;
    if (A->xw_mode == 0 || (A->xw_mode < 0 && A->nrows < kXwAutoRows))
        return NUM_OK;
    if (A->nnz == 0 || s.nblocks == 0 || s.long_rows != 0 ||
        s.rows_per_block > kSpmvBS || A->max_row_nnz > kSellW || s.row_begin != 0 ||
        s.row_end != A->nrows)
        return NUM_OK;
    const int32_t W = std::max(1, A->max_row_nnz);
    if ((int64_t)s.nblocks * W > INT32_MAX / kSpmvBS) return NUM_OK;
    const TileEntry* tiles = static_cast<const TileEntry*>(s.d_tiles);
    const size_t nrec = (size_t)s.nblocks * kXwRec;
    const size_t nval = (size_t)s.nblocks * W * kSpmvBS;
    const size_t npat = (size_t)s.nblocks * 2 * kXwPat;
    int32_t h_fail[2] = {0, 0};
    // the records, then the failure flags (windows, patterns); the FP64
    // values, then the slots, then the row lengths; the pattern tables,
    // then the row pattern ids
    hipError_t e = hipMalloc(&A->xw, (nrec + 2) * sizeof(int32_t));
    int32_t* d_fail = A->xw ? A->xw + nrec : nullptr;
    if (e == hipSuccess) e = hipMemsetAsync(d_fail, 0, 2 * sizeof(int32_t), st);
    if (e == hipSuccess)
        e = hipMalloc(&A->sell_vals, nval * (sizeof(double) + sizeof(uint16_t)) + A->nrows);
    // every tile's table, then the ids (share_tables keeps the distinct tables)
    if (e == hipSuccess)
        e = hipMalloc(&A->sell_pat, npat * sizeof(uint64_t) + (size_t)s.nblocks * (kSpmvBS / 2));
    if (e == hipSuccess) {
        A->sell_codes = reinterpret_cast<uint16_t*>(A->sell_vals + nval);
        A->sell_lens = reinterpret_cast<uint8_t*>(A->sell_codes + nval);
        A->sell_pid = reinterpret_cast<uint8_t*>(A->sell_pat + npat);
        xw_tile_kernel<<<s.nblocks, kSpmvBS, 0, st>>>(A->rowptr, A->colind, tiles, A->ncols, W,
                                                      A->xw, d_fail);
        xw_code_kernel<<<s.nblocks, kSpmvBS, 0, st>>>(A->rowptr, A->colind, tiles, A->xw,
                                                      A->sell_codes, A->sell_lens, A->sell_pat,
                                                      A->sell_pid, d_fail + 1);
        sell_scatter_kernel<double, 1><<<s.nblocks, kSpmvBS, 0, st>>>(A->rowptr, A->vals, tiles,
                                                                      A->xw, A->sell_vals);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_fail, d_fail, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess || h_fail[0]) {
        free_windows(A);
        if (e != hipSuccess)
            return fail(NUM_EHIP, "windowed layout build failed: %s", hipGetErrorString(e));
        return NUM_OK;  // a tile does not fit: no windows
    }
    if (h_fail[1] || A->pat_mode == 0) {
        free_patterns(A);  // the slots serve instead
    } else {
        const int rc = share_tables(A, A->sell_pat, A->sell_pid, st);
        if (rc) {
            free_windows(A);
            return rc;
        }
    }
    A->sell_w = W;
    return build_sell32(A, st);
}

}  // namespace numamd

using namespace numamd;

extern "C" {

const char* num_last_error(void) { return last_error(); }

int32_t num_abi_version(void) { return 1; }

int num_init(int32_t device_id)
{
    int count = 0;
    NUM_HIP(hipGetDeviceCount(&count));
    if (device_id < 0 || device_id >= count)
        return fail(NUM_EINVAL, "device_id %d out of range (%d devices)", device_id, count);
    NUM_HIP(hipSetDevice(device_id));
    // the static scratch of num_dot_noalloc, resolved once per device
    return noalloc_setup();
}

int num_finalize(void) { return NUM_OK; }

int num_csr_create(const int32_t* h_rowptr, const int32_t* d_rowptr, const int32_t* d_colind,
                   const double* d_vals, int32_t nrows, int32_t ncols, num_csr** out)
{
    if (!out) return fail(NUM_EINVAL, "out is NULL");
    *out = nullptr;
    int rc = validate_rowptr(h_rowptr, nrows, ncols);
    if (rc) return rc;
    const int64_t nnz = h_rowptr[nrows];
    if (!d_rowptr || (nnz > 0 && (!d_colind || !d_vals)))
        return fail(NUM_EINVAL, "device CSR pointer is NULL");
    num_csr* A = new (std::nothrow) num_csr();
    if (!A) return fail(NUM_ENOMEM, "out of host memory");
    if (hipGetDevice(&A->device) != hipSuccess) {
        delete A;
        return fail(NUM_EHIP, "hipGetDevice failed");
    }
    A->nrows = nrows;
    A->ncols = ncols;
    A->nnz = nnz;
    A->rowptr = d_rowptr;
    A->colind = d_colind;
    A->vals = d_vals;
    A->h_rowptr.assign(h_rowptr, h_rowptr + nrows + 1);
    rc = finish_create(A);
    if (rc) {
        delete A;
        return rc;
    }
    *out = A;
    return NUM_OK;
}

int num_csr_create_host(const int32_t* h_rowptr, const int32_t* h_colind, const double* h_vals,
                        int32_t nrows, int32_t ncols, num_csr** out)
{
    if (!out) return fail(NUM_EINVAL, "out is NULL");
    *out = nullptr;
    int rc = validate_rowptr(h_rowptr, nrows, ncols);
    if (rc) return rc;
    const int64_t nnz = h_rowptr[nrows];
    if (nnz > 0 && (!h_colind || !h_vals)) return fail(NUM_EINVAL, "host CSR pointer is NULL");
    for (int64_t j = 0; j < nnz; ++j)
        if (h_colind[j] < 0 || h_colind[j] >= ncols)
            return fail(NUM_EINVAL, "column index %d at position %lld outside [0, %d)",
                        h_colind[j], (long long)j, ncols);
    num_csr* A = new (std::nothrow) num_csr();
    if (!A) return fail(NUM_ENOMEM, "out of host memory");
    if (hipGetDevice(&A->device) != hipSuccess) {
        delete A;
        return fail(NUM_EHIP, "hipGetDevice failed");
    }
    A->nrows = nrows;
    A->ncols = ncols;
    A->nnz = nnz;
    A->owns = true;
    int32_t* rp = nullptr;
    int32_t* ci = nullptr;
    double* v = nullptr;
    hipError_t e = hipMalloc(&rp, sizeof(int32_t) * (size_t)(nrows + 1));
    if (e == hipSuccess) e = hipMalloc(&ci, sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1));
    if (e == hipSuccess) e = hipMalloc(&v, sizeof(double) * (size_t)std::max<int64_t>(nnz, 1));
    if (e == hipSuccess)
        e = hipMemcpy(rp, h_rowptr, sizeof(int32_t) * (size_t)(nrows + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz > 0)
        e = hipMemcpy(ci, h_colind, sizeof(int32_t) * (size_t)nnz, hipMemcpyHostToDevice);
    if (e == hipSuccess && nnz > 0)
        e = hipMemcpy(v, h_vals, sizeof(double) * (size_t)nnz, hipMemcpyHostToDevice);
    A->rowptr = rp;
    A->colind = ci;
    A->vals = v;
    if (e != hipSuccess) {
        num_csr_destroy(A);
        return fail(e == hipErrorOutOfMemory ? NUM_ENOMEM : NUM_EHIP, "CSR upload failed: %s",
                    hipGetErrorString(e));
    }
    A->h_rowptr.assign(h_rowptr, h_rowptr + nrows + 1);
    rc = finish_create(A);
    if (rc) {
        num_csr_destroy(A);
        return rc;
    }
    *out = A;
    return NUM_OK;
}

int num_csr_attach_f32(num_csr* A, const float* d_vals32)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (!d_vals32 && A->nnz > 0) return fail(NUM_EINVAL, "d_vals32 is NULL");
    int rc = check_device(A);
    if (rc) return rc;
    A->vals32 = d_vals32;
    return build_sell32(A, nullptr);  // the sliced FP32 copy, if the layout is built
}

int num_csr_attach_rel16(num_csr* A, int16_t* d_rel16, void* stream, int32_t* attached)
{
    if (!A || !attached) return fail(NUM_EINVAL, "matrix or attached is NULL");
    *attached = 0;
    if (!d_rel16 && A->nnz > 0) return fail(NUM_EINVAL, "d_rel16 is NULL");
    int rc = check_device(A);
    if (rc) return rc;
    // only for matrices every row of which the one-thread walk handles
    if (A->nrows != A->ncols || A->max_row_nnz > kShortRow || A->nnz == 0) return NUM_OK;
    hipStream_t st = (hipStream_t)stream;
    int32_t* d_bad = nullptr;
    NUM_HIP(hipMalloc(&d_bad, sizeof(int32_t)));
    int32_t h_bad = 0;
    hipError_t e = hipMemcpyAsync(d_bad, &h_bad, sizeof(int32_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
        rel16_kernel<<<(unsigned)((A->nrows + 255) / 256), 256, 0, st>>>(
            A->nrows, A->rowptr, A->colind, d_rel16, d_bad);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(&h_bad, d_bad, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(d_bad);
    if (e != hipSuccess) return fail(NUM_EHIP, "rel16 build failed: %s", hipGetErrorString(e));
    if (h_bad == 0) {
        A->rel16 = d_rel16;
        *attached = 1;
        return build_relpat(A, st);
    }
    return NUM_OK;
}

int num_csr_attach_d16(num_csr* A, uint16_t* d_codes, void* stream, int32_t* attached)
{
    if (!A || !attached) return fail(NUM_EINVAL, "matrix or attached is NULL");
    *attached = 0;
    if (!d_codes && A->nnz > 0) return fail(NUM_EINVAL, "d_codes is NULL");
    if ((uintptr_t)d_codes % 16 != 0) return fail(NUM_EINVAL, "d_codes must be 16-byte aligned");
    if ((uintptr_t)A->vals % 16 != 0) return NUM_OK;  // the stage loads values in 16-byte units
    int rc = check_device(A);
    if (rc) return rc;
    A->d16_buf = d_codes;
    rc = build_d16(A, (hipStream_t)stream);
    if (rc) {
        A->d16_buf = nullptr;
        return rc;
    }
    if (!A->d16) A->d16_buf = nullptr;
    rc = build_xw(A, (hipStream_t)stream);
    if (rc) return rc;
    *attached = A->d16 ? 1 : 0;
    return NUM_OK;
}

int num_csr_set_windows(num_csr* A, int32_t mode, void* stream)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    int rc = check_device(A);
    if (rc) return rc;
    if (mode < -1 || mode > 2)
        return fail(NUM_EINVAL, "mode must be -1, 0, 1 or 2, got %d", mode);
    A->xw_mode = mode == 2 ? 1 : mode;
    A->pat_mode = mode == 2 ? 0 : -1;
    return build_xw(A, (hipStream_t)stream);
}

int num_csr_window_info(const num_csr* A, int32_t* windowed)
{
    if (!A || !windowed) return fail(NUM_EINVAL, "matrix or windowed is NULL");
    *windowed = A->xw ? (A->sell_pat ? 2 : 1) : 0;
    return NUM_OK;
}

int num_csr_pattern_info(const num_csr* A, int32_t* windows, int32_t* offsets)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (windows) *windows = A->sell_pat ? 1 : 0;
    if (offsets) *offsets = A->rel_pat ? 1 : 0;
    return NUM_OK;
}

int num_csr_index_info(const num_csr* A, int32_t* delta16)
{
    if (!A || !delta16) return fail(NUM_EINVAL, "matrix or delta16 is NULL");
    *delta16 = A->d16 ? 1 : 0;
    return NUM_OK;
}

int num_csr_attach_xcd(num_csr* A, const num_csr* S, int32_t nblocks,
                       const int32_t* h_block_seg_start, const int32_t* d_row_seg_ptr,
                       const int32_t* d_row_seg_idx)
{
    if (!A || !S || !h_block_seg_start) return fail(NUM_EINVAL, "NULL argument");
    if (nblocks < 1) return fail(NUM_EINVAL, "nblocks must be >= 1");
    if (S->ncols != A->ncols || S->nnz != A->nnz)
        return fail(NUM_EINVAL, "segment matrix does not match A (ncols / nnz)");
    if (h_block_seg_start[0] != 0 || h_block_seg_start[nblocks] != S->nrows)
        return fail(NUM_EINVAL, "block segment starts must run from 0 to the segment count");
    for (int32_t j = 0; j < nblocks; ++j)
        if (h_block_seg_start[j] > h_block_seg_start[j + 1])
            return fail(NUM_EINVAL, "block segment starts must be non-decreasing");
    if (!d_row_seg_ptr || (S->nrows > 0 && !d_row_seg_idx))
        return fail(NUM_EINVAL, "row -> segment map is NULL");
    int rc = check_device(A);
    if (rc) return rc;
    // the segment tiles of pass 1: up to kSegRows segments and kSegNnz
    // entries each (the dense-gather walk keeps every lane busy whatever the
    // segment lengths, and a wave sums a long segment's products from LDS),
    // a segment longer than a tile holds alone
    Schedule seg;
    rc = build_schedule(S, 0, S->nrows, kSegRows, kSegNnz, 1, 1, 1, seg, kSegRows, kSegNnz);
    if (rc) return rc;
    const int32_t nt = seg.nblocks;
    std::vector<TileEntry> tiles((size_t)nt + 1);
    if (nt > 0) {
        const hipError_t e = hipMemcpy(tiles.data(), seg.d_tiles, tiles.size() * sizeof(TileEntry),
                                       hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            free_schedule(seg);
            return fail(NUM_EHIP, "segment schedule read-back failed: %s", hipGetErrorString(e));
        }
    }
    std::vector<int32_t> group[kNumXcd];
    int32_t j = 0;
    for (int32_t t = 0; t < nt; ++t) {
        while (j + 1 < nblocks && tiles[t].row >= h_block_seg_start[j + 1]) ++j;
        group[j % kNumXcd].push_back(t);
    }
    std::vector<int32_t> flat;
    flat.reserve((size_t)nt);
    int32_t off[kNumXcd + 1];
    off[0] = 0;
    for (int k = 0; k < kNumXcd; ++k) {
        flat.insert(flat.end(), group[k].begin(), group[k].end());
        off[k + 1] = (int32_t)flat.size();
    }
    int32_t* d_tiles = nullptr;
    if (nt > 0) {
        hipError_t e = hipMalloc(&d_tiles, flat.size() * sizeof(int32_t));
        if (e == hipSuccess)
            e = hipMemcpy(d_tiles, flat.data(), flat.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (d_tiles) (void)hipFree(d_tiles);
            free_schedule(seg);
            return fail(NUM_EHIP, "XCD tile list upload failed: %s", hipGetErrorString(e));
        }
    }
    if (A->xcd_tiles) (void)hipFree(A->xcd_tiles);
    A->xcd_tiles = d_tiles;
    free_schedule(A->xcd_sched);
    A->xcd_sched = seg;
    std::copy(off, off + kNumXcd + 1, A->xcd_off);
    A->xcd_row_seg_ptr = d_row_seg_ptr;
    A->xcd_row_seg_idx = d_row_seg_idx;
    A->xcd_seg = S;
    return NUM_OK;
}

int num_csr_destroy(num_csr* A)
{
    if (!A) return NUM_OK;
    if (A->engine_holds.load() != 0)
        return fail(NUM_ERUNTIME, "matrix is held by %d engine system(s); remove them first",
                    (int)A->engine_holds.load());
    if (A->xcd_tiles) (void)hipFree(A->xcd_tiles);
    free_windows(A);
    free_relpat(A);
    free_schedule(A->xcd_sched);
    free_schedule(A->sched);
    for (auto& kv : A->range_sched) free_schedule(kv.second);
    if (A->owns) {
        if (A->rowptr) (void)hipFree(const_cast<int32_t*>(A->rowptr));
        if (A->colind) (void)hipFree(const_cast<int32_t*>(A->colind));
        if (A->vals) (void)hipFree(const_cast<double*>(A->vals));
    }
    delete A;
    return NUM_OK;
}

int num_csr_shape(const num_csr* A, int32_t* nrows, int32_t* ncols, int64_t* nnz)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (nrows) *nrows = A->nrows;
    if (ncols) *ncols = A->ncols;
    if (nnz) *nnz = A->nnz;
    return NUM_OK;
}

int num_csr_set_schedule(num_csr* A, int32_t rows_per_block, int32_t nnz_per_block,
                         int32_t chunk, int32_t nt64, int32_t nt32)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    // a live engine system expands its launch lists from the current tiles
    if (A->engine_holds.load() != 0)
        return fail(NUM_ERUNTIME, "matrix is held by %d engine system(s); its schedule cannot "
                    "change until they are removed", (int)A->engine_holds.load());
    if (rows_per_block <= 0) rows_per_block = kSpmvBS;
    if (nnz_per_block <= 0) nnz_per_block = default_nnz_per_block(A);
    int rc = build_schedule(A, 0, A->nrows, rows_per_block, nnz_per_block, chunk, nt64, nt32,
                            A->sched);
    if (rc) return rc;
    rc = build_d16(A, nullptr);  // the codes depend on the tiles
    if (rc) return rc;
    if (!A->d16) A->d16_buf = nullptr;
    rc = build_xw(A, nullptr);  // and so do the windows
    if (rc) return rc;
    rc = build_relpat(A, nullptr);  // and the offset patterns (per tile pair)
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(A->mu);
    for (auto& kv : A->range_sched) free_schedule(kv.second);
    A->range_sched.clear();
    return NUM_OK;
}

int num_csr_schedule_info(const num_csr* A, int32_t* nblocks, int32_t* rows_per_block,
                          int32_t* nnz_per_block, int32_t* long_rows, int32_t* max_row_nnz,
                          int32_t* chunk, int32_t* nt64, int32_t* nt32)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (nblocks) *nblocks = A->sched.nblocks;
    if (rows_per_block) *rows_per_block = A->sched.rows_per_block;
    if (nnz_per_block) *nnz_per_block = A->sched.nnz_per_block;
    if (long_rows) *long_rows = A->sched.long_rows;
    if (max_row_nnz) *max_row_nnz = A->max_row_nnz;
    if (chunk) *chunk = A->sched.chunk;
    if (nt64) *nt64 = A->sched.nt64;
    if (nt32) *nt32 = A->sched.nt32;
    return NUM_OK;
}

}  // extern "C"
