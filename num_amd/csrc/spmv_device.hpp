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
//   2. walk: a row of up to kShortRow (32) entries is walked by one thread
//      (consecutive threads own consecutive rows, so for stencil rows the
//      gathers of one neighbour offset are contiguous across the wave) and
//      summed in index order from 0.0 with separately rounded products, the
//      arithmetic of the reference's row loop (src/spmv_kernel.hpp:90-97):
//      bit-identical results.  A longer row of the tile gets a whole wave
//      (64 lanes over its entries, fixed butterfly), so a power-law tile with
//      a few medium rows among hundreds of short ones stays balanced.
//   3. epilogue: a functor consumes (row, value) -- store y, or the fused CG
//      updates -- and can contribute to a deterministic per-tile partial sum.
// A row longer than kLongRow entries (or than the LDS budget) gets a
// workgroup of its own ("long" tile) that accumulates straight from global
// memory at stride BS and reduces with a fixed butterfly.
// Tiles are dealt to the 8 XCDs in chunks of consecutive tiles (common.hpp:
// chunk_tile), so each XCD's L2 sees neighbouring x windows.
#pragma once
#include <type_traits>
#include <utility>

#include "common.hpp"

namespace numamd {

// Optional per-workgroup hook of an x accessor (the mixed CG step kernels
// compute their scalars from the previous kernel's partial sums inside every
// workgroup): issue() runs right after the stage loads are issued (its loads
// overlap the stage latency), resolve() after the stage barrier, before any
// gather; when it returns false the workgroup walks no row.  Both are const
// and keep what they compute in mutable members.
// Optional epilogue entry of the windowed walk (spmv_xw_run): own(row, v,
// ca, cb, xr) receives the accessor's value at the row's own column, read
// from the tile's windows instead of from memory.
template <class E, class = void>
struct has_own : std::false_type {};
template <class E>
struct has_own<E, std::void_t<decltype(std::declval<const E&>().own(
                      0, 0.0, std::declval<double&>(), std::declval<double&>(), 0.0))>>
    : std::true_type {};

template <class G, class = void>
struct has_hook : std::false_type {};
template <class G>
struct has_hook<G, std::void_t<decltype(std::declval<const G&>().resolve())>>
    : std::true_type {};

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

// Body for one tile given its schedule entry and the next one.  Leaves this
// thread's epilogue contributions in ca/cb (zero for threads that finish no
// row).
template <typename T, int BS, int NPT, bool NT, class Epi, class Gather>
__device__ __forceinline__ void spmv_tile_run(
    const TileEntry e0, const TileEntry e1, const int32_t* __restrict__ rowptr,
    const int32_t* __restrict__ colind, const T* __restrict__ vals, const Gather& x,
    const Epi& epi, T* __restrict__ lv, int32_t* __restrict__ lc, double& ca, double& cb)
{
    constexpr int NNZ_MAX = BS * NPT;
    constexpr int NW = BS / kWave;
    const int t = threadIdx.x;
    const int r0 = e0.row, j0 = e0.nz0;
    const int r1 = e1.row, j1 = e1.nz0;
    const int nz = j1 - j0;
    ca = 0.0;
    cb = 0.0;
    if (e0.wide == kWalkLong || nz > NNZ_MAX) {
        // long row: one row, whole workgroup; 8 independent loads in flight
        // per thread (clamped, no per-element branches), then a fixed-shape
        // tree (wave butterfly, waves in order).
        if constexpr (has_hook<Gather>::value) {
            x.issue();
            if (!x.resolve()) return;
        }
        T acc = T(0);
        const int last = j1 - 1;
        for (int j = j0 + t; j < j1; j += 8 * BS) {
            T v[8], xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int jj = min(j + u * BS, last);
                v[u] = ld_mat<NT>(vals + jj);
                xv[u] = x(ld_mat<NT>(colind + jj));
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j + u * BS < j1) acc += v[u] * xv[u];
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if ((t & (kWave - 1)) == 0) lv[t / kWave] = acc;
        __syncthreads();
        if (t == 0) {
            T s = T(0);
#pragma unroll
            for (int w = 0; w < NW; ++w) s += lv[w];
            epi(r0, s, ca, cb, x);
        }
        return;
    }
    // row bounds for the walk phase, fetched together with the stage loads:
    // branch-free (a thread past the tile's rows reads rowptr[r1] twice) so
    // the loads stay in flight while the stage loads are issued
    const int nr = r1 - r0;
    const bool active = t < nr;
    const int tr = min(t, nr);
    const int rb_raw = rowptr[r0 + tr];
    const int re_raw = rowptr[r0 + tr + (active ? 1 : 0)];
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
        if constexpr (has_hook<Gather>::value) x.issue();
#pragma unroll
        for (int s = 0; s < NPT; ++s) {
            const int k = t + s * BS;
            if (k < nz) {
                lv[k] = v[s];
                lc[k] = c[s];
            }
        }
    } else if constexpr (has_hook<Gather>::value) {
        x.issue();
    }
    __syncthreads();
    if constexpr (has_hook<Gather>::value) {
        if (!x.resolve()) return;
    }
    const int rb = rb_raw - j0, re = re_raw - j0;
    const bool wide = e0.wide != 0;  // uniform over the workgroup
    // 2a. short rows: one thread per row, entries in index order from 0, 8
    //     gathers in flight before the (ordered) sums
    if (active && (!wide || re - rb <= kShortRow)) {
        T acc = T(0);
        int k = rb;
        for (; k + 7 < re; k += 8) {
            T vv[8], xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                vv[u] = lv[k + u];
                xv[u] = x(lc[k + u]);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += vv[u] * xv[u];
        }
        T vv[7], xv[7];
#pragma unroll
        for (int u = 0; u < 7; ++u) {
            const int kk = k + u;
            const bool ok = kk < re;
            const int ks = ok ? kk : 0;  // slot 0 is always inside the LDS array
            vv[u] = lv[ks];
            const int c = lc[ks];
            xv[u] = x(ok ? c : 0);       // column 0 always exists
        }
#pragma unroll
        for (int u = 0; u < 7; ++u)
            if (k + u < re) acc += vv[u] * xv[u];
        epi(r0 + t, acc, ca, cb, x);
    }
    if (!wide) return;
    // 2b. rows longer than kShortRow: the m-th such row of the tile (in row
    //     order) goes to wave m % NW, whose 64 lanes stride over its entries;
    //     a fixed butterfly sums the lanes.
    __shared__ unsigned long long wmask[NW];
    const unsigned long long bal = __ballot(active && re - rb > kShortRow);
    if ((t & (kWave - 1)) == 0) wmask[t / kWave] = bal;
    __syncthreads();
    const int w = t / kWave, l = t & (kWave - 1);
    int m = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        unsigned long long bits = wmask[i];
        while (bits) {
            const int bpos = __builtin_ctzll(bits);
            bits &= bits - 1;
            if ((m++ % NW) != w) continue;
            const int row = i * kWave + bpos;
            const int mb = rowptr[r0 + row] - j0, me = rowptr[r0 + row + 1] - j0;
            T acc = T(0);
            int k = mb + l;
            for (; k + 3 * kWave < me; k += 4 * kWave) {
                T vv[4], xv[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    vv[u] = lv[k + u * kWave];
                    xv[u] = x(lc[k + u * kWave]);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) acc += vv[u] * xv[u];
            }
            for (; k < me; k += kWave) acc += lv[k] * x(lc[k]);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
            if (l == 0) {
                double ta = 0.0, tb = 0.0;
                epi(r0 + row, acc, ta, tb, x);
                ca += ta;
                cb += tb;
            }
        }
    }
}

// FP32 walk over a group of consecutive tiles whose rows all hold <= kShortRow
// entries, with column ids stored as int16 offsets from their row (half the
// index bytes of int32).  LDS: lv / lc hold RPT * NPT * BS entries.  Thread t owns rows t, t + BS, ..., t + (RPT-1) BS of
// the group; the group's entries (<= RPT * NPT * BS) are staged at once, so
// one workgroup moves RPT tiles' worth of data behind one latency chain.
// Each row is summed in index order from 0.0 (the row loop's arithmetic).
template <int BS, int NPT, int RPT, bool NT, class Epi, class Gather>
__device__ __forceinline__ void spmv_rel16_run(
    const TileEntry e0, const TileEntry e1, const int32_t* __restrict__ rowptr,
    const int16_t* __restrict__ rel, const float* __restrict__ vals, const Gather& x,
    const Epi& epi, float* __restrict__ lv, int16_t* __restrict__ lc, double& ca, double& cb)
{
    constexpr int M = NPT * RPT;
    const int t = threadIdx.x;
    const int r0 = e0.row, j0 = e0.nz0;
    const int r1 = e1.row, j1 = e1.nz0;
    const int nz = j1 - j0, nr = r1 - r0;
    int rb[RPT], re[RPT];  // raw rowptr values, branch-free loads (see spmv_tile_run)
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int lr = min(t + k * BS, nr);
        rb[k] = rowptr[r0 + lr];
        re[k] = rowptr[r0 + lr + (t + k * BS < nr ? 1 : 0)];
    }
    // stage: each thread issues all M value and offset loads (clamped, no
    // per-element branches) before parking them in LDS
    if (nz > 0) {
        const int last = j1 - 1;
        float v[M];
        int16_t c[M];
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const int j = min(j0 + t + s * BS, last);
            v[s] = ld_mat<NT>(vals + j);
            c[s] = ld_mat<NT>(rel + j);
        }
        if constexpr (has_hook<Gather>::value) x.issue();
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const int k = t + s * BS;
            if (k < nz) {
                lv[k] = v[s];
                lc[k] = c[s];
            }
        }
    } else if constexpr (has_hook<Gather>::value) {
        x.issue();
    }
    __syncthreads();
    ca = 0.0;
    cb = 0.0;
    if constexpr (has_hook<Gather>::value) {
        if (!x.resolve()) return;
    }
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int lr = t + k * BS;
        if (lr >= nr) continue;
        const int row = r0 + lr;
        const int b = rb[k] - j0, e = re[k] - j0;
        float acc = 0.0f;
        int q = b;
        for (; q + 7 < e; q += 8) {
            float vv[8], xv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                vv[u] = lv[q + u];
                xv[u] = x(row + lc[q + u]);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += vv[u] * xv[u];
        }
        float vv[7], xv[7];
#pragma unroll
        for (int u = 0; u < 7; ++u) {
            const int qq = q + u;
            const bool ok = qq < e;
            const int qs = ok ? qq : 0;  // slot 0 is always inside the LDS array
            vv[u] = lv[qs];
            const int c = lc[qs];
            xv[u] = x(row + (ok ? c : 0));  // the row's own column exists (square)
        }
#pragma unroll
        for (int u = 0; u < 7; ++u)
            if (q + u < e) acc += vv[u] * xv[u];
        double ta = 0.0, tb = 0.0;
        epi(row, acc, ta, tb, x);
        ca += ta;
        cb += tb;
    }
}

// spmv_rel16_run with the offsets' row patterns (num_csr_attach_rel16,
// kRelPat): only the values are staged; each row takes its pattern id (one
// byte), its length from the table's length words (loaded up front, uniform)
// and, after the stage barrier, its offsets from the table in LDS (`lpat`,
// staged with the values).  No rowptr: a row's first entry is the tile's
// first (e0.nz0) plus the lengths of the rows before it, a workgroup scan of
// the lengths packed 16 bits per row slot (row t + k*BS in field k), its
// wave totals published before the stage barrier.  Same entries, same
// order, same bits as spmv_rel16_run.  Rows hold at most kRelPatW entries.
// pat: NPW pattern words (at most BS) then relpat_len_words<NPW>() length
// words.  T = double walks a short-row FP64 tile with the same sums as
// spmv_tile_run.
template <int BS, int NPT, int RPT, bool NT, int NPW, typename T, class Epi, class Gather>
__device__ __forceinline__ void spmv_relpat_run(
    const TileEntry e0, const TileEntry e1, const uint8_t* __restrict__ pid,
    const uint64_t* __restrict__ pat, const T* __restrict__ vals, const Gather& x,
    const Epi& epi, T* __restrict__ lv, uint64_t* __restrict__ lpat, double& ca, double& cb)
{
    constexpr int M = NPT * RPT;
    constexpr int NL = relpat_len_words<NPW>();
    constexpr int NW = BS / kWave;
    static_assert(NPW <= BS, "one table word per thread");
    static_assert(RPT <= 2 && BS * kRelPatW < 65536, "row starts packed 16 bits per row slot");
    __shared__ int lscan[NW];
    const int t = threadIdx.x;
    const int r0 = e0.row, j0 = e0.nz0;
    const int r1 = e1.row, j1 = e1.nz0;
    const int nz = j1 - j0, nr = r1 - r0;
    uint64_t lw[NL];
#pragma unroll
    for (int w = 0; w < NL; ++w) lw[w] = pat[NPW + w];
    int id[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) id[k] = t + k * BS < nr ? (int)pid[r0 + t + k * BS] : 0;
    uint64_t tab = 0;
    if (t < NPW) tab = pat[t];
    if (nz > 0) {
        const int last = j1 - 1;
        T v[M];
#pragma unroll
        for (int s = 0; s < M; ++s) v[s] = ld_mat<NT>(vals + min(j0 + t + s * BS, last));
        if constexpr (has_hook<Gather>::value) x.issue();
#pragma unroll
        for (int s = 0; s < M; ++s) {
            const int k = t + s * BS;
            if (k < nz) lv[k] = v[s];
        }
    } else if constexpr (has_hook<Gather>::value) {
        x.issue();
    }
    int len[RPT], pk = 0;
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        uint64_t w = lw[0];
#pragma unroll
        for (int u = 1; u < NL; ++u)
            if ((id[k] >> 4) == u) w = lw[u];
        len[k] = t + k * BS < nr ? (int)((w >> (4 * (id[k] & 15))) & 15) : 0;
        pk |= len[k] << (16 * k);
    }
    const int lane = t & (kWave - 1);
    int inc = pk;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int u = __shfl_up(inc, o, kWave);
        if (lane >= o) inc += u;
    }
    if (lane == kWave - 1) lscan[t / kWave] = inc;
    if (t < NPW) lpat[t] = tab;
    __syncthreads();
    ca = 0.0;
    cb = 0.0;
    if constexpr (has_hook<Gather>::value) {
        if (!x.resolve()) return;
    }
    int before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const int sw = lscan[w];
        if (w < t / kWave) before += sw;
        tot += sw;
    }
    const int ex = before + inc - pk;  // exclusive prefix of every field
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int lr = t + k * BS;
        if (lr >= nr) continue;
        const int row = r0 + lr;
        // fields below k hold whole earlier row slots
        int b = (ex >> (16 * k)) & 0xffff;
#pragma unroll
        for (int f = 0; f < k; ++f) b += (tot >> (16 * f)) & 0xffff;
        const int n = len[k];
        const uint64_t q0 = lpat[2 * id[k]], q1 = lpat[2 * id[k] + 1];
        T vv[kRelPatW], xv[kRelPatW];
#pragma unroll
        for (int u = 0; u < kRelPatW; ++u) {
            const bool ok = u < n;
            vv[u] = lv[ok ? b + u : 0];  // slot 0 is always inside the LDS array
            const int off = (int)(int16_t)(uint16_t)(((u < 4 ? q0 : q1) >> (16 * (u & 3))) & 0xffff);
            xv[u] = x(row + (ok ? off : 0));  // zero offset past the row: its own column
        }
        T acc = T(0);
#pragma unroll
        for (int u = 0; u < kRelPatW; ++u)
            if (u < n) acc += vv[u] * xv[u];
        double ta = 0.0, tb = 0.0;
        epi(row, acc, ta, tb, x);
        ca += ta;
        cb += tb;
    }
}

// FP64 walk with delta-coded column ids (num_csr_attach_d16): a row's first
// code is its column minus the tile's base column (TileEntry.base), every
// later code its column minus the previous entry's, as uint16.  Half the
// index bytes of int32 (192^3 Laplacian: 98.6 of its 733 MB).  Every row of
// the matrix holds <= kShortRow entries in ascending column order, so every
// tile is walked one thread per row, entries summed in index order from 0.0
// as in spmv_tile_run (bit-identical); the running column is rebuilt by
// integer adds.  Stage: the values and codes go straight from memory to LDS
// in 16-byte units (global_load_lds_dwordx4: no VGPR round trip, no LDS
// store instructions; 105.6 -> 104.7 us on 192^3), values from the even
// entry jb <= j0, codes from the entry jc <= j0 that is a multiple of 8
// (the code array is padded to nnz + 8, and 16-byte aligned like the
// values); lanes past the last whole unit are masked off, and an odd last
// value is loaded by one thread, so no value past the tile is read.  Row
// starts come from one length byte per row: a workgroup scan after the
// stage barrier.
template <int BS, int NPT>
constexpr int d16_pairs() { return (NPT + 2) / 2; }  // 16-byte value units per thread
template <int BS, int NPT>
constexpr int d16_code_rounds() { return ((BS * NPT + 14) / 8 + BS - 1) / BS; }
// LDS sized to what a tile can stage, not to the loop capacity: a tile holds
// at most BS*NPT entries, so values from jb = j0 & ~1 span <= BS*NPT + 1
// slots and codes from jc = j0 & ~7 span <= (BS*NPT + 14) / 8 whole units.
// At BS = 256, NPT = 7 this is 17984 bytes against 20512, which lets 8
// workgroups share a CU instead of 7 (192^3: 104.5 -> 102.0 us).
template <int BS, int NPT>
constexpr int d16_value_slots() { return ((BS * NPT + 1) / 2) * 2 + 2; }
template <int BS, int NPT>
constexpr int d16_code_slots() { return ((BS * NPT + 14) / 8 + 1) * 8; }

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void glob_void_t;

template <int BS, int NPT, bool NT, class Epi, class Gather>
__device__ __forceinline__ void spmv_d16_run(
    const TileEntry e0, const TileEntry e1, const uint8_t* __restrict__ lens,
    const uint16_t* __restrict__ code, const double* __restrict__ vals, const Gather& x,
    const Epi& epi, double* __restrict__ lvu, uint16_t* __restrict__ lcu,
    double& ca, double& cb)
{
    constexpr int NPP = d16_pairs<BS, NPT>();
    constexpr int NPC = d16_code_rounds<BS, NPT>();
    constexpr int NW = BS / kWave;
    __shared__ int wtot[NW];
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int w = t / kWave;
    const int r0 = e0.row, j0 = e0.nz0;
    const int r1 = e1.row, j1 = e1.nz0;
    const int nz = j1 - j0, nr = r1 - r0;
    ca = 0.0;
    cb = 0.0;
    const bool active = t < nr;
    const uint8_t len_raw = lens[r0 + min(t, nr - 1)];
    int len = 0, incl = 0;
    auto scan = [&]() {
        len = active ? (int)len_raw : 0;
        incl = len;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
            const int y = __shfl_up(incl, d, kWave);
            if (lane >= d) incl += y;
        }
        if (lane == kWave - 1) wtot[w] = incl;
    };
    const int jb = j0 & ~1, off = j0 - jb;    // values: 16-byte units from jb
    const int jc = j0 & ~7, offc = j0 - jc;   // codes: 16-byte units (8 codes) from jc
    constexpr unsigned kPol = NT ? 2u : 0u;
    if (nz > 0) {
        // whole 16-byte value units only (lanes past them are masked off, so
        // they write no LDS); an odd last value goes through one thread, so
        // nothing past the tile's last value is read
        const int nu = (j1 - jb) >> 1;
        const int nc = ((j1 - 1 - jc) >> 3) + 1;  // code units (may read past j1: padding)
#pragma unroll
        for (int s = 0; s < NPP; ++s) {
            const int ub = s * BS + w * kWave;  // wave-uniform
            if (ub + lane < nu)
                __builtin_amdgcn_global_load_lds((glob_void_t*)(vals + jb + 2 * (ub + lane)),
                                                 (lds_void_t*)(lvu + 2 * ub), 16, 0, kPol);
        }
#pragma unroll
        for (int s = 0; s < NPC; ++s) {
            const int ub = s * BS + w * kWave;
            if (ub + lane < nc)
                __builtin_amdgcn_global_load_lds((glob_void_t*)(code + jc + 8 * (ub + lane)),
                                                 (lds_void_t*)(lcu + 8 * ub), 16, 0, kPol);
        }
        if (((j1 - jb) & 1) && t == 0) lvu[2 * nu] = ld_mat<NT>(vals + j1 - 1);
    }
    if constexpr (has_hook<Gather>::value) x.issue();
    // The direct-to-LDS loads are counted by vmcnt, and s_barrier does not
    // wait for them: each wave must drain its own loads before the barrier,
    // or another wave may read LDS that is not filled yet.  Explicit, so it
    // does not depend on where the compiler places its own waits.
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    // the stage barrier then covers every load; the row-length scan follows it
    __syncthreads();
    scan();
    __syncthreads();
    if constexpr (has_hook<Gather>::value) {
        if (!x.resolve()) return;
    }
    const uint16_t* __restrict__ lc = lcu + offc;
    const double* __restrict__ lv = lvu + off;
    if (!active) return;
    int before = 0;
#pragma unroll
    for (int v = 0; v < NW; ++v)
        if (v < w) before += wtot[v];
    const int rb = before + incl - len, re = rb + len;
    double acc = 0.0;
    int col = e0.base;
    int k = rb;
    for (; k + 7 < re; k += 8) {
        double vv[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            vv[u] = lv[k + u];
            col += lc[k + u];
            xv[u] = x(col);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += vv[u] * xv[u];
    }
    double vv[7], xv[7];
#pragma unroll
    for (int u = 0; u < 7; ++u) {
        const int kk = k + u;
        const bool ok = kk < re;
        const int ks = ok ? kk : 0;
        vv[u] = lv[ks];
        col += ok ? (int)lc[ks] : 0;
        xv[u] = x(col);
    }
#pragma unroll
    for (int u = 0; u < 7; ++u)
        if (k + u < re) acc += vv[u] * xv[u];
    epi(r0 + t, acc, ca, cb, x);
}

// A tile's window record as scalars passed by value: the four windows'
// first columns (wl0..wl3) and cumulative slot ends (we0..we3).  No array or
// object whose address is taken: the compiler may move such a private
// variable to LDS, and every address computation would then wait on an LDS
// read.
static_assert(kXwWin == 4, "the window helpers take four windows");
#define NUMAMD_XW_LOAD(rec)                                                      \
    const int wl0 = (rec)[0], wl1 = (rec)[1], wl2 = (rec)[2], wl3 = (rec)[3];    \
    const int we0 = (rec)[4], we1 = (rec)[5], we2 = (rec)[6], we3 = (rec)[7]
// column of slot s: slot s of window q reads column l_q + s - e_(q-1)
__device__ __forceinline__ int xw_col(int s, int l0, int l1, int l2, int l3, int e0, int e1,
                                      int e2)
{
    int lo = l0, start = 0;
    if (s >= e0) {
        lo = l1;
        start = e0;
    }
    if (s >= e1) {
        lo = l2;
        start = e1;
    }
    if (s >= e2) {
        lo = l3;
        start = e2;
    }
    return lo + s - start;
}
// slot of column c, or -1 when no window holds it
__device__ __forceinline__ int xw_slot(int c, int l0, int l1, int l2, int l3, int e0, int e1,
                                       int e2, int e3)
{
    int s = -1;
    if (c >= l0 && c - l0 < e0) s = c - l0;
    if (c >= l1 && c - l1 < e1 - e0) s = e0 + c - l1;
    if (c >= l2 && c - l2 < e2 - e1) s = e1 + c - l2;
    if (c >= l3 && c - l3 < e3 - e2) s = e2 + c - l3;
    return s;
}

// A row's window slots from its tile's packed slot region p (sell_pos with
// four per group): slots 0-3 in q0, 4-7 in q1, 16 bits each.  At most three
// loads (8 + 4 + 2 bytes per lane), uniform branches on the width.
template <bool NT, int BS>
__device__ __forceinline__ void sell_slots(const uint16_t* __restrict__ p, int w, int t,
                                           uint64_t& q0, uint64_t& q1)
{
    const int n4 = w >> 2, r = w & 3;
    uint64_t g0 = 0, g1 = 0, tl = 0;
    if (n4 >= 1) g0 = ld_mat<NT>(reinterpret_cast<const uint64_t*>(p) + t);
    if (n4 >= 2) g1 = ld_mat<NT>(reinterpret_cast<const uint64_t*>(p + 4 * BS) + t);
    const uint16_t* __restrict__ pr = p + n4 * 4 * BS;
    if (r >= 2) tl = ld_mat<NT>(reinterpret_cast<const uint32_t*>(pr) + t);
    if (r & 1)
        tl |= (uint64_t)ld_mat<NT>(pr + (r >= 2 ? 2 * BS : 0) + t) << (r >= 2 ? 32 : 0);
    q0 = n4 >= 1 ? g0 : tl;
    q1 = n4 >= 2 ? g1 : (n4 == 1 ? tl : 0);
}

// Slot k of a row from sell_slots' q0 / q1.
__device__ __forceinline__ int sell_slot(int k, uint64_t q0, uint64_t q1)
{
    return (int)(((k < 4 ? q0 : q1) >> (16 * (k & 3))) & 0xffffu);
}

// Windowed sliced walk (the layout of num_csr_set_windows).  The tile's
// entries come in sliced order -- entry k of the tile's row i at
// (base + k) * BS + i, its window slot packed with the row's others
// (sell_pos) -- so thread i loads its own row's values and slots straight
// into registers, every load instruction reading BS consecutive elements:
// coalesced without staging the tile in LDS.  In the
// same batch every thread fetches its share of the tile's column windows
// through the x accessor (x[c], or the CG's r[c] + beta*p_old[c]) and parks
// them in LDS, the only LDS the walk uses; after one barrier each row is
// summed from its registers and the LDS windows.  One memory round trip per
// tile, a stencil's neighbour values fetched once per tile instead of once
// per entry.  Entries are summed in index order from 0.0 with separately
// rounded products: the bits of spmv_tile_run and spmv_d16_run.
// With PAT the row's slots and length come from its tile's pattern table
// (a 4-bit id per row, kXwPat; `pid` is the tile's kSpmvBS / 2 id bytes)
// instead of 2 bytes per entry and a length byte: slot = pattern value -
// kXwPatBias + lane (low 12 bits), length = bits 12-15 of the first.
template <int BS, bool NT, bool PAT, class Epi, class Gather>
__device__ __forceinline__ void spmv_sell_run(
    const TileEntry e0, const TileEntry e1, const int32_t* __restrict__ rec,
    const uint8_t* __restrict__ lens, const uint16_t* __restrict__ code,
    const uint64_t* __restrict__ pat, const uint8_t* __restrict__ pid,
    const double* __restrict__ vals, const Gather& x, const Epi& epi, double* __restrict__ xw,
    uint64_t* __restrict__ lpat, double& ca, double& cb)
{
    constexpr int NXW = (kXwSlots + BS - 1) / BS;
    static_assert(!has_hook<Gather>::value, "window fill has no hook point");
    const int t = threadIdx.x;
    const int r0 = e0.row, nr = e1.row - r0;
    ca = 0.0;
    cb = 0.0;
    const bool active = t < nr;
    // the record is uniform over the workgroup (scalar loads)
    NUMAMD_XW_LOAD(rec);
    const int64_t vb = (int64_t)rec[kXwBase] * BS;
    const int width = rec[kXwWidth];
    const int total = we3;
    int len = 0;
    if constexpr (!PAT) len = active ? (int)lens[r0 + t] : 0;
    // this row's entries: the tile's slices [0, width) (padding included;
    // uniform branches) and its slots, packed
    double v[kSellW];
#pragma unroll
    for (int k = 0; k < kSellW; ++k)
        if (k < width) v[k] = ld_mat<NT>(vals + vb + k * BS + t);
    uint64_t q0 = 0, q1 = 0;
    int id = 0;
    if constexpr (PAT) {
        // the row's pattern id now; the tile's table goes to LDS with the
        // windows and is read after the barrier (no dependent global load)
        id = active ? ((int)ld_mat<NT>(pid + (t >> 1)) >> (4 * (t & 1))) & 15 : 0;
        if (t < 2 * kXwPat) lpat[t] = pat[t];
    } else {
        sell_slots<NT, BS>(code + vb, width, t, q0, q1);
    }
    // window slots (clamped to the last one: also branch-free)
    const int sm = max(total - 1, 0);
    double xv_w[NXW];
#pragma unroll
    for (int u = 0; u < NXW; ++u) {
        const int s = min(t + u * BS, sm);
        xv_w[u] = x(xw_col(s, wl0, wl1, wl2, wl3, we0, we1, we2));
    }
#pragma unroll
    for (int u = 0; u < NXW; ++u) {
        const int s = t + u * BS;
        if (s < total) xw[s] = xv_w[u];
    }
    __syncthreads();
    if (!active) return;
    if constexpr (PAT) {
        q0 = lpat[2 * id];
        q1 = lpat[2 * id + 1];
        len = (int)((q0 >> 12) & 15);
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kSellW; ++k) {
        if constexpr (PAT) {
            // slot 0 for the padding keeps every LDS address inside the windows
            const int sl = k < len ? (sell_slot(k, q0, q1) & 0xfff) - kXwPatBias + t : 0;
            if (k < len) acc += v[k] * xw[sl];
        } else {
            if (k < len) acc += v[k] * xw[sell_slot(k, q0, q1)];
        }
    }
    const int row = r0 + t;
    if constexpr (has_own<Epi>::value) {
        // the own rows are a window range of every tile of a square matrix
        const int os = xw_slot(row, wl0, wl1, wl2, wl3, we0, we1, we2, we3);
        if (os >= 0) {
            epi.own(row, acc, ca, cb, xw[os]);
            return;
        }
    }
    epi(row, acc, ca, cb, x);
}

template <typename T, int BS, int NPT, bool NT, class Epi, class Gather>
__device__ __forceinline__ void spmv_tile_body(
    int tile, const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const T* __restrict__ vals, const Gather& x, const TileEntry* __restrict__ sched,
    const Epi& epi, T* __restrict__ lv, int32_t* __restrict__ lc, double& ca, double& cb)
{
    spmv_tile_run<T, BS, NPT, NT, Epi, Gather>(sched[tile], sched[tile + 1], rowptr, colind,
                                               vals, x, epi, lv, lc, ca, cb);
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

// The same over delta-coded column ids (spmv_d16_run).
template <int BS, int NPT, bool NT, class Epi, class Gather>
__global__ __launch_bounds__(BS) void spmv_d16_kernel(
    const uint8_t* __restrict__ lens, const uint16_t* __restrict__ code,
    const double* __restrict__ vals, Gather x, const TileEntry* __restrict__ sched, int ntiles,
    int chunk, Epi epi)
{
    __shared__ __attribute__((aligned(16))) double lv[d16_value_slots<BS, NPT>()];
    __shared__ __attribute__((aligned(16))) uint16_t lc[d16_code_slots<BS, NPT>()];
    __shared__ double red[2 * BS / kWave];
    const int tile = chunk_tile(blockIdx.x, ntiles, chunk);
    if (tile < 0) return;
    if (epi.skip()) return;
    x.prepare();
    double ca, cb;
    spmv_d16_run<BS, NPT, NT, Epi, Gather>(sched[tile], sched[tile + 1], lens, code, vals, x, epi,
                                           lv, lc, ca, cb);
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

// The windowed sliced walk (spmv_sell_run): LDS for the kXwSlots window
// slots only.
template <int BS, bool NT, bool PAT, class Epi, class Gather>
__global__ __launch_bounds__(BS) void spmv_sell_kernel(
    const int32_t* __restrict__ rec, const uint8_t* __restrict__ lens,
    const uint16_t* __restrict__ code, const uint64_t* __restrict__ pat,
    const int32_t* __restrict__ tix, const uint8_t* __restrict__ pid,
    const double* __restrict__ vals, Gather x, const TileEntry* __restrict__ sched, int ntiles,
    int chunk, Epi epi)
{
    __shared__ double xw[kXwSlots];
    __shared__ double red[2 * BS / kWave];
    __shared__ uint64_t lpat[PAT ? 2 * kXwPat : 1];
    const int tile = chunk_tile(blockIdx.x, ntiles, chunk);
    if (tile < 0) return;
    if (epi.skip()) return;
    x.prepare();
    double ca, cb;
    spmv_sell_run<BS, NT, PAT, Epi, Gather>(
        sched[tile], sched[tile + 1], rec + tile * kXwRec, lens, code,
        PAT ? pat + (size_t)tix[tile] * 2 * kXwPat : pat, PAT ? pid + tile * (BS / 2) : pid, vals,
        x, epi, xw, lpat, ca, cb);
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
        // streamed out: no burst of L2 write-backs at the end of the kernel
        // (192^3: 120.9 -> 118.3 us)
        __builtin_nontemporal_store(v, y + row);
    }
};

}  // namespace numamd
