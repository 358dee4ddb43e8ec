// spmv_launch.hpp -- launch the row-tile kernel over a schedule, dispatching
// on the schedule's LDS capacity (entries per thread).
#pragma once
#include "csr.hpp"
#include "spmv_device.hpp"

namespace numamd {

// FP64 values on the whole-matrix schedule take the delta-coded column ids
// when they are attached (their tile bases live in that schedule) and the
// matrix streams from HBM (NT): the codes save HBM bytes at the cost of a
// longer walk, which a matrix served from the Infinity Cache does not repay
// (128^3 diffusion, 117 MB of FP64 matrix: SpMV 32.6 -> 34.8 us with codes;
// 192^3 Laplacian, 621 MB: 117.9 -> 110.4 us).
template <typename T, int NPT, bool NT, class Epi, class Gather>
inline void launch_tiles_npt(const num_csr* A, const Schedule& s, const T* vals,
                             const Gather& x, const Epi& epi, hipStream_t stream)
{
    if constexpr (sizeof(T) == 8 && NT) {
        if (A->d16 && &s == &A->sched && vals == A->vals) {
            spmv_d16_kernel<kSpmvBS, NPT, NT, Epi, Gather>
                <<<chunk_grid(s.nblocks, s.chunk), kSpmvBS, 0, stream>>>(
                    d16_lens(A->d16, A->nnz), A->d16, vals, x,
                    static_cast<const TileEntry*>(s.d_tiles),
                    s.nblocks, s.chunk, epi);
            return;
        }
    }
    spmv_tile_kernel<T, kSpmvBS, NPT, NT, Epi, Gather>
        <<<chunk_grid(s.nblocks, s.chunk), kSpmvBS, 0, stream>>>(
            A->rowptr, A->colind, vals, x, static_cast<const TileEntry*>(s.d_tiles), s.nblocks,
            s.chunk, epi);
}

template <typename T, bool NT, class Epi, class Gather>
inline void launch_tiles_nt(const num_csr* A, const Schedule& s, const T* vals,
                            const Gather& x, const Epi& epi, hipStream_t stream)
{
    switch (s.npt) {
    case 5: launch_tiles_npt<T, 5, NT>(A, s, vals, x, epi, stream); break;
    case 7: launch_tiles_npt<T, 7, NT>(A, s, vals, x, epi, stream); break;
    case 12: launch_tiles_npt<T, 12, NT>(A, s, vals, x, epi, stream); break;
    default: launch_tiles_npt<T, 16, NT>(A, s, vals, x, epi, stream); break;
    }
}

// Streaming policy: the schedule's choice for this value type (an FP64 and
// an FP32 copy of one pattern can differ: only the smaller may fit the
// Infinity Cache).
template <typename T>
inline bool stream_nt(const Schedule& s)
{
    return sizeof(T) == 8 ? s.nt64 != 0 : s.nt32 != 0;
}

// The windowed sliced layout, when built, replaces every index form for the
// FP64 values on the whole-matrix schedule.
template <typename T, bool NT, class Epi, class Gather>
inline bool launch_sell(const num_csr* A, const Schedule& s, const T* vals, const Gather& x,
                        const Epi& epi, hipStream_t stream)
{
    if constexpr (sizeof(T) == 8 && !has_hook<Gather>::value) {
        if (A->xw && &s == &A->sched && vals == A->vals) {
            const TileEntry* tiles = static_cast<const TileEntry*>(s.d_tiles);
            if (A->sell_pat)
                spmv_sell_kernel<kSpmvBS, NT, true, Epi, Gather>
                    <<<chunk_grid(s.nblocks, s.chunk), kSpmvBS, 0, stream>>>(
                        A->xw, A->sell_lens, A->sell_codes, A->sell_pat, A->sell_tix,
                        A->sell_pid, A->sell_vals, x, tiles, s.nblocks, s.chunk, epi);
            else
                spmv_sell_kernel<kSpmvBS, NT, false, Epi, Gather>
                    <<<chunk_grid(s.nblocks, s.chunk), kSpmvBS, 0, stream>>>(
                        A->xw, A->sell_lens, A->sell_codes, nullptr, nullptr, nullptr,
                        A->sell_vals, x, tiles, s.nblocks, s.chunk, epi);
            return true;
        }
    }
    return false;
}

template <typename T, class Epi, class Gather>
inline void launch_tiles(const num_csr* A, const Schedule& s, const T* vals, const Gather& x,
                         const Epi& epi, hipStream_t stream)
{
    if (s.nblocks == 0) return;
    if (stream_nt<T>(s) ? launch_sell<T, true>(A, s, vals, x, epi, stream)
                        : launch_sell<T, false>(A, s, vals, x, epi, stream))
        return;
    if (stream_nt<T>(s)) launch_tiles_nt<T, true>(A, s, vals, x, epi, stream);
    else launch_tiles_nt<T, false>(A, s, vals, x, epi, stream);
}

}  // namespace numamd
