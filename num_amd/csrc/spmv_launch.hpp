// spmv_launch.hpp -- launch the row-tile kernel over a schedule, dispatching
// on the schedule's LDS capacity (entries per thread).
#pragma once
#include "csr.hpp"
#include "spmv_device.hpp"

namespace numamd {

template <typename T, int NPT, class Epi, class Gather>
inline void launch_tiles_npt(const num_csr* A, const Schedule& s, const T* vals,
                             const Gather& x, const Epi& epi, hipStream_t stream)
{
    spmv_tile_kernel<T, kSpmvBS, NPT, Epi, Gather>
        <<<chunk_grid(s.nblocks, s.chunk), kSpmvBS, 0, stream>>>(
            A->rowptr, A->colind, vals, x, static_cast<const TileEntry*>(s.d_tiles), s.nblocks,
            s.chunk, epi);
}

template <typename T, class Epi, class Gather>
inline void launch_tiles(const num_csr* A, const Schedule& s, const T* vals, const Gather& x,
                         const Epi& epi, hipStream_t stream)
{
    if (s.nblocks == 0) return;
    switch (s.npt) {
    case 5: launch_tiles_npt<T, 5>(A, s, vals, x, epi, stream); break;
    case 7: launch_tiles_npt<T, 7>(A, s, vals, x, epi, stream); break;
    case 12: launch_tiles_npt<T, 12>(A, s, vals, x, epi, stream); break;
    default: launch_tiles_npt<T, 16>(A, s, vals, x, epi, stream); break;
    }
}

}  // namespace numamd
