/*
 * num_amd.h -- C ABI of libnum_amd.so: CSR SpMV, multi-RHS SpMV, BLAS-1 and
 * CG on one MI355X (gfx950), the drop-in native layer under the Python module
 * num_amd.kokkos_spmv.
 *
 * Conventions
 *  - Plain pointers and sizes only.  "d_" pointers are device memory
 *    (hipMalloc / torch tensors), "h_" pointers host memory.  Index arrays are
 *    int32 (reference src/crs_matrix.hpp:22-24: Ordinal = Offset = int),
 *    values float64 (Scalar = double) unless the name says f32.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Kernel entry points are asynchronous on that stream; the Python layer
 *    synchronises before returning, as every reference call ends in a
 *    Kokkos::fence (src/spmv_kernel.hpp:160,179; src/vec_ops.hpp:25,41,57).
 *  - Return 0 on success, else a NUM_E* code; num_last_error() gives the
 *    thread-local message.  NUM_EINVAL is the reference's
 *    std::invalid_argument (Python ValueError), the others its
 *    std::runtime_error (Python RuntimeError).
 *  - No entry point falls back to the CPU.
 */
#ifndef NUM_AMD_H
#define NUM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NUM_OK 0
#define NUM_EINVAL 1   /* bad argument: shape, size, pointer, range        */
#define NUM_ERUNTIME 2 /* numerical breakdown, state error                  */
#define NUM_EHIP 3     /* HIP runtime error                                 */
#define NUM_ENOMEM 4   /* allocation failed                                 */

typedef struct num_csr num_csr;

/* CGResult (src/cg_solver.hpp:22-26) plus what this build adds. */
typedef struct {
    int32_t iters;        /* loop SpMVs; 0 on early exit (cg_solver.hpp:158-159) */
    int32_t converged;    /* 1 if sqrt(r.r)/||b|| < tol (strict <)              */
    int32_t breakdown;    /* 1 if |p.Ap| < 1e-300 stopped the solve             */
    int32_t pad_;
    double residual;      /* sqrt(r.r)/||b|| of the recursive residual          */
} num_cg_result;

/* Mixed-precision refined CG (no reference counterpart). */
typedef struct {
    int32_t outer;        /* FP64 true-residual evaluations (refinements)      */
    int32_t inner_total;  /* FP32 CG steps                                     */
    int32_t converged;    /* 1 if the FP64 true residual reached tol          */
    int32_t breakdown;
    double true_residual; /* ||b - A x|| / ||b|| in FP64 at exit              */
} num_mixed_result;

/* ---- runtime ----------------------------------------------------------- */
/* init(device_id) of src/bindings.cpp:53-66: selects the device; idempotent. */
int num_init(int32_t device_id);
/* finalize() of src/bindings.cpp:68-70. */
int num_finalize(void);
const char* num_last_error(void);
int32_t num_abi_version(void);

/* ---- CSR handle (src/crs_matrix.hpp:37-73, CrsMatrixHandle) -------------- */
/* Borrow device CSR arrays owned by the caller (they must outlive the handle)
 * and build the row-block schedule from the host copy of rowptr.
 * Rejects nrows<=0 or ncols<=0 (crs_matrix.hpp:52-53) and, unlike the
 * reference (no range validation, crs_matrix.hpp:50-68), a non-monotone
 * rowptr, rowptr[0] != 0 and a column index outside [0, ncols) (checked on
 * the device, with NUM_EINVAL). */
int num_csr_create(const int32_t* h_rowptr, const int32_t* d_rowptr,
                   const int32_t* d_colind, const double* d_vals, int32_t nrows,
                   int32_t ncols, num_csr** out);
/* Copy host CSR arrays to the device (the CrsMatrixHandle constructor itself,
 * crs_matrix.hpp:45-72); the handle owns the copies.  Also checks every
 * column index is in [0, ncols). */
int num_csr_create_host(const int32_t* h_rowptr, const int32_t* h_colind,
                        const double* h_vals, int32_t nrows, int32_t ncols,
                        num_csr** out);
/* Attach an FP32 copy of the values (same pattern) for the FP32 inner solve.
 * When the windowed layout is built (num_csr_set_windows) this also builds
 * its FP32 sliced values (synchronises the default stream). */
int num_csr_attach_f32(num_csr* A, const float* d_vals32);
/* Build column - row offsets as int16 into d_rel16 [nnz] for the FP32 inner
 * solve (no reference counterpart): attached (*attached = 1) only for a
 * square matrix whose rows all hold <= 32 entries and whose offsets all fit
 * in int16 (stencils, banded matrices); otherwise the FP32 path keeps the
 * int32 columns.  Synchronises the stream. */
int num_csr_attach_rel16(num_csr* A, int16_t* d_rel16, void* stream, int32_t* attached);
/* Build a delta-coded copy of the column ids into d_codes (uint16, 16-byte
 * aligned, nnz + 8 + ceil(nrows / 2) entries: the nnz codes and 8 padding
 * codes, read in 16-byte units, then one length byte per row that replaces
 * rowptr in the kernels) for
 * the FP64 row-tile kernels (no reference counterpart: a data format of this
 * build; the reference reads int32 colind, crs_matrix.hpp:22-34).  Entry j
 * of a row holds its column minus the previous entry's column; a row's first
 * entry holds its column minus its tile's base column.  Attached
 * (*attached = 1) only when every row holds <= 32 entries in ascending
 * column order, every code fits in 16 bits and the values are 16-byte
 * aligned; then the FP64 SpMV and CG on
 * the whole-matrix schedule read 2 index bytes per entry instead of 4 and 1
 * byte per row instead of two 4-byte row bounds, with the same columns in the
 * same order (bit-identical results).  The buffer
 * stays the caller's and must outlive A; num_csr_set_schedule re-codes it
 * (and detaches it if the new tiles do not fit).  Synchronises the stream. */
int num_csr_attach_d16(num_csr* A, uint16_t* d_codes, void* stream, int32_t* attached);
/* Attach a column-partitioned layout of A for num_spmv_xcd (no reference
 * counterpart: a data layout of this build for matrices whose x gathers
 * spread far beyond one XCD's L2, such as random columns).  The columns are
 * cut into nblocks equal blocks; S is a segment matrix (num_csr_create over
 * the caller's device arrays) with one row per (block, row) pair that holds
 * entries, block-major and row-ascending within a block, each holding that
 * row's entries of the block in their original order;
 * h_block_seg_start [nblocks + 1] gives each block's first segment;
 * d_row_seg_ptr [nrows + 1] / d_row_seg_idx [nseg] list each row's segments
 * in block order.  Block j is served by XCD j % 8.  S and the maps stay the
 * caller's and must outlive A; S's schedule must not change afterwards. */
int num_csr_attach_xcd(num_csr* A, const num_csr* S, int32_t nblocks,
                       const int32_t* h_block_seg_start, const int32_t* d_row_seg_ptr,
                       const int32_t* d_row_seg_idx);
/* y = A x through the column-partitioned layout (replaces the kernels behind
 * launch_spmv_kokkoskernels / launch_spmv_custom, spmv_kernel.hpp:102-180,
 * for such matrices): a partial sum per segment into d_work [nseg] doubles,
 * then each row's partials summed in block order.  A row whose entries all
 * lie in one block is bit-identical to the row loop; others agree to
 * rounding.  d_work is per call (concurrent calls need their own). */
int num_spmv_xcd(const num_csr* A, const double* d_x, double* d_y, double* d_work,
                 void* stream);
/* *delta16 = 1 if a delta-coded index is attached (num_csr_attach_d16). */
int num_csr_index_info(const num_csr* A, int32_t* delta16);
/* Windowed sliced layout (no reference counterpart: a data layout of this
 * build for banded and stencil matrices, which the FP64 SpMV and CG kernels
 * and the mixed solver's streamed FP32 steps use in place of the int32 or
 * delta-coded columns).  It applies when every row holds <= 8 entries and
 * every tile of the schedule reads x from at most 4 contiguous column
 * ranges of at most 1152 columns in all: a tile fetches those ranges once
 * into LDS with coalesced loads and reads its rows -- library-owned copies
 * of the values and of each entry's 16-bit slot in the ranges, in sliced
 * order -- straight into registers.  Same entries in the same order:
 * bit-identical results.  Where every tile holds at most 16 distinct row
 * patterns (a row's slots minus its position in the tile: a stencil's
 * interior and boundary rows), each row keeps a 4-bit pattern id into its
 * tile's table (tiles with equal tables share one) instead of a 2-byte slot
 * per entry and a length byte.  mode 1 builds it (when
 * the matrix fits), 2 builds it with a slot per entry only, 0 drops
 * it, -1 (the default, applied at num_csr_attach_d16) builds it for
 * matrices of at least 65,536 rows (below that an SpMV is launch-latency
 * bound and the build costs more than it saves).  num_csr_set_schedule
 * rebuilds it for the new tiles.  The layout copies the values: they must
 * not change while it is built.  Synchronises the stream. */
int num_csr_set_windows(num_csr* A, int32_t mode, void* stream);
/* *windowed = 2 if the column windows are in use with row patterns, 1 with
 * a slot per entry, 0 if not built. */
int num_csr_window_info(const num_csr* A, int32_t* windowed);
/* Which kernel takes the FP32 steps of num_cg_solve_mixed (no reference
 * counterpart): mode 1 the windowed sliced step, two tiles per workgroup
 * over the windowed layout's row patterns or slots (when the windowed
 * layout and the FP32 values are attached), 0 the staged row-tile / pair
 * kernels (the pair kernel reading the int16 offsets' row patterns where
 * they were built: one id byte per row into a table of at most 16 offset
 * patterns and lengths per pair of tiles, no rowptr), 2 the staged kernels
 * with 2 offset bytes per entry, -1
 * (the default) the windowed step when the FP32 copy streams from HBM and
 * the matrix's dominant long-range column offset is at least 16,384 (a 3-D
 * grid plane of that many rows; on 2-D and short-plane grids the staged
 * pair kernel is faster), else as 0.  Never changes which layout is built;
 * every mode gives the same bits as another mode on the same kernel. */
int num_csr_set_mixed_path(num_csr* A, int32_t mode);
/* *windows = 1 if the windowed layout reads row patterns, *offsets = 1 if the
 * int16 column offsets have their row-pattern form (no reference
 * counterpart; either pointer may be NULL). */
int num_csr_pattern_info(const num_csr* A, int32_t* windows, int32_t* offsets);
/* Fails with NUM_ERUNTIME (and frees nothing) while an engine system admitted
 * with A (num_engine_add) has not been removed. */
int num_csr_destroy(num_csr* A);
/* nrows / ncols / nnz read-only attributes (src/bindings.cpp:79-88). */
int num_csr_shape(const num_csr* A, int32_t* nrows, int32_t* ncols, int64_t* nnz);
/* Row-tile schedule: at most rows_per_block rows (<= 256) and nnz_per_block
 * non-zeros per workgroup (rows longer than nnz_per_block get a workgroup of
 * their own); tiles dealt to the 8 XCDs in chunks of `chunk` tiles; nt64 /
 * nt32 = 1 streams the FP64 / FP32 values non-temporally.  Values <= 0 (for
 * the sizes) or < 0 (chunk, nt64, nt32) pick the defaults.  Rebuilds the
 * schedule; used by the autotuner.  Fails with NUM_ERUNTIME while an engine
 * system admitted with A (num_engine_add) has not been removed: the engine
 * reads A's tiles whenever it rebuilds its launch lists. */
int num_csr_set_schedule(num_csr* A, int32_t rows_per_block, int32_t nnz_per_block,
                         int32_t chunk, int32_t nt64, int32_t nt32);
int num_csr_schedule_info(const num_csr* A, int32_t* nblocks, int32_t* rows_per_block,
                          int32_t* nnz_per_block, int32_t* long_rows, int32_t* max_row_nnz,
                          int32_t* chunk, int32_t* nt64, int32_t* nt32);

/* ---- SpMV (src/spmv_kernel.hpp:118-180) ---------------------------------- */
/* y = A x.  Replaces both launch_spmv_kokkoskernels (spmv_kernel.hpp:170-180)
 * and launch_spmv_custom (spmv_kernel.hpp:118-161): one row-binned kernel. */
int num_spmv(const num_csr* A, const double* d_x, double* d_y, void* stream);
/* y[r] = (A x)[r] for rows r in [row_begin, row_end) only (interior/boundary
 * split of one rank's slab; no reference counterpart, the reference runs the
 * whole local SpMV after the halo, src/dist_spmv.hpp:157-162). */
int num_spmv_rows(const num_csr* A, int32_t row_begin, int32_t row_end,
                  const double* d_x, double* d_y, void* stream);
/* Slab SpMV without an x_ext copy: rows [row_begin, row_end) of a rank's
 * local matrix whose columns >= nrows are ghosts; column c reads
 * x_own[c] if c < nrows else x_ghost[c - nrows].  Replaces x_ext_fill_local +
 * local SpMV of distributed_spmv (src/dist_spmv.hpp:118-163). */
int num_spmv_split(const num_csr* A, int32_t row_begin, int32_t row_end,
                   const double* d_x_own, const double* d_x_ghost, double* d_y, void* stream);
/* FP32 y = A32 x (needs num_csr_attach_f32). */
int num_spmv_f32(const num_csr* A, const float* d_x, float* d_y, void* stream);
/* Host round trip: upload, y = A x, download (spmv_host_roundtrip,
 * src/spmv_kernel.hpp:186-217). */
int num_spmv_host(const int32_t* h_rowptr, const int32_t* h_colind, const double* h_vals,
                  int32_t nrows, int32_t ncols, const double* h_x, double* h_y);

/* ---- multi-RHS SpMV (src/batch_spmv_kernel.hpp:120-174) ------------------ */
/* Y = A X, X (ncols, k) and Y (nrows, k) row-major (LayoutRight). */
int num_batch_spmv(const num_csr* A, const double* d_X, double* d_Y, int32_t k,
                   void* stream);
/* batch_spmv_host_roundtrip (src/batch_spmv_kernel.hpp:179-209). */
int num_batch_spmv_host(const int32_t* h_rowptr, const int32_t* h_colind,
                        const double* h_vals, int32_t nrows, int32_t ncols, int32_t k,
                        const double* h_X, double* h_Y);

/* ---- BLAS-1 (src/vec_ops.hpp:16-100) ------------------------------------- */
int num_axpy(int64_t n, double alpha, const double* d_x, const double* d_y, double* d_z,
             void* stream);                                    /* z = a x + y   */
int num_axpy2(int64_t n, double alpha, const double* d_x, double beta, const double* d_y,
              double* d_z, void* stream);                      /* z = a x + b y */
int num_scale(int64_t n, double alpha, double* d_x, void* stream);   /* x *= a  */
int num_fill(int64_t n, double val, double* d_x, void* stream);      /* x = v   */
int num_copy(int64_t n, const double* d_src, double* d_dst, void* stream);
/* Bytes of device workspace num_dot needs for length n. */
int64_t num_dot_workspace_bytes(int64_t n);
/* *d_out = sum x_i y_i, deterministic two-stage reduction, result left on the
 * device (vec_local_dot, vec_ops.hpp:47-59, minus its host read-back). */
int num_dot(int64_t n, const double* d_x, const double* d_y, double* d_out, void* d_work,
            void* stream);
/* num_dot without a workspace argument and without allocating: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The same two
 * kernels as num_dot, bit-identical results, with the stage-1 partials in a
 * static 16 KiB device array (2048 doubles, enough for any n: the partial
 * count is capped) that each GPU holds until the process ends, its address
 * resolved by num_init (else by the first call on a device).  Calls no
 * hipMalloc / hipFree.  All calls on one GPU share that array: calls that
 * overlap on one GPU (different streams, or threads) race and must be
 * serialised by the caller. */
int num_dot_noalloc(int64_t n, const double* d_x, const double* d_y, double* d_out,
                    void* stream);
/* *d_out = sqrt(sum x_i^2) (vec_local_norm2, vec_ops.hpp:64-69). */
int num_norm2(int64_t n, const double* d_x, double* d_out, void* d_work, void* stream);

/* ---- CG (src/cg_solver.hpp:133-196, cg_solve_local) ---------------------- */
/* Workspace bytes for num_cg_solve; every workspace handed to the CG and
 * engine entry points must be 256-byte aligned (hipMalloc and torch
 * allocations are). */
int64_t num_cg_workspace_bytes(const num_csr* A);
/* Solves A x = b; x holds x0 on entry.  Same operation order, ||b||=0 -> 1
 * rule, strict < test and iteration count as the reference; scalars stay on
 * the device.  Returns NUM_ERUNTIME (result filled, breakdown=1) when
 * |p.Ap| < 1e-300, the guard of cg_solve_distributed (cg_solver.hpp:84-85). */
int num_cg_solve(const num_csr* A, const double* d_b, double* d_x, double tol,
                 int32_t max_iter, void* d_work, int64_t work_bytes, num_cg_result* res,
                 void* stream);

/* ---- mixed-precision refined CG (no reference counterpart) -------------- */
int64_t num_cg_mixed_workspace_bytes(const num_csr* A);
/* CG steps in FP32 (FP32 matrix copy and vectors, dot products accumulated
 * in FP64) inside FP64 iterative refinement with reliable updates: whenever
 * the FP32 residual has dropped by `delta` since the last refinement (or
 * looks converged), x += x32 and the true residual r = b - A x are computed
 * in FP64 on the GPU, the FP32 residual is re-seeded from r and the search
 * direction is kept.  Stops when the FP64 true residual ||b - A x||/||b|| <
 * tol (strict) or after max_iter FP32 steps.  Needs num_csr_attach_f32. */
int num_cg_solve_mixed(const num_csr* A, const double* d_b, double* d_x, double tol,
                       int32_t max_iter, double delta, void* d_work, int64_t work_bytes,
                       num_mixed_result* res, void* stream);

/* ---- many systems at once: batched and served CG (no reference
 * counterpart; the reference's batch is k right-hand sides on one matrix).
 * An engine holds up to max_systems independent SPD systems, each a num_csr
 * handle (schedule with <= 7 entries per thread, i.e. nnz_per_block <= 1792)
 * or an assembled stencil, plus b, x and workspace, all resident on the
 * device.  Every iteration advances all live systems with four launches;
 * each system keeps its own scalars, convergence flag and iteration count,
 * and computes exactly what it would alone (bit-identical).  All calls on
 * one engine must use one stream; none of them synchronises the host except
 * num_engine_step, num_engine_poll(wait=1), and num_engine_launch when two
 * steps are already in flight. -------------------------------------------- */
typedef struct num_engine num_engine;
/* nt64 / nt32: stream the FP64 / FP32 matrix values non-temporally. */
int num_engine_create(int32_t max_systems, int32_t nt64, int32_t nt32, num_engine** out);
int num_engine_destroy(num_engine* E);
/* Workspace bytes for one system of n rows; mode 0 = FP64 CG, 1 = mixed. */
int64_t num_engine_work_bytes(int32_t n, int32_t mode);
/* Admit a system (joins the running batch at the next step).  mode 0: the
 * reference CG sequence in FP64 (src/cg_solver.hpp:133-196); mode 1: FP32
 * steps with FP64 reliable-update refinement (needs num_csr_attach_f32), stop
 * on the FP64 true residual.  x holds x0; on completion it holds the
 * solution.  Returns the slot. */
int num_engine_add(num_engine* E, const num_csr* A, const double* d_b, double* d_x, double tol,
                   int32_t max_iter, int32_t mode, double delta, void* d_work,
                   int64_t work_bytes, void* stream, int32_t* slot_out);
/* Device bytes of an assembled stencil system (see num_engine_add_stencil),
 * or -1 for an invalid shape. */
int64_t num_engine_stencil_bytes(int32_t ndim, int32_t nx, int32_t ny, int32_t nz, int32_t mode);
/* Admit A x = b for the Dirichlet stencil of num_stencil_fill on an
 * nx x ny x nz grid, assembled on the device inside d_buf (>= the bytes
 * above, 256-byte aligned): b[i] = 2u - 1 with u the SplitMix64 draw of
 * (seed, i) (generators.hash_uniform), x0 = 0.  The solution is the first n
 * doubles of d_buf.  Grid lines (2-D) and planes (3-D) must hold <= 32767
 * points, so that column offsets fit the int16 copy of the mixed form. */
int num_engine_add_stencil(num_engine* E, int32_t ndim, int32_t nx, int32_t ny, int32_t nz,
                           double shift, uint64_t seed, double tol, int32_t max_iter,
                           int32_t mode, double delta, void* d_buf, int64_t buf_bytes,
                           void* stream, int32_t* slot_out);
/* on != 0: capture a step's kernels into a hipGraph (on a private stream)
 * and replay it on the caller's stream; the graph is re-captured when the
 * live set or the iteration count changes.  Results are identical. */
int num_engine_set_graph(num_engine* E, int32_t on);
/* Tile deal of the engine's SpMV launch: ka_chunk consecutive tiles per XCD
 * turn (0 = chosen from the tile count: min(8, tiles / 64)).  Only the
 * tile-to-workgroup mapping changes, never a result (every system's partial
 * sums keep their order).  Set by the autotuner (no reference counterpart). */
int num_engine_set_schedule(num_engine* E, int32_t ka_chunk);
/* Queue `iters` iterations of every live system and return; at most two
 * steps are in flight (with two queued, the older is waited for first). */
int num_engine_launch(num_engine* E, int32_t iters, void* stream);
/* Retire finished steps (all of them, waiting, if wait != 0): systems they
 * finished leave the launch lists.  *n_live = systems not known finished,
 * *n_inflight = steps still queued. */
int num_engine_poll(num_engine* E, int32_t wait, int32_t* n_live, int32_t* n_inflight);
/* num_engine_launch + num_engine_poll(wait = 1). */
int num_engine_step(num_engine* E, int32_t iters, int32_t* n_active, void* stream);
/* Result of a finished system (mixed: refinements = FP64 evaluations). */
int num_engine_result(num_engine* E, int32_t slot, num_cg_result* res, int32_t* refinements,
                      void* stream);
/* Slots whose system finished and has not been removed yet (up to
 * max_slots written; *count = how many there are). */
int num_engine_finished(const num_engine* E, int32_t* slots, int32_t max_slots, int32_t* count);
/* Free a slot (its buffers are the caller's; a step still in flight may
 * touch them, in stream order). */
int num_engine_remove(num_engine* E, int32_t slot);
int num_engine_info(const num_engine* E, int32_t* n_live, int32_t* n_tiles, int32_t* n_chunks);

/* Build the values and column indices of the 1-/2-/3-D Dirichlet stencil
 * (5-point in 2-D, 7-point in 3-D; diagonal 2*ndim + shift, off-diagonals
 * -1, columns sorted, x fastest) on an nx x ny x nz grid, on the device,
 * given its rowptr (the closed form of python/matrix_generators.py:16-80 with
 * a shift; no reference counterpart as a device routine). */
int num_stencil_fill(int32_t ndim, int32_t nx, int32_t ny, int32_t nz, double shift,
                     const int32_t* d_rowptr, int32_t* d_colind, double* d_vals, void* stream);

/* d_out[i] = 2u - 1 for i < n, u the top 53 bits of SplitMix64 of
 * seed * 0xD1B54A32D192ED03 + (i + 1) * 0x9E3779B97F4A7C15: the right-hand
 * side num_engine_add_stencil assembles, for callers that build the same
 * system themselves (no reference counterpart: a workload generator). */
int num_hash_uniform(int64_t n, uint64_t seed, double* d_out, void* stream);

/* ---- halo pack / scatter of one rank's share (src/halo_exchange.hpp) --- */
/* buf[i] = x[idx[i]] for i < count (halo_pack, halo_exchange.hpp:222-235). */
int num_halo_pack(const int32_t* d_idx, int32_t count, const double* d_x, double* d_buf,
                  void* stream);
/* x_ghost[offset + i] = buf[i] for i < count (halo_scatter_<r>,
 * halo_exchange.hpp:287-300). */
int num_halo_scatter(const double* d_buf, int32_t count, double* d_x_ghost, int32_t offset,
                     void* stream);

#ifdef __cplusplus
}
#endif
#endif /* NUM_AMD_H */
