// engine.hip -- many independent CG solves advancing together on one GPU
// (the batched and served forms; no reference counterpart: the reference's
// "batch" is k right-hand sides on one matrix, src/batch_spmv_kernel.hpp:4).
//
// Each admitted system keeps its matrix (its own row tiles), its vectors and
// a per-system state record resident in HBM.  One engine iteration is two
// launches over ALL live systems:
//   KA  row tiles of every live system (one workgroup per tile): the SpMV
//       of the system's current phase (set-up, FP64 step, FP64 refinement or
//       FP32 step) with its fused epilogue; per-tile partial sums
//   KC  chunks (>= 2048 rows, <= 256 per system) of every live system: each
//       sums its system's tile partials in index order -> alpha, updates its
//       rows and publishes a partial; the system's last chunk (an
//       agent-scope ticket per system, common.hpp) sums the chunk partials
//       -> beta, convergence, the refinement trigger of the mixed form (or
//       finishes the set-up / a refinement, when KA evaluated a residual)
// A step (k iterations) ends with two more launches: the mixed form's
// finished systems fold x64 + x32 into the caller's x, and one record per
// live system (phase, iterations, residual) is packed and copied to pinned
// host memory behind an event.  Optionally a step's kernels are captured
// once into a hipGraph and replayed (num_engine_set_graph), re-captured
// whenever the live set changes.  Steps are pipelined: num_engine_launch
// returns as soon as the step is queued, and the host learns which systems
// finished when the step's event completes, while the next step runs.
//
// The per-system phase is the convergence mask: a finished system's tiles
// and chunks return at once, and the host drops it from the launch lists
// when it retires the step that finished it.  A newly admitted system joins
// at the next step; its state record and descriptor are written by a
// one-thread kernel from launch arguments (no host synchronisation).  The
// launch lists are expanded on the device from per-system prefix sums, so a
// change of the live set costs O(live systems) host work.  Every reduction
// of system s reads only s's partials, laid out in s's own tile / chunk
// order, so a system's result is bit-identical whatever else shares the
// batch and whenever it was admitted.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <new>
#include <vector>

#include "blas1.hpp"
#include "csr.hpp"
#include "spmv_device.hpp"

namespace numamd {

enum : int32_t {
    PH_FREE = 0,
    PH_INIT = 1,    // FP64: next KA evaluates r = b - A x0
    PH_RUN64 = 2,   // FP64 CG steps
    PH_REFINE = 3,  // mixed: next KA refines in FP64
    PH_RUN32 = 4,   // mixed: FP32 CG steps
    PH_DONE = 5,    // finished, x final
    PH_FIN = 6,     // mixed, finished: x = x64 + x32 still to be folded
};

// 16-byte vector for non-temporal float4 stores
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int kEngNpt = 7;        // LDS entries per thread of the engine's tile kernel
constexpr int kChunkRows = 2048;  // rows per KC workgroup (256 threads x 8), at least
constexpr int kMaxChunks = 256;   // chunks per system, at most (each sums the tile partials)
// Rows per tile of an assembled stencil (<= 7 entries per row): 256, or 512
// in the mixed form, whose FP32 steps stage a 512-row tile at once (its FP32
// values fill the LDS of one 256-row FP64 tile) and whose FP64 phases walk it
// as two 256-row halves.
constexpr int kStencilRpt = 256;
constexpr int kStencilRptPair = 512;
// Row patterns of an assembled stencil (FP64 tiles, and the mixed form's FP32
// steps; its rare FP64 refinements read colind): a row's column offsets are
// fixed by which of its six grid neighbours exist, so a row carries one id
// byte (bit 0: -x, 1: +x, 2: -y, 3: +y, 4: -z, 5: +z neighbour present) into
// the system's table of 64 patterns, each kRelPatW int16 offsets in column
// order (-sz, -sy, -1, 0, +1, +sy, +sz for those present; 0 past the row),
// followed by their lengths (relpat_words: 4 bits each, 4 words).
constexpr int kStencilPat = 64;
constexpr int kStencilPatBytes = relpat_words<2 * kStencilPat>() * 8;

struct SysDesc {
    const int32_t* rowptr;
    const int32_t* colind;
    const double* vals;
    const float* vals32;
    const TileEntry* tiles;  // nullptr: uniform tiles of rpt rows
    const double* b;
    double* x;
    double* r;
    double* p0;
    double* p1;
    double* Ap;
    double* x64b;
    float* x32a;
    float* x32b;
    float* r32;
    float* q32;
    float* p32a;
    float* p32b;
    const uint8_t* rpid;   // assembled stencil: row pattern ids
    const uint64_t* rpat;  // and the table of kStencilPat patterns; or nullptr
    int32_t n, ntiles, mode, rpt;
    int32_t crows;         // rows per update-kernel chunk
    int32_t pad_;
};

struct SysState {
    double rr, rr_ref, alpha, beta, bnorm, tol, delta, rel;
    int32_t it, max_iter, phase;
    uint32_t tick;  // tickets of the update kernel's chunks (common.hpp: take_ticket)
    int32_t converged, breakdown, psel, xsel, nref, final_, gen;
    int32_t pad[5];
};
static_assert(sizeof(SysState) == 128, "state record is 128 bytes");

// What the host learns about a live system at the end of a step.
struct EngRec {
    int32_t phase, it, converged, breakdown, nref, gen;
    double rel;
};
static_assert(sizeof(EngRec) == 32, "record is 32 bytes");

// One entry of the tile list: the system and its tile bounds (rows [r0, r1),
// entries [nz0, nz1), walk mode), resolved once when the list is built so a
// KA workgroup starts its stage loads after a single dependent load.
struct BTile {
    int32_t sys, tile, r0, r1, nz0, nz1, wide, pad;
};
// One entry of the chunk list: the system, its rows [row0, row1), and where
// the system's tile partials [b0, b0 + nb) and chunk partials [c0, c0 + nc)
// sit in the launch's partial arrays.
struct Chunk {
    int32_t sys, row0, row1, b0, nb, c0, nc, pad;
};

// ---------------------------------------------------------------- gathers
template <typename T>
struct EGatherPNew {  // p_new = r + beta * p_old
    const T* __restrict__ r;
    const T* __restrict__ p_old;
    T beta;
    __device__ __forceinline__ void prepare() {}
    __device__ __forceinline__ T operator()(int c) const { return r[c] + beta * p_old[c]; }
};

struct EGatherXsum {  // refined iterate x64 + x32
    const double* __restrict__ x64;
    const float* __restrict__ x32;
    __device__ __forceinline__ void prepare() {}
    __device__ __forceinline__ double operator()(int c) const
    {
        return x64[c] + (double)x32[c];
    }
};

// -------------------------------------------------------------- epilogues
struct EEpiInit {  // r = -1*Ax0 + b ; p_old = 0 ; partials r.r, b.b
    const double* __restrict__ b;
    double* __restrict__ r;
    double* __restrict__ p0;
    template <class G>
    __device__ __forceinline__ void operator()(int row, double v, double& ca, double& cb,
                                               const G&) const
    {
        const double bi = b[row];
        const double ri = -1.0 * v + bi;
        r[row] = ri;
        p0[row] = 0.0;
        ca = ri * ri;
        cb = bi * bi;
    }
};

template <typename T>
struct EEpiStep {  // p_new row to the other buffer, Ap, p_new . Ap
    T* __restrict__ pn_out;
    T* __restrict__ Ap;
    __device__ __forceinline__ void operator()(int row, T v, double& ca, double&,
                                               const EGatherPNew<T>& g) const
    {
        const T pn = g(row);
        __builtin_nontemporal_store(pn, pn_out + row);  // streamed (see cg.hip)
        __builtin_nontemporal_store(v, Ap + row);
        ca = (double)pn * (double)v;
    }
};

// x64' = x64 + x32, x32' = 0, r = b - Ax (FP64), r32 = (float)r; partials
// r.r and r32.r32 -- or b.b instead of r32.r32 on the first refinement,
// which also supplies ||b||.
struct EEpiRefine {
    const double* __restrict__ b;
    double* __restrict__ x64n;
    float* __restrict__ x32n;
    float* __restrict__ r32;
    bool first;
    __device__ __forceinline__ void operator()(int row, double v, double& ca, double& cb,
                                               const EGatherXsum& g) const
    {
        x64n[row] = g(row);
        x32n[row] = 0.0f;
        const double bi = b[row];
        const double ri = -1.0 * v + bi;
        const float rf = (float)ri;
        r32[row] = rf;
        ca = ri * ri;
        cb = first ? bi * bi : (double)rf * (double)rf;
    }
};

// A finished mixed system whose iterate lives in x64b / x32 goes through
// PH_FIN so its x is folded before the host hears it is done.
__device__ __forceinline__ int32_t done_phase(const SysDesc* d, const SysState* st)
{
    return (d->mode == 1 && (st->xsel || st->breakdown)) ? PH_FIN : PH_DONE;
}

// ------------------------------------------------------------------ kernels
// Registers held to 7 waves per SIMD, what the LDS allows (otherwise the
// FP64 phases' register count would set the FP32 steps' occupancy too):
// batched mixed 83k -> 88k solves/s, served lambda = 10k 3.08k -> 3.25k.
template <bool NT64, bool NT32>
__global__ __launch_bounds__(kSpmvBS, 7) void eng_ka(const BTile* __restrict__ bts, int nbt,
                                                  int chunk, const SysDesc* __restrict__ sd,
                                                  const SysState* __restrict__ ss,
                                                  double* __restrict__ pa,
                                                  double* __restrict__ pb)
{
    __shared__ double lv[kSpmvBS * kEngNpt];   // FP32 pairs: 2 * NPT * BS floats
    __shared__ __align__(8) int32_t lc[kSpmvBS * kEngNpt];  // FP32 pairs: the pattern table
    __shared__ double red[2 * kSpmvBS / kWave];
    const int b = chunk_tile(blockIdx.x, nbt, chunk);
    if (b < 0) return;
    const BTile e = bts[b];
    const SysDesc* d = sd + e.sys;
    const SysState* st = ss + e.sys;
    const int ph = st->phase;
    if (ph != PH_INIT && ph != PH_RUN64 && ph != PH_REFINE && ph != PH_RUN32) return;
    const TileEntry e0 = {e.r0, e.wide, e.nz0, 0};
    const TileEntry e1 = {e.r1, kWalkShort, e.nz1, 0};
    double ca = 0.0, cb = 0.0;
    if (d->rpat && d->mode == 1) {
        // 512-row tile of a mixed stencil system
        if (ph == PH_RUN32) {
            const int ps = st->psel;
            EGatherPNew<float> g{d->r32, ps ? d->p32b : d->p32a, (float)st->beta};
            EEpiStep<float> epi{ps ? d->p32a : d->p32b, d->q32};
            spmv_relpat_run<kSpmvBS, kEngNpt, 2, NT32, 2 * kStencilPat>(
                e0, e1, d->rpid, d->rpat, d->vals32, g, epi,
                reinterpret_cast<float*>(lv), reinterpret_cast<uint64_t*>(lc), ca, cb);
        } else {
            // FP64 refinement (the only other phase of a mixed system): the two
            // 256-row halves in turn
            const int xs = st->xsel;
            EGatherXsum g{xs ? d->x64b : d->x, xs ? d->x32b : d->x32a};
            EEpiRefine epi{d->b, xs ? d->x : d->x64b, xs ? d->x32a : d->x32b, d->r32,
                           st->nref == 0};
            const int rm = min(e.r1, e.r0 + kStencilRpt);
            const TileEntry em = {rm, kWalkShort, d->rowptr[rm], 0};
            double xa, xb;
            spmv_tile_run<double, kSpmvBS, kEngNpt, NT64>(e0, em, d->rowptr, d->colind, d->vals,
                                                          g, epi, lv, lc, xa, xb);
            __syncthreads();  // the LDS is staged again for the second half
            spmv_tile_run<double, kSpmvBS, kEngNpt, NT64>(em, e1, d->rowptr, d->colind, d->vals,
                                                          g, epi, lv, lc, ca, cb);
            ca += xa;
            cb += xb;
        }
    } else if (ph == PH_INIT || ph == PH_RUN64) {
        // FP64 system: an assembled stencil's 256-row tiles read the row
        // patterns, a CSR system's tiles its columns
        auto run = [&](const auto& g, const auto& epi) {
            if (d->rpat)
                spmv_relpat_run<kSpmvBS, kEngNpt, 1, NT64, 2 * kStencilPat>(
                    e0, e1, d->rpid, d->rpat, d->vals, g, epi, lv,
                    reinterpret_cast<uint64_t*>(lc), ca, cb);
            else
                spmv_tile_run<double, kSpmvBS, kEngNpt, NT64>(e0, e1, d->rowptr, d->colind,
                                                              d->vals, g, epi, lv, lc, ca, cb);
        };
        if (ph == PH_INIT) {
            run(GatherX<double>{d->x}, EEpiInit{d->b, d->r, d->p0});
        } else {
            const int ps = st->psel;
            run(EGatherPNew<double>{d->r, ps ? d->p1 : d->p0, st->beta},
                EEpiStep<double>{ps ? d->p0 : d->p1, d->Ap});
        }
    } else if (ph == PH_REFINE) {
        const int xs = st->xsel;
        EGatherXsum g{xs ? d->x64b : d->x, xs ? d->x32b : d->x32a};
        EEpiRefine epi{d->b, xs ? d->x : d->x64b, xs ? d->x32a : d->x32b, d->r32,
                       st->nref == 0};
        spmv_tile_run<double, kSpmvBS, kEngNpt, NT64>(e0, e1, d->rowptr, d->colind, d->vals, g,
                                                      epi, lv, lc, ca, cb);
    } else {
        const int ps = st->psel;
        EGatherPNew<float> g{d->r32, ps ? d->p32b : d->p32a, (float)st->beta};
        EEpiStep<float> epi{ps ? d->p32a : d->p32b, d->q32};
        spmv_tile_run<float, kSpmvBS, kEngNpt, NT32>(e0, e1, d->rowptr, d->colind, d->vals32, g,
                                                     epi, reinterpret_cast<float*>(lv), lc, ca,
                                                     cb);
    }
    block_sum2<kSpmvBS>(ca, cb, red);
    if (threadIdx.x == 0) {
        pa[b] = ca;
        pb[b] = cb;
    }
}

// One chunk's CG update x += alpha p, r -= alpha q (T = double: q = Ap; T =
// float: the FP32 step), alpha = rr / p.q from the system's tile partials.
// kPre rows per thread are loaded before alpha's sum, the rest kPre at a time
// (all loads of a group in flight together).  acc gets this thread's r.r.
template <typename T>
__device__ __forceinline__ void chunk_update(const Chunk& e, const T* __restrict__ p,
                                             const T* __restrict__ q, T* __restrict__ x,
                                             T* __restrict__ r, double rr,
                                             const double* __restrict__ pa, double* red,
                                             double& s_alpha, int& s_brk, double& acc)
{
    constexpr int kPre = 4;
    const int t = threadIdx.x;
    T pv[kPre], qv[kPre], xv[kPre], rv[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const int i = e.row0 + t + u * kRedBlock;
        if (i < e.row1) {
            pv[u] = p[i];
            qv[u] = q[i];
            xv[u] = x[i];
            rv[u] = r[i];
        }
    }
    const double pq = sum_partials_block<kRedBlock, 32>(pa + e.b0, e.nb, red);
    if (t == 0) {
        const int brk = !(fabs(pq) >= 1e-300);  // src/cg_solver.hpp:84-85
        s_brk = brk;
        s_alpha = brk ? 0.0 : rr / pq;
    }
    __syncthreads();
    if (s_brk) return;
    const T al = (T)s_alpha, nal = -al;
    for (int base = e.row0 + t;; base += kPre * kRedBlock) {
        if (base != e.row0 + t) {
            if (base >= e.row1) break;
#pragma unroll
            for (int u = 0; u < kPre; ++u) {
                const int i = base + u * kRedBlock;
                if (i < e.row1) {
                    pv[u] = p[i];
                    qv[u] = q[i];
                    xv[u] = x[i];
                    rv[u] = r[i];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kPre; ++u) {
            const int i = base + u * kRedBlock;
            if (i < e.row1) {
                __builtin_nontemporal_store(al * pv[u] + xv[u], x + i);
                const T ri = nal * qv[u] + rv[u];
                __builtin_nontemporal_store(ri, r + i);
                acc += (double)ri * (double)ri;
            }
        }
    }
}

// chunk_update for the FP32 step over float4 groups (row0 is a multiple of
// kRedBlock, the vectors are 256-byte aligned); the rows of a last partial
// group go to thread 0 after its groups.
__device__ __forceinline__ void chunk_update_f4(const Chunk& e, const float* __restrict__ p,
                                                const float* __restrict__ q,
                                                float* __restrict__ x, float* __restrict__ r,
                                                double rr, const double* __restrict__ pa,
                                                double* red, double& s_alpha, int& s_brk,
                                                double& acc)
{
    constexpr int kPre = 2;
    const int t = threadIdx.x;
    const int g0 = e.row0 >> 2, g1 = e.row1 >> 2;  // whole float4 groups [g0, g1)
    const float4* __restrict__ p4 = reinterpret_cast<const float4*>(p);
    const float4* __restrict__ q4 = reinterpret_cast<const float4*>(q);
    float4* __restrict__ x4 = reinterpret_cast<float4*>(x);
    float4* __restrict__ r4 = reinterpret_cast<float4*>(r);
    float4 pv[kPre], qv[kPre], xv[kPre], rv[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const int i = g0 + t + u * kRedBlock;
        if (i < g1) {
            pv[u] = p4[i];
            qv[u] = q4[i];
            xv[u] = x4[i];
            rv[u] = r4[i];
        }
    }
    const double pq = sum_partials_block<kRedBlock, 32>(pa + e.b0, e.nb, red);
    if (t == 0) {
        const int brk = !(fabs(pq) >= 1e-300);  // src/cg_solver.hpp:84-85
        s_brk = brk;
        s_alpha = brk ? 0.0 : rr / pq;
    }
    __syncthreads();
    if (s_brk) return;
    const float al = (float)s_alpha, nal = -al;
    auto upd = [&](int i, const float4& pw, const float4& qw, float4 xw, float4 rw) {
        xw.x = al * pw.x + xw.x;
        xw.y = al * pw.y + xw.y;
        xw.z = al * pw.z + xw.z;
        xw.w = al * pw.w + xw.w;
        rw.x = nal * qw.x + rw.x;
        rw.y = nal * qw.y + rw.y;
        rw.z = nal * qw.z + rw.z;
        rw.w = nal * qw.w + rw.w;
        __builtin_nontemporal_store(f4v{xw.x, xw.y, xw.z, xw.w}, reinterpret_cast<f4v*>(x4 + i));
        __builtin_nontemporal_store(f4v{rw.x, rw.y, rw.z, rw.w}, reinterpret_cast<f4v*>(r4 + i));
        acc += (double)rw.x * (double)rw.x;
        acc += (double)rw.y * (double)rw.y;
        acc += (double)rw.z * (double)rw.z;
        acc += (double)rw.w * (double)rw.w;
    };
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const int i = g0 + t + u * kRedBlock;
        if (i < g1) upd(i, pv[u], qv[u], xv[u], rv[u]);
    }
    for (int i = g0 + t + kPre * kRedBlock; i < g1; i += kRedBlock) upd(i, p4[i], q4[i], x4[i], r4[i]);
    if (t == 0) {
        for (int i = g1 * 4; i < e.row1; ++i) {
            x[i] = al * p[i] + x[i];
            const float ri = nal * q[i] + r[i];
            r[i] = ri;
            acc += (double)ri * (double)ri;
        }
    }
}

// The update kernel, one workgroup per chunk of every live system.
//  * PH_INIT / PH_REFINE (the SpMV kernel evaluated a residual; no update):
//    the system's last chunk (ticket) sums its tile partials and finishes the
//    set-up or the refinement (convergence test, beta).
//  * PH_RUN64 / PH_RUN32: every chunk sums the system's tile partials (same
//    order, same bits) -> alpha, updates its rows and publishes its r.r
//    partial; the system's last chunk folds them: beta, convergence, the
//    mixed form's refinement trigger.
// Every chunk reads the state before it draws its ticket, and only the last
// writes it.
__global__ __launch_bounds__(kRedBlock) void eng_kc(const Chunk* __restrict__ chs,
                                                    const SysDesc* __restrict__ sd,
                                                    SysState* __restrict__ ss,
                                                    const double* __restrict__ pa,
                                                    const double* __restrict__ pb,
                                                    double* __restrict__ pc)
{
    __shared__ double red[kRedBlock / kWave];
    __shared__ double s_alpha;
    __shared__ int s_brk, s_last;
    const Chunk e = chs[blockIdx.x];
    const SysDesc* d = sd + e.sys;
    SysState* st = ss + e.sys;
    const int ph = st->phase;
    const int t = threadIdx.x;
    if (ph == PH_INIT) {
        if (!take_ticket(&st->tick, &s_last, e.nc)) return;
        const double rr = sum_partials_block<kRedBlock>(pa + e.b0, e.nb, red);
        const double bb = sum_partials_block<kRedBlock>(pb + e.b0, e.nb, red);
        if (t == 0) {
            double bn = sqrt(bb);
            if (bn == 0.0) bn = 1.0;  // src/cg_solver.hpp:153
            st->bnorm = bn;
            st->rr = rr;
            st->rel = sqrt(rr) / bn;
            st->beta = 0.0;
            st->psel = 0;
            st->it = 0;
            if (sqrt(rr) / bn < st->tol) {  // early exit, src/cg_solver.hpp:158-159
                st->converged = 1;
                st->phase = PH_DONE;
            } else if (st->max_iter <= 0) {
                st->phase = PH_DONE;
            } else {
                st->phase = PH_RUN64;
            }
        }
        return;
    }
    if (ph == PH_REFINE) {
        if (!take_ticket(&st->tick, &s_last, e.nc)) return;
        const double rr64 = sum_partials_block<kRedBlock>(pa + e.b0, e.nb, red);
        double rr32 = sum_partials_block<kRedBlock>(pb + e.b0, e.nb, red);
        if (t == 0) {
            if (st->nref == 0) {
                // first refinement: the second partial was b.b
                double bn = sqrt(rr32);
                if (bn == 0.0) bn = 1.0;  // ||b|| = 0 -> 1, src/cg_solver.hpp:153
                st->bnorm = bn;
                rr32 = rr64;
            }
            const double rel = sqrt(rr64) / st->bnorm;
            st->rel = rel;
            st->xsel ^= 1;
            st->nref += 1;
            if (rel < st->tol) {
                st->converged = 1;
                st->phase = done_phase(d, st);
            } else if (st->final_ || !(rel == rel)) {
                st->phase = done_phase(d, st);
            } else {
                st->beta = rr32 / st->rr;  // rr = r.r the last p was built with
                st->rr = rr32;
                st->rr_ref = rr32;
                st->phase = PH_RUN32;
            }
        }
        return;
    }
    if (ph != PH_RUN64 && ph != PH_RUN32) return;
    const double rr = st->rr;
    const int psel = st->psel, xsel = st->xsel;
    double acc = 0.0;
    bool brk;
    double alpha;
    if (ph == PH_RUN64) {
        chunk_update<double>(e, psel ? d->p0 : d->p1, d->Ap, d->x, d->r, rr, pa, red, s_alpha,
                             s_brk, acc);
    } else {
        chunk_update_f4(e, psel ? d->p32a : d->p32b, d->q32, xsel ? d->x32b : d->x32a, d->r32,
                        rr, pa, red, s_alpha, s_brk, acc);
    }
    brk = s_brk != 0;
    alpha = s_alpha;
    if (!brk) {
        const double tot = block_sum<kRedBlock>(acc, red);
        if (t == 0) publish(pc + blockIdx.x, tot);
    }
    if (!take_ticket(&st->tick, &s_last, e.nc)) return;
    // the system's last chunk
    if (brk) {
        if (t == 0) {
            st->breakdown = 1;
            st->phase = done_phase(d, st);
        }
        return;
    }
    const double rrn = sum_published_block<kRedBlock>(pc + e.c0, e.nc, red);
    if (t != 0) return;
    st->alpha = alpha;
    const int it = st->it + 1;
    st->it = it;
    st->psel = psel ^ 1;
    if (ph == PH_RUN64) {
        const double rel = sqrt(rrn) / st->bnorm;
        st->rel = rel;
        if (rel < st->tol) {  // strict <, src/cg_solver.hpp:180-183
            st->converged = 1;
            st->phase = PH_DONE;
        } else {
            st->beta = rrn / rr;
            st->rr = rrn;
            if (it >= st->max_iter) st->phase = PH_DONE;
        }
    } else {
        const double est = sqrt(rrn) / st->bnorm;
        if (it >= st->max_iter) {
            st->final_ = 1;
            st->phase = PH_REFINE;
        } else if (est < st->tol || sqrt(rrn) < st->delta * sqrt(st->rr_ref) ||
                   !(rrn == rrn)) {
            st->phase = PH_REFINE;  // beta comes from the refined residual
        } else {
            st->beta = rrn / rr;
            st->rr = rrn;
        }
    }
}

// x = x64 + x32 over the chunks of systems in PH_FIN.
__global__ __launch_bounds__(kRedBlock) void eng_fold(const Chunk* __restrict__ chs,
                                                      const SysDesc* __restrict__ sd,
                                                      const SysState* __restrict__ ss)
{
    const Chunk e = chs[blockIdx.x];
    const SysState* st = ss + e.sys;
    if (st->phase != PH_FIN) return;
    const SysDesc* d = sd + e.sys;
    const double* __restrict__ x64 = st->xsel ? d->x64b : d->x;
    const float* __restrict__ x32 = st->xsel ? d->x32b : d->x32a;
    double* x = d->x;  // may alias x64 (elementwise, same index)
    for (int i = e.row0 + threadIdx.x; i < e.row1; i += kRedBlock) x[i] = x64[i] + (double)x32[i];
}

// One record per live system; PH_FIN becomes PH_DONE (eng_fold ran before).
__global__ void eng_pack(const int32_t* __restrict__ act, int nact, SysState* __restrict__ ss,
                         EngRec* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nact) return;
    SysState* st = ss + act[i];
    int32_t ph = st->phase;
    if (ph == PH_FIN) {
        ph = PH_DONE;
        st->phase = ph;
    }
    EngRec r;
    r.phase = ph;
    r.it = st->it;
    r.converged = st->converged;
    r.breakdown = st->breakdown;
    r.nref = st->nref;
    r.gen = st->gen;
    r.rel = st->rel;
    out[i] = r;
}

// Launch lists from the live systems' prefix sums of tiles and chunks.
__device__ __forceinline__ int owner_of(const int32_t* __restrict__ off, int nact, int g)
{
    int lo = 0, hi = nact - 1;  // largest i with off[i] <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void eng_expand(const int32_t* __restrict__ act, const int32_t* __restrict__ toff,
                           const int32_t* __restrict__ coff, int nact, int nbt, int nch,
                           const SysDesc* __restrict__ sd, BTile* __restrict__ bt,
                           Chunk* __restrict__ ch)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nbt) {
        const int i = owner_of(toff, nact, g);
        const int s = act[i];
        const int t = g - toff[i];
        const SysDesc* d = sd + s;
        BTile e;
        e.sys = s;
        e.tile = t;
        e.pad = 0;
        if (d->tiles) {
            const TileEntry a = d->tiles[t], z = d->tiles[t + 1];
            e.r0 = a.row;
            e.wide = a.wide;
            e.nz0 = a.nz0;
            e.r1 = z.row;
            e.nz1 = z.nz0;
        } else {
            e.r0 = t * d->rpt;
            e.r1 = min(d->n, e.r0 + d->rpt);
            e.wide = kWalkShort;
            e.nz0 = d->rowptr[e.r0];
            e.nz1 = d->rowptr[e.r1];
        }
        bt[g] = e;
    }
    if (g < nch) {
        const int i = owner_of(coff, nact, g);
        const int s = act[i];
        const int cr = sd[s].crows;
        const int r0 = (g - coff[i]) * cr;
        ch[g] = {s, r0, min(sd[s].n, r0 + cr), toff[i], toff[i + 1] - toff[i], coff[i],
                 coff[i + 1] - coff[i], 0};
    }
}

__global__ void eng_install(SysDesc* __restrict__ sd, SysState* __restrict__ ss, int slot,
                            SysDesc d, SysState h)
{
    sd[slot] = d;
    ss[slot] = h;
}

__global__ void eng_zero32(int64_t n, float* __restrict__ a, float* __restrict__ b)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        a[i] = 0.0f;
        b[i] = 0.0f;
    }
}

// ------------------------------------------------- assembled stencil systems
// Counter-based uniform draw in [-1, 1): SplitMix64 (Steele, Lea & Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014) of
// seed * 0xD1B54A32D192ED03 + (i + 1) * 0x9E3779B97F4A7C15; the top 53 bits
// make u in [0, 1), b = 2u - 1 (exact).  generators.hash_uniform is the same
// function in numpy.
__device__ __forceinline__ double hash_uniform(uint64_t seed, int64_t i)
{
    uint64_t z = seed * 0xD1B54A32D192ED03ull + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return 2.0 * ((double)(z >> 11) * 0x1.0p-53) - 1.0;
}

__global__ void hash_uniform_kernel(int64_t n, uint64_t seed, double* __restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = hash_uniform(seed, i);
}

// rowptr[i] of the Dirichlet stencil on an nx x ny x nz grid (x fastest):
// i diagonals plus, per axis of stride s and extent L (block B = s*L), the
// rows before i that have a lower neighbour, (i/B)(B-s) + max(0, i%B - s),
// and those that have an upper one, (i/B)(B-s) + min(i%B, B-s).
__device__ __forceinline__ int64_t stencil_rowptr_at(int ndim, int nx, int ny, int nz, int64_t i)
{
    int64_t p = i;
    const int64_t ext[3] = {nx, ny, nz};
    int64_t s = 1;
    for (int a = 0; a < ndim; ++a) {
        const int64_t B = s * ext[a];
        const int64_t q = i / B, m = i % B;
        p += 2 * q * (B - s) + (m > s ? m - s : 0) + (m < B - s ? m : B - s);
        s = B;
    }
    return p;
}

// One thread per row: rowptr, columns (-z, -y, -x, self, +x, +y, +z), values
// (-1 off the diagonal, 2*ndim + shift on it), the row pattern ids and table
// (kStencilPat), b, x = 0 and, for the mixed form, the FP32 copy and the
// zeroed x32a / p32a.  Unused grid dimensions are 1, so their neighbour bits
// are never set.  Offsets fit int16: lines and planes hold <= 32767 points.
__global__ void stencil_request_kernel(int ndim, int nx, int ny, int nz, double shift,
                                       uint64_t seed, int32_t* __restrict__ rowptr,
                                       int32_t* __restrict__ colind, double* __restrict__ vals,
                                       float* __restrict__ vals32, uint8_t* __restrict__ rpid,
                                       uint64_t* __restrict__ rpat, double* __restrict__ b,
                                       double* __restrict__ x, float* __restrict__ z32a,
                                       float* __restrict__ z32b)
{
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int sy = nx, sz = nx * ny;
    if (rpat && i < kStencilPat) {
        // pattern i: the offsets of a row whose neighbour bits are i
        const int f = (int)i;
        const int off[7] = {-sz, -sy, -1, 0, 1, sy, sz};
        const int bit[7] = {16, 4, 1, 0, 2, 8, 32};
        uint64_t q[2] = {0, 0};
        int u = 0;
        for (int k = 0; k < 7; ++k) {
            if (bit[k] && !(f & bit[k])) continue;
            q[u >> 2] |= (uint64_t)(uint16_t)(int16_t)off[k] << (16 * (u & 3));
            ++u;
        }
        rpat[2 * f] = q[0];
        rpat[2 * f + 1] = q[1];
        if (f < relpat_len_words<2 * kStencilPat>()) {
            // lengths of patterns 16f .. 16f + 15: the row itself plus its
            // neighbours
            uint64_t lw = 0;
            for (int g = 0; g < 16; ++g)
                lw |= (uint64_t)(1 + __popc(16 * f + g)) << (4 * g);
            rpat[2 * kStencilPat + f] = lw;
        }
    }
    if (i >= n) return;
    const int ix = (int)(i % nx);
    const int iy = (int)((i / nx) % ny);
    const int iz = (int)(i / ((int64_t)nx * ny));
    const int f = (ix > 0 ? 1 : 0) | (ix < nx - 1 ? 2 : 0) | (iy > 0 ? 4 : 0) |
                  (iy < ny - 1 ? 8 : 0) | (iz > 0 ? 16 : 0) | (iz < nz - 1 ? 32 : 0);
    int p = (int)stencil_rowptr_at(ndim, nx, ny, nz, i);
    rowptr[i] = p;
    if (i == n - 1) rowptr[n] = (int)stencil_rowptr_at(ndim, nx, ny, nz, n);
    const int row = (int)i;
    auto put = [&](int col, double val) {
        colind[p] = col;
        vals[p] = val;
        if (vals32) vals32[p] = (float)val;
        ++p;
    };
    if (f & 16) put(row - sz, -1.0);
    if (f & 4) put(row - sy, -1.0);
    if (f & 1) put(row - 1, -1.0);
    put(row, 2.0 * ndim + shift);
    if (f & 2) put(row + 1, -1.0);
    if (f & 8) put(row + sy, -1.0);
    if (f & 32) put(row + sz, -1.0);
    if (rpid) rpid[i] = (uint8_t)f;
    b[i] = hash_uniform(seed, i);
    x[i] = 0.0;
    if (z32a) {
        z32a[i] = 0.0f;
        z32b[i] = 0.0f;
    }
}

static inline size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

static int64_t eng_work_bytes(int64_t n, int32_t mode)
{
    if (mode == 0) return (int64_t)(4 * al256((size_t)n * 8));
    return (int64_t)(al256((size_t)n * 8) + 6 * al256((size_t)n * 4));
}

// Carve a system's workspace.
static void carve_work(SysDesc& d, char* w, size_t n, int32_t mode)
{
    if (mode == 0) {
        d.r = (double*)w;
        d.p0 = (double*)(w + al256(n * 8));
        d.p1 = (double*)(w + 2 * al256(n * 8));
        d.Ap = (double*)(w + 3 * al256(n * 8));
    } else {
        size_t o = 0;
        d.x64b = (double*)(w + o); o += al256(n * 8);
        d.x32a = (float*)(w + o); o += al256(n * 4);
        d.x32b = (float*)(w + o); o += al256(n * 4);
        d.r32 = (float*)(w + o);  o += al256(n * 4);
        d.q32 = (float*)(w + o);  o += al256(n * 4);
        d.p32a = (float*)(w + o); o += al256(n * 4);
        d.p32b = (float*)(w + o);
    }
}

struct StencilShape {
    int32_t ndim, nx, ny, nz;
    int64_t n, nnz;
};

static int stencil_shape(int32_t ndim, int32_t nx, int32_t ny, int32_t nz, StencilShape& g)
{
    if (ndim < 1 || ndim > 3) return fail(NUM_EINVAL, "ndim must be 1, 2 or 3");
    if (nx < 1 || ny < 1 || nz < 1) return fail(NUM_EINVAL, "grid sizes must be >= 1");
    if ((ndim < 3 && nz != 1) || (ndim < 2 && ny != 1))
        return fail(NUM_EINVAL, "unused grid dimensions must be 1");
    g.ndim = ndim;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.n = (int64_t)nx * ny * nz;
    if (g.n * (2 * ndim + 1) >= (int64_t)1 << 31) return fail(NUM_EINVAL, "grid too large");
    if ((int64_t)nx * ny > 32767 && ndim == 3)
        return fail(NUM_EINVAL, "3-D grid planes above 32767 points are not supported");
    if ((int64_t)nx > 32767 && ndim >= 2)
        return fail(NUM_EINVAL, "grid lines above 32767 points are not supported");
    const int64_t ext[3] = {nx, ny, nz};
    g.nnz = g.n;
    for (int a = 0; a < ndim; ++a) g.nnz += 2 * (ext[a] - 1) * (g.n / ext[a]);
    return NUM_OK;
}

// Device bytes of an assembled stencil system: x, b, rowptr, colind, values
// (+ FP32 values), row pattern ids and table, workspace -- x first, so the
// solution is the first n doubles of the buffer.
static int64_t stencil_bytes(const StencilShape& g, int32_t mode)
{
    const size_t n = (size_t)g.n, nnz = (size_t)g.nnz;
    size_t b = 2 * al256(n * 8) + al256((n + 1) * 4) + al256(nnz * 4) + al256(nnz * 8);
    b += al256(n) + al256(kStencilPatBytes);
    if (mode == 1) b += al256(nnz * 4);
    return (int64_t)b + eng_work_bytes((int64_t)n, mode);
}

}  // namespace numamd

using namespace numamd;

struct num_engine {
    struct Flight {
        int buf;
        std::vector<int32_t> act;
    };
    int device = 0;
    int32_t cap = 0;
    int32_t nt64 = 1, nt32 = 1;
    SysDesc* d_desc = nullptr;
    SysState* d_state = nullptr;
    // per slot, host side
    std::vector<int32_t> n, ntiles, nchunks;
    std::vector<int32_t> used;  // holds a system
    std::vector<int32_t> live;  // in the launch lists (not yet known finished)
    std::vector<int32_t> done;  // finished and retired: res is valid
    std::vector<int32_t> gen;   // admission counter, tags device records
    // the CSR handle a num_engine_add system was admitted with (its hold is
    // released at remove / destroy), else nullptr
    std::vector<const num_csr*> held;
    std::vector<EngRec> res;
    std::vector<int32_t> order;  // admission order of live systems
    bool dirty = true;
    // launch lists
    BTile* d_bt = nullptr;
    Chunk* d_ch = nullptr;
    int32_t* d_lists = nullptr;  // [act cap | toff cap+1 | coff cap+1]
    double *d_pa = nullptr, *d_pb = nullptr, *d_pc = nullptr;
    EngRec* d_rec = nullptr;
    int64_t bt_cap = 0, ch_cap = 0;
    int32_t nbt = 0, nch = 0, nact = 0;
    std::vector<int32_t> act_host;
    // pinned staging for the lists (two, alternating) and step records (two)
    int32_t* h_stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_used[2] = {false, false};
    int stage_next = 0;
    EngRec* h_rec[2] = {nullptr, nullptr};
    hipEvent_t rec_ev[2] = {nullptr, nullptr};
    std::deque<Flight> flights;
    // optional hipGraph of one step (captured on a private stream, replayed
    // on the caller's; re-captured when the launch lists or iters change)
    // XCD chunk of the tile launch (tiles per XCD deal); 0: by tile count
    int32_t ka_chunk = 0;
    bool use_graph = false;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t gexec = nullptr;
    int32_t g_iters = -1;
    bool g_valid = false;
};

namespace numamd {

static int32_t* eng_act(num_engine* E) { return E->d_lists; }
static int32_t* eng_toff(num_engine* E) { return E->d_lists + E->cap; }
static int32_t* eng_coff(num_engine* E) { return E->d_lists + 2 * E->cap + 1; }

static void eng_free(num_engine* E)
{
    if (E->d_bt) (void)hipFree(E->d_bt);
    if (E->d_ch) (void)hipFree(E->d_ch);
    if (E->d_pa) (void)hipFree(E->d_pa);
    if (E->d_pb) (void)hipFree(E->d_pb);
    if (E->d_pc) (void)hipFree(E->d_pc);
    E->d_bt = nullptr;
    E->d_ch = nullptr;
    E->d_pa = E->d_pb = E->d_pc = nullptr;
    E->bt_cap = E->ch_cap = 0;
}

// Retire the oldest in-flight step: fold its records into the host view.
// Returns 1 if a step was retired, 0 if it is still running (non-blocking).
static int eng_retire_one(num_engine* E, bool block, int* rc)
{
    *rc = NUM_OK;
    if (E->flights.empty()) return 0;
    num_engine::Flight& f = E->flights.front();
    if (!block) {
        const hipError_t q = hipEventQuery(E->rec_ev[f.buf]);
        if (q == hipErrorNotReady) {
            (void)hipGetLastError();  // not an error: clear it
            return 0;
        }
        if (q != hipSuccess) {
            *rc = fail(NUM_EHIP, "engine step failed: %s", hipGetErrorString(q));
            return 0;
        }
    } else {
        const hipError_t q = hipEventSynchronize(E->rec_ev[f.buf]);
        if (q != hipSuccess) {
            *rc = fail(NUM_EHIP, "engine step failed: %s", hipGetErrorString(q));
            return 0;
        }
    }
    const EngRec* rec = E->h_rec[f.buf];
    bool changed = false;
    for (size_t i = 0; i < f.act.size(); ++i) {
        const int32_t s = f.act[i];
        if (!E->used[s] || rec[i].gen != E->gen[s] || E->done[s]) continue;
        if (rec[i].phase == PH_DONE) {
            E->done[s] = 1;
            E->res[s] = rec[i];
            if (E->live[s]) {
                E->live[s] = 0;
                changed = true;
            }
        }
    }
    E->flights.pop_front();
    if (changed) {
        E->dirty = true;
        std::vector<int32_t> keep;
        keep.reserve(E->order.size());
        for (int32_t s : E->order)
            if (E->live[s]) keep.push_back(s);
        E->order.swap(keep);
    }
    return 1;
}

static int eng_retire(num_engine* E, bool all)
{
    int rc = NUM_OK;
    while (!E->flights.empty()) {
        if (!eng_retire_one(E, all, &rc)) break;
    }
    return rc;
}

// Rebuild the launch lists of the live systems (admission order).
static int eng_rebuild(num_engine* E, hipStream_t st)
{
    const int b = E->stage_next;
    E->stage_next ^= 1;
    if (E->stage_used[b]) NUM_HIP(hipEventSynchronize(E->stage_ev[b]));
    int32_t* act = E->h_stage[b];
    int32_t* toff = act + E->cap;
    int32_t* coff = act + 2 * E->cap + 1;
    int32_t na = 0;
    toff[0] = 0;
    coff[0] = 0;
    for (int32_t s : E->order) {
        if (!E->live[s]) continue;
        act[na] = s;
        toff[na + 1] = toff[na] + E->ntiles[s];
        coff[na + 1] = coff[na] + E->nchunks[s];
        ++na;
    }
    const int64_t nbt = toff[na], nch = coff[na];
    if (nbt >= ((int64_t)1 << 31) || nch >= ((int64_t)1 << 31))
        return fail(NUM_EINVAL, "too many tiles in the batch");
    if (nbt > E->bt_cap || nch > E->ch_cap) {
        // in-flight steps still read the old lists
        NUM_HIP(hipStreamSynchronize(st));
        eng_free(E);
        E->bt_cap = std::max<int64_t>(4096, nbt + nbt / 2);
        E->ch_cap = std::max<int64_t>(4096, nch + nch / 2);
        NUM_HIP(hipMalloc(&E->d_bt, E->bt_cap * sizeof(BTile)));
        NUM_HIP(hipMalloc(&E->d_ch, E->ch_cap * sizeof(Chunk)));
        NUM_HIP(hipMalloc(&E->d_pa, E->bt_cap * sizeof(double)));
        NUM_HIP(hipMalloc(&E->d_pb, E->bt_cap * sizeof(double)));
        NUM_HIP(hipMalloc(&E->d_pc, E->ch_cap * sizeof(double)));
    }
    if (na > 0) {
        NUM_HIP(hipMemcpyAsync(E->d_lists, act, (size_t)(3 * E->cap + 2) * sizeof(int32_t),
                               hipMemcpyHostToDevice, st));
        NUM_HIP(hipEventRecord(E->stage_ev[b], st));
        E->stage_used[b] = true;
        const int64_t g = std::max(nbt, nch);
        eng_expand<<<(unsigned)((g + 255) / 256), 256, 0, st>>>(
            eng_act(E), eng_toff(E), eng_coff(E), na, (int)nbt, (int)nch, E->d_desc, E->d_bt,
            E->d_ch);
        NUM_LAUNCH_CHECK();
    }
    E->nbt = (int32_t)nbt;
    E->nch = (int32_t)nch;
    E->nact = na;
    E->act_host.assign(act, act + na);
    E->dirty = false;
    E->g_valid = false;  // grids and list contents changed
    return NUM_OK;
}

// The kernels of one step: `iters` engine iterations, then the fold and
// the record pack (device work only: capturable into a hipGraph).
static int eng_enqueue_step(num_engine* E, int32_t iters, hipStream_t st)
{
    const int32_t* act = eng_act(E);
    // XCD chunks of 8 tiles: the launch mixes many systems, so no single
    // plane offset applies (csr.hpp: plane_chunk); small chunks keep the
    // XCDs on neighbouring tiles (served lambda = 10k: 3257 -> 3288 solves/s,
    // batched FP64 52.4k -> 53.2k against 64-tile chunks)
    const int chunk = E->ka_chunk > 0 ? E->ka_chunk : std::min(8, std::max(1, E->nbt / 64));
    const int grid_a = chunk_grid(E->nbt, chunk);
    for (int32_t k = 0; k < iters; ++k) {
        if (E->nt64) {
            if (E->nt32)
                eng_ka<true, true><<<grid_a, kSpmvBS, 0, st>>>(E->d_bt, E->nbt, chunk, E->d_desc,
                                                               E->d_state, E->d_pa, E->d_pb);
            else
                eng_ka<true, false><<<grid_a, kSpmvBS, 0, st>>>(
                    E->d_bt, E->nbt, chunk, E->d_desc, E->d_state, E->d_pa, E->d_pb);
        } else {
            if (E->nt32)
                eng_ka<false, true><<<grid_a, kSpmvBS, 0, st>>>(
                    E->d_bt, E->nbt, chunk, E->d_desc, E->d_state, E->d_pa, E->d_pb);
            else
                eng_ka<false, false><<<grid_a, kSpmvBS, 0, st>>>(
                    E->d_bt, E->nbt, chunk, E->d_desc, E->d_state, E->d_pa, E->d_pb);
        }
        if (E->nch > 0)
            eng_kc<<<E->nch, kRedBlock, 0, st>>>(E->d_ch, E->d_desc, E->d_state, E->d_pa,
                                                 E->d_pb, E->d_pc);
    }
    if (E->nch > 0) eng_fold<<<E->nch, kRedBlock, 0, st>>>(E->d_ch, E->d_desc, E->d_state);
    eng_pack<<<(E->nact + 255) / 256, 256, 0, st>>>(act, E->nact, E->d_state, E->d_rec);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

static int eng_slot(num_engine* E, int32_t* s)
{
    for (int32_t i = 0; i < E->cap; ++i)
        if (!E->used[i]) {
            *s = i;
            return NUM_OK;
        }
    return fail(NUM_ERUNTIME, "engine full (%d systems)", E->cap);
}

// Queue the descriptor / state upload and register the system.
static int eng_admit(num_engine* E, int32_t s, SysDesc& d, SysState& h, int32_t nchunks,
                     hipStream_t st)
{
    E->gen[s] += 1;
    h.gen = E->gen[s];
    eng_install<<<1, 1, 0, st>>>(E->d_desc, E->d_state, s, d, h);
    NUM_LAUNCH_CHECK();
    E->n[s] = d.n;
    E->ntiles[s] = d.ntiles;
    E->nchunks[s] = nchunks;
    E->used[s] = 1;
    E->live[s] = 1;
    E->done[s] = 0;
    E->order.push_back(s);
    E->dirty = true;
    return NUM_OK;
}

static int eng_check(const num_engine* E)
{
    if (!E) return fail(NUM_EINVAL, "engine is NULL");
    int dev = 0;
    NUM_HIP(hipGetDevice(&dev));
    if (dev != E->device) return fail(NUM_EINVAL, "engine lives on device %d", E->device);
    return NUM_OK;
}

// Rows per update-kernel chunk of an n-row system: kChunkRows, or more (a
// multiple of kRedBlock) so that no system has more than kMaxChunks chunks.
static int32_t chunk_rows(int64_t n)
{
    const int64_t per = (n + kMaxChunks - 1) / kMaxChunks;
    const int64_t r = (per + kRedBlock - 1) / kRedBlock * kRedBlock;
    return (int32_t)std::max<int64_t>(kChunkRows, r);
}

static SysState initial_state(double tol, int32_t max_iter, int32_t mode, double delta)
{
    SysState h{};
    h.tol = tol;
    h.max_iter = max_iter;
    h.delta = delta;
    if (mode == 0) {
        h.phase = PH_INIT;
    } else {
        h.phase = PH_REFINE;  // first slot: r = b - A x0 in FP64, and ||b||
        h.rr = INFINITY;      // first refinement: beta = rr'/inf = 0
        h.final_ = max_iter <= 0 ? 1 : 0;
    }
    return h;
}

static int check_solver_args(double tol, int32_t max_iter, int32_t mode, double delta)
{
    if (mode != 0 && mode != 1) return fail(NUM_EINVAL, "mode must be 0 (FP64) or 1 (mixed)");
    if (mode == 1 && !(delta > 0.0 && delta < 1.0))
        return fail(NUM_EINVAL, "delta must be in (0, 1)");
    if (max_iter < 0) return fail(NUM_EINVAL, "max_iter must be >= 0");
    if (!(tol >= 0.0)) return fail(NUM_EINVAL, "tol must be >= 0");
    return NUM_OK;
}

}  // namespace numamd

extern "C" {

int num_engine_create(int32_t max_systems, int32_t nt64, int32_t nt32, num_engine** out)
{
    if (!out) return fail(NUM_EINVAL, "out is NULL");
    *out = nullptr;
    if (max_systems <= 0 || max_systems > (1 << 24))
        return fail(NUM_EINVAL, "max_systems must be in [1, 2^24]");
    num_engine* E = new (std::nothrow) num_engine();
    if (!E) return fail(NUM_ENOMEM, "out of host memory");
    E->cap = max_systems;
    E->nt64 = nt64 != 0;
    E->nt32 = nt32 != 0;
    const size_t c = (size_t)max_systems;
    E->n.assign(c, 0);
    E->ntiles.assign(c, 0);
    E->nchunks.assign(c, 0);
    E->used.assign(c, 0);
    E->live.assign(c, 0);
    E->done.assign(c, 0);
    E->gen.assign(c, 0);
    E->held.assign(c, nullptr);
    E->res.assign(c, EngRec{});
    const size_t lists = (3 * c + 2) * sizeof(int32_t);
    hipError_t e = hipGetDevice(&E->device);
    if (e == hipSuccess) e = hipMalloc(&E->d_desc, c * sizeof(SysDesc));
    if (e == hipSuccess) e = hipMalloc(&E->d_state, c * sizeof(SysState));
    if (e == hipSuccess) e = hipMalloc(&E->d_lists, lists);
    if (e == hipSuccess) e = hipMalloc(&E->d_rec, c * sizeof(EngRec));
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
        e = hipHostMalloc((void**)&E->h_stage[i], lists, hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&E->h_rec[i], c * sizeof(EngRec), hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&E->stage_ev[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&E->rec_ev[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        (void)num_engine_destroy(E);
        return fail(NUM_EHIP, "engine allocation failed: %s", hipGetErrorString(e));
    }
    *out = E;
    return NUM_OK;
}

int num_engine_set_graph(num_engine* E, int32_t on)
{
    if (!E) return fail(NUM_EINVAL, "engine is NULL");
    if (on && !E->cap_stream)
        NUM_HIP(hipStreamCreateWithFlags(&E->cap_stream, hipStreamNonBlocking));
    E->use_graph = on != 0;
    E->g_valid = false;
    return NUM_OK;
}

int num_engine_set_schedule(num_engine* E, int32_t ka_chunk)
{
    if (!E) return fail(NUM_EINVAL, "engine is NULL");
    if (ka_chunk < 0 || ka_chunk > 4096)
        return fail(NUM_EINVAL, "ka_chunk must be in [0, 4096], got %d", ka_chunk);
    if (ka_chunk != E->ka_chunk) {
        E->ka_chunk = ka_chunk;
        E->g_valid = false;  // the captured step's grid changes
    }
    return NUM_OK;
}

int num_engine_destroy(num_engine* E)
{
    if (!E) return NUM_OK;
    // queued steps and installs still read the engine's arrays
    (void)hipDeviceSynchronize();
    if (E->gexec) (void)hipGraphExecDestroy(E->gexec);
    if (E->cap_stream) (void)hipStreamDestroy(E->cap_stream);
    eng_free(E);
    if (E->d_desc) (void)hipFree(E->d_desc);
    if (E->d_state) (void)hipFree(E->d_state);
    if (E->d_lists) (void)hipFree(E->d_lists);
    if (E->d_rec) (void)hipFree(E->d_rec);
    for (const num_csr* A : E->held)
        if (A) A->engine_holds.fetch_sub(1);
    for (int i = 0; i < 2; ++i) {
        if (E->stage_ev[i]) (void)hipEventDestroy(E->stage_ev[i]);
        if (E->rec_ev[i]) (void)hipEventDestroy(E->rec_ev[i]);
        if (E->h_stage[i]) (void)hipHostFree(E->h_stage[i]);
        if (E->h_rec[i]) (void)hipHostFree(E->h_rec[i]);
    }
    delete E;
    return NUM_OK;
}

int64_t num_engine_work_bytes(int32_t n, int32_t mode)
{
    if (n < 0 || (mode != 0 && mode != 1)) return -1;
    return eng_work_bytes(n, mode);
}

int num_engine_add(num_engine* E, const num_csr* A, const double* d_b, double* d_x, double tol,
                   int32_t max_iter, int32_t mode, double delta, void* d_work,
                   int64_t work_bytes, void* stream, int32_t* slot_out)
{
    int rc = eng_check(E);
    if (rc) return rc;
    if (!A || !slot_out) return fail(NUM_EINVAL, "matrix or slot is NULL");
    if (!d_b || !d_x) return fail(NUM_EINVAL, "b or x is NULL");
    if (A->nrows != A->ncols) return fail(NUM_EINVAL, "CG needs a square matrix");
    rc = check_solver_args(tol, max_iter, mode, delta);
    if (rc) return rc;
    if (mode == 1 && !A->vals32)
        return fail(NUM_EINVAL, "mixed mode needs FP32 values (num_csr_attach_f32)");
    if (A->sched.npt > kEngNpt || A->sched.row_begin != 0 || A->sched.row_end != A->nrows)
        return fail(NUM_EINVAL, "matrix schedule holds %d entries per thread; the engine "
                    "handles at most %d (set_schedule with nnz_per_block <= %d)",
                    A->sched.npt, kEngNpt, kEngNpt * kSpmvBS);
    const int64_t need = eng_work_bytes(A->nrows, mode);
    if (!d_work || work_bytes < need)
        return fail(NUM_EINVAL, "workspace too small: need %lld bytes", (long long)need);
    if (((uintptr_t)d_work & 255) != 0)
        return fail(NUM_EINVAL, "workspace must be 256-byte aligned");
    rc = check_device(A);
    if (rc) return rc;
    int32_t s = -1;
    rc = eng_slot(E, &s);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)A->nrows;
    SysDesc d{};
    d.rowptr = A->rowptr;
    d.colind = A->colind;
    d.vals = A->vals;
    d.vals32 = A->vals32;
    d.tiles = static_cast<const TileEntry*>(A->sched.d_tiles);
    d.rpid = nullptr;
    d.rpat = nullptr;
    d.b = d_b;
    d.x = d_x;
    d.n = A->nrows;
    d.ntiles = A->sched.nblocks;
    d.mode = mode;
    d.rpt = 0;
    carve_work(d, (char*)d_work, n, mode);
    SysState h = initial_state(tol, max_iter, mode, delta);
    if (mode == 1 && n > 0) {
        eng_zero32<<<elementwise_grid((int64_t)n), kRedBlock, 0, st>>>((int64_t)n, d.x32a,
                                                                        d.p32a);
        NUM_LAUNCH_CHECK();
    }
    d.crows = chunk_rows(n);
    rc = eng_admit(E, s, d, h, (int32_t)((n + d.crows - 1) / d.crows), st);
    if (rc) return rc;
    // the launch lists are expanded from A's tiles while the system is in
    // the engine: hold the schedule until the slot is removed
    A->engine_holds.fetch_add(1);
    E->held[s] = A;
    *slot_out = s;
    return NUM_OK;
}

int num_hash_uniform(int64_t n, uint64_t seed, double* d_out, void* stream)
{
    if (n < 0) return fail(NUM_EINVAL, "n must be >= 0");
    if (n == 0) return NUM_OK;
    if (!d_out) return fail(NUM_EINVAL, "d_out is NULL");
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hash_uniform_kernel<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(n, seed, d_out);
    NUM_LAUNCH_CHECK();
    return NUM_OK;
}

int64_t num_engine_stencil_bytes(int32_t ndim, int32_t nx, int32_t ny, int32_t nz, int32_t mode)
{
    StencilShape g;
    if (mode != 0 && mode != 1) return -1;
    if (stencil_shape(ndim, nx, ny, nz, g) != NUM_OK) return -1;
    return stencil_bytes(g, mode);
}

int num_engine_add_stencil(num_engine* E, int32_t ndim, int32_t nx, int32_t ny, int32_t nz,
                           double shift, uint64_t seed, double tol, int32_t max_iter,
                           int32_t mode, double delta, void* d_buf, int64_t buf_bytes,
                           void* stream, int32_t* slot_out)
{
    int rc = eng_check(E);
    if (rc) return rc;
    if (!slot_out) return fail(NUM_EINVAL, "slot is NULL");
    rc = check_solver_args(tol, max_iter, mode, delta);
    if (rc) return rc;
    StencilShape g;
    rc = stencil_shape(ndim, nx, ny, nz, g);
    if (rc) return rc;
    const int64_t need = stencil_bytes(g, mode);
    if (!d_buf || buf_bytes < need)
        return fail(NUM_EINVAL, "buffer too small: need %lld bytes", (long long)need);
    if (((uintptr_t)d_buf & 255) != 0) return fail(NUM_EINVAL, "buffer must be 256-byte aligned");
    int32_t s = -1;
    rc = eng_slot(E, &s);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)g.n, nnz = (size_t)g.nnz;
    char* w = (char*)d_buf;
    SysDesc d{};
    size_t o = 0;
    d.x = (double*)(w + o); o += al256(n * 8);
    d.b = (const double*)(w + o);
    double* b = (double*)(w + o); o += al256(n * 8);
    int32_t* rowptr = (int32_t*)(w + o); o += al256((n + 1) * 4);
    int32_t* colind = (int32_t*)(w + o); o += al256(nnz * 4);
    double* vals = (double*)(w + o); o += al256(nnz * 8);
    float* vals32 = nullptr;
    if (mode == 1) {
        vals32 = (float*)(w + o);
        o += al256(nnz * 4);
    }
    uint8_t* rpid = (uint8_t*)(w + o);
    o += al256(n);
    uint64_t* rpat = (uint64_t*)(w + o);
    o += al256(kStencilPatBytes);
    carve_work(d, w + o, n, mode);
    d.rowptr = rowptr;
    d.colind = colind;
    d.vals = vals;
    d.vals32 = vals32;
    d.tiles = nullptr;
    d.rpid = rpid;
    d.rpat = rpat;
    d.rpt = mode == 1 ? kStencilRptPair : kStencilRpt;
    d.n = (int32_t)n;
    d.ntiles = (int32_t)((n + d.rpt - 1) / d.rpt);
    d.mode = mode;
    stencil_request_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
        ndim, nx, ny, nz, shift, seed, rowptr, colind, vals, vals32, rpid, rpat, b, d.x,
        mode == 1 ? d.x32a : nullptr, mode == 1 ? d.p32a : nullptr);
    NUM_LAUNCH_CHECK();
    SysState h = initial_state(tol, max_iter, mode, delta);
    d.crows = chunk_rows(n);
    rc = eng_admit(E, s, d, h, (int32_t)((n + d.crows - 1) / d.crows), st);
    if (rc) return rc;
    *slot_out = s;
    return NUM_OK;
}

int num_engine_launch(num_engine* E, int32_t iters, void* stream)
{
    int rc = eng_check(E);
    if (rc) return rc;
    if (iters < 0) return fail(NUM_EINVAL, "iters must be >= 0");
    hipStream_t st = (hipStream_t)stream;
    rc = eng_retire(E, false);
    if (rc) return rc;
    while (E->flights.size() >= 2) {
        if (!eng_retire_one(E, true, &rc)) return rc ? rc : fail(NUM_ERUNTIME, "retire failed");
    }
    if (E->dirty) {
        rc = eng_rebuild(E, st);
        if (rc) return rc;
    }
    if (E->nact == 0) return NUM_OK;
    if (E->use_graph) {
        if (!E->g_valid || E->g_iters != iters) {
            if (E->gexec) {
                (void)hipGraphExecDestroy(E->gexec);
                E->gexec = nullptr;
            }
            NUM_HIP(hipStreamBeginCapture(E->cap_stream, hipStreamCaptureModeThreadLocal));
            rc = eng_enqueue_step(E, iters, E->cap_stream);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(E->cap_stream, &g);
            if (rc) {
                if (e == hipSuccess && g) (void)hipGraphDestroy(g);
                return rc;
            }
            NUM_HIP(e);
            const hipError_t ei = hipGraphInstantiate(&E->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            NUM_HIP(ei);
            E->g_valid = true;
            E->g_iters = iters;
        }
        NUM_HIP(hipGraphLaunch(E->gexec, st));
    } else {
        rc = eng_enqueue_step(E, iters, st);
        if (rc) return rc;
    }
    // the record buffer not held by the (at most one) step still in flight
    const int buf = (!E->flights.empty() && E->flights.front().buf == 0) ? 1 : 0;
    NUM_HIP(hipMemcpyAsync(E->h_rec[buf], E->d_rec, (size_t)E->nact * sizeof(EngRec),
                           hipMemcpyDeviceToHost, st));
    NUM_HIP(hipEventRecord(E->rec_ev[buf], st));
    E->flights.push_back({buf, E->act_host});
    return NUM_OK;
}

int num_engine_poll(num_engine* E, int32_t wait, int32_t* n_live, int32_t* n_inflight)
{
    int rc = eng_check(E);
    if (rc) return rc;
    rc = eng_retire(E, wait != 0);
    if (rc) return rc;
    if (n_live) {
        int32_t c = 0;
        for (int32_t s : E->order) c += E->live[s];
        *n_live = c;
    }
    if (n_inflight) *n_inflight = (int32_t)E->flights.size();
    return NUM_OK;
}

int num_engine_step(num_engine* E, int32_t iters, int32_t* n_active, void* stream)
{
    int rc = num_engine_launch(E, iters, stream);
    if (rc) return rc;
    return num_engine_poll(E, 1, n_active, nullptr);
}

int num_engine_result(num_engine* E, int32_t slot, num_cg_result* res, int32_t* refinements,
                      void* stream)
{
    (void)stream;
    if (!E || !res) return fail(NUM_EINVAL, "engine or result is NULL");
    if (slot < 0 || slot >= E->cap || !E->used[slot])
        return fail(NUM_EINVAL, "slot %d holds no system", slot);
    if (!E->done[slot])
        return fail(NUM_ERUNTIME, "system %d has not finished (or its step is not retired)",
                    slot);
    const EngRec& r = E->res[slot];
    res->iters = r.it;
    res->converged = r.converged;
    res->breakdown = r.breakdown;
    res->residual = r.rel;
    if (refinements) *refinements = r.nref;
    return NUM_OK;
}

int num_engine_finished(const num_engine* E, int32_t* slots, int32_t max_slots, int32_t* count)
{
    if (!E || !count) return fail(NUM_EINVAL, "engine or count is NULL");
    int32_t c = 0;
    for (int32_t s = 0; s < E->cap; ++s) {
        if (E->used[s] && E->done[s]) {
            if (slots && c < max_slots) slots[c] = s;
            ++c;
        }
    }
    *count = c;
    return NUM_OK;
}

int num_engine_remove(num_engine* E, int32_t slot)
{
    if (!E) return fail(NUM_EINVAL, "engine is NULL");
    if (slot < 0 || slot >= E->cap || !E->used[slot])
        return fail(NUM_EINVAL, "slot %d holds no system", slot);
    if (E->live[slot]) {
        E->live[slot] = 0;
        std::vector<int32_t> keep;
        for (int32_t s : E->order)
            if (s != slot) keep.push_back(s);
        E->order.swap(keep);
        E->dirty = true;
    }
    E->used[slot] = 0;
    E->done[slot] = 0;
    if (E->held[slot]) {
        E->held[slot]->engine_holds.fetch_sub(1);
        E->held[slot] = nullptr;
    }
    return NUM_OK;
}

int num_engine_info(const num_engine* E, int32_t* n_live, int32_t* n_tiles, int32_t* n_chunks)
{
    if (!E) return fail(NUM_EINVAL, "engine is NULL");
    int32_t c = 0;
    for (int32_t v : E->live) c += v;
    if (n_live) *n_live = c;
    if (n_tiles) *n_tiles = E->nbt;
    if (n_chunks) *n_chunks = E->nch;
    return NUM_OK;
}

}  // extern "C"
