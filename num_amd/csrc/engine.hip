// engine.hip -- many independent CG solves advancing together on one GPU
// (the batched and served forms; no reference counterpart: the reference's
// "batch" is k right-hand sides on one matrix, src/batch_spmv_kernel.hpp:4).
//
// Each admitted system keeps its CSR handle (with its own row-tile
// schedule), its vectors and a per-system state record resident in HBM.  One
// engine iteration ("slot") is four launches over ALL active systems:
//   KA  row tiles of every active system (one workgroup per tile): the SpMV
//       of the system's current phase (set-up, FP64 step, FP64 refinement or
//       FP32 step) with its fused epilogue; per-tile partial sums
//   KB  one workgroup per active system: sums that system's tile partials in
//       index order -> alpha (or the set-up / refinement finish)
//   KC  2048-row chunks of every active system: the vector updates, per-chunk
//       partials
//   KD  one workgroup per active system: beta, convergence, the refinement
//       trigger of the mixed form
// The per-system phase is the convergence mask: a finished system's tiles
// and chunks return at once, and the host drops it from the launch lists at
// the next step.  A newly admitted system joins the running batch at the next
// step; its first slot evaluates r = b - A x0.  Every reduction of system s
// reads only s's partials, laid out in s's own tile / chunk order, so a
// system's result is bit-identical whatever else shares the batch.
#include <algorithm>
#include <cmath>
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
    PH_DONE = 5,
};

constexpr int kEngNpt = 7;        // LDS entries per thread of the engine's tile kernel
constexpr int kChunkRows = 2048;  // rows per KC workgroup (256 threads x 8)

struct SysDesc {
    const int32_t* rowptr;
    const int32_t* colind;
    const double* vals;
    const float* vals32;
    const TileEntry* tiles;
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
    int32_t n, ntiles, mode, pad;
    int32_t bt_begin, bt_end, ch_begin, ch_end;
};

struct SysState {
    double rr, rr_ref, alpha, beta, bnorm, tol, delta, rel;
    int32_t it, max_iter, phase, post, converged, breakdown, psel, xsel, nref, final_;
    int32_t pad[6];
};
static_assert(sizeof(SysState) == 128, "state record is 128 bytes");

struct BTile {
    int32_t sys, tile;
};
struct Chunk {
    int32_t sys, row0, row1, pad;
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
        pn_out[row] = pn;
        Ap[row] = v;
        ca = (double)pn * (double)v;
    }
};

struct EEpiRefine {  // x64' = x64 + x32, x32' = 0, r = b - Ax (FP64), r32
    const double* __restrict__ b;
    double* __restrict__ x64n;
    float* __restrict__ x32n;
    float* __restrict__ r32;
    __device__ __forceinline__ void operator()(int row, double v, double& ca, double& cb,
                                               const EGatherXsum& g) const
    {
        x64n[row] = g(row);
        x32n[row] = 0.0f;
        const double ri = -1.0 * v + b[row];
        const float rf = (float)ri;
        r32[row] = rf;
        ca = ri * ri;
        cb = (double)rf * (double)rf;
    }
};

// ------------------------------------------------------------------ kernels
template <bool NT64, bool NT32>
__global__ __launch_bounds__(kSpmvBS) void eng_ka(const BTile* __restrict__ bts, int nbt,
                                                  int chunk, const SysDesc* __restrict__ sd,
                                                  const SysState* __restrict__ ss,
                                                  double* __restrict__ pa,
                                                  double* __restrict__ pb)
{
    __shared__ double lv[kSpmvBS * kEngNpt];
    __shared__ int32_t lc[kSpmvBS * kEngNpt];
    __shared__ double red[2 * kSpmvBS / kWave];
    const int b = chunk_tile(blockIdx.x, nbt, chunk);
    if (b < 0) return;
    const BTile e = bts[b];
    const SysDesc* d = sd + e.sys;
    const SysState* st = ss + e.sys;
    const int ph = st->phase;
    double ca = 0.0, cb = 0.0;
    if (ph == PH_INIT) {
        GatherX<double> g{d->x};
        EEpiInit epi{d->b, d->r, d->p0};
        spmv_tile_body<double, kSpmvBS, kEngNpt, NT64>(e.tile, d->rowptr, d->colind, d->vals,
                                                       g, d->tiles, epi, lv, lc, ca, cb);
    } else if (ph == PH_RUN64) {
        const int ps = st->psel;
        EGatherPNew<double> g{d->r, ps ? d->p1 : d->p0, st->beta};
        EEpiStep<double> epi{ps ? d->p0 : d->p1, d->Ap};
        spmv_tile_body<double, kSpmvBS, kEngNpt, NT64>(e.tile, d->rowptr, d->colind, d->vals,
                                                       g, d->tiles, epi, lv, lc, ca, cb);
    } else if (ph == PH_REFINE) {
        const int xs = st->xsel;
        EGatherXsum g{xs ? d->x64b : d->x, xs ? d->x32b : d->x32a};
        EEpiRefine epi{d->b, xs ? d->x : d->x64b, xs ? d->x32a : d->x32b, d->r32};
        spmv_tile_body<double, kSpmvBS, kEngNpt, NT64>(e.tile, d->rowptr, d->colind, d->vals,
                                                       g, d->tiles, epi, lv, lc, ca, cb);
    } else if (ph == PH_RUN32) {
        const int ps = st->psel;
        EGatherPNew<float> g{d->r32, ps ? d->p32b : d->p32a, (float)st->beta};
        EEpiStep<float> epi{ps ? d->p32a : d->p32b, d->q32};
        spmv_tile_body<float, kSpmvBS, kEngNpt, NT32>(e.tile, d->rowptr, d->colind, d->vals32,
                                                      g, d->tiles, epi,
                                                      reinterpret_cast<float*>(lv), lc, ca, cb);
    } else {
        return;
    }
    block_sum2<kSpmvBS>(ca, cb, red);
    if (threadIdx.x == 0) {
        pa[b] = ca;
        pb[b] = cb;
    }
}

__global__ __launch_bounds__(kSpmvBS) void eng_kb(const int32_t* __restrict__ act, int nact,
                                                  const SysDesc* __restrict__ sd,
                                                  SysState* __restrict__ ss,
                                                  const double* __restrict__ pa,
                                                  const double* __restrict__ pb)
{
    __shared__ double red[kSpmvBS / kWave];
    const int s = act[blockIdx.x];
    SysState* st = ss + s;
    const int ph = st->phase;
    const int b0 = sd[s].bt_begin, nb = sd[s].bt_end - sd[s].bt_begin;
    if (ph == PH_INIT) {
        const double rr = sum_partials_block<kSpmvBS>(pa + b0, nb, red);
        const double bb = sum_partials_block<kSpmvBS>(pb + b0, nb, red);
        if (threadIdx.x == 0) {
            double bn = sqrt(bb);
            if (bn == 0.0) bn = 1.0;  // src/cg_solver.hpp:153
            st->bnorm = bn;
            st->rr = rr;
            st->rel = sqrt(rr) / bn;
            st->beta = 0.0;
            st->psel = 0;
            st->it = 0;
            st->post = 1;
            if (sqrt(rr) / bn < st->tol) {  // early exit, src/cg_solver.hpp:158-159
                st->converged = 1;
                st->phase = PH_DONE;
            } else if (st->max_iter <= 0) {
                st->phase = PH_DONE;
            } else {
                st->phase = PH_RUN64;
            }
        }
    } else if (ph == PH_REFINE) {
        const double rr64 = sum_partials_block<kSpmvBS>(pa + b0, nb, red);
        const double rr32 = sum_partials_block<kSpmvBS>(pb + b0, nb, red);
        if (threadIdx.x == 0) {
            const double rel = sqrt(rr64) / st->bnorm;
            st->rel = rel;
            st->xsel ^= 1;
            st->nref += 1;
            st->post = 1;
            if (rel < st->tol) {
                st->converged = 1;
                st->phase = PH_DONE;
            } else if (st->final_ || !(rel == rel)) {
                st->phase = PH_DONE;
            } else {
                st->beta = rr32 / st->rr;  // rr = r.r the last p was built with
                st->rr = rr32;
                st->rr_ref = rr32;
                st->phase = PH_RUN32;
            }
        }
    } else if (ph == PH_RUN64 || ph == PH_RUN32) {
        const double pq = sum_partials_block<kSpmvBS>(pa + b0, nb, red);
        if (threadIdx.x == 0) {
            if (!(fabs(pq) >= 1e-300)) {  // src/cg_solver.hpp:84-85
                st->breakdown = 1;
                st->phase = PH_DONE;
            } else {
                st->alpha = st->rr / pq;
            }
        }
    }
}

__global__ __launch_bounds__(kRedBlock) void eng_kc(const Chunk* __restrict__ chs, int nch,
                                                    const SysDesc* __restrict__ sd,
                                                    const SysState* __restrict__ ss,
                                                    double* __restrict__ pc)
{
    __shared__ double red[kRedBlock / kWave];
    const Chunk e = chs[blockIdx.x];
    const SysDesc* d = sd + e.sys;
    const SysState* st = ss + e.sys;
    const int ph = st->phase;
    if (st->post || (ph != PH_RUN64 && ph != PH_RUN32)) return;
    double acc = 0.0;
    if (ph == PH_RUN64) {
        const double al = st->alpha, nal = -al;
        const double* __restrict__ p = st->psel ? d->p0 : d->p1;  // p_new
        double* __restrict__ x = d->x;
        double* __restrict__ r = d->r;
        const double* __restrict__ Ap = d->Ap;
        for (int i = e.row0 + threadIdx.x; i < e.row1; i += kRedBlock) {
            x[i] = al * p[i] + x[i];
            const double ri = nal * Ap[i] + r[i];
            r[i] = ri;
            acc += ri * ri;
        }
    } else {
        const float al = (float)st->alpha, nal = -al;
        const float* __restrict__ p = st->psel ? d->p32a : d->p32b;
        float* __restrict__ x32 = st->xsel ? d->x32b : d->x32a;
        float* __restrict__ r = d->r32;
        const float* __restrict__ q = d->q32;
        for (int i = e.row0 + threadIdx.x; i < e.row1; i += kRedBlock) {
            x32[i] = al * p[i] + x32[i];
            const float ri = nal * q[i] + r[i];
            r[i] = ri;
            acc += (double)ri * (double)ri;
        }
    }
    const double s = block_sum<kRedBlock>(acc, red);
    if (threadIdx.x == 0) pc[blockIdx.x] = s;
}

__global__ __launch_bounds__(kSpmvBS) void eng_kd(const int32_t* __restrict__ act, int nact,
                                                  const SysDesc* __restrict__ sd,
                                                  SysState* __restrict__ ss,
                                                  const double* __restrict__ pc)
{
    __shared__ double red[kSpmvBS / kWave];
    const int s = act[blockIdx.x];
    SysState* st = ss + s;
    const int ph = st->phase;
    if (ph != PH_RUN64 && ph != PH_RUN32) return;
    if (st->post) {
        if (threadIdx.x == 0) st->post = 0;
        return;
    }
    const int c0 = sd[s].ch_begin, nc = sd[s].ch_end - sd[s].ch_begin;
    const double rrn = sum_partials_block<kSpmvBS>(pc + c0, nc, red);
    if (threadIdx.x != 0) return;
    const int it = st->it + 1;
    st->it = it;
    st->psel ^= 1;
    if (ph == PH_RUN64) {
        const double rel = sqrt(rrn) / st->bnorm;
        st->rel = rel;
        if (rel < st->tol) {  // strict <, src/cg_solver.hpp:180-183
            st->converged = 1;
            st->phase = PH_DONE;
        } else {
            st->beta = rrn / st->rr;
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
            st->beta = rrn / st->rr;
            st->rr = rrn;
        }
    }
}

__global__ void eng_pack(const int32_t* __restrict__ act, int nact,
                         const SysState* __restrict__ ss, int32_t* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nact) out[i] = ss[act[i]].phase;
}

__global__ void eng_fill(int64_t n, float* __restrict__ a, float* __restrict__ b)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        a[i] = 0.0f;
        b[i] = 0.0f;
    }
}

__global__ void eng_mix_finish(int64_t n, const double* __restrict__ x64,
                               const float* __restrict__ x32, double* __restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = x64[i] + (double)x32[i];
}

static inline size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace numamd

using namespace numamd;

struct num_engine {
    int device = 0;
    int32_t cap = 0;
    int32_t nt64 = 1, nt32 = 1;
    SysDesc* d_desc = nullptr;
    SysState* d_state = nullptr;
    std::vector<SysDesc> h_desc;
    std::vector<int32_t> used;    // 1 if the slot holds a system
    std::vector<int32_t> live;    // 1 if the system is in the launch lists
    std::vector<int32_t> order;   // admission order of live systems
    bool dirty = true;
    BTile* d_bt = nullptr;
    Chunk* d_ch = nullptr;
    int32_t* d_act = nullptr;
    double *d_pa = nullptr, *d_pb = nullptr, *d_pc = nullptr;
    int32_t* d_flags = nullptr;
    int32_t* h_flags = nullptr;  // pinned
    int64_t bt_cap = 0, ch_cap = 0, act_cap = 0;
    int32_t nbt = 0, nch = 0, nact = 0;
    std::vector<int32_t> act_host;
};

namespace numamd {

static void eng_free_lists(num_engine* E)
{
    if (E->d_bt) (void)hipFree(E->d_bt);
    if (E->d_ch) (void)hipFree(E->d_ch);
    if (E->d_act) (void)hipFree(E->d_act);
    if (E->d_pa) (void)hipFree(E->d_pa);
    if (E->d_pb) (void)hipFree(E->d_pb);
    if (E->d_pc) (void)hipFree(E->d_pc);
    if (E->d_flags) (void)hipFree(E->d_flags);
    if (E->h_flags) (void)hipHostFree(E->h_flags);
    E->d_bt = nullptr;
    E->d_ch = nullptr;
    E->d_act = nullptr;
    E->d_pa = E->d_pb = E->d_pc = nullptr;
    E->d_flags = nullptr;
    E->h_flags = nullptr;
    E->bt_cap = E->ch_cap = E->act_cap = 0;
}

// Rebuild the launch lists from the live systems (admission order) and
// upload them with the descriptors.
static int eng_rebuild(num_engine* E, hipStream_t st)
{
    std::vector<BTile> bt;
    std::vector<Chunk> ch;
    std::vector<int32_t> act;
    for (int32_t s : E->order) {
        if (!E->live[s]) continue;
        SysDesc& d = E->h_desc[s];
        d.bt_begin = (int32_t)bt.size();
        for (int32_t t = 0; t < d.ntiles; ++t) bt.push_back({s, t});
        d.bt_end = (int32_t)bt.size();
        d.ch_begin = (int32_t)ch.size();
        for (int32_t r0 = 0; r0 < d.n; r0 += kChunkRows)
            ch.push_back({s, r0, std::min(d.n, r0 + kChunkRows), 0});
        d.ch_end = (int32_t)ch.size();
        act.push_back(s);
    }
    NUM_HIP(hipStreamSynchronize(st));
    if ((int64_t)bt.size() > E->bt_cap || (int64_t)ch.size() > E->ch_cap ||
        (int64_t)act.size() > E->act_cap) {
        eng_free_lists(E);
        E->bt_cap = std::max<int64_t>(1024, (int64_t)bt.size() * 5 / 4);
        E->ch_cap = std::max<int64_t>(1024, (int64_t)ch.size() * 5 / 4);
        E->act_cap = std::max<int64_t>(256, (int64_t)act.size() * 5 / 4);
        NUM_HIP(hipMalloc(&E->d_bt, E->bt_cap * sizeof(BTile)));
        NUM_HIP(hipMalloc(&E->d_ch, E->ch_cap * sizeof(Chunk)));
        NUM_HIP(hipMalloc(&E->d_act, E->act_cap * sizeof(int32_t)));
        NUM_HIP(hipMalloc(&E->d_pa, E->bt_cap * sizeof(double)));
        NUM_HIP(hipMalloc(&E->d_pb, E->bt_cap * sizeof(double)));
        NUM_HIP(hipMalloc(&E->d_pc, E->ch_cap * sizeof(double)));
        NUM_HIP(hipMalloc(&E->d_flags, E->act_cap * sizeof(int32_t)));
        NUM_HIP(hipHostMalloc((void**)&E->h_flags, E->act_cap * sizeof(int32_t),
                              hipHostMallocDefault));
    }
    if (!bt.empty()) NUM_HIP(hipMemcpy(E->d_bt, bt.data(), bt.size() * sizeof(BTile),
                                       hipMemcpyHostToDevice));
    if (!ch.empty()) NUM_HIP(hipMemcpy(E->d_ch, ch.data(), ch.size() * sizeof(Chunk),
                                       hipMemcpyHostToDevice));
    if (!act.empty()) NUM_HIP(hipMemcpy(E->d_act, act.data(), act.size() * sizeof(int32_t),
                                        hipMemcpyHostToDevice));
    NUM_HIP(hipMemcpy(E->d_desc, E->h_desc.data(), (size_t)E->cap * sizeof(SysDesc),
                      hipMemcpyHostToDevice));
    E->nbt = (int32_t)bt.size();
    E->nch = (int32_t)ch.size();
    E->nact = (int32_t)act.size();
    E->act_host = std::move(act);
    E->dirty = false;
    return NUM_OK;
}

static int64_t eng_work_bytes(int64_t n, int32_t mode)
{
    if (mode == 0) return (int64_t)(4 * al256((size_t)n * 8));
    return (int64_t)(al256((size_t)n * 8) + 6 * al256((size_t)n * 4));
}

}  // namespace numamd

extern "C" {

int num_engine_create(int32_t max_systems, int32_t nt64, int32_t nt32, num_engine** out)
{
    if (!out) return fail(NUM_EINVAL, "out is NULL");
    *out = nullptr;
    if (max_systems <= 0) return fail(NUM_EINVAL, "max_systems must be > 0");
    num_engine* E = new (std::nothrow) num_engine();
    if (!E) return fail(NUM_ENOMEM, "out of host memory");
    hipError_t e = hipGetDevice(&E->device);
    E->cap = max_systems;
    E->nt64 = nt64 != 0;
    E->nt32 = nt32 != 0;
    E->h_desc.assign((size_t)max_systems, SysDesc{});
    E->used.assign((size_t)max_systems, 0);
    E->live.assign((size_t)max_systems, 0);
    if (e == hipSuccess) e = hipMalloc(&E->d_desc, (size_t)max_systems * sizeof(SysDesc));
    if (e == hipSuccess) e = hipMalloc(&E->d_state, (size_t)max_systems * sizeof(SysState));
    if (e != hipSuccess) {
        if (E->d_desc) (void)hipFree(E->d_desc);
        if (E->d_state) (void)hipFree(E->d_state);
        delete E;
        return fail(NUM_EHIP, "engine allocation failed: %s", hipGetErrorString(e));
    }
    *out = E;
    return NUM_OK;
}

int num_engine_destroy(num_engine* E)
{
    if (!E) return NUM_OK;
    eng_free_lists(E);
    if (E->d_desc) (void)hipFree(E->d_desc);
    if (E->d_state) (void)hipFree(E->d_state);
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
    if (!E || !A || !slot_out) return fail(NUM_EINVAL, "engine, matrix or slot is NULL");
    if (!d_b || !d_x) return fail(NUM_EINVAL, "b or x is NULL");
    if (A->nrows != A->ncols) return fail(NUM_EINVAL, "CG needs a square matrix");
    if (mode != 0 && mode != 1) return fail(NUM_EINVAL, "mode must be 0 (FP64) or 1 (mixed)");
    if (mode == 1 && !A->vals32)
        return fail(NUM_EINVAL, "mixed mode needs FP32 values (num_csr_attach_f32)");
    if (mode == 1 && !(delta > 0.0 && delta < 1.0))
        return fail(NUM_EINVAL, "delta must be in (0, 1)");
    if (max_iter < 0) return fail(NUM_EINVAL, "max_iter must be >= 0");
    if (A->sched.npt > kEngNpt || A->sched.row_begin != 0 || A->sched.row_end != A->nrows)
        return fail(NUM_EINVAL, "matrix schedule holds %d entries per thread; the engine "
                    "handles at most %d (set_schedule with nnz_per_block <= %d)",
                    A->sched.npt, kEngNpt, kEngNpt * kSpmvBS);
    const int64_t need = eng_work_bytes(A->nrows, mode);
    if (!d_work || work_bytes < need)
        return fail(NUM_EINVAL, "workspace too small: need %lld bytes", (long long)need);
    int rc = check_device(A);
    if (rc) return rc;
    int32_t s = -1;
    for (int32_t i = 0; i < E->cap; ++i)
        if (!E->used[i]) { s = i; break; }
    if (s < 0) return fail(NUM_ERUNTIME, "engine full (%d systems)", E->cap);
    const size_t n = (size_t)A->nrows;
    char* w = (char*)d_work;
    SysDesc d{};
    d.rowptr = A->rowptr;
    d.colind = A->colind;
    d.vals = A->vals;
    d.vals32 = A->vals32;
    d.tiles = static_cast<const TileEntry*>(A->sched.d_tiles);
    d.b = d_b;
    d.x = d_x;
    d.n = A->nrows;
    d.ntiles = A->sched.nblocks;
    d.mode = mode;
    SysState h{};
    h.tol = tol;
    h.max_iter = max_iter;
    h.delta = delta;
    if (mode == 0) {
        d.r = (double*)w;
        d.p0 = (double*)(w + al256(n * 8));
        d.p1 = (double*)(w + 2 * al256(n * 8));
        d.Ap = (double*)(w + 3 * al256(n * 8));
        h.phase = PH_INIT;
    } else {
        size_t o = 0;
        d.x64b = (double*)(w + o); o += al256(n * 8);
        d.x32a = (float*)(w + o); o += al256(n * 4);
        d.x32b = (float*)(w + o); o += al256(n * 4);
        d.r32 = (float*)(w + o);  o += al256(n * 4);
        d.q32 = (float*)(w + o);  o += al256(n * 4);
        d.p32a = (float*)(w + o); o += al256(n * 4);
        d.p32b = (float*)(w + o); o += al256(n * 4);
        h.phase = PH_REFINE;  // first slot: r = b - A x0 in FP64
        h.rr = INFINITY;      // first refinement: beta = rr'/inf = 0
        h.final_ = max_iter <= 0 ? 1 : 0;
        // ||b|| is needed by the first refinement's test.  Scratch: the sum
        // lands in x64b and the partials in q32 (>= 256 bytes, >= one partial
        // per 2048 rows), both overwritten before they are read as vectors.
        double* bb = (double*)(d.x64b);
        rc = launch_dot((int64_t)n, d_b, d_b, bb, (double*)d.q32, 0, (hipStream_t)stream);
        if (rc) return rc;
        double hb = 0.0;
        NUM_HIP(hipMemcpyAsync(&hb, bb, sizeof(double), hipMemcpyDeviceToHost,
                               (hipStream_t)stream));
        NUM_HIP(hipStreamSynchronize((hipStream_t)stream));
        h.bnorm = sqrt(hb) == 0.0 ? 1.0 : sqrt(hb);  // ||b|| = 0 -> 1
        eng_fill<<<elementwise_grid((int64_t)n), kRedBlock, 0, (hipStream_t)stream>>>(
            (int64_t)n, d.x32a, d.p32a);
        NUM_LAUNCH_CHECK();
    }
    NUM_HIP(hipMemcpyAsync(E->d_state + s, &h, sizeof(SysState), hipMemcpyHostToDevice,
                           (hipStream_t)stream));
    NUM_HIP(hipStreamSynchronize((hipStream_t)stream));
    E->h_desc[s] = d;
    E->used[s] = 1;
    E->live[s] = 1;
    E->order.push_back(s);
    E->dirty = true;
    *slot_out = s;
    return NUM_OK;
}

int num_engine_step(num_engine* E, int32_t iters, int32_t* n_active, void* stream)
{
    if (!E) return fail(NUM_EINVAL, "engine is NULL");
    if (iters < 0) return fail(NUM_EINVAL, "iters must be >= 0");
    hipStream_t st = (hipStream_t)stream;
    int dev = 0;
    NUM_HIP(hipGetDevice(&dev));
    if (dev != E->device) return fail(NUM_EINVAL, "engine lives on device %d", E->device);
    if (E->dirty) {
        int rc = eng_rebuild(E, st);
        if (rc) return rc;
    }
    if (E->nact == 0) {
        if (n_active) *n_active = 0;
        return NUM_OK;
    }
    const int chunk = tile_chunk(E->nbt);
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
        eng_kb<<<E->nact, kSpmvBS, 0, st>>>(E->d_act, E->nact, E->d_desc, E->d_state, E->d_pa,
                                            E->d_pb);
        eng_kc<<<E->nch, kRedBlock, 0, st>>>(E->d_ch, E->nch, E->d_desc, E->d_state, E->d_pc);
        eng_kd<<<E->nact, kSpmvBS, 0, st>>>(E->d_act, E->nact, E->d_desc, E->d_state, E->d_pc);
    }
    eng_pack<<<(E->nact + 255) / 256, 256, 0, st>>>(E->d_act, E->nact, E->d_state, E->d_flags);
    NUM_LAUNCH_CHECK();
    NUM_HIP(hipMemcpyAsync(E->h_flags, E->d_flags, (size_t)E->nact * sizeof(int32_t),
                           hipMemcpyDeviceToHost, st));
    NUM_HIP(hipStreamSynchronize(st));
    int32_t still = 0;
    for (int32_t i = 0; i < E->nact; ++i) {
        const int32_t s = E->act_host[i];
        if (E->h_flags[i] == PH_DONE) {
            E->live[s] = 0;
            E->dirty = true;
        } else {
            ++still;
        }
    }
    if (E->dirty) {
        std::vector<int32_t> keep;
        for (int32_t s : E->order)
            if (E->live[s]) keep.push_back(s);
        E->order.swap(keep);
    }
    if (n_active) *n_active = still;
    return NUM_OK;
}

int num_engine_result(num_engine* E, int32_t slot, num_cg_result* res, int32_t* refinements,
                      void* stream)
{
    if (!E || !res) return fail(NUM_EINVAL, "engine or result is NULL");
    if (slot < 0 || slot >= E->cap || !E->used[slot])
        return fail(NUM_EINVAL, "slot %d holds no system", slot);
    hipStream_t st = (hipStream_t)stream;
    SysState h;
    NUM_HIP(hipMemcpyAsync(&h, E->d_state + slot, sizeof(SysState), hipMemcpyDeviceToHost, st));
    NUM_HIP(hipStreamSynchronize(st));
    if (h.phase != PH_DONE) return fail(NUM_ERUNTIME, "system %d is still running", slot);
    const SysDesc& d = E->h_desc[slot];
    if (d.mode == 1 && (h.xsel || h.breakdown)) {
        // the iterate lives in x64[xsel] + x32[xsel]; bring it to the caller's x
        const double* x64 = h.xsel ? d.x64b : d.x;
        const float* x32 = h.xsel ? d.x32b : d.x32a;
        eng_mix_finish<<<elementwise_grid(d.n), kRedBlock, 0, st>>>(d.n, x64, x32, d.x);
        NUM_LAUNCH_CHECK();
        NUM_HIP(hipStreamSynchronize(st));
    }
    res->iters = h.it;
    res->converged = h.converged;
    res->breakdown = h.breakdown;
    res->residual = h.rel;
    if (refinements) *refinements = h.nref;
    return NUM_OK;
}

int num_engine_finished(const num_engine* E, int32_t* slots, int32_t max_slots, int32_t* count)
{
    if (!E || !count) return fail(NUM_EINVAL, "engine or count is NULL");
    int32_t c = 0;
    for (int32_t s = 0; s < E->cap; ++s) {
        if (E->used[s] && !E->live[s]) {
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
    E->h_desc[slot] = SysDesc{};
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
