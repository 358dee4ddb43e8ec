// cg.hip -- unpreconditioned CG (src/cg_solver.hpp:133-196) and the FP32-inner
// / FP64-refined form, with fused kernels and device-resident scalars.
//
// FP64 CG, per iteration (same arithmetic as the reference, in its order):
//   K1 spmv_pap   Ap = A p (row-block SpMV), per-block partials of p_i*Ap_i
//   K2 alpha      pAp = sum partials (fixed order); breakdown guard; alpha = rr/pAp
//   K3 update     x = alpha*p + x ; r = (-alpha)*Ap + r ; partials of r_i*r_i
//   K4 beta       rrn = sum; rel = sqrt(rrn)/||b||; stop if rel < tol (strict),
//                 else beta = rrn/rr, rr = rrn
//   K5 pupdate    p = 1.0*r + beta*p
// Every kernel first reads the device `done` flag and returns at once when it
// is set, so the host can queue iterations ahead and poll the flag
// asynchronously (two chunks in flight) instead of reading back two scalars
// per iteration like the reference.
#include <algorithm>
#include <cmath>

#include "blas1.hpp"
#include "csr.hpp"
#include "spmv_launch.hpp"

namespace numamd {

struct CgState {
    double rr;       // current r.r
    double bnorm;    // ||b|| (0 -> 1)
    double alpha;
    double beta;
    double rel;      // sqrt(rr)/||b|| of the latest residual
    double tol;
    int32_t it;      // completed iterations
    int32_t max_iter;
    int32_t done;
    int32_t converged;
    int32_t breakdown;
    int32_t pad_[5];
};
static_assert(sizeof(CgState) <= 256, "state fits its slot");

static inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// ---------------------------------------------------------------- epilogues
// Set-up: Ap = A x0 ; r = -1*Ap + b ; p = r ; partials r.r and b.b
struct EpiInit {
    static constexpr int kRed = 2;
    const double* __restrict__ b;
    double* __restrict__ r;
    double* __restrict__ p;
    double* partial_a;
    double* partial_b;
    __device__ __forceinline__ bool skip() const { return false; }
    __device__ __forceinline__ void operator()(int row, double v, double& ca, double& cb) const
    {
        const double bi = b[row];
        const double ri = -1.0 * v + bi;
        r[row] = ri;
        p[row] = ri;
        ca = ri * ri;
        cb = bi * bi;
    }
};

struct EpiPAp {
    static constexpr int kRed = 1;
    const double* __restrict__ p;
    double* __restrict__ Ap;
    const CgState* st;
    double* partial_a;
    double* partial_b;
    __device__ __forceinline__ bool skip() const { return st->done != 0; }
    __device__ __forceinline__ void operator()(int row, double v, double& ca, double&) const
    {
        Ap[row] = v;
        ca = p[row] * v;
    }
};

// ------------------------------------------------------------- scalar steps
__global__ __launch_bounds__(kRedBlock) void cg_init_finalize(const double* __restrict__ pa,
                                                              const double* __restrict__ pb,
                                                              int np, CgState* st)
{
    __shared__ double red[2 * kRedBlock / kWave];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedBlock) {
        a += pa[i];
        b += pb[i];
    }
    block_sum2<kRedBlock>(a, b, red);
    if (threadIdx.x == 0) {
        double bn = sqrt(b);
        if (bn == 0.0) bn = 1.0;  // src/cg_solver.hpp:153
        st->bnorm = bn;
        st->rr = a;
        st->rel = sqrt(a) / bn;
        st->it = 0;
        st->breakdown = 0;
        st->converged = 0;
        st->done = 0;
        if (sqrt(a) / bn < st->tol) {  // src/cg_solver.hpp:158-159
            st->converged = 1;
            st->done = 1;
        }
        if (st->max_iter <= 0) st->done = 1;
    }
}

__global__ __launch_bounds__(kRedBlock) void cg_alpha(const double* __restrict__ pa, int np,
                                                      CgState* st)
{
    __shared__ double red[kRedBlock / kWave];
    if (st->done) return;
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedBlock) a += pa[i];
    const double pAp = block_sum<kRedBlock>(a, red);
    if (threadIdx.x == 0) {
        if (fabs(pAp) < 1e-300) {  // src/cg_solver.hpp:84-85
            st->breakdown = 1;
            st->done = 1;
        } else {
            st->alpha = st->rr / pAp;
        }
    }
}

__global__ __launch_bounds__(kRedBlock) void cg_update(int64_t n, const double* __restrict__ p,
                                                       const double* __restrict__ Ap, double* x,
                                                       double* r, const CgState* st,
                                                       double* __restrict__ partial)
{
    __shared__ double red[kRedBlock / kWave];
    if (st->done) return;
    const double alpha = st->alpha;
    const double nalpha = -alpha;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        x[i] = alpha * p[i] + x[i];
        const double ri = nalpha * Ap[i] + r[i];
        r[i] = ri;
        acc += ri * ri;
    }
    const double s = block_sum<kRedBlock>(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedBlock) void cg_beta(const double* __restrict__ pb, int np,
                                                     CgState* st)
{
    __shared__ double red[kRedBlock / kWave];
    if (st->done) return;
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedBlock) a += pb[i];
    const double rrn = block_sum<kRedBlock>(a, red);
    if (threadIdx.x == 0) {
        const int it = st->it + 1;
        st->it = it;
        const double rel = sqrt(rrn) / st->bnorm;
        st->rel = rel;
        if (rel < st->tol) {  // strict <, src/cg_solver.hpp:180-183
            st->converged = 1;
            st->done = 1;
        } else {
            st->beta = rrn / st->rr;
            st->rr = rrn;
            if (it >= st->max_iter) st->done = 1;
        }
    }
}

__global__ __launch_bounds__(kRedBlock) void cg_pupdate(int64_t n, const double* __restrict__ r,
                                                        double* p, const CgState* st)
{
    if (st->done) return;
    const double beta = st->beta;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = 1.0 * r[i] + beta * p[i];
}

__global__ void cg_set_params(CgState* st, double tol, int32_t max_iter)
{
    st->tol = tol;
    st->max_iter = max_iter;
    st->done = 0;
}

// ------------------------------------------------------------ host helpers
struct CgLayout {
    size_t off_r, off_p, off_Ap, off_pa, off_pb, off_state, total;
    int np;  // partial slots
};

static CgLayout cg_layout(const num_csr* A)
{
    CgLayout L;
    const size_t n = (size_t)A->nrows;
    const int np = std::max(A->sched.nblocks, dot_grid(A->nrows));
    size_t o = 0;
    L.off_r = o;  o += align256(n * 8);
    L.off_p = o;  o += align256(n * 8);
    L.off_Ap = o; o += align256(n * 8);
    L.off_pa = o; o += align256((size_t)np * 8);
    L.off_pb = o; o += align256((size_t)np * 8);
    L.off_state = o; o += 256;
    L.total = o;
    L.np = np;
    return L;
}

template <typename T, class Epi>
static void launch_fused_spmv(const num_csr* A, const T* vals, const T* x, const Epi& epi,
                              hipStream_t stream)
{
    launch_tiles<T>(A, A->sched, vals, GatherX<T>{x}, epi, stream);
}

// Pinned status slots for the asynchronous poll (one pair per host thread).
constexpr size_t kSlot = 256;
struct PollSlots {
    unsigned char* h = nullptr;  // 2 slots of kSlot bytes
    hipEvent_t ev[2] = {nullptr, nullptr};
    int device = -1;
};
static thread_local PollSlots g_poll;

static int poll_slots(PollSlots** out)
{
    int dev = 0;
    NUM_HIP(hipGetDevice(&dev));
    if (g_poll.h && g_poll.device != dev) {
        // slots belong to another device: rebuild for this one
        (void)hipHostFree(g_poll.h);
        (void)hipEventDestroy(g_poll.ev[0]);
        (void)hipEventDestroy(g_poll.ev[1]);
        g_poll = PollSlots();
    }
    if (!g_poll.h) {
        NUM_HIP(hipHostMalloc((void**)&g_poll.h, 2 * kSlot, hipHostMallocDefault));
        NUM_HIP(hipEventCreateWithFlags(&g_poll.ev[0], hipEventDisableTiming));
        NUM_HIP(hipEventCreateWithFlags(&g_poll.ev[1], hipEventDisableTiming));
        g_poll.device = dev;
    }
    *out = &g_poll;
    return NUM_OK;
}

// Queue iterations in chunks, keep at most two chunks ahead of the poll.
// `launch_iter()` enqueues one iteration; `stop(const State&)` reads a
// polled copy of the device state.
template <class State, class F, class Stop>
static int drive(hipStream_t stream, const State* d_state, int32_t max_iter, F launch_iter,
                 Stop stop)
{
    static_assert(sizeof(State) <= kSlot, "state fits a poll slot");
    PollSlots* ps = nullptr;
    int rc = poll_slots(&ps);
    if (rc) return rc;
    int32_t launched = 0;
    int chunk = 4;
    int slot = 0;
    bool have_prev = false;
    while (launched < max_iter) {
        const int m = std::min<int32_t>(chunk, max_iter - launched);
        for (int k = 0; k < m; ++k) {
            rc = launch_iter();
            if (rc) return rc;
        }
        launched += m;
        NUM_HIP(hipMemcpyAsync(ps->h + slot * kSlot, d_state, sizeof(State),
                               hipMemcpyDeviceToHost, stream));
        NUM_HIP(hipEventRecord(ps->ev[slot], stream));
        if (have_prev) {
            NUM_HIP(hipEventSynchronize(ps->ev[slot ^ 1]));
            if (stop(*reinterpret_cast<const State*>(ps->h + (slot ^ 1) * kSlot))) break;
        }
        have_prev = true;
        slot ^= 1;
        chunk = std::min(chunk * 2, 32);
    }
    return NUM_OK;
}

// ======================================================================
// Mixed precision: FP64 outer residual / update / test, FP32 inner CG with
// dot products accumulated in FP64.
// ======================================================================
struct MixState {
    double rr;         // inner r.r (FP64 accumulation of FP32 values)
    double rr0;        // inner r.r at the start of this inner solve
    double alpha;
    double beta;
    double bnorm;
    double rel;        // FP64 true residual of the latest outer evaluation
    double tol;
    double inner_tol;
    int32_t inner_it, inner_max, inner_done, inner_total;
    int32_t outer, max_outer, done, converged, breakdown;
    int32_t pad_[3];
};

// Outer residual: Ax = A x (FP64); r = -1*Ax + b; r32 = (float) r; p32 = r32;
// d32 = 0; partials r.r (FP64) and r32.r32 (FP64 accumulation).
struct EpiOuter {
    static constexpr int kRed = 2;
    const double* __restrict__ b;
    float* __restrict__ r32;
    float* __restrict__ p32;
    float* __restrict__ d32;
    double* partial_a;
    double* partial_b;
    __device__ __forceinline__ bool skip() const { return false; }
    __device__ __forceinline__ void operator()(int row, double v, double& ca, double& cb) const
    {
        const double ri = -1.0 * v + b[row];
        const float rf = (float)ri;
        r32[row] = rf;
        p32[row] = rf;
        d32[row] = 0.0f;
        ca = ri * ri;
        cb = (double)rf * (double)rf;
    }
};

struct EpiPq {
    static constexpr int kRed = 1;
    const float* __restrict__ p32;
    float* __restrict__ q32;
    const MixState* st;
    double* partial_a;
    double* partial_b;
    __device__ __forceinline__ bool skip() const { return st->inner_done != 0; }
    __device__ __forceinline__ void operator()(int row, float v, double& ca, double&) const
    {
        q32[row] = v;
        ca = (double)p32[row] * (double)v;
    }
};

__global__ void mix_set_params(MixState* st, double tol, int32_t max_outer, double inner_tol,
                               int32_t inner_max)
{
    st->tol = tol;
    st->inner_tol = inner_tol;
    st->max_outer = max_outer;
    st->inner_max = inner_max;
    st->outer = 0;
    st->inner_total = 0;
    st->done = 0;
    st->converged = 0;
    st->breakdown = 0;
    st->inner_done = 1;
}

__global__ __launch_bounds__(kRedBlock) void mix_outer_check(const double* __restrict__ pa,
                                                             const double* __restrict__ pb,
                                                             int np, MixState* st, int first)
{
    __shared__ double red[2 * kRedBlock / kWave];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedBlock) {
        a += pa[i];
        b += pb[i];
    }
    block_sum2<kRedBlock>(a, b, red);
    if (threadIdx.x == 0) {
        const double rel = sqrt(a) / st->bnorm;
        st->rel = rel;
        if (!first) st->outer += 1;
        if (rel < st->tol) {
            st->converged = 1;
            st->done = 1;
        } else if (st->outer >= st->max_outer || !(rel == rel)) {
            st->done = 1;
        } else {
            st->rr = b;
            st->rr0 = b;
            st->inner_it = 0;
            st->inner_done = (b > 0.0 && st->inner_max > 0) ? 0 : 1;
        }
    }
}

__global__ void mix_set_bnorm(const double* __restrict__ bb, MixState* st)
{
    double bn = sqrt(bb[0]);
    if (bn == 0.0) bn = 1.0;
    st->bnorm = bn;
}

__global__ __launch_bounds__(kRedBlock) void mix_alpha(const double* __restrict__ pa, int np,
                                                       MixState* st)
{
    __shared__ double red[kRedBlock / kWave];
    if (st->inner_done) return;
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedBlock) a += pa[i];
    const double pq = block_sum<kRedBlock>(a, red);
    if (threadIdx.x == 0) {
        if (!(fabs(pq) >= 1e-300)) {
            st->breakdown = 1;
            st->inner_done = 1;
            st->done = 1;
        } else {
            st->alpha = st->rr / pq;
        }
    }
}

__global__ __launch_bounds__(kRedBlock) void mix_update(int64_t n, const float* __restrict__ p,
                                                        const float* __restrict__ q, float* d,
                                                        float* r, const MixState* st,
                                                        double* __restrict__ partial)
{
    __shared__ double red[kRedBlock / kWave];
    if (st->inner_done) return;
    const float a = (float)st->alpha;
    const float na = -a;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        d[i] = a * p[i] + d[i];
        const float ri = na * q[i] + r[i];
        r[i] = ri;
        acc += (double)ri * (double)ri;
    }
    const double s = block_sum<kRedBlock>(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kRedBlock) void mix_beta(const double* __restrict__ pb, int np,
                                                      MixState* st)
{
    __shared__ double red[kRedBlock / kWave];
    if (st->inner_done) return;
    double a = 0.0;
    for (int i = threadIdx.x; i < np; i += kRedBlock) a += pb[i];
    const double rrn = block_sum<kRedBlock>(a, red);
    if (threadIdx.x == 0) {
        st->inner_it += 1;
        st->inner_total += 1;
        if (sqrt(rrn / st->rr0) < st->inner_tol || st->inner_it >= st->inner_max) {
            st->inner_done = 1;
        } else {
            st->beta = rrn / st->rr;
            st->rr = rrn;
        }
    }
}

__global__ __launch_bounds__(kRedBlock) void mix_pupdate(int64_t n, const float* __restrict__ r,
                                                         float* p, const MixState* st)
{
    if (st->inner_done) return;
    const float beta = (float)st->beta;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = r[i] + beta * p[i];
}

__global__ __launch_bounds__(kRedBlock) void mix_xupdate(int64_t n, const float* __restrict__ d,
                                                         double* x)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        x[i] = x[i] + (double)d[i];
}

struct MixLayout {
    size_t off_d, off_r, off_p, off_q, off_pa, off_pb, off_bb, off_state, total;
    int np;
};

static MixLayout mix_layout(const num_csr* A)
{
    MixLayout L;
    const size_t n = (size_t)A->nrows;
    const int np = std::max(A->sched.nblocks, dot_grid(A->nrows));
    size_t o = 0;
    L.off_d = o;  o += align256(n * 4);
    L.off_r = o;  o += align256(n * 4);
    L.off_p = o;  o += align256(n * 4);
    L.off_q = o;  o += align256(n * 4);
    L.off_pa = o; o += align256((size_t)np * 8);
    L.off_pb = o; o += align256((size_t)np * 8);
    L.off_bb = o; o += 256;
    L.off_state = o; o += 256;
    L.total = o;
    L.np = np;
    return L;
}

}  // namespace numamd

using namespace numamd;

extern "C" {

int64_t num_cg_workspace_bytes(const num_csr* A)
{
    if (!A) return -1;
    return (int64_t)cg_layout(A).total;
}

int num_cg_solve(const num_csr* A, const double* d_b, double* d_x, double tol, int32_t max_iter,
                 void* d_work, int64_t work_bytes, num_cg_result* res, void* stream)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (!d_b || !d_x || !res) return fail(NUM_EINVAL, "b, x or result is NULL");
    if (A->nrows != A->ncols) return fail(NUM_EINVAL, "CG needs a square matrix");
    if (max_iter < 0) return fail(NUM_EINVAL, "max_iter must be >= 0");
    int rc = check_device(A);
    if (rc) return rc;
    const CgLayout L = cg_layout(A);
    if (!d_work || work_bytes < (int64_t)L.total)
        return fail(NUM_EINVAL, "workspace too small: need %lld bytes", (long long)L.total);
    hipStream_t st = (hipStream_t)stream;
    char* w = (char*)d_work;
    double* r = (double*)(w + L.off_r);
    double* p = (double*)(w + L.off_p);
    double* Ap = (double*)(w + L.off_Ap);
    double* pa = (double*)(w + L.off_pa);
    double* pb = (double*)(w + L.off_pb);
    CgState* state = (CgState*)(w + L.off_state);
    const int64_t n = A->nrows;
    const int np_spmv = A->sched.nblocks;
    const int gu = dot_grid(n);
    const int ge = elementwise_grid(n);

    cg_set_params<<<1, 1, 0, st>>>(state, tol, max_iter);
    launch_fused_spmv<double>(A, A->vals, d_x, EpiInit{d_b, r, p, pa, pb}, st);
    cg_init_finalize<<<1, kRedBlock, 0, st>>>(pa, pb, np_spmv, state);
    NUM_LAUNCH_CHECK();

    auto iter = [&]() -> int {
        launch_fused_spmv<double>(A, A->vals, p, EpiPAp{p, Ap, state, pa, nullptr}, st);
        cg_alpha<<<1, kRedBlock, 0, st>>>(pa, np_spmv, state);
        cg_update<<<gu, kRedBlock, 0, st>>>(n, p, Ap, d_x, r, state, pb);
        cg_beta<<<1, kRedBlock, 0, st>>>(pb, gu, state);
        cg_pupdate<<<ge, kRedBlock, 0, st>>>(n, r, p, state);
        NUM_LAUNCH_CHECK();
        return NUM_OK;
    };
    rc = drive(st, state, max_iter, iter, [](const CgState& h) { return h.done != 0; });
    if (rc) return rc;
    CgState h;
    NUM_HIP(hipMemcpyAsync(&h, state, sizeof(CgState), hipMemcpyDeviceToHost, st));
    NUM_HIP(hipStreamSynchronize(st));
    res->iters = h.it;
    res->converged = h.converged;
    res->breakdown = h.breakdown;
    res->residual = h.rel;
    if (h.breakdown)
        return fail(NUM_ERUNTIME, "CG: p'Ap is near zero - A may not be SPD");
    return NUM_OK;
}

int64_t num_cg_mixed_workspace_bytes(const num_csr* A)
{
    if (!A) return -1;
    return (int64_t)mix_layout(A).total;
}

int num_cg_solve_mixed(const num_csr* A, const double* d_b, double* d_x, double tol,
                       int32_t max_outer, double inner_tol, int32_t inner_max, void* d_work,
                       int64_t work_bytes, num_mixed_result* res, void* stream)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (!A->vals32) return fail(NUM_EINVAL, "no FP32 values attached (num_csr_attach_f32)");
    if (!d_b || !d_x || !res) return fail(NUM_EINVAL, "b, x or result is NULL");
    if (A->nrows != A->ncols) return fail(NUM_EINVAL, "CG needs a square matrix");
    if (max_outer < 0 || inner_max < 0) return fail(NUM_EINVAL, "iteration limits must be >= 0");
    if (!(inner_tol > 0.0 && inner_tol < 1.0)) return fail(NUM_EINVAL, "inner_tol must be in (0,1)");
    int rc = check_device(A);
    if (rc) return rc;
    const MixLayout L = mix_layout(A);
    if (!d_work || work_bytes < (int64_t)L.total)
        return fail(NUM_EINVAL, "workspace too small: need %lld bytes", (long long)L.total);
    hipStream_t st = (hipStream_t)stream;
    char* w = (char*)d_work;
    float* d32 = (float*)(w + L.off_d);
    float* r32 = (float*)(w + L.off_r);
    float* p32 = (float*)(w + L.off_p);
    float* q32 = (float*)(w + L.off_q);
    double* pa = (double*)(w + L.off_pa);
    double* pb = (double*)(w + L.off_pb);
    double* bb = (double*)(w + L.off_bb);
    MixState* state = (MixState*)(w + L.off_state);
    const int64_t n = A->nrows;
    const int np_spmv = A->sched.nblocks;
    const int gu = dot_grid(n);
    const int ge = elementwise_grid(n);

    mix_set_params<<<1, 1, 0, st>>>(state, tol, max_outer, inner_tol, inner_max);
    rc = launch_dot(n, d_b, d_b, bb, pb, 0, st);
    if (rc) return rc;
    mix_set_bnorm<<<1, 1, 0, st>>>(bb, state);
    NUM_LAUNCH_CHECK();
    PollSlots* ps = nullptr;
    rc = poll_slots(&ps);
    if (rc) return rc;
    MixState h;
    for (int outer = 0;; ++outer) {
        if (outer > 0) mix_xupdate<<<ge, kRedBlock, 0, st>>>(n, d32, d_x);
        launch_fused_spmv<double>(A, A->vals, d_x, EpiOuter{d_b, r32, p32, d32, pa, pb}, st);
        mix_outer_check<<<1, kRedBlock, 0, st>>>(pa, pb, np_spmv, state, outer == 0 ? 1 : 0);
        NUM_LAUNCH_CHECK();
        NUM_HIP(hipMemcpyAsync(&h, state, sizeof(MixState), hipMemcpyDeviceToHost, st));
        NUM_HIP(hipStreamSynchronize(st));
        if (h.done) break;
        auto iter = [&]() -> int {
            launch_fused_spmv<float>(A, A->vals32, p32, EpiPq{p32, q32, state, pa, nullptr}, st);
            mix_alpha<<<1, kRedBlock, 0, st>>>(pa, np_spmv, state);
            mix_update<<<gu, kRedBlock, 0, st>>>(n, p32, q32, d32, r32, state, pb);
            mix_beta<<<1, kRedBlock, 0, st>>>(pb, gu, state);
            mix_pupdate<<<ge, kRedBlock, 0, st>>>(n, r32, p32, state);
            NUM_LAUNCH_CHECK();
            return NUM_OK;
        };
        rc = drive(st, state, inner_max, iter,
                   [](const MixState& s) { return s.inner_done != 0; });
        if (rc) return rc;
    }
    res->outer = h.outer;
    res->inner_total = h.inner_total;
    res->converged = h.converged;
    res->breakdown = h.breakdown;
    res->true_residual = h.rel;
    if (h.breakdown)
        return fail(NUM_ERUNTIME, "CG: p'Ap is near zero - A may not be SPD");
    return NUM_OK;
}

}  // extern "C"
