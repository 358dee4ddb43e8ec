// cg.hip -- unpreconditioned CG (src/cg_solver.hpp:133-196) and its
// FP32-inner / FP64-refined form, with fused kernels and device-resident
// scalars.
//
// Both solvers run as a uniform sequence of 2-launch slots (an SpMV kernel
// KA, an update kernel KC); every scalar (alpha, beta, r.r, iteration count,
// stop flags) lives in a device state record, so the host only queues slots
// and polls the `done` flag asynchronously (two chunks in flight) instead of
// reading back two scalars per iteration like the reference.  No launch is
// spent on a bare global sum: alpha's p.Ap is summed by EVERY workgroup of KC
// from KA's per-tile partials (one fixed order, so all hold the same bits).
// KC's own r.r (<= 256 partials) is folded into the state (convergence test,
// beta) in one of two ways, chosen by measurement:
//  * FP64 CG: by KC's last workgroup (common.hpp: publish / take_ticket),
//    ~2-3 us at the end of KC;
//  * mixed CG: by every workgroup of the next KA (GatherMix's hook), ~0.2 ns
//    of throughput per KA workgroup: cheaper for the 4096 pair workgroups of
//    an FP32 step on 128^3, dearer for the 27k tiles of the FP64 KA on 192^3.
// Separate one-workgroup finalize kernels cost ~4.5 us each per iteration on
// this GPU (two dependent memory round trips plus a launch); a last-workgroup
// ticket in KA itself (27k tickets per launch on 192^3) measured 25% slower.
//
// FP64 CG slot (the reference's arithmetic in its order):
//   KA  p_new = 1*r + beta*p_old is recomputed inside the SpMV gather (the
//       reference's axpy2, cg_solver.hpp:190, fused into the next SpMV), the
//       tile writes its own rows of p_new to the other p buffer, Ap = A p_new,
//       per-tile partials of p_new_i * Ap_i
//   KC  pAp = sum of KA's partials (fixed order, every workgroup); |pAp| <
//       1e-300 guard; alpha = rr/pAp; x = alpha*p_new + x ;
//       r = (-alpha)*Ap + r ; partials of r_i * r_i; the last workgroup:
//       rrn = sum; rel = sqrt(rrn)/||b||; stop if rel < tol (strict), else
//       beta = rrn/rr, rr = rrn; flip the p buffers
//
// Mixed slot (reliable updates): the same two kernels run the CG step in
// FP32 (FP32 matrix copy and vectors, dot products accumulated in FP64,
// scalars in FP64).  The FP32 solution correction x32 accumulates since the
// last refinement.  KA first folds the previous update's r32.r32; when the
// FP32 recursive residual has dropped by a factor delta since the last
// refinement, or looks converged, the fold raises `ru` and that same KA
// refines in FP64 instead of stepping: it computes x64' = x64 + x32 and the
// true residual r = b - A x64' in FP64 (re-seeding the FP32 residual from it
// and restarting x32 at 0); KC then tests the FP64 true residual against tol
// and recomputes beta from the new residual, keeping the search direction
// (the Krylov space survives the refinement).  The mixed state is
// double-buffered: KA reads record A and its workgroup 0 writes B, KC reads B
// and its workgroup 0 writes A, so no workgroup reads a record another
// workgroup of the same launch writes.
// When every row holds <= 32 entries and every column - row offset fits in
// int16 (num_csr_attach_rel16), KA is mix_pair_kernel: one workgroup stages
// two schedule tiles of FP32 values and int16 offsets at once (6 bytes per
// entry instead of 8, twice the data behind one latency chain) and walks a
// refinement's FP64 SpMV as the two tiles in turn.
#include <algorithm>
#include <cmath>

#include "blas1.hpp"
#include "csr.hpp"
#include "spmv_launch.hpp"

namespace numamd {

constexpr int kFin = 1024;  // threads of the update and set-up finalize kernels

static inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// ============================================================== FP64 CG ===
struct CgState {
    double rr;       // r.r of the current residual
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
    int32_t psel;    // p buffer holding the current search direction
    // 1: x still lacks alpha * p of the last iteration (alpha in `alpha`, p
    // the current search direction): the x update runs every other iteration
    int32_t xpend;
    int32_t pad_[3];
};
static_assert(sizeof(CgState) <= 256, "state fits its slot");

// p_new[c] = 1*r[c] + beta*p_old[c], the p update of the previous iteration
template <typename T>
struct GatherPNew {
    const T* __restrict__ r;
    const T* __restrict__ p0;   // p buffer 0
    const T* __restrict__ p1;   // p buffer 1
    const double* beta_src;     // device beta
    const int32_t* sel_src;     // device: which buffer holds p_old
    const T* __restrict__ p_old;
    T beta;
    __device__ __forceinline__ void prepare()
    {
        beta = (T)(*beta_src);
        p_old = *sel_src ? p1 : p0;
    }
    __device__ __forceinline__ T operator()(int c) const { return r[c] + beta * p_old[c]; }
};

// Set-up: Ax = A x0 ; r = -1*Ax + b ; partials r.r and b.b
struct EpiInit {
    static constexpr int kRed = 2;
    const double* __restrict__ b;
    double* __restrict__ r;
    double* partial_a;
    double* partial_b;
    __device__ __forceinline__ bool skip() const { return false; }
    template <class G>
    __device__ __forceinline__ void operator()(int row, double v, double& ca, double& cb,
                                               const G&) const
    {
        const double bi = b[row];
        const double ri = -1.0 * v + bi;
        r[row] = ri;
        ca = ri * ri;
        cb = bi * bi;
    }
};

// Output vectors of the CG kernels are stored non-temporally: the next
// kernel finds them in the Infinity Cache or HBM either way (each XCD's L2 is
// private), and streaming them out avoids a burst of L2 write-backs at the
// end of the kernel (measured: 192^3 CG -2%, 128^3 mixed -4%).
typedef float f4v __attribute__((ext_vector_type(4)));

// KA epilogue: own row of p_new to the other buffer, Ap, p_new_i * Ap_i
struct EpiPAp {
    static constexpr int kRed = 1;
    double* __restrict__ p0;
    double* __restrict__ p1;
    double* __restrict__ Ap;
    const CgState* st;
    double* partial_a;
    double* partial_b;
    __device__ __forceinline__ bool skip() const { return st->done != 0; }
    __device__ __forceinline__ void operator()(int row, double v, double& ca, double& cb,
                                               const GatherPNew<double>& g) const
    {
        own(row, v, ca, cb, g(row));
    }
    // pn = p_new[row], computed by the caller (the windowed walk has it in LDS)
    __device__ __forceinline__ void own(int row, double v, double& ca, double&, double pn) const
    {
        __builtin_nontemporal_store(pn, (st->psel ? p0 : p1) + row);
        __builtin_nontemporal_store(v, Ap + row);
        ca = pn * v;
    }
};

__global__ __launch_bounds__(kFin) void cg_init_finalize(const double* __restrict__ pa,
                                                         const double* __restrict__ pb, int np,
                                                         CgState* st, uint32_t* ticket,
                                                         double tol, int32_t max_iter)
{
    __shared__ double red[kFin / kWave];
    const double a = sum_partials_block<kFin>(pa, np, red);
    const double b = sum_partials_block<kFin>(pb, np, red);
    if (threadIdx.x == 0) {
        double bn = sqrt(b);
        if (bn == 0.0) bn = 1.0;  // src/cg_solver.hpp:153
        *ticket = 0u;
        st->tol = tol;
        st->max_iter = max_iter;
        st->bnorm = bn;
        st->rr = a;
        st->rel = sqrt(a) / bn;
        st->alpha = 0.0;
        st->beta = 0.0;  // first p_new = r + 0*p_old, p_old zero-filled
        st->psel = 0;
        st->xpend = 0;
        st->it = 0;
        st->breakdown = 0;
        st->converged = 0;
        st->done = 0;
        if (sqrt(a) / bn < tol) {  // early exit, src/cg_solver.hpp:158-159
            st->converged = 1;
            st->done = 1;
        }
        if (max_iter <= 0) st->done = 1;
    }
}

// Grid of the update kernels: at most 256 workgroups of kFin threads (their
// tickets and the last workgroup's sum stay cheap).
constexpr int kMaxUpdateGroups = 256;
static int update_grid(int64_t n)
{
    int64_t g = (n + (int64_t)kFin * 4 - 1) / ((int64_t)kFin * 4);
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, kMaxUpdateGroups));
}

// KC: alpha from KA's partials (every workgroup, one fixed order), then
// r = (-alpha)*Ap + r and r.r partials; the last workgroup folds the r.r into
// the state.  The x update x = alpha*p_new + x runs every other iteration,
// for two iterations at once: x = alpha*p_new + (alpha_prev*p_prev + x),
// the same two rounded updates in the same order (bit-identical to updating
// every iteration), reading and writing x once per two iterations (saves 8
// of the 80 vector bytes per row and iteration).  p_prev is the other p
// buffer, which the next KA overwrites only after this KC.  When the solve
// ends with an update pending, cg_x_fixup applies it.  A thread's first
// elements are loaded together with the partials, so alpha's sum overlaps
// their latency.
__global__ __launch_bounds__(kFin) void cg_kc(int64_t n, const double* __restrict__ p0,
                                              const double* __restrict__ p1,
                                              const double* __restrict__ Ap, double* x,
                                              double* r, CgState* st,
                                              const double* __restrict__ pa, int npa,
                                              double* pb, uint32_t* ticket)
{
    __shared__ double red[kFin / kWave];
    __shared__ double s_alpha;
    __shared__ int s_brk, s_last;
    if (st->done) return;  // read by every workgroup before any ticket is drawn
    const int t = threadIdx.x;
    // the fields the last workgroup's fold needs, read up front
    const double rr = st->rr;
    const double bnorm = st->bnorm, tol = st->tol;
    const int it = st->it + 1, max_iter = st->max_iter, psel = st->psel;
    const bool both = st->xpend != 0;  // this iteration completes a pending x update
    const double alpha_prev = st->alpha;
    const double* __restrict__ p = psel ? p0 : p1;   // p_new was written to !psel
    const double* __restrict__ pp = psel ? p1 : p0;  // p of the previous iteration
    const int64_t stride = (int64_t)gridDim.x * kFin;
    const int64_t i0 = (int64_t)blockIdx.x * kFin + t;
    // this thread's first kPre elements, loaded together with the partials
    constexpr int kPre = 4;
    double av[kPre], rv[kPre], pv[kPre], qv[kPre], xv[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const int64_t i = i0 + u * stride;
        if (i < n) {
            av[u] = Ap[i];
            rv[u] = r[i];
            if (both) {
                pv[u] = p[i];
                qv[u] = pp[i];
                xv[u] = x[i];
            }
        }
    }
    const double pAp = sum_partials_block<kFin, 32>(pa, npa, red);
    if (t == 0) {
        // the guard of src/cg_solver.hpp:84-85 (std::abs(pAp) < 1e-300): a NaN
        // p.Ap is no breakdown there, so it runs on to a non-converged result
        // with a NaN residual, as the unguarded cg_solve_local
        // (cg_solver.hpp:165-167) returns
        const int brk = fabs(pAp) < 1e-300;
        s_brk = brk;
        s_alpha = brk ? 0.0 : rr / pAp;
    }
    __syncthreads();
    const bool brk = s_brk != 0;
    const double alpha = s_alpha;
    if (!brk) {
        const double nalpha = -alpha;
        double acc = 0.0;
#pragma unroll
        for (int u = 0; u < kPre; ++u) {
            const int64_t i = i0 + u * stride;
            if (i < n) {
                if (both)
                    __builtin_nontemporal_store(alpha * pv[u] + (alpha_prev * qv[u] + xv[u]),
                                                x + i);
                const double rn = nalpha * av[u] + rv[u];
                __builtin_nontemporal_store(rn, r + i);
                acc += rn * rn;
            }
        }
        int64_t i = i0 + kPre * stride;
        if (both) {
            for (; i + stride < n; i += 2 * stride) {
                const int64_t k = i + stride;
                const double pu = p[i], pw = p[k], qu = pp[i], qw = pp[k];
                const double au = Ap[i], aw = Ap[k];
                const double xu = x[i], xw = x[k], ru = r[i], rw = r[k];
                __builtin_nontemporal_store(alpha * pu + (alpha_prev * qu + xu), x + i);
                __builtin_nontemporal_store(alpha * pw + (alpha_prev * qw + xw), x + k);
                const double rnu = nalpha * au + ru;
                const double rnw = nalpha * aw + rw;
                __builtin_nontemporal_store(rnu, r + i);
                __builtin_nontemporal_store(rnw, r + k);
                acc += rnu * rnu;
                acc += rnw * rnw;
            }
        } else {
            for (; i + stride < n; i += 2 * stride) {
                const int64_t k = i + stride;
                const double au = Ap[i], aw = Ap[k], ru = r[i], rw = r[k];
                const double rnu = nalpha * au + ru;
                const double rnw = nalpha * aw + rw;
                __builtin_nontemporal_store(rnu, r + i);
                __builtin_nontemporal_store(rnw, r + k);
                acc += rnu * rnu;
                acc += rnw * rnw;
            }
        }
        if (i < n) {
            if (both) x[i] = alpha * p[i] + (alpha_prev * pp[i] + x[i]);
            const double rn = nalpha * Ap[i] + r[i];
            r[i] = rn;
            acc += rn * rn;
        }
        const double tot = block_sum<kFin>(acc, red);
        if (t == 0) publish(pb + blockIdx.x, tot);
    }
    if (!take_ticket(ticket, &s_last)) return;
    // the last workgroup: every other one has read the state
    if (brk) {
        // no update this iteration; a pending one (alpha_prev, p of the
        // current direction) is left to cg_x_fixup
        if (t == 0) {
            st->breakdown = 1;
            st->done = 1;
        }
        return;
    }
    const double rrn = sum_published_block<kFin>(pb, gridDim.x, red);
    if (t == 0) {
        st->alpha = alpha;
        st->it = it;
        st->psel = psel ^ 1;
        st->xpend = both ? 0 : 1;
        const double rel = sqrt(rrn) / bnorm;
        st->rel = rel;
        if (rel < tol) {  // strict <, src/cg_solver.hpp:180-183
            st->converged = 1;
            st->done = 1;
        } else {
            st->beta = rrn / rr;
            st->rr = rrn;
            if (it >= max_iter) st->done = 1;
        }
    }
}

// x = alpha*p + x for an x update still pending when the solve ended
// (alpha in the state, p the current search direction).
__global__ __launch_bounds__(kRedBlock) void cg_x_fixup(int64_t n, const double* __restrict__ p0,
                                                        const double* __restrict__ p1,
                                                        double* __restrict__ x,
                                                        const CgState* __restrict__ st)
{
    if (!st->xpend) return;
    const double alpha = st->alpha;
    const double* __restrict__ p = st->psel ? p1 : p0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        x[i] = alpha * p[i] + x[i];
}

struct CgLayout {
    size_t off_r, off_p0, off_p1, off_Ap, off_pa, off_pb, off_state, off_ticket, total;
};

static CgLayout cg_layout(const num_csr* A)
{
    CgLayout L;
    const size_t n = (size_t)A->nrows;
    const int np = std::max(A->sched.nblocks, kMaxUpdateGroups);
    size_t o = 0;
    L.off_r = o;  o += align256(n * 8);
    L.off_p0 = o; o += align256(n * 8);
    L.off_p1 = o; o += align256(n * 8);
    L.off_Ap = o; o += align256(n * 8);
    L.off_pa = o; o += align256((size_t)np * 8);
    L.off_pb = o; o += align256((size_t)np * 8);
    L.off_state = o; o += 256;
    L.off_ticket = o; o += 256;
    L.total = o;
    return L;
}

// =========================================================== polling ===
constexpr size_t kSlot = 256;
struct PollSlots {
    unsigned char* h = nullptr;  // 2 slots of kSlot bytes, pinned
    hipEvent_t ev[2] = {nullptr, nullptr};
    int device = -1;
};
static thread_local PollSlots g_poll;

static int poll_slots(PollSlots** out)
{
    int dev = 0;
    NUM_HIP(hipGetDevice(&dev));
    if (g_poll.h && g_poll.device != dev) {
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

// Queue slots in chunks, keeping at most two chunks ahead of the poll.
// `launch_slot()` enqueues one slot; `stop(const State&)` reads a polled copy
// of the device state.  Kernels skip themselves once the state says done, so
// slots queued past the end cost only their launch.
template <class State, class F, class Stop>
static int drive(hipStream_t stream, const State* d_state, int64_t max_slots, F launch_slot,
                 Stop stop)
{
    static_assert(sizeof(State) <= kSlot, "state fits a poll slot");
    PollSlots* ps = nullptr;
    int rc = poll_slots(&ps);
    if (rc) return rc;
    int64_t launched = 0;
    int chunk = 4;
    int slot = 0;
    bool have_prev = false;
    while (launched < max_slots) {
        const int m = (int)std::min<int64_t>(chunk, max_slots - launched);
        for (int k = 0; k < m; ++k) {
            rc = launch_slot();
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

// ============================================================ mixed CG ===
// A value every lane holds alike, moved to scalar registers.
__device__ __forceinline__ double uniform_d(double v)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ int uniform_i(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Workgroup 0 publishes the folded state: the record goes through LDS as
// 32-bit words and thread 0 folds it in place.
template <class S, class Fold>
__device__ __forceinline__ void publish_state(const S* in, S* out, Fold fold)
{
    __shared__ S s;
    const int t = threadIdx.x;
    constexpr int W = (int)(sizeof(S) / 4);
    if (t < W) reinterpret_cast<uint32_t*>(&s)[t] = reinterpret_cast<const uint32_t*>(in)[t];
    __syncthreads();
    if (t == 0) fold(s);
    __syncthreads();
    if (t < W) reinterpret_cast<uint32_t*>(out)[t] = reinterpret_cast<const uint32_t*>(&s)[t];
}

// State records are copied between device memory and LDS as 32-bit words.
template <class S>
constexpr int state_words() { return (int)(sizeof(S) / 4); }

struct MixState {
    double rr;       // FP32 residual r32.r32 (FP64 accumulation) of the current step
    double rr_ref;   // rr right after the last refinement
    double alpha;
    double beta;
    double bnorm;
    double tol;
    double delta;    // refinement trigger: residual drop since the last one
    double rel;      // FP64 true residual of the latest refinement
    int32_t it;      // FP32 CG steps taken
    int32_t max_iter;
    int32_t done, converged, breakdown;
    int32_t ru;       // 1: the next SpMV kernel refines in FP64 instead of stepping
    int32_t pend;     // 1: the r32.r32 partials of the last update are not folded in yet
    int32_t nref;     // refinements done (FP64 residual evaluations)
    int32_t psel;     // p buffer holding the current search direction
    int32_t xsel;     // x64 / x32 buffer pair holding the current iterate
    int32_t final_;   // 1: the pending refinement ends the solve
    int32_t pad_;
};
static_assert(sizeof(MixState) <= 256, "state fits its slot");

struct MixArgs {
    const double* b;
    double* x64a;   // user x
    double* x64b;   // workspace
    float* x32a;
    float* x32b;
    float* r32;
    float* p0;
    float* p1;
    float* q;
    double* pa;     // SpMV kernel partials: p.q, or the refinement's FP64 r.r
    double* pb;     // update kernel partials: r32.r32
    double* pc;     // the refinement's FP32 r32.r32 partials
    int npb;        // update kernel workgroups (<= kSpmvBS)
    MixState* sa;   // read by the SpMV kernel, written by the update kernel
    MixState* sb;   // read by the update kernel, written by the SpMV kernel
};

// FP64 gather of the refined iterate x64 + x32
struct GatherXsum {
    const double* __restrict__ x64;
    const float* __restrict__ x32;
    __device__ __forceinline__ void prepare() {}
    __device__ __forceinline__ double operator()(int c) const
    {
        return x64[c] + (double)x32[c];
    }
};

// Refinement epilogue: x64' = x64 + x32 (to the other buffer), x32' = 0,
// r = -1*Ax + b in FP64, r32 = (float) r; partials r.r and r32.r32
struct EpiRefine {
    static constexpr int kRed = 2;
    const double* __restrict__ b;
    double* __restrict__ x64n;
    float* __restrict__ x32n;
    float* __restrict__ r32;
    double* partial_a;
    double* partial_b;
    __device__ __forceinline__ bool skip() const { return false; }
    __device__ __forceinline__ void operator()(int row, double v, double& ca, double& cb,
                                               const GatherXsum& g) const
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

// The scalar step after the SpMV kernel: the refinement's FP64 convergence
// test and new beta (ru), else alpha = rr / p.q.
__device__ __forceinline__ void mix_after_ka(bool ru, double sa, double sb, MixState& s)
{
    if (ru) {
        const double rr64 = sa, rr32 = sb;
        const double rel = sqrt(rr64) / s.bnorm;
        s.rel = rel;
        s.xsel ^= 1;
        s.nref += 1;
        s.ru = 0;
        if (rel < s.tol) {
            s.converged = 1;
            s.done = 1;
        } else if (s.final_ || !(rel == rel)) {
            s.done = 1;
        } else {
            s.beta = rr32 / s.rr;  // rr = r.r the last p was built with
            s.rr = rr32;
            s.rr_ref = rr32;
        }
        return;
    }
    const double pq = sa;
    if (!(fabs(pq) >= 1e-300)) {
        s.breakdown = 1;
        s.done = 1;
    } else {
        s.alpha = s.rr / pq;
    }
}

// Fold the r32.r32 of an update into the state: beta, or schedule a
// refinement.  mix_fold_view gives what a workgroup needs (beta, psel, ru);
// mix_after_kc applies the whole fold to the published record.
// The p buffer of the next FP32 step: it flips when an update is pending.
// It does not depend on the update's r.r, so a kernel can issue its loads of
// p_old before it has summed the partials.
__device__ __forceinline__ int mix_next_psel(const MixState& s)
{
    return (s.done || !s.pend) ? s.psel : s.psel ^ 1;
}

__device__ __forceinline__ void mix_fold_view(double rrn, const MixState& s, double& beta,
                                              int& psel, int& ru, int& fin)
{
    beta = s.beta;
    psel = mix_next_psel(s);
    ru = s.ru;
    fin = s.final_;
    if (s.done || !s.pend) return;
    const double est = sqrt(rrn) / s.bnorm;
    if (s.it + 1 >= s.max_iter) {
        ru = 1;
        fin = 1;
    } else if (est < s.tol || sqrt(rrn) < s.delta * sqrt(s.rr_ref) || !(rrn == rrn)) {
        ru = 1;  // beta comes from the refined residual
    } else {
        beta = rrn / s.rr;
    }
}

__device__ __forceinline__ void mix_after_kc(double rrn, MixState& s)
{
    if (s.done || !s.pend) return;
    double beta;
    int psel, ru, fin;
    mix_fold_view(rrn, s, beta, psel, ru, fin);
    s.pend = 0;
    s.it += 1;
    s.psel = psel;
    s.ru = ru;
    s.final_ = fin;
    if (!ru) {
        s.beta = beta;
        s.rr = rrn;
    }
}

// FP32 step gather: p_new[c] = 1*r32[c] + beta*p_old[c].  Its hook folds the
// update kernel's partials into the state read from `in`: every thread sums
// the <= 256 partials (one barrier) and applies mix_fold_view to the record,
// and the results move to scalar registers; workgroup 0 also writes the
// folded record to `out` (publish_state).  It walks only when the result is
// an FP32 step (not done, no refinement due); `done`, `ru` and `xsel` tell
// the kernel what to do otherwise.
struct GatherMix {
    const float* __restrict__ r;
    const float* __restrict__ p0;
    const float* __restrict__ p1;
    const MixState* in;
    MixState* out;
    const double* pb;
    int npb;
    mutable const float* __restrict__ p_old;
    mutable float beta;
    mutable double part;
    mutable int psel, done, ru, xsel;
    __device__ __forceinline__ void prepare() const {}
    __device__ __forceinline__ void issue() const
    {
        const int t = threadIdx.x;
        part = t < npb ? pb[t] : 0.0;
    }
    __device__ __forceinline__ bool resolve() const
    {
        __shared__ double red[kSpmvBS / kWave];
        const double rrn = uniform_d(block_total_once<kSpmvBS>(part, red));
        double b;
        int ps, r_, fin;
        mix_fold_view(rrn, *in, b, ps, r_, fin);
        beta = (float)uniform_d(b);
        psel = uniform_i(ps);
        ru = uniform_i(r_);
        done = uniform_i(in->done);
        xsel = uniform_i(in->xsel);
        p_old = psel ? p1 : p0;
        if (blockIdx.x == 0) publish_state(in, out, [rrn](MixState& st) { mix_after_kc(rrn, st); });
        return !done && !ru;
    }
    __device__ __forceinline__ float operator()(int c) const { return r[c] + beta * p_old[c]; }
};

// FP32 step epilogue: p_new row to the other buffer, q = A p_new, p_new.q
struct EpiPq {
    static constexpr int kRed = 1;
    float* __restrict__ p0;
    float* __restrict__ p1;
    float* __restrict__ q;
    double* partial_a;
    double* partial_b;
    __device__ __forceinline__ bool skip() const { return false; }
    __device__ __forceinline__ void operator()(int row, float v, double& ca, double&,
                                               const GatherMix& g) const
    {
        const float pn = g(row);
        __builtin_nontemporal_store(pn, (g.psel ? p0 : p1) + row);
        __builtin_nontemporal_store(v, q + row);
        ca = (double)pn * (double)v;
    }
};

// The refinement a mixed SpMV workgroup runs instead of its FP32 step over
// tiles [t0, t1) (two at most), one after the other; partials to pa / pc[w].
template <int NPT, bool NT64>
__device__ __forceinline__ void mix_refine_tiles(int t0, int t1, int w,
                                                 const int32_t* __restrict__ rowptr,
                                                 const int32_t* __restrict__ colind,
                                                 const double* __restrict__ v64,
                                                 const TileEntry* __restrict__ sched,
                                                 const MixArgs& a, int xs, double* lv,
                                                 int32_t* lc, double* red)
{
    GatherXsum g{xs ? a.x64b : a.x64a, xs ? a.x32b : a.x32a};
    EpiRefine epi{a.b, xs ? a.x64a : a.x64b, xs ? a.x32a : a.x32b, a.r32, nullptr, nullptr};
    double ca = 0.0, cb = 0.0;
    for (int t = t0; t < t1; ++t) {
        __syncthreads();  // the LDS is staged again for this tile
        double xa, xb;
        spmv_tile_body<double, kSpmvBS, NPT, NT64>(t, rowptr, colind, v64, g, sched, epi, lv, lc,
                                                   xa, xb);
        ca += xa;
        cb += xb;
    }
    block_sum2<kSpmvBS>(ca, cb, red);
    if (threadIdx.x == 0) {
        a.pa[w] = ca;
        a.pc[w] = cb;
    }
}

// SpMV kernel of the mixed slot: FP32 step, or the FP64 refinement.
// Waves per SIMD the LDS of a mixed SpMV workgroup allows (160 KB per CU):
// the register budget is held to it (the FP64 refinement path would
// otherwise set it for the FP32 steps too).
template <int NPT>
constexpr int mix_waves()
{
    constexpr int w = (160 * 1024) / (kSpmvBS * NPT * 12 + 1024) * (kSpmvBS / kWave) / 4;
    return w > 7 ? 7 : (w < 1 ? 1 : w);
}

template <int NPT, bool NT32, bool NT64>
__global__ __launch_bounds__(kSpmvBS, mix_waves<NPT>()) void mix_spmv_kernel(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const double* __restrict__ v64, const float* __restrict__ v32,
    const TileEntry* __restrict__ sched, int ntiles, int chunk, MixArgs a)
{
    __shared__ double lv[kSpmvBS * NPT];
    __shared__ int32_t lc[kSpmvBS * NPT];
    __shared__ double red[2 * kSpmvBS / kWave];
    const int tile = chunk_tile(blockIdx.x, ntiles, chunk);
    if (tile < 0) return;
    GatherMix g{a.r32, a.p0, a.p1, a.sa, a.sb, a.pb, a.npb};
    EpiPq epi{a.p0, a.p1, a.q, nullptr, nullptr};
    double ca, cb;
    spmv_tile_body<float, kSpmvBS, NPT, NT32>(tile, rowptr, colind, v32, g, sched, epi,
                                              reinterpret_cast<float*>(lv), lc, ca, cb);
    if (g.done) return;
    if (g.ru) {
        mix_refine_tiles<NPT, NT64>(tile, tile + 1, tile, rowptr, colind, v64, sched, a, g.xsel,
                                    lv, lc, red);
        return;
    }
    const double s = block_sum<kSpmvBS>(ca, red);
    if (threadIdx.x == 0) a.pa[tile] = s;
}

// SpMV kernel of the mixed slot for a matrix with int16 column offsets and
// short rows only: one workgroup per PAIR of schedule tiles.  FP32 step:
// both tiles are staged at once (FP32 values + int16 offsets fill the LDS of
// one FP64 tile); FP64 refinement: the two tiles in turn.  Partials: one per
// pair.
// PAT: the offsets and lengths come as row patterns (rel_pat / rel_pid: one
// id byte per row into the pair's table) instead of 2 bytes per entry and
// rowptr; same bits.
template <int NPT, bool NT32, bool NT64, bool PAT>
__global__ __launch_bounds__(kSpmvBS, mix_waves<NPT>()) void mix_pair_kernel(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colind,
    const double* __restrict__ v64, const float* __restrict__ v32,
    const int16_t* __restrict__ rel, const uint64_t* __restrict__ rpat,
    const uint8_t* __restrict__ rpid, const TileEntry* __restrict__ sched, int ntiles,
    int npairs, int chunk, MixArgs a)
{
    __shared__ double lv[kSpmvBS * NPT];   // FP32 pairs: 2 * NPT * BS floats
    __shared__ int32_t lc[kSpmvBS * NPT];  // FP32 pairs: 2 * NPT * BS int16
    __shared__ double red[2 * kSpmvBS / kWave];
    __shared__ uint64_t lpat[PAT ? 2 * kRelPat : 1];
    const int p = chunk_tile(blockIdx.x, npairs, chunk);
    if (p < 0) return;
    const int t0 = 2 * p, t1 = min(2 * p + 2, ntiles);
    GatherMix g{a.r32, a.p0, a.p1, a.sa, a.sb, a.pb, a.npb};
    EpiPq epi{a.p0, a.p1, a.q, nullptr, nullptr};
    double ca, cb;
    if constexpr (PAT)
        spmv_relpat_run<kSpmvBS, NPT, 2, NT32, 2 * kRelPat>(
            sched[t0], sched[t1], rpid, rpat + (size_t)p * relpat_words<2 * kRelPat>(), v32, g,
            epi, reinterpret_cast<float*>(lv), lpat, ca, cb);
    else
        spmv_rel16_run<kSpmvBS, NPT, 2, NT32>(sched[t0], sched[t1], rowptr, rel, v32, g, epi,
                                              reinterpret_cast<float*>(lv),
                                              reinterpret_cast<int16_t*>(lc), ca, cb);
    if (g.done) return;
    if (g.ru) {
        mix_refine_tiles<NPT, NT64>(t0, t1, p, rowptr, colind, v64, sched, a, g.xsel, lv, lc,
                                    red);
        return;
    }
    const double s = block_sum<kSpmvBS>(ca, red);
    if (threadIdx.x == 0) a.pa[p] = s;
}

// A tile's window record as scalars (the compiler keeps named fields in
// registers; a record indexed inside loops went to LDS).
struct XwWin {
    int l0, l1, l2, l3, e0, e1, e2, e3;
};
__device__ __forceinline__ XwWin xw_win(const int32_t* rec)
{
    return {rec[0], rec[1], rec[2], rec[3], rec[4], rec[5], rec[6], rec[7]};
}
__device__ __forceinline__ int xw_win_col(int s, const XwWin& w)
{
    return xw_col(s, w.l0, w.l1, w.l2, w.l3, w.e0, w.e1, w.e2);
}

// One tile's share of mix_sell2_kernel's FP32 step for thread t: its row's
// FP32 values and pattern (or slots), and its share of the tile's r32 and
// p_old windows, all issued before the fold's barrier; the window values
// are parked in LDS before the fold (park), so that their registers are free
// while both tiles' values wait for the step.
template <bool NT32, bool PAT>
struct Sell2Tile {
    typedef float f2v __attribute__((ext_vector_type(2)));
    static constexpr int NXW = (kXwSlots + kSpmvBS - 1) / kSpmvBS;
    XwWin w;
    int64_t vb;
    int width, total, r0, nr, id, len;
    uint64_t q0, q1;
    f2v v[kSellW / 2];
    float vl;
    float wr[NXW], wp[NXW];
    // wr, wp to LDS: slots [0, total) of the tile's r32 and p_old windows
    __device__ __forceinline__ void park(float* lr, float* lp) const
    {
#pragma unroll
        for (int u = 0; u < NXW; ++u) {
            const int s = (int)threadIdx.x + u * kSpmvBS;
            if (s < total) {
                lr[s] = wr[u];
                lp[s] = wp[u];
            }
        }
    }
    __device__ __forceinline__ void issue(int tile, const int32_t* __restrict__ recs,
                                          const uint8_t* __restrict__ lens,
                                          const uint16_t* __restrict__ code,
                                          const uint8_t* __restrict__ pid,
                                          const float* __restrict__ v32,
                                          const TileEntry* __restrict__ sched,
                                          const float* __restrict__ r32,
                                          const float* __restrict__ p_old)
    {
        constexpr int BS = kSpmvBS;
        const int t = threadIdx.x;
        const int32_t* rec = recs + tile * kXwRec;
        w = xw_win(rec);
        vb = (int64_t)rec[kXwBase] * BS;
        width = rec[kXwWidth];
        total = w.e3;
        r0 = sched[tile].row;
        nr = sched[tile + 1].row - r0;
        const bool active = t < nr;
        const int n2 = width >> 1;
        const f2v* __restrict__ vp = reinterpret_cast<const f2v*>(v32 + vb);
#pragma unroll
        for (int h = 0; h < kSellW / 2; ++h)
            if (h < n2) v[h] = ld_mat<NT32>(vp + h * BS + t);
        vl = (width & 1) ? ld_mat<NT32>(v32 + vb + 2 * n2 * BS + t) : 0.0f;
        q0 = q1 = 0;
        id = len = 0;
        if constexpr (PAT) {
            id = active ? ((int)ld_mat<NT32>(pid + tile * (BS / 2) + (t >> 1)) >> (4 * (t & 1))) &
                              15
                        : 0;
        } else {
            len = active ? (int)lens[r0 + t] : 0;
            sell_slots<NT32, BS>(code + vb, width, t, q0, q1);
        }
        const int sm = max(total - 1, 0);
#pragma unroll
        for (int u = 0; u < NXW; ++u) {
            const int col = xw_win_col(min(t + u * BS, sm), w);
            wr[u] = r32[col];
            wp[u] = p_old[col];
        }
    }
    // after the barrier: the row's slots and length from the staged table
    __device__ __forceinline__ void resolve(const uint64_t* lpat)
    {
        if constexpr (PAT) {
            q0 = lpat[2 * id];
            q1 = lpat[2 * id + 1];
            len = (int)threadIdx.x < nr ? (int)((q0 >> 12) & 15) : 0;
        }
    }
    __device__ __forceinline__ int slot(int k) const
    {
        if constexpr (PAT) return (sell_slot(k, q0, q1) & 0xfff) - kXwPatBias + (int)threadIdx.x;
        else return sell_slot(k, q0, q1);
    }
    __device__ __forceinline__ float value(int k) const
    {
        return k < 2 * (width >> 1) ? ((k & 1) ? v[k / 2].y : v[k / 2].x) : vl;
    }
};

// SpMV kernel of the mixed slot over the windowed sliced layout (the FP32
// and FP64 values in sliced order; the row patterns, or one window-slot
// code per entry, shared by both; see spmv_sell_run), for a PAIR of
// schedule tiles per workgroup (thread t takes row t of each).  FP32 step:
// everything that does not depend on the fold is issued before its barrier
// -- both tiles' FP32 values and patterns (or slots), and r32 and p_old
// over their windows (parked in LDS) -- and the update kernel's partials
// are folded once per pair; after it, p_new = r + beta*p_old is formed in
// place over the windows and each row is summed from registers and LDS in
// index order (the same bits per row as the staged kernels), p_new of the
// own row taken from the windows.  One tile per workgroup instead ran 2-10
// % slower on every grid measured (its step is bound by its latency chain,
// not its bytes: 78 us per 192^3 step with 2-byte slots or row patterns
// alike).  A refinement (ru) walks the two tiles in turn with the FP64
// values over x64 + x32 windows.  Partials: one per pair.
template <bool NT32, bool NT64, bool PAT>
__global__ __launch_bounds__(kSpmvBS, 7) void mix_sell2_kernel(
    const int32_t* __restrict__ recs, const uint8_t* __restrict__ lens,
    const uint16_t* __restrict__ code, const uint64_t* __restrict__ pat,
    const int32_t* __restrict__ tix, const uint8_t* __restrict__ pid,
    const float* __restrict__ v32, const double* __restrict__ v64,
    const TileEntry* __restrict__ sched, int ntiles, int npairs, int chunk, MixArgs a)
{
    constexpr int BS = kSpmvBS;
    constexpr int NXW = Sell2Tile<NT32, PAT>::NXW;
    constexpr int K = kXwSlots;
    // FP32 step: [r32 | p_old] windows of tile A, then of tile B, as floats
    // (the r32 part becomes p_new); a refinement: one tile's FP64 windows
    __shared__ double xw64[2 * K];
    __shared__ double red[2 * BS / kWave];
    __shared__ uint64_t lpat[PAT ? 4 * kXwPat : 1];
    float* xw32 = reinterpret_cast<float*>(xw64);
    const int p = chunk_tile(blockIdx.x, npairs, chunk);
    if (p < 0) return;
    const int t = threadIdx.x;
    const int ta = 2 * p;
    const bool two = ta + 1 < ntiles;  // uniform: the last pair may hold one tile
    const float* __restrict__ p_old = mix_next_psel(*a.sa) ? a.p1 : a.p0;
    Sell2Tile<NT32, PAT> A, B;
    A.issue(ta, recs, lens, code, pid, v32, sched, a.r32, p_old);
    if (two) B.issue(ta + 1, recs, lens, code, pid, v32, sched, a.r32, p_old);
    if constexpr (PAT) {
        if (t < 2 * kXwPat) {
            lpat[t] = pat[(size_t)tix[ta] * 2 * kXwPat + t];
            if (two) lpat[2 * kXwPat + t] = pat[(size_t)tix[ta + 1] * 2 * kXwPat + t];
        }
    }
    GatherMix g{a.r32, a.p0, a.p1, a.sa, a.sb, a.pb, a.npb};
    g.issue();
    A.park(xw32, xw32 + K);
    if (two) B.park(xw32 + 2 * K, xw32 + 3 * K);
    g.resolve();  // a barrier; workgroup 0 publishes the folded state
    if (g.done) return;
    A.resolve(lpat);
    if (two) B.resolve(lpat + 2 * kXwPat);
    if (g.ru) {
        // FP64 refinement, the tiles in turn over the FP64 windows
        const int xs = g.xsel;
        GatherXsum gx{xs ? a.x64b : a.x64a, xs ? a.x32b : a.x32a};
        EpiRefine epi{a.b, xs ? a.x64a : a.x64b, xs ? a.x32a : a.x32b, a.r32, nullptr, nullptr};
        double ca = 0.0, cb = 0.0;
        for (int h = 0; h < (two ? 2 : 1); ++h) {
            const Sell2Tile<NT32, PAT>& T = h ? B : A;
            double w64[kSellW];
#pragma unroll
            for (int k = 0; k < kSellW; ++k)
                if (k < T.width) w64[k] = ld_mat<NT64>(v64 + T.vb + k * BS + t);
            __syncthreads();  // the LDS is filled again for this tile
#pragma unroll
            for (int u = 0; u < NXW; ++u) {
                const int s = t + u * BS;
                if (s < T.total) xw64[s] = gx(xw_win_col(s, T.w));
            }
            __syncthreads();
            if (t < T.nr) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < kSellW; ++k)
                    if (k < T.len) acc += w64[k] * xw64[T.slot(k)];
                double xa, xb;
                epi(T.r0 + t, acc, xa, xb, gx);
                ca += xa;
                cb += xb;
            }
        }
        block_sum2<BS>(ca, cb, red);
        if (t == 0) {
            a.pa[p] = ca;
            a.pc[p] = cb;
        }
        return;
    }
    const float beta = g.beta;
    // p_new = r + beta * p_old over each window slot, in place of r (each
    // thread the slots it parked)
#pragma unroll
    for (int u = 0; u < NXW; ++u) {
        const int s = t + u * BS;
        if (s < A.total) xw32[s] = xw32[s] + beta * xw32[K + s];
        if (two && s < B.total) xw32[2 * K + s] = xw32[2 * K + s] + beta * xw32[3 * K + s];
    }
    __syncthreads();
    double ca = 0.0;
    for (int h = 0; h < (two ? 2 : 1); ++h) {
        const Sell2Tile<NT32, PAT>& T = h ? B : A;
        const float* xwt = xw32 + 2 * h * K;
        if (t < T.nr) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < kSellW; ++k)
                if (k < T.len) acc += T.value(k) * xwt[T.slot(k)];
            // p_new of the own row: a window slot (the own rows are a
            // window range of every tile of a square matrix)
            const int row = T.r0 + t;
            const int os = xw_slot(row, T.w.l0, T.w.l1, T.w.l2, T.w.l3, T.w.e0, T.w.e1, T.w.e2,
                                   T.w.e3);
            const float pn = os >= 0 ? xwt[os] : g(row);
            __builtin_nontemporal_store(pn, (g.psel ? a.p0 : a.p1) + row);
            __builtin_nontemporal_store(acc, a.q + row);
            ca += (double)pn * (double)acc;
        }
    }
    const double sum = block_sum<BS>(ca, red);
    if (t == 0) a.pa[p] = sum;
}

__global__ void mix_init(const double* __restrict__ bb, MixState* st, double tol, double delta,
                         int32_t max_iter)
{
    if (threadIdx.x != 0) return;
    double bn = sqrt(bb[0]);
    if (bn == 0.0) bn = 1.0;
    st->rr = INFINITY;  // first refinement: beta = rr'/inf = 0
    st->rr_ref = 0.0;
    st->alpha = 0.0;
    st->beta = 0.0;
    st->bnorm = bn;
    st->tol = tol;
    st->delta = delta;
    st->rel = 0.0;
    st->it = 0;
    st->max_iter = max_iter;
    st->done = 0;
    st->converged = 0;
    st->breakdown = 0;
    st->ru = 1;  // slot 0 evaluates r = b - A x0 in FP64
    st->pend = 0;
    st->nref = 0;
    st->psel = 0;
    st->xsel = 0;
    st->final_ = max_iter <= 0 ? 1 : 0;
}

// Update kernel of the mixed slot.  After a refinement: its convergence test
// and beta from the SpMV kernel's partials, no update.  After an FP32 step:
// alpha = rr / p.q, then x32 += alpha p_new, r32 -= alpha q (FP32) over
// float4 groups (n4 = n / 4 of them; the n % 4 tail by thread 0 of
// workgroup 0) and per-workgroup partials of r32.r32 (FP64).
__global__ __launch_bounds__(kFin) void mix_kc(int64_t n, int npa, MixArgs a)
{
    __shared__ double red[kFin / kWave];
    __shared__ MixState s;
    const int t = threadIdx.x;
    const MixState* in = a.sb;
    const bool live = in->done == 0;
    const bool ru = in->ru != 0;
    const bool step = live && !ru;
    if (t < state_words<MixState>())
        reinterpret_cast<uint32_t*>(&s)[t] = reinterpret_cast<const uint32_t*>(in)[t];
    // a thread's first kPre float4 groups, loaded together with the partials
    const float* __restrict__ p = in->psel ? a.p0 : a.p1;
    float* __restrict__ x32 = in->xsel ? a.x32b : a.x32a;
    const float* __restrict__ q = a.q;
    float* __restrict__ r = a.r32;
    const int64_t n4 = n >> 2;
    const int64_t stride = (int64_t)gridDim.x * kFin;
    const int64_t i0 = (int64_t)blockIdx.x * kFin + t;
    constexpr int kPre = 2;
    float4 pv[kPre], qv[kPre], xv[kPre], rv[kPre];
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const int64_t i = i0 + u * stride;
        if (step && i < n4) {
            pv[u] = reinterpret_cast<const float4*>(p)[i];
            qv[u] = reinterpret_cast<const float4*>(q)[i];
            xv[u] = reinterpret_cast<const float4*>(x32)[i];
            rv[u] = reinterpret_cast<const float4*>(r)[i];
        }
    }
    double s1 = 0.0, s2 = 0.0;
    if (live) {
        s1 = sum_partials_block<kFin, 32>(a.pa, npa, red);  // its barriers publish s
        if (ru) s2 = sum_partials_block<kFin, 32>(a.pc, npa, red);
    } else {
        __syncthreads();
    }
    if (t == 0 && live) {
        mix_after_ka(ru, s1, s2, s);
        if (!ru) s.pend = s.done ? 0 : 1;
    }
    __syncthreads();
    if (blockIdx.x == 0 && t < state_words<MixState>())
        reinterpret_cast<uint32_t*>(a.sa)[t] = reinterpret_cast<const uint32_t*>(&s)[t];
    if (!step || s.done) return;
    const float al = (float)s.alpha;
    const float nal = -al;
    double acc = 0.0;
    auto update = [&](int64_t i, const float4& pw, const float4& qw, float4 xw, float4 rw) {
        xw.x = al * pw.x + xw.x;
        xw.y = al * pw.y + xw.y;
        xw.z = al * pw.z + xw.z;
        xw.w = al * pw.w + xw.w;
        rw.x = nal * qw.x + rw.x;
        rw.y = nal * qw.y + rw.y;
        rw.z = nal * qw.z + rw.z;
        rw.w = nal * qw.w + rw.w;
        __builtin_nontemporal_store(f4v{xw.x, xw.y, xw.z, xw.w}, reinterpret_cast<f4v*>(x32) + i);
        __builtin_nontemporal_store(f4v{rw.x, rw.y, rw.z, rw.w}, reinterpret_cast<f4v*>(r) + i);
        acc += (double)rw.x * (double)rw.x;
        acc += (double)rw.y * (double)rw.y;
        acc += (double)rw.z * (double)rw.z;
        acc += (double)rw.w * (double)rw.w;
    };
#pragma unroll
    for (int u = 0; u < kPre; ++u) {
        const int64_t i = i0 + u * stride;
        if (i < n4) update(i, pv[u], qv[u], xv[u], rv[u]);
    }
    for (int64_t i = i0 + kPre * stride; i < n4; i += stride)
        update(i, reinterpret_cast<const float4*>(p)[i], reinterpret_cast<const float4*>(q)[i],
               reinterpret_cast<const float4*>(x32)[i], reinterpret_cast<const float4*>(r)[i]);
    if (blockIdx.x == 0 && t == 0) {
        for (int64_t k = n4 * 4; k < n; ++k) {
            x32[k] = al * p[k] + x32[k];
            const float rk = nal * q[k] + r[k];
            r[k] = rk;
            acc += (double)rk * (double)rk;
        }
    }
    const double tot = block_sum<kFin>(acc, red);
    if (t == 0) a.pb[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kRedBlock) void mix_finish(int64_t n, const double* __restrict__ x64,
                                                        const float* __restrict__ x32,
                                                        double* __restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = x64[i] + (double)x32[i];
}

__global__ void zero_f32(int64_t n, float* __restrict__ a)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        a[i] = 0.0f;
}

struct MixLayout {
    size_t off_x64b, off_x32a, off_x32b, off_r, off_p0, off_p1, off_q, off_pa, off_pb, off_pc,
        off_bb, off_sa, off_sb, total;
};

static MixLayout mix_layout(const num_csr* A)
{
    MixLayout L;
    const size_t n = (size_t)A->nrows;
    // pb also holds the b.b partials of the set-up dot product
    const int np = std::max({A->sched.nblocks, dot_grid(A->nrows), kSpmvBS});
    size_t o = 0;
    L.off_x64b = o; o += align256(n * 8);
    L.off_x32a = o; o += align256(n * 4);
    L.off_x32b = o; o += align256(n * 4);
    L.off_r = o;  o += align256(n * 4);
    L.off_p0 = o; o += align256(n * 4);
    L.off_p1 = o; o += align256(n * 4);
    L.off_q = o;  o += align256(n * 4);
    L.off_pa = o; o += align256((size_t)np * 8);
    L.off_pb = o; o += align256((size_t)np * 8);
    L.off_pc = o; o += align256((size_t)np * 8);
    L.off_bb = o; o += 256;
    L.off_sa = o; o += 256;
    L.off_sb = o; o += 256;
    L.total = o;
    return L;
}

template <int NPT>
static void launch_mix_spmv_npt(const num_csr* A, const MixArgs& a, hipStream_t st)
{
    const Schedule& s = A->sched;
    const int grid = chunk_grid(s.nblocks, s.chunk);
    const TileEntry* tiles = static_cast<const TileEntry*>(s.d_tiles);
    if (s.nt32) {
        if (s.nt64)
            mix_spmv_kernel<NPT, true, true><<<grid, kSpmvBS, 0, st>>>(
                A->rowptr, A->colind, A->vals, A->vals32, tiles, s.nblocks, s.chunk, a);
        else
            mix_spmv_kernel<NPT, true, false><<<grid, kSpmvBS, 0, st>>>(
                A->rowptr, A->colind, A->vals, A->vals32, tiles, s.nblocks, s.chunk, a);
    } else {
        if (s.nt64)
            mix_spmv_kernel<NPT, false, true><<<grid, kSpmvBS, 0, st>>>(
                A->rowptr, A->colind, A->vals, A->vals32, tiles, s.nblocks, s.chunk, a);
        else
            mix_spmv_kernel<NPT, false, false><<<grid, kSpmvBS, 0, st>>>(
                A->rowptr, A->colind, A->vals, A->vals32, tiles, s.nblocks, s.chunk, a);
    }
}

template <int NPT>
static void launch_mix_pair_npt(const num_csr* A, const MixArgs& a, hipStream_t st)
{
    const Schedule& s = A->sched;
    const int npairs = (s.nblocks + 1) / 2;
    const int pc = plane_chunk(A, 2.0 * A->nrows / std::max(1, s.nblocks));
    // plane-aligned from 2 pairs up (1024 rows), as the SpMV schedule from 3
    // tiles: a 2-D 3000^2 stencil (0.7 pairs) solves in 41.5 ms with
    // size-based chunks against 43.3 ms with 1-pair chunks
    const int chunk = pc >= 2 && npairs >= 8 * pc ? pc : tile_chunk(npairs);
    const int grid = chunk_grid(npairs, chunk);
    const TileEntry* tiles = static_cast<const TileEntry*>(s.d_tiles);
    const bool pat = A->rel_pat && A->mix_path != 2;
    auto go = [&](auto k) {
        k<<<grid, kSpmvBS, 0, st>>>(A->rowptr, A->colind, A->vals, A->vals32, A->rel16,
                                    A->rel_pat, A->rel_pid, tiles, s.nblocks, npairs, chunk, a);
    };
    if (pat) {
        if (s.nt32) {
            if (s.nt64) go(mix_pair_kernel<NPT, true, true, true>);
            else go(mix_pair_kernel<NPT, true, false, true>);
        } else {
            if (s.nt64) go(mix_pair_kernel<NPT, false, true, true>);
            else go(mix_pair_kernel<NPT, false, false, true>);
        }
    } else {
        if (s.nt32) {
            if (s.nt64) go(mix_pair_kernel<NPT, true, true, false>);
            else go(mix_pair_kernel<NPT, true, false, false>);
        } else {
            if (s.nt64) go(mix_pair_kernel<NPT, false, true, false>);
            else go(mix_pair_kernel<NPT, false, false, false>);
        }
    }
}

// The paired int16 path applies when the offsets are attached, no tile is a
// single long row and the staging depth is one the pair kernel is built for.
static bool mix_paired(const num_csr* A)
{
    return A->rel16 && A->sched.long_rows == 0 && A->max_row_nnz <= kShortRow &&
           (A->sched.npt == 5 || A->sched.npt == 7);
}

// The windowed sliced path applies (by default) when the layout carries the
// FP32 values, the FP32 matrix streams from HBM and its dominant long-range
// column offset is at least kMixSellOffset rows.  Per FP32 step (both with
// row patterns), windowed against the staged pair kernel: 192^3 (offset
// 36,864) 116 against 144 us, 96x96x1200 (9,216) 180 against 157 us, 2-D
// 4000^2 (4,000) 232 against 206 us.  An FP32 matrix that stays in the
// Infinity Cache runs the staged pair kernel (128^3 diffusion: 35.0 against
// 39.4 us).  num_csr_set_mixed_path overrides the choice.
constexpr int32_t kMixSellOffset = 16384;
static bool mix_sliced(const num_csr* A)
{
    if (!A->xw || !A->sell_vals32) return false;
    if (A->mix_path >= 0) return A->mix_path == 1;  // 0 and 2: the staged kernels
    return A->sched.nt32 != 0 && A->offset_d >= kMixSellOffset;
}

// SpMV-kernel partials the update kernel sums: one per pair for the paired
// kernels (windowed or staged), else one per tile.
static int mix_partials(const num_csr* A)
{
    const int nb = A->sched.nblocks;
    return mix_sliced(A) || mix_paired(A) ? (nb + 1) / 2 : nb;
}

static void launch_mix_spmv(const num_csr* A, const MixArgs& a, hipStream_t st)
{
    if (mix_sliced(A)) {
        const Schedule& s = A->sched;
        const int npairs = (s.nblocks + 1) / 2;
        // plane-aligned pair chunks, as the staged pair kernel
        const int pc = plane_chunk(A, 2.0 * A->nrows / std::max(1, s.nblocks));
        const int chunk = pc >= 2 && npairs >= 8 * pc ? pc : tile_chunk(npairs);
        const int grid = chunk_grid(npairs, chunk);
        const TileEntry* tiles = static_cast<const TileEntry*>(s.d_tiles);
        auto go = [&](auto k) {
            k<<<grid, kSpmvBS, 0, st>>>(A->xw, A->sell_lens, A->sell_codes, A->sell_pat,
                                        A->sell_tix, A->sell_pid, A->sell_vals32, A->sell_vals,
                                        tiles, s.nblocks, npairs, chunk, a);
        };
        // the row patterns of the FP64 layout serve the FP32 values too
        // (same tiles, same slots)
        if (A->sell_pat) {
            if (s.nt32) {
                if (s.nt64) go(mix_sell2_kernel<true, true, true>);
                else go(mix_sell2_kernel<true, false, true>);
            } else {
                if (s.nt64) go(mix_sell2_kernel<false, true, true>);
                else go(mix_sell2_kernel<false, false, true>);
            }
        } else {
            if (s.nt32) {
                if (s.nt64) go(mix_sell2_kernel<true, true, false>);
                else go(mix_sell2_kernel<true, false, false>);
            } else {
                if (s.nt64) go(mix_sell2_kernel<false, true, false>);
                else go(mix_sell2_kernel<false, false, false>);
            }
        }
        return;
    }
    if (mix_paired(A)) {
        if (A->sched.npt == 5) launch_mix_pair_npt<5>(A, a, st);
        else launch_mix_pair_npt<7>(A, a, st);
        return;
    }
    switch (A->sched.npt) {
    case 5: launch_mix_spmv_npt<5>(A, a, st); break;
    case 7: launch_mix_spmv_npt<7>(A, a, st); break;
    case 12: launch_mix_spmv_npt<12>(A, a, st); break;
    default: launch_mix_spmv_npt<16>(A, a, st); break;
    }
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
    if (((uintptr_t)d_work & 255) != 0)
        return fail(NUM_EINVAL, "workspace must be 256-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    char* w = (char*)d_work;
    double* r = (double*)(w + L.off_r);
    double* p0 = (double*)(w + L.off_p0);
    double* p1 = (double*)(w + L.off_p1);
    double* Ap = (double*)(w + L.off_Ap);
    double* pa = (double*)(w + L.off_pa);
    double* pb = (double*)(w + L.off_pb);
    CgState* state = (CgState*)(w + L.off_state);
    uint32_t* ticket = (uint32_t*)(w + L.off_ticket);
    const int64_t n = A->nrows;
    const int np_spmv = A->sched.nblocks;
    const int gk = update_grid(n);

    // p_old of the first step must be finite (it is multiplied by beta = 0)
    rc = num_fill(n, 0.0, p0, stream);
    if (rc) return rc;
    launch_tiles<double>(A, A->sched, A->vals, GatherX<double>{d_x},
                         EpiInit{d_b, r, pa, pb}, st);
    cg_init_finalize<<<1, kFin, 0, st>>>(pa, pb, np_spmv, state, ticket, tol, max_iter);
    NUM_LAUNCH_CHECK();

    auto slot = [&]() -> int {
        GatherPNew<double> g{r, p0, p1, &state->beta, &state->psel, nullptr, 0.0};
        launch_tiles<double>(A, A->sched, A->vals, g, EpiPAp{p0, p1, Ap, state, pa, nullptr},
                             st);
        cg_kc<<<gk, kFin, 0, st>>>(n, p0, p1, Ap, d_x, r, state, pa, np_spmv, pb, ticket);
        NUM_LAUNCH_CHECK();
        return NUM_OK;
    };
    rc = drive(st, state, max_iter, slot, [](const CgState& h) { return h.done != 0; });
    if (rc) return rc;
    if (n > 0) {
        cg_x_fixup<<<elementwise_grid(n), kRedBlock, 0, st>>>(n, p0, p1, d_x, state);
        NUM_LAUNCH_CHECK();
    }
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

int num_csr_set_mixed_path(num_csr* A, int32_t mode)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (mode < -1 || mode > 2)
        return fail(NUM_EINVAL, "mode must be -1, 0, 1 or 2, got %d", mode);
    A->mix_path = mode;
    return NUM_OK;
}

int64_t num_cg_mixed_workspace_bytes(const num_csr* A)
{
    if (!A) return -1;
    return (int64_t)mix_layout(A).total;
}

int num_cg_solve_mixed(const num_csr* A, const double* d_b, double* d_x, double tol,
                       int32_t max_iter, double delta, void* d_work, int64_t work_bytes,
                       num_mixed_result* res, void* stream)
{
    if (!A) return fail(NUM_EINVAL, "matrix is NULL");
    if (!A->vals32) return fail(NUM_EINVAL, "no FP32 values attached (num_csr_attach_f32)");
    if (!d_b || !d_x || !res) return fail(NUM_EINVAL, "b, x or result is NULL");
    if (A->nrows != A->ncols) return fail(NUM_EINVAL, "CG needs a square matrix");
    if (max_iter < 0) return fail(NUM_EINVAL, "max_iter must be >= 0");
    if (!(delta > 0.0 && delta < 1.0)) return fail(NUM_EINVAL, "delta must be in (0, 1)");
    int rc = check_device(A);
    if (rc) return rc;
    const MixLayout L = mix_layout(A);
    if (!d_work || work_bytes < (int64_t)L.total)
        return fail(NUM_EINVAL, "workspace too small: need %lld bytes", (long long)L.total);
    if (((uintptr_t)d_work & 255) != 0)
        return fail(NUM_EINVAL, "workspace must be 256-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    char* w = (char*)d_work;
    MixArgs a;
    a.b = d_b;
    a.x64a = d_x;
    a.x64b = (double*)(w + L.off_x64b);
    a.x32a = (float*)(w + L.off_x32a);
    a.x32b = (float*)(w + L.off_x32b);
    a.r32 = (float*)(w + L.off_r);
    a.p0 = (float*)(w + L.off_p0);
    a.p1 = (float*)(w + L.off_p1);
    a.q = (float*)(w + L.off_q);
    a.pa = (double*)(w + L.off_pa);
    a.pb = (double*)(w + L.off_pb);
    a.pc = (double*)(w + L.off_pc);
    double* bb = (double*)(w + L.off_bb);
    a.sa = (MixState*)(w + L.off_sa);
    a.sb = (MixState*)(w + L.off_sb);
    const int64_t n = A->nrows;
    const int gk = update_grid(n);
    const int ge = elementwise_grid(n);
    a.npb = gk;

    zero_f32<<<ge, kRedBlock, 0, st>>>(n, a.x32a);
    zero_f32<<<ge, kRedBlock, 0, st>>>(n, a.p0);
    rc = launch_dot(n, d_b, d_b, bb, a.pb, 0, st);
    if (rc) return rc;
    mix_init<<<1, 64, 0, st>>>(bb, a.sa, tol, delta, max_iter);
    NUM_LAUNCH_CHECK();
    const int np_spmv = mix_partials(A);
    auto slot = [&]() -> int {
        launch_mix_spmv(A, a, st);
        mix_kc<<<gk, kFin, 0, st>>>(n, np_spmv, a);
        NUM_LAUNCH_CHECK();
        return NUM_OK;
    };
    // FP32 steps plus one slot per refinement (at most one per step, plus
    // the first evaluation and a final one)
    rc = drive(st, a.sa, 2 * (int64_t)max_iter + 2, slot,
               [](const MixState& h) { return h.done != 0; });
    if (rc) return rc;
    MixState h;
    NUM_HIP(hipMemcpyAsync(&h, a.sa, sizeof(MixState), hipMemcpyDeviceToHost, st));
    NUM_HIP(hipStreamSynchronize(st));
    // the iterate lives in x64[xsel] + x32[xsel]; bring it to the caller's x
    const double* xcur = h.xsel ? a.x64b : a.x64a;
    const float* x32cur = h.xsel ? a.x32b : a.x32a;
    if (xcur != d_x || h.breakdown) {
        mix_finish<<<ge, kRedBlock, 0, st>>>(n, xcur, x32cur, d_x);
        NUM_LAUNCH_CHECK();
        NUM_HIP(hipStreamSynchronize(st));
    }
    res->outer = h.nref;
    res->inner_total = h.it;
    res->converged = h.converged;
    res->breakdown = h.breakdown;
    res->true_residual = h.rel;
    if (h.breakdown)
        return fail(NUM_ERUNTIME, "CG: p'Ap is near zero - A may not be SPD");
    return NUM_OK;
}

}  // extern "C"
