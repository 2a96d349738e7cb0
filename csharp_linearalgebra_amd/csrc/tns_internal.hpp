// Internal declarations of the MI355X backend: the context (Tensile's
// Workspace on the device), status plumbing, the device buffer pool, and the
// device-side entry points each kernel file provides. Nothing here crosses
// the C ABI; tensile_hip.h is the public surface.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "tensile_hip.h"
#include "tns_scalar.hpp"

namespace tns {

using cplx = zd;  // (re, im), the layout of System.Numerics.Complex

// A thrown status: the C ABI catches it at the boundary and returns the code,
// so no C++ exception ever crosses extern "C".
struct Error {
    int status;
    std::string message;
};

[[noreturn]] void fail(int status, const std::string& message);
void check_hip(hipError_t e, const char* what);
#define TNS_HIP(call) ::tns::check_hip((call), #call)

// Grow-only caching pool of device buffers, keyed by rounded size. Every
// internal temporary comes from here, so a run of calls (the block steps of
// a factorization, the squarings of an exponential) allocates once. A buffer
// returns to the pool when its handle dies; reuse is ordered by the one
// stream every kernel of a context runs on.
class Pool {
public:
    explicit Pool(size_t limit_bytes) : limit_(limit_bytes) {}
    ~Pool();
    void* take(size_t bytes);
    void give(void* p, size_t bytes);
    void trim();
    void release_all();
    size_t held() const { return held_; }
    size_t limit() const { return limit_; }
    void set_limit(size_t l) { limit_ = l; }
    // Bytes the pool keeps between calls: after a call, free blocks beyond
    // it are returned to the device, the largest first.
    size_t retention() const { return retain_; }
    void set_retention(size_t r) { retain_ = r; }
    void shrink_to_retention();

private:
    static size_t round(size_t bytes);
    std::multimap<size_t, void*> free_;
    std::map<void*, size_t> sizes_;
    size_t held_ = 0;
    size_t limit_;
    size_t retain_ = ~size_t(0);
};

// Host threads a context keeps for the estimators' host steps (many
// independent runs advanced in lockstep between device products): run()
// hands out the indices [0, count) to the caller and the workers and returns
// when all are done, rethrowing the first exception. Started on first use,
// stopped with the context; one run() at a time (the context's lock).
class HostPool {
public:
    HostPool() = default;
    HostPool(const HostPool&) = delete;
    HostPool& operator=(const HostPool&) = delete;
    ~HostPool() { stop(); }
    void run(int count, const std::function<void(int)>& fn);
    void stop();

private:
    void work();
    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable wake_, done_;
    const std::function<void(int)>* job_ = nullptr;
    int count_ = 0, busy_ = 0;
    unsigned long long generation_ = 0;
    bool stopping_ = false;
    std::atomic<int> next_{0};
    std::exception_ptr error_;
};

struct Context;

// A device buffer on loan from the pool.
template <typename T>
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(Context* ctx, size_t count);
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept { *this = std::move(o); }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        release();
        ctx_ = o.ctx_; p_ = o.p_; bytes_ = o.bytes_;
        o.ctx_ = nullptr; o.p_ = nullptr; o.bytes_ = 0;
        return *this;
    }
    T* get() const { return p_; }
    operator T*() const { return p_; }
    void release();

private:
    Context* ctx_ = nullptr;
    T* p_ = nullptr;
    size_t bytes_ = 0;
};

struct Context {
    int device = 0;
    int max_lds = 65536;  // shared memory per workgroup on this device, bytes
    int lp_mode = 0;      // LP_FP32 or LP_BF16: products inside sgetrf/sgetrs
    // While a graph is being captured, released pool buffers are kept here
    // (the graph's nodes still use them; the graph's owner returns them).
    std::vector<std::pair<void*, size_t>>* hold = nullptr;
    // The batch autotuner's choices: batch shape -> (streams, 4M tile).
    std::map<std::string, std::pair<int, int>> batch_tuning;
    hipStream_t stream = nullptr;
    // The LU's look-ahead: a second stream for the trailing updates while
    // the next panel is factored on `stream`, and the events that order
    // them (a panel and its interchanges on `stream` done; the trailing
    // update of the even and of the odd steps done).
    hipStream_t side = nullptr;
    hipEvent_t ev_panel = nullptr;
    hipEvent_t ev_rest[2] = {nullptr, nullptr};
    Pool pool{size_t(64) << 30};
    // Small fixed device areas for reductions and flags.
    double* d_scal = nullptr;   // 64 doubles
    int* d_int = nullptr;       // 64 ints
    // Page-locked staging for host matrices (grow-only, at most
    // kStageMax bytes; larger matrices pass through it in column chunks).
    char* h_stage = nullptr;
    size_t h_stage_bytes = 0;
    // Page-locked host memory for read-backs and the estimators' rounds: a
    // copy between pageable memory and the device costs 70-170 us of wall
    // time for a few kilobytes, against a few for page-locked memory.
    // h_mail (kMailBytes, fixed) takes scalar read-backs; h_pin (grow-only)
    // the panels of one estimator or refinement loop at a time.
    char* h_mail = nullptr;
    char* h_pin = nullptr;
    size_t h_pin_bytes = 0;
    hipEvent_t ev_pin = nullptr;  // the last upload from h_pin
    bool pin_pending = false;
    std::string last_error;
    std::mutex gate;            // the Workspace lock (Workspace.cs:136-139)
    // This device's copy of the no-allocation LU's file-scope scratch.
    int* noalloc_scratch = nullptr;
    // The one-launch triangular solve's exchange words: a pool block kept
    // between calls (grown on demand, zeroed when taken), and the sequence
    // number of its last launch (each launch tags its words with its own).
    unsigned long long* flow = nullptr;
    size_t flow_bytes = 0;
    unsigned flow_seq = 0;
    bool capturing = false;  // a graph capture is recording `stream`
    bool no_alloc = false;   // inside the no-allocation entry point
    HostPool host;           // host threads for the estimators' lockstep steps
};

// Points ctx->stream at s for the guard's life: the helpers launch on
// ctx->stream.
struct OnStream {
    Context* ctx;
    hipStream_t saved;
    OnStream(Context* c, hipStream_t s) : ctx(c), saved(c->stream) { c->stream = s; }
    ~OnStream() { ctx->stream = saved; }
    OnStream(const OnStream&) = delete;
    OnStream& operator=(const OnStream&) = delete;
};

// Points ctx->hold at v for the guard's life: buffers released meanwhile
// are kept there instead of returning to the pool.
struct HoldReleases {
    Context* ctx;
    std::vector<std::pair<void*, size_t>>* saved;
    HoldReleases(Context* c, std::vector<std::pair<void*, size_t>>* v) : ctx(c), saved(c->hold) { c->hold = v; }
    ~HoldReleases() { ctx->hold = saved; }
    HoldReleases(const HoldReleases&) = delete;
    HoldReleases& operator=(const HoldReleases&) = delete;
};

template <typename T>
DevBuf<T>::DevBuf(Context* ctx, size_t count) : ctx_(ctx), bytes_(sizeof(T) * (count ? count : 1))
{
    p_ = static_cast<T*>(ctx->pool.take(bytes_));
}

template <typename T>
void DevBuf<T>::release()
{
    if (ctx_ && p_) {
        if (ctx_->hold) ctx_->hold->emplace_back(p_, bytes_);
        else ctx_->pool.give(p_, bytes_);
    }
    ctx_ = nullptr;
    p_ = nullptr;
    bytes_ = 0;
}

constexpr size_t kMailBytes = 4096;
// Device -> host for at most kMailBytes through the page-locked mailbox;
// waits for the copy (and everything before it on ctx->stream).
void read_back(Context* ctx, void* h, const void* d, size_t bytes);
// The page-locked panel area, grown to at least `bytes`, once the last
// upload from it (write_pinned) has been read; the pointer stays valid
// until the next call.
void* pinned_panel(Context* ctx, size_t bytes);
// Host -> device through the panel area, asynchronous on ctx->stream.
void write_pinned(Context* ctx, void* d, const void* h, size_t bytes);

// Whether p is device (or managed) memory this process can hand to a kernel.
bool is_device_pointer(const void* p);

void sync(Context* ctx);

// ---------------------------------------------------------------------------
// Device-side operations. Every pointer here is device memory; every launch
// goes on ctx->stream and returns without waiting.
// ---------------------------------------------------------------------------

// C := beta*C + alpha*A*B, FP64 MFMA (gemm.hip).
void dgemm(Context* ctx, int m, int n, int k, double alpha, const double* a, int lda,
           const double* b, int ldb, double beta, double* c, int ldc);
// Complex C := beta*C + alpha*A*B by the 4M method on FP64 MFMA (gemm.hip).
void zgemm(Context* ctx, int m, int n, int k, cplx alpha, const cplx* a, int lda,
           const cplx* b, int ldb, cplx beta, cplx* c, int ldc);
// op(A) = A (ta 0), A^T (1) or A^H (2, complex only).
void dgemm_ex(Context* ctx, int ta, int m, int n, int k, double alpha, const double* a, int lda,
              const double* b, int ldb, double beta, double* c, int ldc);
void zgemm_ex(Context* ctx, int ta, int m, int n, int k, cplx alpha, const cplx* a, int lda,
              const cplx* b, int ldb, cplx beta, cplx* c, int ldc);
// `batch` products of one shape, untransposed, product z at a + z sa,
// b + z sb, c + z sc, in one launch. kmode marks a triangular operand, whose
// zero part the tiles skip: 1 A lower, 2 A upper, 3 B upper, 4 B lower
// (0 none); the zero half must be stored as zeros (the tiles skip most of
// it, not all).
void dgemm_batch(Context* ctx, int m, int n, int k, double alpha, const double* a, int lda, long long sa,
                 const double* b, int ldb, long long sb, double beta, double* c, int ldc, long long sc, int batch,
                 int kmode);
void zgemm_batch(Context* ctx, int m, int n, int k, cplx alpha, const cplx* a, int lda, long long sa, const cplx* b,
                 int ldb, long long sb, cplx beta, cplx* c, int ldc, long long sc, int batch, int kmode);
// Tile configuration override for the autotuner: -1 = heuristic.
void gemm_force_config(int real_cfg, int cplx_cfg);
int gemm_config_count(bool complex_);

// The LU phase split (LuPhaseTimings, Lu.cs:92-164), in milliseconds of
// device time; a factorization adds its split to the fields.
struct LuPhases {
    double panel_ms = 0, swaps_ms = 0, trsm_ms = 0, gemm_ms = 0, total_ms = 0;
    int steps = 0;
};

// Device-time split of a sequence of launches on ctx->stream: mark(b)
// records an event, and the time from the previous mark to it is charged
// to bucket b (b < 0: charged to none). read() waits for the last mark.
struct EventClock {
    Context* ctx;
    std::vector<hipEvent_t> ev;
    std::vector<int> bucket;
    explicit EventClock(Context* c) : ctx(c) {}
    EventClock(const EventClock&) = delete;
    EventClock& operator=(const EventClock&) = delete;
    ~EventClock()
    {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    void mark(int b);
    // Adds each bucket's milliseconds to out[b] (b < nb) and the span from
    // the first mark to the last to *total.
    void read(double* out, int nb, double* total);
};

// LU (lu.hip). d_piv: min(m,n) LAPACK-order interchanges (0-based);
// d_perm: m entries, row i of P*A is row d_perm[i] of A; d_info: the first
// exactly-zero pivot column, -1, or LU_ABORTED if a panel's cross-workgroup
// exchange gave up waiting (a device fault, never a property of A).
// phases, if given, receives the phase split (the call then waits for the
// factorization). work, if given, is device scratch of lu_work_ints(m, n,
// nb) ints, 8-byte aligned, used instead of pool memory.
constexpr int LU_ABORTED = -2;
size_t lu_work_ints(int m, int n, int nb);
void dgetrf(Context* ctx, int m, int n, double* a, int lda, int nb, int* d_piv, int* d_perm, int* d_info,
            LuPhases* phases = nullptr, int* work = nullptr);
void zgetrf(Context* ctx, int m, int n, cplx* a, int lda, int nb, int* d_piv, int* d_perm, int* d_info,
            LuPhases* phases = nullptr, int* work = nullptr);
// Solves against the factors; trans 1 = A^T (real) / A^H (complex).
void dgetrs(Context* ctx, int trans, int n, int nrhs, const double* lu, int ldlu, const int* d_perm,
            double* b, int ldb);
void zgetrs(Context* ctx, int trans, int n, int nrhs, const cplx* lu, int ldlu, const int* d_perm,
            cplx* b, int ldb);
// A X = B against the factors with the diagonal blocks applied through
// their explicit inverses (for many right-hand sides of a well-conditioned
// A; rounding differs from dgetrs/zgetrs).
void dgetrs_blockinv(Context* ctx, int n, int nrhs, const double* lu, int ldlu, const int* d_perm, double* b, int ldb);
void zgetrs_blockinv(Context* ctx, int n, int nrhs, const cplx* lu, int ldlu, const int* d_perm, cplx* b, int ldb);
// The KernelEntry triangular solves: kind as tns_tri_kind.
void dtrsm(Context* ctx, int kind, int m, int n, const double* a, int lda, double* b, int ldb);
// min/max |u_ii| into d_out[0], d_out[1] (and a NaN flag).
void dpivot_range(Context* ctx, int n, const double* a, int lda, double* d_out);
void zpivot_range(Context* ctx, int n, const cplx* a, int lda, double* d_out);

// Mixed precision (gemm_lp.hip, lu.hip). Single-precision products on the
// matrix cores with FP32 accumulation; mode LP_FP32 takes the operands as
// they are, LP_BF16 rounds them to bfloat16 first, LP_BF16X3 splits each
// into two bfloat16 parts and sums three products. A untransposed.
enum { LP_FP32 = 0, LP_BF16 = 1, LP_BF16X3 = 2 };
void sgemm_lp(Context* ctx, int mode, int m, int n, int k, float alpha, const float* a, int lda, const float* b,
              int ldb, float beta, float* c, int ldc);
// FP32 LU with partial pivoting; its trailing updates use sgemm_lp in
// ctx->lp_mode. Same outputs as dgetrf.
void sgetrf(Context* ctx, int m, int n, float* a, int lda, int nb, int* d_piv, int* d_perm, int* d_info);
// A X = B against FP32 factors (no transpose).
void sgetrs(Context* ctx, int n, int nrhs, const float* lu, int ldlu, const int* d_perm, float* b, int ldb);
// FP32 LU with block-local partial pivoting (each block step's pivots
// from its B x B diagonal block, B = local_lu_block(); for a randomized
// matrix of order n a multiple of B), products on sgemm_lp in
// ctx->lp_mode; P A = L U with perm as sgetrf's, info the first zero pivot
// or -1; inv (n / B * 2 B^2 floats) receives the inverses of the diagonal
// blocks' triangles, which its solve for few right-hand sides uses.
int local_lu_block();
void sgetrf_local(Context* ctx, int n, float* a, int lda, int* d_piv, int* d_perm, float* inv, int* d_info);
void sgetrs_local(Context* ctx, int n, int nrhs, const float* lu, int ldlu, const int* d_perm, const float* inv,
                  float* b, int ldb);

// Level-1/2 helpers (blas1.hip).
void dnorm(Context* ctx, int kind, int m, int n, const double* a, int lda, double* d_out);
void znorm(Context* ctx, int kind, int m, int n, const cplx* a, int lda, double* d_out);
void dpanel(Context* ctx, int trans, int n, int t, const double* a, int lda, const double* x, int ldx,
            double* y, int ldy);
void zpanel(Context* ctx, int adjoint, int n, int t, const cplx* a, int lda, const cplx* x, int ldx,
            cplx* y, int ldy);
void dcopy2d(Context* ctx, int m, int n, const double* src, int lds, double* dst, int ldd);
void zcopy2d(Context* ctx, int m, int n, const cplx* src, int lds, cplx* dst, int ldd);
void dscale_c(Context* ctx, int m, int n, double beta, double* c, int ldc);
void zscale_c(Context* ctx, int m, int n, cplx beta, cplx* c, int ldc);

}  // namespace tns
