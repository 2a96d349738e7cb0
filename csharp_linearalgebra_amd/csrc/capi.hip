// The C ABI (include/tensile_hip.h): argument checks in the form of
// Tensile's seam (Operand.cs:80-96, KernelEntry.cs:246-269), staging of
// host operands through the device, the context and its device pool.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <iterator>
#include <memory>
#include <new>
#include <numeric>
#include <set>

#include "algos.hpp"
#include "batch.hpp"
#include "blas1.hpp"
#include "mixed.hpp"

struct tns_batch;

struct tns_context {
    tns::Context impl;
    // Read without the lock for a fast refusal, then again under it: a call
    // that was waiting on the lock while the context was destroyed must not
    // run its body (Workspace.cs:243-253 checks disposal inside the lock).
    std::atomic<bool> disposed{false};
    // Live batches: their plans hold pool blocks and graphs, released with
    // the context if the caller destroys the context first.
    std::set<tns_batch*> batches;
};

struct tns_batch {
    tns_context* ctx = nullptr;
    std::unique_ptr<tns::BatchPlan> plan;
};

// Device scratch of tns_dgetrf_noalloc: the row payloads and the panels'
// exchange words and move lists, for callers that cannot allocate. One copy
// per device, held until the process ends; every call on a device shares
// it, so calls that overlap on one device race. Host code never uses this
// name as an address: tns_context_create looks up each device's copy with
// hipGetSymbolAddress and keeps that pointer.
constexpr size_t kNoAllocScratchInts = size_t(1) << 18;  // 1 MiB
static __device__ __align__(256) int scratch[kNoAllocScratchInts];

namespace {
std::mutex g_scratch_mutex;
int* g_scratch_ptr[256] = {nullptr};

// This device's copy of `scratch` (the device is current).
int* device_scratch(int device)
{
    if (device < 0 || device >= 256) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        g_scratch_ptr[device] = static_cast<int*>(p);
    }
    return g_scratch_ptr[device];
}
}  // namespace

namespace tns {

[[noreturn]] void fail(int status, const std::string& message) { throw Error{status, message}; }

void check_hip(hipError_t e, const char* what)
{
    if (e == hipSuccess) return;
    (void)hipGetLastError();
    fail(e == hipErrorOutOfMemory ? TNS_ENOMEM : TNS_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

size_t Pool::round(size_t bytes)
{
    const size_t small = size_t(4) << 10, big = size_t(1) << 20;
    if (bytes <= big) return (bytes + small - 1) / small * small;
    return (bytes + big - 1) / big * big;
}

void* Pool::take(size_t bytes)
{
    size_t want = round(bytes);
    auto it = free_.lower_bound(want);
    if (it != free_.end() && it->first <= 2 * want) {
        void* p = it->second;
        size_t have = it->first;
        free_.erase(it);
        // Remember the block's real size by handing the caller's count back
        // through give(): blocks are re-keyed by the size they were made with.
        (void)have;
        return p;
    }
    if (held_ + want > limit_) trim();
    if (held_ + want > limit_)
        fail(TNS_ENOMEM, "device working memory would exceed the context's limit (" + std::to_string(limit_) +
                             " bytes)");
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(TNS_ENOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
        }
    }
    held_ += want;
    sizes_[p] = want;
    return p;
}

void Pool::give(void* p, size_t)
{
    auto it = sizes_.find(p);
    if (it == sizes_.end()) return;
    free_.emplace(it->second, p);
}

void Pool::trim()
{
    if (free_.empty()) return;
    (void)hipDeviceSynchronize();
    for (auto& kv : free_) {
        (void)hipFree(kv.second);
        held_ -= kv.first;
        sizes_.erase(kv.second);
    }
    free_.clear();
}

// The retention cap of ComplexScratch (ComplexScratch.cs:12-60, 77-122):
// what is kept between calls is bounded, whatever a call needed. Free blocks
// go back largest first until the pool holds at most the cap (blocks still
// lent are not touched; they come back with the next call's end).
void Pool::shrink_to_retention()
{
    if (held_ <= retain_ || free_.empty()) return;
    (void)hipDeviceSynchronize();
    while (held_ > retain_ && !free_.empty()) {
        auto it = std::prev(free_.end());
        (void)hipFree(it->second);
        held_ -= it->first;
        sizes_.erase(it->second);
        free_.erase(it);
    }
}

// Frees every block, lent or not; a handle that later gives one back finds
// it unknown and drops it.
void Pool::release_all()
{
    (void)hipDeviceSynchronize();
    for (auto& kv : sizes_) (void)hipFree(kv.first);
    sizes_.clear();
    free_.clear();
    held_ = 0;
}

Pool::~Pool() { release_all(); }

void HostPool::run(int count, const std::function<void(int)>& fn)
{
    if (count <= 0) return;
    if (count <= 4) {
        // A few steps: waking the workers costs more than the steps.
        for (int i = 0; i < count; i++) fn(i);
        return;
    }
    if (threads_.empty()) {
        const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
        const int workers = std::min(hw, 8) - 1;  // the lockstep estimates are at most 8
        for (int w = 0; w < workers; w++) threads_.emplace_back([this] { work(); });
    }
    {
        std::lock_guard<std::mutex> lock(m_);
        job_ = &fn;
        count_ = count;
        next_.store(0);
        error_ = nullptr;
        busy_ = (int)threads_.size();
        ++generation_;
    }
    wake_.notify_all();
    auto take = [&]() {
        try {
            for (int i; (i = next_.fetch_add(1)) < count;) fn(i);
        } catch (...) {
            std::lock_guard<std::mutex> lock(m_);
            if (!error_) error_ = std::current_exception();
            next_.store(count);
        }
    };
    take();
    std::unique_lock<std::mutex> lock(m_);
    done_.wait(lock, [&] { return busy_ == 0; });
    job_ = nullptr;
    if (error_) std::rethrow_exception(error_);
}

void HostPool::work()
{
    unsigned long long seen = 0;
    for (;;) {
        const std::function<void(int)>* job;
        int count;
        {
            std::unique_lock<std::mutex> lock(m_);
            wake_.wait(lock, [&] { return stopping_ || generation_ != seen; });
            if (stopping_) return;
            seen = generation_;
            job = job_;
            count = count_;
        }
        try {
            for (int i; (i = next_.fetch_add(1)) < count;) (*job)(i);
        } catch (...) {
            std::lock_guard<std::mutex> lock(m_);
            if (!error_) error_ = std::current_exception();
            next_.store(count);
        }
        {
            std::lock_guard<std::mutex> lock(m_);
            if (--busy_ == 0) done_.notify_all();
        }
    }
}

void HostPool::stop()
{
    {
        std::lock_guard<std::mutex> lock(m_);
        stopping_ = true;
    }
    wake_.notify_all();
    for (auto& t : threads_) t.join();
    threads_.clear();
    stopping_ = false;
}

void read_back(Context* ctx, void* h, const void* d, size_t bytes)
{
    if (bytes > kMailBytes) fail(TNS_EINVAL, "read_back: more than the mailbox holds");
    TNS_HIP(hipMemcpyAsync(ctx->h_mail, d, bytes, hipMemcpyDeviceToHost, ctx->stream));
    TNS_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(h, ctx->h_mail, bytes);
}

void* pinned_panel(Context* ctx, size_t bytes)
{
    if (ctx->pin_pending) {
        TNS_HIP(hipEventSynchronize(ctx->ev_pin));
        ctx->pin_pending = false;
    }
    if (ctx->h_pin_bytes < bytes) {
        // Growing frees the old area: its last copies must be done.
        TNS_HIP(hipStreamSynchronize(ctx->stream));
        if (ctx->h_pin) TNS_HIP(hipHostFree(ctx->h_pin));
        ctx->h_pin = nullptr;
        ctx->h_pin_bytes = 0;
        void* h = nullptr;
        const size_t want = bytes < (size_t(1) << 20) ? (size_t(1) << 20) : bytes;
        TNS_HIP(hipHostMalloc(&h, want, hipHostMallocDefault));
        ctx->h_pin = static_cast<char*>(h);
        ctx->h_pin_bytes = want;
    }
    return ctx->h_pin;
}

void write_pinned(Context* ctx, void* d, const void* h, size_t bytes)
{
    void* p = pinned_panel(ctx, bytes);
    std::memcpy(p, h, bytes);
    TNS_HIP(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, ctx->stream));
    TNS_HIP(hipEventRecord(ctx->ev_pin, ctx->stream));
    ctx->pin_pending = true;
}


bool is_device_pointer(const void* p)
{
    if (!p) return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged || attr.type == hipMemoryTypeUnified;
}

void sync(Context* ctx) { TNS_HIP(hipStreamSynchronize(ctx->stream)); }

void EventClock::mark(int b)
{
    hipEvent_t e = nullptr;
    TNS_HIP(hipEventCreate(&e));
    ev.push_back(e);
    bucket.push_back(b);
    TNS_HIP(hipEventRecord(e, ctx->stream));
}

void EventClock::read(double* out, int nb, double* total)
{
    if (ev.empty()) return;
    TNS_HIP(hipEventSynchronize(ev.back()));
    for (size_t i = 1; i < ev.size(); i++) {
        float ms = 0.f;
        TNS_HIP(hipEventElapsedTime(&ms, ev[i - 1], ev[i]));
        if (bucket[i] >= 0 && bucket[i] < nb) out[bucket[i]] += ms;
    }
    float all = 0.f;
    TNS_HIP(hipEventElapsedTime(&all, ev.front(), ev.back()));
    *total += all;
}

}  // namespace tns

namespace {

using tns::Context;
using tns::DevBuf;
using tns::fail;
using tns::zd;
using tns::make_zd;

// A caller's matrix as the device sees it: the caller's own pointer when it
// is device memory, else a packed device copy (ld = rows).
template <typename T>
struct View {
    T* d = nullptr;
    int ld = 1;
    bool staged = false;
    DevBuf<T> buf;
};

constexpr size_t kStageMax = size_t(64) << 20;

// The staging buffer, grown to hold min(bytes, kStageMax).
char* stage(Context* ctx, size_t bytes)
{
    const size_t want = bytes < kStageMax ? bytes : kStageMax;
    if (ctx->h_stage_bytes < want) {
        if (ctx->h_stage) TNS_HIP(hipHostFree(ctx->h_stage));
        ctx->h_stage = nullptr;
        ctx->h_stage_bytes = 0;
        void* h = nullptr;
        TNS_HIP(hipHostMalloc(&h, want, hipHostMallocDefault));
        ctx->h_stage = static_cast<char*>(h);
        ctx->h_stage_bytes = want;
    }
    return ctx->h_stage;
}

// Host (ld) <-> device (packed, ld = rows) through the staging buffer, in
// chunks of whole columns, or, for a column longer than the buffer, in
// pieces of rows. Each chunk waits for its copy before the buffer is reused,
// so the host pointer may be freed as soon as the call returns.
template <typename T>
void host_to_device(Context* ctx, T* d, const T* h, int rows, int cols, int ld)
{
    const size_t col_bytes = sizeof(T) * (size_t)rows;
    char* buf = stage(ctx, col_bytes * cols);
    const size_t cap = ctx->h_stage_bytes;
    if (col_bytes > cap) {
        const size_t piece = cap / sizeof(T);
        for (int c = 0; c < cols; c++)
            for (size_t r0 = 0; r0 < (size_t)rows; r0 += piece) {
                const size_t nr = std::min(piece, (size_t)rows - r0);
                std::memcpy(buf, h + (size_t)c * ld + r0, sizeof(T) * nr);
                TNS_HIP(hipMemcpyAsync(d + (size_t)c * rows + r0, buf, sizeof(T) * nr, hipMemcpyHostToDevice,
                                       ctx->stream));
                TNS_HIP(hipStreamSynchronize(ctx->stream));
            }
        return;
    }
    const int per = (int)(cap / col_bytes);
    for (int c0 = 0; c0 < cols; c0 += per) {
        const int nc = std::min(per, cols - c0);
        for (int c = 0; c < nc; c++) std::memcpy(buf + col_bytes * c, h + (size_t)(c0 + c) * ld, col_bytes);
        TNS_HIP(hipMemcpyAsync(d + (size_t)c0 * rows, buf, col_bytes * nc, hipMemcpyHostToDevice, ctx->stream));
        TNS_HIP(hipStreamSynchronize(ctx->stream));
    }
}

template <typename T>
void device_to_host(Context* ctx, T* h, const T* d, int rows, int cols, int ld)
{
    const size_t col_bytes = sizeof(T) * (size_t)rows;
    char* buf = stage(ctx, col_bytes * cols);
    const size_t cap = ctx->h_stage_bytes;
    if (col_bytes > cap) {
        const size_t piece = cap / sizeof(T);
        for (int c = 0; c < cols; c++)
            for (size_t r0 = 0; r0 < (size_t)rows; r0 += piece) {
                const size_t nr = std::min(piece, (size_t)rows - r0);
                TNS_HIP(hipMemcpyAsync(buf, d + (size_t)c * rows + r0, sizeof(T) * nr, hipMemcpyDeviceToHost,
                                       ctx->stream));
                TNS_HIP(hipStreamSynchronize(ctx->stream));
                std::memcpy(h + (size_t)c * ld + r0, buf, sizeof(T) * nr);
            }
        return;
    }
    const int per = (int)(cap / col_bytes);
    for (int c0 = 0; c0 < cols; c0 += per) {
        const int nc = std::min(per, cols - c0);
        TNS_HIP(hipMemcpyAsync(buf, d + (size_t)c0 * rows, col_bytes * nc, hipMemcpyDeviceToHost, ctx->stream));
        TNS_HIP(hipStreamSynchronize(ctx->stream));
        for (int c = 0; c < nc; c++) std::memcpy(h + (size_t)(c0 + c) * ld, buf + col_bytes * c, col_bytes);
    }
}

template <typename T>
View<T> bring(Context* ctx, const T* p, int rows, int cols, int ld, bool copy_in)
{
    View<T> v;
    if (rows <= 0 || cols <= 0) return v;
    if (tns::is_device_pointer(p)) {
        v.d = const_cast<T*>(p);
        v.ld = ld;
        return v;
    }
    v.buf = DevBuf<T>(ctx, (size_t)rows * cols);
    v.d = v.buf.get();
    v.ld = rows;
    v.staged = true;
    if (copy_in) host_to_device(ctx, v.d, p, rows, cols, ld);
    return v;
}

// Only the rows x cols part of the caller's matrix is written (padding
// rows between columns are left alone).
template <typename T>
void send_back(Context* ctx, const View<T>& v, T* p, int rows, int cols, int ld)
{
    if (!v.staged || rows <= 0 || cols <= 0) return;
    device_to_host(ctx, p, v.d, rows, cols, ld);
}

// Operand.cs:80-96: non-negative shape, stride >= rows, extent in 64 bits.
void check_shape(int rows, int cols, int ld, const char* what)
{
    if (rows < 0 || cols < 0) fail(TNS_EINVAL, std::string(what) + ": negative dimension");
    if (ld < (rows > 1 ? rows : 1)) fail(TNS_EINVAL, std::string(what) + ": leading dimension below the row count");
}

void check_operand(const void* p, int rows, int cols, int ld, const char* what)
{
    check_shape(rows, cols, ld, what);
    if ((long long)rows * cols > 0 && !p) fail(TNS_EINVAL, std::string(what) + ": null data for a non-empty operand");
}

void require(bool cond, const char* what)
{
    if (!cond) fail(TNS_EINVAL, std::string("Kernel operands disagree on ") + what + ".");
}

template <typename F>
int guarded(tns_context* ctx, F&& body)
{
    // After a call: the pool keeps at most its retention in free blocks; the
    // solve's exchange area counts as kept memory and goes back first.
    auto settle = [](Context* c) {
        if (c->flow && c->pool.held() > c->pool.retention()) {
            c->pool.give(c->flow, c->flow_bytes);
            c->flow = nullptr;
            c->flow_bytes = 0;
        }
        c->pool.shrink_to_retention();
    };
    if (!ctx) return TNS_EINVAL;
    if (ctx->disposed.load()) return TNS_EDISPOSED;
    std::lock_guard<std::mutex> lock(ctx->impl.gate);
    if (ctx->disposed.load()) return TNS_EDISPOSED;
    try {
        TNS_HIP(hipSetDevice(ctx->impl.device));
        body(&ctx->impl);
        settle(&ctx->impl);
        ctx->impl.last_error.clear();
        return TNS_OK;
    } catch (const tns::Error& e) {
        ctx->impl.last_error = e.message;
        (void)hipStreamSynchronize(ctx->impl.stream);
        settle(&ctx->impl);
        return e.status;
    } catch (const std::bad_alloc&) {
        ctx->impl.last_error = "host allocation failed";
        return TNS_ENOMEM;
    } catch (const std::exception& e) {
        ctx->impl.last_error = e.what();
        return TNS_EDEVICE;
    }
}

// P from LAPACK-order interchanges (0-based): row i of P*A is row perm[i] of
// A. Checks each interchange names a row at or below its step.
std::vector<int> permutation_of(Context* ctx, const int* pivots, int count, int rows)
{
    std::vector<int> piv(count), perm(rows);
    if (count > 0) {
        if (tns::is_device_pointer(pivots)) {
            TNS_HIP(hipMemcpyAsync(piv.data(), pivots, sizeof(int) * count, hipMemcpyDeviceToHost, ctx->stream));
            tns::sync(ctx);
        } else {
            if (!pivots) fail(TNS_EINVAL, "pivots: null");
            std::memcpy(piv.data(), pivots, sizeof(int) * count);
        }
    }
    std::iota(perm.begin(), perm.end(), 0);
    for (int k = 0; k < count; k++) {
        if (piv[k] < k || piv[k] >= rows) fail(TNS_EINVAL, "pivots: interchange out of range");
        std::swap(perm[k], perm[piv[k]]);
    }
    return perm;
}

DevBuf<int> upload_ints(Context* ctx, const std::vector<int>& h)
{
    DevBuf<int> d(ctx, h.size());
    if (!h.empty())
        TNS_HIP(hipMemcpyAsync(d.get(), h.data(), sizeof(int) * h.size(), hipMemcpyHostToDevice, ctx->stream));
    return d;
}

inline zd c2(tns_complex z) { return make_zd(z.re, z.im); }

// ---- generic bodies -----------------------------------------------------

template <typename T>
void getrf_body(Context* ctx, int m, int n, T* a, int lda, int nb, int* pivots, tns_lu_info* info,
                tns_lu_phases* phases = nullptr)
{
    check_operand(a, m, n, lda, "A");
    const int limit = std::min(m, n);
    if (limit > 0 && !pivots) fail(TNS_EINVAL, "pivots: null");
    View<T> va = bring(ctx, a, m, n, lda, true);
    DevBuf<int> piv(ctx, (size_t)std::max(limit, 1)), perm(ctx, (size_t)std::max(m, 1));
    tns::LuPhases split, *ps = phases ? &split : nullptr;
    if constexpr (sizeof(T) == 8)
        tns::dgetrf(ctx, m, n, (double*)va.d, va.ld, nb, piv.get(), perm.get(), ctx->d_int, ps);
    else
        tns::zgetrf(ctx, m, n, (zd*)va.d, va.ld, nb, piv.get(), perm.get(), ctx->d_int, ps);
    if (limit > 0) {
        if constexpr (sizeof(T) == 8) tns::dpivot_range(ctx, limit, (const double*)va.d, va.ld, ctx->d_scal);
        else tns::zpivot_range(ctx, limit, (const zd*)va.d, va.ld, ctx->d_scal);
    }
    int sing = -1;
    double range[2] = {0.0, 0.0};
    // Through the page-locked mailbox: the pivot range at 0, the info at 16.
    if (limit > 0)
        TNS_HIP(hipMemcpyAsync(ctx->h_mail, ctx->d_scal, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TNS_HIP(hipMemcpyAsync(ctx->h_mail + 16, ctx->d_int, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    tns::sync(ctx);
    std::memcpy(&sing, ctx->h_mail + 16, sizeof(int));
    if (limit > 0) std::memcpy(range, ctx->h_mail, 2 * sizeof(double));
    // Checked before any output leaves the device.
    if (sing == tns::LU_ABORTED) fail(TNS_EDEVICE, "an LU panel's cross-workgroup exchange timed out");
    if (limit > 0) {
        if (tns::is_device_pointer(pivots))
            TNS_HIP(hipMemcpyAsync(pivots, piv.get(), sizeof(int) * limit, hipMemcpyDeviceToDevice, ctx->stream));
        else
            TNS_HIP(hipMemcpyAsync(pivots, piv.get(), sizeof(int) * limit, hipMemcpyDeviceToHost, ctx->stream));
    }
    send_back(ctx, va, a, m, n, lda);
    tns::sync(ctx);
    if (phases) {
        phases->panel_ms += split.panel_ms;
        phases->swaps_ms += split.swaps_ms;
        phases->trsm_ms += split.trsm_ms;
        phases->gemm_ms += split.gemm_ms;
        phases->total_ms += split.total_ms;
        phases->block_steps += split.steps;
    }
    if (info) {
        info->singular_column = sing;
        info->smallest_pivot = range[0];
        info->largest_pivot = range[1];
    }
}

template <typename T>
void getrs_body(Context* ctx, int trans, int n, int nrhs, const T* lu, int ldlu, const int* pivots, int singular,
                T* b, int ldb)
{
    check_operand(lu, n, n, ldlu, "factors");
    check_operand(b, n, nrhs, ldb, "B");
    if (trans != 0 && trans != 1) fail(TNS_EINVAL, "trans must be 0 or 1");
    if (singular >= 0)
        fail(TNS_ESTATE, "Matrix is singular: zero pivot at column " + std::to_string(singular) + ".");
    if (n == 0 || nrhs == 0) return;
    std::vector<int> hperm = permutation_of(ctx, pivots, n, n);
    DevBuf<int> perm = upload_ints(ctx, hperm);
    View<T> vlu = bring(ctx, lu, n, n, ldlu, true);
    View<T> vb = bring(ctx, b, n, nrhs, ldb, true);
    if constexpr (sizeof(T) == 8)
        tns::dgetrs(ctx, trans, n, nrhs, (const double*)vlu.d, vlu.ld, perm.get(), (double*)vb.d, vb.ld);
    else
        tns::zgetrs(ctx, trans, n, nrhs, (const zd*)vlu.d, vlu.ld, perm.get(), (zd*)vb.d, vb.ld);
    send_back(ctx, vb, b, n, nrhs, ldb);
    tns::sync(ctx);
}

template <typename T>
void gecon_body(Context* ctx, int n, const T* lu, int ldlu, const int* pivots, int singular, double anorm,
                int columns, double* rcond)
{
    check_operand(lu, n, n, ldlu, "factors");
    if (!rcond) fail(TNS_EINVAL, "rcond: null");
    if (singular >= 0 || anorm == 0.0 || n == 0) {
        *rcond = 0.0;
        return;
    }
    std::vector<int> hperm = permutation_of(ctx, pivots, n, n);
    DevBuf<int> perm = upload_ints(ctx, hperm);
    View<T> vlu = bring(ctx, lu, n, n, ldlu, true);
    if (vlu.ld != n) {
        // The solves run on packed factors.
        View<T> packed;
        packed.buf = DevBuf<T>(ctx, (size_t)n * n);
        packed.d = packed.buf.get();
        packed.ld = n;
        if constexpr (sizeof(T) == 8) tns::dcopy2d(ctx, n, n, (const double*)vlu.d, vlu.ld, (double*)packed.d, n);
        else tns::zcopy2d(ctx, n, n, (const zd*)vlu.d, vlu.ld, (zd*)packed.d, n);
        *rcond = tns::gecon<T>(ctx, n, packed.d, n, perm.get(), singular, anorm, columns);
    } else {
        *rcond = tns::gecon<T>(ctx, n, vlu.d, vlu.ld, perm.get(), singular, anorm, columns);
    }
}

void mixed_body(Context* ctx, int n, int nrhs, const double* a, int lda, const double* b, int ldb, double* x, int ldx,
                int precision, double tol, int max_iter, tns_mixed_info* info)
{
    if (n < 0 || nrhs < 0) fail(TNS_EINVAL, "negative dimension");
    if (precision != 32 && precision != 16) fail(TNS_EINVAL, "precision must be 32 or 16");
    if (!(tol > 0.0) || max_iter < 1) fail(TNS_ERANGE, "tolerance must be positive and max_iter at least 1");
    check_operand(a, n, n, lda, "A");
    check_operand(b, n, nrhs, ldb, "B");
    check_operand(x, n, nrhs, ldx, "X");
    tns::MixedInfo mi;
    mi.precision = precision;
    if (n > 0 && nrhs > 0) {
        View<double> va = bring(ctx, a, n, n, lda, true);
        View<double> vb = bring(ctx, b, n, nrhs, ldb, true);
        View<double> vx = bring(ctx, x, n, nrhs, ldx, false);
        tns::dgesv_mixed(ctx, n, nrhs, va.d, va.ld, vb.d, vb.ld, vx.d, vx.ld, precision, tol, max_iter, &mi);
        send_back(ctx, vx, x, n, nrhs, ldx);
        tns::sync(ctx);
    }
    if (info)
        *info = tns_mixed_info{mi.precision, mi.iterations, mi.fell_back, mi.backward_error, mi.factor_ms,
                               mi.refine_ms, mi.fallback_ms, mi.first_error};
}

template <typename T>
void expmv_grad_body(Context* ctx, int n, int nb, const T* a, int lda, const T* b, int ldb, double t, const T* g,
                     int ldg, T* ga, int ldga, T* gb, int ldgb, tns_expmv_diag* diag)
{
    if (n < 0 || nb < 0) fail(TNS_EINVAL, "negative dimension");
    if (!std::isfinite(t)) fail(TNS_ERANGE, "t must be finite");
    check_operand(a, n, n, lda, "A");
    check_operand(b, n, nb, ldb, "B");
    check_operand(g, n, nb, ldg, "G");
    check_operand(ga, n, n, ldga, "dA");
    check_operand(gb, n, nb, ldgb, "dB");
    tns::ExpmvDiag d;
    if (n > 0) {
        View<T> va = bring(ctx, a, n, n, lda, true);
        View<T> vb = bring(ctx, b, n, nb, ldb, true);
        View<T> vg = bring(ctx, g, n, nb, ldg, true);
        View<T> vga = bring(ctx, ga, n, n, ldga, false);
        View<T> vgb = bring(ctx, gb, n, nb, ldgb, false);
        tns::expmv_grad<T>(ctx, n, nb, va.d, va.ld, vb.d, vb.ld, t, vg.d, vg.ld, vga.d, vga.ld, vgb.d, vgb.ld, &d);
        send_back(ctx, vga, ga, n, n, ldga);
        send_back(ctx, vgb, gb, n, nb, ldgb);
        tns::sync(ctx);
    }
    if (diag) *diag = tns_expmv_diag{d.degree, d.scaling, d.applications, d.early_exits};
}

template <typename T>
void expm_body(Context* ctx, int n, const T* a, int lda, T* e, int lde, tns_expm_diag* diag,
               tns_expm_phases* phases = nullptr)
{
    if (n < 0) fail(TNS_EINVAL, "negative order");
    check_operand(a, n, n, lda, "A");
    check_operand(e, n, n, lde, "E");
    if (n == 0) {
        if (diag) *diag = tns_expm_diag{0, 0};
        return;
    }
    View<T> va = bring(ctx, a, n, n, lda, true);
    View<T> ve = bring(ctx, e, n, n, lde, false);
    tns::ExpmDiag d;
    tns::ExpmPhases ph;
    tns::expm<T>(ctx, n, va.d, va.ld, ve.d, ve.ld, &d, phases ? &ph : nullptr);
    send_back(ctx, ve, e, n, n, lde);
    tns::sync(ctx);
    if (diag) *diag = tns_expm_diag{d.degree, d.squarings};
    if (phases)
        *phases = tns_expm_phases{ph.ms[0], ph.ms[1], ph.ms[2], ph.ms[3], ph.ms[4], ph.ms[5], ph.total_ms};
}

template <typename T>
void expmv_body(Context* ctx, int n, int nb, const T* a, int lda, const T* b, int ldb, double t, T* f, int ldf,
                tns_expmv_diag* diag)
{
    check_operand(a, n, n, lda, "A");
    check_operand(b, n, nb, ldb, "B");
    check_operand(f, n, nb, ldf, "F");
    View<T> vb = bring(ctx, b, n, nb, ldb, true);
    View<T> vf = bring(ctx, f, n, nb, ldf, false);
    tns::ExpmvDiag d;
    if (n > 0 && nb > 0) {
        View<T> va = bring(ctx, a, n, n, lda, true);
        tns::expmv<T>(ctx, n, nb, va.d, va.ld, vb.d, vb.ld, t, vf.d, vf.ld, &d);
    } else if (n > 0) {
        // No columns: F is B (MatrixExponentialAction.cs:116).
    }
    send_back(ctx, vf, f, n, nb, ldf);
    tns::sync(ctx);
    if (diag) *diag = tns_expmv_diag{d.degree, d.scaling, d.applications, d.early_exits};
}

// The caller's operator, host or device side, as a device operator.
template <typename T, typename Fn>
tns::DevOp<T> callback_operator(Context* ctx, int n, Fn apply, void* user, int device_operator)
{
    tns::DevOp<T> op;
    op.n = n;
    using Pub = typename std::conditional<sizeof(T) == 8, double, tns_complex>::type;
    op.apply = [=](int t, const T* x, T* y) {
        int rc;
        if (device_operator) {
            tns::sync(ctx);
            rc = apply(user, n, t, (const Pub*)x, n, (Pub*)y, n);
        } else {
            std::vector<T> hx((size_t)n * t), hy((size_t)n * t);
            TNS_HIP(hipMemcpyAsync(hx.data(), x, sizeof(T) * hx.size(), hipMemcpyDeviceToHost, ctx->stream));
            tns::sync(ctx);
            rc = apply(user, n, t, (const Pub*)hx.data(), n, (Pub*)hy.data(), n);
            if (rc == 0)
                TNS_HIP(hipMemcpyAsync(y, hy.data(), sizeof(T) * hy.size(), hipMemcpyHostToDevice, ctx->stream));
            tns::sync(ctx);
        }
        if (rc != 0) fail(TNS_ECALLBACK, "the operator callback reported failure (" + std::to_string(rc) + ")");
    };
    op.adjoint = [](int, const T*, T*) { fail(TNS_EINVAL, "the matrix-free path never applies an adjoint"); };
    return op;
}

template <typename T, typename Fn>
void expmv_op_body(Context* ctx, int n, Fn apply, void* user, int device_operator, int nb, const T* b, int ldb,
                   double t, double bound, T* f, int ldf, tns_expmv_diag* diag)
{
    if (n < 0) fail(TNS_EINVAL, "negative order");
    if (!apply) fail(TNS_EINVAL, "operator: null");
    check_operand(b, n, nb, ldb, "B");
    check_operand(f, n, nb, ldf, "F");
    if (!std::isfinite(bound) || bound < 0.0) fail(TNS_ERANGE, "The norm bound must be finite and not negative.");
    View<T> vb = bring(ctx, b, n, nb, ldb, true);
    View<T> vf = bring(ctx, f, n, nb, ldf, false);
    tns::ExpmvDiag d;
    if (n > 0 && nb > 0) {
        tns::DevOp<T> op = callback_operator<T>(ctx, n, apply, user, device_operator);
        tns::expmv_op<T>(ctx, op, nb, vb.d, vb.ld, t, bound, vf.d, vf.ld, &d);
    }
    send_back(ctx, vf, f, n, nb, ldf);
    tns::sync(ctx);
    if (diag) *diag = tns_expmv_diag{d.degree, d.scaling, d.applications, d.early_exits};
}

template <typename T>
void normest_body(Context* ctx, int n, const T* a, int lda, int power, int columns, int itmax, int seed,
                  tns_normest* out)
{
    check_operand(a, n, n, lda, "A");
    if (!out) fail(TNS_EINVAL, "out: null");
    if (power < 1) fail(TNS_ERANGE, "power must be at least 1");
    if (n == 0) {
        *out = tns_normest{0.0, 0, 0};
        return;
    }
    View<T> va = bring(ctx, a, n, n, lda, true);
    tns::NormEst r = tns::normest<T>(ctx, tns::power_operator<T>(ctx, n, va.d, va.ld, power), columns, itmax, seed);
    *out = tns_normest{r.value, r.iterations, r.products};
}

template <typename T>
void panel_body(Context* ctx, int trans, int n, int t, const T* a, int lda, const T* x, int ldx, T* y, int ldy)
{
    check_operand(a, n, n, lda, "A");
    check_operand(x, n, t, ldx, "X");
    check_operand(y, n, t, ldy, "Y");
    if (n == 0 || t == 0) return;
    View<T> va = bring(ctx, a, n, n, lda, true);
    View<T> vx = bring(ctx, x, n, t, ldx, true);
    View<T> vy = bring(ctx, y, n, t, ldy, false);
    if constexpr (sizeof(T) == 8)
        tns::dpanel(ctx, trans, n, t, (const double*)va.d, va.ld, (const double*)vx.d, vx.ld, (double*)vy.d, vy.ld);
    else
        tns::zpanel(ctx, trans, n, t, (const zd*)va.d, va.ld, (const zd*)vx.d, vx.ld, (zd*)vy.d,
                    vy.ld);
    send_back(ctx, vy, y, n, t, ldy);
    if (va.staged || vx.staged || vy.staged) tns::sync(ctx);
}

template <typename T>
void norm_body(Context* ctx, char kind, int m, int n, const T* a, int lda, double* out)
{
    check_operand(a, m, n, lda, "A");
    if (!out) fail(TNS_EINVAL, "out: null");
    const bool ok = kind == '1' || kind == 'I' || (kind == 'F' && sizeof(T) == 8);
    if (!ok) fail(TNS_EINVAL, "unknown norm kind");
    if (m == 0 || n == 0) {
        *out = 0.0;
        return;
    }
    View<T> va = bring(ctx, a, m, n, lda, true);
    if constexpr (sizeof(T) == 8) tns::dnorm(ctx, kind, m, n, (const double*)va.d, va.ld, ctx->d_scal);
    else tns::znorm(ctx, kind, m, n, (const zd*)va.d, va.ld, ctx->d_scal);
    *out = tns::read_double(ctx, ctx->d_scal);
}

}  // namespace

// ===========================================================================
// Exported entry points.
// ===========================================================================
extern "C" {

TNS_API const char* tns_version(void) { return "tensile-hip 0.1 gfx950"; }

TNS_API int tns_gemm_set_config(int real_config, int complex_config)
{
    if (real_config < -1 || real_config >= tns::gemm_config_count(false)) return TNS_EINVAL;
    if (complex_config < -1 || complex_config >= tns::gemm_config_count(true)) return TNS_EINVAL;
    tns::gemm_force_config(real_config, complex_config);
    return TNS_OK;
}

TNS_API int tns_gemm_config_count(int complex_) { return tns::gemm_config_count(complex_ != 0); }

TNS_API int tns_context_create(int device, tns_context** out)
{
    if (!out) return TNS_EINVAL;
    *out = nullptr;
    tns_context* ctx = new (std::nothrow) tns_context();
    if (!ctx) return TNS_ENOMEM;
    tns::Context& c = ctx->impl;
    c.device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking);
    // The two streams' ordering events: only this device's streams wait on
    // them, so their record needs no system-scope fence (the cache write-back
    // it costs sat on the LU's critical path, once per panel).
    const unsigned order_flags = hipEventDisableTiming | hipEventDisableSystemFence;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c.ev_panel, order_flags);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c.ev_rest[0], order_flags);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c.ev_rest[1], order_flags);
    if (e == hipSuccess) e = hipMalloc(&c.d_scal, 64 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&c.d_int, 64 * sizeof(int));
    if (e == hipSuccess) e = hipHostMalloc((void**)&c.h_mail, tns::kMailBytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c.ev_pin, hipEventDisableTiming);
    if (e == hipSuccess) {
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
            c.max_lds = lds;
        else
            (void)hipGetLastError();
    }
    if (e == hipSuccess) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0) c.pool.set_limit(total_b / 4 * 3);
        else (void)hipGetLastError();
    }
    if (e == hipSuccess) c.noalloc_scratch = device_scratch(device);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (c.d_scal) (void)hipFree(c.d_scal);
        if (c.d_int) (void)hipFree(c.d_int);
        if (c.h_mail) (void)hipHostFree(c.h_mail);
        if (c.stream) (void)hipStreamDestroy(c.stream);
        if (c.side) (void)hipStreamDestroy(c.side);
        for (hipEvent_t ev : {c.ev_panel, c.ev_rest[0], c.ev_rest[1], c.ev_pin})
            if (ev) (void)hipEventDestroy(ev);
        delete ctx;
        return e == hipErrorOutOfMemory ? TNS_ENOMEM : TNS_EDEVICE;
    }
    *out = ctx;
    return TNS_OK;
}

// Releases the device resources. The handle itself stays valid so a late
// call on it is refused with TNS_EDISPOSED (the Workspace's
// ObjectDisposedException) instead of touching freed memory.
TNS_API int tns_context_destroy(tns_context* ctx)
{
    if (!ctx) return TNS_EINVAL;
    if (ctx->disposed.load()) return TNS_OK;
    std::lock_guard<std::mutex> lock(ctx->impl.gate);
    if (ctx->disposed.load()) return TNS_OK;
    tns::Context& c = ctx->impl;
    (void)hipSetDevice(c.device);
    (void)hipStreamSynchronize(c.stream);
    (void)hipStreamSynchronize(c.side);
    // Batches still alive give their blocks and graphs back now; their
    // handles stay valid and refuse further calls with TNS_EDISPOSED.
    for (tns_batch* b : ctx->batches) b->plan.reset();
    ctx->batches.clear();
    c.host.stop();
    c.pool.release_all();  // the solve's exchange area among the blocks
    c.flow = nullptr;
    c.flow_bytes = 0;
    if (c.h_stage) (void)hipHostFree(c.h_stage);
    c.h_stage = nullptr;
    c.h_stage_bytes = 0;
    if (c.h_pin) (void)hipHostFree(c.h_pin);
    c.h_pin = nullptr;
    c.h_pin_bytes = 0;
    if (c.h_mail) (void)hipHostFree(c.h_mail);
    c.h_mail = nullptr;
    (void)hipFree(c.d_scal);
    (void)hipFree(c.d_int);
    (void)hipStreamDestroy(c.stream);
    (void)hipStreamDestroy(c.side);
    for (hipEvent_t ev : {c.ev_panel, c.ev_rest[0], c.ev_rest[1], c.ev_pin}) (void)hipEventDestroy(ev);
    c.ev_pin = nullptr;
    c.side = nullptr;
    c.ev_panel = c.ev_rest[0] = c.ev_rest[1] = nullptr;
    c.d_scal = nullptr;
    c.d_int = nullptr;
    c.stream = nullptr;
    ctx->disposed = true;
    return TNS_OK;
}

TNS_API const char* tns_last_error(const tns_context* ctx)
{
    if (!ctx) return "null context";
    if (ctx->disposed.load()) return "the context has been destroyed";
    return ctx->impl.last_error.c_str();
}

TNS_API int tns_context_set_limit(tns_context* ctx, size_t bytes)
{
    return guarded(ctx, [&](Context* c) {
        c->pool.set_limit(bytes);
        c->pool.trim();
    });
}

TNS_API int tns_context_set_retention(tns_context* ctx, size_t bytes)
{
    return guarded(ctx, [&](Context* c) { c->pool.set_retention(bytes); });
}

TNS_API int tns_context_memory(tns_context* ctx, size_t* held, size_t* limit, size_t* retention)
{
    return guarded(ctx, [&](Context* c) {
        if (held) *held = c->pool.held();
        if (limit) *limit = c->pool.limit();
        if (retention) *retention = c->pool.retention();
    });
}

TNS_API int tns_context_stream(tns_context* ctx, void** stream)
{
    return guarded(ctx, [&](Context* c) {
        if (!stream) fail(TNS_EINVAL, "stream: null");
        *stream = (void*)c->stream;
    });
}

TNS_API int tns_synchronize(tns_context* ctx)
{
    return guarded(ctx, [&](Context* c) { tns::sync(c); });
}

TNS_API int tns_dgemm(tns_context* ctx, int m, int n, int k, double alpha, const double* a, int lda, const double* b,
                      int ldb, double beta, double* c, int ldc)
{
    return guarded(ctx, [&](Context* cx) {
        check_shape(m, k, lda, "A");
        check_shape(k, n, ldb, "B");
        check_shape(m, n, ldc, "C");
        // An empty destination returns before anything is read
        // (KernelEntry.cs:41-44), however many columns it nominally has.
        if (m == 0 || n == 0) return;
        check_operand(a, m, k, lda, "A");
        check_operand(b, k, n, ldb, "B");
        check_operand(c, m, n, ldc, "C");
        View<double> va = bring(cx, a, m, k, lda, true);
        View<double> vb = bring(cx, b, k, n, ldb, true);
        View<double> vc = bring(cx, c, m, n, ldc, beta != 0.0);
        tns::dgemm(cx, m, n, k, alpha, va.d, va.ld, vb.d, vb.ld, beta, vc.d, vc.ld);
        send_back(cx, vc, c, m, n, ldc);
        if (va.staged || vb.staged || vc.staged) tns::sync(cx);
    });
}

TNS_API int tns_zgemm(tns_context* ctx, int m, int n, int k, tns_complex alpha, const tns_complex* a, int lda,
                      const tns_complex* b, int ldb, tns_complex beta, tns_complex* c, int ldc)
{
    return guarded(ctx, [&](Context* cx) {
        check_shape(m, k, lda, "A");
        check_shape(k, n, ldb, "B");
        check_shape(m, n, ldc, "C");
        // An empty destination returns before anything is read
        // (KernelEntry.cs:41-44), however many columns it nominally has.
        if (m == 0 || n == 0) return;
        check_operand(a, m, k, lda, "A");
        check_operand(b, k, n, ldb, "B");
        check_operand(c, m, n, ldc, "C");
        const zd *pa = (const zd*)a, *pb = (const zd*)b;
        zd* pc = (zd*)c;
        View<zd> va = bring(cx, pa, m, k, lda, true);
        View<zd> vb = bring(cx, pb, k, n, ldb, true);
        View<zd> vc = bring(cx, (const zd*)pc, m, n, ldc, !(beta.re == 0.0 && beta.im == 0.0));
        tns::zgemm(cx, m, n, k, c2(alpha), va.d, va.ld, vb.d, vb.ld, c2(beta), vc.d, vc.ld);
        send_back(cx, vc, pc, m, n, ldc);
        if (va.staged || vb.staged || vc.staged) tns::sync(cx);
    });
}

TNS_API int tns_dgetrf(tns_context* ctx, int m, int n, double* a, int lda, int nb, int* pivots, tns_lu_info* info)
{
    return guarded(ctx, [&](Context* c) { getrf_body<double>(c, m, n, a, lda, nb, pivots, info); });
}

TNS_API int tns_zgetrf(tns_context* ctx, int m, int n, tns_complex* a, int lda, int nb, int* pivots,
                       tns_lu_info* info)
{
    return guarded(ctx, [&](Context* c) { getrf_body<zd>(c, m, n, (zd*)a, lda, nb, pivots, info); });
}

TNS_API int tns_dgetrf_noalloc(tns_context* ctx, int m, int n, double* a, int lda, int nb, int* pivots, int* info)
{
    return guarded(ctx, [&](Context* c) {
        check_shape(m, n, lda, "A");
        if (nb > 64) fail(TNS_EINVAL, "block size above 64: the no-allocation LU takes nb <= 64 (or < 1 for the default)");
        const size_t need = 8 + (size_t)(m > 0 ? m : 1) + tns::lu_work_ints(m, n, nb);
        if (need > kNoAllocScratchInts)
            fail(TNS_ERANGE, "this factorization needs " + std::to_string(need * sizeof(int)) +
                                 " bytes of scratch; the no-allocation entry point holds " +
                                 std::to_string(kNoAllocScratchInts * sizeof(int)));
        if (!info || !tns::is_device_pointer(info)) fail(TNS_EINVAL, "info must be device memory");
        const int limit = std::min(m, n);
        if (limit > 0 && (!a || !pivots || !tns::is_device_pointer(a) || !tns::is_device_pointer(pivots)))
            fail(TNS_EINVAL, "A and pivots must be device memory");
        if (!c->noalloc_scratch) fail(TNS_EDEVICE, "the device scratch could not be located");
        // Layout: the LU's work area (8-byte aligned, at the start), then
        // the row payloads.
        int* work = c->noalloc_scratch;
        int* perm = work + (tns::lu_work_ints(m, n, nb) + 1) / 2 * 2;
        struct NoAlloc {
            Context* c;
            explicit NoAlloc(Context* x) : c(x) { c->no_alloc = true; }
            ~NoAlloc() { c->no_alloc = false; }
        } no_alloc(c);
        tns::dgetrf(c, m, n, a, lda, nb, pivots, perm, info, nullptr, work);
    });
}

TNS_API int tns_dgetrf_timed(tns_context* ctx, int m, int n, double* a, int lda, int nb, int* pivots,
                             tns_lu_info* info, tns_lu_phases* timings)
{
    return guarded(ctx, [&](Context* c) { getrf_body<double>(c, m, n, a, lda, nb, pivots, info, timings); });
}

TNS_API int tns_zgetrf_timed(tns_context* ctx, int m, int n, tns_complex* a, int lda, int nb, int* pivots,
                             tns_lu_info* info, tns_lu_phases* timings)
{
    return guarded(ctx, [&](Context* c) { getrf_body<zd>(c, m, n, (zd*)a, lda, nb, pivots, info, timings); });
}

TNS_API int tns_dgetrs(tns_context* ctx, int trans, int n, int nrhs, const double* lu, int ldlu, const int* pivots,
                       int singular_column, double* b, int ldb)
{
    return guarded(ctx, [&](Context* c) {
        getrs_body<double>(c, trans, n, nrhs, lu, ldlu, pivots, singular_column, b, ldb);
    });
}

TNS_API int tns_zgetrs(tns_context* ctx, int trans, int n, int nrhs, const tns_complex* lu, int ldlu,
                       const int* pivots, int singular_column, tns_complex* b, int ldb)
{
    return guarded(ctx, [&](Context* c) {
        getrs_body<zd>(c, trans, n, nrhs, (const zd*)lu, ldlu, pivots, singular_column, (zd*)b, ldb);
    });
}

TNS_API int tns_dtrsm(tns_context* ctx, int kind, int m, int n, const double* a, int lda, double* b, int ldb)
{
    return guarded(ctx, [&](Context* c) {
        check_operand(a, m, m, lda, "A");
        check_operand(b, m, n, ldb, "B");
        if (kind < 0 || kind > 5) fail(TNS_EINVAL, "unknown triangular solve kind");
        if (m == 0 || n == 0) return;
        View<double> va = bring(c, a, m, m, lda, true);
        View<double> vb = bring(c, b, m, n, ldb, true);
        tns::dtrsm(c, kind, m, n, va.d, va.ld, vb.d, vb.ld);
        send_back(c, vb, b, m, n, ldb);
        if (va.staged || vb.staged) tns::sync(c);
    });
}

TNS_API int tns_dpanel(tns_context* ctx, int trans, int n, int t, const double* a, int lda, const double* x, int ldx,
                       double* y, int ldy)
{
    return guarded(ctx, [&](Context* c) { panel_body<double>(c, trans, n, t, a, lda, x, ldx, y, ldy); });
}

TNS_API int tns_zpanel(tns_context* ctx, int adjoint, int n, int t, const tns_complex* a, int lda,
                       const tns_complex* x, int ldx, tns_complex* y, int ldy)
{
    return guarded(ctx, [&](Context* c) {
        panel_body<zd>(c, adjoint, n, t, (const zd*)a, lda, (const zd*)x, ldx, (zd*)y, ldy);
    });
}

TNS_API int tns_dnorm(tns_context* ctx, char kind, int m, int n, const double* a, int lda, double* out)
{
    return guarded(ctx, [&](Context* c) { norm_body<double>(c, kind, m, n, a, lda, out); });
}

TNS_API int tns_znorm(tns_context* ctx, char kind, int m, int n, const tns_complex* a, int lda, double* out)
{
    return guarded(ctx, [&](Context* c) { norm_body<zd>(c, kind, m, n, (const zd*)a, lda, out); });
}

TNS_API int tns_dnormest_power(tns_context* ctx, int n, const double* a, int lda, int power, int columns,
                               int max_iterations, int seed, tns_normest* out)
{
    return guarded(ctx, [&](Context* c) {
        normest_body<double>(c, n, a, lda, power, columns, max_iterations, seed, out);
    });
}

TNS_API int tns_znormest_power(tns_context* ctx, int n, const tns_complex* a, int lda, int power, int columns,
                               int max_iterations, int seed, tns_normest* out)
{
    return guarded(ctx, [&](Context* c) {
        normest_body<zd>(c, n, (const zd*)a, lda, power, columns, max_iterations, seed, out);
    });
}

TNS_API int tns_dgecon(tns_context* ctx, int n, const double* lu, int ldlu, const int* pivots, int singular_column,
                       double anorm, int columns, double* rcond)
{
    return guarded(ctx, [&](Context* c) {
        gecon_body<double>(c, n, lu, ldlu, pivots, singular_column, anorm, columns, rcond);
    });
}

TNS_API int tns_zgecon(tns_context* ctx, int n, const tns_complex* lu, int ldlu, const int* pivots,
                       int singular_column, double anorm, int columns, double* rcond)
{
    return guarded(ctx, [&](Context* c) {
        gecon_body<zd>(c, n, (const zd*)lu, ldlu, pivots, singular_column, anorm, columns, rcond);
    });
}

TNS_API int tns_dgesv_mixed(tns_context* ctx, int n, int nrhs, const double* a, int lda, const double* b, int ldb,
                            double* x, int ldx, int precision, double tol, int max_iter, tns_mixed_info* info)
{
    return guarded(ctx, [&](Context* c) { mixed_body(c, n, nrhs, a, lda, b, ldb, x, ldx, precision, tol, max_iter, info); });
}

TNS_API int tns_sgemm(tns_context* ctx, int precision, int m, int n, int k, float alpha, const float* a, int lda,
                      const float* b, int ldb, float beta, float* c, int ldc)
{
    return guarded(ctx, [&](Context* cx) {
        if (precision != 32 && precision != 16 && precision != 48)
            fail(TNS_EINVAL, "precision must be 32, 16 or 48");
        if (m < 0 || n < 0 || k < 0) fail(TNS_EINVAL, "negative dimension");
        check_operand(c, m, n, ldc, "C");
        if (m == 0 || n == 0) return;
        check_operand(a, m, k, lda, "A");
        check_operand(b, k, n, ldb, "B");
        View<float> va = bring(cx, a, m, k, lda, true);
        View<float> vb = bring(cx, b, k, n, ldb, true);
        View<float> vc = bring(cx, c, m, n, ldc, beta != 0.f);
        const int mode = precision == 16 ? tns::LP_BF16 : precision == 48 ? tns::LP_BF16X3 : tns::LP_FP32;
        tns::sgemm_lp(cx, mode, m, n, k, alpha, va.d, va.ld, vb.d, vb.ld, beta, vc.d, vc.ld);
        send_back(cx, vc, c, m, n, ldc);
        tns::sync(cx);
    });
}

TNS_API int tns_sgetrf(tns_context* ctx, int precision, int n, float* a, int lda, int* pivots, int* singular_column)
{
    return guarded(ctx, [&](Context* cx) {
        if (precision != 32 && precision != 16) fail(TNS_EINVAL, "precision must be 32 or 16");
        if (n < 0) fail(TNS_EINVAL, "negative order");
        check_operand(a, n, n, lda, "A");
        if (n == 0) {
            if (singular_column) *singular_column = -1;
            return;
        }
        View<float> va = bring(cx, a, n, n, lda, true);
        DevBuf<int> piv(cx, n), perm(cx, n), info(cx, 1);
        const int saved = cx->lp_mode;
        cx->lp_mode = precision == 16 ? tns::LP_BF16 : tns::LP_FP32;
        tns::sgetrf(cx, n, n, va.d, va.ld, 0, piv.get(), perm.get(), info.get());
        cx->lp_mode = saved;
        send_back(cx, va, a, n, n, lda);
        int h = -1;
        TNS_HIP(hipMemcpyAsync(pivots, piv.get(), sizeof(int) * (size_t)n, hipMemcpyDeviceToHost, cx->stream));
        TNS_HIP(hipMemcpyAsync(&h, info.get(), sizeof(int), hipMemcpyDeviceToHost, cx->stream));
        tns::sync(cx);
        if (singular_column) *singular_column = h;
    });
}

TNS_API int tns_sgetrs(tns_context* ctx, int precision, int n, int nrhs, const float* lu, int ldlu, const int* pivots,
                       float* b, int ldb)
{
    return guarded(ctx, [&](Context* cx) {
        if (precision != 32 && precision != 16) fail(TNS_EINVAL, "precision must be 32 or 16");
        if (n < 0 || nrhs < 0) fail(TNS_EINVAL, "negative dimension");
        check_operand(lu, n, n, ldlu, "LU");
        check_operand(b, n, nrhs, ldb, "B");
        if (n == 0 || nrhs == 0) return;
        std::vector<int> perm = permutation_of(cx, pivots, n, n);
        DevBuf<int> dperm = upload_ints(cx, perm);
        View<float> vlu = bring(cx, lu, n, n, ldlu, true);
        View<float> vb = bring(cx, b, n, nrhs, ldb, true);
        const int saved = cx->lp_mode;
        cx->lp_mode = precision == 16 ? tns::LP_BF16 : tns::LP_FP32;
        tns::sgetrs(cx, n, nrhs, vlu.d, vlu.ld, dperm.get(), vb.d, vb.ld);
        cx->lp_mode = saved;
        send_back(cx, vb, b, n, nrhs, ldb);
        tns::sync(cx);
    });
}

TNS_API int tns_zbatch_create(tns_context* ctx, int count, const int* n, tns_complex* const* a, const int* lda,
                              tns_complex* const* b, const int* ldb, int nrhs, int tune, tns_batch** out)
{
    if (!out) return TNS_EINVAL;
    *out = nullptr;
    return guarded(ctx, [&](Context* c) {
        if (count < 0 || nrhs < 0) fail(TNS_EINVAL, "negative count or right-hand side count");
        if (tune < 0 || tune > 2) fail(TNS_EINVAL, "tune must be 0, 1 or 2");
        if (count > 0 && (!n || !a || !lda || !b || !ldb)) fail(TNS_EINVAL, "null batch arrays");
        for (int i = 0; i < count; i++) {
            check_shape(n[i], n[i], lda[i], "A_i");
            check_shape(n[i], nrhs, ldb[i], "B_i");
            if (n[i] > 0 && (!tns::is_device_pointer(a[i]) || (nrhs > 0 && !tns::is_device_pointer(b[i]))))
                fail(TNS_EINVAL, "batched systems must be in device memory");
        }
        auto batch = std::make_unique<tns_batch>();
        batch->ctx = ctx;
        batch->plan = tns::make_batch(c, count, n, reinterpret_cast<zd* const*>(a), lda,
                                      reinterpret_cast<zd* const*>(b), ldb, nrhs, tune);
        ctx->batches.insert(batch.get());
        *out = batch.release();
    });
}

TNS_API int tns_zbatch_solve(tns_batch* batch, int* singular_columns)
{
    if (!batch) return TNS_EINVAL;
    return guarded(batch->ctx, [&](Context*) { batch->plan->solve(singular_columns); });
}

TNS_API int tns_zbatch_rcond(tns_batch* batch, int columns, double* rcond)
{
    if (!batch || !rcond) return TNS_EINVAL;
    return guarded(batch->ctx, [&](Context*) {
        if (columns < 1) fail(TNS_EINVAL, "columns must be at least 1");
        batch->plan->rcond(rcond, columns);
    });
}

TNS_API int tns_zbatch_pivots(tns_batch* batch, int index, int* pivots)
{
    if (!batch || !pivots) return TNS_EINVAL;
    return guarded(batch->ctx, [&](Context*) { batch->plan->pivots(index, pivots); });
}

TNS_API int tns_zbatch_info(tns_batch* batch, int* block, int* gemm_config, double* tuned_ms)
{
    if (!batch) return TNS_EINVAL;
    return guarded(batch->ctx, [&](Context*) {
        if (block) *block = batch->plan->block;
        if (gemm_config) *gemm_config = batch->plan->gemm_cfg;
        if (tuned_ms) *tuned_ms = batch->plan->tuned_ms;
    });
}

// The plan is released under the context's lock. After tns_context_destroy
// it is already gone (the context released it with its pool), and only the
// handle is freed here.
TNS_API int tns_zbatch_destroy(tns_batch* batch)
{
    if (!batch) return TNS_EINVAL;
    tns_context* ctx = batch->ctx;
    if (ctx) {
        std::lock_guard<std::mutex> lock(ctx->impl.gate);
        if (!ctx->disposed.load()) {
            (void)hipSetDevice(ctx->impl.device);
            (void)hipStreamSynchronize(ctx->impl.stream);
            batch->plan.reset();
            ctx->batches.erase(batch);
        }
    }
    delete batch;
    return TNS_OK;
}

TNS_API int tns_dexpmv_grad(tns_context* ctx, int n, int nb, const double* a, int lda, const double* b, int ldb,
                            double t, const double* g, int ldg, double* ga, int ldga, double* gb, int ldgb,
                            tns_expmv_diag* diag)
{
    return guarded(ctx, [&](Context* c) {
        expmv_grad_body<double>(c, n, nb, a, lda, b, ldb, t, g, ldg, ga, ldga, gb, ldgb, diag);
    });
}

TNS_API int tns_zexpmv_grad(tns_context* ctx, int n, int nb, const tns_complex* a, int lda, const tns_complex* b,
                            int ldb, double t, const tns_complex* g, int ldg, tns_complex* ga, int ldga,
                            tns_complex* gb, int ldgb, tns_expmv_diag* diag)
{
    return guarded(ctx, [&](Context* c) {
        expmv_grad_body<zd>(c, n, nb, (const zd*)a, lda, (const zd*)b, ldb, t, (const zd*)g, ldg, (zd*)ga, ldga,
                            (zd*)gb, ldgb, diag);
    });
}

TNS_API int tns_dexpm(tns_context* ctx, int n, const double* a, int lda, double* e, int lde, tns_expm_diag* diag)
{
    return guarded(ctx, [&](Context* c) { expm_body<double>(c, n, a, lda, e, lde, diag); });
}

TNS_API int tns_zexpm(tns_context* ctx, int n, const tns_complex* a, int lda, tns_complex* e, int lde,
                      tns_expm_diag* diag)
{
    return guarded(ctx, [&](Context* c) {
        expm_body<zd>(c, n, (const zd*)a, lda, (zd*)e, lde, diag);
    });
}

TNS_API int tns_dexpm_timed(tns_context* ctx, int n, const double* a, int lda, double* e, int lde,
                            tns_expm_diag* diag, tns_expm_phases* phases)
{
    return guarded(ctx, [&](Context* c) { expm_body<double>(c, n, a, lda, e, lde, diag, phases); });
}

TNS_API int tns_zexpm_timed(tns_context* ctx, int n, const tns_complex* a, int lda, tns_complex* e, int lde,
                            tns_expm_diag* diag, tns_expm_phases* phases)
{
    return guarded(ctx, [&](Context* c) {
        expm_body<zd>(c, n, (const zd*)a, lda, (zd*)e, lde, diag, phases);
    });
}

TNS_API int tns_dexpmv(tns_context* ctx, int n, int nb, const double* a, int lda, const double* b, int ldb, double t,
                       double* f, int ldf, tns_expmv_diag* diag)
{
    return guarded(ctx, [&](Context* c) { expmv_body<double>(c, n, nb, a, lda, b, ldb, t, f, ldf, diag); });
}

TNS_API int tns_zexpmv(tns_context* ctx, int n, int nb, const tns_complex* a, int lda, const tns_complex* b, int ldb,
                       double t, tns_complex* f, int ldf, tns_expmv_diag* diag)
{
    return guarded(ctx, [&](Context* c) {
        expmv_body<zd>(c, n, nb, (const zd*)a, lda, (const zd*)b, ldb, t, (zd*)f, ldf, diag);
    });
}

TNS_API int tns_dexpmv_op(tns_context* ctx, int n, tns_dapply_fn apply, void* user, int device_operator, int nb,
                          const double* b, int ldb, double t, double one_norm_bound, double* f, int ldf,
                          tns_expmv_diag* diag)
{
    return guarded(ctx, [&](Context* c) {
        expmv_op_body<double>(c, n, apply, user, device_operator, nb, b, ldb, t, one_norm_bound, f, ldf, diag);
    });
}

TNS_API int tns_zexpmv_op(tns_context* ctx, int n, tns_zapply_fn apply, void* user, int device_operator, int nb,
                          const tns_complex* b, int ldb, double t, double one_norm_bound, tns_complex* f, int ldf,
                          tns_expmv_diag* diag)
{
    return guarded(ctx, [&](Context* c) {
        expmv_op_body<zd>(c, n, apply, user, device_operator, nb, (const zd*)b, ldb, t, one_norm_bound,
                               (zd*)f, ldf, diag);
    });
}

}  // extern "C"
