// Batched solves of many complex systems of uneven order (a frequency
// sweep Z(w) X = B): every system's 1-norm, LU with partial pivoting
// (BlockedLu.Factor, BlockedLu.cs:45-141) and solve (BlockedLu.Solve,
// 148-206) are captured once into one hipGraph and run as one launch. The
// capture forks over S streams, so independent systems run side by side
// (a single system's panels occupy one CU); systems go to the stream with
// the least work so far, largest first. The reciprocal condition numbers
// (Condition.ReciprocalOne, NormEstimate.cs:432-466) run the estimator of
// every system in lockstep: each round is one replay of a captured graph
// of inverse products (all systems' A^-1 X, or all A^-H X) between pinned
// host panels, then the host logic of every estimate advances.
//
// The autotuner times the graph for a few stream counts and 4M GEMM tile
// shapes on scratch copies of the systems and keeps the fastest, per
// context and batch shape.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "algos.hpp"
#include "batch.hpp"

namespace tns {

namespace {

// Streams for capture forks; default priority (non-blocking).
struct Streams {
    std::vector<hipStream_t> s;
    explicit Streams(int count)
    {
        s.resize(count, nullptr);
        for (auto& x : s) TNS_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
    }
    ~Streams()
    {
        for (auto x : s)
            if (x) (void)hipStreamDestroy(x);
    }
};

struct Event {
    hipEvent_t e = nullptr;
    Event() { TNS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming)); }
    ~Event()
    {
        if (e) (void)hipEventDestroy(e);
    }
};

// Largest systems first, each to the least-loaded stream (cost n^3).
std::vector<int> assign_streams(const std::vector<int>& n, int streams)
{
    std::vector<int> order(n.size()), where(n.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
    std::vector<double> load(streams, 0.0);
    for (int i : order) {
        const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        where[i] = s;
        load[s] += (double)n[i] * n[i] * n[i];
    }
    return where;
}

// Capture body(i) for every system i on its stream into one graph. Every
// forked stream receives at least one system (never more streams than
// systems), so no branch of the capture is empty.
template <typename F, typename Pre = void (*)(), typename Post = void (*)()>
hipGraphExec_t capture(Context* ctx, const std::vector<int>& n, int streams, std::vector<std::pair<void*, size_t>>* hold,
                       F&& body, Pre pre = [] {}, Post post = [] {})
{
    if (n.empty()) fail(TNS_EINVAL, "an empty batch has nothing to capture");
    streams = std::max(1, std::min(streams, (int)n.size()));
    Streams side(streams);
    std::vector<int> where = assign_streams(n, streams);
    hipStream_t origin = ctx->stream;
    Event fork;
    std::vector<Event> join(streams);
    hipGraph_t graph = nullptr;
    TNS_HIP(hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed));
    ctx->hold = hold;
    try {
        pre();  // on the origin stream, before the fork
        TNS_HIP(hipEventRecord(fork.e, origin));
        for (int s = 0; s < streams; s++) TNS_HIP(hipStreamWaitEvent(side.s[s], fork.e, 0));
        for (size_t i = 0; i < n.size(); i++) {
            ctx->stream = side.s[where[i]];
            body((int)i);
        }
        ctx->stream = origin;
        for (int s = 0; s < streams; s++) {
            TNS_HIP(hipEventRecord(join[s].e, side.s[s]));
            TNS_HIP(hipStreamWaitEvent(origin, join[s].e, 0));
        }
        post();  // on the origin stream, after the join
        TNS_HIP(hipStreamEndCapture(origin, &graph));
    } catch (...) {
        ctx->stream = origin;
        ctx->hold = nullptr;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(origin, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        throw;
    }
    ctx->hold = nullptr;
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    TNS_HIP(e);
    return exec;
}

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

BatchPlan::~BatchPlan() { release(); }

void BatchPlan::release()
{
    for (auto e : {solve_exec, apply_exec, adjoint_exec})
        if (e) (void)hipGraphExecDestroy(e);
    solve_exec = apply_exec = adjoint_exec = nullptr;
    if (ctx) {
        for (auto& h : held) ctx->pool.give(h.first, h.second);
        for (auto& h : rc_held) ctx->pool.give(h.first, h.second);
    }
    held.clear();
    rc_held.clear();
    if (h_panels) (void)hipHostFree(h_panels);
    h_panels = nullptr;
    piv = DevBuf<int>();
    perm = DevBuf<int>();
    info = DevBuf<int>();
    anorm = DevBuf<double>();
    d_panels = DevBuf<zd>();
}

// The per-system work of a solve: ||A||_1 (before the factors overwrite
// it), LU in place, then X in place of B.
void BatchPlan::record_solve(int i, zd* ai, int ldai, zd* bi, int ldbi)
{
    const int ni = n[i];
    znorm(ctx, '1', ni, ni, ai, ldai, anorm.get() + i);
    zgetrf(ctx, ni, ni, ai, ldai, 0, piv.get() + offset[i], perm.get() + offset[i], info.get() + i);
    zgetrs(ctx, 0, ni, nrhs, ai, ldai, perm.get() + offset[i], bi, ldbi);
}

void BatchPlan::build(int streams_, int gemm_cfg_)
{
    if (solve_exec) (void)hipGraphExecDestroy(solve_exec);
    solve_exec = nullptr;
    for (auto& h : held) ctx->pool.give(h.first, h.second);
    held.clear();
    streams = std::max(1, std::min(streams_, count));
    gemm_cfg = gemm_cfg_;
    if (count == 0) return;
    gemm_force_config(-1, gemm_cfg);
    try {
        solve_exec = capture(ctx, n, streams, &held, [&](int i) { record_solve(i, a[i], lda[i], b[i], ldb[i]); });
    } catch (...) {
        gemm_force_config(-1, -1);
        throw;
    }
    gemm_force_config(-1, -1);
}

void BatchPlan::tune()
{
    // Scratch copies: the candidates factor these, never the caller's data.
    std::vector<DevBuf<zd>> sa, sb;
    for (int i = 0; i < count; i++) {
        sa.emplace_back(ctx, (size_t)n[i] * n[i]);
        sb.emplace_back(ctx, (size_t)n[i] * nrhs);
    }
    auto restore = [&]() {
        for (int i = 0; i < count; i++) {
            zcopy2d(ctx, n[i], n[i], a[i], lda[i], sa[i].get(), n[i]);
            zcopy2d(ctx, n[i], nrhs, b[i], ldb[i], sb[i].get(), n[i]);
        }
    };
    const int stream_options[] = {8, 16, 32, 64};
    const int cfg_options[] = {-1, 2};
    double best = 1e300;
    int best_s = 8, best_c = -1;
    for (int s : stream_options)
        for (int c : cfg_options) {
            if (s > count && s != stream_options[0]) continue;  // the same capture as a smaller count
            std::vector<std::pair<void*, size_t>> hold;
            gemm_force_config(-1, c);
            hipGraphExec_t exec = nullptr;
            try {
                exec = capture(ctx, n, s, &hold, [&](int i) { record_solve(i, sa[i].get(), n[i], sb[i].get(), n[i]); });
            } catch (...) {
                gemm_force_config(-1, -1);
                for (auto& h : hold) ctx->pool.give(h.first, h.second);
                throw;
            }
            gemm_force_config(-1, -1);
            double t = 1e300;
            for (int rep = 0; rep < 3; rep++) {
                restore();
                TNS_HIP(hipStreamSynchronize(ctx->stream));
                const double t0 = now_ms();
                TNS_HIP(hipGraphLaunch(exec, ctx->stream));
                TNS_HIP(hipStreamSynchronize(ctx->stream));
                t = std::min(t, now_ms() - t0);
            }
            (void)hipGraphExecDestroy(exec);
            for (auto& h : hold) ctx->pool.give(h.first, h.second);
            if (t < best) {
                best = t;
                best_s = s;
                best_c = c;
            }
        }
    tuned_ms = best;
    build(best_s, best_c);
}

void BatchPlan::solve(int* singular_columns)
{
    if (count == 0) return;
    TNS_HIP(hipGraphLaunch(solve_exec, ctx->stream));
    solved = true;
    if (singular_columns) {
        std::vector<int> h(count);
        TNS_HIP(hipMemcpyAsync(h.data(), info.get(), sizeof(int) * count, hipMemcpyDeviceToHost, ctx->stream));
        TNS_HIP(hipStreamSynchronize(ctx->stream));
        std::copy(h.begin(), h.end(), singular_columns);
    }
}

// Pinned host panels (in and out, n_i x t each) and their device twins,
// and the two graphs of inverse products between them.
void BatchPlan::build_rcond(int t)
{
    rc_t = t;
    size_t total = 0;
    rc_off.assign(count + 1, 0);
    for (int i = 0; i < count; i++) {
        rc_off[i] = total;
        total += (size_t)n[i] * t;
    }
    rc_off[count] = total;
    void* h = nullptr;
    TNS_HIP(hipHostMalloc(&h, sizeof(zd) * 2 * total, hipHostMallocDefault));
    h_panels = static_cast<zd*>(h);
    d_panels = DevBuf<zd>(ctx, 2 * total);
    // One copy of every input panel in, the solves side by side, one copy
    // of every result out (the panels are contiguous in both buffers).
    auto body = [&](int i, int trans) {
        const int ni = n[i];
        zd* dx = d_panels.get() + rc_off[i];
        zd* dy = d_panels.get() + total + rc_off[i];
        zcopy2d(ctx, ni, t, dx, ni, dy, ni);
        zgetrs(ctx, trans, ni, t, a[i], lda[i], perm.get() + offset[i], dy, ni);
    };
    auto pre = [&]() {
        TNS_HIP(hipMemcpyAsync(d_panels.get(), h_panels, sizeof(zd) * total, hipMemcpyHostToDevice, ctx->stream));
    };
    auto post = [&]() {
        TNS_HIP(hipMemcpyAsync(h_panels + total, d_panels.get() + total, sizeof(zd) * total, hipMemcpyDeviceToHost,
                               ctx->stream));
    };
    apply_exec = capture(ctx, n, streams, &rc_held, [&](int i) { body(i, 0); }, pre, post);
    adjoint_exec = capture(ctx, n, streams, &rc_held, [&](int i) { body(i, 1); }, pre, post);
}

void BatchPlan::rcond(double* out, int t)
{
    if (count == 0) return;
    if (!solved) fail(TNS_ESTATE, "the batch has not been factored: call solve first");
    if (!apply_exec || rc_t != t) {
        for (auto e : {apply_exec, adjoint_exec})
            if (e) (void)hipGraphExecDestroy(e);
        apply_exec = adjoint_exec = nullptr;
        for (auto& h : rc_held) ctx->pool.give(h.first, h.second);
        rc_held.clear();
        if (h_panels) (void)hipHostFree(h_panels);
        h_panels = nullptr;
        build_rcond(t);
    }
    std::vector<int> sing(count);
    std::vector<double> an(count);
    TNS_HIP(hipMemcpyAsync(sing.data(), info.get(), sizeof(int) * count, hipMemcpyDeviceToHost, ctx->stream));
    TNS_HIP(hipMemcpyAsync(an.data(), anorm.get(), sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
    TNS_HIP(hipStreamSynchronize(ctx->stream));

    const size_t total = rc_off[count];
    std::vector<NormEstRun<zd>> runs;
    runs.reserve(count);
    for (int i = 0; i < count; i++) runs.emplace_back(n[i], t, 5, 20260915);
    std::vector<char> active(count);
    for (int i = 0; i < count; i++) active[i] = sing[i] < 0 && an[i] != 0.0 && runs[i].want() != 0;
    for (int round = 0;; round++) {
        bool any = false;
        const int kind = round % 2 == 0 ? 1 : 2;
        for (int i = 0; i < count; i++) {
            if (!active[i]) continue;
            if (runs[i].want() != kind) fail(TNS_ESTATE, "estimates out of step");
            std::memcpy(h_panels + rc_off[i], runs[i].input().data(), sizeof(zd) * runs[i].input().size());
            any = true;
        }
        if (!any) break;
        TNS_HIP(hipGraphLaunch(kind == 1 ? apply_exec : adjoint_exec, ctx->stream));
        TNS_HIP(hipStreamSynchronize(ctx->stream));
        for (int i = 0; i < count; i++) {
            if (!active[i]) continue;
            std::vector<zd>& o = runs[i].output();
            std::memcpy(o.data(), h_panels + total + rc_off[i], sizeof(zd) * o.size());
            runs[i].resume();
            if (runs[i].want() == 0) active[i] = 0;
        }
    }
    for (int i = 0; i < count; i++) {
        if (sing[i] >= 0 || an[i] == 0.0 || n[i] == 0) {
            out[i] = 0.0;
            continue;
        }
        const double est = runs[i].result().value;
        out[i] = est == 0.0 ? 0.0 : 1.0 / (an[i] * est);
    }
}

void BatchPlan::pivots(int index, int* out) const
{
    if (index < 0 || index >= count) fail(TNS_EINVAL, "system index out of range");
    TNS_HIP(hipMemcpyAsync(out, piv.get() + offset[index], sizeof(int) * n[index], hipMemcpyDeviceToHost,
                           ctx->stream));
    TNS_HIP(hipStreamSynchronize(ctx->stream));
}

std::unique_ptr<BatchPlan> make_batch(Context* ctx, int count, const int* n, zd* const* a, const int* lda,
                                      zd* const* b, const int* ldb, int nrhs, int tune)
{
    auto p = std::make_unique<BatchPlan>();
    p->ctx = ctx;
    p->count = count;
    p->nrhs = nrhs;
    p->n.assign(n, n + count);
    p->lda.assign(lda, lda + count);
    p->ldb.assign(ldb, ldb + count);
    p->a.assign(a, a + count);
    p->b.assign(b, b + count);
    p->offset.assign(count + 1, 0);
    for (int i = 0; i < count; i++) p->offset[i + 1] = p->offset[i] + n[i];
    const size_t total = (size_t)p->offset[count];
    p->piv = DevBuf<int>(ctx, total);
    p->perm = DevBuf<int>(ctx, total);
    p->info = DevBuf<int>(ctx, count);
    p->anorm = DevBuf<double>(ctx, count);
    // Shape key of the tuner's memory: count, largest order, total n^3.
    double work = 0.0;
    int largest = 0;
    for (int i = 0; i < count; i++) {
        work += (double)n[i] * n[i] * n[i];
        largest = std::max(largest, n[i]);
    }
    const std::string key = std::to_string(count) + ":" + std::to_string(largest) + ":" +
                            std::to_string((long long)(work / 1e6)) + ":" + std::to_string(nrhs);
    auto it = ctx->batch_tuning.find(key);
    if (count == 0) {
        p->build(1, -1);
    } else if (it != ctx->batch_tuning.end() && tune != 2) {
        p->build(it->second.first, it->second.second);
    } else if (tune) {
        p->tune();
        ctx->batch_tuning[key] = {p->streams, p->gemm_cfg};
    } else {
        p->build(16, -1);
    }
    return p;
}

}  // namespace tns
