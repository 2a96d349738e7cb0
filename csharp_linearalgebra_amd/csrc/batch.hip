// Batched solves of many complex systems of uneven order (a frequency
// sweep Z(w) X = B): every system's 1-norm, LU with partial pivoting
// (BlockedLu.Factor, BlockedLu.cs:45-141) and solve (BlockedLu.Solve,
// 148-206), as one captured hipGraph run as one launch.
//
// The graph is a chain of grouped launches: each block step of the LU is
// one launch per kernel for all systems at once (their panels side by side,
// one workgroup each; their interchanges, triangular solves and trailing
// updates as flat grids over every system's columns or tiles). Systems are
// taken largest first; step j covers the systems of order > j * nb, so the
// work of one launch spreads over the whole chip instead of one system's
// panel occupying one CU while the rest wait, and the graph has a few
// hundred nodes instead of a hundred per system. Each system's part of a
// step is the right-looking block step of its own factorization (getrf_impl
// and panel() in lu.hip): the same panel kernel, chunk widths and move
// lists; the block row U12 = L11^-1 A12 is solved in diagonal blocks of 32
// with a product between them (trsm_impl's blocked order), and the trailing
// update is one grouped 4M product.
//
// The reciprocal condition numbers (Condition.ReciprocalOne,
// NormEstimate.cs:432-466) run the estimator of every system in lockstep:
// each round is one replay of a captured graph (one copy of all probe
// panels in, one grouped solve of all systems, one copy out), then the host
// logic of every estimate advances.
//
// The autotuner times the graph for panel widths {32, 64} and two grouped
// GEMM tiles on scratch copies of the systems and keeps the fastest, per
// context and batch shape.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstring>
#include <exception>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <vector>

#include "algos.hpp"
#include "batch.hpp"
#include "grouped.hpp"

namespace tns {

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Ints per system for a panel's move list: 1 + 2 * (2 * widest chunk).
constexpr int LIST_INTS = 1 + 2 * 64;

// Capture body() on ctx->stream into a graph and instantiate it.
template <typename F>
hipGraphExec_t capture(Context* ctx, std::vector<std::pair<void*, size_t>>* hold, F&& body)
{
    hipStream_t origin = ctx->stream;
    hipGraph_t graph = nullptr;
    TNS_HIP(hipStreamBeginCapture(origin, hipStreamCaptureModeRelaxed));
    ctx->hold = hold;
    ctx->capturing = true;
    try {
        body();
        ctx->capturing = false;
        TNS_HIP(hipStreamEndCapture(origin, &graph));
    } catch (...) {
        ctx->stream = origin;
        ctx->hold = nullptr;
        ctx->capturing = false;
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

// Flat launches: give each record its first workgroup; returns the total.
template <typename R, typename Blocks>
int assign_starts(std::vector<R>& recs, Blocks&& blocks_of)
{
    int total = 0;
    for (auto& r : recs) {
        r.start = total;
        total += blocks_of(r);
    }
    return total;
}

}  // namespace

void Schedule::upload(Context* ctx)
{
    dev = DevBuf<unsigned char>(ctx, host.empty() ? 16 : host.size());
    if (!host.empty())
        TNS_HIP(hipMemcpyAsync(dev.get(), host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
    TNS_HIP(hipStreamSynchronize(ctx->stream));
}

void Schedule::record(Context* ctx) const
{
    for (const Launch& l : launches) {
        const void* p = dev.get() + l.offset;
        const int* starts = reinterpret_cast<const int*>(dev.get() + l.starts);
        switch (l.kind) {
        case INIT: zgrouped_init(ctx, static_cast<const InitRec<zd>*>(p), l.count); break;
        case PANEL: zgrouped_panel(ctx, static_cast<const PanelRec<zd>*>(p), l.count, l.param); break;
        case MOVES:
            zgrouped_moves(ctx, static_cast<const MovesRec<zd>*>(p), starts, l.count, l.blocks, l.param);
            break;
        case TRSM: zgrouped_trsm(ctx, static_cast<const TrsmRec<zd>*>(p), starts, l.count, l.blocks); break;
        case GEMM: zgrouped_gemm(ctx, static_cast<const GemmRec<zd>*>(p), starts, l.count, l.blocks, l.param); break;
        case SOLVE:
            zgrouped_solve(ctx, static_cast<const SolveRec<zd>*>(p), l.count, l.param, l.param2, l.param3);
            break;
        default: fail(TNS_EINVAL, "unknown grouped launch");
        }
    }
}

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
    solve_sched = Schedule();
    apply_sched = Schedule();
    adjoint_sched = Schedule();
    piv = DevBuf<int>();
    perm = DevBuf<int>();
    info = DevBuf<int>();
    lists = DevBuf<int>();
    anorm = DevBuf<double>();
    d_panels = DevBuf<zd>();
}

// getrf_impl + getrs_impl of every grouped system, step by step. Block step
// jb factors the m x w panel at (jb, jb) in column chunks of the width the
// register kernel allows for its height (panel() in lu.hip): chunk c0
// factors m - c0 rows, moves its interchanges to every other column,
// solves and updates the panel columns to its right; then U12 := L11^-1 A12
// and A22 -= A21 U12. Chunks start at multiples of their width, so the
// launches of step jb go by chunk offset c0 = 0, 4, 8, ... and, within it,
// panel, moves, solve, update: each system meets its own operations in its
// own order.
void BatchPlan::plan_solve(Schedule& s, const std::vector<zd*>& A, const std::vector<zd*>& B,
                           const std::vector<int>& LDA, const std::vector<int>& LDB, int nb, int cfg) const
{
    s.launches.clear();
    s.host.clear();
    if (order.empty()) return;
    auto add = [&](int kind, size_t off, int cnt, int blocks, int param, int param2 = 0, int param3 = 0,
                   size_t starts = 0) {
        if (cnt > 0) s.launches.push_back({kind, off, cnt, blocks, param, param2, param3, starts});
    };
    const int maxn = n[order[0]];
    int* d_piv = piv.get();
    int* d_perm = perm.get();
    int* d_info = info.get();
    int* d_lists = lists.get();
    double* d_anorm = anorm.get();
    int bm = 64, bn = 32;
    grouped_gemm_tile(cfg, bm, bn);

    {
        std::vector<InitRec<zd>> init;
        for (int i : order) init.push_back({A[i], d_perm + offset[i], d_info + i, d_anorm + i, LDA[i], n[i]});
        add(Schedule::INIT, s.put(init), (int)init.size(), 0, 0);
    }
    auto emit_trsm = [&](std::vector<TrsmRec<zd>>& tr) {
        const int tb = assign_starts(tr, [](const TrsmRec<zd>& r) {
            return (r.ncols + GROUPED_TRSM_COLUMNS - 1) / GROUPED_TRSM_COLUMNS;
        });
        const size_t at = s.put(tr);
        add(Schedule::TRSM, at, (int)tr.size(), tb, 0, 0, 0, s.put_starts(tr));
    };
    auto emit_gemm = [&](std::vector<GemmRec<zd>>& gm) {
        const int gb = assign_starts(gm, [&](const GemmRec<zd>& r) {
            return ((r.m + bm - 1) / bm) * ((r.n + bn - 1) / bn);
        });
        const size_t at = s.put(gm);
        add(Schedule::GEMM, at, (int)gm.size(), gb, cfg, 0, 0, s.put_starts(gm));
    };
    constexpr int TB = 32;  // diagonal block of the complex triangular solves

    for (int jb = 0; jb < maxn; jb += nb) {
        for (int c0 = 0; c0 < nb; c0 += 4) {
            std::vector<PanelRec<zd>> pr[4];
            int maxm[4] = {0, 0, 0, 0};
            std::vector<MovesRec<zd>> mv;
            int longest = 0;
            std::vector<TrsmRec<zd>> tr;
            std::vector<GemmRec<zd>> gm;
            for (int i : order) {
                const int ni = n[i];
                if (ni <= jb) break;
                const int m = ni - jb, w = nb < m ? nb : m;
                const int cw = zpanel_chunk_width(m);
                if (c0 >= w || c0 % cw != 0) continue;
                const int cwid = cw < w - c0 ? cw : w - c0;
                const int ld = LDA[i];
                zd* top = A[i] + jb + (size_t)jb * ld;
                zd* sub = top + c0 + (size_t)c0 * ld;
                const int mc = m - c0;
                const int cls = zpanel_class(mc);
                int* list = d_lists + (size_t)i * LIST_INTS;
                pr[cls].push_back({sub, d_piv + offset[i] + jb + c0, d_perm + offset[i] + jb + c0, list, d_info + i, ld,
                                   mc, cwid, jb + c0});
                maxm[cls] = std::max(maxm[cls], mc);
                if (ni - cwid > 0) {
                    mv.push_back({A[i] + jb + c0, list, ld, jb + c0, jb + c0 + cwid, ni, 0});
                    longest = std::max(longest, 2 * cwid);
                }
                if (c0 + cwid < w) {
                    tr.push_back({sub, sub + (size_t)cwid * ld, ld, ld, cwid, w - c0 - cwid, 0});
                    if (mc > cwid)
                        gm.push_back({sub + cwid, sub + (size_t)cwid * ld, sub + cwid + (size_t)cwid * ld, ld, ld, ld,
                                      mc - cwid, w - c0 - cwid, cwid, 0});
                }
            }
            for (int c = 0; c < 4; c++) add(Schedule::PANEL, s.put(pr[c]), (int)pr[c].size(), 0, maxm[c]);
            if (!mv.empty()) {
                const int cpb = grouped_moves_columns(longest);
                const int blocks = assign_starts(mv, [&](const MovesRec<zd>& r) {
                    return (r.c0 + (r.ncols - r.c1) + cpb - 1) / cpb;
                });
                const size_t at = s.put(mv);
                add(Schedule::MOVES, at, (int)mv.size(), blocks, longest, 0, 0, s.put_starts(mv));
            }
            emit_trsm(tr);
            emit_gemm(gm);
        }
        // U12 := L11^-1 A12 by diagonal blocks of TB rows, each solved and
        // then subtracted from the rows below it (trsm_impl's order), then
        // A22 -= A21 U12.
        for (int q0 = 0; q0 < nb; q0 += TB) {
            std::vector<TrsmRec<zd>> tr;
            std::vector<GemmRec<zd>> gm;
            for (int i : order) {
                const int ni = n[i];
                if (ni <= jb) break;
                const int m = ni - jb, w = nb < m ? nb : m;
                if (jb + w >= ni || q0 >= w) continue;
                const int kb = TB < w - q0 ? TB : w - q0;
                const int ld = LDA[i];
                zd* top = A[i] + jb + (size_t)jb * ld;
                zd* bq = top + q0 + (size_t)w * ld;
                tr.push_back({top + q0 + (size_t)q0 * ld, bq, ld, ld, kb, ni - jb - w, 0});
                if (q0 + kb < w)
                    gm.push_back({top + q0 + kb + (size_t)q0 * ld, bq, bq + kb, ld, ld, ld, w - q0 - kb, ni - jb - w, kb,
                                  0});
            }
            emit_trsm(tr);
            emit_gemm(gm);
        }
        std::vector<GemmRec<zd>> gm;
        for (int i : order) {
            const int ni = n[i];
            if (ni <= jb) break;
            const int m = ni - jb, w = nb < m ? nb : m;
            if (jb + w >= ni) continue;
            const int ld = LDA[i];
            zd* top = A[i] + jb + (size_t)jb * ld;
            gm.push_back({top + w, top + (size_t)w * ld, top + w + (size_t)w * ld, ld, ld, ld, m - w, ni - jb - w, w, 0});
        }
        emit_gemm(gm);
    }

    std::vector<SolveRec<zd>> sv;
    for (int i : order)
        sv.push_back({A[i], d_perm + offset[i], B[i], B[i], d_info + i, LDA[i], LDB[i], LDB[i], n[i], nrhs});
    add(Schedule::SOLVE, s.put(sv), (int)sv.size(), 0, 0, grouped_solve_group(ctx, maxn), maxn);
}

hipGraphExec_t BatchPlan::capture_solve(Schedule& s, const std::vector<zd*>& A, const std::vector<zd*>& B,
                                        const std::vector<int>& LDA, const std::vector<int>& LDB, int nb, int cfg,
                                        std::vector<std::pair<void*, size_t>>* hold)
{
    plan_solve(s, A, B, LDA, LDB, nb, cfg);
    s.upload(ctx);
    return capture(ctx, hold, [&]() {
        s.record(ctx);
        // Systems too large to group: the one-system path, in the same graph.
        for (int i : single) {
            znorm(ctx, '1', n[i], n[i], A[i], LDA[i], anorm.get() + i);
            zgetrf(ctx, n[i], n[i], A[i], LDA[i], 0, piv.get() + offset[i], perm.get() + offset[i], info.get() + i);
            zgetrs(ctx, 0, n[i], nrhs, A[i], LDA[i], perm.get() + offset[i], B[i], LDB[i]);
        }
    });
}

void BatchPlan::build(int nb, int cfg)
{
    if (solve_exec) (void)hipGraphExecDestroy(solve_exec);
    solve_exec = nullptr;
    for (auto& h : held) ctx->pool.give(h.first, h.second);
    held.clear();
    block = nb;
    gemm_cfg = cfg;
    if (count == 0) return;
    solve_exec = capture_solve(solve_sched, a, b, lda, ldb, nb, cfg, &held);
}

void BatchPlan::tune()
{
    // Scratch copies: the candidates factor these, never the caller's data.
    std::vector<DevBuf<zd>> sa, sb;
    std::vector<zd*> pa(count), pb(count);
    std::vector<int> la(count), lb(count);
    for (int i = 0; i < count; i++) {
        la[i] = lb[i] = std::max(1, n[i]);
        sa.emplace_back(ctx, (size_t)la[i] * std::max(1, n[i]));
        sb.emplace_back(ctx, (size_t)lb[i] * std::max(1, nrhs));
        pa[i] = sa[i].get();
        pb[i] = sb[i].get();
    }
    auto restore = [&]() {
        for (int i = 0; i < count; i++) {
            zcopy2d(ctx, n[i], n[i], a[i], lda[i], pa[i], la[i]);
            zcopy2d(ctx, n[i], nrhs, b[i], ldb[i], pb[i], lb[i]);
        }
    };
    double best = 1e300;
    int best_nb = 32, best_cfg = 2;
    for (int nb : {32, 64})
        for (int cfg : {1, 2}) {
            std::vector<std::pair<void*, size_t>> hold;
            Schedule sched;
            hipGraphExec_t exec = nullptr;
            try {
                exec = capture_solve(sched, pa, pb, la, lb, nb, cfg, &hold);
            } catch (...) {
                for (auto& h : hold) ctx->pool.give(h.first, h.second);
                throw;
            }
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
                best_nb = nb;
                best_cfg = cfg;
            }
        }
    tuned_ms = best;
    build(best_nb, best_cfg);
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
        for (int v : h)
            if (v == LU_ABORTED) fail(TNS_EDEVICE, "a system's LU panel exchange timed out");
        std::copy(h.begin(), h.end(), singular_columns);
    }
}

// Pinned host panels (in and out, n_i x t each) and their device twins,
// and the two graphs of inverse products between them: one copy of every
// input panel in, one grouped solve of every system, one copy of every
// result out (the panels are contiguous in both buffers).
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
    TNS_HIP(hipHostMalloc(&h, sizeof(zd) * std::max<size_t>(1, 2 * total), hipHostMallocDefault));
    h_panels = static_cast<zd*>(h);
    d_panels = DevBuf<zd>(ctx, std::max<size_t>(1, 2 * total));
    int maxn = 0;
    for (int i : order) maxn = std::max(maxn, n[i]);
    // Right-hand sides the grouped solve takes at a time: no more than the
    // estimates' t (each column gets the same operations whatever the group).
    int group = grouped_solve_group(ctx, maxn);
    if (group > 2 && t <= 2) group = 2;
    if (group > 1 && t <= 1) group = 1;
    auto graph = [&](Schedule& s, int trans) {
        std::vector<SolveRec<zd>> sv;
        for (int i : order) {
            zd* dx = d_panels.get() + rc_off[i];
            zd* dy = d_panels.get() + total + rc_off[i];
            const int ld = std::max(1, n[i]);
            sv.push_back({a[i], perm.get() + offset[i], dx, dy, nullptr, lda[i], ld, ld, n[i], t});
        }
        s.launches.clear();
        s.host.clear();
        if (!sv.empty()) s.launches.push_back({Schedule::SOLVE, s.put(sv), (int)sv.size(), 0, trans, group, maxn, 0});
        s.upload(ctx);
        return capture(ctx, &rc_held, [&]() {
            TNS_HIP(hipMemcpyAsync(d_panels.get(), h_panels, sizeof(zd) * total, hipMemcpyHostToDevice, ctx->stream));
            s.record(ctx);
            for (int i : single) {
                zd* dx = d_panels.get() + rc_off[i];
                zd* dy = d_panels.get() + total + rc_off[i];
                zcopy2d(ctx, n[i], t, dx, n[i], dy, n[i]);
                zgetrs(ctx, trans, n[i], t, a[i], lda[i], perm.get() + offset[i], dy, n[i]);
            }
            TNS_HIP(hipMemcpyAsync(h_panels + total, d_panels.get() + total, sizeof(zd) * total,
                                   hipMemcpyDeviceToHost, ctx->stream));
        });
    };
    apply_exec = graph(apply_sched, 0);
    adjoint_exec = graph(adjoint_sched, 1);
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
    for (int v : sing)
        if (v == LU_ABORTED) fail(TNS_EDEVICE, "a system's LU panel exchange timed out");

    const size_t total = rc_off[count];
    std::vector<std::unique_ptr<NormEstRun<zd>>> runs(count);
    std::vector<char> active(count);
    ctx->host.run(count, [&](int i) {
        runs[i] = std::make_unique<NormEstRun<zd>>(n[i], t, 5, 20260915);
        active[i] = sing[i] < 0 && an[i] != 0.0 && runs[i]->want() != 0;
    });
    for (int round = 0;; round++) {
        bool any = false;
        const int kind = round % 2 == 0 ? 1 : 2;
        for (int i = 0; i < count; i++) {
            if (!active[i]) continue;
            if (runs[i]->want() != kind) fail(TNS_ESTATE, "estimates out of step");
            any = true;
        }
        if (!any) break;
        ctx->host.run(count, [&](int i) {
            if (active[i])
                std::memcpy(h_panels + rc_off[i], runs[i]->input().data(), sizeof(zd) * runs[i]->input().size());
        });
        TNS_HIP(hipGraphLaunch(kind == 1 ? apply_exec : adjoint_exec, ctx->stream));
        TNS_HIP(hipStreamSynchronize(ctx->stream));
        ctx->host.run(count, [&](int i) {
            if (!active[i]) return;
            std::vector<zd>& o = runs[i]->output();
            std::memcpy(o.data(), h_panels + total + rc_off[i], sizeof(zd) * o.size());
            runs[i]->resume();
            if (runs[i]->want() == 0) active[i] = 0;
        });
    }
    for (int i = 0; i < count; i++) {
        if (sing[i] >= 0 || an[i] == 0.0 || n[i] == 0) {
            out[i] = 0.0;
            continue;
        }
        const double est = runs[i]->result().value;
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
    // Grouped: orders the register panel kernel holds and whose solve fits
    // one workgroup's LDS; the rest go one by one.
    for (int i = 0; i < count; i++) {
        const bool grouped = zpanel_chunk_width(std::max(1, n[i])) > 0 && grouped_solve_group(ctx, n[i]) > 0;
        (grouped ? p->order : p->single).push_back(i);
    }
    std::stable_sort(p->order.begin(), p->order.end(), [&](int x, int y) { return n[x] > n[y]; });
    const size_t total = (size_t)p->offset[count];
    p->piv = DevBuf<int>(ctx, std::max<size_t>(1, total));
    p->perm = DevBuf<int>(ctx, std::max<size_t>(1, total));
    p->info = DevBuf<int>(ctx, (size_t)std::max(1, count));
    p->lists = DevBuf<int>(ctx, (size_t)std::max(1, count) * LIST_INTS);
    p->anorm = DevBuf<double>(ctx, (size_t)std::max(1, count));
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
        p->build(32, 2);
    } else if (it != ctx->batch_tuning.end() && tune != 2) {
        p->build(it->second.first, it->second.second);
    } else if (tune) {
        p->tune();
        ctx->batch_tuning[key] = {p->block, p->gemm_cfg};
    } else {
        p->build(32, 2);
    }
    // The estimator's graphs for the reference's two probe columns, built
    // with the plan so rcond() only runs them.
    if (count > 0) p->build_rcond(2);
    return p;
}

}  // namespace tns
