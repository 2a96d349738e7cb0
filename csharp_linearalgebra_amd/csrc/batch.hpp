// Batched solves of uneven complex systems as one captured graph of grouped
// launches (batch.hip).
#pragma once

#include <cstring>
#include <memory>
#include <utility>
#include <vector>

#include "tns_internal.hpp"

namespace tns {

// The kernels of a grouped schedule, in launch order, with their records
// in one device buffer.
struct Schedule {
    enum Kind { INIT, PANEL, MOVES, TRSM, GEMM, SOLVE };
    struct Launch {
        int kind;
        size_t offset;  // of the records, in bytes
        int count;      // records
        int blocks;     // workgroups (GEMM: tiles) of a flat launch
        int param;      // PANEL: tallest m; MOVES: longest list; GEMM: tile; SOLVE: trans
        int param2;     // SOLVE: right-hand sides per workgroup pass; others 0
        int param3;     // SOLVE: largest order
        size_t starts;  // flat launches: offset of the records' first workgroups (ints)
    };
    std::vector<Launch> launches;
    std::vector<unsigned char> host;
    DevBuf<unsigned char> dev;

    template <typename R>
    size_t put(const std::vector<R>& recs)
    {
        const size_t at = (host.size() + 15) & ~size_t(15);
        host.resize(at + sizeof(R) * recs.size());
        if (!recs.empty()) std::memcpy(host.data() + at, recs.data(), sizeof(R) * recs.size());
        return at;
    }
    // The records' `start` fields as an int array (flat launches).
    template <typename R>
    size_t put_starts(const std::vector<R>& recs)
    {
        std::vector<int> st;
        st.reserve(recs.size());
        for (const R& r : recs) st.push_back(r.start);
        return put(st);
    }
    // Copy the records to the device (before any capture that uses them).
    void upload(Context* ctx);
    // Issue every launch on ctx->stream.
    void record(Context* ctx) const;
};

struct BatchPlan {
    Context* ctx = nullptr;
    int count = 0, nrhs = 0;
    std::vector<int> n, lda, ldb, offset;  // offset: prefix sums of n
    std::vector<zd*> a, b;                 // the caller's device matrices
    std::vector<int> order;                // grouped systems, largest first
    std::vector<int> single;               // systems factored one by one (too large to group)
    DevBuf<int> piv, perm, info, lists;
    DevBuf<double> anorm;
    int block = 32, gemm_cfg = 2;          // panel width and grouped GEMM tile
    double tuned_ms = 0.0;                 // the tuner's best time, if it ran
    bool solved = false;
    Schedule solve_sched;
    hipGraphExec_t solve_exec = nullptr;
    std::vector<std::pair<void*, size_t>> held;  // scratch the one-by-one systems' nodes use
    // Reciprocal condition numbers.
    int rc_t = 0;
    std::vector<size_t> rc_off;
    zd* h_panels = nullptr;
    DevBuf<zd> d_panels;
    Schedule apply_sched, adjoint_sched;
    hipGraphExec_t apply_exec = nullptr, adjoint_exec = nullptr;
    std::vector<std::pair<void*, size_t>> rc_held;

    ~BatchPlan();
    void release();
    // Schedule of norm + LU + solve for matrices A (lda) and right-hand
    // sides B (ldb) with panel width nb and GEMM tile cfg.
    void plan_solve(Schedule& s, const std::vector<zd*>& A, const std::vector<zd*>& B, const std::vector<int>& LDA,
                    const std::vector<int>& LDB, int nb, int cfg) const;
    hipGraphExec_t capture_solve(Schedule& s, const std::vector<zd*>& A, const std::vector<zd*>& B,
                                 const std::vector<int>& LDA, const std::vector<int>& LDB, int nb, int cfg,
                                 std::vector<std::pair<void*, size_t>>* hold);
    void build(int nb, int cfg);
    void tune();
    // One graph launch: ||A_i||_1, A_i := LU, B_i := X_i. Waits only if
    // singular_columns is wanted.
    void solve(int* singular_columns);
    void build_rcond(int t);
    // rcond of every system from its factors (after solve()).
    void rcond(double* out, int t);
    void pivots(int index, int* out) const;
};

// tune: 0 build with the defaults (panel width 32, 64 x 32 tiles); 1 use the
// context's remembered choice for this batch shape, tuning if there is
// none; 2 tune again.
std::unique_ptr<BatchPlan> make_batch(Context* ctx, int count, const int* n, zd* const* a, const int* lda,
                                      zd* const* b, const int* ldb, int nrhs, int tune);

}  // namespace tns
