// Batched solves of uneven complex systems in one captured graph (batch.hip).
#pragma once

#include <memory>
#include <utility>
#include <vector>

#include "tns_internal.hpp"

namespace tns {

struct BatchPlan {
    Context* ctx = nullptr;
    int count = 0, nrhs = 0;
    std::vector<int> n, lda, ldb, offset;  // offset: prefix sums of n
    std::vector<zd*> a, b;                 // the caller's device matrices
    DevBuf<int> piv, perm, info;
    DevBuf<double> anorm;
    int streams = 16, gemm_cfg = -1;       // the chosen capture fan-out and 4M tile
    double tuned_ms = 0.0;                 // the tuner's best time, if it ran
    bool solved = false;
    hipGraphExec_t solve_exec = nullptr;
    std::vector<std::pair<void*, size_t>> held;  // scratch the solve graph uses
    // Reciprocal condition numbers.
    int rc_t = 0;
    std::vector<size_t> rc_off;
    zd* h_panels = nullptr;
    DevBuf<zd> d_panels;
    hipGraphExec_t apply_exec = nullptr, adjoint_exec = nullptr;
    std::vector<std::pair<void*, size_t>> rc_held;

    ~BatchPlan();
    void release();
    void record_solve(int i, zd* ai, int ldai, zd* bi, int ldbi);
    void build(int streams, int gemm_cfg);
    void tune();
    // One graph launch: ||A_i||_1, A_i := LU, B_i := X_i. Waits only if
    // singular_columns is wanted.
    void solve(int* singular_columns);
    void build_rcond(int t);
    // rcond of every system from its factors (after solve()).
    void rcond(double* out, int t);
    void pivots(int index, int* out) const;
};

// tune: 0 build with the defaults (16 streams, heuristic tiles); 1 use the
// context's remembered choice for this batch shape, tuning if there is
// none; 2 tune again.
std::unique_ptr<BatchPlan> make_batch(Context* ctx, int count, const int* n, zd* const* a, const int* lda,
                                      zd* const* b, const int* ldb, int nrhs, int tune);

}  // namespace tns
