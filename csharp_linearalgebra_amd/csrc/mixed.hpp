// Mixed-precision solve with FP64 iterative refinement (mixed.hip).
#pragma once

#include "tns_internal.hpp"

namespace tns {

struct MixedInfo {
    int precision = 32;         // 32: FP32 factors; 16: BF16 products, FP32 factors
    int iterations = 0;         // refinement steps taken
    int fell_back = 0;          // 1: the FP64 LU produced the result
    double backward_error = 0;  // max_j ||b_j - A x_j||inf / (||A||inf ||x_j||inf + ||b_j||inf)
    double factor_ms = 0, refine_ms = 0, fallback_ms = 0;
    double first_error = 0;     // after the first, plain low-precision solve
};

// A X = B, all device pointers; A and B are left unchanged.
void dgesv_mixed(Context* ctx, int n, int t, const double* a, int lda, const double* b, int ldb, double* x, int ldx,
                 int precision, double tol, int max_iter, MixedInfo* info);

}  // namespace tns
