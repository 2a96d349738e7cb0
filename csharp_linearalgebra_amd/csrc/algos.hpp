// Host-side drivers of the device algorithms: the norm estimator and its
// operators, condition estimation, the exponential and its action. The
// control flow (iteration tests, parameter selection, branch choices) runs
// on the host exactly as the reference's does; every O(n^2) or O(n^3) step
// runs on the device.
#pragma once

#include <functional>
#include <vector>

#include "tns_internal.hpp"

namespace tns {

// A square operator applied to n x t panels that live on the device (ld n).
template <typename T>
struct DevOp {
    int n = 0;
    std::function<void(int t, const T* x, T* y)> apply;
    std::function<void(int t, const T* x, T* y)> adjoint;
};

struct NormEst {
    double value = 0.0;
    int iterations = 0;
    int products = 0;
};

// NormEstimate.Estimate (NormEstimate.cs:104-274).
template <typename T>
NormEst normest(Context* ctx, const DevOp<T>& op, int columns, int max_iterations, int seed);

// DenseMatrixOperator / ComplexDenseOperator raised to a power
// (LinearOperators.cs:89-293) over a device matrix a (n x n, ld lda).
template <typename T>
DevOp<T> power_operator(Context* ctx, int n, const T* a, int lda, int power);

// LuInverseOperator (LinearOperators.cs:303-350) over device factors and
// the row permutation produced by getrf.
template <typename T>
DevOp<T> inverse_operator(Context* ctx, int n, const T* lu, int ldlu, const int* d_perm);

// Condition.ReciprocalOne (NormEstimate.cs:432-466).
template <typename T>
double gecon(Context* ctx, int n, const T* lu, int ldlu, const int* d_perm, int singular_column, double anorm,
             int columns);

struct ExpmDiag {
    int degree = 0;
    int squarings = 0;
};

// MatrixExponential.ScalingAndSquaring (MatrixExponential.cs:181-270) on a
// device matrix a (n x n); writes exp(A) to e (ld lde). Throws TNS_ESTATE if
// the Pade denominator is exactly singular.
template <typename T>
void expm(Context* ctx, int n, const T* a, int lda, T* e, int lde, ExpmDiag* diag);

struct ExpmvDiag {
    int degree = 0;
    int scaling = 0;
    int applications = 0;
    int early_exits = 0;
};

// MatrixExponentialAction.Expmv, dense (MatrixExponentialAction.cs:102-178).
template <typename T>
void expmv(Context* ctx, int n, int nb, const T* a, int lda, const T* b, int ldb, double t, T* f, int ldf,
           ExpmvDiag* diag);

// MatrixExponentialAction.Run over an arbitrary operator with a 1-norm
// bound: the matrix-free path (MatrixExponentialAction.cs:203-230).
template <typename T>
void expmv_op(Context* ctx, const DevOp<T>& op, int nb, const T* b, int ldb, double t, double one_norm_bound, T* f,
              int ldf, ExpmvDiag* diag);

// Read device scalars back (waits for the stream).
double read_double(Context* ctx, const double* d);
void read_doubles(Context* ctx, const double* d, double* h, int count);

}  // namespace tns
