// Host-side drivers of the device algorithms: the norm estimator and its
// operators, condition estimation, the exponential and its action. The
// control flow (iteration tests, parameter selection, branch choices) runs
// on the host exactly as the reference's does; every O(n^2) or O(n^3) step
// runs on the device.
#pragma once

#include <climits>
#include <cstdlib>
#include <functional>
#include <vector>

#include "dotnet_bcl.hpp"
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

// NormEstimate.Estimate (NormEstimate.cs:104-274) as a resumable run.
// want(): 0 finished, 1 needs out = A * input, 2 needs out = A^H * input
// (n x columns panels, column-major, ld n); resume() consumes output().
template <typename T>
class NormEstRun {
public:
    NormEstRun(int n, int columns, int max_iterations, int seed);
    int want() const { return want_; }
    int columns() const { return t_; }
    const std::vector<T>& input() const { return want_ == 2 ? s_ : x_; }
    std::vector<T>& output() { return out_; }
    void resume();
    NormEst result() const;

private:
    int n_ = 0, t_ = 1, itmax_ = 2, want_ = 0;
    int k_ = 0, iterations_ = 0, products_ = 0, best_index_ = -1, used_count_ = 0;
    double est_ = 0.0, est_old_ = 0.0;
    std::vector<T> x_, s_, s_old_, out_;
    std::vector<double> h_, keys_;
    std::vector<int> order_, scratch_, current_;
    std::vector<char> used_;
    SeededRandom rng_;
};

// NormEstimate.Estimate driven to completion over a device operator.
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

// Device milliseconds of one expm by phase: the powers A^2, A^4, A^6 and
// their norms; the degree and scaling selection; the Pade products; the
// denominator's LU; its solve; the squarings.
struct ExpmPhases {
    enum { POWERS = 0, SELECT = 1, PADE = 2, LU = 3, SOLVE = 4, SQUARE = 5 };
    double ms[6] = {0, 0, 0, 0, 0, 0};
    double total_ms = 0;
};

// MatrixExponential.ScalingAndSquaring (MatrixExponential.cs:181-270) on a
// device matrix a (n x n); writes exp(A) to e (ld lde). Throws TNS_ESTATE if
// the Pade denominator is exactly singular. phases, if given, receives the
// phase split (the call then waits for the result).
template <typename T>
void expm(Context* ctx, int n, const T* a, int lda, T* e, int lde, ExpmDiag* diag, ExpmPhases* phases = nullptr);

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

// Gradients through F = exp(tA) B (computed as expmv computes it) for an
// upstream gradient G = dL/dF (n x nb): dL/dA into ga (n x n), dL/dB into gb
// (n x nb). All device pointers.
template <typename T>
void expmv_grad(Context* ctx, int n, int nb, const T* a, int lda, const T* b, int ldb, double t, const T* g, int ldg,
                T* ga, int ldga, T* gb, int ldgb, ExpmvDiag* diag);

// Read device scalars back (waits for the stream).
double read_double(Context* ctx, const double* d);
void read_doubles(Context* ctx, const double* d, double* h, int count);

}  // namespace tns
