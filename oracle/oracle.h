/*
 * CPU restatement of Tensile's dense hot path: the ORACLE.
 *
 * This is test infrastructure. Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark load it, and only to check or time the
 * device library against it. The shipped library never links or calls it.
 *
 * Every function restates the reference algorithm it names (file:line in the
 * reference repository, C# sources). Matrices are column-major with an
 * explicit leading dimension, like Tensile's MatrixView<T>. Complex values are
 * two adjacent doubles (re, im), the layout of System.Numerics.Complex.
 *
 * Arithmetic follows .NET's: no fused multiply-add outside the GEMM
 * micro-kernel (the file is built with -ffp-contract=off), Complex products,
 * quotients (Smith's method) and moduli (scaled hypot) restated from their
 * published algorithms, and System.Random's seeded Knuth subtractive
 * generator and Array.Sort's introspective sort restated for the norm
 * estimator.
 */
#ifndef TENSILE_ORACLE_H
#define TENSILE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } ocplx;

/* ---- GEMM ------------------------------------------------------------- */
/* C := beta*C + alpha*A*B (Gemm.cs:117-203 semantics: beta==0 overwrites,
 * k==0 or alpha==0 only scales C). Blocked, OpenMP-threaded, AVX2 FMA
 * micro-kernel when available. */
void or_dgemm(int m, int n, int k, double alpha, const double* a, int lda,
              const double* b, int ldb, double beta, double* c, int ldc);
/* Textbook triple loop (Reference.cs:10-28), single thread. */
void or_dgemm_ref(int m, int n, int k, double alpha, const double* a, int lda,
                  const double* b, int ldb, double beta, double* c, int ldc);
/* 4M complex product over four real GEMMs plus Combine
 * (ComplexKernels.cs:49-108, 187-240). */
void or_zgemm(int m, int n, int k, ocplx alpha, const ocplx* a, int lda,
              const ocplx* b, int ldb, ocplx beta, ocplx* c, int ldc);
/* The 3M product, kept only so a test can show the small-imaginary bound
 * rejects it (ComplexMultiplyTests.cs:199-210). */
void or_zgemm_3m(int m, int n, int k, const ocplx* a, int lda,
                 const ocplx* b, int ldb, ocplx* c, int ldc);
void or_set_threads(int threads);
int or_get_threads(void);

/* ---- norms and panel products ---------------------------------------- */
double or_dnorm1(int m, int n, const double* a, int lda);      /* Norms.cs:11-26 */
double or_dnorminf(int m, int n, const double* a, int lda);    /* Norms.cs:47-66 */
double or_dnormfro(int m, int n, const double* a, int lda);    /* Norms.cs:36-45 */
double or_znorm1(int m, int n, const ocplx* a, int lda);       /* ComplexKernels.cs:144-160 */
double or_znorminf(int m, int n, const ocplx* a, int lda);     /* ComplexKernels.cs:163-180 */
/* Y := A*X or A^T*X for square A (PanelProduct.cs:25-55). */
void or_dpanel(int trans, int n, int t, const double* a, int lda,
               const double* x, int ldx, double* y, int ldy);
/* Complex modulus / quotient / exponential as .NET computes them. */
double or_cabs(ocplx z);
ocplx or_cdiv(ocplx x, ocplx y);
ocplx or_cexp(ocplx z);

/* ---- LU ----------------------------------------------------------------
 * Blocked right-looking getrf (Lu.cs:230-309, BlockedLu.cs:45-101).
 * nb<1 selects Lu.DefaultBlockSizeFor (Lu.cs:221-222). pivots: min(m,n)
 * 0-based LAPACK-order interchanges. Returns the first exactly-zero pivot
 * column or -1. smallest/largest: range of |u_ii| (may be NULL). */
int or_dgetrf(int m, int n, double* a, int lda, int nb, int* pivots,
              double* smallest, double* largest);
int or_zgetrf(int m, int n, ocplx* a, int lda, int nb, int* pivots,
              double* smallest, double* largest);
/* trans: 0 = A X = B; 1 = A^T X = B (real) / A^H X = B (complex)
 * (Lu.cs:336-378, BlockedLu.cs:108-141). */
void or_dgetrs(int trans, int n, int nrhs, const double* lu, int ldlu,
               const int* pivots, double* b, int ldb);
void or_zgetrs(int trans, int n, int nrhs, const ocplx* lu, int ldlu,
               const int* pivots, ocplx* b, int ldb);
/* The six triangular solves of KernelEntry (Triangular.cs:26-206).
 * kind: 0 LowerUnit, 1 Upper, 2 Lower, 3 UpperT, 4 LowerT, 5 LowerUnitT. */
void or_dtrsv(int kind, int m, int n, const double* l, int ldl, double* b, int ldb);

/* ---- norm estimation (NormEstimate.cs:104-411) ------------------------ */
typedef struct { double value; int iterations; int products; } or_normest;
typedef void (*or_dapply)(void* ctx, int t, const double* x, double* y);
typedef void (*or_zapply)(void* ctx, int t, const ocplx* x, ocplx* y);
or_normest or_dnormest(int n, or_dapply apply, or_dapply adjoint, void* ctx,
                       int columns, int max_iterations, int seed);
or_normest or_znormest(int n, or_zapply apply, or_zapply adjoint, void* ctx,
                       int columns, int max_iterations, int seed);
/* ||A^p||_1 estimate through the dense power operator
 * (LinearOperators.cs:89-170, 194-293). */
or_normest or_dnormest_power(int n, const double* a, int lda, int power,
                             int columns, int max_iterations, int seed);
or_normest or_znormest_power(int n, const ocplx* a, int lda, int power,
                             int columns, int max_iterations, int seed);
/* rcond = 1/(||A||_1 ||A^-1||_1) from factors (NormEstimate.cs:432-466). */
double or_dgecon(int n, const double* lu, int ldlu, const int* pivots,
                 int singular_column, double anorm, int columns);
double or_zgecon(int n, const ocplx* lu, int ldlu, const int* pivots,
                 int singular_column, double anorm, int columns);

/* ---- .NET System.Random / Array.Sort restatements --------------------- */
typedef struct { int seed_array[56]; int inext, inextp; } or_random;
void or_random_init(or_random* r, int seed);
double or_random_sample(or_random* r);
int or_random_next(or_random* r, int max_value);
void or_sort_keys_items(int n, double* keys, int* items);

/* ---- exponentials ------------------------------------------------------ */
typedef struct { int degree; int squarings; } or_expm_diag;
typedef struct { int degree; int scaling; int applications; int early_exits; } or_expmv_diag;
/* Padé scaling and squaring (MatrixExponential.cs:181-505). Returns 0, or
 * -1 if the Padé denominator is exactly singular. E is n x n, lde >= n. */
int or_dexpm(int n, const double* a, int lda, double* e, int lde, or_expm_diag* diag);
int or_zexpm(int n, const ocplx* a, int lda, ocplx* e, int lde, or_expm_diag* diag);
/* Taylor action exp(tA)B (MatrixExponentialAction.cs:149-416). Returns 0,
 * or -2 if the selected cost is not finite. */
int or_dexpmv(int n, int nb, const double* a, int lda, const double* b, int ldb,
              double t, double* f, int ldf, or_expmv_diag* diag);
int or_zexpmv(int n, int nb, const ocplx* a, int lda, const ocplx* b, int ldb,
              double t, ocplx* f, int ldf, or_expmv_diag* diag);
/* Matrix-free action with a caller's operator and a 1-norm bound
 * (MatrixExponentialAction.cs:203-230): no shift, no sharp bound. */
int or_dexpmv_op(int n, or_dapply apply, void* ctx, int nb, const double* b, int ldb,
                 double t, double one_norm_bound, double* f, int ldf, or_expmv_diag* diag);
int or_zexpmv_op(int n, or_zapply apply, void* ctx, int nb, const ocplx* b, int ldb,
                 double t, double one_norm_bound, ocplx* f, int ldf, or_expmv_diag* diag);

/* Extended-precision (x87 long double, 64-bit mantissa) Taylor scaling and
 * squaring: the accuracy yardstick for error statements at small n
 * (the MatrixExponentialTests.cs:454-480 oracle, at higher precision). */
void or_dexpm_taylor_ld(int n, const double* a, int lda, double* e, int lde);
void or_zexpm_taylor_ld(int n, const ocplx* a, int lda, ocplx* e, int lde);
/* exp(tA)B in extended precision by scaled Taylor sums on the columns
 * (n^2 work per term and column), for error statements at large n. */
void or_zexpmv_taylor_ld(int n, const ocplx* a, int lda, double t, int nb, const ocplx* b, int ldb, ocplx* f,
                         int ldf);

#ifdef __cplusplus
}
#endif
#endif
