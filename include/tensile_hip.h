/*
 * tensile_hip.h -- the C ABI of the MI355X backend for Tensile's dense hot
 * path (xfmrexpert/CSharp_LinearAlgebra).
 *
 * The entry points are what Tensile's pinning seam, KernelEntry
 * (src/Tensile.Kernels/KernelEntry.cs:31-270), and the public assembly's
 * device-resident algorithms would bind through P/Invoke. Each declaration
 * names the reference interface it replaces.
 *
 * Conventions (Operand.cs:15-97, MatrixShape.cs:20-152):
 *   - matrices are column-major with int32 rows, columns and leading
 *     dimension ld >= max(1, rows); complex elements are two doubles (re, im),
 *     the layout of System.Numerics.Complex;
 *   - every pointer may be host memory (the C# caller's pinned span) or
 *     device memory. Host-memory calls stage through the device and return
 *     when the results are back (the seam is synchronous, KernelEntry.cs:19-24).
 *     Device-memory calls run on the context's stream; those that return a
 *     value through a host pointer (norms, estimates, LU info, diagnostics)
 *     wait for it, the rest return once the work is enqueued
 *     (tns_synchronize waits);
 *   - padding rows between rows and ld are never written (GemmTests.cs:107-129);
 *   - on any failure the destination is left untouched (invariant I5,
 *     ComplexKernels.cs:103-106) and the status says why; no C++ exception
 *     crosses this boundary. The exception is an in-place call on device
 *     memory (tns_dgetrf / tns_zgetrf on a device matrix): the factorization
 *     overwrites the matrix as it runs, so a failure reported at its end
 *     (TNS_EDEVICE) leaves the matrix partly factored.
 *
 * A context is not thread-safe; the C# Workspace keeps its lock around it
 * (Workspace.cs:136-139).
 */
#ifndef TENSILE_HIP_H
#define TENSILE_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
#define TNS_API __declspec(dllexport)
#else
#define TNS_API __attribute__((visibility("default")))
#endif

/* Status codes; the C# binding maps each to the reference's exception. */
#define TNS_OK 0
#define TNS_EINVAL 1      /* ArgumentException: operands disagree (KernelEntry.cs:250-269) */
#define TNS_ERANGE 2      /* ArgumentOutOfRangeException (MatrixExponentialAction.cs:215-220, 385-393) */
#define TNS_ESTATE 3      /* InvalidOperationException: singular or non-square solve (Lu.cs:338-342),
                             singular Pade denominator (MatrixExponential.cs:397-402) */
#define TNS_ENOMEM 4      /* AllocationLimitException (TensileLimits.cs:78-96) */
#define TNS_EDISPOSED 5   /* ObjectDisposedException (GemmDispatch.cs:107) */
#define TNS_EDEVICE 6     /* the HIP runtime failed */
#define TNS_ECALLBACK 7   /* a caller's operator callback returned non-zero (the binding rethrows its exception) */

typedef struct tns_context tns_context;
typedef struct { double re, im; } tns_complex;

/* Triangular solve kinds: the six KernelEntry.Solve* (KernelEntry.cs:98-173). */
enum {
    TNS_TRI_LOWER_UNIT = 0,   /* SolveLowerUnit */
    TNS_TRI_UPPER = 1,        /* SolveUpper */
    TNS_TRI_LOWER = 2,        /* SolveLower */
    TNS_TRI_UPPER_T = 3,      /* SolveUpperTransposed */
    TNS_TRI_LOWER_T = 4,      /* SolveLowerTransposed */
    TNS_TRI_LOWER_UNIT_T = 5  /* SolveLowerUnitTransposed */
};

/* LuFactorization's diagnostics (Lu.cs:17-67). */
typedef struct {
    int singular_column;     /* first exactly-zero pivot column, or -1 */
    double smallest_pivot;   /* min |u_ii| */
    double largest_pivot;    /* max |u_ii| */
} tns_lu_info;

/* NormEstimateResult (NormEstimate.cs:9). */
typedef struct {
    double value;
    int iterations;
    int products;
} tns_normest;

/* ExpmDiagnostics (MatrixExponential.cs:575-582). */
typedef struct {
    int degree;
    int squarings;
} tns_expm_diag;

/* ExpmvDiagnostics (MatrixExponentialAction.cs:498-511). */
typedef struct {
    int degree;
    int scaling;
    int applications;
    int early_exits;
} tns_expmv_diag;

/* ---- context: the device side of Tensile's Workspace (Workspace.cs:36-268) */
TNS_API int tns_context_create(int device, tns_context** out);
TNS_API int tns_context_destroy(tns_context* ctx);
/* Message for the last failed call on this context (owned by the context). */
TNS_API const char* tns_last_error(const tns_context* ctx);
/* Ceiling on the device memory the context's pool may hold, the analogue of
 * TensileLimits.MaxElements (TensileLimits.cs:27-147). */
TNS_API int tns_context_set_limit(tns_context* ctx, size_t bytes);
/* Bytes of free device memory the pool keeps between calls (default: no
 * cap below the limit); after each call, free blocks beyond it go back to
 * the device, largest first. The retention cap of ComplexScratch's retained
 * storage (ComplexScratch.cs:12-60, 77-122). */
TNS_API int tns_context_set_retention(tns_context* ctx, size_t bytes);
/* The pool's bytes held (lent and free), its limit and its retention cap. */
TNS_API int tns_context_memory(tns_context* ctx, size_t* held, size_t* limit, size_t* retention);
/* The context's HIP stream (hipStream_t), for callers chaining device work. */
TNS_API int tns_context_stream(tns_context* ctx, void** stream);
TNS_API int tns_synchronize(tns_context* ctx);
/* Library build tag. */
TNS_API const char* tns_version(void);
/* Tile configuration of the GEMM kernels, for the library's autotuner:
 * -1 restores the size heuristic; otherwise an index below
 * tns_gemm_config_count(complex). Process-wide. */
TNS_API int tns_gemm_set_config(int real_config, int complex_config);
TNS_API int tns_gemm_config_count(int complex_);

/* ---- GEMM: KernelEntry.Multiply (KernelEntry.cs:34-56) ------------------
 * C := beta*C + alpha*A*B; A m x k, B k x n, C m x n. beta == 0 overwrites
 * (a NaN in C does not survive); k == 0 or alpha == 0 only scales C; an empty
 * C returns at once (KernelEntry.cs:41-44, Gemm.cs:117-203). */
TNS_API int tns_dgemm(tns_context* ctx, int m, int n, int k, double alpha, const double* a, int lda,
                      const double* b, int ldb, double beta, double* c, int ldc);
/* Complex product by the 4M method (ComplexKernels.Multiply,
 * ComplexKernels.cs:49-108), with Combine's special cases for alpha = 1,
 * real alpha, beta = 0 and beta = 1 (ComplexKernels.cs:213-240). */
TNS_API int tns_zgemm(tns_context* ctx, int m, int n, int k, tns_complex alpha, const tns_complex* a, int lda,
                      const tns_complex* b, int ldb, tns_complex beta, tns_complex* c, int ldc);

/* ---- LU: KernelEntry.FactorLu (KernelEntry.cs:59-67) / Lu.Factor
 * (Lu.cs:230-309); complex: BlockedLu.Factor (BlockedLu.cs:45-101) ---------
 * Factors A in place as P*A = L*U. nb < 1 selects Lu.DefaultBlockSizeFor
 * (32 below min(m,n) = 1024, else 64; Lu.cs:221-222). pivots receives
 * min(m,n) 0-based LAPACK-order interchanges. */
TNS_API int tns_dgetrf(tns_context* ctx, int m, int n, double* a, int lda, int nb, int* pivots,
                       tns_lu_info* info);
TNS_API int tns_zgetrf(tns_context* ctx, int m, int n, tns_complex* a, int lda, int nb, int* pivots,
                       tns_lu_info* info);

/* ---- LU that allocates nothing. No counterpart in the reference: it is
 * tns_dgetrf (the same kernels, pivots and factors) for callers that cannot
 * call hipMalloc (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation). A, pivots (min(m,n) ints) and info (one
 * int: the first zero-pivot column, -1, or -2 if a panel's cross-workgroup
 * exchange timed out) are all device memory; the call only enqueues work on
 * the context's stream and returns. Its scratch (1 MiB per device, held until
 * the process ends) is shared by every call on a device, so calls that
 * overlap on one device race. nb <= 64 (< 1: the default); a size whose
 * scratch would exceed 1 MiB (about min(m,n) > 16384 at the default block
 * size) is refused with TNS_ERANGE. */
TNS_API int tns_dgetrf_noalloc(tns_context* ctx, int m, int n, double* a, int lda, int nb, int* pivots, int* info);

/* ---- the same with the phase split: KernelEntry.FactorLu's optional
 * LuPhaseTimings (KernelEntry.cs:59-67, Lu.cs:92-164, hooked in at
 * Lu.cs:251-305). The device time of the panel factorizations, the row
 * interchanges either side of the panels, the U12 triangular solves and the
 * trailing GEMMs is ADDED to *timings (pooling over runs, as
 * LuPhaseTimings.Add); total_ms spans the whole factorization, so
 * total_ms - (panel + swaps + trsm + gemm) is the time no phase accounts
 * for. block_steps counts the panels. Events are recorded between phases,
 * so the call waits for the factorization. timings may be null. */
typedef struct {
    double panel_ms;
    double swaps_ms;
    double trsm_ms;
    double gemm_ms;
    double total_ms;
    int block_steps;
} tns_lu_phases;
TNS_API int tns_dgetrf_timed(tns_context* ctx, int m, int n, double* a, int lda, int nb, int* pivots,
                             tns_lu_info* info, tns_lu_phases* timings);
TNS_API int tns_zgetrf_timed(tns_context* ctx, int m, int n, tns_complex* a, int lda, int nb, int* pivots,
                             tns_lu_info* info, tns_lu_phases* timings);

/* ---- LU solves: KernelEntry.SolveLu / SolveLuTransposed (KernelEntry.cs:70-95),
 * Lu.Solve / SolveTransposed (Lu.cs:336-378); complex BlockedLu.Solve /
 * SolveAdjoint (BlockedLu.cs:108-141) -------------------------------------
 * trans 0: A X = B; 1: A^T X = B (real) or A^H X = B (complex). B n x nrhs is
 * overwritten. singular_column >= 0 refuses with TNS_ESTATE. */
TNS_API int tns_dgetrs(tns_context* ctx, int trans, int n, int nrhs, const double* lu, int ldlu,
                       const int* pivots, int singular_column, double* b, int ldb);
TNS_API int tns_zgetrs(tns_context* ctx, int trans, int n, int nrhs, const tns_complex* lu, int ldlu,
                       const int* pivots, int singular_column, tns_complex* b, int ldb);

/* ---- triangular solves: KernelEntry.Solve* (KernelEntry.cs:98-173),
 * Triangular.cs:26-206. A m x m, B m x n overwritten; kind TNS_TRI_*. */
TNS_API int tns_dtrsm(tns_context* ctx, int kind, int m, int n, const double* a, int lda, double* b, int ldb);

/* ---- panel products: KernelEntry.MultiplyPanel / MultiplyPanelTransposed
 * (KernelEntry.cs:176-201), PanelProduct.cs:25-55. Y := A X or A^T X for a
 * square A and an n x t panel X. Complex: A X or A^H X, the product of
 * ComplexDenseOperator (LinearOperators.cs:232-283). */
TNS_API int tns_dpanel(tns_context* ctx, int trans, int n, int t, const double* a, int lda, const double* x,
                       int ldx, double* y, int ldy);
TNS_API int tns_zpanel(tns_context* ctx, int adjoint, int n, int t, const tns_complex* a, int lda,
                       const tns_complex* x, int ldx, tns_complex* y, int ldy);

/* ---- norms: KernelEntry.OneNorm / InfinityNorm / FrobeniusNorm
 * (KernelEntry.cs:204-234), Norms.cs:11-66; complex ComplexKernels.OneNorm /
 * InfinityNorm (ComplexKernels.cs:144-180). kind '1', 'I' or 'F' (complex:
 * '1' or 'I'). An empty operand gives 0. */
TNS_API int tns_dnorm(tns_context* ctx, char kind, int m, int n, const double* a, int lda, double* out);
TNS_API int tns_znorm(tns_context* ctx, char kind, int m, int n, const tns_complex* a, int lda, double* out);

/* ---- 1-norm estimation: NormEstimate.Of over DenseMatrixOperator(a, power)
 * (NormEstimate.cs:69-274, LinearOperators.cs:89-170); MatrixOperations.
 * EstimateOneNorm. Complex: ComplexDenseOperator (LinearOperators.cs:194-293).
 * columns/max_iterations/seed as NormEstimate.DefaultColumns (2),
 * DefaultMaxIterations (5), DefaultSeed (20260915). */
TNS_API int tns_dnormest_power(tns_context* ctx, int n, const double* a, int lda, int power, int columns,
                               int max_iterations, int seed, tns_normest* out);
TNS_API int tns_znormest_power(tns_context* ctx, int n, const tns_complex* a, int lda, int power, int columns,
                               int max_iterations, int seed, tns_normest* out);

/* ---- condition: LuDecomposition.ReciprocalCondition (MatrixOperations.cs:106-114),
 * Condition.ReciprocalOne (NormEstimate.cs:432-466): 1/(||A||_1 ||A^-1||_1)
 * with ||A^-1||_1 estimated through the LU inverse operator
 * (LinearOperators.cs:303-350). anorm is ||A||_1 of the matrix before it was
 * factored. 0 for a singular factorization or a zero norm. */
TNS_API int tns_dgecon(tns_context* ctx, int n, const double* lu, int ldlu, const int* pivots,
                       int singular_column, double anorm, int columns, double* rcond);
TNS_API int tns_zgecon(tns_context* ctx, int n, const tns_complex* lu, int ldlu, const int* pivots,
                       int singular_column, double anorm, int columns, double* rcond);

/* ---- matrix exponential: MatrixExponential.Expm (MatrixExponential.cs:143-270,
 * MatrixExponential.Complex.cs:35-53). E := exp(A), Al-Mohy & Higham 2009
 * scaling and squaring with Pade degrees 3/5/7/9/13. TNS_ESTATE if the Pade
 * denominator is exactly singular. */
TNS_API int tns_dexpm(tns_context* ctx, int n, const double* a, int lda, double* e, int lde,
                      tns_expm_diag* diag);
TNS_API int tns_zexpm(tns_context* ctx, int n, const tns_complex* a, int lda, tns_complex* e, int lde,
                      tns_expm_diag* diag);
/* The same with its device-time split, from HIP events between the phases
 * (MatrixExponential.cs:181-270): the powers A^2, A^4, A^6 and their norms;
 * the selection of degree and scaling (normest1 of the higher powers and the
 * ell corrections); the Pade polynomial products and combinations; the LU
 * of the denominator; its solve; the squarings. The diagnostics tool's
 * phase report (tools/Tensile.Diagnostics/Program.cs:196-284) for Expm has no
 * other counterpart in the reference. */
typedef struct {
    double powers_ms;
    double select_ms;
    double pade_ms;
    double lu_ms;
    double solve_ms;
    double squaring_ms;
    double total_ms;
} tns_expm_phases;
TNS_API int tns_dexpm_timed(tns_context* ctx, int n, const double* a, int lda, double* e, int lde,
                            tns_expm_diag* diag, tns_expm_phases* phases);
TNS_API int tns_zexpm_timed(tns_context* ctx, int n, const tns_complex* a, int lda, tns_complex* e, int lde,
                            tns_expm_diag* diag, tns_expm_phases* phases);

/* ---- exponential action: MatrixExponentialAction.Expmv
 * (MatrixExponentialAction.cs:102-178, MatrixExponentialAction.Complex.cs:34-63).
 * F := exp(tA) B, B n x nb; Al-Mohy & Higham 2011 truncated Taylor with the
 * trace shift and the sharp (m, s) selection. TNS_ERANGE if the selected
 * cost is not a finite computation. */
TNS_API int tns_dexpmv(tns_context* ctx, int n, int nb, const double* a, int lda, const double* b, int ldb,
                       double t, double* f, int ldf, tns_expmv_diag* diag);
TNS_API int tns_zexpmv(tns_context* ctx, int n, int nb, const tns_complex* a, int lda, const tns_complex* b,
                       int ldb, double t, tns_complex* f, int ldf, tns_expmv_diag* diag);

/* ---- matrix-free action over an ILinearOperator<T> (LinearOperators.cs:37-57,
 * MatrixExponentialAction.cs:203-230, MatrixExponentialAction.Complex.cs:85-111).
 * The operator is the caller's callback: Y := Op * X for an n x t panel,
 * returning 0 on success. With device_operator == 0 the panels are host
 * memory (a C# ILinearOperator); with 1 they are device memory on the
 * context's stream (a device-resident operator). No shift, no sharp bound:
 * the caller's one_norm_bound >= ||A||_1 drives the selection; it must be
 * finite and not negative (TNS_ERANGE). */
typedef int (*tns_dapply_fn)(void* user, int n, int t, const double* x, int ldx, double* y, int ldy);
typedef int (*tns_zapply_fn)(void* user, int n, int t, const tns_complex* x, int ldx, tns_complex* y, int ldy);
TNS_API int tns_dexpmv_op(tns_context* ctx, int n, tns_dapply_fn apply, void* user, int device_operator, int nb,
                          const double* b, int ldb, double t, double one_norm_bound, double* f, int ldf,
                          tns_expmv_diag* diag);
TNS_API int tns_zexpmv_op(tns_context* ctx, int n, tns_zapply_fn apply, void* user, int device_operator, int nb,
                          const tns_complex* b, int ldb, double t, double one_norm_bound, tns_complex* f, int ldf,
                          tns_expmv_diag* diag);

/* ---- mixed precision: A X = B by an FP32 LU and FP64 iterative refinement.
 * No counterpart in the reference: it stands beside the FP64 solve
 * KernelEntry.FactorLu + SolveLu (KernelEntry.cs:59-95, Lu.cs:189-378),
 * which it falls back to. The low-precision factorisation is of a random
 * butterfly transformation U^T (R A C) V of A (depth 2, after power-of-two
 * row and column equilibration R, C when the row or column scales spread
 * over more than 2^8), in FP32 with each block step's pivots
 * chosen inside its 64 x 64 diagonal block. precision 32: its products on
 * the FP32 matrix cores; 16: its trailing updates on the BF16 matrix cores
 * (FP32 accumulation, operands split into two bfloat16 parts, three
 * products each).
 * Refinement (residuals and updates in FP64; after the first solve, FGMRES
 * restart cycles of at most 30 steps, right-preconditioned by the
 * low-precision solve) runs until the backward error
 * max_j ||b_j - A x_j||inf / (||A||inf ||x_j||inf + ||b_j||inf) <= tol, at
 * most max_iter low-precision solves in all. A cycle that fails to halve the
 * error, residuals that shrink too slowly for the remaining steps to cost
 * less than about 3.5 factorizations, or an exactly zero low-precision pivot
 * switch to the FP64 LU with partial pivoting (fell_back = 1). A and B are
 * not modified. TNS_ESTATE if the FP64 LU also meets a zero pivot. */
typedef struct {
    int precision;          /* 32 or 16, as requested */
    int iterations;         /* low-precision solves taken */
    int fell_back;          /* 1: the result comes from the FP64 LU */
    double backward_error;  /* of the returned X */
    double factor_ms;       /* low-precision factorisation, host wall time */
    double refine_ms;       /* refinement */
    double fallback_ms;     /* FP64 LU + solve, when taken */
    double first_error;     /* backward error after the first (plain low-precision) solve */
} tns_mixed_info;
TNS_API int tns_dgesv_mixed(tns_context* ctx, int n, int nrhs, const double* a, int lda, const double* b, int ldb,
                            double* x, int ldx, int precision, double tol, int max_iter, tns_mixed_info* info);
/* Its low-precision parts, exported so they can be checked on their own
 * (no counterparts in the reference). tns_sgemm: C := beta*C + alpha*A*B
 * in FP32 with the products on the FP32 (precision 32) or BF16 (16)
 * matrix cores, or on the BF16 ones with each operand split into two
 * bfloat16 parts and three products per block (48: about 2^-16 relative
 * error; the trailing updates of the mixed solve's precision 16). The FP32
 * LU with partial pivoting (pivots and diagnostics as tns_dgetrf) and the
 * solve against its factors take precision 32 or 16. */
TNS_API int tns_sgemm(tns_context* ctx, int precision, int m, int n, int k, float alpha, const float* a, int lda,
                      const float* b, int ldb, float beta, float* c, int ldc);
TNS_API int tns_sgetrf(tns_context* ctx, int precision, int n, float* a, int lda, int* pivots, int* singular_column);
TNS_API int tns_sgetrs(tns_context* ctx, int precision, int n, int nrhs, const float* lu, int ldlu, const int* pivots,
                       float* b, int ldb);

/* ---- gradients through exp(tA)B: the adjoint of tns_{d,z}expmv. No
 * counterpart in the reference; it differentiates MatrixExponentialAction's
 * dense path (MatrixExponentialAction.cs:102-178, the complex trace shift of
 * MatrixExponentialAction.Complex.cs:53-63). Given G = dL/dF for F = exp(tA)B
 * (n x nb), writes dL/dA (n x n) to ga and dL/dB (n x nb) to gb: the
 * gradients of the computed F, with the same (m, s) and Taylor truncation
 * (for complex data, dL = Re tr(ga^H dA) + Re tr(gb^H dB)). The Taylor
 * stages are recomputed from per-stage checkpoints, not stored.
 * Mathematically dL/dB = exp(tA^H) G and dL/dA = t L(tA^H, G B^H), L the
 * Frechet derivative of exp. diag reports the forward's parameters. */
TNS_API int tns_dexpmv_grad(tns_context* ctx, int n, int nb, const double* a, int lda, const double* b, int ldb,
                            double t, const double* g, int ldg, double* ga, int ldga, double* gb, int ldgb,
                            tns_expmv_diag* diag);
TNS_API int tns_zexpmv_grad(tns_context* ctx, int n, int nb, const tns_complex* a, int lda, const tns_complex* b,
                            int ldb, double t, const tns_complex* g, int ldg, tns_complex* ga, int ldga,
                            tns_complex* gb, int ldgb, tns_expmv_diag* diag);

/* ---- batched frequency sweep: many complex systems Z_i X_i = B_i of uneven
 * order, factored and solved in one launch of a captured hipGraph. No single
 * counterpart in the reference: per system it is BlockedLu.Factor + Solve
 * (BlockedLu.cs:45-206) with ||Z_i||_1 taken first (LuDecomposition's
 * OneNormOfA, LuDecomposition.cs:37-80), and Condition.ReciprocalOne
 * (NormEstimate.cs:432-466) for rcond. All matrices must be device memory;
 * the batch records the pointers, so they must stay valid until
 * tns_zbatch_destroy. A batch belongs to its context: tns_context_destroy
 * releases the device memory of every batch still alive, whose handle then
 * refuses calls with TNS_EDISPOSED until tns_zbatch_destroy frees it. tune: 0 defaults, 1 the context's remembered choice for
 * this batch shape (tuning on scratch copies if there is none), 2 tune
 * again. tns_zbatch_solve overwrites every A_i with its LU factors and every
 * B_i with X_i; refill them before solving again. singular_columns (host,
 * count entries, may be null) receives each system's first zero pivot or
 * -1. tns_zbatch_rcond (after a solve) writes count reciprocal condition
 * numbers (host), estimating with `columns` probe columns (2 in the
 * reference). */
typedef struct tns_batch tns_batch;
TNS_API int tns_zbatch_create(tns_context* ctx, int count, const int* n, tns_complex* const* a, const int* lda,
                              tns_complex* const* b, const int* ldb, int nrhs, int tune, tns_batch** out);
TNS_API int tns_zbatch_solve(tns_batch* batch, int* singular_columns);
TNS_API int tns_zbatch_rcond(tns_batch* batch, int columns, double* rcond);
TNS_API int tns_zbatch_pivots(tns_batch* batch, int index, int* pivots);
/* What the autotuner chose: the LU panel width (32 or 64), the tile of the
 * grouped 4M trailing update (1: 64 x 64, 2: 64 x 32), and the best graph
 * time it measured (0 if it did not run). No counterpart in the reference. */
TNS_API int tns_zbatch_info(tns_batch* batch, int* block, int* gemm_config, double* tuned_ms);
TNS_API int tns_zbatch_destroy(tns_batch* batch);

#ifdef __cplusplus
}
#endif
#endif
