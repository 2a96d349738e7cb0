/*
 * Oracle algorithms: LU, solves, norm estimation, condition estimation,
 * exponential and exponential action, each instantiated for double and for
 * complex from oracle_generic.inc. Test infrastructure only (see oracle.h).
 */
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

/* ---- double: DoubleKernels (ElementKernels.cs:125-170) ---------------- */
static inline double add_d(double a, double b) { return a + b; }
static inline double sub_d(double a, double b) { return a - b; }
static inline double mul_d(double a, double b) { return a * b; }
static inline double div_d(double a, double b) { return a / b; }
static inline double neg_d(double a) { return -a; }
static inline int iszero_d(double a) { return a == 0.0; }
static inline int eq_d(double a, double b) { return a == b; }
static inline double mag_d(double a) { return fabs(a); }
static inline double pivmag_d(double a) { return fabs(a); }
static inline double conj_d(double a) { return a; }
/* Anything not at least zero, NaN included, is negative (ElementKernels.cs:147-148). */
static inline double sign_d(double a) { return a >= 0.0 ? 1.0 : -1.0; }
static inline double scale_d(double a, double f) { return a * f; }
static inline double exp_d(double a) { return exp(a); }
static inline double from_d_d(double a) { return a; }
static double onenorm_d(int m, int n, const double* a, int lda) { return or_dnorm1(m, n, a, lda); }
static double infnorm_d(int m, int n, const double* a, int lda) { return or_dnorminf(m, n, a, lda); }
static void gemm_d(int m, int n, int k, double alpha, const double* a, int lda, const double* b, int ldb,
                   double beta, double* c, int ldc)
{
    or_dgemm(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
static double trace_mean_d(int n, const double* a, int lda)
{
    double total = 0.0;
    for (int i = 0; i < n; i++) total += a[(size_t)i * lda + i];
    return total / n;
}
/* DenseMatrixOperator.Repeat (LinearOperators.cs:124-151): y = A^p x or (A^T)^p x. */
static void panel_power_d(int n, const double* a, int lda, int power, int adjoint, int t,
                          const double* x, double* y)
{
    or_dpanel(adjoint, n, t, a, lda, x, n, y, n);
    if (power == 1) return;
    double* staging = (double*)malloc(sizeof(double) * (size_t)n * t);
    for (int step = 1; step < power; step++) {
        memcpy(staging, y, sizeof(double) * (size_t)n * t);
        or_dpanel(adjoint, n, t, a, lda, staging, n, y, n);
    }
    free(staging);
}

/* ---- complex: ComplexKernels (ComplexKernels.cs:40-180) --------------- */
static inline ocplx mk(double re, double im) { ocplx z = { re, im }; return z; }
static inline ocplx add_z(ocplx a, ocplx b) { return mk(a.re + b.re, a.im + b.im); }
static inline ocplx sub_z(ocplx a, ocplx b) { return mk(a.re - b.re, a.im - b.im); }
static inline ocplx mul_z(ocplx a, ocplx b) { return mk(a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im); }
static inline ocplx div_z(ocplx a, ocplx b) { return or_cdiv(a, b); }
static inline ocplx neg_z(ocplx a) { return mk(-a.re, -a.im); }
static inline int iszero_z(ocplx a) { return a.re == 0.0 && a.im == 0.0; }
static inline int eq_z(ocplx a, ocplx b) { return a.re == b.re && a.im == b.im; }
static inline double mag_z(ocplx a) { return or_cabs(a); }
/* |Re| + |Im|, LAPACK's cabs1 (ElementKernels.cs:61-69). */
static inline double pivmag_z(ocplx a) { return fabs(a.re) + fabs(a.im); }
static inline ocplx conj_z(ocplx a) { return mk(a.re, -a.im); }
static inline ocplx sign_z(ocplx a)
{
    if (iszero_z(a)) return mk(1.0, 0.0);
    double m = or_cabs(a);
    return mk(a.re / m, a.im / m);
}
static inline ocplx scale_z(ocplx a, double f) { return mk(a.re * f, a.im * f); }
static inline ocplx exp_z(ocplx a) { return or_cexp(a); }
static inline ocplx from_d_z(double a) { return mk(a, 0.0); }
static double onenorm_z(int m, int n, const ocplx* a, int lda) { return or_znorm1(m, n, a, lda); }
static double infnorm_z(int m, int n, const ocplx* a, int lda) { return or_znorminf(m, n, a, lda); }
static void gemm_z(int m, int n, int k, ocplx alpha, const ocplx* a, int lda, const ocplx* b, int ldb,
                   ocplx beta, ocplx* c, int ldc)
{
    or_zgemm(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
static ocplx trace_mean_z(int n, const ocplx* a, int lda)
{
    ocplx total = mk(0.0, 0.0);
    for (int i = 0; i < n; i++) total = add_z(total, a[(size_t)i * lda + i]);
    return mk(total.re / n, total.im / n);
}
/* ComplexDenseOperator.Repeat (LinearOperators.cs:232-283): two real panel
 * products over the stacked [Xr | Xi], recombined; the adjoint uses A^T on
 * the split parts and recombines with the conjugate's signs. */
static void panel_power_z(int n, const ocplx* a, int lda, int power, int adjoint, int t,
                          const ocplx* x, ocplx* y)
{
    size_t nn = (size_t)n * n, nt2 = (size_t)n * 2 * t;
    double* ar = (double*)malloc(sizeof(double) * nn);
    double* ai = (double*)malloc(sizeof(double) * nn);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            ar[(size_t)j * n + i] = a[(size_t)j * lda + i].re;
            ai[(size_t)j * n + i] = a[(size_t)j * lda + i].im;
        }
    double* st = (double*)malloc(sizeof(double) * nt2);
    double* p = (double*)malloc(sizeof(double) * nt2);
    double* q = (double*)malloc(sizeof(double) * nt2);
    for (int step = 0; step < power; step++) {
        const ocplx* src = step == 0 ? x : y;
        for (int j = 0; j < t; j++)
            for (int i = 0; i < n; i++) {
                st[(size_t)j * n + i] = src[(size_t)j * n + i].re;
                st[(size_t)(t + j) * n + i] = src[(size_t)j * n + i].im;
            }
        or_dpanel(adjoint, n, 2 * t, ar, n, st, n, p, n);
        or_dpanel(adjoint, n, 2 * t, ai, n, st, n, q, n);
        for (int j = 0; j < t; j++)
            for (int i = 0; i < n; i++) {
                double pr = p[(size_t)j * n + i], pi = p[(size_t)(t + j) * n + i];
                double qr = q[(size_t)j * n + i], qi = q[(size_t)(t + j) * n + i];
                y[(size_t)j * n + i] = adjoint ? mk(pr + qi, pi - qr) : mk(pr - qi, pi + qr);
            }
    }
    free(ar); free(ai); free(st); free(p); free(q);
}

/* ---- instantiate ------------------------------------------------------ */
#define T double
#define OP(x) CAT(x, _d)
#define DISCRETE_SIGNS 1
#include "oracle_generic.inc"
#undef T
#undef OP
#undef DISCRETE_SIGNS

#define T ocplx
#define OP(x) CAT(x, _z)
#define DISCRETE_SIGNS 0
#include "oracle_generic.inc"
#undef T
#undef OP
#undef DISCRETE_SIGNS

/* ---- exported names ---------------------------------------------------- */
int or_dgetrf(int m, int n, double* a, int lda, int nb, int* p, double* s, double* l) { return getrf_d(m, n, a, lda, nb, p, s, l); }
int or_zgetrf(int m, int n, ocplx* a, int lda, int nb, int* p, double* s, double* l) { return getrf_z(m, n, a, lda, nb, p, s, l); }
void or_dgetrs(int tr, int n, int r, const double* lu, int ld, const int* p, double* b, int ldb) { getrs_d(tr, n, r, lu, ld, p, b, ldb); }
void or_zgetrs(int tr, int n, int r, const ocplx* lu, int ld, const int* p, ocplx* b, int ldb) { getrs_z(tr, n, r, lu, ld, p, b, ldb); }

or_normest or_dnormest(int n, or_dapply ap, or_dapply adj, void* ctx, int c, int it, int seed) { return normest_d(n, ap, adj, ctx, c, it, seed); }
or_normest or_znormest(int n, or_zapply ap, or_zapply adj, void* ctx, int c, int it, int seed) { return normest_z(n, ap, adj, ctx, c, it, seed); }
or_normest or_dnormest_power(int n, const double* a, int lda, int p, int c, int it, int seed) { return normest_power_d(n, a, lda, p, c, it, seed); }
or_normest or_znormest_power(int n, const ocplx* a, int lda, int p, int c, int it, int seed) { return normest_power_z(n, a, lda, p, c, it, seed); }
double or_dgecon(int n, const double* lu, int ld, const int* p, int sc, double an, int c) { return gecon_d(n, lu, ld, p, sc, an, c); }
double or_zgecon(int n, const ocplx* lu, int ld, const int* p, int sc, double an, int c) { return gecon_z(n, lu, ld, p, sc, an, c); }

int or_dexpm(int n, const double* a, int lda, double* e, int lde, or_expm_diag* d) { return expm_d(n, a, lda, e, lde, d); }
int or_zexpm(int n, const ocplx* a, int lda, ocplx* e, int lde, or_expm_diag* d) { return expm_z(n, a, lda, e, lde, d); }
int or_dexpmv(int n, int nb, const double* a, int lda, const double* b, int ldb, double t, double* f, int ldf, or_expmv_diag* d)
{
    return expmv_d(n, nb, a, lda, b, ldb, t, f, ldf, d);
}
int or_zexpmv(int n, int nb, const ocplx* a, int lda, const ocplx* b, int ldb, double t, ocplx* f, int ldf, or_expmv_diag* d)
{
    return expmv_z(n, nb, a, lda, b, ldb, t, f, ldf, d);
}
int or_dexpmv_op(int n, or_dapply ap, void* ctx, int nb, const double* b, int ldb, double t, double bound, double* f, int ldf, or_expmv_diag* d)
{
    return expmv_op_d(n, ap, ctx, nb, b, ldb, t, bound, f, ldf, d);
}
int or_zexpmv_op(int n, or_zapply ap, void* ctx, int nb, const ocplx* b, int ldb, double t, double bound, ocplx* f, int ldf, or_expmv_diag* d)
{
    return expmv_op_z(n, ap, ctx, nb, b, ldb, t, bound, f, ldf, d);
}

/* The six KernelEntry triangular solves, restated from Triangular.cs:26-206. */
void or_dtrsv(int kind, int m, int n, const double* l, int ldl, double* b, int ldb)
{
    for (int j = 0; j < n; j++) {
        double* x = b + (size_t)j * ldb;
        switch (kind) {
        case 0: /* LowerUnit */
            for (int p = 0; p < m - 1; p++) {
                double v = x[p];
                if (v == 0.0) continue;
                for (int i = p + 1; i < m; i++) x[i] = x[i] + (-v) * l[(size_t)p * ldl + i];
            }
            break;
        case 1: /* Upper */
            for (int p = m - 1; p >= 0; p--) {
                x[p] /= l[(size_t)p * ldl + p];
                double v = x[p];
                if (v == 0.0 || p == 0) continue;
                for (int i = 0; i < p; i++) x[i] = x[i] + (-v) * l[(size_t)p * ldl + i];
            }
            break;
        case 2: /* Lower, explicit diagonal */
            for (int p = 0; p < m; p++) {
                x[p] /= l[(size_t)p * ldl + p];
                double v = x[p];
                if (v == 0.0 || p + 1 >= m) continue;
                for (int i = p + 1; i < m; i++) x[i] = x[i] + (-v) * l[(size_t)p * ldl + i];
            }
            break;
        case 3: /* Upper transposed: forward, dots down columns of U */
            for (int i = 0; i < m; i++) {
                double s = 0.0;
                for (int k = 0; k < i; k++) s += l[(size_t)i * ldl + k] * x[k];
                x[i] = (x[i] - s) / l[(size_t)i * ldl + i];
            }
            break;
        case 4: /* Lower transposed */
            for (int i = m - 1; i >= 0; i--) {
                double s = 0.0;
                for (int k = i + 1; k < m; k++) s += l[(size_t)i * ldl + k] * x[k];
                x[i] = (x[i] - s) / l[(size_t)i * ldl + i];
            }
            break;
        default: /* Lower unit transposed */
            for (int i = m - 1; i >= 0; i--) {
                double s = 0.0;
                for (int k = i + 1; k < m; k++) s += l[(size_t)i * ldl + k] * x[k];
                x[i] -= s;
            }
            break;
        }
    }
}

/* ---- extended-precision Taylor scaling and squaring --------------------
 * The reference's test oracle (MatrixExponentialTests.cs:454-480): scale to
 * ||A||_1 <= 1/32, 40 Taylor terms, square back; here every operation is in
 * x87 extended precision (64-bit mantissa), so its own rounding error sits
 * about three orders of magnitude below double's. */
typedef struct { long double re, im; } ldc;

static void taylor_ld(int n, const ldc* a, ldc* e)
{
    size_t nn = (size_t)n * n;
    long double norm = 0.0L;
    for (int j = 0; j < n; j++) {
        long double s = 0.0L;
        for (int i = 0; i < n; i++) s += sqrtl(a[(size_t)j * n + i].re * a[(size_t)j * n + i].re + a[(size_t)j * n + i].im * a[(size_t)j * n + i].im);
        if (s > norm) norm = s;
    }
    int s = 0;
    while (norm > 0.03125L) { norm *= 0.5L; s++; }
    long double f = ldexpl(1.0L, -s);
    ldc* b = (ldc*)malloc(sizeof(ldc) * nn);
    ldc* term = (ldc*)calloc(nn, sizeof(ldc));
    ldc* next = (ldc*)calloc(nn, sizeof(ldc));
    for (size_t i = 0; i < nn; i++) { b[i].re = a[i].re * f; b[i].im = a[i].im * f; }
    for (size_t i = 0; i < nn; i++) { e[i].re = 0; e[i].im = 0; }
    for (int i = 0; i < n; i++) { e[(size_t)i * n + i].re = 1.0L; term[(size_t)i * n + i].re = 1.0L; }
    for (int k = 1; k <= 40; k++) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                long double sr = 0.0L, si = 0.0L;
                for (int p = 0; p < n; p++) {
                    ldc x = term[(size_t)p * n + i], y = b[(size_t)j * n + p];
                    sr += x.re * y.re - x.im * y.im;
                    si += x.re * y.im + x.im * y.re;
                }
                next[(size_t)j * n + i].re = sr / k;
                next[(size_t)j * n + i].im = si / k;
            }
        ldc* tmp = term; term = next; next = tmp;
        for (size_t i = 0; i < nn; i++) { e[i].re += term[i].re; e[i].im += term[i].im; }
    }
    for (int q = 0; q < s; q++) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < n; i++) {
                long double sr = 0.0L, si = 0.0L;
                for (int p = 0; p < n; p++) {
                    ldc x = e[(size_t)p * n + i], y = e[(size_t)j * n + p];
                    sr += x.re * y.re - x.im * y.im;
                    si += x.re * y.im + x.im * y.re;
                }
                next[(size_t)j * n + i].re = sr;
                next[(size_t)j * n + i].im = si;
            }
        memcpy(e, next, sizeof(ldc) * nn);
    }
    free(b); free(term); free(next);
}

/* exp(tA)B in extended precision, column by column: tA/s with
 * s = ceil(||tA||_1) (so ||tA/s||_1 <= 1), each of the s steps a Taylor sum
 * of x <- sum_k (tA/s)^k x / k! until a term is below 1e-22 of the sum in
 * max-norm. The yardstick for the error of exp(A)V and exp(tA)B at orders
 * where the full Taylor squaring above is too slow: only n^2 operations per
 * term and column. */
void or_zexpmv_taylor_ld(int n, const ocplx* a, int lda, double t, int nb, const ocplx* b, int ldb, ocplx* f,
                         int ldf)
{
    size_t nn = (size_t)n * n;
    ldc* x = (ldc*)malloc(sizeof(ldc) * (nn ? nn : 1));
    long double norm = 0.0L;
    for (int j = 0; j < n; j++) {
        long double s = 0.0L;
        for (int i = 0; i < n; i++) {
            long double re = a[(size_t)j * lda + i].re, im = a[(size_t)j * lda + i].im;
            s += sqrtl(re * re + im * im);
        }
        if (s > norm) norm = s;
    }
    norm *= fabsl((long double)t);
    int steps = (int)ceill(norm);
    if (steps < 1) steps = 1;
    const long double f_scale = (long double)t / steps;
    for (size_t i = 0; i < nn; i++) {
        x[i].re = a[i % n + (i / n) * (size_t)lda].re * f_scale;
        x[i].im = a[i % n + (i / n) * (size_t)lda].im * f_scale;
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < nb; c++) {
        ldc* v = (ldc*)malloc(sizeof(ldc) * (n ? n : 1));
        ldc* term = (ldc*)malloc(sizeof(ldc) * (n ? n : 1));
        ldc* next = (ldc*)malloc(sizeof(ldc) * (n ? n : 1));
        for (int i = 0; i < n; i++) {
            v[i].re = b[(size_t)c * ldb + i].re;
            v[i].im = b[(size_t)c * ldb + i].im;
        }
        for (int q = 0; q < steps; q++) {
            memcpy(term, v, sizeof(ldc) * n);
            for (int k = 1; k < 200; k++) {
                for (int i = 0; i < n; i++) { next[i].re = 0.0L; next[i].im = 0.0L; }
                for (int p = 0; p < n; p++) {
                    const ldc y = term[p];
                    const ldc* col = x + (size_t)p * n;
                    for (int i = 0; i < n; i++) {
                        next[i].re += col[i].re * y.re - col[i].im * y.im;
                        next[i].im += col[i].re * y.im + col[i].im * y.re;
                    }
                }
                long double tmax = 0.0L, smax = 0.0L;
                for (int i = 0; i < n; i++) {
                    term[i].re = next[i].re / k;
                    term[i].im = next[i].im / k;
                    v[i].re += term[i].re;
                    v[i].im += term[i].im;
                    long double tm = fabsl(term[i].re) + fabsl(term[i].im);
                    long double sm = fabsl(v[i].re) + fabsl(v[i].im);
                    if (tm > tmax) tmax = tm;
                    if (sm > smax) smax = sm;
                }
                if (tmax <= 1e-22L * smax) break;
            }
        }
        for (int i = 0; i < n; i++) {
            f[(size_t)c * ldf + i].re = (double)v[i].re;
            f[(size_t)c * ldf + i].im = (double)v[i].im;
        }
        free(v); free(term); free(next);
    }
    free(x);
}

void or_dexpm_taylor_ld(int n, const double* a, int lda, double* e, int lde)
{
    size_t nn = (size_t)n * n;
    ldc* x = (ldc*)calloc(nn ? nn : 1, sizeof(ldc));
    ldc* r = (ldc*)calloc(nn ? nn : 1, sizeof(ldc));
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) x[(size_t)j * n + i].re = a[(size_t)j * lda + i];
    if (n > 0) taylor_ld(n, x, r);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) e[(size_t)j * lde + i] = (double)r[(size_t)j * n + i].re;
    free(x); free(r);
}

void or_zexpm_taylor_ld(int n, const ocplx* a, int lda, ocplx* e, int lde)
{
    size_t nn = (size_t)n * n;
    ldc* x = (ldc*)calloc(nn ? nn : 1, sizeof(ldc));
    ldc* r = (ldc*)calloc(nn ? nn : 1, sizeof(ldc));
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            x[(size_t)j * n + i].re = a[(size_t)j * lda + i].re;
            x[(size_t)j * n + i].im = a[(size_t)j * lda + i].im;
        }
    if (n > 0) taylor_ld(n, x, r);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < n; i++) {
            e[(size_t)j * lde + i].re = (double)r[(size_t)j * n + i].re;
            e[(size_t)j * lde + i].im = (double)r[(size_t)j * n + i].im;
        }
    free(x); free(r);
}
