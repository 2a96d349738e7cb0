/*
 * Oracle core: .NET arithmetic restatements, GEMM, norms, panel products.
 * Test infrastructure only (see oracle.h).
 */
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#if defined(__x86_64__)
#include <immintrin.h>
#endif

/* ------------------------------------------------------------------------
 * System.Numerics.Complex arithmetic as .NET evaluates it.
 * ---------------------------------------------------------------------- */

/* Complex.Abs: the larger part factored out, |z| = L*sqrt(1 + (S/L)^2), so
 * the square cannot overflow (the scaled hypot of the .NET runtime). */
double or_cabs(ocplx z)
{
    double a = fabs(z.re), b = fabs(z.im);
    double small, large;
    if (a < b) { small = a; large = b; } else { small = b; large = a; }
    if (small == 0.0) return large;
    if (isinf(large) && !isnan(small)) return INFINITY;
    double ratio = small / large;
    return large * sqrt(1.0 + ratio * ratio);
}

/* Complex division by Smith's method (R. L. Smith, CACM 5(8), 1962), the
 * form the .NET runtime uses. */
ocplx or_cdiv(ocplx x, ocplx y)
{
    double a = x.re, b = x.im, c = y.re, d = y.im;
    ocplx r;
    if (fabs(d) < fabs(c)) {
        double doc = d / c;
        r.re = (a + b * doc) / (c + d * doc);
        r.im = (b - a * doc) / (c + d * doc);
    } else {
        double cod = c / d;
        r.re = (b + a * cod) / (d + c * cod);
        r.im = (-a + b * cod) / (d + c * cod);
    }
    return r;
}

/* Complex.Exp: e^re (cos im, sin im). */
ocplx or_cexp(ocplx z)
{
    double e = exp(z.re);
    ocplx r = { e * cos(z.im), e * sin(z.im) };
    return r;
}

/* ------------------------------------------------------------------------
 * System.Random(seed): the seeded generator is Knuth's subtractive method
 * (TAOCP vol. 2, 3.6; Numerical Recipes "ran3"), in the form .NET keeps for
 * seeded instances. Used by NormEstimate.cs:163, 286-305.
 * ---------------------------------------------------------------------- */
#define RNG_MBIG INT_MAX
#define RNG_MSEED 161803398

void or_random_init(or_random* r, int seed)
{
    int* s = r->seed_array;
    int subtraction = (seed == INT_MIN) ? INT_MAX : abs(seed);
    int mj = RNG_MSEED - subtraction;
    s[55] = mj;
    int mk = 1;
    int ii = 0;
    for (int i = 1; i < 55; i++) {
        if ((ii += 21) >= 55) ii -= 55;
        s[ii] = mk;
        mk = mj - mk;
        if (mk < 0) mk += RNG_MBIG;
        mj = s[ii];
    }
    for (int k = 1; k < 5; k++) {
        for (int i = 1; i < 56; i++) {
            int n = i + 30;
            if (n >= 55) n -= 55;
            s[i] -= s[1 + n];
            if (s[i] < 0) s[i] += RNG_MBIG;
        }
    }
    r->inext = 0;
    r->inextp = 21;
}

static int random_internal(or_random* r)
{
    int a = r->inext, b = r->inextp;
    if (++a >= 56) a = 1;
    if (++b >= 56) b = 1;
    int v = r->seed_array[a] - r->seed_array[b];
    if (v == RNG_MBIG) v--;
    if (v < 0) v += RNG_MBIG;
    r->seed_array[a] = v;
    r->inext = a;
    r->inextp = b;
    return v;
}

double or_random_sample(or_random* r) { return random_internal(r) * (1.0 / RNG_MBIG); }

int or_random_next(or_random* r, int max_value) { return (int)(or_random_sample(r) * max_value); }

/* ------------------------------------------------------------------------
 * Array.Sort(double[] keys, int[] items): introspective sort (Musser 1997)
 * in the form of the .NET runtime's generic sort helper: NaNs moved to the
 * front first, median-of-three quicksort, insertion sort at <= 16 elements,
 * heapsort past a depth limit of 2*(floor(log2 n)+1). The estimator sorts
 * -h with the indices as items (NormEstimate.cs:376-385); the order of equal
 * keys depends on exactly this procedure.
 * ---------------------------------------------------------------------- */
static void swap_kv(double* k, int* v, int i, int j)
{
    double tk = k[i]; k[i] = k[j]; k[j] = tk;
    int tv = v[i]; v[i] = v[j]; v[j] = tv;
}

static void swap_if_greater(double* k, int* v, int i, int j)
{
    if (k[i] > k[j]) swap_kv(k, v, i, j);
}

static void insertion_sort(double* k, int* v, int n)
{
    for (int i = 0; i < n - 1; i++) {
        double t = k[i + 1];
        int tv = v[i + 1];
        int j = i;
        while (j >= 0 && t < k[j]) {
            k[j + 1] = k[j];
            v[j + 1] = v[j];
            j--;
        }
        k[j + 1] = t;
        v[j + 1] = tv;
    }
}

static void down_heap(double* k, int* v, int i, int n)
{
    double d = k[i - 1];
    int dv = v[i - 1];
    while (i <= n >> 1) {
        int child = 2 * i;
        if (child < n && k[child - 1] < k[child]) child++;
        if (!(d < k[child - 1])) break;
        k[i - 1] = k[child - 1];
        v[i - 1] = v[child - 1];
        i = child;
    }
    k[i - 1] = d;
    v[i - 1] = dv;
}

static void heap_sort(double* k, int* v, int n)
{
    for (int i = n >> 1; i >= 1; i--) down_heap(k, v, i, n);
    for (int i = n; i > 1; i--) {
        swap_kv(k, v, 0, i - 1);
        down_heap(k, v, 1, i - 1);
    }
}

static int pick_pivot_and_partition(double* k, int* v, int n)
{
    int hi = n - 1;
    int middle = hi >> 1;
    swap_if_greater(k, v, 0, middle);
    swap_if_greater(k, v, 0, hi);
    swap_if_greater(k, v, middle, hi);
    double pivot = k[middle];
    swap_kv(k, v, middle, hi - 1);
    int left = 0, right = hi - 1;
    while (left < right) {
        while (k[++left] < pivot) { }
        while (pivot < k[--right]) { }
        if (left >= right) break;
        swap_kv(k, v, left, right);
    }
    if (left != hi - 1) swap_kv(k, v, left, hi - 1);
    return left;
}

static void intro_sort(double* k, int* v, int n, int depth)
{
    int size = n;
    while (size > 1) {
        if (size <= 16) {
            if (size == 2) { swap_if_greater(k, v, 0, 1); return; }
            if (size == 3) {
                swap_if_greater(k, v, 0, 1);
                swap_if_greater(k, v, 0, 2);
                swap_if_greater(k, v, 1, 2);
                return;
            }
            insertion_sort(k, v, size);
            return;
        }
        if (depth == 0) { heap_sort(k, v, size); return; }
        depth--;
        int p = pick_pivot_and_partition(k, v, size);
        intro_sort(k + p + 1, v + p + 1, size - (p + 1), depth);
        size = p;
    }
}

void or_sort_keys_items(int n, double* keys, int* items)
{
    if (n <= 1) return;
    int left = 0;
    for (int i = 0; i < n; i++) {
        if (isnan(keys[i])) {
            swap_kv(keys, items, left, i);
            left++;
        }
    }
    n -= left;
    keys += left;
    items += left;
    if (n <= 1) return;
    int log2n = 31 - __builtin_clz((unsigned)n);
    intro_sort(keys, items, n, 2 * (log2n + 1));
}

/* ------------------------------------------------------------------------
 * GEMM. Semantics of Gemm.cs:117-203: scale C by beta first (beta==0
 * overwrites, beta==1 skips), return if k==0 or alpha==0, then accumulate
 * alpha*A*B with alpha applied at write-back. Blocking as Tensile's threaded
 * path (MC=144, KC=256, NC=4096; ParallelGemm.cs:16-250), micro-kernel 8x6
 * with FMA as Avx2Kernel8x6 (MicroKernels.cs:149-223).
 * ---------------------------------------------------------------------- */
static int g_threads = 0;

void or_set_threads(int threads) { g_threads = threads; }

int or_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

static void scale_c(int m, int n, double beta, double* c, int ldc)
{
    if (beta == 1.0) return;
    for (int j = 0; j < n; j++) {
        double* col = c + (size_t)j * ldc;
        if (beta == 0.0) memset(col, 0, sizeof(double) * (size_t)m);
        else for (int i = 0; i < m; i++) col[i] *= beta;
    }
}

void or_dgemm_ref(int m, int n, int k, double alpha, const double* a, int lda,
                  const double* b, int ldb, double beta, double* c, int ldc)
{
    if (m == 0 || n == 0) return;
    scale_c(m, n, beta, c, ldc);
    if (k == 0 || alpha == 0.0) return;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double sum = 0.0;
            for (int p = 0; p < k; p++) sum += a[(size_t)p * lda + i] * b[(size_t)j * ldb + p];
            c[(size_t)j * ldc + i] += alpha * sum;
        }
}

#define MR 8
#define NR 6
#define MC 144
#define KC 256
#define NC 4096

static void pack_a(int mc, int kc, const double* a, int lda, double* ap)
{
    for (int ir = 0; ir < mc; ir += MR) {
        int mr = mc - ir < MR ? mc - ir : MR;
        for (int p = 0; p < kc; p++) {
            const double* src = a + (size_t)p * lda + ir;
            for (int i = 0; i < mr; i++) ap[i] = src[i];
            for (int i = mr; i < MR; i++) ap[i] = 0.0;
            ap += MR;
        }
    }
}

static void pack_b(int kc, int nc, const double* b, int ldb, double* bp)
{
    for (int jr = 0; jr < nc; jr += NR) {
        int nr = nc - jr < NR ? nc - jr : NR;
        double* dst = bp + (size_t)(jr / NR) * NR * kc;
        for (int p = 0; p < kc; p++) {
            for (int j = 0; j < nr; j++) dst[j] = b[(size_t)(jr + j) * ldb + p];
            for (int j = nr; j < NR; j++) dst[j] = 0.0;
            dst += NR;
        }
    }
}

static void kernel_scalar(int kc, const double* ap, const double* bp, double* acc)
{
    for (int x = 0; x < MR * NR; x++) acc[x] = 0.0;
    for (int p = 0; p < kc; p++)
        for (int j = 0; j < NR; j++)
            for (int i = 0; i < MR; i++) acc[j * MR + i] = fma(ap[p * MR + i], bp[p * NR + j], acc[j * MR + i]);
}

#if defined(__x86_64__)
__attribute__((target("avx2,fma")))
static void kernel_avx2(int kc, const double* ap, const double* bp, double* acc)
{
    __m256d c00 = _mm256_setzero_pd(), c01 = _mm256_setzero_pd();
    __m256d c10 = _mm256_setzero_pd(), c11 = _mm256_setzero_pd();
    __m256d c20 = _mm256_setzero_pd(), c21 = _mm256_setzero_pd();
    __m256d c30 = _mm256_setzero_pd(), c31 = _mm256_setzero_pd();
    __m256d c40 = _mm256_setzero_pd(), c41 = _mm256_setzero_pd();
    __m256d c50 = _mm256_setzero_pd(), c51 = _mm256_setzero_pd();
    for (int p = 0; p < kc; p++) {
        __m256d a0 = _mm256_loadu_pd(ap), a1 = _mm256_loadu_pd(ap + 4);
        __m256d b;
        b = _mm256_broadcast_sd(bp + 0); c00 = _mm256_fmadd_pd(a0, b, c00); c01 = _mm256_fmadd_pd(a1, b, c01);
        b = _mm256_broadcast_sd(bp + 1); c10 = _mm256_fmadd_pd(a0, b, c10); c11 = _mm256_fmadd_pd(a1, b, c11);
        b = _mm256_broadcast_sd(bp + 2); c20 = _mm256_fmadd_pd(a0, b, c20); c21 = _mm256_fmadd_pd(a1, b, c21);
        b = _mm256_broadcast_sd(bp + 3); c30 = _mm256_fmadd_pd(a0, b, c30); c31 = _mm256_fmadd_pd(a1, b, c31);
        b = _mm256_broadcast_sd(bp + 4); c40 = _mm256_fmadd_pd(a0, b, c40); c41 = _mm256_fmadd_pd(a1, b, c41);
        b = _mm256_broadcast_sd(bp + 5); c50 = _mm256_fmadd_pd(a0, b, c50); c51 = _mm256_fmadd_pd(a1, b, c51);
        ap += MR;
        bp += NR;
    }
    _mm256_storeu_pd(acc + 0, c00); _mm256_storeu_pd(acc + 4, c01);
    _mm256_storeu_pd(acc + 8, c10); _mm256_storeu_pd(acc + 12, c11);
    _mm256_storeu_pd(acc + 16, c20); _mm256_storeu_pd(acc + 20, c21);
    _mm256_storeu_pd(acc + 24, c30); _mm256_storeu_pd(acc + 28, c31);
    _mm256_storeu_pd(acc + 32, c40); _mm256_storeu_pd(acc + 36, c41);
    _mm256_storeu_pd(acc + 40, c50); _mm256_storeu_pd(acc + 44, c51);
}
#endif

typedef void (*micro_fn)(int, const double*, const double*, double*);

static micro_fn pick_kernel(void)
{
#if defined(__x86_64__)
    if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return kernel_avx2;
#endif
    return kernel_scalar;
}

void or_dgemm(int m, int n, int k, double alpha, const double* a, int lda,
              const double* b, int ldb, double beta, double* c, int ldc)
{
    if (m == 0 || n == 0) return;
    scale_c(m, n, beta, c, ldc);
    if (k == 0 || alpha == 0.0) return;

    micro_fn kern = pick_kernel();
    int threads = or_get_threads();
    /* Tensile's dispatch threshold: threaded only from m*n*k >= 2^24
     * (GemmDispatch.cs:45, 100-119). */
    if ((double)m * n * k < 16777216.0) threads = 1;

    int ncap = n < NC ? n : NC;
    double* bp = (double*)aligned_alloc(64, sizeof(double) * (size_t)KC * (size_t)(((ncap + NR - 1) / NR) * NR));

    for (int jc = 0; jc < n; jc += NC) {
        int nc = n - jc < NC ? n - jc : NC;
        for (int pc = 0; pc < k; pc += KC) {
            int kc = k - pc < KC ? k - pc : KC;
            pack_b(kc, nc, b + (size_t)jc * ldb + pc, ldb, bp);
            int blocks = (m + MC - 1) / MC;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) if (threads > 1)
            for (int blk = 0; blk < blocks; blk++) {
                int ic = blk * MC;
                int mc = m - ic < MC ? m - ic : MC;
                double ap[MC * KC] __attribute__((aligned(64)));
                double acc[MR * NR] __attribute__((aligned(64)));
                pack_a(mc, kc, a + (size_t)pc * lda + ic, lda, ap);
                for (int jr = 0; jr < nc; jr += NR) {
                    int nr = nc - jr < NR ? nc - jr : NR;
                    const double* bpan = bp + (size_t)(jr / NR) * NR * kc;
                    for (int ir = 0; ir < mc; ir += MR) {
                        int mr = mc - ir < MR ? mc - ir : MR;
                        kern(kc, ap + (size_t)(ir / MR) * MR * kc, bpan, acc);
                        double* ct = c + (size_t)(jc + jr) * ldc + ic + ir;
                        for (int j = 0; j < nr; j++)
                            for (int i = 0; i < mr; i++) ct[(size_t)j * ldc + i] += alpha * acc[j * MR + i];
                    }
                }
            }
        }
    }
    free(bp);
}

/* ------------------------------------------------------------------------
 * 4M complex GEMM (ComplexKernels.cs:49-108): split, four real products
 *   real = Ar*Br - Ai*Bi,  imag = Ar*Bi + Ai*Br,
 * then Combine (ComplexKernels.cs:213-240) with its special cases.
 * ---------------------------------------------------------------------- */
static int cplx_eq(ocplx a, double re, double im) { return a.re == re && a.im == im; }

static ocplx cmul(ocplx x, ocplx y)
{
    ocplx r = { x.re * y.re - x.im * y.im, x.im * y.re + x.re * y.im };
    return r;
}

void or_zgemm(int m, int n, int k, ocplx alpha, const ocplx* a, int lda,
              const ocplx* b, int ldb, ocplx beta, ocplx* c, int ldc)
{
    if (m == 0 || n == 0) return;
    size_t mk = (size_t)m * k, kn = (size_t)k * n, mn = (size_t)m * n;
    double* ar = (double*)malloc(sizeof(double) * (mk ? mk : 1));
    double* ai = (double*)malloc(sizeof(double) * (mk ? mk : 1));
    double* br = (double*)malloc(sizeof(double) * (kn ? kn : 1));
    double* bi = (double*)malloc(sizeof(double) * (kn ? kn : 1));
    double* p = (double*)malloc(sizeof(double) * mn);
    double* q = (double*)malloc(sizeof(double) * mn);
    for (int j = 0; j < k; j++)
        for (int i = 0; i < m; i++) {
            ar[(size_t)j * m + i] = a[(size_t)j * lda + i].re;
            ai[(size_t)j * m + i] = a[(size_t)j * lda + i].im;
        }
    for (int j = 0; j < n; j++)
        for (int i = 0; i < k; i++) {
            br[(size_t)j * k + i] = b[(size_t)j * ldb + i].re;
            bi[(size_t)j * k + i] = b[(size_t)j * ldb + i].im;
        }
    int ldk = k > 0 ? k : 1;
    or_dgemm(m, n, k, 1.0, ar, m, br, ldk, 0.0, p, m);
    or_dgemm(m, n, k, -1.0, ai, m, bi, ldk, 1.0, p, m);
    or_dgemm(m, n, k, 1.0, ar, m, bi, ldk, 0.0, q, m);
    or_dgemm(m, n, k, 1.0, ai, m, br, ldk, 1.0, q, m);

    int unit_alpha = cplx_eq(alpha, 1.0, 0.0);
    int real_alpha = alpha.im == 0.0;
    int zero_beta = cplx_eq(beta, 0.0, 0.0);
    int unit_beta = cplx_eq(beta, 1.0, 0.0);
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double pv = p[(size_t)j * m + i], qv = q[(size_t)j * m + i];
            ocplx prod;
            if (unit_alpha) { prod.re = pv; prod.im = qv; }
            else if (real_alpha) { prod.re = alpha.re * pv; prod.im = alpha.re * qv; }
            else { ocplx pq = { pv, qv }; prod = cmul(alpha, pq); }
            ocplx* t = c + (size_t)j * ldc + i;
            if (zero_beta) *t = prod;
            else if (unit_beta) { t->re = t->re + prod.re; t->im = t->im + prod.im; }
            else { ocplx bt = cmul(beta, *t); t->re = bt.re + prod.re; t->im = bt.im + prod.im; }
        }
    free(ar); free(ai); free(br); free(bi); free(p); free(q);
}

void or_zgemm_3m(int m, int n, int k, const ocplx* a, int lda,
                 const ocplx* b, int ldb, ocplx* c, int ldc)
{
    /* Im(AB) = (Ar+Ai)(Br+Bi) - ArBr - AiBi: the shortcut 4M refuses. */
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            double t1 = 0.0, t2 = 0.0, t3 = 0.0;
            for (int p = 0; p < k; p++) {
                ocplx x = a[(size_t)p * lda + i], y = b[(size_t)j * ldb + p];
                t1 += x.re * y.re;
                t2 += x.im * y.im;
                t3 += (x.re + x.im) * (y.re + y.im);
            }
            c[(size_t)j * ldc + i].re = t1 - t2;
            c[(size_t)j * ldc + i].im = t3 - t1 - t2;
        }
}

/* ------------------------------------------------------------------------
 * Exact norms and panel products.
 * ---------------------------------------------------------------------- */
double or_dnorm1(int m, int n, const double* a, int lda)
{
    if (m == 0 || n == 0) return 0.0;
    double best = 0.0;
    for (int j = 0; j < n; j++) {
        double sum = 0.0;
        for (int i = 0; i < m; i++) sum += fabs(a[(size_t)j * lda + i]);
        best = fmax(best, sum);
    }
    return best;
}

double or_dnorminf(int m, int n, const double* a, int lda)
{
    if (m == 0 || n == 0) return 0.0;
    double* sums = (double*)calloc((size_t)m, sizeof(double));
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) sums[i] += fabs(a[(size_t)j * lda + i]);
    double best = 0.0;
    for (int i = 0; i < m; i++) best = fmax(best, sums[i]);
    free(sums);
    return best;
}

double or_dnormfro(int m, int n, const double* a, int lda)
{
    if (m == 0 || n == 0) return 0.0;
    double total = 0.0;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) total += a[(size_t)j * lda + i] * a[(size_t)j * lda + i];
    return sqrt(total);
}

double or_znorm1(int m, int n, const ocplx* a, int lda)
{
    if (m == 0 || n == 0) return 0.0;
    double best = 0.0;
    for (int j = 0; j < n; j++) {
        double sum = 0.0;
        for (int i = 0; i < m; i++) sum += or_cabs(a[(size_t)j * lda + i]);
        best = fmax(best, sum);
    }
    return best;
}

double or_znorminf(int m, int n, const ocplx* a, int lda)
{
    if (m == 0 || n == 0) return 0.0;
    double* sums = (double*)calloc((size_t)m, sizeof(double));
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) sums[i] += or_cabs(a[(size_t)j * lda + i]);
    double best = 0.0;
    for (int i = 0; i < m; i++) best = fmax(best, sums[i]);
    free(sums);
    return best;
}

void or_dpanel(int trans, int n, int t, const double* a, int lda,
               const double* x, int ldx, double* y, int ldy)
{
    for (int col = 0; col < t; col++) {
        const double* src = x + (size_t)col * ldx;
        double* dst = y + (size_t)col * ldy;
        if (!trans) {
            /* Axpys over the columns of A, zero scales skipped. */
            for (int i = 0; i < n; i++) dst[i] = 0.0;
            for (int j = 0; j < n; j++) {
                double s = src[j];
                if (s == 0.0) continue;
                const double* aj = a + (size_t)j * lda;
                for (int i = 0; i < n; i++) dst[i] = dst[i] + s * aj[i];
            }
        } else {
            /* Inner products down the columns of A. */
            for (int j = 0; j < n; j++) {
                const double* aj = a + (size_t)j * lda;
                double sum = 0.0;
                for (int i = 0; i < n; i++) sum += aj[i] * src[i];
                dst[j] = sum;
            }
        }
    }
}
