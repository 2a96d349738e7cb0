// Memory-bound device kernels: exact norms (Norms.cs:11-66,
// ComplexKernels.cs:144-180), narrow panel products (PanelProduct.cs:25-55,
// LinearOperators.cs:232-283), and the element-wise steps of the
// exponentials (MatrixExponential.cs:518-558, MatrixExponentialAction.cs:
// 286-317, 446-496), fused where the reference makes several passes.
//
// Every reduction is deterministic: partial results go to a buffer in a
// fixed layout and are combined in a fixed order, so repeated calls give the
// same bits (the norm estimator and the (m, s) selection depend on that).
// Maxima are combined with NaN propagating, as Math.Max does.
#include "blas1.hpp"

#include "tns_scalar.hpp"

namespace tns {
namespace {

__device__ __forceinline__ double nan_max(double a, double b)
{
    if (isnan(a) || isnan(b)) return NAN;
    return a > b ? a : b;
}

__device__ __forceinline__ double wave_sum(double v) { return wave_xor_sum(v); }

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = nan_max(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ zd wave_sum(zd v)
{
    return make_zd(wave_sum(v.x), wave_sum(v.y));
}

template <typename T>
__device__ __forceinline__ double sq(T x);
template <>
__device__ __forceinline__ double sq<double>(double x) { return x * x; }
template <>
__device__ __forceinline__ double sq<zd>(zd x) { return x.x * x.x + x.y * x.y; }

// One wavefront per column: sum of |a_ij| (kind 0) or a_ij^2 (kind 2).
template <typename T>
__global__ void __launch_bounds__(256) col_reduce_kernel(int m, int n, const T* __restrict__ a, int lda, int kind,
                                                         double* __restrict__ out)
{
    int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    int lane = threadIdx.x & 63;
    if (j >= n) return;
    const T* col = a + (size_t)j * lda;
    double s = 0.0;
    if (kind == 0)
        for (int i = lane; i < m; i += 64) s += s_abs(col[i]);
    else
        for (int i = lane; i < m; i += 64) s += sq(col[i]);
    s = wave_sum(s);
    if (lane == 0) out[j] = s;
}

// Partial row sums of |a_ij| over column chunks of 64: part[chunk][i].
template <typename T>
__global__ void __launch_bounds__(256) row_partial_kernel(int m, int n, const T* __restrict__ a, int lda,
                                                          double* __restrict__ part)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    int chunk = blockIdx.y;
    if (i >= m) return;
    int j0 = chunk * 64, j1 = j0 + 64 < n ? j0 + 64 : n;
    double s = 0.0;
    for (int j = j0; j < j1; j++) s += s_abs(a[i + (size_t)j * lda]);
    part[(size_t)chunk * m + i] = s;
}

__global__ void __launch_bounds__(256) row_finish_kernel(int m, int chunks, const double* __restrict__ part,
                                                         double* __restrict__ out)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    double s = 0.0;
    for (int c = 0; c < chunks; c++) s += part[(size_t)c * m + i];
    out[i] = s;
}

// One workgroup: out[0] = max (mode 0) or sqrt(sum) (mode 1) of v[0..n).
__global__ void __launch_bounds__(256) vec_finish_kernel(int n, const double* __restrict__ v, int mode,
                                                         double* __restrict__ out)
{
    __shared__ double red[4];
    double acc = 0.0;
    if (mode == 0)
        for (int i = threadIdx.x; i < n; i += 256) acc = nan_max(acc, v[i]);
    else
        for (int i = threadIdx.x; i < n; i += 256) acc += v[i];
    acc = mode == 0 ? wave_max(acc) : wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int q = 1; q < 4; q++) r = mode == 0 ? nan_max(r, red[q]) : r + red[q];
        out[0] = mode == 0 ? r : sqrt(r);
    }
}

// ---------------------------------------------------------------------------
// Narrow panel products. Y = op(A) X with A n x n and X n x t, t small.
// ---------------------------------------------------------------------------
constexpr int PN_ROWS = 64;   // rows per workgroup
constexpr int PN_KL = 4;      // k-lanes per workgroup (one wavefront each)
constexpr int PN_SUB = 64;    // columns of the chunk staged at a time
constexpr int PN_PER = PN_SUB / PN_KL;  // a wavefront's columns per stage

// Y = A X, split over k: part[ks][c][i] = sum over the ks-th column chunk.
// Wavefront kl takes columns j0 + kl, j0 + kl + 4, ... of the chunk, its
// lanes 64 consecutive rows (coalesced), and adds in that column order.
// The chunk is taken 64 columns at a time: the wavefront's 16 columns of A
// are loaded at once while the workgroup copies X's 64 rows to LDS, from
// where every lane reads the same entry (a broadcast).
// FULL: t == TMAX, so no column test guards a load.
template <typename T, int TMAX, bool FULL>
__device__ __forceinline__ void panel_n_body(int n, int t, const T* __restrict__ a, int lda, const T* __restrict__ x,
                                             int ldx, int kchunk, T* __restrict__ part, T (&red)[PN_KL][TMAX][PN_ROWS],
                                             T (&xs)[TMAX][PN_SUB])
{
    constexpr int XPER = (TMAX * PN_SUB + 255) / 256;  // X entries a thread stages
    const int lane = threadIdx.x & 63;
    const int kl = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = blockIdx.x * PN_ROWS + lane;
    const int ir = i < n ? i : n - 1;  // rows past the end load row n - 1; never stored
    const int ks = blockIdx.y;
    const int j0 = ks * kchunk, j1 = j0 + kchunk < n ? j0 + kchunk : n;
    T acc[TMAX];
#pragma unroll
    for (int c = 0; c < TMAX; c++) acc[c] = s_zero(T{});
    for (int s0 = j0; s0 < j1; s0 += PN_SUB) {
        const int cnt = j1 - s0 < PN_SUB ? j1 - s0 : PN_SUB;
        T xv[XPER];
#pragma unroll
        for (int q = 0; q < XPER; q++) {
            const int e = threadIdx.x + 256 * q, c = e / PN_SUB, jj = e % PN_SUB;
            xv[q] = e < TMAX * PN_SUB && jj < cnt && (FULL || c < t) ? x[s0 + jj + (size_t)c * ldx] : s_zero(T{});
        }
        T av[PN_PER];
#pragma unroll
        for (int u = 0; u < PN_PER; u++) {
            const int jj = kl + PN_KL * u;
            av[u] = a[ir + (size_t)(s0 + (jj < cnt ? jj : 0)) * lda];
        }
        __syncthreads();  // the previous stage's readers are done with xs
#pragma unroll
        for (int q = 0; q < XPER; q++) {
            const int e = threadIdx.x + 256 * q;
            if (e < TMAX * PN_SUB) xs[e / PN_SUB][e % PN_SUB] = xv[q];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PN_PER; u++) {
            const int jj = kl + PN_KL * u;
            if (jj < cnt) {
#pragma unroll
                for (int c = 0; c < TMAX; c++)
                    if (FULL || c < t) acc[c] = s_add(acc[c], s_mul(av[u], xs[c][jj]));
            }
        }
    }
#pragma unroll
    for (int c = 0; c < TMAX; c++) red[kl][c][lane] = acc[c];
    __syncthreads();
    if (kl == 0 && i < n) {
#pragma unroll
        for (int c = 0; c < TMAX; c++) {
            if (FULL || c < t) {
                T s = red[0][c][lane];
#pragma unroll
                for (int q = 1; q < PN_KL; q++) s = s_add(s, red[q][c][lane]);
                part[((size_t)ks * t + c) * n + i] = s;
            }
        }
    }
}

// About one wavefront per SIMD at the orders these serve: registers for
// loads in flight matter more than occupancy.
template <typename T, int TMAX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) panel_n_kernel(int n, int t, const T* __restrict__ a, int lda,
                                                      const T* __restrict__ x, int ldx, int kchunk,
                                                      T* __restrict__ part, const int* __restrict__ skip)
{
    if (skip != nullptr && *skip) return;  // a finished Taylor stage: no work
    __shared__ T red[PN_KL][TMAX][PN_ROWS];
    __shared__ T xs[TMAX][PN_SUB];
    if (t == TMAX)
        panel_n_body<T, TMAX, true>(n, t, a, lda, x, ldx, kchunk, part, red, xs);
    else
        panel_n_body<T, TMAX, false>(n, t, a, lda, x, ldx, kchunk, part, red, xs);
}

template <typename T>
__global__ void __launch_bounds__(256) panel_finish_kernel(int n, int t, int ksplit, const T* __restrict__ part,
                                                           T* __restrict__ y, int ldy)
{
    size_t e = blockIdx.x * (size_t)256 + threadIdx.x;
    if (e >= (size_t)n * t) return;
    int c = (int)(e / n), i = (int)(e % n);
    const T* p = part + (size_t)c * n + i;
    const size_t qs = (size_t)t * n;  // from one k-chunk's partials to the next
    T s = p[0];
    int q = 1;
    // Eight chunks' loads in flight at once, summed in chunk order.
    for (; q + 7 < ksplit; q += 8) {
        T v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = p[(size_t)(q + u) * qs];
#pragma unroll
        for (int u = 0; u < 8; u++) s = s_add(s, v[u]);
    }
    for (; q < ksplit; q++) s = s_add(s, p[(size_t)q * qs]);
    y[i + (size_t)c * ldy] = s;
}

// Rows a lane takes per stage of the transposed product (R, at most 16
// real or 8 complex): X's stage, 64 R rows x TMAX columns, stays within
// 64 KB of LDS.
template <typename T, int TMAX>
constexpr int pt_rows()
{
    constexpr int cap = sizeof(T) == 8 ? 16 : 8;
    constexpr int fit = 65536 / (64 * TMAX * (int)sizeof(T));
    return fit < cap ? fit : cap;
}

// Y = op(A)^T-form: y_j = sum_i conj?(a_ij) x_i, one wavefront per column j,
// four columns to a workgroup. Lane l adds rows l, l + 64, ... in that
// order, then the wavefront sums its lanes. The rows are taken 64 R at a
// time: the workgroup copies X's rows of the stage to LDS, shared by its
// four columns, and each lane's R loads of A for the next stage are issued
// before the current stage's products.
// FULL: t == TMAX, so no column test guards a load.
template <typename T, int TMAX, bool CONJ, bool FULL>
__device__ __forceinline__ void panel_t_body(int n, int t, const T* __restrict__ a, int lda, const T* __restrict__ x,
                                             int ldx, T* __restrict__ y, int ldy, T (&xs)[TMAX][64 * pt_rows<T, TMAX>()])
{
    constexpr int R = pt_rows<T, TMAX>(), RC = 64 * R;
    constexpr int XPER = TMAX * RC / 256;  // X entries a thread stages
    const int j = blockIdx.x * 4 + (int)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    // A workgroup's columns past the end read column n - 1 and store nothing
    // (every wavefront takes part in the stages' barriers).
    const T* col = a + (size_t)(j < n ? j : n - 1) * lda;
    T acc[TMAX];
#pragma unroll
    for (int c = 0; c < TMAX; c++) acc[c] = s_zero(T{});
    T av[R];
#pragma unroll
    for (int u = 0; u < R; u++) av[u] = col[lane + 64 * u < n ? lane + 64 * u : 0];
    for (int r0 = 0; r0 < n; r0 += RC) {
        T xv[XPER];
#pragma unroll
        for (int q = 0; q < XPER; q++) {
            const int e = threadIdx.x + 256 * q, c = e / RC, ii = r0 + e % RC;
            xv[q] = ii < n && (FULL || c < t) ? x[ii + (size_t)c * ldx] : s_zero(T{});
        }
        __syncthreads();  // the previous stage's readers are done with xs
#pragma unroll
        for (int q = 0; q < XPER; q++) {
            const int e = threadIdx.x + 256 * q;
            xs[e / RC][e % RC] = xv[q];
        }
        __syncthreads();
        T an[R];
#pragma unroll
        for (int u = 0; u < R; u++) {
            const int ii = r0 + RC + lane + 64 * u;
            an[u] = col[ii < n ? ii : 0];
        }
#pragma unroll
        for (int u = 0; u < R; u++) {
            if (r0 + lane + 64 * u < n) {
                const T v = CONJ ? s_conj(av[u]) : av[u];
#pragma unroll
                for (int c = 0; c < TMAX; c++)
                    if (FULL || c < t) acc[c] = s_add(acc[c], s_mul(v, xs[c][lane + 64 * u]));
            }
        }
#pragma unroll
        for (int u = 0; u < R; u++) av[u] = an[u];
    }
    if (j < n) {
#pragma unroll
        for (int c = 0; c < TMAX; c++) {
            if (FULL || c < t) {
                T s = wave_sum(acc[c]);
                if (lane == 0) y[j + (size_t)c * ldy] = s;
            }
        }
    }
}

template <typename T, int TMAX, bool CONJ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2))) panel_t_kernel(int n, int t, const T* __restrict__ a, int lda,
                                                      const T* __restrict__ x, int ldx, T* __restrict__ y, int ldy)
{
    __shared__ T xs[TMAX][64 * pt_rows<T, TMAX>()];
    if (t == TMAX)
        panel_t_body<T, TMAX, CONJ, true>(n, t, a, lda, x, ldx, y, ldy, xs);
    else
        panel_t_body<T, TMAX, CONJ, false>(n, t, a, lda, x, ldx, y, ldy, xs);
}

// ---------------------------------------------------------------------------
// Element-wise.
// ---------------------------------------------------------------------------

// out := c0*I + sum_q coef[q] * src[q], accumulated in the reference's
// order (Identity, then AddScaled one operand at a time), n x n, ld n.
template <typename T>
__global__ void lincomb_kernel(int n, double c0, int terms, LinTerms<T> lt, T* __restrict__ out)
{
    size_t total = (size_t)n * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        size_t j = e / n, i = e % n;
        T v = s_zero(T{});
        if (i == j) v = s_add(v, s_scale(s_one(T{}), c0));
        for (int q = 0; q < terms; q++) v = s_add(v, s_scale(lt.src[q][e], lt.coef[q]));
        out[e] = v;
    }
}

// den := v - u, num := v + u.
template <typename T>
__global__ void sub_add_kernel(size_t count, const T* __restrict__ u, const T* __restrict__ v, T* __restrict__ den,
                               T* __restrict__ num)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
        den[e] = s_sub(v[e], u[e]);
        num[e] = s_add(v[e], u[e]);
    }
}

template <typename T>
__global__ void scale_kernel(size_t count, const T* __restrict__ src, double f, T* __restrict__ dst)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
        dst[e] = s_scale(src[e], f);
}

template <typename T>
__global__ void mul_scalar_kernel(size_t count, T* __restrict__ x, T s)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
        x[e] = s_mul(x[e], s);
}

template <typename T>
__global__ void shift_diag_kernel(int n, T* __restrict__ a, int lda, T delta)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i + (size_t)i * lda] = s_add(a[i + (size_t)i * lda], delta);
}

// Sequential diagonal sum, one thread: the reference's Trace order.
template <typename T>
__global__ void __launch_bounds__(1024) trace_kernel(int n, const T* __restrict__ a, int lda, T* __restrict__ out)
{
    // The diagonal is read by the whole workgroup a slab at a time, then
    // summed by one thread in index order (the reference's sequential sum).
    __shared__ T diag[1024];
    T s = s_zero(T{});
    for (int i0 = 0; i0 < n; i0 += 1024) {
        const int i = i0 + (int)threadIdx.x;
        if (i < n) diag[threadIdx.x] = a[i + (size_t)i * lda];
        __syncthreads();
        if (threadIdx.x == 0) {
            const int cnt = n - i0 < 1024 ? n - i0 : 1024;
            for (int q = 0; q < cnt; q++) s = s_add(s, diag[q]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s;
}

template <typename T>
__global__ void abs_kernel(size_t count, const T* __restrict__ a, double* __restrict__ out)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
        out[e] = s_abs(a[e]);
}

// next_j = sum_i row_i * m_ij (the e^T |A|^p recurrence of Ell), one
// wavefront per column.
__global__ void __launch_bounds__(256) rowvec_kernel(int n, const double* __restrict__ mags,
                                                     const double* __restrict__ row, double* __restrict__ next)
{
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= n) return;
    const double* col = mags + (size_t)j * n;
    double s = 0.0;
    // Sixteen rows' loads in flight at once (past the end a lane reloads
    // its first row and skips the add), added in the same order as one at a
    // time.
    for (int i = lane; i < n; i += 16 * 64) {
        double r[16], c[16];
#pragma unroll
        for (int u = 0; u < 16; u++) {
            const int k = i + 64 * u < n ? i + 64 * u : i;
            r[u] = row[k];
            c[u] = col[k];
        }
#pragma unroll
        for (int u = 0; u < 16; u++)
            if (i + 64 * u < n) s += r[u] * c[u];
    }
    s = wave_sum(s);
    if (lane == 0) next[j] = s;
}

__global__ void fill_kernel(int n, double* __restrict__ v, double value)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = value;
}

// Bits of a non-negative double (or +NaN) order as the integers do, so the
// maximum over rows is an integer atomic max: order-independent.
__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v)
{
    if (isnan(v)) v = NAN;
    atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

// The Taylor step of MatrixExponentialAction.Run (MatrixExponentialAction.cs:
// 292-303), fused with the second pass of the panel product:
//   next = sum of the k-split partials, work = coef*next, f += work,
//   norms[0] = ||work||_inf, norms[1] = ||f||_inf.
// A workgroup takes 32 rows, eight threads to a row, each on one column
// of a group of eight at a time (the partials' loads eight at once); the
// row sums of |work| and |f| are then added by the row's first thread in
// column order, as one thread per row would add them.
template <typename T>
__global__ void __launch_bounds__(256) taylor_step_kernel(int n, int nb, int ksplit, const T* __restrict__ part,
                                                          double coef, T* __restrict__ work, T* __restrict__ f,
                                                          double* __restrict__ norms, const int* __restrict__ skip)
{
    if (skip != nullptr && *skip) return;  // a finished Taylor stage: no work
    __shared__ double aw[32][9], af[32][9];
    const int r = threadIdx.x >> 3, slot = threadIdx.x & 7;
    const int i = blockIdx.x * 32 + r;
    const size_t qs = (size_t)nb * n;  // between one k-chunk's partials and the next
    double sw = 0.0, sf = 0.0;
    for (int c0 = 0; c0 < nb; c0 += 8) {
        const int c = c0 + slot;
        double mw = 0.0, mf = 0.0;
        if (i < n && c < nb) {
            const T* p = part + (size_t)c * n + i;
            T s = p[0];
            int q = 1;
            for (; q + 7 < ksplit; q += 8) {
                T v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) v[u] = p[(size_t)(q + u) * qs];
#pragma unroll
                for (int u = 0; u < 8; u++) s = s_add(s, v[u]);
            }
            for (; q < ksplit; q++) s = s_add(s, p[(size_t)q * qs]);
            const T w = s_scale(s, coef);
            const size_t e = i + (size_t)c * n;
            work[e] = w;
            const T fv = s_add(f[e], w);
            f[e] = fv;
            mw = s_abs(w);
            mf = s_abs(fv);
        }
        aw[r][slot] = mw;
        af[r][slot] = mf;
        __syncthreads();
        if (slot == 0) {
            const int cn = nb - c0 < 8 ? nb - c0 : 8;
            for (int k = 0; k < cn; k++) {
                sw += aw[r][k];
                sf += af[r][k];
            }
        }
        __syncthreads();
    }
    if (slot == 0 && i < n) {
        atomic_max_nonneg(norms + 0, sw);
        atomic_max_nonneg(norms + 1, sf);
    }
}

// ||x||_inf of an n x nb panel (ld n) into out[0] (which must start at 0).
template <typename T>
__global__ void __launch_bounds__(256) panel_infnorm_kernel(int n, int nb, const T* __restrict__ x,
                                                            double* __restrict__ out)
{
    int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < nb; c++) s += s_abs(x[i + (size_t)c * n]);
    atomic_max_nonneg(out, s);
}

// After a Taylor step of a stage still running (state[0] == 0): one more
// application; the stage ends (state[0] = 1) when c1 + c2 <= unit * ||F||
// (c1 the previous step's ||work||inf, c2 this one's, in norms[2] and
// norms[0], ||F||inf in norms[1]), else c1 := c2. norms[0..1] are zeroed
// for the next step either way. One thread.
__global__ void taylor_exit_kernel(double* norms, int* state, double unit)
{
    if (state[0] == 0) {
        const double c2 = norms[0];
        state[1] += 1;
        if (norms[2] + c2 <= unit * norms[1]) state[0] = 1;
        else norms[2] = c2;
    }
    norms[0] = 0.0;
    norms[1] = 0.0;
}

__global__ void zero_doubles_kernel(double* p, int count)
{
    int i = threadIdx.x;
    if (i < count) p[i] = 0.0;
}

inline int grid_for(size_t total, int block = 256, int cap = 16384)
{
    size_t g = (total + block - 1) / block;
    if (g < 1) g = 1;
    return (int)(g < (size_t)cap ? g : cap);
}

template <typename T>
void norm_impl(Context* ctx, int kind, int m, int n, const T* a, int lda, double* d_out)
{
    if (m == 0 || n == 0) {
        hipLaunchKernelGGL(zero_doubles_kernel, dim3(1), dim3(64), 0, ctx->stream, d_out, 1);
        return;
    }
    if (kind == '1' || kind == 'F') {
        DevBuf<double> col(ctx, (size_t)n);
        hipLaunchKernelGGL((col_reduce_kernel<T>), dim3((n + 3) / 4), dim3(256), 0, ctx->stream, m, n, a, lda,
                           kind == '1' ? 0 : 2, col.get());
        hipLaunchKernelGGL(vec_finish_kernel, dim3(1), dim3(256), 0, ctx->stream, n, col.get(), kind == '1' ? 0 : 1,
                           d_out);
    } else {
        int chunks = (n + 63) / 64;
        DevBuf<double> part(ctx, (size_t)chunks * m), rows(ctx, (size_t)m);
        hipLaunchKernelGGL((row_partial_kernel<T>), dim3((m + 255) / 256, chunks), dim3(256), 0, ctx->stream, m, n, a,
                           lda, part.get());
        hipLaunchKernelGGL(row_finish_kernel, dim3((m + 255) / 256), dim3(256), 0, ctx->stream, m, chunks, part.get(),
                           rows.get());
        hipLaunchKernelGGL(vec_finish_kernel, dim3(1), dim3(256), 0, ctx->stream, m, rows.get(), 0, d_out);
    }
    TNS_HIP(hipGetLastError());
}

// k-split for the n-form panel product: enough workgroups to cover the chip.
inline int ksplit_for(int n)
{
    int rowblocks = (n + PN_ROWS - 1) / PN_ROWS;
    int ks = (512 + rowblocks - 1) / rowblocks;
    int maxks = (n + 63) / 64;  // at least 64 columns per chunk
    if (ks > maxks) ks = maxks;
    if (ks > 64) ks = 64;
    if (ks < 1) ks = 1;
    return ks;
}

template <typename T>
void panel_partials(Context* ctx, int n, int t, const T* a, int lda, const T* x, int ldx, T* part, int ks,
                    const int* skip = nullptr)
{
    int kchunk = (n + ks - 1) / ks;
    dim3 grid((n + PN_ROWS - 1) / PN_ROWS, ks);
    if (t <= 2)
        hipLaunchKernelGGL((panel_n_kernel<T, 2>), grid, dim3(256), 0, ctx->stream, n, t, a, lda, x, ldx, kchunk, part,
                           skip);
    else if (t <= 4)
        hipLaunchKernelGGL((panel_n_kernel<T, 4>), grid, dim3(256), 0, ctx->stream, n, t, a, lda, x, ldx, kchunk, part,
                           skip);
    else if (t <= 8 || sizeof(T) != 8)
        hipLaunchKernelGGL((panel_n_kernel<T, 8>), grid, dim3(256), 0, ctx->stream, n, t, a, lda, x, ldx, kchunk, part,
                           skip);
    else
        hipLaunchKernelGGL((panel_n_kernel<T, (sizeof(T) == 8 ? 16 : 8)>), grid, dim3(256), 0, ctx->stream, n, t, a,
                           lda, x, ldx, kchunk, part, skip);
}

template <typename T>
void panel_impl(Context* ctx, int trans, int n, int t, const T* a, int lda, const T* x, int ldx, T* y, int ldy)
{
    if (n == 0 || t == 0) return;
    // Wide panels are GEMMs (complex beyond 8 columns: the k-split kernel's
    // LDS holds 8 complex columns).
    if (t > 16 || (sizeof(T) != 8 && t > 8)) {
        if constexpr (sizeof(T) == 8)
            dgemm_ex(ctx, trans ? 1 : 0, n, t, n, 1.0, (const double*)a, lda, (const double*)x, ldx, 0.0, (double*)y,
                     ldy);
        else
            zgemm_ex(ctx, trans ? 2 : 0, n, t, n, make_zd(1.0, 0.0), (const zd*)a, lda, (const zd*)x,
                     ldx, make_zd(0.0, 0.0), (zd*)y, ldy);
        return;
    }
    if (!trans) {
        int ks = ksplit_for(n);
        DevBuf<T> part(ctx, (size_t)ks * t * n);
        panel_partials<T>(ctx, n, t, a, lda, x, ldx, part, ks);
        hipLaunchKernelGGL((panel_finish_kernel<T>), dim3(grid_for((size_t)n * t, 256, 1 << 30)), dim3(256), 0,
                           ctx->stream, n, t, ks, part.get(), y, ldy);
    } else {
        dim3 grid((n + 3) / 4);
        constexpr bool C = sizeof(T) != 8;
        if (t <= 2)
            hipLaunchKernelGGL((panel_t_kernel<T, 2, C>), grid, dim3(256), 0, ctx->stream, n, t, a, lda, x, ldx, y, ldy);
        else if (t <= 4)
            hipLaunchKernelGGL((panel_t_kernel<T, 4, C>), grid, dim3(256), 0, ctx->stream, n, t, a, lda, x, ldx, y, ldy);
        else if (t <= 8)
            hipLaunchKernelGGL((panel_t_kernel<T, 8, C>), grid, dim3(256), 0, ctx->stream, n, t, a, lda, x, ldx, y, ldy);
        else
            hipLaunchKernelGGL((panel_t_kernel<T, 16, C>), grid, dim3(256), 0, ctx->stream, n, t, a, lda, x, ldx, y,
                               ldy);
    }
    TNS_HIP(hipGetLastError());
}

}  // namespace

void dnorm(Context* ctx, int kind, int m, int n, const double* a, int lda, double* d_out)
{
    norm_impl<double>(ctx, kind, m, n, a, lda, d_out);
}

void znorm(Context* ctx, int kind, int m, int n, const cplx* a, int lda, double* d_out)
{
    norm_impl<zd>(ctx, kind, m, n, a, lda, d_out);
}

void dpanel(Context* ctx, int trans, int n, int t, const double* a, int lda, const double* x, int ldx, double* y,
            int ldy)
{
    panel_impl<double>(ctx, trans, n, t, a, lda, x, ldx, y, ldy);
}

void zpanel(Context* ctx, int adjoint, int n, int t, const cplx* a, int lda, const cplx* x, int ldx, cplx* y, int ldy)
{
    panel_impl<zd>(ctx, adjoint, n, t, a, lda, x, ldx, y, ldy);
}

template <typename T>
void lincomb(Context* ctx, int n, double c0, int terms, const LinTerms<T>& lt, T* out)
{
    hipLaunchKernelGGL((lincomb_kernel<T>), dim3(grid_for((size_t)n * n)), dim3(256), 0, ctx->stream, n, c0, terms, lt,
                       out);
    TNS_HIP(hipGetLastError());
}
template void lincomb<double>(Context*, int, double, int, const LinTerms<double>&, double*);
template void lincomb<zd>(Context*, int, double, int, const LinTerms<zd>&, zd*);

template <typename T>
void sub_add(Context* ctx, size_t count, const T* u, const T* v, T* den, T* num)
{
    hipLaunchKernelGGL((sub_add_kernel<T>), dim3(grid_for(count)), dim3(256), 0, ctx->stream, count, u, v, den, num);
}
template void sub_add<double>(Context*, size_t, const double*, const double*, double*, double*);
template void sub_add<zd>(Context*, size_t, const zd*, const zd*, zd*, zd*);

template <typename T>
void scale_to(Context* ctx, size_t count, const T* src, double f, T* dst)
{
    if (!count) return;
    hipLaunchKernelGGL((scale_kernel<T>), dim3(grid_for(count)), dim3(256), 0, ctx->stream, count, src, f, dst);
}
template void scale_to<double>(Context*, size_t, const double*, double, double*);
template void scale_to<zd>(Context*, size_t, const zd*, double, zd*);

template <typename T>
void mul_scalar(Context* ctx, size_t count, T* x, T s)
{
    if (!count) return;
    hipLaunchKernelGGL((mul_scalar_kernel<T>), dim3(grid_for(count)), dim3(256), 0, ctx->stream, count, x, s);
}
template void mul_scalar<double>(Context*, size_t, double*, double);
template void mul_scalar<zd>(Context*, size_t, zd*, zd);

template <typename T>
void shift_diag(Context* ctx, int n, T* a, int lda, T delta)
{
    hipLaunchKernelGGL((shift_diag_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, a, lda, delta);
}
template void shift_diag<double>(Context*, int, double*, int, double);
template void shift_diag<zd>(Context*, int, zd*, int, zd);

template <typename T>
void trace(Context* ctx, int n, const T* a, int lda, T* d_out)
{
    hipLaunchKernelGGL((trace_kernel<T>), dim3(1), dim3(1024), 0, ctx->stream, n, a, lda, d_out);
}
template void trace<double>(Context*, int, const double*, int, double*);
template void trace<zd>(Context*, int, const zd*, int, zd*);

template <typename T>
void magnitudes(Context* ctx, size_t count, const T* a, double* out)
{
    hipLaunchKernelGGL((abs_kernel<T>), dim3(grid_for(count)), dim3(256), 0, ctx->stream, count, a, out);
}
template void magnitudes<double>(Context*, size_t, const double*, double*);
template void magnitudes<zd>(Context*, size_t, const zd*, double*);

void rowvec(Context* ctx, int n, const double* mags, const double* row, double* next)
{
    hipLaunchKernelGGL(rowvec_kernel, dim3((n + 3) / 4), dim3(256), 0, ctx->stream, n, mags, row, next);
}

void fill(Context* ctx, int n, double* v, double value)
{
    hipLaunchKernelGGL(fill_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, v, value);
}

void vec_max(Context* ctx, int n, const double* v, double* d_out)
{
    hipLaunchKernelGGL(vec_finish_kernel, dim3(1), dim3(256), 0, ctx->stream, n, v, 0, d_out);
}

void zero_scalars(Context* ctx, double* d, int count)
{
    hipLaunchKernelGGL(zero_doubles_kernel, dim3(1), dim3(64), 0, ctx->stream, d, count);
}

template <typename T>
int panel_ksplit(int n)
{
    return ksplit_for(n);
}
template int panel_ksplit<double>(int);
template int panel_ksplit<zd>(int);

template <typename T>
void panel_n_partials(Context* ctx, int n, int t, const T* a, int lda, const T* x, int ldx, T* part, int ks,
                      const int* skip)
{
    panel_partials<T>(ctx, n, t, a, lda, x, ldx, part, ks, skip);
}
template void panel_n_partials<double>(Context*, int, int, const double*, int, const double*, int, double*, int,
                                       const int*);
template void panel_n_partials<zd>(Context*, int, int, const zd*, int, const zd*, int, zd*, int, const int*);

void taylor_exit(Context* ctx, double* d_norms, int* d_state, double unit)
{
    hipLaunchKernelGGL(taylor_exit_kernel, dim3(1), dim3(1), 0, ctx->stream, d_norms, d_state, unit);
}

template <typename T>
void taylor_step(Context* ctx, int n, int nb, int ks, const T* part, double coef, T* work, T* f, double* d_norms,
                 const int* skip)
{
    hipLaunchKernelGGL((taylor_step_kernel<T>), dim3((n + 31) / 32), dim3(256), 0, ctx->stream, n, nb, ks, part,
                       coef, work, f, d_norms, skip);
}
template void taylor_step<double>(Context*, int, int, int, const double*, double, double*, double*, double*,
                                  const int*);
template void taylor_step<zd>(Context*, int, int, int, const zd*, double, zd*, zd*, double*, const int*);

template <typename T>
void panel_infnorm(Context* ctx, int n, int nb, const T* x, double* d_out)
{
    hipLaunchKernelGGL((panel_infnorm_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, nb, x, d_out);
}
template void panel_infnorm<double>(Context*, int, int, const double*, double*);
template void panel_infnorm<zd>(Context*, int, int, const zd*, double*);

}  // namespace tns
