// Mixed-precision solve A X = B: power-of-two row and column equilibration
// (when the scales spread widely), a random butterfly transformation, the
// result rounded to FP32 and
// factored in FP32 with the pivots of each block step taken inside its
// 64 x 64 diagonal block (products on the FP32 matrix cores, or on the BF16
// ones with each operand split into two bfloat16 parts), then
// iterative refinement in FP64 — residuals r = b - A x
// and the updates of x in FP64, corrections by FGMRES right-preconditioned
// with the low-precision factors — until the normwise backward error
// max_j ||r_j||inf / (||A||inf ||x_j||inf + ||b_j||inf) is at most the
// tolerance. If a restart cycle fails to halve it, the step budget runs
// out, the residuals shrink too slowly to beat the FP64 solve's cost, or a
// pivot is exactly zero in low precision, the solve falls back to the FP64
// LU with partial pivoting of lu.hip.
//
// The randomization (Parker, "Random butterfly transformations with
// applications in computational linear algebra", UCLA CSD-950023, 1995;
// Baboulin, Dongarra, Herrmann, Tomov, "Accelerating linear system
// solutions using randomization techniques", ACM TOMS 39(2), 2013):
// A_r = U^T A V with U, V recursive butterflies of depth 2, whose diagonal
// factors are exp(r / 10), r uniform on [-1/2, 1/2). With probability close
// to one A_r needs no pivoting across blocks, so no pivot search spans a
// whole column: each panel is a 64 x 64 diagonal block factored with
// partial pivoting inside it and a triangular solve for the rows below,
// and the time goes to the trailing products. Any growth this leaves is
// the refinement's to remove, and the FP64 fallback bounds the worst case.
// The transformed system is A_r (V^-1 x) = U^T b, so x = V y.
//
// The reference has no mixed-precision path; this one is pinned by the
// backward error of its result and by the FP64 solve it falls back to
// (Lu.cs:189-378).
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <vector>

#include "algos.hpp"
#include "mixed.hpp"

namespace tns {
namespace {

// Power-of-two scale 2^-e with |x| = f 2^e, f in [0.5, 1): exact to apply.
__device__ __forceinline__ double pow2_inverse(double x)
{
    if (!(x > 0.0) || isinf(x)) return 1.0;
    int e;
    frexp(x, &e);
    return ldexp(1.0, -e);
}

// Row maxima of |a_ij| in two passes: workgroup (x, y) takes rows x*256 ..
// and the columns of chunk y, and folds its partial maximum into rmax[i]
// (bits of a non-negative double order as unsigned integers; NaNs are
// skipped, as fmax skips them); then rs[i] = 2^-e for the row's maximum.
constexpr int ROW_CHUNK = 64;

__global__ void __launch_bounds__(256) row_max_kernel(int n, const double* __restrict__ a, int lda,
                                                      unsigned long long* __restrict__ rmax)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int j0 = blockIdx.y * ROW_CHUNK, j1 = j0 + ROW_CHUNK < n ? j0 + ROW_CHUNK : n;
    double mx = 0.0;
    for (int j = j0; j < j1; j++) mx = fmax(mx, fabs(a[i + (size_t)j * lda]));
    atomicMax(rmax + i, (unsigned long long)__double_as_longlong(mx));
}

__global__ void row_scale_kernel(int n, const unsigned long long* __restrict__ rmax, double* __restrict__ rs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rs[i] = pow2_inverse(__longlong_as_double((long long)rmax[i]));
}

// cs[j] = 2^-e for the largest |rs_i a_ij| of column j (a block per column).
__global__ void __launch_bounds__(256) col_scale_kernel(int n, const double* __restrict__ a, int lda,
                                                        const double* __restrict__ rs, double* __restrict__ cs)
{
    __shared__ double red[4];
    const int j = blockIdx.x;
    double mx = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) mx = fmax(mx, fabs(rs[i] * a[i + (size_t)j * lda]));
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) cs[j] = pow2_inverse(fmax(fmax(red[0], red[1]), fmax(red[2], red[3])));
}


// Per column j of an n x t matrix: out[j] = max_i |x_ij| (a block per column).
__global__ void __launch_bounds__(256) col_absmax_kernel(int n, const double* __restrict__ x, int ldx,
                                                         double* __restrict__ out)
{
    __shared__ double red[4];
    const int j = blockIdx.x;
    double mx = 0.0;
    bool nan = false;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double v = fabs(x[i + (size_t)j * ldx]);
        nan |= v != v;
        mx = fmax(mx, v);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        mx = fmax(mx, __shfl_xor(mx, off, 64));
        nan |= __shfl_xor((int)nan, off, 64) != 0;
    }
    __shared__ int rnan[4];
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6] = mx;
        rnan[threadIdx.x >> 6] = nan;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const bool any = rnan[0] | rnan[1] | rnan[2] | rnan[3];
        out[j] = any ? NAN : fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    }
}



// A butterfly of order n4 = 4q, depth 2, stored as its diagonal factors:
// level 1 (order n4) R0[0,h) R1[0,h); level 2 (two of order h) R0 and R1 of
// the upper one, then of the lower one, q each: 2 n4 values. For a quad of
// rows i0 + {0, q, 2q, 3q}:
//   rbt_t: the transposed butterfly, level 1 then level 2 (U^T x, and the
//          column side of A V, which has the same form);
//   rbt_n: the butterfly itself, level 2 then level 1 (V y).
// Both leave out the factor 1/2 of the two levels' 1/sqrt(2).
__device__ __forceinline__ void rbt_t(const double* __restrict__ b, int q, int i0, double& x0, double& x1, double& x2,
                                      double& x3)
{
    const int h = 2 * q;
    const double y0 = b[i0] * (x0 + x2), y2 = b[h + i0] * (x0 - x2);
    const double y1 = b[i0 + q] * (x1 + x3), y3 = b[h + i0 + q] * (x1 - x3);
    const double* lv = b + 2 * h;
    x0 = lv[i0] * (y0 + y1);
    x1 = lv[q + i0] * (y0 - y1);
    x2 = lv[2 * q + i0] * (y2 + y3);
    x3 = lv[3 * q + i0] * (y2 - y3);
}

__device__ __forceinline__ void rbt_n(const double* __restrict__ b, int q, int i0, double& y0, double& y1, double& y2,
                                      double& y3)
{
    const int h = 2 * q;
    const double* lv = b + 2 * h;
    const double w0 = lv[i0] * y0 + lv[q + i0] * y1, w1 = lv[i0] * y0 - lv[q + i0] * y1;
    const double w2 = lv[2 * q + i0] * y2 + lv[3 * q + i0] * y3, w3 = lv[2 * q + i0] * y2 - lv[3 * q + i0] * y3;
    y0 = b[i0] * w0 + b[h + i0] * w2;
    y2 = b[i0] * w0 - b[h + i0] * w2;
    y1 = b[i0 + q] * w1 + b[h + i0 + q] * w3;
    y3 = b[i0 + q] * w1 - b[h + i0 + q] * w3;
}

// s = float(U^T (R A C) V) of order n4 >= n, A bordered by the identity;
// rs, cs may be null (unit). Thread (i0, j0) owns the 4 x 4 entries in rows
// i0 + kq and columns j0 + lq, the only ones that mix.
__global__ void rbt_matrix_kernel(int n, int n4, const double* __restrict__ a, int lda, const double* __restrict__ rs,
                                  const double* __restrict__ cs, const double* __restrict__ ub,
                                  const double* __restrict__ vb, float* __restrict__ s, int lds)
{
    const int q = n4 / 4;
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y;
    if (i0 >= q) return;
    double m[4][4];
#pragma unroll
    for (int l = 0; l < 4; l++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int r = i0 + k * q, c = j0 + l * q;
            double v = r == c ? 1.0 : 0.0;
            if (r < n && c < n) {
                v = a[r + (size_t)c * lda];
                if (rs) v *= rs[r] * cs[c];
            }
            m[k][l] = v;
        }
#pragma unroll
    for (int l = 0; l < 4; l++) rbt_t(ub, q, i0, m[0][l], m[1][l], m[2][l], m[3][l]);
#pragma unroll
    for (int k = 0; k < 4; k++) rbt_t(vb, q, j0, m[k][0], m[k][1], m[k][2], m[k][3]);
#pragma unroll
    for (int l = 0; l < 4; l++)
#pragma unroll
        for (int k = 0; k < 4; k++) s[i0 + k * q + (size_t)(j0 + l * q) * lds] = (float)(0.25 * m[k][l]);
}

// d = float(U^T (R r) / sc_j) of order n4 (r zero below row n), column j.
__global__ void rbt_rhs_kernel(int n, int n4, const double* __restrict__ r, int ldr, const double* __restrict__ rs,
                               const double* __restrict__ sc, const double* __restrict__ ub, float* __restrict__ d)
{
    const int q = n4 / 4;
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i0 >= q) return;
    double x[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int row = i0 + k * q;
        double v = 0.0;
        if (row < n) {
            v = r[row + (size_t)j * ldr];
            if (rs) v *= rs[row];
        }
        x[k] = v;
    }
    rbt_t(ub, q, i0, x[0], x[1], x[2], x[3]);
    const double scale = 0.5 * (sc[j] > 0.0 ? 1.0 / sc[j] : 1.0);
#pragma unroll
    for (int k = 0; k < 4; k++) d[i0 + k * q + (size_t)j * n4] = (float)(x[k] * scale);
}

// x[0,n) += C V y * sc_j, y = d (order n4), column j.
__global__ void rbt_accumulate_kernel(int n, int n4, const float* __restrict__ d, const double* __restrict__ vb,
                                      const double* __restrict__ cs, const double* __restrict__ sc,
                                      double* __restrict__ x, int ldx)
{
    const int q = n4 / 4;
    const int i0 = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    if (i0 >= q) return;
    double y[4];
#pragma unroll
    for (int k = 0; k < 4; k++) y[k] = (double)d[i0 + k * q + (size_t)j * n4];
    rbt_n(vb, q, i0, y[0], y[1], y[2], y[3]);
    const double scale = 0.5 * (sc[j] > 0.0 ? sc[j] : 1.0);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int row = i0 + k * q;
        if (row >= n) continue;
        double v = y[k] * scale;
        if (cs) v *= cs[row];
        x[row + (size_t)j * ldx] += v;
    }
}

int grid_for(size_t total)
{
    const size_t b = (total + 255) / 256;
    return (int)(b < 8192 ? b : 8192);
}

// Krylov blocks: `count` n x t FP64 matrices with leading dimension n, vs
// apart. Workgroup (i, j): h[i + j ldh] = <v_i[:, j], w[:, j]>, one fixed
// summation order.
__global__ void __launch_bounds__(256) block_dots_kernel(int n, const double* __restrict__ v, size_t vs,
                                                         const double* __restrict__ w, int ldw, double* __restrict__ h,
                                                         int ldh)
{
    __shared__ double red[4];
    const int i = blockIdx.x, j = blockIdx.y;
    const double* vi = v + i * vs + (size_t)j * n;
    const double* wj = w + (size_t)j * ldw;
    double s = 0.0;
    for (int row = threadIdx.x; row < n; row += 256) s += vi[row] * wj[row];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) h[i + (size_t)j * ldh] = (red[0] + red[1]) + (red[2] + red[3]);
}
// w[:, j] += sign * sum_{i < count} v_i[:, j] c[i + j ldc] (sign +1 or -1).
__global__ void __launch_bounds__(256) block_combine_kernel(int n, const double* __restrict__ v, size_t vs, int count,
                                                            const double* __restrict__ c, int ldc, int sign,
                                                            double* __restrict__ w, int ldw)
{
    const int row = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (row >= n) return;
    double acc = 0.0;
    for (int i = 0; i < count; i++) acc += v[i * vs + (size_t)j * n + row] * c[i + (size_t)j * ldc];
    double* p = w + row + (size_t)j * ldw;
    *p = sign > 0 ? *p + acc : *p - acc;
}
// w[:, j] *= s[j].
__global__ void __launch_bounds__(256) col_mul_kernel(int n, double* __restrict__ w, int ldw,
                                                      const double* __restrict__ s)
{
    const int row = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y;
    if (row < n) w[row + (size_t)j * ldw] *= s[j];
}

// The low-precision solve z := M^-1 v for t columns: M^-1 = C V (LU)^-1 U^T R
// applied through the FP32 factors of the transformed matrix, each column
// scaled by its largest entry on the way into FP32.
struct LowSolve {
    Context* ctx;
    int n, n4, t;
    const float* s;
    const int* perm;
    const float* inv;
    const double *rs, *cs, *ub, *vb;
    float* d;    // n4 x t
    double* sc;  // t
    // accumulate: z += M^-1 v; otherwise z := M^-1 v.
    void apply(const double* v, int ldv, double* z, int ldz, bool accumulate) const
    {
        const int q = n4 / 4;
        const dim3 cols((q + 255) / 256, t);
        hipLaunchKernelGGL(col_absmax_kernel, dim3(t), dim3(256), 0, ctx->stream, n, v, ldv, sc);
        hipLaunchKernelGGL(rbt_rhs_kernel, cols, dim3(256), 0, ctx->stream, n, n4, v, ldv, rs, sc, ub, d);
        const int saved = ctx->lp_mode;
        ctx->lp_mode = LP_FP32;  // the solves' products in FP32 in both precisions
        try {
            sgetrs_local(ctx, n4, t, s, n4, perm, inv, d, n4);
        } catch (...) {
            ctx->lp_mode = saved;
            throw;
        }
        ctx->lp_mode = saved;
        if (!accumulate) TNS_HIP(hipMemset2DAsync(z, sizeof(double) * (size_t)ldz, 0, sizeof(double) * (size_t)n, t,
                                                  ctx->stream));
        hipLaunchKernelGGL(rbt_accumulate_kernel, cols, dim3(256), 0, ctx->stream, n, n4, d, vb, cs, sc, z, ldz);
        TNS_HIP(hipGetLastError());
    }
};

// One restart cycle of FGMRES (Saad, "A flexible inner-outer preconditioned
// GMRES algorithm", SIAM J. Sci. Comput. 14(2), 1993) on A e = r, right-
// preconditioned by the low-precision solve, every column its own Krylov
// space, all columns in step: z_k = M^-1 v_k, w = A z_k (FP64), w
// orthogonalized against v_0..v_k by classical Gram-Schmidt applied twice,
// the small least-squares problems by Givens rotations on the host; then
// x += Z y. Because e is taken from the span of the z_k actually computed,
// the low-precision solves may differ from one step to the next, and the
// residual is minimized for the true A in FP64. A column stops when its
// residual estimate |g_{k+1}| is at most target[j] or its Krylov space is
// exhausted. Returns the steps (low-precision solves) taken, at most m.
//
// From the fourth step on, the cycle gives up (hopeless = true, x left as
// it was) when the worst column's residual-to-target ratio, shrinking at
// its rate over the last three steps, would take longer to reach 1 at the
// measured time per step than fallback_ms, the estimated cost of the FP64
// solve that replaces it.
int fgmres_cycle(const LowSolve& M, const double* a, int lda, const double* r, double* x, int ldx, int m,
                 const std::vector<double>& target, double* V, double* Z, double* d_h, double fallback_ms,
                 bool* hopeless)
{
    using clock = std::chrono::steady_clock;
    const auto t_start = clock::now();
    *hopeless = false;
    std::vector<double> ratio;
    Context* ctx = M.ctx;
    const int n = M.n, t = M.t, ldh = m + 1;
    const size_t vs = (size_t)n * t;
    const dim3 rows((n + 255) / 256, t);
    // d_h: h1 and h2 ((m+1) x t each: the two Gram-Schmidt passes), the
    // squared norms (t), the scales (t), y (m x t).
    double* h1 = d_h;
    double* h2 = h1 + (size_t)ldh * t;
    double* hn = h2 + (size_t)ldh * t;
    double* dsc = hn + t;
    double* dy = dsc + t;
    const size_t back = 2 * (size_t)ldh * t + t;  // read back per step

    // v_0 = r / ||r||.
    dcopy2d(ctx, n, t, r, n, V, n);
    hipLaunchKernelGGL(block_dots_kernel, dim3(1, t), dim3(256), 0, ctx->stream, n, V, vs, V, n, hn, 1);
    std::vector<double> hb(back), scale(t);
    read_doubles(ctx, hn, hb.data(), t);
    std::vector<double> R((size_t)ldh * m * t, 0.0), cr((size_t)m * t), sr((size_t)m * t), g((size_t)ldh * t, 0.0);
    std::vector<int> basis(t, 0);
    std::vector<char> active(t);
    for (int j = 0; j < t; j++) {
        const double beta = std::sqrt(hb[j]);
        active[j] = beta > target[j] && beta == beta;
        g[(size_t)j * ldh] = beta;
        scale[j] = active[j] ? 1.0 / beta : 0.0;
    }
    int steps = 0;
    for (int k = 0; k < m; k++) {
        bool any = false;
        for (int j = 0; j < t; j++) any |= active[j] != 0;
        if (!any) break;
        double* vk = V + (size_t)k * vs;
        double* w = V + (size_t)(k + 1) * vs;
        double* zk = Z + (size_t)k * vs;
        write_pinned(ctx, dsc, scale.data(), sizeof(double) * t);
        hipLaunchKernelGGL(col_mul_kernel, rows, dim3(256), 0, ctx->stream, n, vk, n, dsc);
        M.apply(vk, n, zk, n, false);
        dgemm_ex(ctx, 0, n, t, n, 1.0, a, lda, zk, n, 0.0, w, n);
        hipLaunchKernelGGL(block_dots_kernel, dim3(k + 1, t), dim3(256), 0, ctx->stream, n, V, vs, w, n, h1, ldh);
        hipLaunchKernelGGL(block_combine_kernel, rows, dim3(256), 0, ctx->stream, n, V, vs, k + 1, h1, ldh, -1, w, n);
        hipLaunchKernelGGL(block_dots_kernel, dim3(k + 1, t), dim3(256), 0, ctx->stream, n, V, vs, w, n, h2, ldh);
        hipLaunchKernelGGL(block_combine_kernel, rows, dim3(256), 0, ctx->stream, n, V, vs, k + 1, h2, ldh, -1, w, n);
        hipLaunchKernelGGL(block_dots_kernel, dim3(1, t), dim3(256), 0, ctx->stream, n, w, vs, w, n, hn, 1);
        TNS_HIP(hipGetLastError());
        read_doubles(ctx, h1, hb.data(), (int)back);
        steps++;
        double worst = 0.0;
        for (int j = 0; j < t; j++) {
            if (!active[j]) {
                scale[j] = 0.0;
                continue;
            }
            double* col = R.data() + ((size_t)j * m + k) * ldh;
            for (int i = 0; i <= k; i++) col[i] = hb[i + (size_t)j * ldh] + hb[(size_t)ldh * t + i + (size_t)j * ldh];
            const double sub = std::sqrt(hb[2 * (size_t)ldh * t + j]);
            double* cj = cr.data() + (size_t)j * m;
            double* sj = sr.data() + (size_t)j * m;
            for (int i = 0; i < k; i++) {
                const double u = cj[i] * col[i] + sj[i] * col[i + 1];
                col[i + 1] = -sj[i] * col[i] + cj[i] * col[i + 1];
                col[i] = u;
            }
            const double den = std::hypot(col[k], sub);
            if (!(den > 0.0)) {  // no progress possible in this column
                active[j] = 0;
                scale[j] = 0.0;
                continue;
            }
            cj[k] = col[k] / den;
            sj[k] = sub / den;
            col[k] = den;
            double* gj = g.data() + (size_t)j * ldh;
            gj[k + 1] = -sj[k] * gj[k];
            gj[k] = cj[k] * gj[k];
            basis[j] = k + 1;
            const bool done = !(std::fabs(gj[k + 1]) > target[j]) || !(sub > 0.0);
            if (!done) worst = std::max(worst, std::fabs(gj[k + 1]) / target[j]);
            active[j] = !done;
            scale[j] = done ? 0.0 : 1.0 / sub;
        }
        ratio.push_back(worst);
        if (k >= 3 && worst > 1.0) {
            const double rate = std::pow(worst / ratio[k - 3], 1.0 / 3.0);
            const double need = rate < 1.0 ? std::log(worst) / -std::log(rate) : INFINITY;
            const double per_step = std::chrono::duration<double, std::milli>(clock::now() - t_start).count() / steps;
            if (!(need * per_step <= fallback_ms)) {
                *hopeless = true;
                return steps;
            }
        }
    }
    // y = R^-1 g per column, then x += Z y.
    int kmax = 0;
    std::vector<double> y((size_t)m * t, 0.0);
    for (int j = 0; j < t; j++) {
        const int kb = basis[j];
        kmax = std::max(kmax, kb);
        const double* Rj = R.data() + (size_t)j * m * ldh;
        double* yj = y.data() + (size_t)j * m;
        for (int i = kb - 1; i >= 0; i--) {
            double s = g[i + (size_t)j * ldh];
            for (int l = i + 1; l < kb; l++) s -= Rj[i + (size_t)l * ldh] * yj[l];
            yj[i] = s / Rj[i + (size_t)i * ldh];
        }
    }
    if (kmax > 0) {
        write_pinned(ctx, dy, y.data(), sizeof(double) * y.size());
        hipLaunchKernelGGL(block_combine_kernel, rows, dim3(256), 0, ctx->stream, n, Z, vs, kmax, dy, m, 1, x, ldx);
        TNS_HIP(hipGetLastError());
    }
    return steps;
}

// Backward error of the current x: r := b - A x, then
// max_j ||r_j|| / (||A|| ||x_j|| + ||b_j||). NaN if any part is NaN.
double backward_error(Context* ctx, int n, int t, const double* a, int lda, const double* b, int ldb, const double* x,
                      int ldx, double* r, double anorm, const std::vector<double>& bnorm, double* d_tmp,
                      std::vector<double>* xnorm = nullptr)
{
    dcopy2d(ctx, n, t, b, ldb, r, n);
    dgemm_ex(ctx, 0, n, t, n, -1.0, a, lda, x, ldx, 1.0, r, n);
    hipLaunchKernelGGL(col_absmax_kernel, dim3(t), dim3(256), 0, ctx->stream, n, r, n, d_tmp);
    hipLaunchKernelGGL(col_absmax_kernel, dim3(t), dim3(256), 0, ctx->stream, n, x, ldx, d_tmp + t);
    TNS_HIP(hipGetLastError());
    std::vector<double> h(2 * (size_t)t);
    read_doubles(ctx, d_tmp, h.data(), 2 * t);
    if (xnorm) xnorm->assign(h.begin() + t, h.end());
    double worst = 0.0;
    for (int j = 0; j < t; j++) {
        const double den = anorm * h[t + j] + bnorm[j];
        const double e = den > 0.0 ? h[j] / den : (h[j] > 0.0 ? INFINITY : 0.0);
        if (e != e) return NAN;
        worst = std::max(worst, e);
    }
    return worst;
}

}  // namespace

void dgesv_mixed(Context* ctx, int n, int t, const double* a, int lda, const double* b, int ldb, double* x, int ldx,
                 int precision, double tol, int max_iter, MixedInfo* info)
{
    using clock = std::chrono::steady_clock;
    MixedInfo out;
    out.precision = precision;
    if (n <= 0 || t <= 0) {
        if (info) *info = out;
        return;
    }
    const auto t0 = clock::now();
    // BF16: the three-part products (operands split into two bfloat16
    // parts): single-part BF16 products leave the block-locally pivoted
    // factors too coarse for the refinement to contract.
    const int mode = precision == 16 ? LP_BF16X3 : LP_FP32;

    DevBuf<double> d_tmp(ctx, 2 * (size_t)t + 8);
    DevBuf<double> r(ctx, (size_t)n * t);
    // ||A||inf and ||b_j||inf.
    dnorm(ctx, 'I', n, n, a, lda, d_tmp.get());
    const double anorm = read_double(ctx, d_tmp.get());
    hipLaunchKernelGGL(col_absmax_kernel, dim3(t), dim3(256), 0, ctx->stream, n, b, ldb, d_tmp.get());
    std::vector<double> bnorm(t);
    read_doubles(ctx, d_tmp.get(), bnorm.data(), t);

    bool converged = false, singular = false;
    int iters = 0;
    double berr = INFINITY;
    {
        // Row and column equilibration by powers of two (R A C is exact),
        // when the rows' or the columns' scales spread over more than 2^8:
        // the butterflies mix rows and columns four at a time, which on a
        // badly scaled matrix buries the small ones in the rounding of the
        // large. Milder spreads are left alone: there equilibration only
        // moves the factorization's error onto the rows that weigh least in
        // the normwise backward error, and slows the refinement.
        DevBuf<double> rs(ctx, n), cs(ctx, n);
        {
            DevBuf<unsigned long long> rmax(ctx, n);
            TNS_HIP(hipMemsetAsync(rmax.get(), 0, sizeof(unsigned long long) * n, ctx->stream));
            hipLaunchKernelGGL(row_max_kernel, dim3((n + 255) / 256, (n + ROW_CHUNK - 1) / ROW_CHUNK), dim3(256), 0,
                               ctx->stream, n, a, lda, rmax.get());
            hipLaunchKernelGGL(row_scale_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, rmax.get(),
                               rs.get());
        }
        hipLaunchKernelGGL(col_scale_kernel, dim3(n), dim3(256), 0, ctx->stream, n, a, lda, rs.get(), cs.get());
        std::vector<double> hs(2 * (size_t)n);
        read_doubles(ctx, rs.get(), hs.data(), n);
        read_doubles(ctx, cs.get(), hs.data() + n, n);
        auto spread = [](const double* v, int len) {
            const auto mm = std::minmax_element(v, v + len);
            return *mm.second / *mm.first;
        };
        const bool equilibrate = std::max(spread(hs.data(), n), spread(hs.data() + n, n)) > 256.0;
        const double* prs = equilibrate ? rs.get() : nullptr;
        const double* pcs = equilibrate ? cs.get() : nullptr;
        // The randomized matrix, order n4 >= n (the border is the
        // identity), a multiple of the factorization's block.
        const int blk = local_lu_block();
        const int n4 = (n + blk - 1) / blk * blk, q = n4 / 4;
        std::vector<double> hb(4 * (size_t)n4);
        SeededRandom rng(20261203);
        for (double& v : hb) v = std::exp((rng.sample() - 0.5) / 10.0);
        DevBuf<double> bf(ctx, hb.size());
        TNS_HIP(hipMemcpyAsync(bf.get(), hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice, ctx->stream));
        const double* ub = bf.get();
        const double* vb = bf.get() + 2 * (size_t)n4;
        DevBuf<float> s(ctx, (size_t)n4 * n4);
        const dim3 quads((q + 255) / 256, q);
        hipLaunchKernelGGL(rbt_matrix_kernel, quads, dim3(256), 0, ctx->stream, n, n4, a, lda, prs, pcs, ub, vb, s.get(),
                           n4);
        TNS_HIP(hipGetLastError());
        DevBuf<int> piv(ctx, n4), perm(ctx, n4), dinfo(ctx, 1);
        DevBuf<float> inv(ctx, (size_t)(n4 / blk) * 2 * blk * blk);
        {
            // The factorization's products in the requested precision.
            struct Mode {
                Context* c;
                int saved;
                ~Mode() { c->lp_mode = saved; }
            } restore{ctx, ctx->lp_mode};
            ctx->lp_mode = mode;
            sgetrf_local(ctx, n4, s.get(), n4, piv.get(), perm.get(), inv.get(), dinfo.get());
        }
        int hinfo = -1;
        read_back(ctx, &hinfo, dinfo.get(), sizeof(int));
        out.factor_ms = std::chrono::duration<double, std::milli>(clock::now() - t0).count();
        singular = hinfo >= 0;

        if (!singular) {
            // x = M^-1 b by the low-precision solve; then, while the
            // backward error is above tol, FGMRES restart cycles on
            // A e = b - A x (residuals and updates in FP64), each of which
            // must at least halve the backward error.
            DevBuf<float> d(ctx, (size_t)n4 * t);
            DevBuf<double> sc(ctx, t);
            const LowSolve M{ctx, n, n4, t, s.get(), perm.get(), inv.get(), prs, pcs, ub, vb, d.get(), sc.get()};
            TNS_HIP(hipMemset2DAsync(x, sizeof(double) * (size_t)ldx, 0, sizeof(double) * (size_t)n, t, ctx->stream));
            M.apply(b, ldb, x, ldx, true);
            iters = 1;
            std::vector<double> xnorm;
            berr = backward_error(ctx, n, t, a, lda, b, ldb, x, ldx, r.get(), anorm, bnorm, d_tmp.get(), &xnorm);
            out.first_error = berr;
            converged = berr <= tol;
            if (!converged && iters < max_iter && berr == berr) {
                // Restart length: at most 30 steps, and the Krylov blocks
                // (2m + 1 of n x t) within 2 GiB.
                const double blocks = 2147483648.0 / (8.0 * n * t);
                const int mcap = std::max(2, std::min(30, (int)((blocks - 1.0) / 2.0)));
                const int mall = std::min(mcap, max_iter - iters);
                DevBuf<double> V(ctx, (size_t)(mall + 1) * n * t), Z(ctx, (size_t)mall * n * t);
                DevBuf<double> h(ctx, (size_t)(2 * (mall + 1) + 2 + mall) * t);
                std::vector<double> target(t);
                // The FP64 LU and solve take about 2.5 times as long as the
                // low-precision factorization (measured at n = 8192: 48 ms
                // against 17.5 ms); refinement steps that would cost more
                // are not taken.
                const double fallback_ms = 2.5 * out.factor_ms;
                double prev = berr;
                while (iters < max_iter) {
                    const int m = std::min(mall, max_iter - iters);
                    for (int j = 0; j < t; j++) target[j] = 0.5 * tol * (anorm * xnorm[j] + bnorm[j]);
                    bool hopeless = false;
                    iters += fgmres_cycle(M, a, lda, r.get(), x, ldx, m, target, V.get(), Z.get(), h.get(),
                                          fallback_ms, &hopeless);
                    if (hopeless) break;
                    berr = backward_error(ctx, n, t, a, lda, b, ldb, x, ldx, r.get(), anorm, bnorm, d_tmp.get(), &xnorm);
                    if (berr <= tol) {
                        converged = true;
                        break;
                    }
                    if (!(berr <= 0.5 * prev)) break;  // stalled, or NaN
                    prev = berr;
                }
            }
        }
    }
    out.iterations = iters;
    out.refine_ms = std::chrono::duration<double, std::milli>(clock::now() - t0).count() - out.factor_ms;

    if (!converged) {
        // Fall back to the FP64 LU and solve.
        const auto t1 = clock::now();
        DevBuf<double> f(ctx, (size_t)n * n);
        dcopy2d(ctx, n, n, a, lda, f.get(), n);
        DevBuf<int> piv(ctx, n), perm(ctx, n), dinfo(ctx, 1);
        dgetrf(ctx, n, n, f.get(), n, 0, piv.get(), perm.get(), dinfo.get());
        int hinfo = -1;
        read_back(ctx, &hinfo, dinfo.get(), sizeof(int));
        if (hinfo == LU_ABORTED) fail(TNS_EDEVICE, "the FP64 fallback's panel exchange timed out");
        if (hinfo >= 0) fail(TNS_ESTATE, "the matrix is singular: the FP64 fallback met a zero pivot");
        dcopy2d(ctx, n, t, b, ldb, x, ldx);
        dgetrs(ctx, 0, n, t, f.get(), n, perm.get(), x, ldx);
        berr = backward_error(ctx, n, t, a, lda, b, ldb, x, ldx, r.get(), anorm, bnorm, d_tmp.get());
        out.fell_back = 1;
        out.fallback_ms = std::chrono::duration<double, std::milli>(clock::now() - t1).count();
    }
    out.backward_error = berr;
    if (info) *info = out;
}

}  // namespace tns
