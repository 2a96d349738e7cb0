// Mixed-precision solve A X = B: LU in FP32 (products on the FP32 matrix
// cores, or on the BF16 ones after row and column equilibration), then
// iterative refinement in FP64 — residuals r = b - A x and the updates of
// x in FP64, corrections from the low-precision factors — until the
// normwise backward error max_j ||r_j||inf / (||A||inf ||x_j||inf +
// ||b_j||inf) is at most the tolerance. If a step fails to halve it, or a
// pivot is exactly zero in low precision, the solve falls back to the FP64
// LU of lu.hip.
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

// rs[i] = 2^-e for the largest |a_ij| of row i (one thread per row).
__global__ void row_scale_kernel(int n, const double* __restrict__ a, int lda, double* __restrict__ rs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double mx = 0.0;
    for (int j = 0; j < n; j++) mx = fmax(mx, fabs(a[i + (size_t)j * lda]));
    rs[i] = pow2_inverse(mx);
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

// s[i + j*n] = float(rs_i * a_ij * cs_j); rs, cs may be null (unit).
__global__ void to_single_kernel(int n, const double* __restrict__ a, int lda, const double* __restrict__ rs,
                                 const double* __restrict__ cs, float* __restrict__ s)
{
    const size_t total = (size_t)n * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % n), j = (int)(e / n);
        double v = a[i + (size_t)j * lda];
        if (rs) v *= rs[i] * cs[j];
        s[e] = (float)v;
    }
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

// The low-precision right-hand side of the correction: d[i,j] =
// float(rs_i * r_ij / sc_j) with sc_j = max_i |r_ij| (so the single-precision
// solve sees unit-sized data).
__global__ void rhs_to_single_kernel(int n, int t, const double* __restrict__ r, int ldr,
                                     const double* __restrict__ rs, const double* __restrict__ sc,
                                     float* __restrict__ d)
{
    const size_t total = (size_t)n * t;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % n), j = (int)(e / n);
        const double s = sc[j] > 0.0 ? 1.0 / sc[j] : 1.0;
        double v = r[i + (size_t)j * ldr] * s;
        if (rs) v *= rs[i];
        d[e] = (float)v;
    }
}

// x[i,j] += cs_i * double(d[i,j]) * sc_j.
__global__ void accumulate_kernel(int n, int t, const float* __restrict__ d, const double* __restrict__ cs,
                                  const double* __restrict__ sc, double* __restrict__ x, int ldx)
{
    const size_t total = (size_t)n * t;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % n), j = (int)(e / n);
        double v = (double)d[e] * (sc[j] > 0.0 ? sc[j] : 1.0);
        if (cs) v *= cs[i];
        x[i + (size_t)j * ldx] += v;
    }
}

int grid_for(size_t total)
{
    const size_t b = (total + 255) / 256;
    return (int)(b < 8192 ? b : 8192);
}

// Backward error of the current x: r := b - A x, then
// max_j ||r_j|| / (||A|| ||x_j|| + ||b_j||). NaN if any part is NaN.
double backward_error(Context* ctx, int n, int t, const double* a, int lda, const double* b, int ldb, const double* x,
                      int ldx, double* r, double anorm, const std::vector<double>& bnorm, double* d_tmp)
{
    dcopy2d(ctx, n, t, b, ldb, r, n);
    dgemm_ex(ctx, 0, n, t, n, -1.0, a, lda, x, ldx, 1.0, r, n);
    hipLaunchKernelGGL(col_absmax_kernel, dim3(t), dim3(256), 0, ctx->stream, n, r, n, d_tmp);
    hipLaunchKernelGGL(col_absmax_kernel, dim3(t), dim3(256), 0, ctx->stream, n, x, ldx, d_tmp + t);
    TNS_HIP(hipGetLastError());
    std::vector<double> h(2 * (size_t)t);
    read_doubles(ctx, d_tmp, h.data(), 2 * t);
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
    const int mode = precision == 16 ? LP_BF16 : LP_FP32;

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
        // Equilibration for BF16: powers of two, so R A C is exact.
        DevBuf<double> rs, cs;
        const double* prs = nullptr;
        const double* pcs = nullptr;
        if (mode == LP_BF16) {
            rs = DevBuf<double>(ctx, n);
            cs = DevBuf<double>(ctx, n);
            hipLaunchKernelGGL(row_scale_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, a, lda, rs.get());
            hipLaunchKernelGGL(col_scale_kernel, dim3(n), dim3(256), 0, ctx->stream, n, a, lda, rs.get(), cs.get());
            prs = rs.get();
            pcs = cs.get();
        }
        DevBuf<float> s(ctx, (size_t)n * n);
        hipLaunchKernelGGL(to_single_kernel, dim3(grid_for((size_t)n * n)), dim3(256), 0, ctx->stream, n, a, lda, prs,
                           pcs, s.get());
        DevBuf<int> piv(ctx, n), perm(ctx, n), dinfo(ctx, 1);
        const int saved = ctx->lp_mode;
        ctx->lp_mode = mode;
        try {
            sgetrf(ctx, n, n, s.get(), n, 0, piv.get(), perm.get(), dinfo.get());
        } catch (...) {
            ctx->lp_mode = saved;
            throw;
        }
        int hinfo = -1;
        TNS_HIP(hipMemcpyAsync(&hinfo, dinfo.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        TNS_HIP(hipStreamSynchronize(ctx->stream));
        out.factor_ms = std::chrono::duration<double, std::milli>(clock::now() - t0).count();
        singular = hinfo >= 0;

        if (!singular) {
            // x = 0, r = b; each step: d = LU \ r (low precision), x += d,
            // r = b - A x (FP64).
            TNS_HIP(hipMemset2DAsync(x, sizeof(double) * (size_t)ldx, 0, sizeof(double) * (size_t)n, t, ctx->stream));
            dcopy2d(ctx, n, t, b, ldb, r.get(), n);
            DevBuf<float> d(ctx, (size_t)n * t);
            DevBuf<double> sc(ctx, t);
            double prev = INFINITY;
            for (iters = 1; iters <= max_iter; iters++) {
                hipLaunchKernelGGL(col_absmax_kernel, dim3(t), dim3(256), 0, ctx->stream, n, r.get(), n, sc.get());
                hipLaunchKernelGGL(rhs_to_single_kernel, dim3(grid_for((size_t)n * t)), dim3(256), 0, ctx->stream, n,
                                   t, r.get(), n, prs, sc.get(), d.get());
                ctx->lp_mode = mode;
                sgetrs(ctx, n, t, s.get(), n, perm.get(), d.get(), n);
                hipLaunchKernelGGL(accumulate_kernel, dim3(grid_for((size_t)n * t)), dim3(256), 0, ctx->stream, n, t,
                                   d.get(), pcs, sc.get(), x, ldx);
                TNS_HIP(hipGetLastError());
                berr = backward_error(ctx, n, t, a, lda, b, ldb, x, ldx, r.get(), anorm, bnorm, d_tmp.get());
                if (iters == 1) out.first_error = berr;
                if (berr <= tol) {
                    converged = true;
                    break;
                }
                if (!(berr <= 0.5 * prev)) break;  // stalled (or NaN)
                prev = berr;
            }
            if (iters > max_iter) iters = max_iter;
        }
        ctx->lp_mode = saved;
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
        TNS_HIP(hipMemcpyAsync(&hinfo, dinfo.get(), sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        TNS_HIP(hipStreamSynchronize(ctx->stream));
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
