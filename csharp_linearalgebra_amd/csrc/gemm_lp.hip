// Single-precision products for the mixed-precision factorisation: the
// trailing updates C := beta*C + alpha*A*B of an FP32 LU (and of its
// solves), on the matrix cores with FP32 accumulation. Three input modes:
//   LP_FP32:   v_mfma_f32_16x16x4f32, operands as stored;
//   LP_BF16:   v_mfma_f32_16x16x32_bf16, operands rounded to bfloat16 (to
//              nearest, ties to even) as they are staged to LDS;
//   LP_BF16X3: the same instruction three times per block, each operand
//              split as x = hi + lo into two bfloat16 parts and the
//              product taken as hi*hi + hi*lo + lo*hi (relative error
//              about 2^-16, near FP32's, at the BF16 rate).
// C and the accumulators stay FP32 in both. A is never transposed here.
//
// Tiles of 64 x 64 x 32 per workgroup of four waves (2 x 2, 32 x 32 each,
// four 16 x 16 MFMA blocks per wave); the tile order is the XCD-banded,
// grouped order of the FP64 kernels. Loads are clamped into the matrix and
// zeroed at the edges, never branched around.
#include <type_traits>

#include "tile_order.hpp"
#include "tns_internal.hpp"

namespace tns {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int LBM = 64, LBN = 64, LBK = 32, LNT = 256;

// Round to nearest, ties to even; NaN stays NaN (quiet).
__device__ __forceinline__ unsigned short to_bf16(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

template <int MODE>
struct LpPlanes;

// FP32: k-contiguous rows padded to 33 floats (the 16 rows of a fragment
// read hit distinct banks).
template <>
struct LpPlanes<LP_FP32> {
    static constexpr int KP = LBK + 1;
    static constexpr int WORDS = LBM * KP;  // floats per plane
};

// BF16: k-contiguous rows of 40 bf16 (80 bytes: 16-byte aligned, and a
// lane's eight consecutive k are one 16-byte read).
template <>
struct LpPlanes<LP_BF16> {
    static constexpr int KP = LBK + 8;
    static constexpr int WORDS = LBM * KP / 2;  // 32-bit words per plane
};

// BF16 x 3: the same rows, a plane of leading parts and, after it, a plane
// of trailing parts.
template <>
struct LpPlanes<LP_BF16X3> {
    static constexpr int KP = LBK + 8;
    static constexpr int WORDS = LBM * KP;  // both parts
    static constexpr int LO = LBM * KP;     // offset of the trailing parts, in bf16
};

// x = hi + lo + O(2^-16 |x|): hi = bf16(x), lo = bf16(x - hi) (the
// difference is exact in FP32).
__device__ __forceinline__ void split_bf16(float x, unsigned short& hi, unsigned short& lo)
{
    hi = to_bf16(x);
    lo = to_bf16(x - __uint_as_float((unsigned)hi << 16));
}

// VEC: A, B and their leading dimensions 16-byte aligned, so a slab is
// loaded four floats at a time (A down a column, B along k).
template <int MODE, bool VEC>
__global__ void __launch_bounds__(LNT, 2)
lpgemm_kernel(int m, int n, int k, float alpha, const float* __restrict__ A, int lda, const float* __restrict__ B,
              int ldb, float beta, float* __restrict__ C, int ldc, int tiles_m, int tiles_n, int kc, long long cs)
{
    // Split over k (gridDim.y > 1): part blockIdx.y takes k in [y kc,
    // (y + 1) kc) and writes its product to its own C at C + y cs.
    {
        const int kbeg = blockIdx.y * kc;
        k = (k < kbeg + kc ? k : kbeg + kc) - kbeg;
        A += (size_t)kbeg * lda;
        B += kbeg;
        C += blockIdx.y * cs;
    }
    using P = LpPlanes<MODE>;
    __shared__ __align__(16) unsigned As[P::WORDS];
    __shared__ __align__(16) unsigned Bs[P::WORDS];

    int tmi, tni;
    tile_of(blockIdx.x, gridDim.x, tiles_m, tiles_n, tmi, tni);
    const int m0 = tmi * LBM, n0 = tni * LBN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;

    v4f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    constexpr int LA = LBM * LBK / LNT;  // 8 elements of A and of B per thread
    float ra[LA], rb[LA];

    auto gload = [&](int k0) {
        if constexpr (VEC) {
            // Thread t: A rows 4(t%16).. at depth t/16 (+16), B depths
            // 4(t%8).. of column t/8 (+32); ra/rb hold them in that order.
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int e4 = tid + h * LNT;
                {
                    const int i4 = (e4 % (LBM / 4)) * 4, kk = e4 / (LBM / 4);
                    const int gi = m0 + i4, gk = k0 + kk;
                    if (gi + 3 < m && gk < k) {
                        const float4 v = *reinterpret_cast<const float4*>(A + gi + (size_t)gk * lda);
                        ra[4 * h] = v.x;
                        ra[4 * h + 1] = v.y;
                        ra[4 * h + 2] = v.z;
                        ra[4 * h + 3] = v.w;
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const int ci = gi + q < m ? gi + q : m - 1, ck = gk < k ? gk : k - 1;
                            const float va = A[ci + (size_t)ck * lda];
                            ra[4 * h + q] = (gi + q < m && gk < k) ? va : 0.f;
                        }
                    }
                }
                {
                    const int kb4 = (e4 % (LBK / 4)) * 4, j = e4 / (LBK / 4);
                    const int gkb = k0 + kb4, gj = n0 + j;
                    if (gkb + 3 < k && gj < n) {
                        const float4 v = *reinterpret_cast<const float4*>(B + gkb + (size_t)gj * ldb);
                        rb[4 * h] = v.x;
                        rb[4 * h + 1] = v.y;
                        rb[4 * h + 2] = v.z;
                        rb[4 * h + 3] = v.w;
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const int ckb = gkb + q < k ? gkb + q : k - 1, cj = gj < n ? gj : n - 1;
                            const float vb = B[ckb + (size_t)cj * ldb];
                            rb[4 * h + q] = (gkb + q < k && gj < n) ? vb : 0.f;
                        }
                    }
                }
            }
            return;
        }
#pragma unroll
        for (int r = 0; r < LA; r++) {
            const int e = tid + r * LNT;
            // A: i fastest (coalesced down a column).
            const int i = e % LBM, kk = e / LBM;
            const int gi = m0 + i, gk = k0 + kk;
            const int ci = gi < m ? gi : m - 1, ck = gk < k ? gk : k - 1;
            const float va = A[ci + (size_t)ck * lda];
            ra[r] = (gi < m && gk < k) ? va : 0.f;
            // B: k fastest.
            const int kb = e % LBK, j = e / LBK;
            const int gkb = k0 + kb, gj = n0 + j;
            const int ckb = gkb < k ? gkb : k - 1, cj = gj < n ? gj : n - 1;
            const float vb = B[ckb + (size_t)cj * ldb];
            rb[r] = (gkb < k && gj < n) ? vb : 0.f;
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int r = 0; r < LA; r++) {
            int i, kk, kb, j;
            if constexpr (VEC) {
                const int e4 = tid + (r / 4) * LNT, q = r % 4;
                i = (e4 % (LBM / 4)) * 4 + q;
                kk = e4 / (LBM / 4);
                kb = (e4 % (LBK / 4)) * 4 + q;
                j = e4 / (LBK / 4);
            } else {
                const int e = tid + r * LNT;
                i = e % LBM;
                kk = e / LBM;
                kb = e % LBK;
                j = e / LBK;
            }
            if constexpr (MODE == LP_FP32) {
                As[i * P::KP + kk] = __float_as_uint(ra[r]);
                Bs[j * P::KP + kb] = __float_as_uint(rb[r]);
            } else if constexpr (MODE == LP_BF16) {
                reinterpret_cast<unsigned short*>(As)[i * P::KP + kk] = to_bf16(ra[r]);
                reinterpret_cast<unsigned short*>(Bs)[j * P::KP + kb] = to_bf16(rb[r]);
            } else {
                unsigned short* as = reinterpret_cast<unsigned short*>(As);
                unsigned short* bs = reinterpret_cast<unsigned short*>(Bs);
                split_bf16(ra[r], as[i * P::KP + kk], as[P::LO + i * P::KP + kk]);
                split_bf16(rb[r], bs[j * P::KP + kb], bs[P::LO + j * P::KP + kb]);
            }
        }
    };

    const int ktiles = (k + LBK - 1) / LBK;
    gload(0);
    sstore();
    __syncthreads();

    const int frow = lane & 15, kq = lane >> 4;
    for (int t = 0; t < ktiles; ++t) {
        if (t + 1 < ktiles) gload((t + 1) * LBK);
        if constexpr (MODE == LP_FP32) {
#pragma unroll
            for (int kk = 0; kk < LBK; kk += 4) {
                float af[2], bf[2];
#pragma unroll
                for (int i = 0; i < 2; i++) af[i] = __uint_as_float(As[(wm * 32 + i * 16 + frow) * P::KP + kk + kq]);
#pragma unroll
                for (int j = 0; j < 2; j++) bf[j] = __uint_as_float(Bs[(wn * 32 + j * 16 + frow) * P::KP + kk + kq]);
#pragma unroll
                for (int i = 0; i < 2; i++)
#pragma unroll
                    for (int j = 0; j < 2; j++)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i], bf[j], acc[i][j], 0, 0, 0);
            }
        } else if constexpr (MODE == LP_BF16) {
            const unsigned short* as = reinterpret_cast<const unsigned short*>(As);
            const unsigned short* bs = reinterpret_cast<const unsigned short*>(Bs);
            bf16x8 af[2], bfr[2];
#pragma unroll
            for (int i = 0; i < 2; i++)
                af[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(as + (wm * 32 + i * 16 + frow) * P::KP + 8 * kq));
#pragma unroll
            for (int j = 0; j < 2; j++)
                bfr[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(bs + (wn * 32 + j * 16 + frow) * P::KP + 8 * kq));
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
        } else {
            // Three products per block: the trailing parts first, then the
            // leading ones (lo x lo, below 2^-16 of the product, is left out).
            const unsigned short* as = reinterpret_cast<const unsigned short*>(As);
            const unsigned short* bs = reinterpret_cast<const unsigned short*>(Bs);
            bf16x8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                const int o = (wm * 32 + i * 16 + frow) * P::KP + 8 * kq;
                ah[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(as + o));
                al[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(as + P::LO + o));
            }
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int o = (wn * 32 + j * 16 + frow) * P::KP + 8 * kq;
                bh[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(bs + o));
                bl[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4*>(bs + P::LO + o));
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
        if (t + 1 < ktiles) {
            sstore();
            __syncthreads();
        }
    }

    // D layout of both 16x16 FP32-accumulating forms: lane l, register r
    // holds row 4*(l/16) + r, column l%16 of the block. The old C values
    // are all read first (their loads in flight together) and every new
    // value formed before the first store; tiles inside C take straight-line
    // code, so no store waits for the one before it.
    auto write_back = [&](auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        float val[2][2][4];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) val[i][j][r] = 0.f;
        if (beta != 0.f) {
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    const int col = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int row = m0 + wm * 32 + i * 16 + 4 * kq + r;
                        const float* p = C + row + (size_t)col * ldc;
                        val[i][j][r] = FULL ? *p : load_if(p, row < m && col < n);
                    }
                }
        }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const float v = alpha * acc[i][j][r];
                    val[i][j][r] = beta == 0.f ? v : beta * val[i][j][r] + v;
                }
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int col = n0 + wn * 32 + j * 16 + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = m0 + wm * 32 + i * 16 + 4 * kq + r;
                    if (FULL || (row < m && col < n)) C[row + (size_t)col * ldc] = val[i][j][r];
                }
            }
    };
    if (m0 + LBM <= m && n0 + LBN <= n) write_back(std::true_type{});
    else write_back(std::false_type{});
}

// Sum of the split-k parts: c = beta c + alpha sum_p w_p (p in order).
__global__ void lp_split_reduce_kernel(int m, int n, int parts, const float* __restrict__ w, float alpha, float beta,
                                       float* __restrict__ c, int ldc)
{
    const size_t total = (size_t)m * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        float acc = 0.f;
        for (int p = 0; p < parts; p++) acc += w[p * total + e];
        float* q = c + e % m + (e / m) * ldc;
        const float v = alpha * acc;
        *q = beta == 0.f ? v : beta * *q + v;
    }
}

__global__ void lp_scale_kernel(int m, int n, float beta, float* __restrict__ c, int ldc)
{
    const size_t total = (size_t)m * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % m), j = (int)(e / m);
        float* p = c + i + (size_t)j * ldc;
        *p = beta == 0.f ? 0.f : beta * *p;
    }
}

}  // namespace

void sgemm_lp(Context* ctx, int mode, int m, int n, int k, float alpha, const float* a, int lda, const float* b,
              int ldb, float beta, float* c, int ldc)
{
    if (m <= 0 || n <= 0) return;
    if (k == 0 || alpha == 0.f) {
        const size_t total = (size_t)m * n;
        const int blocks = (int)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096);
        hipLaunchKernelGGL(lp_scale_kernel, dim3(blocks), dim3(256), 0, ctx->stream, m, n, beta, c, ldc);
        TNS_HIP(hipGetLastError());
        return;
    }
    const int tm = (m + LBM - 1) / LBM, tn = (n + LBN - 1) / LBN;
    const int tiles = tm * tn;
    const bool vec = ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0 && lda % 4 == 0 &&
                     ldb % 4 == 0;
    auto kernel = vec ? (mode == LP_BF16 ? lpgemm_kernel<LP_BF16, true>
                         : mode == LP_BF16X3 ? lpgemm_kernel<LP_BF16X3, true>
                                             : lpgemm_kernel<LP_FP32, true>)
                      : (mode == LP_BF16 ? lpgemm_kernel<LP_BF16, false>
                         : mode == LP_BF16X3 ? lpgemm_kernel<LP_BF16X3, false>
                                             : lpgemm_kernel<LP_FP32, false>);
    if (tiles <= 256 && k >= 256) {
        // Few tiles over a long k (the thin products of the triangular
        // solves): split k into parts written to scratch, then summed.
        int parts = 1024 / tiles;
        parts = parts < 32 ? parts : 32;
        parts = parts < k / 128 ? parts : k / 128;
        const int kc = ((k + parts - 1) / parts + LBK - 1) / LBK * LBK;
        parts = (k + kc - 1) / kc;
        DevBuf<float> w(ctx, (size_t)parts * m * n);
        hipLaunchKernelGGL(kernel, dim3(tiles, parts), dim3(LNT), 0, ctx->stream, m, n, k, 1.f, a, lda, b, ldb, 0.f,
                           w.get(), m, tm, tn, kc, (long long)m * n);
        const size_t total = (size_t)m * n;
        const int blocks = (int)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096);
        hipLaunchKernelGGL(lp_split_reduce_kernel, dim3(blocks), dim3(256), 0, ctx->stream, m, n, parts, w.get(),
                           alpha, beta, c, ldc);
        TNS_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(kernel, dim3(tiles), dim3(LNT), 0, ctx->stream, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc,
                       tm, tn, k, 0LL);
    TNS_HIP(hipGetLastError());
}

}  // namespace tns
