// Real and complex GEMM on gfx950 FP64 matrix cores.
//
// Replaces Tensile's BLIS-style path (Packing.cs, Gemm.cs, ParallelGemm.cs,
// MicroKernels.cs) and the 4M complex product (ComplexKernels.cs:49-108).
//
// Design, per workgroup: a BM x BN tile of C, 4 waves each owning a WM x WN
// sub-tile built from 16x16 v_mfma_f64_16x16x4_f64 blocks held in
// registers. A and B stream through LDS in 16-deep slabs: the next slab is
// fetched into registers (clamped, branch-free loads, all in flight at once)
// while the current one feeds the matrix cores, then written to LDS between
// two barriers.
// Complex operands are split into real and imaginary LDS planes as they are
// staged, and each complex 16x16 block takes four real MFMAs,
//     Re += Ar*Br - Ai*Bi,   Im += Ar*Bi + Ai*Br,
// the 4M method (never 3M): every part is a signed sum of exactly the
// products the conventional method forms.
//
// Workgroups are dispatched round-robin over the 8 XCDs; the tile order is
// remapped so each XCD receives a contiguous band of tiles, grouped 8 tile
// rows at a time, and the A row-panels and B column-panels a band shares
// stay in that XCD's L2.
#include <type_traits>

#include "grouped.hpp"
#include "tile_order.hpp"
#include "tns_internal.hpp"

namespace tns {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <int BM, int BN, int BK, int WM, int WN>
struct Shape {
    static constexpr int WAVES_M = BM / WM;
    static constexpr int WAVES_N = BN / WN;
    static constexpr int NT = WAVES_M * WAVES_N * 64;
    static constexpr int TM = WM / 16;
    static constexpr int TN = WN / 16;
    static constexpr int LA = BM * BK / NT;
    static constexpr int LB = BK * BN / NT;
    static constexpr int KP = BK + 2;  // padded k-stride of the k-contiguous LDS planes
    static_assert((BM * BK) % NT == 0 && (BK * BN) % NT == 0, "slab must split evenly over threads");
    static_assert(BK % 4 == 0 && WM % 16 == 0 && WN % 16 == 0, "MFMA granularity");
    static_assert(LA % 2 == 0 && LB % 2 == 0, "paired loads need an even count per thread");
};

// LDS planes of one slab. A is held [k][i] (i contiguous) when it is read
// untransposed and [i][k] (k contiguous, padded) when op(A) = A^T, so the
// global loads stay coalesced either way; B is [j][k] (k contiguous,
// padded). With the k-stride BK + 2 the 16 lanes of a fragment read touch
// distinct bank pairs.
template <int BM, int BN, int BK, int TA>
struct Planes {
    static constexpr int KP = BK + 2;
    static constexpr int A_SIZE = TA == 0 ? BK * BM : BM * KP;
    static constexpr int B_SIZE = BN * KP;
    __device__ static int a_index(int i, int k) { return TA == 0 ? k * BM + i : i * KP + k; }
    __device__ static int b_index(int k, int j) { return j * KP + k; }
};

// Global index of the e-th element of an A slab (rows i, depth kk), laid
// out so consecutive threads read consecutive memory.
template <int BM, int BK, int TA>
__device__ __forceinline__ void a_coords(int e, int& i, int& kk)
{
    if (TA == 0) {
        i = e % BM;
        kk = e / BM;
    } else {
        kk = e % BK;
        i = e / BK;
    }
}

// Products with a triangular operand (op(A) untransposed): the k range of
// the tile at (m0, n0) shrinks to where the triangle can be nonzero.
// kmode 1: A lower (k < m0 + BM); 2: A upper (k >= m0); 3: B upper
// (k < n0 + BN); 4: B lower (k >= n0). The range keeps at least one slab
// and starts at a multiple of BK, so aligned operands stay aligned; the
// parts of the k range it still covers beyond the triangle are zeros of
// the operand, which add nothing.
template <int BK, int TA, typename PA, typename PB>
__device__ __forceinline__ void tri_k_range(int kmode, int m0, int n0, int BM, int BN, int& k, PA& A, int lda, PB& B)
{
    if (kmode == 0 || TA != 0) return;
    if (kmode == 1 || kmode == 3) {
        const int end = kmode == 1 ? m0 + BM : n0 + BN;
        k = k < end ? k : end;
    } else {
        int kb = kmode == 2 ? m0 : n0;
        const int last = (k - 1) / BK * BK;
        kb = kb < last ? kb / BK * BK : last;
        A += (size_t)kb * lda;
        B += kb;
        k -= kb;
    }
}

template <int BM, int BN, int BK, int WM, int WN, int TA>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64, (BM / WM) * (BN / WN) * 64 > 256 ? 1 : 2)
dgemm_kernel(int m, int n, int k, double alpha, const double* __restrict__ A, int lda,
             const double* __restrict__ B, int ldb, double beta, double* __restrict__ C, int ldc,
             int tiles_m, int tiles_n, int aligned, int kc, long long cs, int kmode, long long sa, long long sb,
             long long sc)
{
    // Split over k (gridDim.y > 1): part blockIdx.y takes k in [y kc,
    // (y + 1) kc) and writes its product to its own C at C + y cs.
    {
        const int kbeg = blockIdx.y * kc;
        k = (k < kbeg + kc ? k : kbeg + kc) - kbeg;
        A += TA == 0 ? (size_t)kbeg * lda : (size_t)kbeg;
        B += kbeg;
        C += blockIdx.y * cs;
    }
    // A batch of products (gridDim.z > 1): product z at A + z sa, B + z sb,
    // C + z sc.
    A += blockIdx.z * sa;
    B += blockIdx.z * sb;
    C += blockIdx.z * sc;
    using S = Shape<BM, BN, BK, WM, WN>;
    using P = Planes<BM, BN, BK, TA>;
    __shared__ double As[P::A_SIZE];
    __shared__ double Bs[P::B_SIZE];

    int tmi, tni;
    tile_of(blockIdx.x, gridDim.x, tiles_m, tiles_n, tmi, tni);
    const int m0 = tmi * BM, n0 = tni * BN;
    tri_k_range<BK, TA>(kmode, m0, n0, BM, BN, k, A, lda, B);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % S::WAVES_M, wn = wave / S::WAVES_M;

    v4d acc[S::TM][S::TN];
#pragma unroll
    for (int i = 0; i < S::TM; i++)
#pragma unroll
        for (int j = 0; j < S::TN; j++) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    double ra[S::LA], rb[S::LB];

    // Interior slabs of aligned operands load 16 bytes per lane (two
    // neighbours along the contiguous dimension); elsewhere loads are
    // clamped into the matrix and zeroed after, never branched around, so
    // every load of a slab is in flight at once.
    const bool m_inside = m0 + BM <= m, n_inside = n0 + BN <= n;
    bool vec = false, vec_a = false, vec_b = false;
    auto gload = [&](int k0) {
        const bool k_inside = k0 + BK <= k;
        if (aligned && k_inside && (TA == 0 ? m_inside : m_inside)) {
#pragma unroll
            for (int r = 0; r < S::LA / 2; r++) {
                const int e = tid + r * S::NT;
                int i, kk;
                if (TA == 0) { i = 2 * (e % (BM / 2)); kk = e / (BM / 2); }
                else { kk = 2 * (e % (BK / 2)); i = e / (BK / 2); }
                const double2 v = TA == 0 ? *reinterpret_cast<const double2*>(A + (m0 + i) + (size_t)(k0 + kk) * lda)
                                          : *reinterpret_cast<const double2*>(A + (k0 + kk) + (size_t)(m0 + i) * lda);
                ra[2 * r] = v.x;
                ra[2 * r + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int r = 0; r < S::LA; r++) {
                int i, kk;
                a_coords<BM, BK, TA>(tid + r * S::NT, i, kk);
                const int gi = m0 + i, gk = k0 + kk;
                const int ci = gi < m ? gi : m - 1, ck = gk < k ? gk : k - 1;
                const double v = TA == 0 ? A[ci + (size_t)ck * lda] : A[ck + (size_t)ci * lda];
                ra[r] = (gi < m && gk < k) ? v : 0.0;
            }
        }
        if (aligned && k_inside && n_inside) {
#pragma unroll
            for (int r = 0; r < S::LB / 2; r++) {
                const int e = tid + r * S::NT;
                const int kk = 2 * (e % (BK / 2)), j = e / (BK / 2);
                const double2 v = *reinterpret_cast<const double2*>(B + (k0 + kk) + (size_t)(n0 + j) * ldb);
                rb[2 * r] = v.x;
                rb[2 * r + 1] = v.y;
            }
        } else {
#pragma unroll
            for (int r = 0; r < S::LB; r++) {
                const int e = tid + r * S::NT;
                const int kk = e % BK, j = e / BK;
                const int gk = k0 + kk, gj = n0 + j;
                const int ck = gk < k ? gk : k - 1, cj = gj < n ? gj : n - 1;
                const double v = B[ck + (size_t)cj * ldb];
                rb[r] = (gk < k && gj < n) ? v : 0.0;
            }
        }
        vec = aligned && k_inside;
        vec_a = vec && m_inside;
        vec_b = vec && n_inside;
    };
    auto sstore = [&]() {
        if (vec_a) {
#pragma unroll
            for (int r = 0; r < S::LA / 2; r++) {
                const int e = tid + r * S::NT;
                int i, kk;
                if (TA == 0) { i = 2 * (e % (BM / 2)); kk = e / (BM / 2); }
                else { kk = 2 * (e % (BK / 2)); i = e / (BK / 2); }
                As[P::a_index(i, kk)] = ra[2 * r];
                As[P::a_index(TA == 0 ? i + 1 : i, TA == 0 ? kk : kk + 1)] = ra[2 * r + 1];
            }
        } else {
#pragma unroll
            for (int r = 0; r < S::LA; r++) {
                int i, kk;
                a_coords<BM, BK, TA>(tid + r * S::NT, i, kk);
                As[P::a_index(i, kk)] = ra[r];
            }
        }
        if (vec_b) {
#pragma unroll
            for (int r = 0; r < S::LB / 2; r++) {
                const int e = tid + r * S::NT;
                const int kk = 2 * (e % (BK / 2)), j = e / (BK / 2);
                Bs[P::b_index(kk, j)] = rb[2 * r];
                Bs[P::b_index(kk + 1, j)] = rb[2 * r + 1];
            }
        } else {
#pragma unroll
            for (int r = 0; r < S::LB; r++) {
                const int e = tid + r * S::NT;
                Bs[P::b_index(e % BK, e / BK)] = rb[r];
            }
        }
    };

    const int ktiles = (k + BK - 1) / BK;
    gload(0);
    sstore();
    __syncthreads();

    const int arow = wm * WM + (lane & 15);
    const int bcol = wn * WN + (lane & 15);
    const int kq = lane >> 4;

    for (int t = 0; t < ktiles; ++t) {
        if (t + 1 < ktiles) gload((t + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double af[S::TM], bf[S::TN];
#pragma unroll
            for (int i = 0; i < S::TM; i++) af[i] = As[P::a_index(arow + i * 16, kk + kq)];
#pragma unroll
            for (int j = 0; j < S::TN; j++) bf[j] = Bs[P::b_index(kk + kq, bcol + j * 16)];
#pragma unroll
            for (int i = 0; i < S::TM; i++)
#pragma unroll
                for (int j = 0; j < S::TN; j++) acc[i][j] = mfma_f64(af[i], bf[j], acc[i][j]);
        }
        __syncthreads();
        if (t + 1 < ktiles) {
            sstore();
            __syncthreads();
        }
    }

    // D block layout of v_mfma_f64_16x16x4_f64 (one-element groups): lane l,
    // register r holds row 4*r + l/16, column l%16 of the 16x16 block.
    // C is written in groups of GB blocks: the group's old values are all
    // read first, so their loads are in flight together, and every new value
    // is formed before the first store (a read-modify-write per element
    // waits out one memory latency each, which dominates the thin-k updates
    // of a factorization). Tiles inside C take straight-line code: with
    // per-element branches between them, every store waits for the one
    // before it.
    auto write_back = [&](auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        constexpr int NBLK = S::TM * S::TN;
        constexpr int GB = NBLK < 4 ? NBLK : 4;
#pragma unroll
        for (int g0 = 0; g0 < NBLK; g0 += GB) {
            double val[GB][4];
#pragma unroll
            for (int g = 0; g < GB; g++)
#pragma unroll
                for (int r = 0; r < 4; r++) val[g][r] = 0.0;
            if (beta != 0.0) {
#pragma unroll
                for (int g = 0; g < GB; g++) {
                    const int i = (g0 + g) / S::TN, j = (g0 + g) % S::TN;
                    const int col = n0 + wn * WN + j * 16 + (lane & 15);
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int row = m0 + wm * WM + i * 16 + 4 * r + kq;
                        const double* p = C + row + (size_t)col * ldc;
                        val[g][r] = FULL ? *p : load_if(p, row < m && col < n);
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < GB; g++) {
                const int i = (g0 + g) / S::TN, j = (g0 + g) % S::TN;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const double v = alpha * acc[i][j][r];
                    val[g][r] = beta == 0.0 ? v : beta * val[g][r] + v;
                }
            }
#pragma unroll
            for (int g = 0; g < GB; g++) {
                const int i = (g0 + g) / S::TN, j = (g0 + g) % S::TN;
                const int col = n0 + wn * WN + j * 16 + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = m0 + wm * WM + i * 16 + 4 * r + kq;
                    if (FULL || (row < m && col < n)) C[row + (size_t)col * ldc] = val[g][r];
                }
            }
        }
    };
    if (m0 + BM <= m && n0 + BN <= n) write_back(std::true_type{});
    else write_back(std::false_type{});
}

// Complex, with ComplexKernels.Combine's scalar special cases folded into
// the write-back (ComplexKernels.cs:213-240). flags: 1 unit alpha, 2 real
// alpha, 4 zero beta, 8 unit beta. TA: 0 A, 1 A^T, 2 A^H.
template <int BM, int BN, int BK, int WM, int WN, int TA, typename PA, typename PB, typename PC>
__device__ __forceinline__ void zgemm_tile(int m, int n, int k, double ar_, double ai_, PA __restrict__ A, int lda,
                                           PB __restrict__ B, int ldb, double br_, double bi_, PC __restrict__ C,
                                           int ldc, int flags, int m0, int n0)
{
    using S = Shape<BM, BN, BK, WM, WN>;
    using P = Planes<BM, BN, BK, TA>;
    __shared__ double Ar[P::A_SIZE];
    __shared__ double Ai[P::A_SIZE];
    __shared__ double Br[P::B_SIZE];
    __shared__ double Bi[P::B_SIZE];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave % S::WAVES_M, wn = wave / S::WAVES_M;

    v4d re[S::TM][S::TN], im[S::TM][S::TN];
#pragma unroll
    for (int i = 0; i < S::TM; i++)
#pragma unroll
        for (int j = 0; j < S::TN; j++) {
            re[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
            im[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
        }

    double ra_r[S::LA], ra_i[S::LA], rb_r[S::LB], rb_i[S::LB];

    auto gload = [&](int k0) {
#pragma unroll
        for (int r = 0; r < S::LA; r++) {
            int i, kk;
            a_coords<BM, BK, TA>(tid + r * S::NT, i, kk);
            const int gi = m0 + i, gk = k0 + kk;
            const int ci = gi < m ? gi : m - 1, ck = gk < k ? gk : k - 1;
            const double2 v = TA == 0 ? A[ci + (size_t)ck * lda] : A[ck + (size_t)ci * lda];
            const bool ok = gi < m && gk < k;
            ra_r[r] = ok ? v.x : 0.0;
            ra_i[r] = ok ? (TA == 2 ? -v.y : v.y) : 0.0;
        }
#pragma unroll
        for (int r = 0; r < S::LB; r++) {
            const int e = tid + r * S::NT;
            const int kk = e % BK, j = e / BK;
            const int gk = k0 + kk, gj = n0 + j;
            const int ck = gk < k ? gk : k - 1, cj = gj < n ? gj : n - 1;
            const double2 v = B[ck + (size_t)cj * ldb];
            const bool ok = gk < k && gj < n;
            rb_r[r] = ok ? v.x : 0.0;
            rb_i[r] = ok ? v.y : 0.0;
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int r = 0; r < S::LA; r++) {
            int i, kk;
            a_coords<BM, BK, TA>(tid + r * S::NT, i, kk);
            Ar[P::a_index(i, kk)] = ra_r[r];
            Ai[P::a_index(i, kk)] = ra_i[r];
        }
#pragma unroll
        for (int r = 0; r < S::LB; r++) {
            const int e = tid + r * S::NT;
            Br[P::b_index(e % BK, e / BK)] = rb_r[r];
            Bi[P::b_index(e % BK, e / BK)] = rb_i[r];
        }
    };

    const int ktiles = (k + BK - 1) / BK;
    gload(0);
    sstore();
    __syncthreads();

    const int arow = wm * WM + (lane & 15);
    const int bcol = wn * WN + (lane & 15);
    const int kq = lane >> 4;

    for (int t = 0; t < ktiles; ++t) {
        if (t + 1 < ktiles) gload((t + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double a_r[S::TM], a_i[S::TM], a_n[S::TM], b_r[S::TN], b_i[S::TN];
#pragma unroll
            for (int i = 0; i < S::TM; i++) {
                a_r[i] = Ar[P::a_index(arow + i * 16, kk + kq)];
                a_i[i] = Ai[P::a_index(arow + i * 16, kk + kq)];
                a_n[i] = -a_i[i];
            }
#pragma unroll
            for (int j = 0; j < S::TN; j++) {
                b_r[j] = Br[P::b_index(kk + kq, bcol + j * 16)];
                b_i[j] = Bi[P::b_index(kk + kq, bcol + j * 16)];
            }
#pragma unroll
            for (int i = 0; i < S::TM; i++)
#pragma unroll
                for (int j = 0; j < S::TN; j++) {
                    re[i][j] = mfma_f64(a_r[i], b_r[j], re[i][j]);
                    im[i][j] = mfma_f64(a_r[i], b_i[j], im[i][j]);
                    re[i][j] = mfma_f64(a_n[i], b_i[j], re[i][j]);
                    im[i][j] = mfma_f64(a_i[i], b_r[j], im[i][j]);
                }
        }
        __syncthreads();
        if (t + 1 < ktiles) {
            sstore();
            __syncthreads();
        }
    }

    const bool unit_alpha = flags & 1, real_alpha = flags & 2, zero_beta = flags & 4, unit_beta = flags & 8;
    // Written as in dgemm_kernel: groups of GB blocks, old C read first,
    // new values formed before the first store, straight-line code for
    // tiles inside C.
    auto write_back = [&](auto full_tag) {
        constexpr bool FULL = decltype(full_tag)::value;
        constexpr int NBLK = S::TM * S::TN;
        constexpr int GB = NBLK < 2 ? NBLK : 2;
#pragma unroll
        for (int g0 = 0; g0 < NBLK; g0 += GB) {
            double vr[GB][4], vi[GB][4];
#pragma unroll
            for (int g = 0; g < GB; g++)
#pragma unroll
                for (int r = 0; r < 4; r++) vr[g][r] = vi[g][r] = 0.0;
            if (!zero_beta) {
#pragma unroll
                for (int g = 0; g < GB; g++) {
                    const int i = (g0 + g) / S::TN, j = (g0 + g) % S::TN;
                    const int col = n0 + wn * WN + j * 16 + (lane & 15);
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int row = m0 + wm * WM + i * 16 + 4 * r + kq;
                        double2 c = make_double2(0.0, 0.0);
                        if (FULL || (row < m && col < n)) c = C[row + (size_t)col * ldc];
                        vr[g][r] = c.x;
                        vi[g][r] = c.y;
                    }
                }
            }
#pragma unroll
            for (int g = 0; g < GB; g++) {
                const int i = (g0 + g) / S::TN, j = (g0 + g) % S::TN;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const double p = re[i][j][r], q = im[i][j][r];
                    double pr, pi;
                    if (unit_alpha) { pr = p; pi = q; }
                    else if (real_alpha) { pr = ar_ * p; pi = ar_ * q; }
                    else { pr = ar_ * p - ai_ * q; pi = ai_ * p + ar_ * q; }
                    const double cr = vr[g][r], ci = vi[g][r];
                    if (zero_beta) { vr[g][r] = pr; vi[g][r] = pi; }
                    else if (unit_beta) { vr[g][r] = cr + pr; vi[g][r] = ci + pi; }
                    else { vr[g][r] = (br_ * cr - bi_ * ci) + pr; vi[g][r] = (bi_ * cr + br_ * ci) + pi; }
                }
            }
#pragma unroll
            for (int g = 0; g < GB; g++) {
                const int i = (g0 + g) / S::TN, j = (g0 + g) % S::TN;
                const int col = n0 + wn * WN + j * 16 + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = m0 + wm * WM + i * 16 + 4 * r + kq;
                    if (FULL || (row < m && col < n)) C[row + (size_t)col * ldc] = make_double2(vr[g][r], vi[g][r]);
                }
            }
        }
    };
    if (m0 + BM <= m && n0 + BN <= n) write_back(std::true_type{});
    else write_back(std::false_type{});
}

template <int BM, int BN, int BK, int WM, int WN, int TA>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64, 2)
zgemm_kernel(int m, int n, int k, double ar_, double ai_, const double2* __restrict__ A, int lda,
             const double2* __restrict__ B, int ldb, double br_, double bi_, double2* __restrict__ C,
             int ldc, int flags, int tiles_m, int tiles_n, int kmode, long long sa, long long sb, long long sc)
{
    // A batch of products (gridDim.y > 1): product y at A + y sa, B + y sb,
    // C + y sc.
    A += blockIdx.y * sa;
    B += blockIdx.y * sb;
    C += blockIdx.y * sc;
    int tmi, tni;
    tile_of(blockIdx.x, gridDim.x, tiles_m, tiles_n, tmi, tni);
    tri_k_range<BK, TA>(kmode, tmi * BM, tni * BN, BM, BN, k, A, lda, B);
    zgemm_tile<BM, BN, BK, WM, WN, TA>(m, n, k, ar_, ai_, A, lda, B, ldb, br_, bi_, C, ldc, flags, tmi * BM,
                                       tni * BN);
}

// C_g -= A_g B_g for many independent products of different shapes in one
// launch (the trailing updates of a batch of LUs). Record g owns tiles
// [starts[g], starts[g + 1]), in column-major order; each workgroup walks
// the tiles t = blockIdx.x, blockIdx.x + gridDim.x, ..., so the grid is
// sized to the chip, not to the tiles, and the record search moves forward
// only. alpha = -1 and beta = 1 take the same write-back as zgemm_ex gives
// them.
template <int BM, int BN, int BK, int WM, int WN>
__global__ void __launch_bounds__((BM / WM) * (BN / WN) * 64, 2)
zgemm_grouped_kernel(const GemmRec<zd>* __restrict__ recs, const int* __restrict__ starts, int count, int total)
{
    int g = 0;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        int hi = count - 1;
        while (g < hi) {
            const int mid = (g + hi + 1) >> 1;
            if (starts[mid] <= t) g = mid;
            else hi = mid - 1;
        }
        const GemmRec<zd> r = recs[g];
        const int local = t - starts[g];
        const int tiles_m = (r.m + BM - 1) / BM;
        zgemm_tile<BM, BN, BK, WM, WN, 0>(r.m, r.n, r.k, -1.0, 0.0, as_global(reinterpret_cast<const double2*>(r.a)),
                                          r.lda, as_global(reinterpret_cast<const double2*>(r.b)), r.ldb, 1.0, 0.0,
                                          as_global(reinterpret_cast<double2*>(r.c)), r.ldc, 2 | 8,
                                          (local % tiles_m) * BM, (local / tiles_m) * BN);
    }
}

__global__ void dscale_kernel(int m, int n, double beta, double* __restrict__ c, int ldc)
{
    size_t total = (size_t)m * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        size_t j = e / m, i = e % m;
        double* p = c + i + j * ldc;
        *p = beta == 0.0 ? 0.0 : beta * *p;
    }
}

__global__ void zscale_kernel(int m, int n, double br, double bi, double2* __restrict__ c, int ldc)
{
    size_t total = (size_t)m * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        size_t j = e / m, i = e % m;
        double2* p = c + i + j * ldc;
        double2 v = *p;
        *p = (br == 0.0 && bi == 0.0) ? make_double2(0.0, 0.0) : make_double2(br * v.x - bi * v.y, bi * v.x + br * v.y);
    }
}

int g_real_cfg = -1, g_cplx_cfg = -1;

// Sum of the split-k parts: c = beta c + alpha sum_p w_p (p in order).
__global__ void dsplit_reduce_kernel(int m, int n, int parts, const double* __restrict__ w, double alpha, double beta,
                                     double* __restrict__ c, int ldc)
{
    const size_t total = (size_t)m * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        double acc = 0.0;
        for (int p = 0; p < parts; p++) acc += w[p * total + e];
        double* q = c + e % m + (e / m) * ldc;
        const double v = alpha * acc;
        *q = beta == 0.0 ? v : beta * *q + v;
    }
}

// One launch, or, when the tiles of C leave most of the chip idle and k is
// long (the thin products of refinement residuals), the product split over
// k into parts written to scratch and summed by dsplit_reduce_kernel.
template <int BM, int BN, int BK, int WM, int WN, int TA>
void launch_d(Context* ctx, int m, int n, int k, double alpha, const double* a, int lda, const double* b, int ldb,
              double beta, double* c, int ldc)
{
    int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    const int aligned = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) && lda % 2 == 0 && ldb % 2 == 0;
    const int tiles = tm * tn;
    if (tiles <= 256 && k >= 2048) {
        int parts = 1024 / tiles;
        parts = parts < 32 ? parts : 32;
        parts = parts < k / 256 ? parts : k / 256;
        const int kc = ((k + parts - 1) / parts + BK - 1) / BK * BK;
        parts = (k + kc - 1) / kc;
        DevBuf<double> w(ctx, (size_t)parts * m * n);
        hipLaunchKernelGGL((dgemm_kernel<BM, BN, BK, WM, WN, TA>), dim3(tiles, parts),
                           dim3(Shape<BM, BN, BK, WM, WN>::NT), 0, ctx->stream, m, n, k, 1.0, a, lda, b, ldb, 0.0,
                           w.get(), m, tm, tn, aligned, kc, (long long)m * n, 0, 0LL, 0LL, 0LL);
        const size_t total = (size_t)m * n;
        const int blocks = (int)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096);
        hipLaunchKernelGGL(dsplit_reduce_kernel, dim3(blocks), dim3(256), 0, ctx->stream, m, n, parts, w.get(), alpha,
                           beta, c, ldc);
        return;
    }
    hipLaunchKernelGGL((dgemm_kernel<BM, BN, BK, WM, WN, TA>), dim3(tiles), dim3(Shape<BM, BN, BK, WM, WN>::NT), 0,
                       ctx->stream, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tm, tn, aligned, k, 0LL, 0, 0LL,
                       0LL, 0LL);
}

// A batch of products of one shape at fixed strides, with an optional
// triangular operand (tri_k_range), untransposed, in one launch.
template <int BM, int BN, int BK, int WM, int WN>
void launch_d_batch(Context* ctx, int m, int n, int k, double alpha, const double* a, int lda, long long sa,
                    const double* b, int ldb, long long sb, double beta, double* c, int ldc, long long sc, int batch,
                    int kmode)
{
    const int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    const int aligned = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) && lda % 2 == 0 && ldb % 2 == 0 &&
                        sa % 2 == 0 && sb % 2 == 0;
    hipLaunchKernelGGL((dgemm_kernel<BM, BN, BK, WM, WN, 0>), dim3(tm * tn, 1, batch),
                       dim3(Shape<BM, BN, BK, WM, WN>::NT), 0, ctx->stream, m, n, k, alpha, a, lda, b, ldb, beta, c,
                       ldc, tm, tn, aligned, k, 0LL, kmode, sa, sb, sc);
}

template <int BM, int BN, int BK, int WM, int WN, int TA>
void launch_z(Context* ctx, int m, int n, int k, cplx alpha, const cplx* a, int lda, const cplx* b, int ldb,
              cplx beta, cplx* c, int ldc, int flags)
{
    int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    hipLaunchKernelGGL((zgemm_kernel<BM, BN, BK, WM, WN, TA>), dim3(tm * tn), dim3(Shape<BM, BN, BK, WM, WN>::NT), 0,
                       ctx->stream, m, n, k, alpha.x, alpha.y, reinterpret_cast<const double2*>(a), lda,
                       reinterpret_cast<const double2*>(b), ldb, beta.x, beta.y, reinterpret_cast<double2*>(c), ldc,
                       flags, tm, tn, 0, 0LL, 0LL, 0LL);
}

template <int BM, int BN, int BK, int WM, int WN>
void launch_z_batch(Context* ctx, int m, int n, int k, cplx alpha, const cplx* a, int lda, long long sa,
                    const cplx* b, int ldb, long long sb, cplx beta, cplx* c, int ldc, long long sc, int batch,
                    int flags, int kmode)
{
    const int tm = (m + BM - 1) / BM, tn = (n + BN - 1) / BN;
    hipLaunchKernelGGL((zgemm_kernel<BM, BN, BK, WM, WN, 0>), dim3(tm * tn, batch),
                       dim3(Shape<BM, BN, BK, WM, WN>::NT), 0, ctx->stream, m, n, k, alpha.x, alpha.y,
                       reinterpret_cast<const double2*>(a), lda, reinterpret_cast<const double2*>(b), ldb, beta.x,
                       beta.y, reinterpret_cast<double2*>(c), ldc, flags, tm, tn, kmode, sa, sb, sc);
}

long tiles(int m, int n, int bm, int bn) { return (long)((m + bm - 1) / bm) * ((n + bn - 1) / bn); }

}  // namespace

void gemm_force_config(int real_cfg, int cplx_cfg)
{
    g_real_cfg = real_cfg;
    g_cplx_cfg = cplx_cfg;
}

int gemm_config_count(bool complex_) { return complex_ ? 5 : 4; }

void dscale_c(Context* ctx, int m, int n, double beta, double* c, int ldc)
{
    if (beta == 1.0 || m == 0 || n == 0) return;
    size_t total = (size_t)m * n;
    int blocks = (int)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096);
    hipLaunchKernelGGL(dscale_kernel, dim3(blocks), dim3(256), 0, ctx->stream, m, n, beta, c, ldc);
}

void zscale_c(Context* ctx, int m, int n, cplx beta, cplx* c, int ldc)
{
    if ((beta.x == 1.0 && beta.y == 0.0) || m == 0 || n == 0) return;
    size_t total = (size_t)m * n;
    int blocks = (int)((total + 255) / 256 < 4096 ? (total + 255) / 256 : 4096);
    hipLaunchKernelGGL(zscale_kernel, dim3(blocks), dim3(256), 0, ctx->stream, m, n, beta.x, beta.y,
                       reinterpret_cast<double2*>(c), ldc);
}

// Real: 0 = 128x128 (4 waves of 64x64), 1 = 128x64 (4 of 64x32),
// 2 = 64x64 (4 of 32x32). Measured on MI355X the 64x64 tile is the fastest
// or level up to n = 4096 and for the thin-k updates of the factorisation;
// the 128x128 tile pays only once there are four of its tiles per CU.
void dgemm_ex(Context* ctx, int ta, int m, int n, int k, double alpha, const double* a, int lda, const double* b,
              int ldb, double beta, double* c, int ldc)
{
    if (m == 0 || n == 0) return;
    if (k == 0 || alpha == 0.0) {
        dscale_c(ctx, m, n, beta, c, ldc);
        return;
    }
    // 3: 128 x 32 tiles for products at most 32 columns wide (refinement
    // residuals, few-column solves, the LU's panel-chunk updates): no idle
    // half tile, twice the rows per workgroup (8192 x 32 x 8192: 234 -> 120
    // us, A streamed at 4.5 TB/s).
    int cfg = g_real_cfg;
    if (cfg < 0) {
        if (n <= 32) cfg = 3;
        else cfg = tiles(m, n, 128, 128) >= 1024 && k >= 1024 ? 0 : 2;
    }
#define TNS_D(TA_)                                                                                          \
    switch (cfg) {                                                                                         \
    case 0: launch_d<128, 128, 16, 64, 64, TA_>(ctx, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); break; \
    case 1: launch_d<128, 64, 16, 64, 32, TA_>(ctx, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); break;  \
    case 3: launch_d<128, 32, 16, 32, 32, TA_>(ctx, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); break;  \
    default: launch_d<64, 64, 16, 32, 32, TA_>(ctx, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc); break; \
    }
    if (ta == 0) { TNS_D(0) } else { TNS_D(1) }
#undef TNS_D
    TNS_HIP(hipGetLastError());
}

// Complex: 0 = 128x64 (4 waves of 64x32), 1 = 64x64 (4 of 32x32),
// 2 = 64x32 (4 of 32x16), all with slabs of depth 16; 3 = 64x32 and
// 4 = 64x64 with slabs of depth 32 (half the barriers per k): measured at
// n = 1000 / 2000 / 4096, config 3 54.0 TFLOP/s against 51.8 for config 2,
// config 4 62.7 / 67.3 against 58.1 / 66.0 for config 0.
void zgemm_ex(Context* ctx, int ta, int m, int n, int k, cplx alpha, const cplx* a, int lda, const cplx* b, int ldb,
              cplx beta, cplx* c, int ldc)
{
    if (m == 0 || n == 0) return;
    if (k == 0 || (alpha.x == 0.0 && alpha.y == 0.0)) {
        zscale_c(ctx, m, n, beta, c, ldc);
        return;
    }
    int flags = 0;
    if (alpha.x == 1.0 && alpha.y == 0.0) flags |= 1;
    if (alpha.y == 0.0) flags |= 2;
    if (beta.x == 0.0 && beta.y == 0.0) flags |= 4;
    if (beta.x == 1.0 && beta.y == 0.0) flags |= 8;
    int cfg = g_cplx_cfg;
    if (cfg < 0) cfg = tiles(m, n, 64, 64) >= 512 ? 4 : 3;
#define TNS_Z(TA_)                                                                                                  \
    switch (cfg) {                                                                                                 \
    case 0: launch_z<128, 64, 16, 64, 32, TA_>(ctx, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, flags); break;  \
    case 1: launch_z<64, 64, 16, 32, 32, TA_>(ctx, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, flags); break;   \
    case 3: launch_z<64, 32, 32, 32, 16, TA_>(ctx, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, flags); break;  \
    case 4: launch_z<64, 64, 32, 32, 32, TA_>(ctx, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, flags); break;  \
    default: launch_z<64, 32, 16, 32, 16, TA_>(ctx, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, flags); break;  \
    }
    if (ta == 0) { TNS_Z(0) } else if (ta == 1) { TNS_Z(1) } else { TNS_Z(2) }
#undef TNS_Z
    TNS_HIP(hipGetLastError());
}

void dgemm_batch(Context* ctx, int m, int n, int k, double alpha, const double* a, int lda, long long sa,
                 const double* b, int ldb, long long sb, double beta, double* c, int ldc, long long sc, int batch,
                 int kmode)
{
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0) return;
    launch_d_batch<64, 64, 16, 32, 32>(ctx, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch, kmode);
    TNS_HIP(hipGetLastError());
}

void zgemm_batch(Context* ctx, int m, int n, int k, cplx alpha, const cplx* a, int lda, long long sa, const cplx* b,
                 int ldb, long long sb, cplx beta, cplx* c, int ldc, long long sc, int batch, int kmode)
{
    if (m <= 0 || n <= 0 || k <= 0 || batch <= 0) return;
    int flags = 0;
    if (alpha.x == 1.0 && alpha.y == 0.0) flags |= 1;
    if (alpha.y == 0.0) flags |= 2;
    if (beta.x == 0.0 && beta.y == 0.0) flags |= 4;
    if (beta.x == 1.0 && beta.y == 0.0) flags |= 8;
    if (tiles(m, n, 64, 64) * batch >= 512)
        launch_z_batch<64, 64, 32, 32, 32>(ctx, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch, flags,
                                           kmode);
    else
        launch_z_batch<64, 32, 32, 32, 16>(ctx, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch, flags,
                                           kmode);
    TNS_HIP(hipGetLastError());
}

void dgemm(Context* ctx, int m, int n, int k, double alpha, const double* a, int lda, const double* b, int ldb,
           double beta, double* c, int ldc)
{
    dgemm_ex(ctx, 0, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void grouped_gemm_tile(int cfg, int& bm, int& bn)
{
    bm = 64;
    bn = cfg == 1 ? 64 : 32;
}

void zgrouped_gemm(Context* ctx, const GemmRec<zd>* recs, const int* starts, int count, int blocks, int cfg)
{
    if (count <= 0 || blocks <= 0) return;
    // Resident workgroups: 256 CUs, 2 (64 x 64) or 4 (64 x 32) per CU.
    const int resident = cfg == 1 ? 512 : 1024;
    const int grid = blocks < resident ? blocks : resident;
    if (cfg == 1)
        hipLaunchKernelGGL((zgemm_grouped_kernel<64, 64, 16, 32, 32>), dim3(grid), dim3(256), 0, ctx->stream, recs,
                           starts, count, blocks);
    else
        hipLaunchKernelGGL((zgemm_grouped_kernel<64, 32, 16, 32, 16>), dim3(grid), dim3(256), 0, ctx->stream, recs,
                           starts, count, blocks);
    TNS_HIP(hipGetLastError());
}

void zgemm(Context* ctx, int m, int n, int k, cplx alpha, const cplx* a, int lda, const cplx* b, int ldb, cplx beta,
           cplx* c, int ldc)
{
    zgemm_ex(ctx, 0, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

}  // namespace tns
