// LU with partial pivoting and its solves on the device, real and complex.
//
// Replaces Lu.Factor/Solve/SolveTransposed (Lu.cs:230-378), the generic
// BlockedLu (BlockedLu.cs:45-381) and the triangular solves
// (Triangular.cs:26-206). The algorithm is the reference's right-looking
// blocked getrf: factor an m x nb panel, apply its interchanges to the
// columns either side, U12 := L11^-1 A12, A22 -= A21 U12 on the matrix
// cores. What moves to the device:
//
//  * The panel (getf2) runs in one workgroup with the panel held in
//    registers: thread t owns rows t, t+T, ... across every panel column.
//    Each column is a wavefront argmax (shuffles, then LDS across waves), one
//    barrier, the pivot row published through LDS, a second barrier, then
//    every thread scales and updates its own rows. The pivot rule is the
//    reference's exactly: largest |x| (|Re|+|Im| for complex), ties to the
//    lowest row, a NaN never wins unless it sits on the diagonal row, where
//    the sequential scan starts. A panel wider than the register file holds
//    for its height is factored in column chunks, right-looking (factor a
//    chunk, move its interchanges, solve and update the rest of the panel).
//  * Row interchanges are carried with the rows as payloads: the row's
//    original index (the permutation P, for the solves) and its index at the
//    start of the panel kernel. At its end the kernel lists the rows whose
//    payload moved as (destination, source) pairs, and one kernel applies
//    that list to every matrix column outside the kernel's columns: a gather
//    of at most 2w rows per column instead of w dependent swaps.
//  * Triangular solves with few right-hand sides run in one launch
//    (trsv_flow_kernel): chains of workgroups on one XCD, each owning
//    64-row diagonal blocks, streaming the triangle's tiles ahead and
//    exchanging only the solved x blocks through the L2. Where that launch
//    cannot run, one workgroup per column (or one launch per diagonal
//    block at large orders). Many right-hand sides (the Pade denominator's
//    n) use blocked substitution: each diagonal block
//    solved by workgroups of 32 columns with the block in LDS (the unit
//    lower one, the LU's U12 solve, by one wavefront per column with x_k
//    broadcast by readlane), the off-diagonal update as a GEMM.
#include <algorithm>
#include <climits>
#include <type_traits>
#include <utility>
#include <vector>

#include "grouped.hpp"
#include "tns_internal.hpp"
#include "tns_scalar.hpp"

namespace tns {
namespace {

// Smallest normal number of the working precision: below it the reference
// divides by the pivot instead of multiplying by its reciprocal
// (Lu.cs:407-412). The float form serves the mixed-precision factorisation.
template <typename T>
constexpr double safe_min()
{
    return sizeof(T) == 4 ? 1.1754943508222875e-38 : 2.2250738585072014e-308;
}

// Larger key wins; equal keys go to the lower row. Written with bitwise
// operators so it compiles to compares and selects, not branches.
__device__ __forceinline__ void best_combine(double& bv, int& bi, double v, int i)
{
    const bool take = (v > bv) | ((v == bv) & (i < bi));
    bv = take ? v : bv;
    bi = take ? i : bi;
}

__device__ __forceinline__ void wave_best(double& bv, int& bi)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        double ov = __shfl_xor(bv, off, 64);
        int oi = __shfl_xor(bi, off, 64);
        best_combine(bv, bi, ov, oi);
    }
}

// One step of the argmax all-reduce across lanes paired by a DPP pattern
// (register-to-register, no LDS round trip).
template <int CTRL>
__device__ __forceinline__ void dpp_best(double& bv, int& bi)
{
    const long long b = __double_as_longlong(bv);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    const int oi = __builtin_amdgcn_update_dpp(0, bi, CTRL, 0xF, 0xF, false);
    best_combine(bv, bi, __longlong_as_double(((long long)hi << 32) | (unsigned)lo), oi);
}

// wave_best without LDS: the in-row all-reduce as DPP moves (quad_perm
// [1,0,3,2] and [2,3,0,1], row_ror 4 and 8), then row_bcast 15 and 31 fold
// rows 0..3 into lane 63, which every lane reads back. The combine (largest
// key, lowest row on ties) is associative and commutative, so the pairing
// order does not change the result.
template <int CTRL, int ROWS>
__device__ __forceinline__ void dpp_best_rows(double& bv, int& bi)
{
    const long long b = __double_as_longlong(bv);
    const long long idv = __double_as_longlong(-2.0);
    const int lo = __builtin_amdgcn_update_dpp((int)idv, (int)b, CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(idv >> 32), (int)(b >> 32), CTRL, ROWS, 0xF, false);
    const int oi = __builtin_amdgcn_update_dpp(INT_MAX, bi, CTRL, ROWS, 0xF, false);
    best_combine(bv, bi, __longlong_as_double(((long long)hi << 32) | (unsigned)lo), oi);
}

__device__ __forceinline__ void wave_best_dpp(double& bv, int& bi)
{
    dpp_best<0xB1>(bv, bi);
    dpp_best<0x4E>(bv, bi);
    dpp_best<0x124>(bv, bi);
    dpp_best<0x128>(bv, bi);
    dpp_best_rows<0x142, 0xA>(bv, bi);  // row_bcast:15 into rows 1 and 3
    dpp_best_rows<0x143, 0xC>(bv, bi);  // row_bcast:31 into rows 2 and 3
    const long long b = __double_as_longlong(bv);
    const int lo = __builtin_amdgcn_readlane((int)b, 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    bi = __builtin_amdgcn_readlane(bi, 63);
    bv = __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Pivot-search key of row `row` in column j: the pivot measure, with a NaN
// never winning except on the diagonal row, where the reference's sequential
// scan starts and so returns it.
template <typename T>
__device__ __forceinline__ double pivot_key(T x, int row, int j)
{
    double v = s_pivmag(x);
    if (isnan(v)) v = (row == j) ? INFINITY : -1.0;
    return v;
}

template <typename T>
__device__ __forceinline__ T shfl_t(T v, int src);
template <>
__device__ __forceinline__ double shfl_t<double>(double v, int src)
{
    return __shfl(v, src, 64);
}
template <>
__device__ __forceinline__ zd shfl_t<zd>(zd v, int src)
{
    return make_zd(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

template <>
__device__ __forceinline__ float shfl_t<float>(float v, int src)
{
    return __shfl(v, src, 64);
}

__device__ __forceinline__ float shfl_xor_t(float v, int off) { return __shfl_xor(v, off, 64); }
__device__ __forceinline__ double shfl_xor_t(double v, int off) { return __shfl_xor(v, off, 64); }
__device__ __forceinline__ zd shfl_xor_t(zd v, int off)
{
    return make_zd(__shfl_xor(v.x, off, 64), __shfl_xor(v.y, off, 64));
}

// v[c] for a runtime c, and v[c] := x, by bitwise selection over
// compile-time indices, so the array stays in registers. (A chain of
// conditional moves is folded by the compiler back into an indexed load,
// which demotes the array to scratch memory; masks are not.)
__device__ __forceinline__ unsigned long long bits_of(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ __forceinline__ double double_of(unsigned long long b) { return __longlong_as_double((long long)b); }

template <int N>
__device__ __forceinline__ double pick(const double (&v)[N], int c)
{
    unsigned long long acc = 0;
#pragma unroll
    for (int q = 0; q < N; q++) acc |= bits_of(v[q]) & (0ull - (unsigned long long)(q == c));
    return double_of(acc);
}

template <int N>
__device__ __forceinline__ zd pick(const zd (&v)[N], int c)
{
    unsigned long long re = 0, im = 0;
#pragma unroll
    for (int q = 0; q < N; q++) {
        const unsigned long long mask = 0ull - (unsigned long long)(q == c);
        re |= bits_of(v[q].x) & mask;
        im |= bits_of(v[q].y) & mask;
    }
    return zd{double_of(re), double_of(im)};
}

// Diagonal block order of the triangular solves: 64 real, 32 complex (a
// 64-element complex column does not fit the register file, nor a 64 x 64
// complex block in one workgroup's LDS).
template <typename T>
constexpr int tri_block() { return sizeof(T) == 16 ? 32 : 64; }

// y if take, else x, by masks.
__device__ __forceinline__ double blend(double x, double y, bool take)
{
    const unsigned long long mask = 0ull - (unsigned long long)take;
    return double_of((bits_of(x) & ~mask) | (bits_of(y) & mask));
}
__device__ __forceinline__ zd blend(zd x, zd y, bool take) { return zd{blend(x.x, y.x, take), blend(x.y, y.y, take)}; }

// ---------------------------------------------------------------------------
// Panel factorization, register resident. m <= blockDim.x * R, w <= NB.
// piv[j] receives the global pivot row (off + local row); perm (indexed by
// local row, offset by the caller) is the payload of original row indices
// that travels with the rows. At the end the kernel writes the move list
// (destination, source) of the rows that changed place to `list`, and
// info[0] is min-reduced with the first zero-pivot column.
//
// Thread t holds rows t, t + nt, ... (its slots) across every column. Row
// interchanges do not move data: each slot carries the row position it has
// in the current order (pos), and step j marks the pivot's slot done (it is
// row j of U from then on) and gives the slot that was at position j the
// pivot's old position, which is the reference's swap of rows j and p. The
// slots are stored to their final positions at the end.
//
// The column loop is not unrolled, yet every register access uses a
// compile-time index: the loop over a row's columns is unrolled and the
// runtime column j is met by uniform conditions. The next pivot column is
// copied aside (nx) as it is updated, so the pivot search and the
// multiplier need no runtime-indexed read. Updates with a zero entry of the
// pivot row are skipped, as the reference's loop does (one ballot gives the
// nonzero pattern); done slots are masked off.
// ---------------------------------------------------------------------------
template <typename T, int NB, int R, int NT, typename PA, typename PI>
__device__ __forceinline__ void getf2_reg_body(int m, int w, PA __restrict__ a, int lda, PI __restrict__ piv, int off,
                                               PI __restrict__ perm, PI __restrict__ list, int cap, PI __restrict__ info)
{
    static_assert(NB <= 64 && NB % 4 == 0, "the nonzero mask is one ballot; columns go in groups of 4");
    constexpr int GW = NB < 8 ? NB : 8;  // columns per group of pivot-row reads
    __shared__ double red_v[2][NT / 64];
    __shared__ int red_i[2][NT / 64];
    __shared__ T cand[2][NT / 64][NB];  // each wave's best row, before the vote
    __shared__ int spiv[NB];
    __shared__ int moves;

    const int nt = blockDim.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    int first_zero = -1;

    T v[R][NB];
    T nx[R];
    int pos[R];
    unsigned done = 0;  // bit r: slot r is a pivot row already, or padding
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int row = tid + r * nt;
        const bool ok = row < m;
#pragma unroll
        for (int c = 0; c < NB; c++) v[r][c] = load_if(a + row + (size_t)c * lda, ok && c < w);
        nx[r] = v[r][0];
        pos[r] = row;
        if (!ok) done |= 1u << r;
    }
    if (tid == 0) moves = 0;

    const int jl = m < w ? m : w;
#pragma unroll 1
    for (int j = 0; j < jl; ++j) {
        const int buf = j & 1;
        double bv = -2.0;
        int bi = INT_MAX;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const bool ok = !((done >> r) & 1u);
            const double key = pivot_key(nx[r], pos[r], j);
            best_combine(bv, bi, ok ? key : -2.0, ok ? pos[r] : INT_MAX);
        }
        wave_best_dpp(bv, bi);
        // Every wave publishes its best row with its vote, so one barrier
        // serves the whole step: the winning wave's row is the pivot row.
        if (lane == 0) {
            red_v[buf][wave] = bv;
            red_i[buf][wave] = bi;
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const bool live = !((done >> r) & 1u);
            if (live && pos[r] == bi) {
#pragma unroll
                for (int c = 0; c < NB; c++) cand[buf][wave][c] = v[r][c];
            }
        }
        __syncthreads();
        bv = red_v[buf][0];
        bi = red_i[buf][0];
        int wq = 0;
        for (int q = 1; q < nw; ++q) {
            const double qv = red_v[buf][q];
            const int qi = red_i[buf][q];
            const bool take = (qv > bv) | ((qv == bv) & (qi < bi));
            bv = take ? qv : bv;
            bi = take ? qi : bi;
            wq = take ? q : wq;
        }
        const int p = bi;
        const T* prow = cand[buf][wq];

        const T pv = prow[j];
        const bool zero_pivot = s_iszero(pv);
        // Pivots and the first zero pivot stay on chip until the end: a
        // global store here would make the next barrier wait for it.
        if (tid == 0) spiv[j] = off + p;
        if (zero_pivot && first_zero < 0) first_zero = off + j;
        // The pivot's slot retires as row j; the slot at position j takes
        // the pivot's place. A zero pivot leaves the rows where they are.
        const int pe = zero_pivot ? j : p;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const bool live = !((done >> r) & 1u);
            const bool is_p = live & (pos[r] == pe);
            const bool is_j = live & (pos[r] == j) & !is_p;
            done |= is_p ? (1u << r) : 0u;
            pos[r] = is_p ? j : (is_j ? pe : pos[r]);
        }
        // Multipliers, in nx: x * (1/pivot) when |pivot| >= the safe
        // minimum, x / pivot below it, x itself for a zero pivot. |pivot|
        // >= max(|Re|, |Im|), so the modulus is needed only when both parts
        // are below the safe minimum.
        if (!zero_pivot) {
            bool safe = s_pivmax(pv) >= safe_min<T>();
            if (!safe) safe = s_abs(pv) >= safe_min<T>();
            if (safe) {
                const T recip = s_div(s_one(T{}), pv);
#pragma unroll
                for (int r = 0; r < R; r++) nx[r] = s_mul(nx[r], recip);
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) nx[r] = s_div(nx[r], pv);
            }
        }
        // Columns to update: right of j, inside w, nonzero in the pivot row
        // (one ballot: lane c looks at column c).
        const bool nz = lane < NB && lane < w && lane > j && !s_iszero(prow[lane < NB ? lane : 0]);
        const unsigned long long mask = __ballot(nz);

        // Groups of GW columns: the group's pivot-row entries are read from
        // LDS together, then every live slot of the thread is updated;
        // groups left of j are skipped by a uniform branch. A group whose
        // columns are all to be updated (right of j, nonzero pivot-row
        // entries: the usual case) takes plain arithmetic; otherwise each
        // column is chosen by a uniform select. The multiplier (nx) is
        // column j's new value; column j receives it, and nx moves on to
        // column j + 1, after every group.
#pragma unroll
        for (int g = 0; g < NB; g += GW) {
            if (g + GW <= j) continue;
            T u[GW];
#pragma unroll
            for (int q = 0; q < GW; q++) u[q] = prow[g + q];
            const unsigned long long gbits = (mask >> g) & ((1ull << GW) - 1);
            const bool full = gbits == (1ull << GW) - 1;
#pragma unroll
            for (int r = 0; r < R; r++) {
                if ((done >> r) & 1u) continue;
                const T l = nx[r];
                if (full) {
#pragma unroll
                    for (int q = 0; q < GW; q++) v[r][g + q] = s_add(v[r][g + q], s_mul(s_neg(u[q]), l));
                } else {
#pragma unroll
                    for (int q = 0; q < GW; q++) {
                        const T x = v[r][g + q];
                        v[r][g + q] = s_sel(((gbits >> q) & 1ull) != 0, s_add(x, s_mul(s_neg(u[q]), l)), x);
                    }
                }
            }
        }
        // Column j takes the multipliers; nx moves on to column j + 1.
#pragma unroll
        for (int g = 0; g < NB; g += GW) {
            if (j < g || j >= g + GW) continue;
#pragma unroll
            for (int r = 0; r < R; r++) {
                if ((done >> r) & 1u) continue;
#pragma unroll
                for (int q = 0; q < GW; q++) v[r][g + q] = s_sel(g + q == j, nx[r], v[r][g + q]);
            }
        }
#pragma unroll
        for (int g = 0; g < NB; g += GW) {
            if (j + 1 < g || j + 1 >= g + GW) continue;
#pragma unroll
            for (int r = 0; r < R; r++)
#pragma unroll
                for (int q = 0; q < GW; q++) nx[r] = s_sel(g + q == j + 1, v[r][g + q], nx[r]);
        }
    }

    // The perm payloads move with the rows: every slot reads its own before
    // any is overwritten.
    int p0[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int row = tid + r * nt;
        p0[r] = row < m ? perm[row] : 0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int row = tid + r * nt;
        if (row < m) {
            const int dst = pos[r];
#pragma unroll
            for (int c = 0; c < NB; c++)
                if (c < w) a[dst + (size_t)c * lda] = v[r][c];
            perm[dst] = p0[r];
            if (dst != row) {
                const int at = atomicAdd(&moves, 1);
                if (at < cap) {
                    list[1 + 2 * at] = dst;
                    list[2 + 2 * at] = row;
                }
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        list[0] = moves;
        for (int j = 0; j < jl; ++j) piv[j] = spiv[j];
        if (first_zero >= 0) atomicMin(info, first_zero);
    }
}


template <typename T, int NB, int R, int NT>
__global__ void __launch_bounds__(NT)
getf2_reg_kernel(int m, int w, T* __restrict__ a, int lda, int* __restrict__ piv, int off, int* __restrict__ perm,
                 int* __restrict__ list, int cap, int* __restrict__ info)
{
    getf2_reg_body<T, NB, R, NT>(m, w, a, lda, piv, off, perm, list, cap, info);
}

// The same panel for many matrices: workgroup g factors record g.
template <typename T, int NB, int R, int NT>
__global__ void __launch_bounds__(NT) getf2_grouped_kernel(const PanelRec<T>* __restrict__ recs)
{
    const PanelRec<T> r = recs[blockIdx.x];
    getf2_reg_body<T, NB, R, NT>(r.m, r.w, as_global(r.a), r.lda, as_global(r.piv), r.off, as_global(r.perm),
                                 as_global(r.list), 2 * r.w, as_global(r.info));
}

// ---------------------------------------------------------------------------
// Panel factorization spread over P workgroups (the whole panel width in one
// launch, no column chunks). Workgroup g holds panel rows [g*512, g*512+512)
// in registers, one row per thread, across every panel column, and does the
// work of getf2_reg_body for its rows with the same pivot rule, interchange
// bookkeeping and update order. Per column:
//  * each wavefront finds its best row by one argmax over order-preserving
//    integer keys (DPP maximum of the high words; the low words, then the
//    row positions, only on ties) and publishes it in LDS; one barrier;
//  * every wavefront reduces the eight wavefront winners the same way, and
//    the lane that holds the workgroup's best row writes that row to LDS;
//    a second barrier;
//  * when P > 1, the workgroups join an all-gather through device memory:
//    each publishes its best key (integer key and row position) and,
//    speculatively, that row's values; one wavefront polls the P keys, picks
//    the winner by the same rule and fetches the winner's row into LDS for a
//    third barrier.
// Every published 64-bit word carries 32 bits of payload and the column's
// sequence number (global column + 1) in its other half, so a reader knows
// a word is current without any fence. Words are written whole by
// workgroup-scope atomic stores, which go to the L2 of the writer's XCD, and
// read whole by agent-scope atomic loads, which skip the reader's L1; the
// working workgroups are blockIdx.x = 8g, which round-robin dispatch places
// on one XCD, so every store and load meets in one L2 (a measured 0.27 us
// one way, against 0.5 us for agent-scope stores, which write through to
// memory). A word that does not arrive keeps its old sequence number, so a
// placement that broke this could only make a wait give up, never return a
// stale value. Records alternate between two buffers by column parity: a
// workgroup can be at most one column ahead of any other (it needs every key
// of a column to leave it). The others exit at once. Every wait gives up
// after SPREAD_SPIN polls or when another workgroup has raised the abort
// word (xbuf[0], written with agent scope), so the kernel always ends; the
// caller turns a raised abort word into an error. P = 1 keeps everything in
// LDS. Rows are stored to their final positions at the end; moved rows are
// appended to `list` (count in list[0], zero at entry) by atomic
// reservation.
// ---------------------------------------------------------------------------
// f(integral_constant<int, 0>) ... f(integral_constant<int, N - 1>), each
// call inlined: a loop whose index is a constant in every iteration.
template <typename F, int... J>
__device__ __forceinline__ void static_for_seq(F& f, std::integer_sequence<int, J...>)
{
    (f(std::integral_constant<int, J>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F& f)
{
    static_for_seq(f, std::make_integer_sequence<int, N>{});
}

// f(integral_constant<int, j>) for a run-time j in [LO, HI), by a binary
// search of uniform branches: the callee sees j as a constant.
template <int LO, int HI, typename F>
__device__ __forceinline__ void dispatch(int j, F& f)
{
    if constexpr (HI - LO == 1) {
        f(std::integral_constant<int, LO>{});
    } else {
        constexpr int MID = (LO + HI) / 2;
        if (j < MID) dispatch<LO, MID>(j, f);
        else dispatch<MID, HI>(j, f);
    }
}

constexpr int LU_OUTER_FROM = 8192;      // order from which getrf_impl takes outer steps
constexpr int SPREAD_NT = 512;           // threads per workgroup, one panel row each
constexpr int SPREAD_MAXP = 32;          // workgroups: panels of up to 16384 rows
constexpr int SPREAD_ROW_WORDS = 128;    // exchange words of one row: 64 doubles' halves
constexpr int SPREAD_REC = 3 + SPREAD_ROW_WORDS;  // key (lo, hi; row) + row
constexpr int SPREAD_SPIN = 1 << 22;     // polls before a wait gives up
// Exchange buffer length in 64-bit words: the abort word, then two parities
// of SPREAD_MAXP records.
constexpr size_t SPREAD_XWORDS = 1 + 2 * (size_t)SPREAD_MAXP * SPREAD_REC;
// One panel's move list in ints: the count, then (destination, source) for
// each moved row, at most two per column of a 64-column panel.
constexpr int SPREAD_LIST = 1 + 2 * (2 * 64);

__device__ __forceinline__ void xstore(unsigned long long* p, unsigned seq, unsigned payload)
{
    __hip_atomic_store(p, ((unsigned long long)seq << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ unsigned long long xload(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until `count` consecutive words at p carry sequence number seq;
// false if it gave up (too many polls, or the abort word is raised).
template <int N>
__device__ __forceinline__ bool xwait(const unsigned long long* p, unsigned seq, const unsigned long long* abort_word,
                                      unsigned (&payload)[N])
{
    for (int it = 0; it < SPREAD_SPIN; ++it) {
        bool all = true;
#pragma unroll
        for (int q = 0; q < N; q++) {
            const unsigned long long x = xload(p + q);
            payload[q] = (unsigned)x;
            all &= (unsigned)(x >> 32) == seq;
        }
        if (all) return true;
        if ((it & 63) == 63 && xload(abort_word) != 0) return false;
    }
    return false;
}

// One DPP step of an unsigned maximum (minimum) across lanes; 0 (all ones)
// is the identity, so lanes a row mask leaves out keep their value and the
// compiler folds the move into the maximum.
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ unsigned dpp_umax(unsigned x)
{
    return max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xF, false));
}
template <int CTRL, int ROWS = 0xF>
__device__ __forceinline__ unsigned dpp_umin(unsigned x)
{
    return min(x, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)x, CTRL, ROWS, 0xF, false));
}

// Maximum (minimum) over the wavefront as a uniform value: quad_perm
// [1,0,3,2] and [2,3,0,1], row_ror 4 and 8 reduce each row of 16 lanes;
// row_bcast 15 and 31 fold the rows into lane 63. ROW: only lanes 0..15
// hold values (the others the identity), so the row reduction suffices.
template <bool ROW>
__device__ __forceinline__ unsigned wave_umax(unsigned x)
{
    x = dpp_umax<0xB1>(x);
    x = dpp_umax<0x4E>(x);
    x = dpp_umax<0x124>(x);
    x = dpp_umax<0x128>(x);
    if constexpr (ROW) return __builtin_amdgcn_readfirstlane(x);
    x = dpp_umax<0x142, 0xA>(x);
    x = dpp_umax<0x143, 0xC>(x);
    return __builtin_amdgcn_readlane(x, 63);
}
template <bool ROW>
__device__ __forceinline__ unsigned wave_umin(unsigned x)
{
    x = dpp_umin<0xB1>(x);
    x = dpp_umin<0x4E>(x);
    x = dpp_umin<0x124>(x);
    x = dpp_umin<0x128>(x);
    if constexpr (ROW) return __builtin_amdgcn_readfirstlane(x);
    x = dpp_umin<0x142, 0xA>(x);
    x = dpp_umin<0x143, 0xC>(x);
    return __builtin_amdgcn_readlane(x, 63);
}

// Order-preserving integer form of a pivot key: rows out of the search 0, a
// NaN off the diagonal 1, a measure v >= 0 (or +inf) its bits + 2. Integer
// order is then the order of pivot_key, ties included.
__device__ __forceinline__ unsigned long long key_bits(double v, bool out)
{
    return out ? 0ull : (v < 0.0 ? 1ull : bits_of(v) + 2ull);
}

// The wavefront's largest key, ties to the lowest position: the high words
// first, the low words only where high words tie, positions only where
// whole keys tie. kmax and pmax receive the winner (pmax = INT_MAX when
// every key is 0); returns its lane.
template <bool ROW>
__device__ __forceinline__ int wave_argmax_u(unsigned long long key, int pos, unsigned long long& kmax, int& pmax)
{
    const unsigned hi = (unsigned)(key >> 32), lo = (unsigned)key;
    const unsigned mh = wave_umax<ROW>(hi);
    bool c = hi == mh;
    unsigned long long hit = __ballot(c);
    if (__popcll(hit) != 1) {
        const unsigned ml = wave_umax<ROW>(c ? lo : 0u);
        c = c && lo == ml;
        hit = __ballot(c);
        if (__popcll(hit) != 1) {
            const unsigned mp = wave_umin<ROW>(c ? (unsigned)pos : 0xFFFFFFFFu);
            hit = __ballot(c && (unsigned)pos == mp);
        }
    }
    const int ln = __ffsll((long long)hit) - 1;
    kmax = ((unsigned long long)mh << 32) | (unsigned)__builtin_amdgcn_readlane((int)lo, ln);
    pmax = kmax == 0 ? INT_MAX : __builtin_amdgcn_readlane(pos, ln);
    return ln;
}

// v[c] -= prow[c] * l for c in [C1, NB): prow (LDS) read in chunks of 32
// bytes, each chunk's reads issued before the previous chunk's arithmetic,
// so the LDS latency is paid once per chunk pair instead of once per read.
template <int C1, typename T, int NB>
__device__ __forceinline__ void update_tail(T (&v)[NB], const T* __restrict__ prow, T l)
{
    constexpr int K = 32 / (int)sizeof(T);
    constexpr int NCH = C1 < NB ? (NB - C1 + K - 1) / K : 0;
    if constexpr (NCH > 0) {
        T u[2][K];
#pragma unroll
        for (int q = 0; q < K; q++) u[0][q] = C1 + q < NB ? prow[C1 + q < NB ? C1 + q : 0] : s_zero(T{});
        auto chunk = [&](auto kc) __attribute__((always_inline)) {
            constexpr int KI = decltype(kc)::value;
            if constexpr (KI + 1 < NCH) {
#pragma unroll
                for (int q = 0; q < K; q++) {
                    const int c = C1 + (KI + 1) * K + q;
                    u[(KI + 1) & 1][q] = c < NB ? prow[c < NB ? c : 0] : s_zero(T{});
                }
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < K; q++) {
                const int c = C1 + KI * K + q;
                if (c < NB) v[c < NB ? c : 0] = s_add(v[c < NB ? c : 0], s_mul(s_neg(u[KI & 1][q]), l));
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        static_for<NCH>(chunk);
    }
}

template <typename T, int NB>
__global__ void __launch_bounds__(SPREAD_NT)
getf2_spread_kernel(int m, int w, T* __restrict__ a, int lda, int* __restrict__ piv, int off, int* __restrict__ perm,
                    int* __restrict__ list, int P, unsigned long long* __restrict__ xbuf, int* __restrict__ info)
{
    static_assert(NB * sizeof(T) <= 4 * SPREAD_ROW_WORDS, "a panel row fits 128 exchange words");
    static_assert(NB <= 64 && NB % 8 == 0, "one ballot covers the row; columns go in groups of 8");
    constexpr int NT = SPREAD_NT, NW = NT / 64, GW = 8, NG = NB / GW;
    constexpr int WPE = (int)sizeof(T) / 4;  // exchange words per element
    constexpr int RW = NB * WPE;              // exchange words per row
    int g = 0;
    if (P > 1) {
        if (blockIdx.x & 7) return;
        g = blockIdx.x >> 3;
    }
    __shared__ unsigned long long red_k[2][NW];  // each wavefront's best key
    __shared__ int red_i[2][NW];                 // and its row position
    __shared__ T wrow[2][NB];                    // the workgroup's best row
    __shared__ T xrow[2][NB];                    // the winning row, from the exchange
    __shared__ int xp[2];                        // its position
    __shared__ int xabort;
    __shared__ int spiv[NB];
    __shared__ int moves;
    __shared__ int base;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int row = g * NT + tid;
    const bool in = row < m;
    T v[NB];
#pragma unroll
    for (int c = 0; c < NB; c++) v[c] = load_if(a + row + (size_t)c * lda, in && c < w);
    const int p0 = in ? perm[row] : 0;
    T nx = v[0];
    int pos = row;
    bool done = !in;
    int first_zero = -1;
    if (tid == 0) {
        moves = 0;
        xabort = 0;
    }
    const unsigned long long* abort_word = xbuf;
    const int jl = m < w ? m : w;
    int steps = 0;
    bool stop = false;

    // Columns in groups of GW: the group index G is a constant, the column
    // inside it (jj) a run-time index. Columns of later groups are all right
    // of j, so the pivot row and the update reach them with constant indices
    // and no test; the current group's columns are reached through a switch
    // on jj. nx is the row's column j, for the pivot search.
    auto group = [&](auto gc) __attribute__((always_inline)) {
        constexpr int G = decltype(gc)::value;
        constexpr int C0 = G * GW;
#pragma unroll 1
        for (int jj = 0; jj < GW; ++jj) {
            const int j = C0 + jj;
            if (stop || j >= jl) {
                stop = true;
                break;
            }
            const int buf = j & 1;
            unsigned long long kw = 0ull;
            int pw = INT_MAX;
            if (__ballot(!done) != 0) wave_argmax_u<false>(key_bits(done ? 0.0 : pivot_key(nx, pos, j), done), pos, kw, pw);
            if (lane == 0) {
                red_k[buf][wave] = kw;
                red_i[buf][wave] = pw;
            }
            __syncthreads();
            // Workgroup best: lane q reads wavefront q's record.
            unsigned long long kg;
            int p;
            wave_argmax_u<true>(lane < NW ? red_k[buf][lane < NW ? lane : 0] : 0ull,
                                lane < NW ? red_i[buf][lane < NW ? lane : 0] : INT_MAX, kg, p);
            // The lane holding that row writes it (from column C0 on).
            if (!done && pos == p) {
#pragma unroll
                for (int c = C0; c < NB; c++) wrow[buf][c] = v[c];
            }
            __syncthreads();
            const T* prow = wrow[buf];
            if (P > 1) {
                const unsigned seq = (unsigned)(off + j + 1);
                unsigned long long* recs = xbuf + 1 + (size_t)buf * SPREAD_MAXP * SPREAD_REC;
                unsigned long long* mine = recs + (size_t)g * SPREAD_REC;
                const unsigned* words = reinterpret_cast<const unsigned*>(prow);
                // Only the columns from j on are read from a fetched row,
                // in pairs of words from an even word on.
                const int lo = (j * WPE) & ~1;
                if (tid >= lo && tid < RW) {
                    xstore(mine + 3 + tid, seq, words[tid]);
                } else if (tid >= SPREAD_ROW_WORDS && tid < SPREAD_ROW_WORDS + 3) {
                    const int q = tid - SPREAD_ROW_WORDS;
                    xstore(mine + q, seq, q == 0 ? (unsigned)kg : q == 1 ? (unsigned)(kg >> 32) : (unsigned)p);
                }
                if (wave == 0) {
                    unsigned long long gk = 0ull;
                    int gi = INT_MAX;
                    bool ok = true;
                    if (lane < P) {
                        unsigned kk[3];
                        ok = xwait<3>(recs + (size_t)lane * SPREAD_REC, seq, abort_word, kk);
                        gk = ((unsigned long long)kk[1] << 32) | kk[0];
                        gi = (int)kk[2];
                    }
                    unsigned long long gkm;
                    int gim;
                    const int gw = wave_argmax_u<false>(gk, gi, gkm, gim);
                    unsigned rw[2] = {0u, 0u};
                    const bool need = 2 * lane >= lo && 2 * lane < RW;
                    if (__all(ok) && need)
                        ok = xwait<2>(recs + (size_t)gw * SPREAD_REC + 3 + 2 * lane, seq, abort_word, rw);
                    if (need)
                        reinterpret_cast<unsigned long long*>(xrow[buf])[lane] =
                            ((unsigned long long)rw[1] << 32) | rw[0];
                    if (lane == 0) xp[buf] = gim;
                    if (!__all(ok) && lane == 0) {
                        xabort = 1;
                        __hip_atomic_store(xbuf, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                __syncthreads();
                if (xabort) {
                    stop = true;
                    break;
                }
                p = xp[buf];
                prow = xrow[buf];
            }

            const T pv = prow[j];
            const bool zero_pivot = s_iszero(pv);
            if (tid == 0) spiv[j] = off + p;
            if (zero_pivot && first_zero < 0) first_zero = off + j;
            // The pivot's row retires as row j of U; the row at position j
            // takes the pivot's place. A zero pivot leaves the rows in place.
            const int pe = zero_pivot ? j : p;
            {
                const bool is_p = !done & (pos == pe);
                const bool is_j = !done & (pos == j) & !is_p;
                done = done | is_p;
                pos = is_p ? j : (is_j ? pe : pos);
            }
            steps = j + 1;
            if (!done) {
                // Multiplier: x * (1/pivot) when |pivot| >= the safe minimum,
                // x / pivot below it, x itself for a zero pivot.
                T l = nx;
                if (!zero_pivot) {
                    bool safe = s_pivmax(pv) >= safe_min<T>();
                    if (!safe) safe = s_abs(pv) >= safe_min<T>();
                    l = safe ? s_mul(l, s_div(s_one(T{}), pv)) : s_div(l, pv);
                }
                // Column j takes the multiplier and every column right of it
                // is updated. For a finite multiplier a zero pivot-row entry
                // leaves the value as it is, as the reference's skip does; a
                // non-finite one times zero is NaN, so there the skip is
                // followed exactly (the nonzero entries by one ballot).
                unsigned long long keep = ~0ull;
                if (!s_isfinite(l)) {
                    const bool nz = lane < NB && lane < w && lane > j && !s_iszero(prow[lane < NB ? lane : 0]);
                    keep = __ballot(nz);
                }
                if (keep == ~0ull) {
                    // The group's columns right of j by a switch on jj (the
                    // empty asm keeps the eight cases apart, so v stays in
                    // registers), then the later groups.
                    auto cols = [&](auto jc) __attribute__((always_inline)) {
                        constexpr int JJ = decltype(jc)::value;
#pragma unroll
                        for (int q = JJ + 1; q < GW; q++) v[C0 + q] = s_add(v[C0 + q], s_mul(s_neg(prow[C0 + q]), l));
                        v[C0 + JJ] = l;
                        asm volatile("; column %0 of the group" ::"n"(JJ));
                    };
                    dispatch<0, GW>(jj, cols);
                    update_tail<C0 + GW>(v, prow, l);
                } else {
#pragma unroll
                    for (int q = 0; q < GW; q++) {
                        const int c = C0 + q;
                        const T nv = s_add(v[c], s_mul(s_neg(prow[c]), l));
                        const bool upd = q > jj && ((keep >> c) & 1ull) != 0;
                        v[c] = s_sel(q == jj, l, s_sel(upd, nv, v[c]));
                    }
#pragma unroll
                    for (int c = C0 + GW; c < NB; c++)
                        v[c] = s_sel(((keep >> c) & 1ull) != 0, s_add(v[c], s_mul(s_neg(prow[c]), l)), v[c]);
                }
            }
            // nx moves on to column j + 1.
            T nn = C0 + GW < NB ? v[C0 + GW < NB ? C0 + GW : 0] : nx;
#pragma unroll
            for (int q = 1; q < GW; q++) nn = s_sel(q == jj + 1, v[C0 + q], nn);
            nx = nn;
        }
    };
    static_for<NG>(group);

    // Rows to their final positions. Every workgroup loaded its rows and
    // payloads before column 0, and none gets past column 0 before all have
    // joined it, so no store here can overtake another workgroup's load.
    int at = -1;
    if (in) {
        const int dst = pos;
#pragma unroll
        for (int c = 0; c < NB; c++)
            if (c < w) a[dst + (size_t)c * lda] = v[c];
        perm[dst] = p0;
        if (dst != row) at = atomicAdd(&moves, 1);
    }
    __syncthreads();
    if (tid == 0) base = moves > 0 ? atomicAdd(list, moves) : 0;
    __syncthreads();
    if (at >= 0 && base + at < (SPREAD_LIST - 1) / 2) {
        list[1 + 2 * (base + at)] = pos;
        list[2 + 2 * (base + at)] = row;
    }
    if (g == 0 && tid == 0) {
        for (int q = 0; q < steps; ++q) piv[q] = spiv[q];
        if (first_zero >= 0) atomicMin(info, first_zero);
    }
}

// Zeroes the spread panel's exchange words and move-list counters and sets
// up the row payloads (d_perm = identity, rc = identity when given).
__global__ void spread_init_kernel(int m, int* __restrict__ perm, int* __restrict__ rc,
                                   unsigned long long* __restrict__ xbuf, int xwords, int* __restrict__ lists,
                                   int nlist, int list_stride)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        perm[i] = i;
        if (rc) rc[i] = i;
    }
    if (i < xwords) xbuf[i] = 0ull;
    if (i < nlist) lists[(size_t)i * list_stride] = 0;
}

// Sets info to -2 when a spread panel gave up waiting (the abort word).
__global__ void spread_check_kernel(const unsigned long long* __restrict__ xbuf, int* __restrict__ info)
{
    if (xbuf[0] != 0ull) *info = -2;
}

// Panel factorization straight from global memory, for panels taller than
// the register-resident kernel holds. One workgroup, column by column.
template <typename T>
__global__ void __launch_bounds__(1024)
getf2_global_kernel(int m, int w, T* __restrict__ a, int lda, int* __restrict__ piv, int off, int* __restrict__ perm,
                    int* __restrict__ rc, int* __restrict__ info)
{
    __shared__ double red_v[16];
    __shared__ int red_i[16];
    __shared__ int s_p;
    const int nt = blockDim.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = nt >> 6;
    const int jl = m < w ? m : w;
    for (int j = 0; j < jl; ++j) {
        T* col = a + (size_t)j * lda;
        double bv = -2.0;
        int bi = INT_MAX;
        for (int row = j + tid; row < m; row += nt) best_combine(bv, bi, pivot_key(col[row], row, j), row);
        wave_best(bv, bi);
        if (lane == 0) {
            red_v[wave] = bv;
            red_i[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double b = red_v[0];
            int i = red_i[0];
            for (int q = 1; q < nw; ++q) best_combine(b, i, red_v[q], red_i[q]);
            s_p = i;
        }
        __syncthreads();
        const int p = s_p;
        const T pv0 = col[p];
        const bool zero_pivot = s_iszero(pv0);
        if (tid == 0) {
            piv[j] = off + p;
            if (zero_pivot) atomicMin(info, off + j);
        }
        if (!zero_pivot && p != j) {
            for (int c = tid; c < w; c += nt) {
                T* cc = a + (size_t)c * lda;
                T t = cc[j];
                cc[j] = cc[p];
                cc[p] = t;
            }
            if (tid == 0) {
                int t = perm[j]; perm[j] = perm[p]; perm[p] = t;
                t = rc[j]; rc[j] = rc[p]; rc[p] = t;
            }
        }
        __syncthreads();
        const T pv = col[j];
        const double mag = s_abs(pv);
        const T recip = s_div(s_one(T{}), pv);
        if (!zero_pivot)
            for (int row = j + 1 + tid; row < m; row += nt)
                col[row] = (mag >= safe_min<T>()) ? s_mul(col[row], recip) : s_div(col[row], pv);
        __syncthreads();
        for (int c = j + 1; c < w; ++c) {
            T* tc = a + (size_t)c * lda;
            const T u = tc[j];
            if (s_iszero(u)) continue;
            const T nu = s_neg(u);
            for (int row = j + 1 + tid; row < m; row += nt) tc[row] = s_add(tc[row], s_mul(nu, col[row]));
        }
        __syncthreads();
    }
}

__global__ void iota_kernel(int n, int* __restrict__ a, int* __restrict__ b)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] = i;
        if (b) b[i] = i;
    }
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

__global__ void finish_info_kernel(int* info)
{
    if (*info == INT_MAX) *info = -1;
}

// Rows of a slice [0, count) whose payload (a global row index; the slice
// starts at global row `base`) says they moved: append (row, source) in
// slice-local indices and reset the payload for the next step. One
// workgroup: the counter starts at zero in LDS and list[0] receives it at
// the end; at most `cap` moves are stored. The list order is irrelevant to
// apply_moves_kernel, which gathers every source before any write.
__global__ void __launch_bounds__(1024)
compact_moves_kernel(int count, int base, int* __restrict__ rowid, int* __restrict__ list, int cap)
{
    __shared__ int moves;
    if (threadIdx.x == 0) moves = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        const int src = rowid[i] - base;
        if (src != i) {
            const int at = atomicAdd(&moves, 1);
            if (at < cap) {
                list[1 + 2 * at] = i;
                list[2 + 2 * at] = src;
            }
            rowid[i] = i + base;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) list[0] = moves;
}

// Apply a move list (rows relative to `a`) to columns [0, c0) and [c1, ncols).
// A workgroup owns 256/QM columns and QM >= count move slots per column:
// every thread reads its source element, one barrier, every thread writes
// its destination, so all reads of a column precede its writes.
template <typename T, int QM, typename PA, typename PL>
__device__ __forceinline__ void apply_moves_body(PA __restrict__ a, int lda, int c0, int c1, int ncols,
                                                 PL __restrict__ list, int block)
{
    constexpr int CPB = 256 / QM;
    const int q = threadIdx.x % QM;
    const int idx = block * CPB + threadIdx.x / QM;
    const int left = c0, right = ncols - c1;
    const int count = list[0] < QM ? list[0] : QM;
    const bool active = idx < left + right && q < count;
    T val{};
    PA col = a;
    if (active) {
        const int j = idx < left ? idx : c1 + (idx - left);
        col = a + (size_t)j * lda;
        val = col[list[2 + 2 * q]];
    }
    __syncthreads();
    if (active) col[list[1 + 2 * q]] = val;
}

template <typename T, int QM>
__global__ void __launch_bounds__(256)
apply_moves_kernel(T* __restrict__ a, int lda, int c0, int c1, int ncols, const int* __restrict__ list)
{
    apply_moves_body<T, QM>(a, lda, c0, c1, ncols, list, blockIdx.x);
}

// Index of the record whose workgroup range holds `bid`: the last record
// with starts[g] <= bid (starts increasing, starts[0] = 0).
__device__ __forceinline__ int find_record(const int* __restrict__ starts, int count, int bid)
{
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= bid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T, int QM>
__global__ void __launch_bounds__(256)
apply_moves_grouped_kernel(const MovesRec<T>* __restrict__ recs, const int* __restrict__ starts, int count)
{
    const MovesRec<T> r = recs[find_record(starts, count, blockIdx.x)];
    apply_moves_body<T, QM>(as_global(r.a), r.lda, r.c0, r.c1, r.ncols, as_global(r.list), blockIdx.x - r.start);
}

// Larger move lists (never produced by nb <= 64, kept for generality).
template <typename T>
__global__ void apply_moves_slow_kernel(T* __restrict__ a, int lda, int c0, int c1, int ncols,
                                        const int* __restrict__ list, T* __restrict__ scratch, int cap)
{
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int left = c0, right = ncols - c1;
    if (idx >= left + right) return;
    int j = idx < left ? idx : c1 + (idx - left);
    const int count = list[0] < cap ? list[0] : cap;
    T* col = a + (size_t)j * lda;
    T* s = scratch + (size_t)idx * count;
    for (int q = 0; q < count; q++) s[q] = col[list[2 + 2 * q]];
    for (int q = 0; q < count; q++) col[list[1 + 2 * q]] = s[q];
}

// Triangular solve op(A) X = B for one diagonal block (order w <= BS) and
// many right-hand sides, in place; TR: 0 none, 1 transpose, 2 conjugate
// transpose. A workgroup of 256 threads owns CB columns of B: thread (g, c)
// holds rows g, g + G, g + 2G, ... of column c in registers (G = 256 / CB
// row groups), op(A) sits in LDS. Step p finalizes x_p (divided by the
// diagonal unless UNIT), passes it through LDS, and every thread subtracts
// x_p times column p of op(A) from its rows still to come: the reference's
// column-oriented substitution (Triangular.cs:45-57, 126-140), in the same
// order for every element. The step loop is unrolled, so the owner's
// register index p / G is a constant.
template <typename T, int BS, int CB, bool LOWER, int TR, bool UNIT, typename PA, typename PB>
__device__ __forceinline__ void trsm_block_body(int w, int ncols, PA __restrict__ A, int lda, PB __restrict__ B,
                                                int ldb, int block)
{
    constexpr int G = 256 / CB;
    constexpr int Q = BS / G;
    static_assert(256 % CB == 0 && BS % G == 0, "block must split over the row groups");
    constexpr bool FORWARD = (LOWER && TR == 0) || (!LOWER && TR != 0);
    __shared__ T D[BS][BS + 1];  // D[i][k] = op(A)(i, k)
    __shared__ T X[BS][CB + 1];
    const int tid = threadIdx.x;
    const int col0 = block * CB;
    const T zero = s_zero(T{});

    for (int e = tid; e < BS * BS; e += 256) {
        const int i = e % BS, k = e / BS;
        T v = zero;
        if (i < w && k < w) {
            v = TR ? A[k + (size_t)i * lda] : A[i + (size_t)k * lda];
            if (TR == 2) v = s_conj(v);
        }
        D[i][k] = v;
    }
    for (int e = tid; e < BS * CB; e += 256) {
        const int i = e % BS, c = e / BS;
        X[i][c] = load_if(B + i + (size_t)(col0 + c) * ldb, i < w && col0 + c < ncols);
    }
    __syncthreads();

    const int c = tid % CB, g = tid / CB;
    T x[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) x[q] = X[g + q * G][c];

#pragma unroll
    for (int s = 0; s < BS; s++) {
        const int p = FORWARD ? s : BS - 1 - s;
        if (p < w) {
            if (g == p % G) {
                T xp = x[p / G];
                if (!UNIT) xp = s_div(xp, D[p][p]);
                x[p / G] = xp;
                X[p][c] = xp;
            }
            __syncthreads();
            const T xp = X[p][c];
            if (!s_iszero(xp)) {
                const T nxp = s_neg(xp);
#pragma unroll
                for (int q = 0; q < Q; q++) {
                    const int i = g + q * G;
                    const bool ahead = FORWARD ? (i > p && i < w) : (i < p);
                    if (ahead) x[q] = s_add(x[q], s_mul(nxp, D[i][p]));
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; q++) X[g + q * G][c] = x[q];
    __syncthreads();
    for (int e = tid; e < BS * CB; e += 256) {
        const int i = e % BS, cc = e / BS;
        if (i < w && col0 + cc < ncols) B[i + (size_t)(col0 + cc) * ldb] = X[i][cc];
    }
}

template <typename T, int BS, int CB, bool LOWER, int TR, bool UNIT>
__global__ void __launch_bounds__(256)
trsm_block_kernel(int w, int ncols, const T* __restrict__ A, int lda, T* __restrict__ B, int ldb)
{
    trsm_block_body<T, BS, CB, LOWER, TR, UNIT>(w, ncols, A, lda, B, ldb, blockIdx.x);
}

// x_k of the wavefront, a uniform value: lane k's register by readlane.
__device__ __forceinline__ double lane_value(double x, int k)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)b, k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ zd lane_value(zd x, int k) { return zd{lane_value(x.x, k), lane_value(x.y, k)}; }
__device__ __forceinline__ float lane_value(float x, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), k));
}

// L X = B in place for the unit lower triangle L of order w <= BS (64 real,
// 32 complex): one wavefront per column of B, lane i holding x_i, the
// columns of L in LDS (D[k][i] = L(i, k), staged once per workgroup).
// Step k takes x_k from lane k by readlane, so no barrier and no LDS round
// trip separates the steps, and the lanes below k update with column k of
// L. Each x_i takes its updates in k order and none for a zero x_k: the
// operations and order of trsm_block_body. Workgroups walk the columns in
// strides of the grid.
template <typename T, int BS>
__global__ void __launch_bounds__(256) trsm_llnu_wave_kernel(int w, int ncols, const T* __restrict__ L, int ldl,
                                                             T* __restrict__ B, int ldb)
{
    __shared__ T D[BS][BS];
    const int tid = threadIdx.x, lane = tid & 63;
    for (int e = tid; e < BS * BS; e += 256) {
        const int i = e % BS, k = e / BS;
        D[k][i] = load_if(L + i + (size_t)k * ldl, i < w && k < w && i > k);
    }
    __syncthreads();
    const bool live = lane < w;
    for (int c = blockIdx.x * 4 + (tid >> 6); c < ncols; c += gridDim.x * 4) {
        T* b = B + (size_t)c * ldb;
        T x = load_if(b + lane, live);
#pragma unroll
        for (int k = 0; k < BS - 1; k++) {
            if (k + 1 >= w) break;
            const T xk = lane_value(x, k);
            if (!s_iszero(xk)) {
                const T d = D[k][lane < BS ? lane : 0];
                if (lane > k && live) x = s_add(x, s_mul(s_neg(xk), d));
            }
        }
        if (live) b[lane] = x;
    }
}

typedef double v4d_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d_t mfma_f64x4(double a, double b, v4d_t c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// One look-ahead step's update of the next panel's nc columns in one launch:
// A22 -= L21 U12 for the m2 rows below the block row, with U12 = L11^-1 A12
// (L11 the unit lower w x w triangle at pa, A12 at b12, L21 at pa + w, A22
// at b12 + w). Every workgroup forms U12 in LDS from A12, one wavefront per
// column with x_k by readlane (the operations and order of
// trsm_llnu_wave_kernel), and leaves A12 as it is: a workgroup storing U12
// there could overtake another's read of A12. The caller solves A12 in
// place later (trsm_llnu, the same bits); nothing reads those rows before.
// Sixteen wavefronts: four columns each for the solve, then one 16 x 16
// block each of the workgroup's 64 rows of A22 on the matrix cores, with
// the operand mapping, k order (groups of
// four from k = 0, zeros past w), MFMA sequence (complex: the 4M order of
// zgemm_tile) and write-back (C + (-acc)) of the GEMM kernels, so every
// element takes the bits the solve and the product give it: the launch
// replaces those two on the critical path, and the gap between them.
template <typename T, int W>
__global__ void __launch_bounds__(1024) next_panel_update_kernel(int m2, int w, int nc, const T* __restrict__ pa,
                                                                 int lda, T* __restrict__ b12)
{
    constexpr bool CPLX = sizeof(T) == 16;
    constexpr int NWAVE = 16;
    constexpr int TN = W / 16;  // 16-column blocks of the widest update
    __shared__ T Ls[W][W];      // Ls[k][i] = L(i, k) below the diagonal
    __shared__ T Us[W][W + 1];  // Us[c][k] = U12(k, c)
    // The wavefront index as a uniform value (scalar loop control below).
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    for (int e = tid; e < W * W; e += 64 * NWAVE) {
        const int i = e % W, k = e / W;
        Ls[k][i] = load_if(pa + i + (size_t)k * lda, i < w && k < w && i > k);
    }
    __syncthreads();
    const bool live = lane < w;
    for (int c = wave; c < nc; c += NWAVE) {
        T x = load_if(b12 + lane + (size_t)c * lda, live);
#pragma unroll
        for (int k = 0; k < W - 1; k++) {
            if (k + 1 >= w) break;
            const T xk = lane_value(x, k);
            // A zero x_k leaves the column as it is (selected).
            const T d = Ls[k][lane < W ? lane : 0];
            x = s_sel(lane > k && live && !s_iszero(xk), s_add(x, s_mul(s_neg(xk), d)), x);
        }
        if (live) Us[c][lane] = x;
    }
    for (int e = tid; e < W * W; e += 64 * NWAVE) {
        // Columns past nc and rows past w are the products' zero padding.
        const int k = e % W, c = e / W;
        if (c >= nc || k >= w) Us[c][k] = s_zero(T{});
    }
    __syncthreads();
    // Wavefront (rb, j): rows r0 + 16 rb of the workgroup's 64, columns 16 j.
    const int rb = wave & 3, j = wave >> 2;
    const int r0 = blockIdx.x * 64 + rb * 16;
    if (j >= TN || 16 * j >= nc || r0 >= m2) return;
    const T* l21 = pa + w;
    T* a22 = b12 + w;
    const int kq = lane >> 4, col_in = lane & 15;
    const int arow = r0 + col_in;
    v4d_t re = v4d_t{0.0, 0.0, 0.0, 0.0}, im = v4d_t{0.0, 0.0, 0.0, 0.0};
    // This lane's entries of L21, all loaded before the products (indexes
    // clamped into the panel; entries past it zeroed by selection): a
    // guarded load per step waited for its data, one memory latency a step.
    constexpr int KQ = W / 4;
    T avs[KQ];
    {
        const int ar = arow < m2 ? arow : m2 - 1;
#pragma unroll
        for (int q = 0; q < KQ; q++) {
            const int k = 4 * q + kq;
            avs[q] = l21[ar + (size_t)(k < w ? k : w - 1) * lda];
        }
    }
    const int kqs = (w + 3) / 4;
#pragma unroll
    for (int q = 0; q < KQ; q++) {
        if (q >= kqs) break;
        const int k = 4 * q + kq;
        const T av = s_sel(arow < m2 && k < w, avs[q], s_zero(T{}));
        const T bv = Us[16 * j + col_in][k < W ? k : 0];
        if constexpr (CPLX) {
            const double a_r = av.x, a_i = av.y, a_n = -a_i;
            re = mfma_f64x4(a_r, bv.x, re);
            im = mfma_f64x4(a_r, bv.y, im);
            re = mfma_f64x4(a_n, bv.y, re);
            im = mfma_f64x4(a_i, bv.x, im);
        } else {
            re = mfma_f64x4(av, bv, re);
        }
    }
    // D layout of v_mfma_f64_16x16x4_f64: lane l, register r holds row
    // 4r + l/16, column l%16 of the block.
    const int col = 16 * j + col_in;
    if (col >= nc) return;
    // The four old entries of C first (all loads in flight), then the stores.
    T cv[4];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int row = r0 + 4 * r + kq;
        cv[r] = a22[(row < m2 ? row : m2 - 1) + (size_t)col * lda];
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int row = r0 + 4 * r + kq;
        if (row >= m2) continue;
        T* p = a22 + row + (size_t)col * lda;
        if constexpr (CPLX) {
            *p = make_zd(cv[r].x + (-re[r]), cv[r].y + (-im[r]));
        } else {
            *p = cv[r] + (-re[r]);
        }
    }
}

// U12 := L11^-1 A12 of many matrices in one launch.
template <typename T>
__global__ void __launch_bounds__(256)
trsm_grouped_kernel(const TrsmRec<T>* __restrict__ recs, const int* __restrict__ starts, int count)
{
    const TrsmRec<T> r = recs[find_record(starts, count, blockIdx.x)];
    trsm_block_body<T, tri_block<T>(), GROUPED_TRSM_COLUMNS, true, 0, true>(
        r.w, r.ncols, as_global(r.l), r.ldl, as_global(r.b), r.ldb, blockIdx.x - r.start);
}

// Whole triangular solve for few right-hand sides: one workgroup of 256
// threads per column of B. op(A) lower runs forward, upper backward, in
// BS-row blocks. Each block's diagonal part is staged in LDS by all threads
// and solved by wavefront 0 with shuffles; the coupling to the other blocks
// is done by all threads, right-looking (in the reference's axpy order) for
// op = none, left-looking dot products for the transposed forms.
template <typename T, bool LOWER, int TR, bool UNIT>
__global__ void __launch_bounds__(256)
trsv_kernel(int n, const T* __restrict__ A, int lda, T* __restrict__ B, int ldb)
{
    constexpr int BS = tri_block<T>();
    __shared__ T D[BS][BS + 1];  // D[i][k] = op(A)(k0 + i, k0 + k)
    __shared__ T xb[BS];
    __shared__ T part[BS];
    T* x = B + (size_t)blockIdx.x * ldb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    auto at = [&](int i, int k) -> T {
        T v = TR ? A[k + (size_t)i * lda] : A[i + (size_t)k * lda];
        return TR == 2 ? s_conj(v) : v;
    };
    constexpr bool FORWARD = (LOWER && TR == 0) || (!LOWER && TR != 0);
    const int nblk = (n + BS - 1) / BS;
    for (int bq = 0; bq < nblk; bq++) {
        const int blk = FORWARD ? bq : nblk - 1 - bq;
        const int k0 = blk * BS;
        const int kb = n - k0 < BS ? n - k0 : BS;

        for (int e = tid; e < BS * BS; e += 256) {
            const int i = e % BS, k = e / BS;
            D[i][k] = (i < kb && k < kb) ? at(k0 + i, k0 + k) : s_zero(T{});
        }
        if (TR != 0) {
            // Left-looking: subtract the already-solved part, a dot product
            // down column (k0 + i) of A per block row i.
            const int lo = FORWARD ? 0 : k0 + kb;
            const int hi = FORWARD ? k0 : n;
            for (int i = wave; i < kb; i += 4) {
                T s0 = s_zero(T{}), s1 = s_zero(T{});
                int k = lo + lane;
                for (; k + 64 < hi; k += 128) {
                    s0 = s_add(s0, s_mul(at(k0 + i, k), x[k]));
                    s1 = s_add(s1, s_mul(at(k0 + i, k + 64), x[k + 64]));
                }
                if (k < hi) s0 = s_add(s0, s_mul(at(k0 + i, k), x[k]));
                T sum = s_add(s0, s1);
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) sum = s_add(sum, shfl_xor_t(sum, off));
                if (lane == 0) part[i] = sum;
            }
        }
        __syncthreads();

        if (wave == 0 && lane < BS) {
            T xi = load_if(x + k0 + lane, lane < kb);
            if (TR != 0 && lane < kb) xi = s_sub(xi, part[lane]);
            if (FORWARD) {
                for (int p = 0; p < kb; p++) {
                    if (!UNIT && lane == p) xi = s_div(xi, D[p][p]);
                    const T xp = shfl_t(xi, p);
                    if (lane > p && lane < kb && !s_iszero(xp)) xi = s_add(xi, s_mul(s_neg(xp), D[lane][p]));
                }
            } else {
                for (int p = kb - 1; p >= 0; p--) {
                    if (!UNIT && lane == p) xi = s_div(xi, D[p][p]);
                    const T xp = shfl_t(xi, p);
                    if (lane < p && !s_iszero(xp)) xi = s_add(xi, s_mul(s_neg(xp), D[lane][p]));
                }
            }
            if (lane < kb) {
                x[k0 + lane] = xi;
                xb[lane] = xi;
            }
        }
        __syncthreads();

        if (TR == 0) {
            // Right-looking: push this block into the rows still to come,
            // eight independent loads in flight per step.
            const int r0 = FORWARD ? k0 + kb : 0;
            const int r1 = FORWARD ? n : k0;
            for (int i = r0 + tid; i < r1; i += 256) {
                T acc = x[i];
                for (int q = 0; q < kb; q += 8) {
                    T av[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int p = FORWARD ? q + u : kb - 1 - (q + u);
                        av[u] = (q + u < kb) ? at(i, k0 + p) : s_zero(T{});
                    }
#pragma unroll
                    for (int u = 0; u < 8; u++) {
                        const int p = FORWARD ? q + u : kb - 1 - (q + u);
                        if (q + u < kb && !s_iszero(xb[p])) acc = s_add(acc, s_mul(s_neg(xb[p]), av[u]));
                    }
                }
                x[i] = acc;
            }
            __syncthreads();
        }
    }
}

// trsv_kernel with the right-hand side held in LDS for the whole solve
// (dynamic shared memory, n elements) and each diagonal block solved from
// registers: lane i of wavefront 0 holds row i of the block of op(A) and
// its x_i; step p divides on lane p (unless UNIT), reads x_p back with
// readlane and updates the lanes ahead. Same operation order as trsv_kernel.
template <typename T>
__device__ __forceinline__ T readlane_t(T v, int lane);
template <>
__device__ __forceinline__ double readlane_t<double>(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <>
__device__ __forceinline__ float readlane_t<float>(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
template <>
__device__ __forceinline__ zd readlane_t<zd>(zd v, int lane)
{
    return make_zd(readlane_t(v.x, lane), readlane_t(v.y, lane));
}

// NT threads per column and UL entries of A loaded per row before they are
// applied: 512 and 16 (8 complex) from order 256 on, where the latency of
// those loads, paid once per group of UL, sets the time (512: wavefront 0
// holds a whole diagonal block in registers, 128 of them, which 1024
// threads per workgroup would not leave it).
template <typename T, bool LOWER, int TR, bool UNIT, int NT = 256, int UL = (sizeof(T) == 16 ? 4 : 8)>
__global__ void __launch_bounds__(NT)
trsv_lds_kernel(int n, const T* __restrict__ A, int lda, T* __restrict__ B, int ldb)
{
    constexpr int BS = tri_block<T>();
    extern __shared__ unsigned char trsv_smem[];
    T* xs = reinterpret_cast<T*>(trsv_smem);
    __shared__ T part[BS];
    T* xg = B + (size_t)blockIdx.x * ldb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int NW = NT / 64;
    auto at = [&](int i, int k) -> T {
        T v = TR ? A[k + (size_t)i * lda] : A[i + (size_t)k * lda];
        return TR == 2 ? s_conj(v) : v;
    };
    constexpr bool FORWARD = (LOWER && TR == 0) || (!LOWER && TR != 0);
    for (int i = tid; i < n; i += NT) xs[i] = xg[i];
    __syncthreads();

    const int nblk = (n + BS - 1) / BS;
    for (int bq = 0; bq < nblk; bq++) {
        const int blk = FORWARD ? bq : nblk - 1 - bq;
        const int k0 = blk * BS;
        const int kb = n - k0 < BS ? n - k0 : BS;
        if (TR != 0) {
            // Left-looking: the solved part, a dot product down column
            // (k0 + i) of A per block row i.
            const int lo = FORWARD ? 0 : k0 + kb;
            const int hi = FORWARD ? k0 : n;
            for (int i = wave; i < kb; i += NW) {
                T s0 = s_zero(T{}), s1 = s_zero(T{});
                int k = lo + lane;
#pragma unroll 4
                for (; k + 64 < hi; k += 128) {
                    s0 = s_add(s0, s_mul(at(k0 + i, k), xs[k]));
                    s1 = s_add(s1, s_mul(at(k0 + i, k + 64), xs[k + 64]));
                }
                if (k < hi) s0 = s_add(s0, s_mul(at(k0 + i, k), xs[k]));
                T sum = s_add(s0, s1);
#pragma unroll
                for (int off = 32; off >= 1; off >>= 1) sum = s_add(sum, shfl_xor_t(sum, off));
                if (lane == 0) part[i] = sum;
            }
            __syncthreads();
        }
        if (wave == 0) {
            const bool mine = lane < kb;
            T d[BS];
#pragma unroll
            for (int k = 0; k < BS; k++) d[k] = (mine && k < kb) ? at(k0 + lane, k0 + k) : s_zero(T{});
            T xi = load_if(xs + k0 + lane, mine);
            if (TR != 0 && mine) xi = s_sub(xi, part[lane]);
#pragma unroll
            for (int s = 0; s < BS; s++) {
                const int p = FORWARD ? s : BS - 1 - s;
                if (p >= kb) continue;
                if (!UNIT && lane == p) xi = s_div(xi, d[p]);
                const T xp = readlane_t(xi, p);
                if (s_iszero(xp)) continue;
                const bool ahead = FORWARD ? (lane > p && lane < kb) : (lane < p);
                if (ahead) xi = s_add(xi, s_mul(s_neg(xp), d[p]));
            }
            if (mine) xs[k0 + lane] = xi;
        }
        __syncthreads();
        if (TR == 0) {
            // Right-looking: push this block into the rows still to come, in
            // the reference's axpy order, two rows at a time, the next U
            // entries of each row loaded together before they are applied.
            constexpr int U = UL;
            const int r0 = FORWARD ? k0 + kb : 0;
            const int r1 = FORWARD ? n : k0;
            for (int i = r0 + tid; i < r1; i += 2 * NT) {
                const int i2 = i + NT;
                const bool two = i2 < r1;
                T acc = xs[i];
                T acc2 = load_if(xs + i2, two);
                for (int q = 0; q < kb; q += U) {
                    T av[U], av2[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int p = FORWARD ? q + u : kb - 1 - (q + u);
                        const bool in = q + u < kb;
                        av[u] = load_if(A + i + (size_t)(k0 + p) * lda, in);
                        av2[u] = load_if(A + i2 + (size_t)(k0 + p) * lda, in && two);
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        if (q + u >= kb) break;
                        const int p = FORWARD ? q + u : kb - 1 - (q + u);
                        const T xp = xs[k0 + p];
                        if (s_iszero(xp)) continue;
                        const T nx = s_neg(xp);
                        acc = s_add(acc, s_mul(nx, av[u]));
                        if (two) acc2 = s_add(acc2, s_mul(nx, av2[u]));
                    }
                }
                xs[i] = acc;
                if (two) xs[i2] = acc2;
            }
            __syncthreads();
        }
    }
    for (int i = tid; i < n; i += NT) xg[i] = xs[i];
}

// One step of a few-right-hand-side solve op(A) X = B, op(A) lower
// (forward) or upper (backward), op(A) = A, A^T (TR 1) or A^H (TR 2),
// spread over the chip: the
// rows still to come are split into slabs of 256, one workgroup each, and
// every row takes block [k0, k0 + kb)'s final x in the order trsv_lds_kernel
// applies it (column by column, skipping zero entries of x); then the
// workgroup of slab 0, which holds the next diagonal block [next0, next0 +
// nextb), solves it as trsv_lds_kernel does (one wavefront per column,
// readlane substitution). One launch per diagonal block, every element
// computed by the same operations in the same order as trsv_lds_kernel, but
// the updates read the triangle with the whole chip instead of one
// workgroup per column. Columns blockIdx.y * NC ...
template <typename T, bool LOWER, int TR, bool UNIT, int NC>
__global__ void __launch_bounds__(256) trsv_step_kernel(int n, const T* __restrict__ A, int lda, T* __restrict__ B,
                                                        int ldb, int nrhs, int k0, int kb, int next0, int nextb)
{
    // op(A)(i, k), read where the stored triangle keeps it (a transposed
    // form reads along the stored row, contiguous in the row update).
    auto opa = [&](int i, int k, bool ok) __attribute__((always_inline)) {
        T v = load_if(TR == 0 ? A + i + (size_t)k * lda : A + k + (size_t)i * lda, ok);
        if (TR == 2) v = s_conj(v);
        return v;
    };
    constexpr int BS = tri_block<T>();
    constexpr int U = sizeof(T) == 16 ? 4 : 8;
    static_assert(NC <= 4, "one wavefront per column for the diagonal block");
    __shared__ T xb[BS][NC];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c0 = blockIdx.y * NC;
    const int nc = nrhs - c0 < NC ? nrhs - c0 : NC;
    T* X = B + (size_t)c0 * ldb;
    for (int e = tid; e < BS * NC; e += 256) {
        const int p = e % BS, c = e / BS;
        xb[p][c] = load_if(X + k0 + p + (size_t)c * ldb, p < kb && c < nc);
    }
    __syncthreads();
    if (kb > 0) {
        // This slab's rows: below the block (forward) or above it.
        const int lo = LOWER ? k0 + kb + (int)blockIdx.x * 256 : k0 - ((int)blockIdx.x + 1) * 256;
        const int i = lo + tid;
        const bool row = LOWER ? i < n : (i >= 0 && i < k0 - (int)blockIdx.x * 256);
        if (row) {
            T acc[NC];
#pragma unroll
            for (int c = 0; c < NC; c++) acc[c] = load_if(X + i + (size_t)c * ldb, c < nc);
            for (int q = 0; q < kb; q += U) {
                T av[U];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int p = LOWER ? q + u : kb - 1 - (q + u);
                    av[u] = opa(i, k0 + p, q + u < kb);
                }
#pragma unroll
                for (int u = 0; u < U; u++) {
                    if (q + u < kb) {
                        const int p = LOWER ? q + u : kb - 1 - (q + u);
#pragma unroll
                        for (int c = 0; c < NC; c++) {
                            if (c >= nc) continue;
                            const T xp = xb[p][c];
                            if (s_iszero(xp)) continue;
                            acc[c] = s_add(acc[c], s_mul(s_neg(xp), av[u]));
                        }
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < NC; c++)
                if (c < nc) X[i + (size_t)c * ldb] = acc[c];
        }
    }
    if (blockIdx.x != 0 || nextb <= 0) return;
    __syncthreads();
    if (wave < nc) {
        T* x = X + (size_t)wave * ldb;
        const bool mine = lane < nextb;
        T d[BS];
#pragma unroll
        for (int k = 0; k < BS; k++) d[k] = opa(next0 + lane, next0 + k, mine && k < nextb);
        T xi = load_if(x + next0 + lane, mine);
#pragma unroll
        for (int st = 0; st < BS; st++) {
            const int p = LOWER ? st : BS - 1 - st;
            if (p >= nextb) continue;
            if (!UNIT && lane == p) xi = s_div(xi, d[p]);
            const T xp = readlane_t(xi, p);
            if (s_iszero(xp)) continue;
            const bool ahead = LOWER ? (lane > p && lane < nextb) : (lane < p);
            if (ahead) xi = s_add(xi, s_mul(s_neg(xp), d[p]));
        }
        if (mine) x[next0 + lane] = xi;
    }
}

// ---------------------------------------------------------------------------
// Few-right-hand-side solve op(A) X = B, op(A) lower (FORWARD) or upper, in
// one launch. The right-hand sides go in chains of up to FLOW_COLS columns
// (one per wavefront: more columns per chain lengthen every block's step,
// 16 took twice as long a block as 4);
// chain h runs on the workgroups blockIdx.x = 8 g + h, g < G, which
// round-robin dispatch places on one XCD, so its exchange words meet in one
// L2 (as the spread panel's do). Workgroup g takes the diagonal blocks g,
// g + G, ... in solve order. For a block it streams the tiles of op(A) in
// its rows, one per block solved before it, through LDS, each tile loaded
// before the x it multiplies is known, and applies a tile as soon as that x
// block is published: every row takes the blocks in solve order and each
// block column by column (ascending forward, descending backward), skipping
// zero entries of x, the operations trsv_lds_kernel (op(A) = A) and
// trsv_step_kernel (every form) give each element. Then each wavefront
// solves the diagonal block for its column by readlane substitution
// (trsv_lds_kernel's steps), publishes x as sequence-tagged words (the
// launch's own number, so a word left by an earlier launch is never taken
// for current) and stores x to B. A reader's lanes poll their own words
// until they carry the sequence number, so one round trip both sees a
// block arrive and delivers it (one lane polling, then every lane reading,
// took two: 4.65 -> 4.47 us a block). Only x crosses workgroups, and
// the triangle is read once, by G workgroups at a time instead of one per
// column: the chain's time is the diagonal solves and one L2 hop a block.
// The G <= FLOW_MAXG workgroups of a chain fit on one XCD together (two per
// CU: 66.5 KB of LDS and 256 VGPRs a lane at most for the real forms) and wait
// only on each other's earlier blocks, so every wait ends; each also gives
// up after SPREAD_SPIN polls or once the abort word (xbuf[0]) is raised, and
// a block solved after a wait gave up is published and stored as NaN.
constexpr int FLOW_COLS = 4;   // right-hand sides per chain, one per wavefront
constexpr int FLOW_MAXG = 64;  // workgroups per chain: two per CU of one XCD

// a / b, correctly rounded, from y = RN(1/b), by Markstein's correction
// (P. Markstein, IBM J. Res. Dev. 34(1), 1990, as in Muller et al.,
// Handbook of Floating-Point Arithmetic, the theorem on FMA division): q =
// RN(a y) lies within one ulp of a/b, r = a - b q is exact (one FMA), and
// RN(q + r y) is RN(a/b), for operands whose quotient and residual stay
// clear of underflow and overflow (both in [2^-400, 2^400] for double,
// [2^-50, 2^50] for float); other operands, zeros and non-finite values
// take the division itself. Three dependent operations instead of the
// division's sequence, with y computed once per divisor off the chain.
__device__ __forceinline__ double div_known(double a, double b, double y)
{
    const double aa = fabs(a), ab = fabs(b);
    if (aa >= 0x1p-400 && aa <= 0x1p400 && ab >= 0x1p-400 && ab <= 0x1p400) {
        const double q = a * y;
        return fma(fma(-q, b, a), y, q);
    }
    return a / b;
}
__device__ __forceinline__ float div_known(float a, float b, float y)
{
    const float aa = fabsf(a), ab = fabsf(b);
    if (aa >= 0x1p-50f && aa <= 0x1p50f && ab >= 0x1p-50f && ab <= 0x1p50f) {
        const float q = a * y;
        return fmaf(fmaf(-q, b, a), y, q);
    }
    return a / b;
}

// A divisor prepared once for many divisions giving s_div's results bit for
// bit: real, the divisor and its correctly rounded reciprocal; complex,
// Smith's scaling (which part is the larger, the ratio r and the scale den,
// as s_div computes them) and den's reciprocal, so each quotient part is
// one division by den, done as div_known does it.
template <typename T>
struct Divisor {
    T b, y;
};
template <>
struct Divisor<zd> {
    double r, den, y;
    int small_d;  // |Im| < |Re|
};
__device__ __forceinline__ Divisor<double> prep_div(double b) { return {b, 1.0 / b}; }
__device__ __forceinline__ Divisor<float> prep_div(float b) { return {b, 1.0f / b}; }
__device__ __forceinline__ Divisor<zd> prep_div(zd v)
{
    const double c = v.x, d = v.y;
    Divisor<zd> q;
    q.small_d = fabs(d) < fabs(c);
    if (q.small_d) {
        q.r = d / c;
        q.den = c + d * q.r;
    } else {
        q.r = c / d;
        q.den = d + c * q.r;
    }
    q.y = 1.0 / q.den;
    return q;
}
__device__ __forceinline__ double div_prep(double a, const Divisor<double>& v) { return div_known(a, v.b, v.y); }
__device__ __forceinline__ float div_prep(float a, const Divisor<float>& v) { return div_known(a, v.b, v.y); }
__device__ __forceinline__ zd div_prep(zd x, const Divisor<zd>& v)
{
    const double a = x.x, b = x.y;
    if (v.small_d) return make_zd(div_known(a + b * v.r, v.den, v.y), div_known(b - a * v.r, v.den, v.y));
    return make_zd(div_known(b + a * v.r, v.den, v.y), div_known(-a + b * v.r, v.den, v.y));
}
// Lane p's prepared divisor, as a uniform value.
template <typename T>
__device__ __forceinline__ Divisor<T> readlane_div(const Divisor<T>& v, int p)
{
    return {readlane_t(v.b, p), readlane_t(v.y, p)};
}
template <>
__device__ __forceinline__ Divisor<zd> readlane_div(const Divisor<zd>& v, int p)
{
    return {readlane_t(v.r, p), readlane_t(v.den, p), readlane_t(v.y, p), __builtin_amdgcn_readlane(v.small_d, p)};
}

template <typename T>
__device__ __forceinline__ T from_words(const unsigned (&w)[sizeof(T) / 4])
{
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

template <typename T, bool FORWARD, int TR, bool UNIT>
__global__ void __launch_bounds__(256)
trsv_flow_kernel(int n, const T* __restrict__ A, int lda, T* __restrict__ B, int ldb, int nrhs, int G,
                 unsigned long long* __restrict__ xbuf, int slot_words, unsigned seq)
{
    constexpr int BS = tri_block<T>();
    constexpr int WPE = (int)sizeof(T) / 4;  // exchange words per element
    constexpr int NCW = FLOW_COLS / 4;       // columns per wavefront
    constexpr int R = BS * BS / 256;         // tile elements per thread
    const int chain = blockIdx.x & 7, g = blockIdx.x >> 3;
    const int c0 = chain * FLOW_COLS;
    if (c0 >= nrhs || g >= G) return;
    const int ncc = nrhs - c0 < FLOW_COLS ? nrhs - c0 : FLOW_COLS;
    const int nblk = (n + BS - 1) / BS;
    unsigned long long* abort_word = xbuf;
    unsigned long long* slots = xbuf + 1 + (size_t)chain * nblk * slot_words;
    __shared__ T tile[2][BS][BS + 1];  // tile[p][i] = op(A)(row i of the block, column p of the source)
    // The wavefront index as a uniform value, so the per-column tests below
    // are scalar branches rather than lane masks.
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int cw0 = wave * NCW;  // this wavefront's first column in the chain
    const int ncw = ncc - cw0 <= 0 ? 0 : (ncc - cw0 < NCW ? ncc - cw0 : NCW);
    T* X = B + (size_t)(c0 + cw0) * ldb;
    // op(A)(i, k), read where the stored triangle keeps it.
    auto opa = [&](int i, int k, bool ok) __attribute__((always_inline)) {
        T v = load_if(TR == 0 ? A + i + (size_t)k * lda : A + k + (size_t)i * lda, ok);
        if (TR == 2) v = s_conj(v);
        return v;
    };
    // Tile elements by thread, consecutive threads along the stored column.
    auto tile_ip = [&](int r, int& i, int& p) __attribute__((always_inline)) {
        const int e = tid + 256 * r;
        i = TR == 0 ? e % BS : e / BS;
        p = TR == 0 ? e / BS : e % BS;
    };
    auto source = [&](int t, int& ks, int& kbs) __attribute__((always_inline)) {
        const int sb = FORWARD ? t : nblk - 1 - t;
        ks = sb * BS;
        kbs = n - ks < BS ? n - ks : BS;
    };
    // Loads are unconditional, from indexes clamped into the block, and the
    // entries past it are zeroed by selection: a guarded load compiles to a
    // branch that waits for its own data, which serialized a tile's loads
    // (about 10 us a tile).
    T reg[R];
    auto load_tile = [&](int t, int k0, int kb) __attribute__((always_inline)) {
        int ks, kbs;
        source(t, ks, kbs);
#pragma unroll
        for (int r = 0; r < R; r++) {
            int i, p;
            tile_ip(r, i, p);
            const int ic = i < kb ? i : kb - 1, pc = p < kbs ? p : kbs - 1;
            const T v = opa(k0 + ic, ks + pc, true);
            reg[r] = s_sel(i < kb && p < kbs, v, s_zero(T{}));
        }
    };
    // x of block sb for this wavefront's NC columns, lane l holding entry l
    // (zero past the block): every lane polls its own words until they carry
    // the launch's sequence number, so the poll that sees them current also
    // delivers them. False if a wait gave up.
    auto fetch_x = [&](auto nc_, int sb, int kbs, T (&xv)[NCW]) __attribute__((always_inline)) -> bool {
        constexpr int NC = decltype(nc_)::value;
        const unsigned long long* xs = slots + (size_t)sb * slot_words + (size_t)cw0 * BS * WPE;
        bool ok = true;
#pragma unroll
        for (int c = 0; c < NC; c++) {
            unsigned w[WPE];
            if (lane < kbs) ok = ok && xwait<WPE>(xs + ((size_t)c * BS + lane) * WPE, seq, abort_word, w);
            xv[c] = s_sel(lane < kbs && ok, from_words<T>(w), s_zero(T{}));
        }
        return __all(ok);
    };
    for (int bq = g; bq < nblk; bq += G) {
        const int blk = FORWARD ? bq : nblk - 1 - bq;
        const int k0 = blk * BS, kb = n - k0 < BS ? n - k0 : BS;
        const bool mine = lane < kb;
        // Every wavefront runs the same steps and barriers; NC (its column
        // count, 0 for a wavefront without columns) is a constant inside.
        auto block = [&](auto nc_) __attribute__((always_inline)) {
            constexpr int NC = decltype(nc_)::value;
            T acc[NCW];
            if constexpr (NC > 0) {
                const int lc = mine ? lane : kb - 1;
#pragma unroll
                for (int c = 0; c < NC; c++) acc[c] = s_sel(mine, X[k0 + lc + (size_t)c * ldb], s_zero(T{}));
            }
            bool bad = false;
            // The diagonal block (tile bq) to the buffer bq & 1, visible
            // after the barrier, and (non-unit) each lane's reciprocal of its
            // own diagonal entry, its divisor in the substitution: staged
            // before the last x is awaited.
            Divisor<T> dvs{};
            auto stage_diagonal = [&]() __attribute__((always_inline)) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    int i, p;
                    tile_ip(r, i, p);
                    tile[bq & 1][p][i] = reg[r];
                }
                __syncthreads();
                if (!UNIT && NC > 0) dvs = prep_div(tile[bq & 1][mine ? lane : 0][mine ? lane : 0]);
            };
            // Tiles 0 .. bq - 1 (the blocks solved before), then the
            // diagonal block, each loaded one step ahead.
            load_tile(0, k0, kb);
            if (bq == 0) stage_diagonal();
            for (int t = 0; t < bq; ++t) {
                int ks, kbs;
                source(t, ks, kbs);
#pragma unroll
                for (int r = 0; r < R; r++) {
                    int i, p;
                    tile_ip(r, i, p);
                    tile[t & 1][p][i] = reg[r];
                }
                __syncthreads();
                load_tile(t + 1, k0, kb);
                // Buffer bq & 1 last held tile bq - 2, done before the barrier.
                if (t + 1 == bq) stage_diagonal();
                if constexpr (NC > 0) {
                    T xv[NCW];
                    if (!fetch_x(nc_, ks / BS, kbs, xv)) {
                        bad = true;
                        if (lane == 0) __hip_atomic_store(abort_word, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
                        for (int c = 0; c < NC; c++) xv[c] = s_zero(T{});
                    }
                    // acc -= x_p a_p, p in order, a zero x_p leaving acc as it
                    // was. Without a zero x in the tile the skip cannot fire,
                    // so the whole-tile path then runs without the select.
                    bool zero = false;
#pragma unroll
                    for (int c = 0; c < NC; c++) zero = zero || (lane < kbs && s_iszero(xv[c]));
                    if (kbs == BS && !__any(zero)) {
                        constexpr int U = 8;
#pragma unroll
                        for (int u0 = 0; u0 < BS; u0 += U) {
                            T av[U];
#pragma unroll
                            for (int v = 0; v < U; v++) av[v] = tile[t & 1][FORWARD ? u0 + v : BS - 1 - (u0 + v)][lane];
#pragma unroll
                            for (int v = 0; v < U; v++) {
                                const int p = FORWARD ? u0 + v : BS - 1 - (u0 + v);
#pragma unroll
                                for (int c = 0; c < NC; c++)
                                    acc[c] = s_add(acc[c], s_mul(s_neg(readlane_t(xv[c], p)), av[v]));
                            }
                        }
                    } else {
                        for (int u = 0; u < kbs; ++u) {
                            const int p = FORWARD ? u : kbs - 1 - u;
                            const T a = tile[t & 1][p][lane];
#pragma unroll
                            for (int c = 0; c < NC; c++) {
                                const T xp = readlane_t(xv[c], p);
                                acc[c] = s_sel(s_iszero(xp), acc[c], s_add(acc[c], s_mul(s_neg(xp), a)));
                            }
                        }
                    }
                }
            }
            if constexpr (NC > 0) {
                // The diagonal block by substitution, a zero x_p skipping its
                // update (selected).
                // Entries of op(A) read from LDS eight steps at a time.
                constexpr int U = 8;
#pragma unroll
                for (int s0 = 0; s0 < BS; s0 += U) {
                    T dv[U];
#pragma unroll
                    for (int v = 0; v < U; v++) dv[v] = tile[bq & 1][FORWARD ? s0 + v : BS - 1 - (s0 + v)][lane];
#pragma unroll
                    for (int v = 0; v < U; v++) {
                        const int p = FORWARD ? s0 + v : BS - 1 - (s0 + v);
                        if (p >= kb) continue;
                        // x_p: lane p's entry, read out and (non-unit) divided
                        // as a value uniform over the wavefront, so the
                        // division's range test is a scalar branch; the
                        // divisor and its reciprocal come from lane p too,
                        // independent of the chain.
                        Divisor<T> dp{};
                        if (!UNIT) dp = readlane_div(dvs, p);
#pragma unroll
                        for (int c = 0; c < NC; c++) {
                            T xp = readlane_t(acc[c], p);
                            if (!UNIT) {
                                xp = div_prep(xp, dp);
                                acc[c] = s_sel(lane == p, xp, acc[c]);
                            }
                            const bool ahead = FORWARD ? (lane > p && lane < kb) : (lane < p);
                            acc[c] = s_sel(ahead && !s_iszero(xp), s_add(acc[c], s_mul(s_neg(xp), dv[v])), acc[c]);
                        }
                    }
                }
                if (bad) {
#pragma unroll
                    for (int c = 0; c < NC; c++) acc[c] = s_nan(T{});
                }
                if (bq + 1 < nblk) {
                    // Blocks solved later read it: publish before the store to B.
                    unsigned long long* xo = slots + (size_t)blk * slot_words + (size_t)cw0 * BS * WPE;
#pragma unroll
                    for (int c = 0; c < NC; c++) {
                        if (mine) {
                            unsigned w[WPE];
                            __builtin_memcpy(w, &acc[c], sizeof(T));
#pragma unroll
                            for (int q = 0; q < WPE; q++) xstore(xo + ((size_t)c * BS + lane) * WPE + q, seq, w[q]);
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < NC; c++)
                    if (mine) X[k0 + lane + (size_t)c * ldb] = acc[c];
            }
            __syncthreads();  // both tile buffers free for the next block
        };
        dispatch<0, NCW + 1>(ncw, block);
    }
}

// trsv_lds_kernel for G right-hand sides at once, over columns already in
// LDS (xs[c * n + i], c < nr <= G): every load of op(A) serves all of them.
// Each column receives exactly trsv_lds_kernel's operations, in its order:
// the left-looking dot products keep its lane split and reduction, wave c
// solves column c's diagonal block, and the right-looking update keeps its
// per-column skip of zero entries.
template <typename T, bool LOWER, int TR, bool UNIT, int G, int NT, typename PA>
__device__ void trsv_lds_group(int n, PA __restrict__ A, int lda, T* xs, int nr)
{
    constexpr int BS = tri_block<T>();
    static_assert(NT % 64 == 0 && NT / 64 >= G, "a wavefront per right-hand side");
    static_assert(G >= 1 && G <= 4, "one wavefront per column for the diagonal blocks");
    __shared__ T part[G][BS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    auto at = [&](int i, int k) -> T {
        T v = TR ? A[k + (size_t)i * lda] : A[i + (size_t)k * lda];
        return TR == 2 ? s_conj(v) : v;
    };
    constexpr bool FORWARD = (LOWER && TR == 0) || (!LOWER && TR != 0);
    const T zero = s_zero(T{});
    const int nblk = (n + BS - 1) / BS;
    for (int bq = 0; bq < nblk; bq++) {
        const int blk = FORWARD ? bq : nblk - 1 - bq;
        const int k0 = blk * BS;
        const int kb = n - k0 < BS ? n - k0 : BS;
        if (TR != 0) {
            const int lo = FORWARD ? 0 : k0 + kb;
            const int hi = FORWARD ? k0 : n;
            // Wavefront w takes the block's rows w, w + NW, ... together: one
            // pass over k with every row's loads in flight, then their
            // reductions side by side. Per row and column, s0 adds entries
            // lo + lane + 128 m and s1 those 64 further on, each in
            // increasing order (entries past hi are read at lo and not
            // added), then s0 + s1 and the xor tree over the lanes.
            constexpr int NW = NT / 64, RW = (BS + NW - 1) / NW;
            T s0[RW][G], s1[RW][G];
#pragma unroll
            for (int r = 0; r < RW; r++)
#pragma unroll
                for (int c = 0; c < G; c++) s0[r][c] = s1[r][c] = zero;
            const int iters = (hi - lo + 127) / 128;
            int k = lo + lane;
            // One pass at a time (every row's loads of the pass in flight
            // together); unrolling a second pass spills the accumulators.
#pragma unroll 1
            for (int m = 0; m < iters; m++, k += 128) {
                const bool v0 = k < hi, v1 = k + 64 < hi;
                const int ka = v0 ? k : lo, kb1 = v1 ? k + 64 : lo;
                T a0[RW], a1[RW];
#pragma unroll
                for (int r = 0; r < RW; r++) {
                    const int i = wave + NW * r;
                    const int ir = i < kb ? i : 0;
                    a0[r] = at(k0 + ir, ka);
                    a1[r] = at(k0 + ir, kb1);
                }
#pragma unroll
                for (int c = 0; c < G; c++) {
                    if (c >= nr) continue;
                    const T x0 = xs[c * n + ka], x1 = xs[c * n + kb1];
#pragma unroll
                    for (int r = 0; r < RW; r++) {
                        s0[r][c] = s_sel(v0, s_add(s0[r][c], s_mul(a0[r], x0)), s0[r][c]);
                        s1[r][c] = s_sel(v1, s_add(s1[r][c], s_mul(a1[r], x1)), s1[r][c]);
                    }
                }
            }
            // Each row and column: s0 + s1, then the xor tree over the lanes
            // (offsets 32, 16, ..., 1, as shfl_xor_t's loop takes them).
#pragma unroll
            for (int r = 0; r < RW; r++) {
                const int i = wave + NW * r;
#pragma unroll
                for (int c = 0; c < G; c++) {
                    T sum = s_add(s0[r][c], s1[r][c]);
                    sum = s_add(sum, xor_lane<32>(sum));
                    sum = s_add(sum, xor_lane<16>(sum));
                    sum = s_add(sum, xor_lane<8>(sum));
                    sum = s_add(sum, xor_lane<4>(sum));
                    sum = s_add(sum, xor_lane<2>(sum));
                    sum = s_add(sum, xor_lane<1>(sum));
                    if (lane == 0 && i < kb && c < nr) part[c][i] = sum;
                }
            }
            __syncthreads();
        }
        if (wave < nr) {
            const int c = wave;
            const bool mine = lane < kb;
            T d[BS];
#pragma unroll
            for (int k = 0; k < BS; k++) d[k] = (mine && k < kb) ? at(k0 + lane, k0 + k) : zero;
            T xi = load_if(xs + c * n + k0 + lane, mine);
            if (TR != 0 && mine) xi = s_sub(xi, part[c][lane]);
#pragma unroll
            for (int s = 0; s < BS; s++) {
                const int p = FORWARD ? s : BS - 1 - s;
                if (p >= kb) continue;
                if (!UNIT && lane == p) xi = s_div(xi, d[p]);
                const T xp = readlane_t(xi, p);
                if (s_iszero(xp)) continue;
                const bool ahead = FORWARD ? (lane > p && lane < kb) : (lane < p);
                if (ahead) xi = s_add(xi, s_mul(s_neg(xp), d[p]));
            }
            if (mine) xs[c * n + k0 + lane] = xi;
        }
        __syncthreads();
        if (TR == 0) {
            const int r0 = FORWARD ? k0 + kb : 0;
            const int r1 = FORWARD ? n : k0;
            for (int i = r0 + tid; i < r1; i += 2 * NT) {
                const int i2 = i + NT;
                const bool two = i2 < r1;
                T acc[G], acc2[G];
#pragma unroll
                for (int c = 0; c < G; c++) {
                    acc[c] = load_if(xs + c * n + i, c < nr);
                    acc2[c] = load_if(xs + c * n + i2, c < nr && two);
                }
                constexpr int U = 8;
                for (int q = 0; q < kb; q += U) {
                    T av[U], av2[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int p = FORWARD ? q + u : kb - 1 - (q + u);
                        const bool in = q + u < kb;
                        av[u] = load_if(A + i + (size_t)(k0 + p) * lda, in);
                        av2[u] = load_if(A + i2 + (size_t)(k0 + p) * lda, in && two);
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        if (q + u >= kb) break;
                        const int p = FORWARD ? q + u : kb - 1 - (q + u);
#pragma unroll
                        for (int c = 0; c < G; c++) {
                            if (c >= nr) continue;
                            const T xp = xs[c * n + k0 + p];
                            if (s_iszero(xp)) continue;
                            const T nx = s_neg(xp);
                            acc[c] = s_add(acc[c], s_mul(nx, av[u]));
                            if (two) acc2[c] = s_add(acc2[c], s_mul(nx, av2[u]));
                        }
                    }
                }
#pragma unroll
                for (int c = 0; c < G; c++) {
                    if (c >= nr) continue;
                    xs[c * n + i] = acc[c];
                    if (two) xs[c * n + i2] = acc2[c];
                }
            }
            __syncthreads();
        }
    }
}

// LU solves of many systems, one workgroup per system, G right-hand sides
// at a time in LDS (Lu.cs:336-378, BlockedLu.cs:126-141): P x gathered on
// the way in, L forward (unit), U backward; or, for the adjoint, U^H
// forward, L^H backward (unit) and the permutation scattered on the way
// out. The same operations per column as getrs_impl's.
constexpr int GS_NT = 512;  // threads of a grouped solve's workgroup

template <typename T, int TRANS, int G>
__global__ void __launch_bounds__(GS_NT) getrs_grouped_kernel(const SolveRec<T>* __restrict__ recs)
{
    extern __shared__ unsigned char getrs_smem[];
    T* xs = reinterpret_cast<T*>(getrs_smem);
    const SolveRec<T> r = recs[blockIdx.x];
    const auto lu = as_global(r.lu);
    const auto perm = as_global(r.perm);
    const auto x = as_global(r.x);
    const auto y = as_global(r.y);
    const int tid = threadIdx.x, n = r.n;
    if (tid == 0 && r.info != nullptr && *as_global(r.info) == INT_MAX) *as_global(r.info) = -1;
    constexpr int TR = sizeof(T) == 16 ? 2 : 1;
    for (int c0 = 0; c0 < r.nrhs; c0 += G) {
        const int nr = r.nrhs - c0 < G ? r.nrhs - c0 : G;
        for (int e = tid; e < nr * n; e += GS_NT) {
            const int c = e / n, i = e % n;
            const int src = TRANS ? i : perm[i];
            xs[c * n + i] = x[src + (size_t)(c0 + c) * r.ldx];
        }
        __syncthreads();
        if (!TRANS) {
            trsv_lds_group<T, true, 0, true, G, GS_NT>(n, lu, r.ldlu, xs, nr);
            trsv_lds_group<T, false, 0, false, G, GS_NT>(n, lu, r.ldlu, xs, nr);
        } else {
            trsv_lds_group<T, false, TR, false, G, GS_NT>(n, lu, r.ldlu, xs, nr);
            trsv_lds_group<T, true, TR, true, G, GS_NT>(n, lu, r.ldlu, xs, nr);
        }
        for (int e = tid; e < nr * n; e += GS_NT) {
            const int c = e / n, i = e % n;
            const int dst = TRANS ? perm[i] : i;
            y[dst + (size_t)(c0 + c) * r.ldy] = xs[c * n + i];
        }
        __syncthreads();
    }
}

__device__ __forceinline__ double lu_nan_max(double a, double b)
{
    if (isnan(a) || isnan(b)) return NAN;
    return a > b ? a : b;
}

// The start of many LUs: ||A||_1 (each column summed as the norm kernels
// sum it: lanes strided by 64, then a butterfly), perm = identity, info =
// the no-zero-pivot sentinel. One workgroup per matrix.
template <typename T>
__global__ void __launch_bounds__(1024) lu_init_grouped_kernel(const InitRec<T>* __restrict__ recs)
{
    constexpr int W = 16;  // wavefronts
    __shared__ double red[W];
    const InitRec<T> rec = recs[blockIdx.x];
    struct {
        decltype(as_global(rec.a)) a;
        int lda, n;
        decltype(as_global(rec.perm)) perm;
        decltype(as_global(rec.anorm)) anorm;
        decltype(as_global(rec.info)) info;
    } r{as_global(rec.a), rec.lda, rec.n, as_global(rec.perm), as_global(rec.anorm), as_global(rec.info)};
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < r.n; i += 64 * W) r.perm[i] = i;
    double best = 0.0;
    for (int j = wave; j < r.n; j += W) {
        const auto col = r.a + (size_t)j * r.lda;
        double s = 0.0;
        // Four entries loaded together, summed in row order.
        int i = lane;
        for (; i + 192 < r.n; i += 256) {
            const T x0 = col[i], x1 = col[i + 64], x2 = col[i + 128], x3 = col[i + 192];
            s += s_abs(x0);
            s += s_abs(x1);
            s += s_abs(x2);
            s += s_abs(x3);
        }
        for (; i < r.n; i += 64) s += s_abs(col[i]);
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
        best = lu_nan_max(best, s);
    }
    if (lane == 0) red[wave] = best;
    __syncthreads();
    if (tid == 0) {
        double v = red[0];
        for (int q = 1; q < W; q++) v = lu_nan_max(v, red[q]);
        *r.anorm = v;
        *r.info = INT_MAX;
    }
}

// Inverses of the W x W diagonal blocks of LU factors (the last block may
// be smaller; it is bordered by the identity): workgroup b inverts block
// b0 + b, writing L_kk^-1 (unit lower) and then U_kk^-1 (upper), each W x W
// column-major, at inv + (b0 + b) * 2 W^2. Thread j < W forms column j of
// L^-1, thread W + j column j of U^-1, by substitution against the block
// in LDS (every thread reads the same entry at once).
template <typename T, int W>
__global__ void __launch_bounds__(2 * W) tri_inverse_kernel(int n, const T* __restrict__ lu, int ld, int b0,
                                                            T* __restrict__ inv)
{
    __shared__ T As[W][W + 1];
    const int blk = b0 + blockIdx.x, k0 = blk * W;
    const int kb = W < n - k0 ? W : n - k0;
    const T* a = lu + k0 + (size_t)k0 * ld;
    for (int e = threadIdx.x; e < W * W; e += 2 * W) {
        const int i = e % W, k = e / W;
        const bool in = i < kb && k < kb;
        As[i][k] = in ? a[i + (size_t)k * ld] : (i == k ? s_one(T{}) : s_zero(T{}));
    }
    __syncthreads();
    const bool upper = threadIdx.x >= W;
    const int j = upper ? threadIdx.x - W : threadIdx.x;
    T y[W];
#pragma unroll
    for (int i = 0; i < W; i++) y[i] = i == j ? s_one(T{}) : s_zero(T{});
    if (!upper) {
#pragma unroll
        for (int p = 0; p < W; p++) {
            const T ny = s_neg(y[p]);
#pragma unroll
            for (int i = p + 1; i < W; i++) y[i] = s_add(y[i], s_mul(ny, As[i][p]));
        }
    } else {
#pragma unroll
        for (int p = W - 1; p >= 0; p--) {
            y[p] = s_div(y[p], As[p][p]);
            const T ny = s_neg(y[p]);
#pragma unroll
            for (int i = 0; i < p; i++) y[i] = s_add(y[i], s_mul(ny, As[i][p]));
        }
    }
    T* out = inv + (size_t)blk * 2 * W * W + (upper ? W * W : 0) + (size_t)j * W;
#pragma unroll
    for (int i = 0; i < W; i++) out[i] = y[i];
}

// The same two inverses per diagonal block as tri_inverse_kernel, one
// workgroup per block (blockIdx.x), the columns of the inverses kept in
// LDS instead of registers: a W-element complex column in registers, with
// the unrolled substitution, spills. Same operations per element, in the
// same order.
template <typename T, int W>
__global__ void __launch_bounds__(2 * W) tri_inverse_lds_kernel(int n, const T* __restrict__ lu, int ld,
                                                                T* __restrict__ inv)
{
    __shared__ T As[W][W + 1];
    __shared__ T Y[2][W][W + 1];  // Y[0]: inv(L) (unit lower), Y[1]: inv(U); Y[h][i][j] = row i, column j
    const int blk = blockIdx.x, k0 = blk * W;
    const int kb = W < n - k0 ? W : n - k0;
    const T* a = lu + k0 + (size_t)k0 * ld;
    for (int e = threadIdx.x; e < W * W; e += 2 * W) {
        const int i = e % W, k = e / W;
        const bool in = i < kb && k < kb;
        As[i][k] = in ? a[i + (size_t)k * ld] : (i == k ? s_one(T{}) : s_zero(T{}));
    }
    const int h = threadIdx.x >= W ? 1 : 0;
    const int j = threadIdx.x - h * W;
    for (int i = 0; i < W; i++) Y[h][i][j] = i == j ? s_one(T{}) : s_zero(T{});
    __syncthreads();
    if (h == 0) {
        for (int p = 0; p < W; p++) {
            const T ny = s_neg(Y[0][p][j]);
            for (int i = p + 1; i < W; i++) Y[0][i][j] = s_add(Y[0][i][j], s_mul(ny, As[i][p]));
        }
    } else {
        for (int p = W - 1; p >= 0; p--) {
            const T yp = s_div(Y[1][p][j], As[p][p]);
            Y[1][p][j] = yp;
            const T ny = s_neg(yp);
            for (int i = 0; i < p; i++) Y[1][i][j] = s_add(Y[1][i][j], s_mul(ny, As[i][p]));
        }
    }
    T* out = inv + (size_t)blk * 2 * W * W + (h ? W * W : 0) + (size_t)j * W;
    for (int i = 0; i < W; i++) out[i] = Y[h][i][j];
}

// dst[i] = src[perm[i]] (gather) or dst[perm[i]] = src[i] (scatter).
template <typename T>
__global__ void permute_rows_kernel(int n, int ncols, const int* __restrict__ perm, const T* __restrict__ src, int lds,
                                    T* __restrict__ dst, int ldd, int scatter)
{
    size_t total = (size_t)n * ncols;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        int j = (int)(e / n), i = (int)(e % n);
        if (scatter) dst[perm[i] + (size_t)j * ldd] = src[i + (size_t)j * lds];
        else dst[i + (size_t)j * ldd] = src[perm[i] + (size_t)j * lds];
    }
}

template <typename T>
__global__ void __launch_bounds__(256) pivot_range_kernel(int n, const T* __restrict__ a, int lda, double* __restrict__ out)
{
    __shared__ double smin[4], smax[4];
    __shared__ int snan[4];
    double lo = INFINITY, hi = 0.0;
    int nan_seen = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        double v = s_abs(a[i + (size_t)i * lda]);
        if (isnan(v)) nan_seen = 1;
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
        nan_seen |= __shfl_xor(nan_seen, off, 64);
    }
    int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smin[wave] = lo;
        smax[wave] = hi;
        snan[wave] = nan_seen;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
            lo = fmin(lo, smin[q]);
            hi = fmax(hi, smax[q]);
            nan_seen |= snan[q];
        }
        // Math.Min/Math.Max propagate a NaN; fmin/fmax drop it.
        if (nan_seen) lo = hi = NAN;
        if (n == 0) lo = 0.0;
        out[0] = lo;
        out[1] = hi;
    }
}

// ---------------------------------------------------------------------------
// Host-side drivers.
// ---------------------------------------------------------------------------
// The LU phase split (LuPhaseTimings, Lu.cs:92-164): an event is recorded on
// the stream after each phase, and the interval it closes is charged to that
// phase's bucket; `total` is measured from the first event to the last, so
// time no phase accounts for stays visible as the difference.
struct PhaseClock : EventClock {
    enum { NONE = -1, PANEL = 0, SWAPS = 1, TRSM = 2, GEMM = 3 };
    using EventClock::EventClock;
    // Adds this factorization's split to *out (pooling, as LuPhaseTimings.Add).
    void read(LuPhases* out)
    {
        double ms[4] = {0, 0, 0, 0};
        EventClock::read(ms, 4, &out->total_ms);
        out->panel_ms += ms[0];
        out->swaps_ms += ms[1];
        out->trsm_ms += ms[2];
        out->gemm_ms += ms[3];
    }
};

template <typename T>
struct Ops;
template <>
struct Ops<double> {
    static void gemm(Context* c, int ta, int m, int n, int k, double alpha, const double* a, int lda, const double* b,
                     int ldb, double beta, double* cc, int ldc)
    {
        dgemm_ex(c, ta, m, n, k, alpha, a, lda, b, ldb, beta, cc, ldc);
    }
    static void batch(Context* c, int m, int n, int k, double alpha, const double* a, int lda, long long sa,
                      const double* b, int ldb, long long sb, double beta, double* cc, int ldc, long long sc, int count,
                      int kmode)
    {
        dgemm_batch(c, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, cc, ldc, sc, count, kmode);
    }
    static double one() { return 1.0; }
    static double neg_one() { return -1.0; }
};
template <>
struct Ops<float> {
    static void gemm(Context* c, int ta, int m, int n, int k, float alpha, const float* a, int lda, const float* b,
                     int ldb, float beta, float* cc, int ldc)
    {
        if (ta != 0) fail(TNS_EINVAL, "transposed single-precision products are not provided");
        sgemm_lp(c, c->lp_mode, m, n, k, alpha, a, lda, b, ldb, beta, cc, ldc);
    }
    static float one() { return 1.0f; }
    static float neg_one() { return -1.0f; }
};
template <>
struct Ops<zd> {
    static void gemm(Context* c, int ta, int m, int n, int k, zd alpha, const zd* a, int lda,
                     const zd* b, int ldb, zd beta, zd* cc, int ldc)
    {
        zgemm_ex(c, ta, m, n, k, alpha, a, lda, b, ldb, beta, cc, ldc);
    }
    static void batch(Context* c, int m, int n, int k, zd alpha, const zd* a, int lda, long long sa, const zd* b,
                      int ldb, long long sb, zd beta, zd* cc, int ldc, long long sc, int count, int kmode)
    {
        zgemm_batch(c, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, cc, ldc, sc, count, kmode);
    }
    static zd one() { return make_zd(1.0, 0.0); }
    static zd neg_one() { return make_zd(-1.0, 0.0); }
};

template <typename T, int NB, int R, int NT = 512>
void launch_getf2_reg(Context* ctx, int m, int w, T* a, int lda, int* piv, int off, int* perm, int* list, int* info)
{
    int need = (m + R - 1) / R;
    int nt = ((need + 63) / 64) * 64;
    if (nt < 64) nt = 64;
    hipLaunchKernelGGL((getf2_reg_kernel<T, NB, R, NT>), dim3(1), dim3(nt), 0, ctx->stream, m, w, a, lda, piv, off,
                       perm, list, 2 * w, info);
}

// Register-resident panel shapes: R rows per thread (512 threads), NB
// columns, 64 doubles' worth per thread (128 VGPRs), which leaves the
// compiler room for the rest without spilling.
template <typename T>
int reg_panel_width(int m)
{
    if constexpr (sizeof(T) != 16) {
        // The panel's cost per column grows with the values each thread
        // holds, so float keeps double's shapes (half the registers).
        if (m <= 512) return 64;
        if (m <= 1024) return 32;
        if (m <= 2048) return 16;
        if (m <= 4096) return 8;
        if (m <= 8192) return 4;
        return 0;
    } else {
        if (m <= 512) return 32;
        if (m <= 1024) return 16;
        if (m <= 2048) return 8;
        if (m <= 4096) return 4;
        return 0;
    }
}

// Factors m x w at a and leaves its move list, at most 2w moves, in `list`.
template <typename T>
void getf2_reg(Context* ctx, int m, int w, T* a, int lda, int* piv, int off, int* perm, int* list, int* info)
{
    if constexpr (sizeof(T) != 16) {
        if (m <= 512 && w <= 32) launch_getf2_reg<T, 32, 1>(ctx, m, w, a, lda, piv, off, perm, list, info);
        else if (m <= 512) launch_getf2_reg<T, 64, 1>(ctx, m, w, a, lda, piv, off, perm, list, info);
        else if (m <= 1024) launch_getf2_reg<T, 32, 2>(ctx, m, w, a, lda, piv, off, perm, list, info);
        else if (m <= 2048) launch_getf2_reg<T, 16, 4>(ctx, m, w, a, lda, piv, off, perm, list, info);
        else if (m <= 4096) launch_getf2_reg<T, 8, 8>(ctx, m, w, a, lda, piv, off, perm, list, info);
        else launch_getf2_reg<T, 4, 16>(ctx, m, w, a, lda, piv, off, perm, list, info);
    } else {
        if (m <= 512) launch_getf2_reg<T, 32, 1>(ctx, m, w, a, lda, piv, off, perm, list, info);
        else if (m <= 1024) launch_getf2_reg<T, 16, 2>(ctx, m, w, a, lda, piv, off, perm, list, info);
        else if (m <= 2048) launch_getf2_reg<T, 8, 4>(ctx, m, w, a, lda, piv, off, perm, list, info);
        else launch_getf2_reg<T, 4, 8>(ctx, m, w, a, lda, piv, off, perm, list, info);
    }
}

template <typename T, int QM>
void launch_apply_moves(Context* ctx, T* a, int lda, int c0, int c1, int ncols, const int* list, int count)
{
    constexpr int CPB = 256 / QM;
    hipLaunchKernelGGL((apply_moves_kernel<T, QM>), dim3((count + CPB - 1) / CPB), dim3(256), 0, ctx->stream, a, lda,
                       c0, c1, ncols, list);
}

template <typename T>
void apply_moves(Context* ctx, T* a, int lda, int c0, int c1, int ncols, const int* list, int max_moves)
{
    int count = c0 + (ncols - c1);
    if (count <= 0) return;
    if (max_moves <= 32) launch_apply_moves<T, 32>(ctx, a, lda, c0, c1, ncols, list, count);
    else if (max_moves <= 64) launch_apply_moves<T, 64>(ctx, a, lda, c0, c1, ncols, list, count);
    else if (max_moves <= 128) launch_apply_moves<T, 128>(ctx, a, lda, c0, c1, ncols, list, count);
    else if (max_moves <= 256) launch_apply_moves<T, 256>(ctx, a, lda, c0, c1, ncols, list, count);
    else {
        DevBuf<T> scratch(ctx, (size_t)count * max_moves);
        hipLaunchKernelGGL((apply_moves_slow_kernel<T>), dim3((count + 63) / 64), dim3(64), 0, ctx->stream, a, lda, c0,
                           c1, ncols, list, scratch.get(), max_moves);
    }
}

// Interchanges of rows [0, m) of a slice starting at global row `base`,
// recorded in rowid: move the listed rows in columns [0,c0) and [c1,ncols)
// of a (rows relative to the slice top). At most max_moves rows can have
// moved (two per interchange); the list buffer holds 1 + 2*max_moves ints.
template <typename T>
void sync_columns(Context* ctx, int m, int base, T* a, int lda, int c0, int c1, int ncols, int* rowid, int* list,
                  int max_moves)
{
    hipLaunchKernelGGL(compact_moves_kernel, dim3(1), dim3(1024), 0, ctx->stream, m, base, rowid, list, max_moves);
    apply_moves(ctx, a, lda, c0, c1, ncols, list, max_moves);
}

template <typename T, bool LOWER, int TR, bool UNIT>
void trsm_diag(Context* ctx, int w, int ncols, const T* A, int lda, T* B, int ldb)
{
    if constexpr (LOWER && TR == 0 && UNIT) {
        // The LU's U12 solve (and getrs' forward pass with many columns).
        const int waves = (ncols + 3) / 4;
        const int blocks = waves < 1024 ? waves : 1024;
        hipLaunchKernelGGL((trsm_llnu_wave_kernel<T, tri_block<T>()>), dim3(blocks), dim3(256), 0, ctx->stream, w,
                           ncols, A, lda, B, ldb);
        return;
    }
    constexpr int CB = 32;
    const int blocks = (ncols + CB - 1) / CB;
    hipLaunchKernelGGL((trsm_block_kernel<T, tri_block<T>(), CB, LOWER, TR, UNIT>), dim3(blocks), dim3(256), 0,
                       ctx->stream, w, ncols, A, lda, B, ldb);
}

// op(A) X = B in place, A n x n triangular.
// Few right-hand sides take trsv_flow_kernel (one launch) wherever it can
// run; where it cannot (graph capture, the LU's second stream, the
// no-allocation entry point), this is the order from which they take
// trsv_step_kernel (one launch per diagonal block, the triangle read by the
// whole chip) instead of one workgroup per column. Measured at 1024 (2 and 16 right-hand sides):
// the 16 dependent launches cost more than one workgroup streams (gecon
// 2.1 -> 2.8 ms); at 8192 the step form is 5x faster.
constexpr int TRSV_STEP_FROM = 2048;

// The exchange area of trsv_flow_kernel, at least `words` long: the context
// keeps one pool block between calls; a larger need returns it (once the
// stream that used it is idle) and takes a zeroed one.
inline unsigned long long* flow_area(Context* ctx, size_t words)
{
    const size_t bytes = words * sizeof(unsigned long long);
    if (bytes > ctx->flow_bytes) {
        if (ctx->flow) {
            TNS_HIP(hipStreamSynchronize(ctx->stream));
            ctx->pool.give(ctx->flow, ctx->flow_bytes);
            ctx->flow = nullptr;
            ctx->flow_bytes = 0;
        }
        const size_t want = bytes > ((size_t)1 << 20) ? bytes : ((size_t)1 << 20);
        ctx->flow = static_cast<unsigned long long*>(ctx->pool.take(want));
        ctx->flow_bytes = want;
        TNS_HIP(hipMemsetAsync(ctx->flow, 0, want, ctx->stream));
    }
    return ctx->flow;
}

// op(A) X = B by trsv_flow_kernel, when it applies: more than one diagonal
// block, at most 8 chains of right-hand sides, and a launch its sequence
// number cannot be replayed under (no graph capture), on the context's own
// stream (the exchange area is one per context), outside the no-allocation
// entry point. Returns false otherwise.
template <typename T, bool FORWARD, int TR, bool UNIT>
bool trsv_flow(Context* ctx, int n, int nrhs, const T* A, int lda, T* B, int ldb)
{
    constexpr int BS = tri_block<T>();
    if (ctx->capturing || ctx->no_alloc || ctx->stream == ctx->side || n <= BS || nrhs > 8 * FLOW_COLS) return false;
    const int nblk = (n + BS - 1) / BS;
    const int ncc = nrhs < FLOW_COLS ? nrhs : FLOW_COLS;
    const int slot_words = ncc * BS * (int)(sizeof(T) / 4);
    const int chains = (nrhs + FLOW_COLS - 1) / FLOW_COLS;
    unsigned long long* xbuf = flow_area(ctx, 1 + (size_t)chains * nblk * slot_words);
    if (++ctx->flow_seq == 0) ++ctx->flow_seq;  // zero is the area's initial tag
    // Workgroups of one chain: at most what one XCD holds at once (its CUs
    // times the kernel's workgroups per CU), so all of them can be resident.
    static const int per_xcd = [ctx] {
        int per_cu = 0, cus = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, trsv_flow_kernel<T, FORWARD, TR, UNIT>, 256, 0) !=
                hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess) {
            (void)hipGetLastError();
            return 1;
        }
        const int v = per_cu * (cus / 8);
        return v > 0 ? v : 1;
    }();
    int G = nblk < FLOW_MAXG ? nblk : FLOW_MAXG;
    if (G > per_xcd) G = per_xcd;
    hipLaunchKernelGGL((trsv_flow_kernel<T, FORWARD, TR, UNIT>), dim3(8 * G), dim3(256), 0, ctx->stream, n, A, lda, B,
                       ldb, nrhs, G, xbuf, slot_words, ctx->flow_seq);
    return true;
}

template <typename T, bool LOWER, int TR, bool UNIT>
void trsm_impl(Context* ctx, int n, int nrhs, const T* A, int lda, T* B, int ldb)
{
    if (n <= 0 || nrhs <= 0) return;
    constexpr bool FORWARD = (LOWER && TR == 0) || (!LOWER && TR != 0);
    if ((nrhs <= 32 || n <= 64) && trsv_flow<T, FORWARD, TR, UNIT>(ctx, n, nrhs, A, lda, B, ldb)) return;
    if ((nrhs <= 32 || n <= 64) && n >= TRSV_STEP_FROM) {
        // Spread over the chip, one launch per diagonal block (the same
        // operations per element as trsv_lds_kernel).
        constexpr int BS = tri_block<T>(), NC = 4;
        const int nblk = (n + BS - 1) / BS;
        const int gy = (nrhs + NC - 1) / NC;
        int k0 = 0, kb = 0;
        for (int bq = 0; bq <= nblk; bq++) {
            int next0 = 0, nextb = 0;
            if (bq < nblk) {
                const int blk = FORWARD ? bq : nblk - 1 - bq;
                next0 = blk * BS;
                nextb = n - next0 < BS ? n - next0 : BS;
            }
            const int rows = kb == 0 ? 0 : (FORWARD ? n - k0 - kb : k0);
            if (nextb == 0 && rows == 0) break;
            const int slabs = rows > 0 ? (rows + 255) / 256 : 1;
            hipLaunchKernelGGL((trsv_step_kernel<T, FORWARD, TR, UNIT, NC>), dim3(slabs, gy), dim3(256), 0, ctx->stream, n, A,
                               lda, B, ldb, nrhs, k0, kb, next0, nextb);
            k0 = next0;
            kb = nextb;
            if (nextb == 0) break;
        }
        return;
    }
    if (nrhs <= 32 || n <= 64) {
        const size_t xbytes = sizeof(T) * (size_t)n;

        if (xbytes + sizeof(T) * tri_block<T>() <= (size_t)ctx->max_lds && n >= 256) {
            constexpr int UL = sizeof(T) == 16 ? 8 : 16;
            hipLaunchKernelGGL((trsv_lds_kernel<T, LOWER, TR, UNIT, 512, UL>), dim3(nrhs), dim3(512), xbytes,
                               ctx->stream, n, A, lda, B, ldb);
        } else if (xbytes + sizeof(T) * tri_block<T>() <= (size_t)ctx->max_lds) {
            hipLaunchKernelGGL((trsv_lds_kernel<T, LOWER, TR, UNIT>), dim3(nrhs), dim3(256), xbytes, ctx->stream, n, A,
                               lda, B, ldb);
        } else {
            hipLaunchKernelGGL((trsv_kernel<T, LOWER, TR, UNIT>), dim3(nrhs), dim3(256), 0, ctx->stream, n, A, lda, B,
                               ldb);
        }
        return;
    }
    constexpr int BS = tri_block<T>();
    const int nblk = (n + BS - 1) / BS;
    const int ta = TR;
    for (int bq = 0; bq < nblk; bq++) {
        const int blk = FORWARD ? bq : nblk - 1 - bq;
        const int k0 = blk * BS;
        const int kb = n - k0 < BS ? n - k0 : BS;
        if (TR != 0) {
            // Left-looking: B_k -= op(A)[k, solved] X_solved.
            if (FORWARD && k0 > 0)
                Ops<T>::gemm(ctx, ta, kb, nrhs, k0, Ops<T>::neg_one(), A + (size_t)k0 * lda, lda, B, ldb, Ops<T>::one(),
                             B + k0, ldb);
            if (!FORWARD && k0 + kb < n)
                Ops<T>::gemm(ctx, ta, kb, nrhs, n - k0 - kb, Ops<T>::neg_one(), A + (k0 + kb) + (size_t)k0 * lda, lda,
                             B + k0 + kb, ldb, Ops<T>::one(), B + k0, ldb);
        }
        trsm_diag<T, LOWER, TR, UNIT>(ctx, kb, nrhs, A + k0 + (size_t)k0 * lda, lda, B + k0, ldb);
        if (TR == 0) {
            // Right-looking: B_rest -= A[rest, k] X_k.
            if (FORWARD && k0 + kb < n)
                Ops<T>::gemm(ctx, 0, n - k0 - kb, nrhs, kb, Ops<T>::neg_one(), A + (k0 + kb) + (size_t)k0 * lda, lda,
                             B + k0, ldb, Ops<T>::one(), B + k0 + kb, ldb);
            if (!FORWARD && k0 > 0)
                Ops<T>::gemm(ctx, 0, k0, nrhs, kb, Ops<T>::neg_one(), A + (size_t)k0 * lda, lda, B + k0, ldb,
                             Ops<T>::one(), B, ldb);
        }
    }
}

// U12 := L11^-1 A12 for a block step: one diagonal-block solve for
// w <= 64 real / 32 complex, blocked substitution beyond.
template <typename T>
void trsm_llnu(Context* ctx, int w, int ncols, const T* L, int ldl, T* B, int ldb)
{
    if (w <= 0 || ncols <= 0) return;
    if (w <= tri_block<T>()) trsm_diag<T, true, 0, true>(ctx, w, ncols, L, ldl, B, ldb);
    else trsm_impl<T, true, 0, true>(ctx, w, ncols, L, ldl, B, ldb);
}

template <typename T>
void permute_rows(Context* ctx, int n, int ncols, const int* perm, const T* src, int lds, T* dst, int ldd, int scatter)
{
    size_t total = (size_t)n * ncols;
    if (!total) return;
    int blocks = (int)((total + 255) / 256 < 8192 ? (total + 255) / 256 : 8192);
    hipLaunchKernelGGL((permute_rows_kernel<T>), dim3(blocks), dim3(256), 0, ctx->stream, n, ncols, perm, src, lds, dst,
                       ldd, scatter);
}

template <typename T>
void copy2d(Context* ctx, int m, int n, const T* src, int lds, T* dst, int ldd)
{
    if (m <= 0 || n <= 0) return;
    TNS_HIP(hipMemcpy2DAsync(dst, sizeof(T) * (size_t)ldd, src, sizeof(T) * (size_t)lds, sizeof(T) * (size_t)m, n,
                             hipMemcpyDeviceToDevice, ctx->stream));
}

// Factor the m x w panel at a (global top row/column `off`): min(m, w)
// pivot steps. The panel's rows span `left` matrix columns to its left and
// `right` to its right; every interchange reaches all of them (the
// reference applies a block step's interchanges either side of the panel,
// Lu.cs:270-278; applying each chunk's as it is made is the same row
// permutation). Pivots and payloads are indexed from the panel top. A panel
// wider than the register kernel holds is factored in column chunks,
// right-looking: factor a chunk, move its interchanges to every other
// column, solve and update the panel columns to its right.
template <typename T>
void panel(Context* ctx, int m, int w, T* a, int lda, int* piv, int off, int* perm, int* rc, int* info, int* list,
           int left, int right, int cap = 0)
{
    if (m <= 0 || w <= 0) return;
    const int jl = m < w ? m : w;
    const int ncols = left + w + right;
    T* rows = a - (size_t)left * lda;  // column 0 of the matrix, panel's top row
    int cw = reg_panel_width<T>(m);
    if (cap > 0 && cw > cap) cw = cap;  // narrower chunks: a shorter update per column
    if (cw == 0) {
        hipLaunchKernelGGL((getf2_global_kernel<T>), dim3(1), dim3(1024), 0, ctx->stream, m, w, a, lda, piv, off, perm,
                           rc, info);
        sync_columns<T>(ctx, m, off, rows, lda, left, left + w, ncols, rc, list, 2 * jl);
        return;
    }
    for (int c0 = 0; c0 < jl; c0 += cw) {
        const int cwid = cw < jl - c0 ? cw : jl - c0;
        T* sub = a + c0 + (size_t)c0 * lda;
        getf2_reg<T>(ctx, m - c0, cwid, sub, lda, piv + c0, off + c0, perm + c0, list, info);
        apply_moves<T>(ctx, rows + c0, lda, left + c0, left + c0 + cwid, ncols, list, 2 * cwid);
        if (c0 + cwid >= w) continue;
        trsm_llnu<T>(ctx, cwid, w - c0 - cwid, sub, lda, sub + (size_t)cwid * lda, lda);
        if (m > c0 + cwid)
            Ops<T>::gemm(ctx, 0, m - c0 - cwid, w - c0 - cwid, cwid, Ops<T>::neg_one(), sub + cwid, lda,
                         sub + (size_t)cwid * lda, lda, Ops<T>::one(), sub + cwid + (size_t)cwid * lda, lda);
    }
}

// Outer step of getrf_impl: a multiple of nb. Measured on MI355X (FP64 LU,
// nb = 64): at order 8192 four panels per outer step, the trailing products
// of depth 256 instead of 64, 56.2 -> 53.5 ms (the products 16.9 -> 10.3 ms,
// the block-row solves 4.9 -> 9.1 ms); at 4096 slower (20.0 -> 22.1 ms), so
// below 8192 one.
template <typename T>
int lu_outer_block(int m, int n, int nb)
{
    const int limit = m < n ? m : n;
    if (sizeof(T) >= 8 && limit >= LU_OUTER_FROM && nb <= 64) return 4 * nb;
    return nb;
}

// Widest panel the spread kernel takes for T.
template <typename T>
constexpr int spread_nb()
{
    return sizeof(T) == 16 ? 32 : 64;
}

// Factors the m x w panel at a with getf2_spread_kernel and moves its
// interchanges to the `left` columns before it and `right` after it.
template <typename T>
void panel_spread(Context* ctx, int m, int w, T* a, int lda, int* piv, int off, int* perm, int* info, int* list,
                  unsigned long long* xbuf, int left, int right, PhaseClock* clk)
{
    constexpr int NB = spread_nb<T>();
    if constexpr (NB > 0) {
        const int P = (m + SPREAD_NT - 1) / SPREAD_NT;
        hipLaunchKernelGGL((getf2_spread_kernel<T, NB>), dim3(P > 1 ? 8 * P : 1), dim3(SPREAD_NT), 0, ctx->stream, m,
                           w, a, lda, piv, off, perm, list, P, xbuf, info);
        if (clk) clk->mark(PhaseClock::PANEL);
        const int jl = m < w ? m : w;
        apply_moves<T>(ctx, a - (size_t)left * lda, lda, left, left + w, left + w + right, list, 2 * jl);
        if (clk) clk->mark(PhaseClock::SWAPS);
    }
}

// getrf_impl's blocked loop with look-ahead (spread panels, panel width ib
// inside outer steps of ob columns). Step k:
//   on ctx->stream: factor panel k; once the side stream's step k - 1 is
//     done, move its interchanges into the outer step's columns left of it
//     and into the next panel's columns; give the next panel's columns
//     step k's update (the U12 solve and the trailing product of depth ib;
//     at the outer step's last panel, the outer step's solve and product of
//     depth ob) - the next panel can start;
//   on ctx->side, after those moves: move the interchanges into every other
//     column (earlier outer steps, and past the next panel) and give the
//     columns past the next panel their update.
// Every column takes the same updates in the same order as in the loop
// without look-ahead, and no update reads rows that another stream is
// moving: the interchanges of step k reach the outer step's L columns only
// after the side stream's step k - 1 (which reads them) is done, and the
// side stream's outer update starts after them. So the pivots are the same
// and every element sees the same arithmetic. The trailing updates of step
// k run while panel k + 1 is factored.
template <typename T>
void getrf_lookahead(Context* ctx, int m, int n, T* a, int lda, int ib, int ob, int* d_piv, int* d_perm,
                     int* d_info, int* lists, int lstride, unsigned long long* xbuf)
{
    constexpr int NB = spread_nb<T>();
    if constexpr (NB > 0) {
        const int limit = m < n ? m : n;
        hipStream_t side = ctx->side;
        int step = 0;
        // A block row left unsolved by next_panel_update_kernel: solved in
        // place on the next step's side stream (after the event that follows
        // the kernel), or at the end.
        struct Pending {
            int w = 0, nc = 0;
            const T* l = nullptr;
            T* b = nullptr;
        } pending;
        auto flush_pending = [&]() {
            if (pending.nc > 0) trsm_llnu<T>(ctx, pending.w, pending.nc, pending.l, lda, pending.b, lda);
            pending = Pending{};
        };
        for (int jo = 0; jo < limit; jo += ob) {
            const int wo = ob < limit - jo ? ob : limit - jo;
            const int end = jo + wo < n ? jo + wo : n;  // columns the inner steps update
            T* po = a + jo + (size_t)jo * lda;
            for (int jb = jo; jb < jo + wo; jb += ib, ++step) {
                const int w = ib < jo + wo - jb ? ib : jo + wo - jb;
                T* pa = a + jb + (size_t)jb * lda;
                int* list = lists + (size_t)step * lstride;
                const int mp = m - jb;
                const int P = (mp + SPREAD_NT - 1) / SPREAD_NT;
                hipLaunchKernelGGL((getf2_spread_kernel<T, NB>), dim3(P > 1 ? 8 * P : 1), dim3(SPREAD_NT), 0,
                                   ctx->stream, mp, w, pa, lda, d_piv + jb, jb, d_perm + jb, list, P, xbuf, d_info);
                const int moves = 2 * (mp < w ? mp : w);
                const bool last = jb + w >= jo + wo;  // the outer step's last panel
                // The next panel's columns [c1, c2).
                const int c1 = jb + w;
                int c2 = c1;
                if (c1 < limit) {
                    const int room = last ? limit - c1 : jo + wo - c1;
                    c2 = c1 + (ib < room ? ib : room);
                }
                if (step > 0) TNS_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_rest[(step - 1) & 1], 0));
                // Columns [jo, jb) and [c1, c2), rows from jb.
                apply_moves<T>(ctx, a + jb + (size_t)jo * lda, lda, jb - jo, c1 - jo, c2 - jo, list, moves);
                TNS_HIP(hipEventRecord(ctx->ev_panel, ctx->stream));
                // Step k's update of the columns [lo, hi): its own solve and
                // product inside the outer step, the outer step's past it.
                auto update = [&](int lo, int hi) {
                    if (hi <= lo) return;
                    // The step's own block width (the panel's, or at an outer
                    // step's last panel the outer step's) and top row.
                    const int bw = last ? wo : w, top = last ? jo : jb;
                    if constexpr (sizeof(T) == 16) {
                        // Complex: the next panel's columns in one launch; their
                        // block row's in-place solve follows on the side stream.
                        // (Measured: the Expm step's LU 3.69 -> 3.58 ms. The real
                        // LU keeps the two launches: its 64-column form needs
                        // 1024-thread, 66 KB workgroups, which wait for CUs while
                        // the side stream's product fills them: n = 8192 41.5 ->
                        // 43.1 ms, n = 1024 2.76 -> 2.99 ms.)
                        if (bw <= NB && hi - lo <= NB && ctx->stream != side && m > top + bw) {
                            const int m2 = m - top - bw;
                            hipLaunchKernelGGL((next_panel_update_kernel<T, NB>), dim3((m2 + 63) / 64), dim3(1024),
                                               0, ctx->stream, m2, bw, hi - lo, last ? po : pa, lda,
                                               a + top + (size_t)lo * lda);
                            pending = Pending{bw, hi - lo, last ? po : pa, a + top + (size_t)lo * lda};
                            return;
                        }
                    }
                    if (!last) {
                        trsm_llnu<T>(ctx, w, hi - lo, pa, lda, a + jb + (size_t)lo * lda, lda);
                        if (jb + w < m)
                            Ops<T>::gemm(ctx, 0, m - jb - w, hi - lo, w, Ops<T>::neg_one(), pa + w, lda,
                                         a + jb + (size_t)lo * lda, lda, Ops<T>::one(), a + jb + w + (size_t)lo * lda,
                                         lda);
                    } else {
                        trsm_llnu<T>(ctx, wo, hi - lo, po, lda, a + jo + (size_t)lo * lda, lda);
                        if (jo + wo < m)
                            Ops<T>::gemm(ctx, 0, m - jo - wo, hi - lo, wo, Ops<T>::neg_one(), po + wo, lda,
                                         a + jo + (size_t)lo * lda, lda, Ops<T>::one(),
                                         a + jo + wo + (size_t)lo * lda, lda);
                    }
                };
                {
                    // The rest, on the side stream.
                    OnStream on(ctx, side);
                    TNS_HIP(hipStreamWaitEvent(side, ctx->ev_panel, 0));
                    flush_pending();
                    apply_moves<T>(ctx, a + jb, lda, jo, c2, n, list, moves);
                    update(c2, last ? n : end);
                    TNS_HIP(hipEventRecord(ctx->ev_rest[step & 1], side));
                }
                update(c1, c2);  // the next panel's columns, on the main stream
            }
        }
        if (step > 0) TNS_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_rest[(step - 1) & 1], 0));
        flush_pending();
    }
}

// work: scratch carved by the caller (lu_work_ints(m, n, nb) ints, 8-byte
// aligned), or null to take it from the pool.
template <typename T>
void getrf_impl(Context* ctx, int m, int n, T* a, int lda, int nb, int* d_piv, int* d_perm, int* d_info,
                LuPhases* phases = nullptr, int* work = nullptr)
{
    PhaseClock clock(ctx), *clk = phases ? &clock : nullptr;
    if (clk) clk->mark(PhaseClock::NONE);
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, ctx->stream, d_info, INT_MAX);
    const int limit = m < n ? m : n;
    bool spread_used = false;
    unsigned long long* xbuf = nullptr;
    DevBuf<int> pooled;
    if (limit > 0) {
        if (nb < 1) nb = limit < 1024 ? 32 : 64;
        const int maxw = nb < limit ? nb : n;
        const bool whole = nb >= limit;  // small problems are all panel (Lu.cs:244-249)
        constexpr int SNB = spread_nb<T>();
        // Panels of the spread kernel's width inside outer steps of the block
        // size when the block is wider than the kernel takes (complex at
        // nb = 64); the pivots are the same whatever the panel width.
        const int ib = (SNB > 0 && !whole && nb > SNB && m <= SPREAD_NT * SPREAD_MAXP) ? SNB : nb;
        const int steps = whole ? 1 : (limit + ib - 1) / ib;
        const bool spread = SNB > 0 && (whole ? maxw : ib) <= SNB && m <= SPREAD_NT * SPREAD_MAXP;
        spread_used = spread;
        const size_t need = lu_work_ints(m, n, nb);
        if (!work) {
            pooled = DevBuf<int>(ctx, need);
            work = pooled.get();
        }
        // Layout: the exchange words first (8-byte aligned), then the
        // per-panel move lists, the chunked panels' list, the row ids.
        xbuf = reinterpret_cast<unsigned long long*>(work);
        int* lists = work + 2 * SPREAD_XWORDS;
        const int lstride = SPREAD_LIST;
        int* list = lists + (size_t)steps * lstride;
        int* rc = list + 4 * (size_t)maxw + 4;
        const int xwords = spread ? (int)SPREAD_XWORDS : 0;
        const int init = std::max(std::max(m, xwords), spread ? steps : 0);
        hipLaunchKernelGGL(spread_init_kernel, dim3((init + 255) / 256), dim3(256), 0, ctx->stream, m, d_perm, rc,
                           xbuf, xwords, lists, spread ? steps : 0, lstride);
        if (clk) clk->mark(PhaseClock::NONE);
        const int ob = lu_outer_block<T>(m, n, nb);
        // Look-ahead needs the second stream, no phase clock (its marks time
        // one stream) and no graph capture.
        const bool ahead = spread && !whole && steps > 1 && !clk && ctx->side && !ctx->hold;
        if (ahead) {
            getrf_lookahead<T>(ctx, m, n, a, lda, ib, ob, d_piv, d_perm, d_info, lists, lstride, xbuf);
        } else if (whole) {
            if (spread) {
                panel_spread<T>(ctx, m, n, a, lda, d_piv, 0, d_perm, d_info, lists, xbuf, 0, 0, clk);
            } else {
                panel<T>(ctx, m, n, a, lda, d_piv, 0, d_perm, rc, d_info, list, 0, 0);
                if (clk) clk->mark(PhaseClock::PANEL);
            }
            if (phases) phases->steps += 1;
        } else {
            // Outer steps of ob columns: inside one, each panel's U12 solve
            // and trailing update reach only the outer step's columns; the
            // columns beyond take their block row by one solve with the outer
            // step's L11 and one product of depth ob, so the trailing matrix
            // is read and written once per outer step instead of once per
            // panel. Every panel sees the same updated columns as with
            // ob = nb, so the pivots are the same.
            int step = 0;
            for (int jo = 0; jo < limit; jo += ob) {
                const int wo = ob < limit - jo ? ob : limit - jo;
                const int end = jo + wo < n ? jo + wo : n;  // columns the inner steps update
                for (int jb = jo; jb < jo + wo; jb += ib, ++step) {
                    const int w = ib < jo + wo - jb ? ib : jo + wo - jb;
                    T* pa = a + jb + (size_t)jb * lda;
                    if (spread) {
                        panel_spread<T>(ctx, m - jb, w, pa, lda, d_piv + jb, jb, d_perm + jb, d_info,
                                        lists + (size_t)step * lstride, xbuf, jb, n - jb - w, clk);
                    } else {
                        panel<T>(ctx, m - jb, w, pa, lda, d_piv + jb, jb, d_perm + jb, rc + jb, d_info, list, jb,
                                 n - jb - w);
                        if (clk) clk->mark(PhaseClock::PANEL);
                    }
                    if (phases) phases->steps += 1;
                    const int right = end - jb - w;
                    if (right <= 0) continue;
                    trsm_llnu<T>(ctx, w, right, pa, lda, pa + (size_t)w * lda, lda);
                    if (clk) clk->mark(PhaseClock::TRSM);
                    if (jb + w >= m) continue;
                    Ops<T>::gemm(ctx, 0, m - jb - w, right, w, Ops<T>::neg_one(), pa + w, lda, pa + (size_t)w * lda,
                                 lda, Ops<T>::one(), pa + w + (size_t)w * lda, lda);
                    if (clk) clk->mark(PhaseClock::GEMM);
                }
                const int beyond = n - end;
                if (beyond <= 0) continue;
                T* po = a + jo + (size_t)jo * lda;
                trsm_llnu<T>(ctx, wo, beyond, po, lda, po + (size_t)wo * lda, lda);
                if (clk) clk->mark(PhaseClock::TRSM);
                if (jo + wo >= m) continue;
                Ops<T>::gemm(ctx, 0, m - jo - wo, beyond, wo, Ops<T>::neg_one(), po + wo, lda, po + (size_t)wo * lda,
                             lda, Ops<T>::one(), po + wo + (size_t)wo * lda, lda);
                if (clk) clk->mark(PhaseClock::GEMM);
            }
        }
    }
    hipLaunchKernelGGL(finish_info_kernel, dim3(1), dim3(1), 0, ctx->stream, d_info);
    if (spread_used) hipLaunchKernelGGL(spread_check_kernel, dim3(1), dim3(1), 0, ctx->stream, xbuf, d_info);
    TNS_HIP(hipGetLastError());
    if (clk) {
        clk->mark(PhaseClock::NONE);
        clk->read(phases);
    }
}

// LU with the pivots of each block step chosen inside its NB x NB
// diagonal block (block-local partial pivoting), right-looking in blocks of
// NB = 64 columns: the diagonal block by local_block_lu_kernel, its
// interchanges moved across the whole rows, the inverses of its two
// triangles, then the rows below it L21 = A21 U11^-1 and the block row
// U12 = L11^-1 A12 as products with those inverses (in place: each product
// is one tile wide in the dimension it overwrites, so every tile reads its
// operand before it writes), and the trailing update, all on sgemm_lp.
// The block row and the trailing update reach only the columns of the
// current outer panel of LOCAL_OB columns; the columns beyond it take
// their block row by block substitution and one product of depth LOCAL_OB
// at the end of the outer step. The
// pivot search sees 64 rows instead of the whole column, so no step waits
// on a tall panel; it is for matrices whose randomized form needs no wider
// search (the mixed-precision solve), where the refinement absorbs the
// growth that remains. P A = L U with perm as in getrf_impl; inv keeps
// every diagonal block's inverses for the solves; a zero pivot sets info.
constexpr int LOCAL_NB = 64;
constexpr int LOCAL_OB = 256;  // outer step of getrf_local_impl, a multiple of LOCAL_NB

// The LOCAL_NB x LOCAL_NB diagonal block of getrf_local_impl factored by
// four wavefronts: lane i of every wavefront holds row i of the block, and
// wavefront q its columns [16q, 16q + 16), in registers throughout. As in
// getf2_reg_body, interchanges move no data (each lane carries its row's
// current position; the pivot takes position j, the row at j takes the
// pivot's), the pivot is the largest |x| with ties to the lower position,
// a zero pivot leaves the rows in place, and the multipliers and updates
// are formed the same way. Column j's wavefront finds the pivot (DPP
// argmax), forms the multipliers and leaves them in LDS with the pivot's
// lane and position; after one barrier every wavefront updates its columns
// right of j, reading the pivot row's entries from its own registers by
// readlane. Rows already pivoted (this step's pivot row included) take a
// zero multiplier, which leaves them unchanged, so the update needs no
// per-element select; unlike the register panel kernel, zero entries of
// the pivot row are not skipped (x + (-0) l = x for the finite l of a
// nonzero pivot). The rows are stored to their final positions at the
// end; piv and perm (from row jb) and the move list (block-relative) are as
// getf2_reg writes them, and info is min-reduced with the first zero-pivot
// column.
constexpr int LOCAL_WAVES = 4;
constexpr int LOCAL_WC = LOCAL_NB / LOCAL_WAVES;  // columns per wavefront

template <int WQ>
__device__ __forceinline__ void local_block_body(float (&v)[LOCAL_WC], int lane, int& pos, bool& done, int& my_piv,
                                                 int& first_zero, int jb, float (*lm)[LOCAL_NB], int (*pk)[4])
{
    constexpr int W = LOCAL_NB, WC = LOCAL_WC;
    auto column = [&](auto jc) __attribute__((always_inline)) {
        constexpr int j = decltype(jc)::value;
        constexpr int buf = j & 1;
        if constexpr (j / WC == WQ) {
            constexpr int cl = j % WC;
            double bv = done ? -2.0 : pivot_key(v[cl], pos, j);
            int bi = done ? INT_MAX : pos;
            wave_best_dpp(bv, bi);
            const int pl = __ffsll((unsigned long long)__ballot(!done && pos == bi)) - 1;  // the pivot's lane
            const float pv = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[cl]), pl));
            const bool zero_pivot = pv == 0.0f;
            // Multipliers: x * (1/pivot) when |pivot| >= the safe minimum,
            // x / pivot below it, x itself for a zero pivot.
            float l = v[cl];
            if (!zero_pivot) {
                if (fabsf(pv) >= (float)safe_min<float>()) l = l * (1.0f / pv);
                else l = l / pv;
            }
            const int pe = zero_pivot ? j : bi;
            const bool now_done = done || (pos == pe);
            lm[buf][lane] = now_done ? 0.0f : l;
            if (!now_done) v[cl] = l;
            if (lane == 0) {
                pk[buf][0] = pl;
                pk[buf][1] = bi;
                pk[buf][2] = zero_pivot;
            }
        }
        __syncthreads();
        const int pl = pk[buf][0], bi = pk[buf][1];
        const bool zero_pivot = pk[buf][2] != 0;
        const float lmv = lm[buf][lane];
        if (lane == j) my_piv = jb + bi;
        if (zero_pivot && first_zero < 0) first_zero = jb + j;
        const int pe = zero_pivot ? j : bi;
        const bool is_p = !done && pos == pe;
        const bool is_j = !done && pos == j && !is_p;
        done = done || is_p;
        pos = is_p ? j : (is_j ? pe : pos);
#pragma unroll
        for (int cl = 0; cl < WC; cl++) {
            if constexpr (true) {
                const int c = WQ * WC + cl;
                if (c > j) {
                    const float u = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[cl]), pl));
                    v[cl] = v[cl] + (-u) * lmv;
                }
            }
        }
    };
    static_for<W>(column);
}

__global__ void __launch_bounds__(64 * LOCAL_WAVES) local_block_lu_kernel(float* __restrict__ a, int lda, int jb,
                                                                         int* __restrict__ piv, int* __restrict__ perm,
                                                                         int* __restrict__ list, int* __restrict__ info)
{
    constexpr int WC = LOCAL_WC;
    __shared__ float lm[2][LOCAL_NB];
    __shared__ int pk[2][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    float* blk = a + jb + (size_t)jb * lda;
    float v[WC];
#pragma unroll
    for (int cl = 0; cl < WC; cl++) v[cl] = blk[lane + (size_t)(wave * WC + cl) * lda];
    const int p0 = perm[jb + lane];
    int pos = lane;
    bool done = false;
    int first_zero = -1;
    int my_piv = 0;  // lane j keeps piv[jb + j]
    switch (wave) {
    case 0: local_block_body<0>(v, lane, pos, done, my_piv, first_zero, jb, lm, pk); break;
    case 1: local_block_body<1>(v, lane, pos, done, my_piv, first_zero, jb, lm, pk); break;
    case 2: local_block_body<2>(v, lane, pos, done, my_piv, first_zero, jb, lm, pk); break;
    default: local_block_body<3>(v, lane, pos, done, my_piv, first_zero, jb, lm, pk); break;
    }
    __syncthreads();  // every lane's perm entry read before any is written
#pragma unroll
    for (int cl = 0; cl < WC; cl++) blk[pos + (size_t)(wave * WC + cl) * lda] = v[cl];
    if (wave == 0) {
        perm[jb + pos] = p0;
        piv[jb + lane] = my_piv;
        const unsigned long long moved = __ballot(pos != lane);
        if (pos != lane) {
            const int at = __popcll(moved & ((1ull << lane) - 1ull));
            list[1 + 2 * at] = pos;
            list[2 + 2 * at] = lane;
        }
        if (lane == 0) {
            list[0] = __popcll(moved);
            if (first_zero >= 0) atomicMin(info, first_zero);
        }
    }
}

template <typename T>
void getrf_local_impl(Context* ctx, int n, T* a, int lda, int* d_piv, int* d_perm, T* inv, int* d_info)
{
    constexpr int NB = LOCAL_NB;
    static_assert(sizeof(T) == 4, "the register kernels hold 64 columns of float");
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, ctx->stream, d_info, INT_MAX);
    if (n <= 0) {
        hipLaunchKernelGGL(finish_info_kernel, dim3(1), dim3(1), 0, ctx->stream, d_info);
        return;
    }
    DevBuf<int> rc(ctx, (size_t)n);
    DevBuf<int> list(ctx, 4 * (size_t)NB + 4);
    hipLaunchKernelGGL(iota_kernel, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, d_perm, rc.get());
    const T one = Ops<T>::one(), zero{}, neg = Ops<T>::neg_one();
    // Outer steps of OB columns: inside one, each NB block step updates only
    // the outer panel's columns; the columns beyond it get one product of
    // depth OB per outer step (a quarter of the trailing matrix's reads and
    // writes of depth-NB updates). Interchanges still move whole rows at
    // once: they permute rows of the pending update's L and A alike.
    // Below n = 8192 the saved traffic is smaller than the cost of the
    // extra launches (measured at 4096: 18.4 -> 19.5 ms with outer steps),
    // and an outer step of one block is the plain right-looking order.
    const int OB = n >= 8192 ? LOCAL_OB : NB;
    for (int J = 0; J < n; J += OB) {
        const int ow = OB < n - J ? OB : n - J;
        for (int jb = J; jb < J + ow; jb += NB) {
            const int w = NB < n - jb ? NB : n - jb;
            T* pa = a + jb + (size_t)jb * lda;
            if (w == NB) {
                // Four wavefronts for the whole diagonal block, then its
                // interchanges across the rest of the rows.
                hipLaunchKernelGGL(local_block_lu_kernel, dim3(1), dim3(64 * LOCAL_WAVES), 0, ctx->stream, a, lda,
                                   jb, d_piv, d_perm, list.get(), d_info);
                apply_moves<T>(ctx, a + jb, lda, jb, jb + w, n, list, 2 * w);
            } else {
                panel<T>(ctx, w, w, pa, lda, d_piv + jb, jb, d_perm + jb, rc.get() + jb, d_info, list, jb,
                         n - jb - w, 32);
            }
            hipLaunchKernelGGL((tri_inverse_kernel<T, NB>), dim3(1), dim3(2 * NB), 0, ctx->stream, n, a, lda, jb / NB,
                               inv);
            const int rest = n - jb - w;
            if (rest <= 0) continue;
            const T* inv_l = inv + (size_t)(jb / NB) * 2 * NB * NB;
            const T* inv_u = inv_l + NB * NB;
            // The products with the inverses stay FP32 in both modes.
            sgemm_lp(ctx, LP_FP32, rest, w, w, one, pa + w, lda, inv_u, NB, zero, pa + w, lda);
            const int inner = J + ow - (jb + w);  // the outer panel's columns still to come
            if (inner <= 0) continue;
            sgemm_lp(ctx, LP_FP32, w, inner, w, one, inv_l, NB, pa + (size_t)w * lda, lda, zero, pa + (size_t)w * lda,
                     lda);
            Ops<T>::gemm(ctx, 0, rest, inner, w, neg, pa + w, lda, pa + (size_t)w * lda, lda, one,
                         pa + w + (size_t)w * lda, lda);
        }
        const int cols = n - (J + ow);  // beyond the outer panel
        if (cols <= 0) continue;
        // U12 of the outer panel by block forward substitution with the
        // diagonal blocks' inverses: U_j = inv(L_jj) (A_j - L_j,<j U_<j).
        T* a12 = a + J + (size_t)(J + ow) * lda;
        for (int r0 = J; r0 < J + ow; r0 += NB) {
            const int w = NB < J + ow - r0 ? NB : J + ow - r0;
            T* uj = a + r0 + (size_t)(J + ow) * lda;
            if (r0 > J)
                Ops<T>::gemm(ctx, 0, w, cols, r0 - J, neg, a + r0 + (size_t)J * lda, lda, a12, lda, one, uj, lda);
            const T* inv_l = inv + (size_t)(r0 / NB) * 2 * NB * NB;
            sgemm_lp(ctx, LP_FP32, w, cols, w, one, inv_l, NB, uj, lda, zero, uj, lda);
        }
        const int below = n - (J + ow);
        Ops<T>::gemm(ctx, 0, below, cols, ow, neg, a + (J + ow) + (size_t)J * lda, lda, a12, lda, one,
                     a + (J + ow) + (size_t)(J + ow) * lda, lda);
    }
    hipLaunchKernelGGL(finish_info_kernel, dim3(1), dim3(1), 0, ctx->stream, d_info);
    TNS_HIP(hipGetLastError());
}

// One diagonal super block (rows r0 .. r0 + rows, rows a multiple of W <=
// 1024) of a triangular solve with the factors of getrf_local_impl, for
// TC right-hand sides per workgroup (columns blockIdx.x * TC ...): the
// block's part of x sits in LDS while its W-blocks are solved in turn,
// x_j := inv_jj x_j (the diagonal blocks' inverses), and each result is
// subtracted from the block's rows still to come. One launch does the work
// of 2 rows / W launches of the block-by-block form. LOWER: unit lower,
// forward; else upper, backward.
template <typename T, int W, int TC, bool LOWER>
__global__ void __launch_bounds__(256) trsv_super_kernel(int r0, int rows, int t, const T* __restrict__ lu, int ld,
                                                         const T* __restrict__ inv, T* __restrict__ x, int ldx)
{
    extern __shared__ unsigned char super_smem[];
    T* xs = reinterpret_cast<T*>(super_smem);  // xs[i * TC + c], i < rows
    __shared__ T iv[2][W][W + 1];              // this block's inverse, and the next one's
    constexpr int IVT = W * W / 256;           // inverse entries per thread
    const int tid = threadIdx.x, c0 = blockIdx.x * TC;
    const int nc = t - c0 < TC ? t - c0 : TC;
    const int nb = rows / W;
    auto inverse_of = [&](int s) {
        const int jb = LOWER ? s : nb - 1 - s;
        return inv + (size_t)((r0 + jb * W) / W) * 2 * W * W + (LOWER ? 0 : W * W);
    };
    for (int e = tid; e < rows * TC; e += 256) {
        const int i = e / TC, c = e % TC;
        xs[e] = load_if(x + r0 + i + (size_t)(c0 + c) * ldx, c < nc);
    }
    {
        const T* src = inverse_of(0);
#pragma unroll
        for (int u = 0; u < IVT; u++) {
            const int e = tid + u * 256;
            iv[0][e % W][e / W] = src[e];
        }
    }
    __syncthreads();
    for (int s = 0; s < nb; s++) {
        const int jb = LOWER ? s : nb - 1 - s;
        const int k0 = jb * W, buf = s & 1;
        // The next block's inverse, in flight while this block is worked.
        T nxt[IVT];
        if (s + 1 < nb) {
            const T* src = inverse_of(s + 1);
#pragma unroll
            for (int u = 0; u < IVT; u++) nxt[u] = src[tid + u * 256];
        }
        // x_j := inv_jj x_j.
        T y[(W * TC + 255) / 256];
#pragma unroll
        for (int u = 0; u < (W * TC + 255) / 256; u++) {
            const int e = tid + u * 256;
            const int i = e / TC, c = e % TC;
            T acc = s_zero(T{});
            if (e < W * TC)
                for (int p = 0; p < W; p++) acc = s_add(acc, s_mul(iv[buf][i][p], xs[(k0 + p) * TC + c]));
            y[u] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < (W * TC + 255) / 256; u++) {
            const int e = tid + u * 256;
            if (e < W * TC) xs[(k0 + e / TC) * TC + e % TC] = y[u];
        }
        __syncthreads();
        // The block's other rows still to come: one row per thread, all TC
        // columns, the row's W entries of the factor loaded together.
        const int lo = LOWER ? k0 + W : 0, hi = LOWER ? rows : k0;
        for (int i = lo + tid; i < hi; i += 256) {
            const T* f = lu + (r0 + i) + (size_t)(r0 + k0) * ld;
            T a[W];
#pragma unroll
            for (int p = 0; p < W; p++) a[p] = f[(size_t)p * ld];
            T acc[TC];
#pragma unroll
            for (int c = 0; c < TC; c++) acc[c] = xs[i * TC + c];
#pragma unroll
            for (int p = 0; p < W; p++) {
                const T na = s_neg(a[p]);
#pragma unroll
                for (int c = 0; c < TC; c++) acc[c] = s_add(acc[c], s_mul(na, xs[(k0 + p) * TC + c]));
            }
#pragma unroll
            for (int c = 0; c < TC; c++) xs[i * TC + c] = acc[c];
        }
        if (s + 1 < nb) {
#pragma unroll
            for (int u = 0; u < IVT; u++) {
                const int e = tid + u * 256;
                iv[buf ^ 1][e % W][e / W] = nxt[u];
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < rows * TC; e += 256) {
        const int i = e / TC, c = e % TC;
        if (c < nc) x[r0 + i + (size_t)(c0 + c) * ldx] = xs[e];
    }
}

// P A = L U solve with the factors of getrf_local_impl (order N a multiple
// of LOCAL_NB): P b, then L forward and U backward in super blocks of up to
// 512 rows, each by trsv_super_kernel and the rows beyond it by one
// product.
template <typename T>
void getrs_local_impl(Context* ctx, int N, int nrhs, const T* lu, int ldlu, const int* d_perm, const T* inv, T* b,
                      int ldb)
{
    constexpr int W = LOCAL_NB, TC = 1, SUPER = 512;
    if (N <= 0 || nrhs <= 0) return;
    if (N % W) fail(TNS_EINVAL, "the local LU's order is a multiple of its block");
    DevBuf<T> bx(ctx, (size_t)N * nrhs);
    T* x = bx.get();
    const T one = Ops<T>::one();
    permute_rows<T>(ctx, N, nrhs, d_perm, b, ldb, x, N, 0);
    const dim3 grid((nrhs + TC - 1) / TC);
    const int nsup = (N + SUPER - 1) / SUPER;
    for (int q = 0; q < nsup; q++) {
        const int r0 = q * SUPER, rows = SUPER < N - r0 ? SUPER : N - r0;
        hipLaunchKernelGGL((trsv_super_kernel<T, W, TC, true>), grid, dim3(256), sizeof(T) * rows * TC, ctx->stream,
                           r0, rows, nrhs, lu, ldlu, inv, x, N);
        if (r0 + rows < N)
            Ops<T>::gemm(ctx, 0, N - r0 - rows, nrhs, rows, Ops<T>::neg_one(), lu + (r0 + rows) + (size_t)r0 * ldlu,
                         ldlu, x + r0, N, one, x + r0 + rows, N);
    }
    for (int q = nsup - 1; q >= 0; q--) {
        const int r0 = q * SUPER, rows = SUPER < N - r0 ? SUPER : N - r0;
        hipLaunchKernelGGL((trsv_super_kernel<T, W, TC, false>), grid, dim3(256), sizeof(T) * rows * TC, ctx->stream,
                           r0, rows, nrhs, lu, ldlu, inv, x, N);
        if (r0 > 0)
            Ops<T>::gemm(ctx, 0, r0, nrhs, rows, Ops<T>::neg_one(), lu + (size_t)r0 * ldlu, ldlu, x + r0, N, one, x,
                         N);
    }
    copy2d<T>(ctx, N, nrhs, x, N, b, ldb);
    TNS_HIP(hipGetLastError());
}

template <typename T>
void getrs_impl(Context* ctx, int trans, int n, int nrhs, const T* lu, int ldlu, const int* d_perm, T* b, int ldb)
{
    if (n <= 0 || nrhs <= 0) return;
    DevBuf<T> tmp(ctx, (size_t)n * nrhs);
    if (!trans) {
        // P b, then L forward, then U backward (Lu.cs:336-351).
        permute_rows<T>(ctx, n, nrhs, d_perm, b, ldb, tmp, n, 0);
        trsm_impl<T, true, 0, true>(ctx, n, nrhs, lu, ldlu, tmp, n);
        trsm_impl<T, false, 0, false>(ctx, n, nrhs, lu, ldlu, tmp, n);
        copy2d<T>(ctx, n, nrhs, tmp, n, b, ldb);
    } else {
        // U^H forward, L^H backward, then the permutation undone
        // (Lu.cs:366-378, BlockedLu.cs:126-141).
        copy2d<T>(ctx, n, nrhs, b, ldb, tmp, n);
        constexpr int TR = sizeof(T) == 8 ? 1 : 2;
        trsm_impl<T, false, TR, false>(ctx, n, nrhs, lu, ldlu, tmp, n);
        trsm_impl<T, true, TR, true>(ctx, n, nrhs, lu, ldlu, tmp, n);
        permute_rows<T>(ctx, n, nrhs, d_perm, tmp, n, b, ldb, 1);
    }
    TNS_HIP(hipGetLastError());
}

// A X = B (op = none) against the LU factors, for many right-hand sides
// of a well-conditioned A (the Pade denominator): the inverses of every
// diagonal block of L and U come from one launch (the blocks are
// independent), and each block step is then two products on the matrix
// cores, X_k = inv(D_kk) B_k and the update of the blocks still to come,
// where the substitution form spends a launch with one barrier per row
// of the block. The products go between two buffers, so none is in place.
// The rounding differs from the substitution's; the backward error of a
// well-conditioned block's explicit inverse is of the same order.
// Writes every diagonal block's two inverses (tri_inverse_lds_kernel's
// layout: per block, inv(L_kk) then inv(U_kk), each BS x BS column-major)
// into the diagonal blocks of the n x n matrices il and iu.
template <typename T, int BS>
__global__ void scatter_block_inverses_kernel(int n, const T* __restrict__ inv, T* __restrict__ il,
                                              T* __restrict__ iu)
{
    const int blk = blockIdx.x, k0 = blk * BS;
    const int kb = BS < n - k0 ? BS : n - k0;
    const T* src = inv + (size_t)blk * 2 * BS * BS;
    for (int e = threadIdx.x; e < BS * BS; e += blockDim.x) {
        const int i = e % BS, j = e / BS;
        if (i < kb && j < kb) {
            il[(k0 + i) + (size_t)(k0 + j) * n] = src[e];
            iu[(k0 + i) + (size_t)(k0 + j) * n] = src[BS * BS + e];
        }
    }
}

template <typename T>
__global__ void zero_fill_kernel(size_t count, T* __restrict__ p)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
        p[e] = s_zero(T{});
}

// X = (LU)^-1 P B for many right-hand sides (the Pade denominator's
// n-column solve) by blocked substitution over diagonal blocks of order
// HB = 128, with those blocks' inverses. The inverses of the BS-order
// diagonal blocks come from one launch; pairs of them are merged into
// inverses of twice the order until HB,
//     inv(L) = [inv(L11) 0; -inv(L22) L21 inv(L11)  inv(L22)],
//     inv(U) = [inv(U11) -inv(U11) U12 inv(U22); 0  inv(U22)],
// each merge level two batched products over its pairs (the last, shorter
// pair on its own), every product skipping the zero half of its
// triangular factor. Then per HB block, forward: X_k = inv(L_kk) B_k,
// B_rest -= L_rest,k X_k; backward likewise with U. Updates of depth HB
// instead of BS: a quarter of the launches and of the passes over the
// right-hand sides. The denominator is well conditioned at the degrees and
// scalings the selection allows, so the explicit inverses cost no accuracy
// the tests can see.
template <typename T>
void getrs_inverse_impl(Context* ctx, int n, int nrhs, const T* lu, int ldlu, const int* d_perm, T* b, int ldb)
{
    if (n <= 0 || nrhs <= 0) return;
    constexpr int BS = tri_block<T>();
    const int nblk = (n + BS - 1) / BS;
    DevBuf<T> inv(ctx, (size_t)nblk * 2 * BS * BS);
    hipLaunchKernelGGL((tri_inverse_lds_kernel<T, BS>), dim3(nblk), dim3(2 * BS), 0, ctx->stream, n, lu, ldlu,
                       inv.get());
    const size_t nn = (size_t)n * n;
    DevBuf<T> il(ctx, nn), iu(ctx, nn), tmp(ctx, (size_t)n * (nrhs > n ? nrhs : n));
    {
        const int blocks = (int)((2 * nn + 255) / 256 < 8192 ? (2 * nn + 255) / 256 : 8192);
        hipLaunchKernelGGL((zero_fill_kernel<T>), dim3(blocks), dim3(256), 0, ctx->stream, nn, il.get());
        hipLaunchKernelGGL((zero_fill_kernel<T>), dim3(blocks), dim3(256), 0, ctx->stream, nn, iu.get());
    }
    hipLaunchKernelGGL((scatter_block_inverses_kernel<T, BS>), dim3(nblk), dim3(256), 0, ctx->stream, n, inv.get(),
                       il.get(), iu.get());
    const T one = Ops<T>::one(), neg = Ops<T>::neg_one(), zero{};
    T* IL = il.get();
    T* IU = iu.get();
    T* W = tmp.get();
    constexpr int HB = 128;
    for (int sz = BS; sz < HB && sz < n; sz *= 2) {
        // Merge pairs of inverted blocks of order sz that lie inside one
        // HB block (HB is a multiple of 2 sz, so every pair does).
        const int full = n / (2 * sz);  // pairs of two whole blocks of order sz
        auto level = [&](int lo, int r, int count) {
            // Pair(s) from row/column lo: first block [lo, lo + sz), second
            // [lo + sz, lo + sz + r); consecutive pairs 2 sz apart.
            const int mid = lo + sz;
            const long long sl = 2LL * sz * (1 + (long long)ldlu), sn = 2LL * sz * (1 + (long long)n);
            // W21 = L21 inv(L11) (inv(L11) lower: B lower), then
            // inv(L)21 = -inv(L22) W21 (A lower).
            Ops<T>::batch(ctx, r, sz, sz, one, lu + mid + (size_t)lo * ldlu, ldlu, sl, IL + lo + (size_t)lo * n, n, sn,
                          zero, W + mid + (size_t)lo * n, n, sn, count, 4);
            Ops<T>::batch(ctx, r, sz, r, neg, IL + mid + (size_t)mid * n, n, sn, W + mid + (size_t)lo * n, n, sn, zero,
                          IL + mid + (size_t)lo * n, n, sn, count, 1);
            // W12 = inv(U11) U12 (A upper), then inv(U)12 = -W12 inv(U22)
            // (B upper).
            Ops<T>::batch(ctx, sz, r, sz, one, IU + lo + (size_t)lo * n, n, sn, lu + lo + (size_t)mid * ldlu, ldlu, sl,
                          zero, W + lo + (size_t)mid * n, n, sn, count, 2);
            Ops<T>::batch(ctx, sz, r, r, neg, W + lo + (size_t)mid * n, n, sn, IU + mid + (size_t)mid * n, n, sn, zero,
                          IU + lo + (size_t)mid * n, n, sn, count, 3);
        };
        if (full > 0) level(0, sz, full);
        const int lo = full * 2 * sz;
        if (lo + sz < n) level(lo, n - lo - sz, 1);
    }
    // Blocked substitution over the HB blocks, ping-ponging t1 and W.
    DevBuf<T> t1(ctx, (size_t)n * nrhs);
    T* x1 = t1.get();
    permute_rows<T>(ctx, n, nrhs, d_perm, b, ldb, x1, n, 0);
    const int hblk = (n + HB - 1) / HB;
    for (int k = 0; k < hblk; k++) {
        const int k0 = k * HB, kb = n - k0 < HB ? n - k0 : HB;
        Ops<T>::batch(ctx, kb, nrhs, kb, one, IL + k0 + (size_t)k0 * n, n, 0, x1 + k0, n, 0, zero, W + k0, n, 0, 1, 1);
        if (k0 + kb < n)
            Ops<T>::gemm(ctx, 0, n - k0 - kb, nrhs, kb, neg, lu + (k0 + kb) + (size_t)k0 * ldlu, ldlu, W + k0, n, one,
                         x1 + k0 + kb, n);
    }
    for (int k = hblk - 1; k >= 0; k--) {
        const int k0 = k * HB, kb = n - k0 < HB ? n - k0 : HB;
        Ops<T>::batch(ctx, kb, nrhs, kb, one, IU + k0 + (size_t)k0 * n, n, 0, W + k0, n, 0, zero, x1 + k0, n, 0, 1, 2);
        if (k0 > 0)
            Ops<T>::gemm(ctx, 0, k0, nrhs, kb, neg, lu + (size_t)k0 * ldlu, ldlu, x1 + k0, n, one, W, n);
    }
    copy2d<T>(ctx, n, nrhs, x1, n, b, ldb);
    TNS_HIP(hipGetLastError());
}

}  // namespace

size_t lu_work_ints(int m, int n, int nb)
{
    const int limit = m < n ? m : n;
    if (nb < 1) nb = limit < 1024 ? 32 : 64;
    const int maxw = nb < limit ? nb : n;
    // Panels are at least 32 columns wide (the complex spread kernel's width).
    const int steps = nb >= limit ? 1 : (limit + 31) / 32 + (limit + nb - 1) / nb;
    return (size_t)m + 4 * (size_t)maxw + 4 + (size_t)steps * SPREAD_LIST + 2 * SPREAD_XWORDS;
}

void dgetrf(Context* ctx, int m, int n, double* a, int lda, int nb, int* d_piv, int* d_perm, int* d_info,
            LuPhases* phases, int* work)
{
    getrf_impl<double>(ctx, m, n, a, lda, nb, d_piv, d_perm, d_info, phases, work);
}

void zgetrf(Context* ctx, int m, int n, cplx* a, int lda, int nb, int* d_piv, int* d_perm, int* d_info,
            LuPhases* phases, int* work)
{
    getrf_impl<zd>(ctx, m, n, a, lda, nb, d_piv, d_perm, d_info, phases, work);
}

void sgetrf(Context* ctx, int m, int n, float* a, int lda, int nb, int* d_piv, int* d_perm, int* d_info)
{
    getrf_impl<float>(ctx, m, n, a, lda, nb, d_piv, d_perm, d_info);
}

void sgetrs(Context* ctx, int n, int nrhs, const float* lu, int ldlu, const int* d_perm, float* b, int ldb)
{
    getrs_impl<float>(ctx, 0, n, nrhs, lu, ldlu, d_perm, b, ldb);
}

int local_lu_block() { return LOCAL_NB; }

void sgetrf_local(Context* ctx, int n, float* a, int lda, int* d_piv, int* d_perm, float* inv, int* d_info)
{
    getrf_local_impl<float>(ctx, n, a, lda, d_piv, d_perm, inv, d_info);
}

void sgetrs_local(Context* ctx, int n, int nrhs, const float* lu, int ldlu, const int* d_perm, const float* inv,
                  float* b, int ldb)
{
    getrs_local_impl<float>(ctx, n, nrhs, lu, ldlu, d_perm, inv, b, ldb);
}

void dgetrs(Context* ctx, int trans, int n, int nrhs, const double* lu, int ldlu, const int* d_perm, double* b, int ldb)
{
    getrs_impl<double>(ctx, trans, n, nrhs, lu, ldlu, d_perm, b, ldb);
}

void zgetrs(Context* ctx, int trans, int n, int nrhs, const cplx* lu, int ldlu, const int* d_perm, cplx* b, int ldb)
{
    getrs_impl<zd>(ctx, trans, n, nrhs, lu, ldlu, d_perm, b, ldb);
}

void dgetrs_blockinv(Context* ctx, int n, int nrhs, const double* lu, int ldlu, const int* d_perm, double* b, int ldb)
{
    getrs_inverse_impl<double>(ctx, n, nrhs, lu, ldlu, d_perm, b, ldb);
}

void zgetrs_blockinv(Context* ctx, int n, int nrhs, const cplx* lu, int ldlu, const int* d_perm, cplx* b, int ldb)
{
    getrs_inverse_impl<zd>(ctx, n, nrhs, lu, ldlu, d_perm, b, ldb);
}

void dtrsm(Context* ctx, int kind, int m, int n, const double* a, int lda, double* b, int ldb)
{
    switch (kind) {
    case TNS_TRI_LOWER_UNIT: trsm_impl<double, true, 0, true>(ctx, m, n, a, lda, b, ldb); break;
    case TNS_TRI_UPPER: trsm_impl<double, false, 0, false>(ctx, m, n, a, lda, b, ldb); break;
    case TNS_TRI_LOWER: trsm_impl<double, true, 0, false>(ctx, m, n, a, lda, b, ldb); break;
    case TNS_TRI_UPPER_T: trsm_impl<double, false, 1, false>(ctx, m, n, a, lda, b, ldb); break;
    case TNS_TRI_LOWER_T: trsm_impl<double, true, 1, false>(ctx, m, n, a, lda, b, ldb); break;
    case TNS_TRI_LOWER_UNIT_T: trsm_impl<double, true, 1, true>(ctx, m, n, a, lda, b, ldb); break;
    default: fail(TNS_EINVAL, "unknown triangular solve kind");
    }
    TNS_HIP(hipGetLastError());
}

void dpivot_range(Context* ctx, int n, const double* a, int lda, double* d_out)
{
    hipLaunchKernelGGL((pivot_range_kernel<double>), dim3(1), dim3(256), 0, ctx->stream, n, a, lda, d_out);
}

void zpivot_range(Context* ctx, int n, const cplx* a, int lda, double* d_out)
{
    hipLaunchKernelGGL((pivot_range_kernel<zd>), dim3(1), dim3(256), 0, ctx->stream, n, a, lda, d_out);
}

void dcopy2d(Context* ctx, int m, int n, const double* src, int lds, double* dst, int ldd)
{
    copy2d<double>(ctx, m, n, src, lds, dst, ldd);
}

void zcopy2d(Context* ctx, int m, int n, const cplx* src, int lds, cplx* dst, int ldd)
{
    copy2d<zd>(ctx, m, n, src, lds, dst, ldd);
}

// ---------------------------------------------------------------------------
// Grouped launches (grouped.hpp).
// ---------------------------------------------------------------------------
int zpanel_chunk_width(int m) { return reg_panel_width<zd>(m); }

// The getf2_reg variant for height m, in getf2_reg's order.
int zpanel_class(int m)
{
    if (m <= 512) return 0;
    if (m <= 1024) return 1;
    if (m <= 2048) return 2;
    if (m <= 4096) return 3;
    return -1;
}

void zgrouped_init(Context* ctx, const InitRec<zd>* recs, int count)
{
    if (count <= 0) return;
    hipLaunchKernelGGL((lu_init_grouped_kernel<zd>), dim3(count), dim3(1024), 0, ctx->stream, recs);
    TNS_HIP(hipGetLastError());
}

namespace {
template <int NB, int R>
void launch_grouped_panel(Context* ctx, const PanelRec<zd>* recs, int count, int max_m)
{
    const int need = (max_m + R - 1) / R;
    int nt = ((need + 63) / 64) * 64;
    if (nt < 64) nt = 64;
    hipLaunchKernelGGL((getf2_grouped_kernel<zd, NB, R, 512>), dim3(count), dim3(nt), 0, ctx->stream, recs);
}
}  // namespace

void zgrouped_panel(Context* ctx, const PanelRec<zd>* recs, int count, int max_m)
{
    if (count <= 0) return;
    switch (zpanel_class(max_m)) {
    case 0: launch_grouped_panel<32, 1>(ctx, recs, count, max_m); break;
    case 1: launch_grouped_panel<16, 2>(ctx, recs, count, max_m); break;
    case 2: launch_grouped_panel<8, 4>(ctx, recs, count, max_m); break;
    case 3: launch_grouped_panel<4, 8>(ctx, recs, count, max_m); break;
    default: fail(TNS_EINVAL, "a panel taller than the register kernel holds");
    }
    TNS_HIP(hipGetLastError());
}

int grouped_moves_columns(int max_moves) { return max_moves <= 32 ? 256 / 32 : 256 / 64; }

void zgrouped_moves(Context* ctx, const MovesRec<zd>* recs, const int* starts, int count, int blocks, int max_moves)
{
    if (count <= 0 || blocks <= 0) return;
    if (max_moves <= 32)
        hipLaunchKernelGGL((apply_moves_grouped_kernel<zd, 32>), dim3(blocks), dim3(256), 0, ctx->stream, recs, starts,
                           count);
    else if (max_moves <= 64)
        hipLaunchKernelGGL((apply_moves_grouped_kernel<zd, 64>), dim3(blocks), dim3(256), 0, ctx->stream, recs, starts,
                           count);
    else
        fail(TNS_EINVAL, "a move list longer than the grouped kernel holds");
    TNS_HIP(hipGetLastError());
}

void zgrouped_trsm(Context* ctx, const TrsmRec<zd>* recs, const int* starts, int count, int blocks)
{
    if (count <= 0 || blocks <= 0) return;
    hipLaunchKernelGGL((trsm_grouped_kernel<zd>), dim3(blocks), dim3(256), 0, ctx->stream, recs, starts, count);
    TNS_HIP(hipGetLastError());
}

namespace {
// Static LDS of getrs_grouped_kernel beside its right-hand sides: the two
// solves' partial sums (G x diagonal block each), with room to spare.
constexpr size_t GROUPED_SOLVE_STATIC = 2 * 4 * 32 * sizeof(zd) + 1024;
}  // namespace

int grouped_solve_group(Context* ctx, int nmax)
{
    for (int g : {4, 2, 1})
        if ((size_t)g * nmax * sizeof(zd) + GROUPED_SOLVE_STATIC <= (size_t)ctx->max_lds) return g;
    return 0;
}

void zgrouped_solve(Context* ctx, const SolveRec<zd>* recs, int count, int trans, int group, int nmax)
{
    if (count <= 0) return;
    const size_t bytes = (size_t)group * (nmax > 0 ? nmax : 1) * sizeof(zd);
#define TNS_GS(TRANS_, G_)                                                                                         \
    hipLaunchKernelGGL((getrs_grouped_kernel<zd, TRANS_, G_>), dim3(count), dim3(GS_NT), bytes, ctx->stream, recs)
    if (!trans) {
        if (group >= 4) TNS_GS(0, 4);
        else if (group == 2) TNS_GS(0, 2);
        else TNS_GS(0, 1);
    } else {
        if (group >= 4) TNS_GS(1, 4);
        else if (group == 2) TNS_GS(1, 2);
        else TNS_GS(1, 1);
    }
#undef TNS_GS
    TNS_HIP(hipGetLastError());
}

}  // namespace tns
