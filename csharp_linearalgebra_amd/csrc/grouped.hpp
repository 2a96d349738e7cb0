// Grouped launches: one kernel launch carries the same step of many
// independent problems of different sizes. Each problem's part of the step
// is a record in device memory, filled in once when the schedule is built;
// the kernel finds its record from the workgroup id (one workgroup per
// record, or a flat range of workgroups per record located by binary search
// over the records' `start`).
//
// The records say exactly what the per-problem launch would have been told
// (pointers, sizes, offsets), so a grouped step computes what the
// per-problem step computes, element for element.
#pragma once

#include "tns_internal.hpp"

namespace tns {

// Column 0 of a matrix's 1-norm, the identity permutation, the zero-pivot
// sentinel: the start of an LU (getrf_impl) for one matrix.
template <typename T>
struct InitRec {
    const T* a;
    int* perm;
    int* info;
    double* anorm;
    int lda, n;
};

// One register-resident panel factorization (getf2_reg): m x w at a.
template <typename T>
struct PanelRec {
    T* a;
    int* piv;
    int* perm;
    int* list;
    int* info;
    int lda, m, w, off;
};

// A panel's move list applied to columns [0, c0) and [c1, ncols) of the rows
// starting at a (column 0). Workgroups [start, start + blocks).
template <typename T>
struct MovesRec {
    T* a;
    const int* list;
    int lda, c0, c1, ncols, start;
};

// B := L^-1 B, L the w x w unit lower triangle at l (w <= the diagonal
// block), B w x ncols.
template <typename T>
struct TrsmRec {
    const T* l;
    T* b;
    int ldl, ldb, w, ncols, start;
};

// C := C - A B, A m x k, B k x n.
template <typename T>
struct GemmRec {
    const T* a;
    const T* b;
    T* c;
    int lda, ldb, ldc, m, n, k, start;
};

// y := A^-1 x (trans 0) or A^-H x (trans 1) from the LU factors and the row
// permutation, nrhs columns; x and y may be the same. A non-null info is
// finished (no zero pivot: -1) on the way.
template <typename T>
struct SolveRec {
    const T* lu;
    const int* perm;
    const T* x;
    T* y;
    int* info;
    int ldlu, ldx, ldy, n, nrhs;
};

// Launches (lu.hip, gemm.hip). `recs` is device memory; `count` records;
// `blocks` the total workgroups (GEMM: tiles) of a flat launch, and
// `starts` (device, count ints) each record's first one.
void zgrouped_init(Context* ctx, const InitRec<zd>* recs, int count);
// The register panel kernel for height m: its column-chunk width (0 if no
// register kernel holds m rows) and its variant (records of one launch share
// it). max_m, the tallest record, sizes the workgroup.
int zpanel_chunk_width(int m);
int zpanel_class(int m);
void zgrouped_panel(Context* ctx, const PanelRec<zd>* recs, int count, int max_m);
int grouped_moves_columns(int max_moves);  // columns per workgroup
void zgrouped_moves(Context* ctx, const MovesRec<zd>* recs, const int* starts, int count, int blocks, int max_moves);
constexpr int GROUPED_TRSM_COLUMNS = 32;
void zgrouped_trsm(Context* ctx, const TrsmRec<zd>* recs, const int* starts, int count, int blocks);
// Tile of grouped GEMM configuration cfg (1: 64 x 64, 2: 64 x 32).
void grouped_gemm_tile(int cfg, int& bm, int& bn);
void zgrouped_gemm(Context* ctx, const GemmRec<zd>* recs, const int* starts, int count, int blocks, int cfg);
// Right-hand-side columns held in LDS at once for systems up to order nmax
// (0: too large for one workgroup's LDS).
int grouped_solve_group(Context* ctx, int nmax);
void zgrouped_solve(Context* ctx, const SolveRec<zd>* recs, int count, int trans, int group, int nmax);

}  // namespace tns
