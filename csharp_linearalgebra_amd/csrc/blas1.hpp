// Declarations of the memory-bound device helpers in blas1.hip, used by the
// host-side algorithm drivers (estimators, exponentials).
#pragma once

#include "tns_internal.hpp"

namespace tns {

// Up to 8 operands of a linear combination, passed by value to the kernel.
template <typename T>
struct LinTerms {
    const T* src[8];
    double coef[8];
};

// out := c0*I + sum coef[q]*src[q] (n x n, ld n), in the order of
// Identity followed by successive AddScaled calls.
template <typename T>
void lincomb(Context* ctx, int n, double c0, int terms, const LinTerms<T>& lt, T* out);
// den := v - u, num := v + u (count elements).
template <typename T>
void sub_add(Context* ctx, size_t count, const T* u, const T* v, T* den, T* num);
// dst := f * src, part by part.
template <typename T>
void scale_to(Context* ctx, size_t count, const T* src, double f, T* dst);
// x := x * s (element type product).
template <typename T>
void mul_scalar(Context* ctx, size_t count, T* x, T s);
template <typename T>
void shift_diag(Context* ctx, int n, T* a, int lda, T delta);
// d_out[0] := trace(A), summed in row order.
template <typename T>
void trace(Context* ctx, int n, const T* a, int lda, T* d_out);
// out := |a| element-wise (the modulus for complex), count elements.
template <typename T>
void magnitudes(Context* ctx, size_t count, const T* a, double* out);
// next := row^T * M for an n x n real M (ld n).
void rowvec(Context* ctx, int n, const double* mags, const double* row, double* next);
void fill(Context* ctx, int n, double* v, double value);
// d_out[0] := max(v) with NaN propagating.
void vec_max(Context* ctx, int n, const double* v, double* d_out);
void zero_scalars(Context* ctx, double* d, int count);

// The n-form panel product in two passes, exposed so a caller can fuse the
// second pass: partials into part (ks * t * n), then a finishing kernel.
template <typename T>
int panel_ksplit(int n);
// skip (device, may be null): when *skip != 0 the launch does nothing.
template <typename T>
void panel_n_partials(Context* ctx, int n, int t, const T* a, int lda, const T* x, int ldx, T* part, int ks,
                      const int* skip = nullptr);
// Second pass fused with the Taylor step: work = coef * sum(part),
// f += work, d_norms[0] = max(d_norms[0], ||work||_inf),
// d_norms[1] = max(d_norms[1], ||f||_inf). Panels n x nb with ld n.
template <typename T>
void taylor_step(Context* ctx, int n, int nb, int ks, const T* part, double coef, T* work, T* f, double* d_norms,
                 const int* skip = nullptr);
// The early-exit test of a Taylor stage on the device (one thread):
// d_state[0] the stage's finished flag, d_state[1] its applications;
// d_norms[0..2] = this step's ||work||, ||F||, the previous ||work||.
void taylor_exit(Context* ctx, double* d_norms, int* d_state, double unit);
// d_out[0] = max(d_out[0], ||x||_inf) for an n x nb panel with ld n.
template <typename T>
void panel_infnorm(Context* ctx, int n, int nb, const T* x, double* d_out);

}  // namespace tns
