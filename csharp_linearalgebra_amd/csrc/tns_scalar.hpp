// Element arithmetic shared by the device kernels, overloaded for double and
// for complex (double2), with the semantics of Tensile's element seam
// (ElementKernels.cs:125-170 for double, ComplexKernels.cs:110-140 for
// System.Numerics.Complex). The library is compiled without floating-point
// contraction, so a*b+c here is two roundings, as in .NET; fused forms are
// written out where they are wanted.
#pragma once

#include <hip/hip_runtime.h>

namespace tns {

__host__ __device__ inline double s_zero(double) { return 0.0; }
__host__ __device__ inline double2 s_zero(double2) { return make_double2(0.0, 0.0); }
__host__ __device__ inline double s_one(double) { return 1.0; }
__host__ __device__ inline double2 s_one(double2) { return make_double2(1.0, 0.0); }

__host__ __device__ inline double s_add(double a, double b) { return a + b; }
__host__ __device__ inline double2 s_add(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double s_sub(double a, double b) { return a - b; }
__host__ __device__ inline double2 s_sub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double s_mul(double a, double b) { return a * b; }
__host__ __device__ inline double2 s_mul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y);
}
__host__ __device__ inline double s_neg(double a) { return -a; }
__host__ __device__ inline double2 s_neg(double2 a) { return make_double2(-a.x, -a.y); }
__host__ __device__ inline double s_conj(double a) { return a; }
__host__ __device__ inline double2 s_conj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline bool s_iszero(double a) { return a == 0.0; }
__host__ __device__ inline bool s_iszero(double2 a) { return a.x == 0.0 && a.y == 0.0; }
__host__ __device__ inline double s_scale(double a, double f) { return a * f; }
__host__ __device__ inline double2 s_scale(double2 a, double f) { return make_double2(a.x * f, a.y * f); }
__host__ __device__ inline double s_re(double a) { return a; }
__host__ __device__ inline double s_re(double2 a) { return a.x; }

__host__ __device__ inline double s_div(double a, double b) { return a / b; }
// Smith's method (CACM 5(8), 1962): scale by the larger part of the divisor.
__host__ __device__ inline double2 s_div(double2 x, double2 y)
{
    double a = x.x, b = x.y, c = y.x, d = y.y;
    if (fabs(d) < fabs(c)) {
        double doc = d / c;
        double den = c + d * doc;
        return make_double2((a + b * doc) / den, (b - a * doc) / den);
    }
    double cod = c / d;
    double den = d + c * cod;
    return make_double2((b + a * cod) / den, (-a + b * cod) / den);
}

// Modulus: |x| for real; for complex the larger part factored out, so the
// square cannot overflow.
__host__ __device__ inline double s_abs(double a) { return fabs(a); }
__host__ __device__ inline double s_abs(double2 z)
{
    double a = fabs(z.x), b = fabs(z.y);
    double small = a < b ? a : b, large = a < b ? b : a;
    if (small == 0.0) return large;
    if (__builtin_isinf(large) && !__builtin_isnan(small)) return __builtin_inf();
    double r = small / large;
    return large * sqrt(1.0 + r * r);
}

// The pivot measure: |x|, or |Re| + |Im| as LAPACK's izamax (cabs1).
__host__ __device__ inline double s_pivmag(double a) { return fabs(a); }
__host__ __device__ inline double s_pivmag(double2 a) { return fabs(a.x) + fabs(a.y); }

}  // namespace tns
