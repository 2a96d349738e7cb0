// Element arithmetic shared by the device kernels, overloaded for double and
// for complex (zd), with the semantics of Tensile's element seam
// (ElementKernels.cs:125-170 for double, ComplexKernels.cs:110-140 for
// System.Numerics.Complex). The library is compiled without floating-point
// contraction, so a*b+c here is two roundings, as in .NET; fused forms are
// written out where they are wanted.
#pragma once

#include <hip/hip_runtime.h>

namespace tns {

// Complex element of the device kernels: a plain aggregate (re, im), 16-byte
// aligned, the layout of System.Numerics.Complex. Plain so that arrays of it
// held per thread stay in registers.
struct alignas(16) zd {
    double x, y;
};
__host__ __device__ inline zd make_zd(double x, double y) { return zd{x, y}; }

// float: the working precision of the mixed-precision factorisation.
__host__ __device__ inline float s_zero(float) { return 0.0f; }
__host__ __device__ inline float s_one(float) { return 1.0f; }
__host__ __device__ inline float s_add(float a, float b) { return a + b; }
__host__ __device__ inline float s_sub(float a, float b) { return a - b; }
__host__ __device__ inline float s_mul(float a, float b) { return a * b; }
__host__ __device__ inline float s_neg(float a) { return -a; }
__host__ __device__ inline float s_conj(float a) { return a; }
__host__ __device__ inline bool s_iszero(float a) { return a == 0.0f; }
__host__ __device__ inline float s_div(float a, float b) { return a / b; }
__host__ __device__ inline double s_abs(float a) { return fabsf(a); }
__host__ __device__ inline double s_pivmag(float a) { return fabsf(a); }
__host__ __device__ inline double s_pivmax(float a) { return fabsf(a); }
__host__ __device__ inline float s_sel(bool c, float a, float b) { return c ? a : b; }

__host__ __device__ inline double s_zero(double) { return 0.0; }
__host__ __device__ inline zd s_zero(zd) { return make_zd(0.0, 0.0); }
__host__ __device__ inline double s_one(double) { return 1.0; }
__host__ __device__ inline zd s_one(zd) { return make_zd(1.0, 0.0); }

__host__ __device__ inline double s_add(double a, double b) { return a + b; }
__host__ __device__ inline zd s_add(zd a, zd b) { return make_zd(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double s_sub(double a, double b) { return a - b; }
__host__ __device__ inline zd s_sub(zd a, zd b) { return make_zd(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double s_mul(double a, double b) { return a * b; }
__host__ __device__ inline zd s_mul(zd a, zd b)
{
    return make_zd(a.x * b.x - a.y * b.y, a.y * b.x + a.x * b.y);
}
__host__ __device__ inline double s_neg(double a) { return -a; }
__host__ __device__ inline zd s_neg(zd a) { return make_zd(-a.x, -a.y); }
__host__ __device__ inline double s_conj(double a) { return a; }
__host__ __device__ inline zd s_conj(zd a) { return make_zd(a.x, -a.y); }
__host__ __device__ inline bool s_iszero(double a) { return a == 0.0; }
__host__ __device__ inline bool s_iszero(zd a) { return a.x == 0.0 && a.y == 0.0; }
__host__ __device__ inline float s_nan(float) { return __builtin_nanf(""); }
__host__ __device__ inline double s_nan(double) { return __builtin_nan(""); }
__host__ __device__ inline zd s_nan(zd) { return make_zd(__builtin_nan(""), __builtin_nan("")); }
__host__ __device__ inline double s_scale(double a, double f) { return a * f; }
__host__ __device__ inline zd s_scale(zd a, double f) { return make_zd(a.x * f, a.y * f); }
__host__ __device__ inline double s_re(double a) { return a; }
__host__ __device__ inline double s_re(zd a) { return a.x; }

__host__ __device__ inline double s_div(double a, double b) { return a / b; }
// Smith's method (CACM 5(8), 1962): scale by the larger part of the divisor.
__host__ __device__ inline zd s_div(zd x, zd y)
{
    double a = x.x, b = x.y, c = y.x, d = y.y;
    if (fabs(d) < fabs(c)) {
        double doc = d / c;
        double den = c + d * doc;
        return make_zd((a + b * doc) / den, (b - a * doc) / den);
    }
    double cod = c / d;
    double den = d + c * cod;
    return make_zd((b + a * cod) / den, (-a + b * cod) / den);
}

// Modulus: |x| for real; for complex the larger part factored out, so the
// square cannot overflow.
__host__ __device__ inline double s_abs(double a) { return fabs(a); }
__host__ __device__ inline double s_abs(zd z)
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
__host__ __device__ inline double s_pivmag(zd a) { return fabs(a.x) + fabs(a.y); }
// max(|Re|, |Im|), a lower bound of the modulus; NaN if either part is.
__host__ __device__ inline double s_pivmax(double a) { return fabs(a); }
__host__ __device__ inline double s_pivmax(zd a)
{
    return (a.x != a.x || a.y != a.y) ? a.x + a.y : fmax(fabs(a.x), fabs(a.y));
}

// Both parts finite.
__host__ __device__ inline bool s_isfinite(double a) { return __builtin_isfinite(a); }
__host__ __device__ inline bool s_isfinite(zd a) { return __builtin_isfinite(a.x) && __builtin_isfinite(a.y); }
__host__ __device__ inline bool s_isfinite(float a) { return __builtin_isfinite(a); }

// c ? a : b, component by component (a conditional expression over
// struct lvalues would select an address; see load_if).
__host__ __device__ inline double s_sel(bool c, double a, double b) { return c ? a : b; }
__host__ __device__ inline zd s_sel(bool c, zd a, zd b) { return make_zd(c ? a.x : b.x, c ? a.y : b.y); }

// p as a pointer into global memory, by type: a pointer read from a record
// in memory (the grouped kernels' arguments) is otherwise generic, and its
// loads become flat loads, which the compiler must also count as LDS
// operations (every wait on one waits on both).
// (The host pass compiles these bodies too, where the pointer stays as it is.)
template <typename T>
__device__ __forceinline__ auto as_global(T* p)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (T __attribute__((address_space(1)))*)p;
#else
    return p;
#endif
}

// *p if ok, else zero. A conditional expression between two lvalues of a
// struct type selects an address, and an address of a local puts the local
// in scratch memory; returning by value keeps everything in registers.
template <typename T>
__device__ __forceinline__ T load_if(const T* p, bool ok)
{
    if (ok) return *p;
    return s_zero(T{});
}


// The value lane ^ OFF holds, as __shfl_xor(v, OFF, 64) gives it. OFF 1 and
// 2: a DPP quad permutation; 4 and 8: DPP row shifts both ways inside a row
// of 16 lanes (row_shl gives lane L the value of L + OFF, row_shr that of
// L - OFF; a lane takes the one whose bit OFF it lacks or has); 16 and 32
// cross rows and go through the LDS crossbar. The DPP forms run on the
// VALU, so a workgroup's lane reductions do not queue on the CU's one LDS
// unit.
template <int OFF>
__device__ __forceinline__ unsigned xor_lane_u32(unsigned v)
{
    static_assert(OFF == 1 || OFF == 2 || OFF == 4 || OFF == 8, "inside a row of 16 lanes");
    if constexpr (OFF == 1) {
        return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
    } else if constexpr (OFF == 2) {
        return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
    } else {
        const unsigned up = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x100 + OFF, 0xF, 0xF, false);
        const unsigned dn = (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x110 + OFF, 0xF, 0xF, false);
        return (__lane_id() & OFF) ? dn : up;
    }
}
template <int OFF>
__device__ __forceinline__ double xor_lane(double v)
{
    if constexpr (OFF >= 16) {
        return __shfl_xor(v, OFF, 64);
    } else {
        const unsigned long long b = (unsigned long long)__double_as_longlong(v);
        const unsigned lo = xor_lane_u32<OFF>((unsigned)b), hi = xor_lane_u32<OFF>((unsigned)(b >> 32));
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    }
}
template <int OFF>
__device__ __forceinline__ zd xor_lane(zd v)
{
    return make_zd(xor_lane<OFF>(v.x), xor_lane<OFF>(v.y));
}
// Sum over the wavefront's 64 lanes by the xor tree (offsets 32, 16, ...,
// 1): every lane ends with the same sum, in the same order as
// v += __shfl_xor(v, off) over those offsets.
__device__ __forceinline__ double wave_xor_sum(double v)
{
    v += xor_lane<32>(v);
    v += xor_lane<16>(v);
    v += xor_lane<8>(v);
    v += xor_lane<4>(v);
    v += xor_lane<2>(v);
    v += xor_lane<1>(v);
    return v;
}

}  // namespace tns
