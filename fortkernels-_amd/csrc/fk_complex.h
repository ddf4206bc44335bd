// fk_complex.h -- complex arithmetic on HIP double2 / float2 (x = re, y = im).
//
// Replaces the reference's hand-rolled complex classes: GPUComplex
// (BGW-Kernels/GPUComplex.h:18-247), its device variant
// (BGW-Kernels/GPUComplex_target.cpp:18-257), the FF GPUComplex
// (FFKer_cori/Complex.h:208-311) and CustomComplex<re,im>
// (FFKer_cori/CustomComplex.h:17-214).  These are free __host__ __device__
// functions on the native vector types, with the mathematically intended
// meaning throughout (the reference's reversed operator-=, in-place unary
// minus and mutating binary +/-/'/' are not reproduced).
//
// Products expand as (ac - bd, ad + bc), the order std::complex<double>
// uses; hipcc may fuse a multiply into the following add on the device.
#pragma once

#include <hip/hip_runtime.h>

namespace fk {

template <class T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };
template <class T> using cplx = typename cvec<T>::type;

__host__ __device__ inline double2 cmake(double re, double im) { return make_double2(re, im); }
__host__ __device__ inline float2 cmake(float re, float im) { return make_float2(re, im); }

// a*b                                   GPUComplex_product (GPUComplex.h:164)
__host__ __device__ inline double2 cmul(double2 a, double2 b)
{ return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__host__ __device__ inline float2 cmul(float2 a, float2 b)
{ return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

// conj(a)                               GPUComplex_conj (GPUComplex.h:150)
__host__ __device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

// |a|^2                                 target GPUComplex_abs (GPUComplex_target.cpp:185-192)
__host__ __device__ inline double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }
__host__ __device__ inline float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }

// |a|                                   GPUComplex_abs (GPUComplex.h:176)
__host__ __device__ inline double cabs(double2 a) { return sqrt(cabs2(a)); }

// a*s                                   GPUComplex_mult(a, b) (GPUComplex.h:197)
__host__ __device__ inline double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// a*s*t                                 GPUComplex_mult(a, b, c) (GPUComplex.h:187)
__host__ __device__ inline double2 cscale(double2 a, double s, double t)
{ return make_double2(a.x * s * t, a.y * s * t); }

// acc + a*b / acc - a*b                 GPUComplex_fma / _fms (GPUComplex.h:207,220)
__host__ __device__ inline double2 cfma(double2 acc, double2 a, double2 b)
{ double2 p = cmul(a, b); return make_double2(acc.x + p.x, acc.y + p.y); }
__host__ __device__ inline double2 cfms(double2 acc, double2 a, double2 b)
{ double2 p = cmul(a, b); return make_double2(acc.x - p.x, acc.y - p.y); }

__host__ __device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
// a - b                                 GPUComplex_minus (FFKer_cori/Complex.h:307)
__host__ __device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// s - a / s + a                         doubleMinusGPUComplex / doublePlusGPUComplex (GPUComplex.h:231,236)
__host__ __device__ inline double2 real_minus(double s, double2 a) { return make_double2(s - a.x, -a.y); }
__host__ __device__ inline double2 real_plus(double s, double2 a) { return make_double2(s + a.x, a.y); }

// a*a                                   GPUComplex_square (GPUComplex.h:138)
__host__ __device__ inline double2 csquare(double2 a) { return cmul(a, a); }

// a/b = a*conj(b)/|b|^2                 GPUComplex_divide (FFKer_cori/Complex.h:208)
__host__ __device__ inline double2 cdiv(double2 a, double2 b)
{ return cscale(cmul(a, cconj(b)), 1.0 / cabs2(b)); }
// a/s                                   GPUComplex_divide(a, double) (FFKer_cori/Complex.h:255)
__host__ __device__ inline double2 cdiv(double2 a, double s) { return make_double2(a.x / s, a.y / s); }

}  // namespace fk
