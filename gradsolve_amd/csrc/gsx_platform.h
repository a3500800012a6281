// gsx_platform.h -- the three ways the kernel sources are compiled, in one place:
//   * hipcc, for the library (host and gfx950 device passes);
//   * hiprtc, for fields registered at run time (gsx_runtime.hip): device only, with no C++
//     standard library headers; its built-in runtime header supplies the fixed-width integer
//     types (in __hip_internal) and the math functions (global namespace), which are brought
//     into std here so that the shared sources can call std::fma and the like;
//   * a plain C++ compiler, for the CPU test harness of the per-lane code.
#pragma once

#if defined(__HIPCC_RTC__)
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::int16_t int16_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
namespace std {
using ::atan;
using ::cos;
using ::cosh;
using ::exp;
using ::exp2;
using ::expm1;
using ::fabs;
using ::fma;
using ::fmax;
using ::fmin;
using ::log;
using ::log1p;
using ::log2;
using ::pow;
using ::rint;
using ::sin;
using ::sinh;
using ::sqrt;
using ::tan;
using ::tanh;
}  // namespace std
#ifndef INT32_MAX
#define INT32_MAX 2147483647
#endif
#else
#include <cmath>
#include <cstdint>
#endif

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>
#endif
#define GSX_HD __host__ __device__ __forceinline__
#define GSX_UNROLL _Pragma("unroll")
#else
#define GSX_HD inline
#define GSX_UNROLL
#endif

// Finite test usable in all three compilations (hiprtc's std has no isfinite).
GSX_HD bool gsx_finite(double x) { return __builtin_isfinite(x); }
