// Host-compiler stand-in for the HIP qualifiers and the gfx950 math
// builtins, so that the device headers' arithmetic can be compiled and
// checked on the CPU by the tests.  Not part of the product build.
#pragma once
#include <cmath>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
using std::copysign;
using std::fabs;
using std::fmax;
using std::fmin;
using std::sqrt;
inline float __builtin_amdgcn_rcpf(float x) { return 1.0f / x; }
inline float __builtin_amdgcn_sqrtf(float x) { return std::sqrt(x); }
inline double __builtin_amdgcn_rcp(double x) { return 1.0 / x; }
inline double __builtin_amdgcn_rsq(double x) { return 1.0 / std::sqrt(x); }
