// Host stand-in for the few HIP names fvs_physics.hpp uses, so that its
// device arithmetic can be compiled and checked by a host compiler (CPU
// tests only; the hardware-estimate builtins map to exact host arithmetic).
#pragma once
#include <cmath>
#include <cstring>
#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __builtin_amdgcn_rcp(x) (1.0 / (x))
#define __builtin_amdgcn_rsq(x) (1.0 / std::sqrt(x))
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#define __builtin_amdgcn_sqrtf(x) std::sqrt(x)
inline long long __double_as_longlong(double x) {
  long long r;
  std::memcpy(&r, &x, 8);
  return r;
}
inline double __longlong_as_double(long long x) {
  double r;
  std::memcpy(&r, &x, 8);
  return r;
}
