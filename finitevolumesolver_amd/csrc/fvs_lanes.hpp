// Lane utilities of the wave64 kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fvs {

// Whole-wave shifts by one lane (DPP wave_shr:1 / wave_shl:1): lane i gets
// lane i - 1 (ShiftUp) or lane i + 1 (ShiftDown); the end lane keeps its own
// value, as __shfl_up / __shfl_down(x, 1) do.  A DPP move is one VALU
// instruction with no LDS round trip.  A lane whose source lane is inactive
// also keeps its own value, so a shift must run where the whole wave is
// active: never inside a divergent branch or a short-circuit operand.
__device__ __forceinline__ int DppUp(int x) {
  return __builtin_amdgcn_update_dpp(x, x, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int DppDown(int x) {
  return __builtin_amdgcn_update_dpp(x, x, 0x130, 0xf, 0xf, false);
}
template <bool UP> __device__ __forceinline__ double Shift(double x) {
  const int64_t v = __double_as_longlong(x);
  const int lo = static_cast<int>(v);
  const int hi = static_cast<int>(v >> 32);
  const int rlo = UP ? DppUp(lo) : DppDown(lo);
  const int rhi = UP ? DppUp(hi) : DppDown(hi);
  return __longlong_as_double((static_cast<int64_t>(rhi) << 32) | static_cast<uint32_t>(rlo));
}
template <bool UP> __device__ __forceinline__ float Shift(float x) {
  const int v = __float_as_int(x);
  return __int_as_float(UP ? DppUp(v) : DppDown(v));
}
template <class T> __device__ __forceinline__ T ShiftUp(T x) { return Shift<true>(x); }
template <class T> __device__ __forceinline__ T ShiftDown(T x) { return Shift<false>(x); }

} // namespace fvs
