// Strided device views shared by the view-based operators of the library.
#pragma once

#include "fvs_internal.hpp"

namespace fvs {

// Device copy of an fvs_view: element (i, j, k, c) of the box lo + [0, ext).
struct DevView {
  double* data;
  int64_t lo[3];
  int64_t ext[3];
  int64_t stride[3];
  int64_t comp_stride;
  __device__ __forceinline__ double* at(int64_t i, int64_t j, int64_t k, int c) const {
    return data + (i - lo[0]) * stride[0] + (j - lo[1]) * stride[1] +
           (k - lo[2]) * stride[2] + c * comp_stride;
  }
  __device__ __forceinline__ double* at(const int64_t* idx, int c) const {
    return at(idx[0], idx[1], idx[2], c);
  }
  __host__ __device__ __forceinline__ int64_t count() const { return ext[0] * ext[1] * ext[2]; }
  __device__ __forceinline__ void index(int64_t q, int64_t* idx) const {
    idx[0] = lo[0] + q % ext[0];
    idx[1] = lo[1] + (q / ext[0]) % ext[1];
    idx[2] = lo[2] + q / (ext[0] * ext[1]);
  }
};

// Validates v (non-null data, at least min_comp components, non-empty box)
// and fills out; sets the thread's error text on failure.
int MakeDevView(const fvs_view* v, int min_comp, DevView* out);

// True when the box [lo, hi] lies inside v's box.
bool ContainsBox(const fvs_view* v, const int64_t lo[3], const int64_t hi[3]);

} // namespace fvs
