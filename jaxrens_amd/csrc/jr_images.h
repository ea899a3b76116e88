// Periodic images of a cubic cell worth testing for a pair, shared by the
// judge's neighbour lists and the surrogate's descriptor passes.
#pragma once

#include <hip/hip_runtime.h>

namespace jr {

// For a minimum-image separation d (fractional, |d_k| <= 1/2) and a cutoff
// rcf (fractional, <= 1): bit k is set when the image shifted along axis k,
// whose coordinate has magnitude 1 - |d_k|, can lie within rcf.  The 1e-5
// margin keeps every image the distance test could accept, so enumerating
// only the subsets of these axes (in ascending order) changes no result.
__device__ __forceinline__ int image_axes(const float (&d)[3], float rcf) {
  const float lim = 1.0f - rcf - 1e-5f;
  return (fabsf(d[0]) > lim ? 1 : 0) | (fabsf(d[1]) > lim ? 2 : 0) |
         (fabsf(d[2]) > lim ? 4 : 0);
}

}  // namespace jr
