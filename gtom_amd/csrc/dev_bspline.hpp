// Cubic B-spline weight, bspline() of include/CubicInterp.cuh:119-130.
#pragma once
#include <hip/hip_runtime.h>

namespace gx {

__device__ __forceinline__ float bspline(float t) {
  t = fabsf(t);
  const float a = 2.0f - t;
  if (t < 1.0f) return 2.0f / 3.0f - 0.5f * t * t * a;
  if (t < 2.0f) return a * a * a / 6.0f;
  return 0.0f;
}

}  // namespace gx
