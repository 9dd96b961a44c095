// Fourier-space trilinear projector sample, the per-bin body of
// Project3DArraytoNDKernel<3> (src/Relion/Project.cu:89-177):
//  q = M k; if q.x < -1e-5, q = -q and the result is conjugated (only the
//  x >= 0 half is stored); x0 = floor(q.x + 1e-5), y0 = floor(q.y),
//  z0 = floor(q.z) with negative y/z indices wrapped by +p; trilinear
//  interpolation with lerp(a, b, t) = fma(b - a, t, a) along x, then y, then z.
// The projector is read at the linear address ((z*p + y)*(p/2+1) + x), as in
// the reference (so x0 + 1 = p/2 + 1 reads the next row's x = 0 bin, weighted
// by a fraction that is zero up to 1e-5); the address is clamped to the array
// so that no sample can read outside it.
#pragma once
#include <hip/hip_runtime.h>

namespace gx {

struct Projector {
  const float2* __restrict__ data;  // (p/2+1) * p * p complex, FFT order
  int p;                            // side of the projector box
  int ph;                           // p/2 + 1
  int last;                         // number of bins - 1
};

__device__ __forceinline__ float2 lerp2(float2 a, float2 b, float t) {
  return make_float2(fmaf(b.x - a.x, t, a.x), fmaf(b.y - a.y, t, a.y));
}

__device__ __forceinline__ float2 proj_load(const Projector& P, int i) {
  i = i < 0 ? 0 : (i > P.last ? P.last : i);
  return P.data[i];
}

// M: row-major 3x3 (rotation^T * oversample); (x, y, z): signed integer frequency.
__device__ __forceinline__ float2 project_sample(const Projector& P, const float* M, int x, int y, int z) {
  float px = M[0] * x + M[1] * y + M[2] * z;
  float py = M[3] * x + M[4] * y + M[5] * z;
  float pz = M[6] * x + M[7] * y + M[8] * z;
  float neg = 1.f;
  if (px < -1e-5f) {
    px = fabsf(px);
    py = -py;
    pz = -pz;
    neg = -1.f;
  }
  const float fx = floorf(px + 1e-5f);
  const int x0 = (int)fx;
  const float tx = px - fx;
  const float fy = floorf(py);
  int y0 = (int)fy, y1 = y0 + 1;
  if (y0 < 0) y0 += P.p;
  if (y1 < 0) y1 += P.p;
  const float ty = py - fy;
  const float fz = floorf(pz);
  int z0 = (int)fz, z1 = z0 + 1;
  if (z0 < 0) z0 += P.p;
  if (z1 < 0) z1 += P.p;
  const float tz = pz - fz;
  const int r00 = (z0 * P.p + y0) * P.ph + x0;
  const int r01 = (z0 * P.p + y1) * P.ph + x0;
  const int r10 = (z1 * P.p + y0) * P.ph + x0;
  const int r11 = (z1 * P.p + y1) * P.ph + x0;
  const float2 d000 = proj_load(P, r00), d001 = proj_load(P, r00 + 1);
  const float2 d010 = proj_load(P, r01), d011 = proj_load(P, r01 + 1);
  const float2 d100 = proj_load(P, r10), d101 = proj_load(P, r10 + 1);
  const float2 d110 = proj_load(P, r11), d111 = proj_load(P, r11 + 1);
  const float2 dx00 = lerp2(d000, d001, tx), dx01 = lerp2(d010, d011, tx);
  const float2 dx10 = lerp2(d100, d101, tx), dx11 = lerp2(d110, d111, tx);
  const float2 dxy0 = lerp2(dx00, dx01, ty), dxy1 = lerp2(dx10, dx11, ty);
  float2 val = lerp2(dxy0, dxy1, tz);
  val.y *= neg;
  return val;
}

// Signed frequency of FFT-order index i along an axis of length n
// (src/Relion/Project.cu:107-109).
__device__ __forceinline__ int signed_freq(int i, int n) { return i <= n / 2 ? i : i - n; }

}  // namespace gx
