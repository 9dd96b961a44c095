// Fourier-space trilinear projector sample, the per-bin body of
// Project3DArraytoNDKernel<3> (src/Relion/Project.cu:89-177):
//  q = M k; if q.x < -1e-5, q = -q and the result is conjugated (only the
//  x >= 0 half is stored); x0 = floor(q.x + 1e-5), y0 = floor(q.y),
//  z0 = floor(q.z) with negative y/z indices wrapped by +p; trilinear
//  interpolation with lerp(a, b, t) = fma(b - a, t, a) along x, then y, then z.
// The projector is read at the linear address ((z*p + y)*(p/2+1) + x), as in
// the reference, so x0 + 1 = p/2 + 1 reads the next row's x = 0 bin with a
// weight that is zero up to 1e-5. Reads go through a buffer resource whose
// hardware range check returns 0 for any address outside the projector, so no
// sample can read outside it (the reference does not check).
#pragma once
#include <hip/hip_runtime.h>

namespace gx {

// dword 3 of a raw (untyped, stride 0) buffer resource on gfx9-family GPUs:
// 32-bit data format, so range checking is done in bytes against num_records.
constexpr int BUFFER_RSRC_WORD3 = 0x00020000;

struct Projector {
  const float2* __restrict__ data;  // (p/2+1) * p * p complex, FFT order
  int p;                            // side of the projector box
  int ph;                           // p/2 + 1
  int bytes;                        // size of the projector in bytes (< 2^31)
};

__device__ __forceinline__ __amdgpu_buffer_rsrc_t projector_rsrc(const Projector& P) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(P.data), (short)0, P.bytes, BUFFER_RSRC_WORD3);
}

// Raw buffer resource over `bytes` bytes at p (bytes < 2^31). Accesses use a
// 32-bit per-lane offset plus a wave-uniform offset, both in bytes.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, BUFFER_RSRC_WORD3);
}
__device__ __forceinline__ float2 buf_load2(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
  float2 f;
  __builtin_memcpy(&f, &v, 8);
  return f;
}
__device__ __forceinline__ float buf_load1(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  auto v = __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0);
  float f;
  __builtin_memcpy(&f, &v, 4);
  return f;
}
__device__ __forceinline__ void buf_store2(float2 f, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __attribute__((ext_vector_type(2))) unsigned v;
  __builtin_memcpy(&v, &f, 8);
  __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, 0);
}

__device__ __forceinline__ float2 lerp2(float2 a, float2 b, float t) {
  return make_float2(fmaf(b.x - a.x, t, a.x), fmaf(b.y - a.y, t, a.y));
}

// Two adjacent bins (x0, x0 + 1) at byte offset `off`; 0 outside the projector.
__device__ __forceinline__ void proj_pair(__amdgpu_buffer_rsrc_t rs, int off, float2& a, float2& b) {
  auto va = __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0);
  auto vb = __builtin_amdgcn_raw_buffer_load_b64(rs, off + 8, 0, 0);
  __builtin_memcpy(&a, &va, 8);
  __builtin_memcpy(&b, &vb, 8);
}

// Byte offset beyond any projector (< 2^31 bytes): a load there returns 0
// without touching memory.
constexpr int OUT_OF_RANGE = 0x7ffffff0;

// Sample at q = (px, py, pz) = M k (already rotated and scaled); 0 when
// `inside` is false. Branch-free, so that the loads of several samples can be
// in flight together.
__device__ __forceinline__ float2 project_at(__amdgpu_buffer_rsrc_t rs, int p, int ph, float px, float py, float pz,
                                             bool inside = true) {
  const float sgn = px < -1e-5f ? -1.f : 1.f;
  px *= sgn;
  py *= sgn;
  pz *= sgn;
  const float fx = floorf(px + 1e-5f);
  const float fy = floorf(py);
  const float fz = floorf(pz);
  const float tx = px - fx, ty = py - fy, tz = pz - fz;
  const int x0 = (int)fx;
  int y0 = (int)fy, z0 = (int)fz;
  int y1 = y0 + 1, z1 = z0 + 1;
  if (y0 < 0) y0 += p;
  if (y1 < 0) y1 += p;
  if (z0 < 0) z0 += p;
  if (z1 < 0) z1 += p;
  const int r00 = __mul24(z0, p) + y0, r01 = __mul24(z0, p) + y1;
  const int r10 = __mul24(z1, p) + y0, r11 = __mul24(z1, p) + y1;
  const int o00 = inside ? (__mul24(r00, ph) + x0) * 8 : OUT_OF_RANGE;
  const int o01 = inside ? (__mul24(r01, ph) + x0) * 8 : OUT_OF_RANGE;
  const int o10 = inside ? (__mul24(r10, ph) + x0) * 8 : OUT_OF_RANGE;
  const int o11 = inside ? (__mul24(r11, ph) + x0) * 8 : OUT_OF_RANGE;
  float2 d000, d001, d010, d011, d100, d101, d110, d111;
  proj_pair(rs, o00, d000, d001);
  proj_pair(rs, o01, d010, d011);
  proj_pair(rs, o10, d100, d101);
  proj_pair(rs, o11, d110, d111);
  const float2 dx00 = lerp2(d000, d001, tx), dx01 = lerp2(d010, d011, tx);
  const float2 dx10 = lerp2(d100, d101, tx), dx11 = lerp2(d110, d111, tx);
  const float2 dxy0 = lerp2(dx00, dx01, ty), dxy1 = lerp2(dx10, dx11, ty);
  float2 val = lerp2(dxy0, dxy1, tz);
  val.y *= sgn;
  return val;
}

// The projector re-laid for the search kernels, so that the sampler needs no
// wrap logic and one address per sample: planes and rows in centred order
// with a one-voxel margin (row c holds the FFT-order row (c - OFF) mod p,
// OFF = p/2 + 1, c in [0, p + 3)), and one extra column x = p/2 + 1 holding
// the element that follows the row at the reference's linear address (x = 0
// of the next FFT-order row, or 0 past the end of the array). For every q
// with |q| < p/2 + 1/2 this reads exactly what project_at reads. Built by
// k_centre_projector in the search plan.
struct CentredProjector {
  const float2* __restrict__ data;
  int bytes;     // PX * PY * PY * 8
  int row8;      // PX * 8: byte step of y
  int plane8;    // PX * PY * 8: byte step of z
  float base8;   // (OFF * PY + OFF) * PX * 8: byte offset of (x, y, z) = (0, 0, 0)
};

__device__ __forceinline__ void proj_quad(__amdgpu_buffer_rsrc_t rs, int off, int soff, float2& a, float2& b) {
  auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, soff, 0);
  __builtin_memcpy(&a, &v, 8);
  __builtin_memcpy(&b, reinterpret_cast<const char*>(&v) + 8, 8);
}

// project_at on the centred layout (same flip, floors, lerp order and result).
// The byte offset of the (x0, y0, z0) corner is computed in float, exact since
// every term is an integer below 2^24.
__device__ __forceinline__ float2 project_centred(__amdgpu_buffer_rsrc_t rs, const CentredProjector& C, float px,
                                                  float py, float pz, bool inside = true) {
  const float sgn = px < -1e-5f ? -1.f : 1.f;
  px *= sgn;
  py *= sgn;
  pz *= sgn;
  const float fx = floorf(px + 1e-5f);
  const float fy = floorf(py);
  const float fz = floorf(pz);
  const float tx = px - fx, ty = py - fy, tz = pz - fz;
  const float fo = fmaf(fz, (float)C.plane8, fmaf(fy, (float)C.row8, fmaf(fx, 8.f, C.base8)));
  const int o = inside ? (int)fo : OUT_OF_RANGE;
  float2 d000, d001, d010, d011, d100, d101, d110, d111;
  proj_quad(rs, o, 0, d000, d001);
  proj_quad(rs, o, C.row8, d010, d011);
  proj_quad(rs, o, C.plane8, d100, d101);
  proj_quad(rs, o, C.plane8 + C.row8, d110, d111);
  const float2 dx00 = lerp2(d000, d001, tx), dx01 = lerp2(d010, d011, tx);
  const float2 dx10 = lerp2(d100, d101, tx), dx11 = lerp2(d110, d111, tx);
  const float2 dxy0 = lerp2(dx00, dx01, ty), dxy1 = lerp2(dx10, dx11, ty);
  float2 val = lerp2(dxy0, dxy1, tz);
  val.y *= sgn;
  return val;
}

// M: row-major 3x3 (rotation^T * oversample); (x, y, z): signed integer frequency.
__device__ __forceinline__ float2 project_sample(__amdgpu_buffer_rsrc_t rs, int p, int ph, const float* M, int x,
                                                 int y, int z) {
  const float px = M[0] * x + M[1] * y + M[2] * z;
  const float py = M[3] * x + M[4] * y + M[5] * z;
  const float pz = M[6] * x + M[7] * y + M[8] * z;
  return project_at(rs, p, ph, px, py, pz);
}

// Signed frequency of FFT-order index i along an axis of length n
// (src/Relion/Project.cu:107-109).
__device__ __forceinline__ int signed_freq(int i, int n) { return i <= n / 2 ? i : i - n; }

}  // namespace gx
