// d_Rotate3DFT with linear interpolation (src/Transformation/Rotation.cu:263-374,
// Rotate3DFTKernel :541-654): half spectra of cubes rotated by Euler angles.
//
// Input: the zero-centred half layout of a cube of side n, element (tx, ty, tz)
// of (n/2+1, n, n) holding frequency (tx - n/2, ty - n/2, tz - n/2), tx <= n/2
// (the non-positive x half). Output bin (idx, idy, idz) of (n/2+1, n, n) for
// orientation b:
//   (x, y, z) = (idx, idy, idz) when the output is zero-centred, else
//   (n/2 - idx, n-1 - ((idy + n/2 - 1) % n), n-1 - ((idz + n/2 - 1) % n));
//   p = transpose(Euler(angles_b)) (x - n/2, y - n/2, z - n/2);
//   0 when |p|^2 >= (n*n/4); p = -p (and the imaginary part negated) when
//   p.x > 1e-6; the value is the trilinear interpolation at texel p + n/2 with
//   texel indices clamped to the array.
// The reference reads textures with hardware filtering, whose weights CUDA
// stores with 8 fractional bits, through cudaAddressModeWrap on unnormalised
// coordinates, which CUDA documents as unsupported there (clamp applies).
// Here the weights are exact float and the indices clamped. The position is
// evaluated in float as fma(M0, x, fma(M1, y, M2 z)) and |p|^2 as fma(px, px,
// fma(py, py, pz pz)), a fixed order, so that the radius cut and the flip, which
// the reference decides in float, are reproducible.
//
// One thread per output bin; the 8 corners are gathered from the (L2-resident)
// input, interleaved complex or real. d_Rotate2DFT (:200-256) below is the
// same scheme on images.
#include <cmath>
#include <vector>

#include "runtime.hpp"

namespace gx {
namespace {

template <bool COMPLEX>
__global__ __launch_bounds__(256) void k_rotate3d_ft(const float* __restrict__ in, float* __restrict__ out, int n,
                                                     const float* __restrict__ mats, int zerocentred) {
  using V = typename std::conditional<COMPLEX, float2, float>::type;
  const V* src = reinterpret_cast<const V*>(in);
  V* dst = reinterpret_cast<V*>(out);
  const int nh = n / 2 + 1;
  const long long per = (long long)nh * n * n;
  const int b = blockIdx.y;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= per) return;
  const int idx = (int)(i % nh), idy = (int)((i / nh) % n), idz = (int)(i / ((long long)nh * n));
  int x = idx, y = idy, z = idz;
  if (!zerocentred) {
    x = n / 2 - idx;
    y = n - 1 - ((idy + n / 2 - 1) % n);
    z = n - 1 - ((idz + n / 2 - 1) % n);
  }
  const float* M = mats + 9 * b;
  const float cx = (float)(x - n / 2), cy = (float)(y - n / 2), cz = (float)(z - n / 2);
  // float evaluation in a fixed order (the radius cut and the flip are decided here)
  float px = fmaf(M[0], cx, fmaf(M[1], cy, M[2] * cz));
  float py = fmaf(M[3], cx, fmaf(M[4], cy, M[5] * cz));
  float pz = fmaf(M[6], cx, fmaf(M[7], cy, M[8] * cz));
  V* o = dst + (size_t)b * per + i;
  if (fmaf(px, px, fmaf(py, py, pz * pz)) >= (float)(n * n / 4)) {
    if constexpr (COMPLEX) *o = make_float2(0.f, 0.f);
    else *o = 0.f;
    return;
  }
  const bool neg = px > 1e-6f;
  if (neg) {
    px = -px;
    py = -py;
    pz = -pz;
  }
  const float c = (float)(n / 2);
  px += c;
  py += c;
  pz += c;
  const float fx0 = floorf(px), fy0 = floorf(py), fz0 = floorf(pz);
  const float ax = px - fx0, ay = py - fy0, az = pz - fz0;
  auto clampi = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  const int x0 = clampi((int)fx0, nh - 1), x1 = clampi((int)fx0 + 1, nh - 1);
  const int y0 = clampi((int)fy0, n - 1), y1 = clampi((int)fy0 + 1, n - 1);
  const int z0 = clampi((int)fz0, n - 1), z1 = clampi((int)fz0 + 1, n - 1);
  auto at = [&](int zz, int yy, int xx) { return src[((size_t)zz * n + yy) * nh + xx]; };
  if constexpr (COMPLEX) {
    auto lerp2 = [](float2 a, float2 b, float t) { return make_float2(fmaf(t, b.x - a.x, a.x), fmaf(t, b.y - a.y, a.y)); };
    const float2 c00 = lerp2(at(z0, y0, x0), at(z0, y0, x1), ax), c01 = lerp2(at(z0, y1, x0), at(z0, y1, x1), ax);
    const float2 c10 = lerp2(at(z1, y0, x0), at(z1, y0, x1), ax), c11 = lerp2(at(z1, y1, x0), at(z1, y1, x1), ax);
    float2 v = lerp2(lerp2(c00, c01, ay), lerp2(c10, c11, ay), az);
    if (neg) v.y = -v.y;
    *o = v;
  } else {
    auto lerp = [](float a, float b, float t) { return fmaf(t, b - a, a); };
    const float c00 = lerp(at(z0, y0, x0), at(z0, y0, x1), ax), c01 = lerp(at(z0, y1, x0), at(z0, y1, x1), ax);
    const float c10 = lerp(at(z1, y0, x0), at(z1, y0, x1), ax), c11 = lerp(at(z1, y1, x0), at(z1, y1, x1), ax);
    *o = lerp(lerp(c00, c01, ay), lerp(c10, c11, ay), az);
  }
}

// Rotate2DFTKernel (src/Transformation/Rotation.cu:485-539) with linear
// interpolation: input the zero-centred half layout of an nx x ny image
// ((nx/2+1) x ny, element (tx, ty) = frequency (tx - nx/2, ty - ny/2)); the
// position comes from the thread's own (idx, idy), pos = R(-angle) (idx - nx/2,
// idy - ny/2), 0 when |pos| > maxfreq, flipped to the Hermitian mate when pos.x
// > 1e-5, and the value is written at (idx, idy) or, for FFT-order output, at
// (nx/2 - idx, ny-1 - ((idy + ny/2 - 1) % ny)). mats holds (c, s) per item with
// c = cos(-angle), s = sin(-angle) rounded to float, as glm::mat2(c, s, -s, c).
// in_stride 0 rotates one input by every angle.
__global__ __launch_bounds__(256) void k_rotate2d_ft(const float2* __restrict__ in, size_t in_stride,
                                                     float2* __restrict__ out, int nx, int ny,
                                                     const float2* __restrict__ cs, float maxfreq, int zerocentred) {
  const int nh = nx / 2 + 1;
  const int per = nh * ny;
  const int b = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= per) return;
  const int idx = i % nh, idy = i / nh;
  int x = idx, y = idy;
  if (!zerocentred) {
    x = nx / 2 - idx;
    y = ny - 1 - ((idy + ny / 2 - 1) % ny);
  }
  const float2 r = cs[b];
  const float vx = (float)(idx - nx / 2), vy = (float)(idy - ny / 2);
  // float evaluation in a fixed order (the radius cut and the flip are decided here)
  float px = fmaf(r.x, vx, -(r.y * vy));
  float py = fmaf(r.y, vx, r.x * vy);
  float2* o = out + (size_t)b * per + (size_t)y * nh + x;
  if (sqrtf(fmaf(px, px, py * py)) > maxfreq) {
    *o = make_float2(0.f, 0.f);
    return;
  }
  const bool neg = px > 1e-5f;
  if (neg) {
    px = -px;
    py = -py;
  }
  px += (float)(nx / 2);
  py += (float)(ny / 2);
  const float fx0 = floorf(px), fy0 = floorf(py);
  const float ax = px - fx0, ay = py - fy0;
  auto clampi = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  const int x0 = clampi((int)fx0, nh - 1), x1 = clampi((int)fx0 + 1, nh - 1);
  const int y0 = clampi((int)fy0, ny - 1), y1 = clampi((int)fy0 + 1, ny - 1);
  const float2* src = in + (size_t)b * in_stride;
  auto lerp2 = [](float2 a, float2 c, float t) { return make_float2(fmaf(t, c.x - a.x, a.x), fmaf(t, c.y - a.y, a.y)); };
  float2 v = lerp2(lerp2(src[(size_t)y0 * nh + x0], src[(size_t)y0 * nh + x1], ax),
                   lerp2(src[(size_t)y1 * nh + x0], src[(size_t)y1 * nh + x1], ax), ay);
  if (neg) v.y = -v.y;
  *o = v;
}

}  // namespace

int rotate3d_ft_batch(const float2* in, float2* out, int n, const float* d_mats, unsigned batch, bool zerocentred,
                      hipStream_t s) {
  const long long per = (long long)(n / 2 + 1) * n * n;
  const dim3 grid((unsigned)((per + 255) / 256), batch);
  hipLaunchKernelGGL((k_rotate3d_ft<true>), grid, dim3(256), 0, s, reinterpret_cast<const float*>(in),
                     reinterpret_cast<float*>(out), n, d_mats, zerocentred ? 1 : 0);
  GX_LAUNCHED("k_rotate3d_ft");
  return GTOMX_OK;
}

int rotate2d_ft_batch(const float2* in, size_t in_stride, float2* out, int nx, int ny, const float2* d_cs,
                      unsigned batch, float maxfreq, bool zerocentred, hipStream_t s) {
  const int per = (nx / 2 + 1) * ny;
  hipLaunchKernelGGL(k_rotate2d_ft, dim3((unsigned)((per + 255) / 256), batch), dim3(256), 0, s, in, in_stride, out,
                     nx, ny, d_cs, maxfreq, zerocentred ? 1 : 0);
  GX_LAUNCHED("k_rotate2d_ft");
  return GTOMX_OK;
}

void rotate2d_coefficients(float angle, float2* cs) {
  const double c = std::cos((double)-angle), sn = std::sin((double)-angle);
  *cs = make_float2((float)c, (float)sn);
}

namespace {

int rotate_launch(const float* in, float* out, gtomx_int3 dims, const gtomx_float3* h_angles, unsigned nangles,
                  int zerocentred, bool complex_values, void* stream) {
  if (!in || !out || (nangles > 0 && !h_angles)) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (dims.x < 2 || dims.y != dims.x || dims.z != dims.x)
    return fail(GTOMX_ERR_UNSUPPORTED, "cubes of side >= 2 only (the reference offsets every axis by dims.x/2)");
  if (nangles == 0) return ok();
  if (nangles > 65535) return fail(GTOMX_ERR_UNSUPPORTED, "more than 65535 orientations per call");
  const long long per = (long long)(dims.x / 2 + 1) * dims.y * dims.z;
  if (per >= (1ll << 31)) return fail(GTOMX_ERR_UNSUPPORTED, "volume too large");
  std::vector<float> h((size_t)nangles * 9);
  GX_TRY(gtomx_projector_matrices(h_angles, nangles, 1.0f, h.data()));
  hipStream_t s = as_stream(stream);
  DevBuf dm;
  GX_TRY(dm.alloc(h.size() * sizeof(float)));
  GX_HIP(hipMemcpyAsync(dm.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, s));
  const dim3 grid((unsigned)((per + 255) / 256), nangles);
  if (complex_values)
    hipLaunchKernelGGL((k_rotate3d_ft<true>), grid, dim3(256), 0, s, in, out, dims.x, dm.as<float>(), zerocentred);
  else
    hipLaunchKernelGGL((k_rotate3d_ft<false>), grid, dim3(256), 0, s, in, out, dims.x, dm.as<float>(), zerocentred);
  GX_LAUNCHED("k_rotate3d_ft");
  GX_HIP(hipStreamSynchronize(s));
  return ok();
}

int rotate2d_entry(const gtomx_complex* d_input, gtomx_complex* d_output, gtomx_int3 dims, const float* h_angles,
                   float maxfreq, int outputzerocentered, unsigned batch, void* stream) {
  if (!d_input || !d_output || (batch > 0 && !h_angles)) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (dims.x < 2 || dims.y < 2 || dims.z != 1) return fail(GTOMX_ERR_UNSUPPORTED, "2D images (dims.z == 1) of side >= 2");
  if (batch == 0) return ok();
  if (batch > 65535) return fail(GTOMX_ERR_UNSUPPORTED, "more than 65535 images per call");
  if ((long long)(dims.x / 2 + 1) * dims.y >= (1ll << 31)) return fail(GTOMX_ERR_UNSUPPORTED, "image too large");
  std::vector<float2> h(batch);
  for (unsigned b = 0; b < batch; ++b) rotate2d_coefficients(h_angles[b], &h[b]);
  hipStream_t s = as_stream(stream);
  DevBuf dm;
  GX_TRY(dm.alloc(h.size() * sizeof(float2)));
  GX_HIP(hipMemcpyAsync(dm.p, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice, s));
  const size_t per = elements_ft(dims);
  GX_TRY(rotate2d_ft_batch(reinterpret_cast<const float2*>(d_input), per, reinterpret_cast<float2*>(d_output), dims.x,
                           dims.y, dm.as<float2>(), batch, maxfreq, outputzerocentered != 0, s));
  GX_HIP(hipStreamSynchronize(s));
  return ok();
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" int gtomx_rotate2d_ft(const gtomx_complex* d_input, gtomx_complex* d_output, gtomx_int3 dims,
                                 const float* h_angles, float maxfreq, int outputzerocentered, unsigned batch,
                                 void* stream) {
  return rotate2d_entry(d_input, d_output, dims, h_angles, maxfreq, outputzerocentered, batch, stream);
}

extern "C" int gtomx_rotate3d_ft(const gtomx_complex* d_volume, gtomx_complex* d_output, gtomx_int3 dims,
                                 const gtomx_float3* h_angles, unsigned nangles, int outputzerocentered,
                                 void* stream) {
  return rotate_launch(reinterpret_cast<const float*>(d_volume), reinterpret_cast<float*>(d_output), dims, h_angles,
                       nangles, outputzerocentered ? 1 : 0, true, stream);
}

extern "C" int gtomx_rotate3d_ft_real(const float* d_volume, float* d_output, gtomx_int3 dims,
                                      const gtomx_float3* h_angles, unsigned nangles, int outputzerocentered,
                                      void* stream) {
  return rotate_launch(d_volume, d_output, dims, h_angles, nangles, outputzerocentered ? 1 : 0, false, stream);
}
