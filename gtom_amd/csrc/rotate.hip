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

#include "dev_bspline.hpp"
#include "runtime.hpp"

namespace gx {
namespace {

// ConvertToInterpolationCoefficients (src/CubicSplines/internal/cubicPrefilter_kernel.cu):
// the cubic B-spline prefilter of one line in place, pole z = sqrt(3) - 2, gain
// (1 - z)(1 - 1/z), causal start c0 + sum_{n < min(12, len)} z^(n+1) c_n, anticausal
// start z/(z - 1) c_(len-1), in float; one thread per (line, channel) along
// `axis` (SamplesToCoefficients3D{X,Y,Z} and the 2D forms).
__global__ __launch_bounds__(256) void k_bspline_lines(float* __restrict__ v, int3 d, int axis, unsigned batch,
                                                       int ch) {
  const size_t V = (size_t)d.x * d.y * d.z;
  const int len = axis == 0 ? d.x : (axis == 1 ? d.y : d.z);
  const size_t stride = (axis == 0 ? 1 : (axis == 1 ? (size_t)d.x : (size_t)d.x * d.y)) * ch;
  const size_t per = V / len;
  const size_t id = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (id >= per * batch * ch) return;
  const int c0 = (int)(id % ch);
  const size_t line = id / ch;
  const size_t b = line / per, l = line % per;
  size_t start;
  if (axis == 0) start = l * d.x;
  else if (axis == 1) start = (l / d.x) * ((size_t)d.x * d.y) + l % d.x;
  else start = l;
  float* c = v + (b * V + start) * ch + c0;
  const float z = sqrtf(3.0f) - 2.0f;
  const float lambda = (1.0f - z) * (1.0f - 1.0f / z);
  const int horizon = len < 12 ? len : 12;
  float zn = z, sum = c[0];
  for (int n = 0; n < horizon; ++n) {
    sum += zn * c[(size_t)n * stride];
    zn *= z;
  }
  float prev = lambda * sum;
  c[0] = prev;
  for (int n = 1; n < len; ++n) {
    prev = lambda * c[(size_t)n * stride] + z * prev;
    c[(size_t)n * stride] = prev;
  }
  prev = (z / (z - 1.0f)) * prev;
  c[(size_t)(len - 1) * stride] = prev;
  for (int n = len - 2; n >= 0; --n) {
    prev = z * (prev - c[(size_t)n * stride]);
    c[(size_t)n * stride] = prev;
  }
}

// Cubic B-spline of coefficients src (dims nxs x ny x nz, V values per voxel)
// at texel coordinate (tx, ty, tz) = p + n/2: taps floor(t) - 1 .. floor(t) + 2
// per axis, clamped (cubicTex3D/2D, include/CubicInterp.cuh, with exact weights).
template <class V>
__device__ __forceinline__ V cubic_at(const V* __restrict__ src, int nxs, int ny, int nz, float tx, float ty,
                                      float tz) {
  auto clampi = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  const float fx0 = floorf(tx), fy0 = floorf(ty), fz0 = floorf(tz);
  const float ax = tx - fx0, ay = ty - fy0, az = tz - fz0;
  V r{};
  const bool three = nz > 1;
  for (int kz = three ? -1 : 0; kz <= (three ? 2 : 0); ++kz) {
    const float wz = three ? bspline((float)kz - az) : 1.0f;
    const int zz = three ? clampi((int)fz0 + kz, nz - 1) : 0;
    for (int ky = -1; ky <= 2; ++ky) {
      const float wyz = bspline((float)ky - ay) * wz;
      const V* row = src + ((size_t)zz * ny + clampi((int)fy0 + ky, ny - 1)) * nxs;
      for (int kx = -1; kx <= 2; ++kx) {
        const float w = bspline((float)kx - ax) * wyz;
        const V t = row[clampi((int)fx0 + kx, nxs - 1)];
        if constexpr (std::is_same<V, float2>::value) {
          r.x = fmaf(w, t.x, r.x);
          r.y = fmaf(w, t.y, r.y);
        } else {
          r = fmaf(w, t, r);
        }
      }
    }
  }
  return r;
}

template <bool COMPLEX, bool CUBIC = false>
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
  if constexpr (CUBIC) {
    V v = cubic_at<V>(src, nh, n, n, px, py, pz);
    if constexpr (COMPLEX) {
      if (neg) v.y = -v.y;
    }
    *o = v;
    return;
  }
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
template <bool CUBIC = false>
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
  if constexpr (CUBIC) {
    float2 v = cubic_at<float2>(in + (size_t)b * in_stride, nh, ny, 1, px, py, 0.f);
    if (neg) v.y = -v.y;
    *o = v;
    return;
  }
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
  hipLaunchKernelGGL(k_rotate2d_ft<false>, dim3((unsigned)((per + 255) / 256), batch), dim3(256), 0, s, in, in_stride,
                     out, nx, ny, d_cs, maxfreq, zerocentred ? 1 : 0);
  GX_LAUNCHED("k_rotate2d_ft");
  return GTOMX_OK;
}

void rotate2d_coefficients(float angle, float2* cs) {
  const double c = std::cos((double)-angle), sn = std::sin((double)-angle);
  *cs = make_float2((float)c, (float)sn);
}

int bspline_prefilter(float* v, gtomx_int3 d, int ndims, unsigned batch, int channels, hipStream_t s) {
  const size_t V = elements(d);
  for (int axis = 0; axis < ndims; ++axis) {
    const size_t len = axis == 0 ? d.x : (axis == 1 ? d.y : d.z);
    const size_t n = V / len * batch * channels;
    hipLaunchKernelGGL(k_bspline_lines, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, v,
                       make_int3(d.x, d.y, d.z), axis, batch, channels);
    GX_LAUNCHED("k_bspline_lines");
  }
  return GTOMX_OK;
}

namespace {

int check_mode(int mode) {
  if (mode != GTOMX_INTERP_LINEAR && mode != GTOMX_INTERP_CUBIC)
    return fail(GTOMX_ERR_UNSUPPORTED, "interpolation is GTOMX_INTERP_LINEAR or GTOMX_INTERP_CUBIC");
  return GTOMX_OK;
}

// For cubic interpolation the input is copied and prefiltered (the reference
// prefilters its own Re and Im copies, src/Transformation/Rotation.cu:274-279).
int rotate_launch(const float* in, float* out, gtomx_int3 dims, const gtomx_float3* h_angles, unsigned nangles,
                  int mode, int zerocentred, bool complex_values, void* stream) {
  if (!in || !out || (nangles > 0 && !h_angles)) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  GX_TRY(check_mode(mode));
  if (dims.x < 2 || dims.y != dims.x || dims.z != dims.x)
    return fail(GTOMX_ERR_UNSUPPORTED, "cubes of side >= 2 only (the reference offsets every axis by dims.x/2)");
  if (nangles == 0) return ok();
  if (nangles > 65535) return fail(GTOMX_ERR_UNSUPPORTED, "more than 65535 orientations per call");
  const long long per = (long long)(dims.x / 2 + 1) * dims.y * dims.z;
  if (per >= (1ll << 31)) return fail(GTOMX_ERR_UNSUPPORTED, "volume too large");
  std::vector<float> h((size_t)nangles * 9);
  GX_TRY(gtomx_projector_matrices(h_angles, nangles, 1.0f, h.data()));
  hipStream_t s = as_stream(stream);
  DevBuf dm, coef;
  GX_TRY(dm.alloc(h.size() * sizeof(float)));
  GX_HIP(hipMemcpyAsync(dm.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, s));
  const bool cubic = mode == GTOMX_INTERP_CUBIC;
  const int ch = complex_values ? 2 : 1;
  const float* src = in;
  if (cubic) {
    GX_TRY(coef.alloc((size_t)per * ch * sizeof(float)));
    GX_HIP(hipMemcpyAsync(coef.p, in, (size_t)per * ch * sizeof(float), hipMemcpyDeviceToDevice, s));
    GX_TRY(bspline_prefilter(coef.as<float>(), gtomx_int3{dims.x / 2 + 1, dims.y, dims.z}, 3, 1, ch, s));
    src = coef.as<float>();
  }
  const dim3 grid((unsigned)((per + 255) / 256), nangles);
  if (complex_values && cubic)
    hipLaunchKernelGGL((k_rotate3d_ft<true, true>), grid, dim3(256), 0, s, src, out, dims.x, dm.as<float>(),
                       zerocentred);
  else if (complex_values)
    hipLaunchKernelGGL((k_rotate3d_ft<true>), grid, dim3(256), 0, s, src, out, dims.x, dm.as<float>(), zerocentred);
  else if (cubic)
    hipLaunchKernelGGL((k_rotate3d_ft<false, true>), grid, dim3(256), 0, s, src, out, dims.x, dm.as<float>(),
                       zerocentred);
  else
    hipLaunchKernelGGL((k_rotate3d_ft<false>), grid, dim3(256), 0, s, src, out, dims.x, dm.as<float>(), zerocentred);
  GX_LAUNCHED("k_rotate3d_ft");
  GX_HIP(hipStreamSynchronize(s));
  return ok();
}

int rotate2d_entry(const gtomx_complex* d_input, gtomx_complex* d_output, gtomx_int3 dims, const float* h_angles,
                   float maxfreq, int mode, int outputzerocentered, unsigned batch, void* stream) {
  if (!d_input || !d_output || (batch > 0 && !h_angles)) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  GX_TRY(check_mode(mode));
  if (dims.x < 2 || dims.y < 2 || dims.z != 1) return fail(GTOMX_ERR_UNSUPPORTED, "2D images (dims.z == 1) of side >= 2");
  if (batch == 0) return ok();
  if (batch > 65535) return fail(GTOMX_ERR_UNSUPPORTED, "more than 65535 images per call");
  if ((long long)(dims.x / 2 + 1) * dims.y >= (1ll << 31)) return fail(GTOMX_ERR_UNSUPPORTED, "image too large");
  std::vector<float2> h(batch);
  for (unsigned b = 0; b < batch; ++b) rotate2d_coefficients(h_angles[b], &h[b]);
  hipStream_t s = as_stream(stream);
  DevBuf dm, coef;
  GX_TRY(dm.alloc(h.size() * sizeof(float2)));
  GX_HIP(hipMemcpyAsync(dm.p, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice, s));
  const size_t per = elements_ft(dims);
  const float2* src = reinterpret_cast<const float2*>(d_input);
  if (mode == GTOMX_INTERP_CUBIC) {
    GX_TRY(coef.alloc(per * batch * sizeof(float2)));
    GX_HIP(hipMemcpyAsync(coef.p, d_input, per * batch * sizeof(float2), hipMemcpyDeviceToDevice, s));
    GX_TRY(bspline_prefilter(coef.as<float>(), gtomx_int3{dims.x / 2 + 1, dims.y, 1}, 2, batch, 2, s));
    src = coef.as<float2>();
    const int nh = dims.x / 2 + 1;
    hipLaunchKernelGGL(k_rotate2d_ft<true>, dim3((unsigned)((nh * dims.y + 255) / 256), batch), dim3(256), 0, s, src,
                       per, reinterpret_cast<float2*>(d_output), dims.x, dims.y, dm.as<float2>(), maxfreq,
                       outputzerocentered != 0 ? 1 : 0);
    GX_LAUNCHED("k_rotate2d_ft");
  } else {
    GX_TRY(rotate2d_ft_batch(src, per, reinterpret_cast<float2*>(d_output), dims.x, dims.y, dm.as<float2>(), batch,
                             maxfreq, outputzerocentered != 0, s));
  }
  GX_HIP(hipStreamSynchronize(s));
  return ok();
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" int gtomx_rotate2d_ft(const gtomx_complex* d_input, gtomx_complex* d_output, gtomx_int3 dims,
                                 const float* h_angles, float maxfreq, int mode, int outputzerocentered,
                                 unsigned batch, void* stream) {
  return rotate2d_entry(d_input, d_output, dims, h_angles, maxfreq, mode, outputzerocentered, batch, stream);
}

extern "C" int gtomx_rotate3d_ft(const gtomx_complex* d_volume, gtomx_complex* d_output, gtomx_int3 dims,
                                 const gtomx_float3* h_angles, unsigned nangles, int mode, int outputzerocentered,
                                 void* stream) {
  return rotate_launch(reinterpret_cast<const float*>(d_volume), reinterpret_cast<float*>(d_output), dims, h_angles,
                       nangles, mode, outputzerocentered ? 1 : 0, true, stream);
}

extern "C" int gtomx_rotate3d_ft_real(const float* d_volume, float* d_output, gtomx_int3 dims,
                                      const gtomx_float3* h_angles, unsigned nangles, int mode,
                                      int outputzerocentered, void* stream) {
  return rotate_launch(d_volume, d_output, dims, h_angles, nangles, mode, outputzerocentered ? 1 : 0, false, stream);
}
