// Projector, CTF and element-wise pieces of the template-matching path, plus
// the local-standard-deviation chain (d_LocalStd).
#include <cmath>
#include <vector>

#include "dev_fft.hpp"
#include "dev_project.hpp"
#include "runtime.hpp"

namespace gx {

namespace {

constexpr int TPB = 256;

inline unsigned grid_for(size_t n, unsigned cap = 1u << 20) {
  size_t g = (n + TPB - 1) / TPB;
  if (g < 1) g = 1;
  return (unsigned)std::min<size_t>(g, cap);
}

// ---- element-wise kernels (grid-stride loops) ----
__global__ void k_fill_f32(float* d, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] = v;
}
__global__ void k_fill_i32(int32_t* d, size_t n, int32_t v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] = v;
}
__global__ void k_square(const float* in, float* out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float v = in[i];
    out[i] = v * v;
  }
}
__global__ void k_mul_conj_inplace(float2* a, const float2* b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    a[i] = cmulc(a[i], b[i]);
}
// ComplexMultiplyByVectorKernel (src/Generics/Arithmetics.cu:275-286)
__global__ void k_cmul_real_vec(const float2* in, const float* v, float2* out, size_t n, unsigned batch) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float m = v[i];
    for (unsigned b = 0; b < batch; ++b) out[b * n + i] = cscale(in[b * n + i], m);
  }
}
// ComplexMultiplyByConjVectorKernel (src/Generics/Arithmetics.cu:334-344)
__global__ void k_cmul_conj_vec(const float2* in, const float2* v, float2* out, size_t n, unsigned batch) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float2 m = v[i];
    for (unsigned b = 0; b < batch; ++b) out[b * n + i] = cmulc(in[b * n + i], m);
  }
}
// DivideSafeByVectorKernel (src/Generics/Arithmetics.cu:478-493)
__global__ void k_divide_safe(const float* in, const float* div, float* out, size_t n, unsigned batch) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float d = div[i];
    bool okd = fabsf(d) > 1e-15f;
    for (unsigned b = 0; b < batch; ++b) out[b * n + i] = okd ? in[b * n + i] / d : 0.f;
  }
}
// UpdateCorrelationKernel (src/Correlation/SubTomograms.cu:138-158), int32 index
__global__ void k_update_max_argmax(const float* corr, size_t n, unsigned batch, int offset, float* best,
                                    int32_t* idx) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    float b = best[i];
    int32_t a = idx[i];
    for (unsigned k = 0; k < batch; ++k) {
      float c = corr[k * n + i];
      if (c > b) {
        b = c;
        a = (int32_t)k + offset;
      }
    }
    best[i] = b;
    idx[i] = a;
  }
}

// PadValueKernel (src/Generics/Padding.cu:97-...): centred copy with fill value.
__global__ void k_pad_value(const float* in, float* out, int3 id, int3 od, int3 off, float value, unsigned batch) {
  size_t total = (size_t)od.x * od.y * od.z;
  size_t itotal = (size_t)id.x * id.y * id.z;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total * batch; i += (size_t)gridDim.x * blockDim.x) {
    size_t b = i / total, r = i % total;
    int x = (int)(r % od.x), y = (int)((r / od.x) % od.y), z = (int)(r / ((size_t)od.x * od.y));
    int ix = x + off.x, iy = y + off.y, iz = z + off.z;
    bool in_range = ix >= 0 && ix < id.x && iy >= 0 && iy < id.y && iz >= 0 && iz < id.z;
    out[i] = in_range ? in[b * itotal + ((size_t)iz * id.y + iy) * id.x + ix] : value;
  }
}

// RemapFull2FullFFTKernel (src/FFT/FFTRemap.cu:227-240): out[ifftshift(r)] = in[r].
__global__ void k_remap_full2fullfft(const float* in, float* out, int3 d, unsigned batch) {
  size_t total = (size_t)d.x * d.y * d.z;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total * batch; i += (size_t)gridDim.x * blockDim.x) {
    size_t b = i / total, r = i % total;
    int x = (int)(r % d.x), y = (int)((r / d.x) % d.y), z = (int)(r / ((size_t)d.x * d.y));
    int rx = (x + (d.x + 1) / 2) % d.x, ry = (y + (d.y + 1) / 2) % d.y, rz = (z + (d.z + 1) / 2) % d.z;
    out[b * total + ((size_t)rz * d.y + ry) * d.x + rx] = in[i];
  }
}

// Soft sphere around (d/2) (SphereMaskKernel, src/Masking/SphereMask.cu:49-106,
// decentered = false), evaluated in double.
__device__ __forceinline__ float sphere_value(int x, int y, int z, int3 d, double radius, double sigma) {
  double xx = x - d.x / 2, yy = y - d.y / 2, zz = d.z > 1 ? z - d.z / 2 : 0;
  double len = sqrt(xx * xx + yy * yy + zz * zz);
  if (len < radius) return 1.f;
  if (sigma > 0) {
    double v = (cos(fmin(1.0, (len - radius) / sigma) * 3.14159265358979323846) + 1.0) * 0.5;
    return (float)fmax(0.0, v);
  }
  return 0.f;
}

__global__ void k_sphere_mask(const float* in, float* out, int3 d, double radius, double sigma, unsigned batch) {
  size_t total = (size_t)d.x * d.y * d.z;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total * batch; i += (size_t)gridDim.x * blockDim.x) {
    size_t r = i % total;
    int x = (int)(r % d.x), y = (int)((r / d.x) % d.y), z = (int)(r / ((size_t)d.x * d.y));
    out[i] = in[i] * sphere_value(x, y, z, d, radius, sigma);
  }
}

// SphereMaskKernel in full (src/Masking/SphereMask.cu:49-106): per axis the
// distance is the FFT-order signed index (i < n/2+1 ? i : i - n) when decentered,
// else (int)(i - c) with the float centre c (C truncation of the difference);
// z contributes only when n.z > 1. Real or interleaved complex data.
template <bool COMPLEX>
__global__ void k_sphere_mask_ex(const float* in, float* out, int3 d, double radius, double sigma, float3 c,
                                 int decentered, unsigned batch) {
  const size_t total = (size_t)d.x * d.y * d.z;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total * batch; i += (size_t)gridDim.x * blockDim.x) {
    const size_t r = i % total;
    const int x = (int)(r % d.x), y = (int)((r / d.x) % d.y), z = (int)(r / ((size_t)d.x * d.y));
    auto axis = [&](int v, int n, float cv) { return decentered ? (v < n / 2 + 1 ? v : v - n) : (int)((float)v - cv); };
    const long long xx = axis(x, d.x, c.x), yy = axis(y, d.y, c.y), zz = d.z > 1 ? axis(z, d.z, c.z) : 0;
    const double len = sqrt((double)(xx * xx + yy * yy + zz * zz));
    double m = 1.0;
    if (!(len < radius)) m = sigma > 0 ? fmax(0.0, (cos(fmin(1.0, (len - radius) / sigma) * 3.14159265358979323846) + 1.0) * 0.5) : 0.0;
    if constexpr (COMPLEX) {
      out[2 * i] = in[2 * i] * (float)m;
      out[2 * i + 1] = in[2 * i + 1] * (float)m;
    } else {
      out[i] = in[i] * (float)m;
    }
  }
}

// The soft sphere already moved to the origin: value at index i is the mask at
// the position r with ifftshift(r) = i (d_SphereMask then d_RemapFull2FullFFT,
// src/Generics/LocalStd.cu:24-26).
__global__ void k_sphere_mask_ifftshift(float* out, int3 d, double radius, double sigma) {
  size_t total = (size_t)d.x * d.y * d.z;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    int x = (int)(i % d.x), y = (int)((i / d.x) % d.y), z = (int)(i / ((size_t)d.x * d.y));
    int sx = (x + d.x - (d.x + 1) / 2) % d.x, sy = (y + d.y - (d.y + 1) / 2) % d.y,
        sz = (z + d.z - (d.z + 1) / 2) % d.z;
    out[i] = sphere_value(sx, sy, sz, d, radius, sigma);
  }
}

// Two-stage deterministic double sum: stage 1 writes one partial per block.
__global__ void k_sum_stage1(const float* d, size_t n, double* partial) {
  __shared__ double red[TPB];
  double s = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) s += d[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}
__global__ void k_sum_stage2(const double* partial, int np, double* result) {
  __shared__ double red[TPB];
  double s = 0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = TPB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *result = red[0];
}

// out = in - sum/N, the mean taken in double.
__global__ void k_minus_mean(const float* in, const double* sum, double invN, float* out, size_t n) {
  const double mu = *sum * invN;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (float)((double)in[i] - mu);
}

// The mask spectrum times a real half-spectrum Fourier mask (d_ComplexMultiplyByVector,
// src/Generics/LocalStd.cu:38-41), and the mask sum that the reference takes as
// the sum of the renormalised C2R of that product (:43-52): that sum is the
// real part of the product's DC bin, the old sum times fouriermask[0].
__global__ void k_mask_fourier(float2* mft, const float* fm, size_t n, double* masksum) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float f = fm[i];
    mft[i] = make_float2(mft[i].x * f, mft[i].y * f);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *masksum *= (double)fm[0];
}

// sigma = sqrt(max(0, M*conv2 - conv1^2)) / M with conv = C2R_u(...)/N, in double
// (src/Generics/LocalStd.cu:98-109); optionally fused with DivideSafeByVector
// of the best correlation (src/Correlation/SubTomograms.cu:117).
// The convolutions are of the map minus its global mean mu = mapsum / n, which
// leaves M*conv2 - conv1^2 unchanged for any mask and keeps it from cancelling
// in float when the map's offset is large against its local spread; the mean
// gets mu back.
__global__ void k_local_std_combine(const float* c1, const float* c2, const double* masksum, size_t n,
                                    double invN, const double* mapsum, float* std_out, float* mean_out, float* best,
                                    const int32_t* idx, float* angle_out) {
  const double M = *masksum;
  const double mu = mapsum ? *mapsum * invN : 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double a = (double)c1[i] * invN, b = (double)c2[i] * invN;
    double v = fmax(M * b - a * a, 0.0);
    float sd = (float)(sqrt(v) / M);
    if (std_out) std_out[i] = sd;
    if (mean_out) mean_out[i] = (float)(a / M + mu);
    if (best) best[i] = fabsf(sd) > 1e-15f ? best[i] / sd : 0.f;
    if (angle_out) angle_out[i] = (float)idx[i];
  }
}

// Standalone projector (Project3DArraytoNDKernel<3>, src/Relion/Project.cu:89-177);
// nz = 1 gives the 2D central sections of Project3DArraytoNDKernel<2> (z = 0).
__global__ void k_project(Projector P, float2* __restrict__ out, int n, int nz, int rmax2,
                          const float* __restrict__ mats, unsigned batch) {
  const int nh = n / 2 + 1;
  const size_t H = (size_t)nh * n * nz;
  const unsigned b = blockIdx.y;
  float M[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) M[k] = mats[9 * b + k];
  float2* o = out + b * H;
  const __amdgpu_buffer_rsrc_t rs = projector_rsrc(P);
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < H; id += (size_t)gridDim.x * blockDim.x) {
    int ix = (int)(id % nh), iy = (int)((id / nh) % n), iz = (int)(id / ((size_t)nh * n));
    int x = ix, y = signed_freq(iy, n), z = nz > 1 ? signed_freq(iz, n) : 0;
    if (x * x + y * y + z * z > rmax2) {
      o[id] = make_float2(0.f, 0.f);
      continue;
    }
    o[id] = project_sample(rs, P.p, P.ph, M, x, y, z);
  }
}

// 3D CTF (d_GetCTF<false,false>, include/CTF.cuh:195-223) on the half spectrum,
// frequency k = |r| / (n * pixelsize[A]) (src/CTF/Wiener.cu:99-110), in double.
struct CtfLean {
  double ny, defocus, K1, K2, K3, phaseshift, B, scale;
};
__global__ void k_ctf3(const CtfLean* __restrict__ params, int n, int nz, float* __restrict__ out) {
  const int nh = n / 2 + 1;
  const size_t H = (size_t)nh * n * nz;
  const CtfLean p = params[blockIdx.y];
  float* o = out + blockIdx.y * H;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < H; id += (size_t)gridDim.x * blockDim.x) {
    int ix = (int)(id % nh), iy = (int)((id / nh) % n), iz = (int)(id / ((size_t)nh * n));
    int y = ((iy + n / 2) % n) - n / 2, z = nz > 1 ? ((iz + n / 2) % n) - n / 2 : 0;
    double r = sqrt((double)ix * ix + (double)y * y + (double)z * z);
    double k = r * p.ny, k2 = k * k, k4 = k2 * k2;
    double gamma = p.K1 * p.defocus * k2 + p.K2 * k4 - p.phaseshift - p.K3;
    double v = -sin(gamma);
    if (p.B != 0) v *= exp(p.B * k2);
    o[id] = (float)(v * p.scale);
  }
}

// CTFParams in the units d_GetCTF works in (CTFParamsLean, include/CTF.cuh:100-134).
struct CtfFullLean {
  double px, pxdelta, pxangle, defocus, dfdelta, dfangle, K1, K2, K3, phaseshift, B, Bdelta, Bangle, scale;
};

// d_GetCTF (include/CTF.cuh:195-223) at k [cycles/pixel] and angle, evaluated in double.
__device__ double ctf_full(const CtfFullLean& p, double k, double angle, double gammacorr, bool ampsq,
                           bool ignorefirst) {
  k /= p.px + p.pxdelta * cos(2.0 * (angle - p.pxangle));
  const double k2 = k * k, k4 = k2 * k2;
  const double deltaf = p.defocus + p.dfdelta * cos(2.0 * (angle - p.dfangle));
  const double gamma = p.K1 * deltaf * k2 + p.K2 * k4 - p.phaseshift - p.K3 + gammacorr;
  double v = (ignorefirst && fabs(gamma) < 1.57079632679489661923) ? 1.0 : -sin(gamma);
  if (p.B != 0 || p.Bdelta != 0) {
    double b = p.B;
    if (p.Bdelta != 0) b += p.Bdelta * cos(2.0 * (angle - p.Bangle));
    v *= exp(b * k2);
  }
  if (ampsq) v = fabs(v);
  return v * p.scale;
}

// CTFSimulateKernel (src/CTF/Simulate.cu:117-139): addresses (k, angle).
__global__ void k_ctf_simulate(const CtfFullLean* __restrict__ params, const float2* __restrict__ addr,
                               const float* __restrict__ gcorr, float* __restrict__ out, unsigned n, int ampsq,
                               int ignorefirst) {
  const CtfFullLean p = params[blockIdx.y];
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float2 a = addr[i];
    out[(size_t)blockIdx.y * n + i] =
        (float)ctf_full(p, a.x, a.y, gcorr ? gcorr[i] : 0.0, ampsq != 0, ignorefirst != 0);
  }
}

// Half-spectrum CTF images: bin (x, y) at k = |(x, y)|/n cycles per pixel, angle atan2(y, x),
// y signed in [-n/2, n/2).
__global__ void k_ctf_2d(const CtfFullLean* __restrict__ params, int n, float* __restrict__ out) {
  const int nh = n / 2 + 1;
  const CtfFullLean p = params[blockIdx.y];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nh * n; i += gridDim.x * blockDim.x) {
    const int x = i % nh, iy = i / nh;
    const int y = ((iy + n / 2) % n) - n / 2;  // the Nyquist row as -n/2, as k_ctf3 does
    out[(size_t)blockIdx.y * nh * n + i] =
        (float)ctf_full(p, sqrt((double)x * x + (double)y * y) / n, atan2((double)y, (double)x), 0.0, false, false);
  }
}

CtfFullLean lean_of(const gtomx_ctf_params_full& q) {
  const double lambda = 12.2643247 / std::sqrt((double)q.voltage * (1.0 + (double)q.voltage * 0.978466e-6));
  CtfFullLean l;
  l.px = (double)q.pixelsize * 1e10;
  l.pxdelta = (double)q.pixeldelta * 0.5e10;
  l.pxangle = q.pixelangle;
  l.defocus = (double)q.defocus * 1e10;
  l.dfdelta = (double)q.defocusdelta * 0.5e10;
  l.dfangle = q.astigmatismangle;
  l.K1 = 3.14159265358979323846 * lambda;
  l.K2 = 1.57079632679489661923 * ((double)q.Cs * 1e10) * lambda * lambda * lambda;
  l.K3 = std::atan((double)q.amplitude / std::sqrt(1.0 - (double)q.amplitude * q.amplitude));
  l.phaseshift = q.phaseshift;
  l.B = (double)q.Bfactor * 0.25e20;
  l.Bdelta = (double)q.Bfactordelta * 0.25e20;
  l.Bangle = q.Bfactorangle;
  l.scale = q.scale;
  return l;
}

int upload_full(const gtomx_ctf_params_full* h, unsigned batch, DevBuf& dp, hipStream_t s) {
  std::vector<CtfFullLean> lean(batch);
  for (unsigned b = 0; b < batch; ++b) {
    const gtomx_ctf_params_full& q = h[b];
    if (!(q.pixelsize > 0) || !(q.voltage > 0) || !(q.amplitude > -1 && q.amplitude < 1))
      return fail(GTOMX_ERR_INVALID_ARGUMENT, "CTF parameters out of range");
    lean[b] = lean_of(q);
  }
  GX_TRY(dp.alloc(sizeof(CtfFullLean) * batch));
  GX_HIP(hipMemcpyAsync(dp.p, lean.data(), sizeof(CtfFullLean) * batch, hipMemcpyHostToDevice, s));
  return GTOMX_OK;
}

}  // namespace

int fill_f32(float* d, size_t n, float v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n)), dim3(TPB), 0, s, d, n, v);
  GX_LAUNCHED("k_fill_f32");
  return GTOMX_OK;
}
int fill_i32(int32_t* d, size_t n, int32_t v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(TPB), 0, s, d, n, v);
  GX_LAUNCHED("k_fill_i32");
  return GTOMX_OK;
}
int square_f32(const float* in, float* out, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_square, dim3(grid_for(n)), dim3(TPB), 0, s, in, out, n);
  GX_LAUNCHED("k_square");
  return GTOMX_OK;
}
int mul_conj_inplace(float2* a, const float2* b, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_mul_conj_inplace, dim3(grid_for(n)), dim3(TPB), 0, s, a, b, n);
  GX_LAUNCHED("k_mul_conj_inplace");
  return GTOMX_OK;
}
int sphere_mask_ifftshift(float* out, gtomx_int3 d, float radius, float sigma, hipStream_t s) {
  size_t n = elements(d);
  hipLaunchKernelGGL(k_sphere_mask_ifftshift, dim3(grid_for(n)), dim3(TPB), 0, s, out, make_int3(d.x, d.y, d.z),
                     (double)radius, (double)sigma);
  GX_LAUNCHED("k_sphere_mask_ifftshift");
  return GTOMX_OK;
}
// scratch must hold 257 doubles; the sum lands in scratch[256].
int sum_f64(const float* d, size_t n, double* scratch, hipStream_t s) {
  hipLaunchKernelGGL(k_sum_stage1, dim3(256), dim3(TPB), 0, s, d, n, scratch);
  GX_LAUNCHED("k_sum_stage1");
  hipLaunchKernelGGL(k_sum_stage2, dim3(1), dim3(TPB), 0, s, scratch, 256, scratch + 256);
  GX_LAUNCHED("k_sum_stage2");
  return GTOMX_OK;
}
int local_std_combine(const float* c1, const float* c2, const double* d_masksum, size_t n, float* std_out,
                      float* mean_out, float* best, const int32_t* idx, float* angle_out, hipStream_t s,
                      const double* d_mapsum) {
  hipLaunchKernelGGL(k_local_std_combine, dim3(grid_for(n)), dim3(TPB), 0, s, c1, c2, d_masksum, n,
                     1.0 / (double)n, d_mapsum, std_out, mean_out, best, idx, angle_out);
  GX_LAUNCHED("k_local_std_combine");
  return GTOMX_OK;
}

// Local standard deviation (d_LocalStd, src/Generics/LocalStd.cu:10-118).
// Scratch layout: maskft, aft, bft (H complex each), rbuf (mask, then map^2),
// c1, c2 (V floats each), then 257 doubles.

size_t local_std_scratch_bytes(gtomx_int3 d) {
  size_t H = elements_ft(d), V = elements(d);
  return 3 * H * sizeof(float2) + 3 * V * sizeof(float) + LS_DOUBLES * sizeof(double) + 256;
}

LocalStdScratch local_std_carve(void* base, gtomx_int3 d) {
  size_t H = elements_ft(d), V = elements(d);
  char* p = static_cast<char*>(base);
  LocalStdScratch sc;
  sc.maskft = reinterpret_cast<float2*>(p); p += H * sizeof(float2);
  sc.aft = reinterpret_cast<float2*>(p); p += H * sizeof(float2);
  sc.bft = reinterpret_cast<float2*>(p); p += H * sizeof(float2);
  sc.rbuf = reinterpret_cast<float*>(p); p += V * sizeof(float);
  sc.c1 = reinterpret_cast<float*>(p); p += V * sizeof(float);
  sc.c2 = reinterpret_cast<float*>(p); p += V * sizeof(float);
  size_t off = (size_t)(p - static_cast<char*>(base));
  off = (off + 255) & ~(size_t)255;
  sc.dsum = reinterpret_cast<double*>(static_cast<char*>(base) + off);
  return sc;
}

// Mask FT and mask sum for a radius (once per call), optionally under a Fourier mask.
int local_std_prepare(const LocalStdScratch& sc, gtomx_int3 d, float radius, hipStream_t s,
                      const float* fouriermask) {
  GX_TRY(sphere_mask_ifftshift(sc.rbuf, d, radius, 1.0f, s));
  GX_TRY(sum_f64(sc.rbuf, elements(d), sc.dsum, s));
  GX_TRY(fft_r2c(sc.rbuf, sc.maskft, d, 1, s));
  if (fouriermask) {
    const size_t H = elements_ft(d);
    hipLaunchKernelGGL(k_mask_fourier, dim3(grid_for(H)), dim3(TPB), 0, s, sc.maskft, fouriermask, H,
                       sc.dsum + 256);
    GX_LAUNCHED("k_mask_fourier");
  }
  return GTOMX_OK;
}

// sigma (and optional mean) of one map; with best != nullptr the best map is divided in place.
// The map's global mean (double sum) is subtracted before the convolutions.
int local_std_apply(const LocalStdScratch& sc, const float* map, gtomx_int3 d, float* std_out, float* mean_out,
                    float* best, const int32_t* idx, float* angle_out, hipStream_t s) {
  size_t V = elements(d), H = elements_ft(d);
  double* mapsum = sc.dsum + LS_MAPSUM;
  GX_TRY(sum_f64(map, V, mapsum - 256, s));
  hipLaunchKernelGGL(k_minus_mean, dim3(grid_for(V)), dim3(TPB), 0, s, map, mapsum, 1.0 / (double)V, sc.rbuf, V);
  GX_LAUNCHED("k_minus_mean");
  GX_TRY(fft_r2c(sc.rbuf, sc.aft, d, 1, s));
  GX_TRY(square_f32(sc.rbuf, sc.rbuf, V, s));
  GX_TRY(fft_r2c(sc.rbuf, sc.bft, d, 1, s));
  GX_TRY(mul_conj_inplace(sc.aft, sc.maskft, H, s));
  GX_TRY(mul_conj_inplace(sc.bft, sc.maskft, H, s));
  GX_TRY(fft_c2r(sc.aft, sc.aft, sc.c1, d, 1, 1.0f, s));
  GX_TRY(fft_c2r(sc.bft, sc.bft, sc.c2, d, 1, 1.0f, s));
  GX_TRY(local_std_combine(sc.c1, sc.c2, sc.dsum + 256, V, std_out, mean_out, best, idx, angle_out, s, mapsum));
  return GTOMX_OK;
}

int project_launch(const float2* vol, int p, float2* out, int n, unsigned rmax, const float* d_mats,
                   unsigned batch, hipStream_t s, int nz, int volz) {
  if (nz < 0) nz = n;
  if (volz < 0) volz = p;
  Projector P{vol, p, p / 2 + 1, (p / 2 + 1) * p * volz * 8};
  size_t H = (size_t)(n / 2 + 1) * n * nz;
  unsigned gx_ = (unsigned)std::min<size_t>((H + TPB - 1) / TPB, 4096);
  hipLaunchKernelGGL(k_project, dim3(gx_, batch), dim3(TPB), 0, s, P, out, n, nz, (int)(rmax * rmax), d_mats,
                     batch);
  GX_LAUNCHED("k_project");
  return GTOMX_OK;
}

}  // namespace gx

using namespace gx;

namespace {
int check_vol(gtomx_int3 d) {
  if (d.x < 1 || d.y < 1 || d.z < 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "dims must be positive");
  return GTOMX_OK;
}
}  // namespace

extern "C" {

int gtomx_pad(const float* d_input, float* d_output, gtomx_int3 id, gtomx_int3 od, float value, unsigned batch,
              void* stream) {
  GX_TRY(check_vol(id));
  GX_TRY(check_vol(od));
  if (!d_input || !d_output || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  int3 off = make_int3(id.x / 2 - od.x / 2, id.y / 2 - od.y / 2, id.z / 2 - od.z / 2);
  size_t n = elements(od) * batch;
  hipLaunchKernelGGL(k_pad_value, dim3(grid_for(n)), dim3(TPB), 0, as_stream(stream), d_input, d_output,
                     make_int3(id.x, id.y, id.z), make_int3(od.x, od.y, od.z), off, value, batch);
  GX_LAUNCHED("k_pad_value");
  return GTOMX_OK;
}

int gtomx_remap_full2fullfft(const float* d_input, float* d_output, gtomx_int3 dims, unsigned batch, void* stream) {
  GX_TRY(check_vol(dims));
  if (!d_input || !d_output || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  if (d_input == d_output) return fail(GTOMX_ERR_INVALID_ARGUMENT, "in-place remap is not supported");
  size_t n = elements(dims) * batch;
  hipLaunchKernelGGL(k_remap_full2fullfft, dim3(grid_for(n)), dim3(TPB), 0, as_stream(stream), d_input, d_output,
                     make_int3(dims.x, dims.y, dims.z), batch);
  GX_LAUNCHED("k_remap_full2fullfft");
  return GTOMX_OK;
}

int gtomx_project(const gtomx_complex* d_volumeft, gtomx_int3 dimsvolume, gtomx_complex* d_proj,
                  gtomx_int3 dimsproj, unsigned rmax, const float* d_matrices, unsigned batch, void* stream) {
  if (!d_volumeft || !d_proj || !d_matrices || batch == 0)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  if (dimsvolume.x != dimsvolume.y || dimsvolume.x != dimsvolume.z || dimsvolume.x < 2 || dimsvolume.x % 2 ||
      dimsvolume.x > 512)
    return fail(GTOMX_ERR_UNSUPPORTED, "projector volume must be an even cube of side <= 512");
  const bool is2d = dimsproj.z == 1;
  if (dimsproj.x != dimsproj.y || (!is2d && dimsproj.x != dimsproj.z) || dimsproj.x < 2 || dimsproj.x % 2 ||
      dimsproj.x > 1024)
    return fail(GTOMX_ERR_UNSUPPORTED, "projections must be even squares or cubes of side <= 1024");
  rmax = std::min<unsigned>(rmax, (unsigned)dimsproj.x / 2);  // src/Relion/Project.cu:32
  return project_launch(reinterpret_cast<const float2*>(d_volumeft), dimsvolume.x, reinterpret_cast<float2*>(d_proj),
                        dimsproj.x, rmax, d_matrices, batch, as_stream(stream), is2d ? 1 : dimsproj.x);
}

int gtomx_project_angles(const gtomx_complex* d_volumeft, gtomx_int3 dimsvolume, gtomx_complex* d_proj,
                         gtomx_int3 dimsproj, const gtomx_float3* h_angles, float supersample, unsigned batch,
                         void* stream) {
  if (!h_angles || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL angles or zero batch");
  std::vector<float> h(9 * (size_t)batch);
  GX_TRY(gtomx_projector_matrices(h_angles, batch, supersample, h.data()));
  DevBuf dm;
  GX_TRY(dm.alloc(h.size() * sizeof(float)));
  hipStream_t s = as_stream(stream);
  GX_HIP(hipMemcpyAsync(dm.p, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, s));
  GX_TRY(gtomx_project(d_volumeft, dimsvolume, d_proj, dimsproj, (unsigned)dimsproj.x / 2, dm.as<float>(), batch,
                       stream));
  GX_HIP(hipStreamSynchronize(s));
  return GTOMX_OK;
}

int gtomx_ctf_3d(const gtomx_ctf_params* h_params, unsigned batch, gtomx_int3 dims, float* d_ctf, void* stream) {
  if (!h_params || !d_ctf || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  if (dims.x != dims.y || (dims.z != 1 && dims.x != dims.z) || dims.x < 2 || dims.x % 2)
    return fail(GTOMX_ERR_UNSUPPORTED, "the CTF needs an even square or cube");
  std::vector<CtfLean> lean(batch);
  for (unsigned b = 0; b < batch; ++b) {
    const gtomx_ctf_params& q = h_params[b];
    if (!(q.pixelsize > 0) || !(q.voltage > 0) || !(q.amplitude > -1 && q.amplitude < 1))
      return fail(GTOMX_ERR_INVALID_ARGUMENT, "CTF parameters out of range");
    double lambda = 12.2643247 / std::sqrt((double)q.voltage * (1.0 + (double)q.voltage * 0.978466e-6));
    CtfLean& l = lean[b];
    l.ny = 1.0 / ((double)dims.x * ((double)q.pixelsize * 1e10));
    l.defocus = (double)q.defocus * 1e10;
    l.K1 = 3.14159265358979323846 * lambda;
    l.K2 = 1.57079632679489661923 * ((double)q.Cs * 1e10) * lambda * lambda * lambda;
    l.K3 = std::atan((double)q.amplitude / std::sqrt(1.0 - (double)q.amplitude * q.amplitude));
    l.phaseshift = q.phaseshift;
    l.B = (double)q.Bfactor * 0.25e20;
    l.scale = q.scale;
  }
  DevBuf dp;
  GX_TRY(dp.alloc(sizeof(CtfLean) * batch));
  hipStream_t s = as_stream(stream);
  GX_HIP(hipMemcpyAsync(dp.p, lean.data(), sizeof(CtfLean) * batch, hipMemcpyHostToDevice, s));
  size_t H = elements_ft(dims);
  unsigned gx_ = (unsigned)std::min<size_t>((H + TPB - 1) / TPB, 4096);
  hipLaunchKernelGGL(k_ctf3, dim3(gx_, batch), dim3(TPB), 0, s, dp.as<CtfLean>(), dims.x, dims.z, d_ctf);
  GX_LAUNCHED("k_ctf3");
  GX_HIP(hipStreamSynchronize(s));
  return GTOMX_OK;
}

int gtomx_ctf_simulate(const gtomx_ctf_params_full* h_params, const float* d_addresses, const float* d_gammacorrection,
                       float* d_output, unsigned n, int amplitudesquared, int ignorefirstpeak, unsigned batch,
                       void* stream) {
  if (!h_params || !d_addresses || !d_output || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  if (n == 0) return GTOMX_OK;
  hipStream_t s = as_stream(stream);
  DevBuf dp;
  GX_TRY(upload_full(h_params, batch, dp, s));
  const unsigned g = (unsigned)std::min<size_t>((n + TPB - 1) / TPB, 4096);
  hipLaunchKernelGGL(k_ctf_simulate, dim3(g, batch), dim3(TPB), 0, s, dp.as<CtfFullLean>(),
                     reinterpret_cast<const float2*>(d_addresses), d_gammacorrection, d_output, n, amplitudesquared,
                     ignorefirstpeak);
  GX_LAUNCHED("k_ctf_simulate");
  GX_HIP(hipStreamSynchronize(s));
  return GTOMX_OK;
}

int gtomx_ctf_tilt_params_grid(const gtomx_ctf_tilt_params* tilt, gtomx_int3 dimsimage, gtomx_int3 dimsregion,
                               const gtomx_int3* h_positions, unsigned npoints, gtomx_ctf_params_full* h_params) {
  if (!tilt || (npoints > 0 && (!h_positions || !h_params))) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  // Rotation about the in-plane axis (cos a, sin a, 0) by r, the reference's
  // Matrix3RotationCustom as written (glm column-major, m[col][row]).
  auto rotation = [](double a, double r, double m[3][3]) {
    const double ax = std::cos(a), ay = std::sin(a), c = std::cos(r), c1 = 1.0 - c, s = std::sin(r);
    m[0][0] = c + ax + ax * c1;
    m[0][1] = ay * ax * c1;
    m[0][2] = -ay * s;
    m[1][0] = ax * ay * c1;
    m[1][1] = c + ay * ay * c1;
    m[1][2] = ax * s;
    m[2][0] = ay * s;
    m[2][1] = -ax * s;
    m[2][2] = c;
  };
  double st[3][3], sp[3][3], normal[3];
  rotation(tilt->stageangle[0], tilt->stageangle[1], st);
  rotation(tilt->specimenangle[0], tilt->specimenangle[1], sp);
  for (int r = 0; r < 3; ++r) normal[r] = st[0][r] * sp[2][0] + st[1][r] * sp[2][1] + st[2][r] * sp[2][2];
  if (normal[2] == 0.0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "specimen plane is vertical");
  const double ci = std::cos(tilt->imageangle), si = std::sin(tilt->imageangle);
  const double px = tilt->centerparams.pixelsize;
  for (unsigned i = 0; i < npoints; ++i) {
    const double fx = (h_positions[i].x + dimsregion.x / 2.0 - dimsimage.x / 2.0) * px;
    const double fy = (h_positions[i].y + dimsregion.y / 2.0 - dimsimage.y / 2.0) * px;
    const double gx = ci * fx - si * fy, gy = si * fx + ci * fy;
    h_params[i] = tilt->centerparams;
    h_params[i].defocus = (float)(tilt->centerparams.defocus + (-gx * normal[0] - gy * normal[1]) / normal[2]);
  }
  return ok();
}

int gtomx_ctf_2d(const gtomx_ctf_params_full* h_params, unsigned batch, gtomx_int3 dims, float* d_ctf, void* stream) {
  if (!h_params || !d_ctf || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  if (dims.x != dims.y || dims.z != 1 || dims.x < 2 || dims.x % 2)
    return fail(GTOMX_ERR_UNSUPPORTED, "2D CTFs need an even square");
  hipStream_t s = as_stream(stream);
  DevBuf dp;
  GX_TRY(upload_full(h_params, batch, dp, s));
  const size_t H = elements_ft(dims);
  const unsigned g = (unsigned)std::min<size_t>((H + TPB - 1) / TPB, 4096);
  hipLaunchKernelGGL(k_ctf_2d, dim3(g, batch), dim3(TPB), 0, s, dp.as<CtfFullLean>(), dims.x, d_ctf);
  GX_LAUNCHED("k_ctf_2d");
  GX_HIP(hipStreamSynchronize(s));
  return GTOMX_OK;
}

int gtomx_complex_multiply_by_vector(const gtomx_complex* d_input, const float* d_mult, gtomx_complex* d_output,
                                     size_t elements_, unsigned batch, void* stream) {
  if (!d_input || !d_mult || !d_output || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  hipLaunchKernelGGL(k_cmul_real_vec, dim3(grid_for(elements_)), dim3(TPB), 0, as_stream(stream),
                     reinterpret_cast<const float2*>(d_input), d_mult, reinterpret_cast<float2*>(d_output), elements_,
                     batch);
  GX_LAUNCHED("k_cmul_real_vec");
  return GTOMX_OK;
}

int gtomx_complex_multiply_by_conj_vector(const gtomx_complex* d_input, const gtomx_complex* d_mult,
                                          gtomx_complex* d_output, size_t elements_, unsigned batch, void* stream) {
  if (!d_input || !d_mult || !d_output || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  hipLaunchKernelGGL(k_cmul_conj_vec, dim3(grid_for(elements_)), dim3(TPB), 0, as_stream(stream),
                     reinterpret_cast<const float2*>(d_input), reinterpret_cast<const float2*>(d_mult),
                     reinterpret_cast<float2*>(d_output), elements_, batch);
  GX_LAUNCHED("k_cmul_conj_vec");
  return GTOMX_OK;
}

int gtomx_update_max_argmax(const float* d_correlation, size_t vectorlength, unsigned batch, int batchoffset,
                            float* d_best, int32_t* d_bestindex, void* stream) {
  if (!d_correlation || !d_best || !d_bestindex || batch == 0)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  hipLaunchKernelGGL(k_update_max_argmax, dim3(grid_for(vectorlength)), dim3(TPB), 0, as_stream(stream),
                     d_correlation, vectorlength, batch, batchoffset, d_best, d_bestindex);
  GX_LAUNCHED("k_update_max_argmax");
  return GTOMX_OK;
}

int gtomx_divide_safe_by_vector(const float* d_input, const float* d_divisors, float* d_output, size_t elements_,
                                unsigned batch, void* stream) {
  if (!d_input || !d_divisors || !d_output || batch == 0)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  hipLaunchKernelGGL(k_divide_safe, dim3(grid_for(elements_)), dim3(TPB), 0, as_stream(stream), d_input, d_divisors,
                     d_output, elements_, batch);
  GX_LAUNCHED("k_divide_safe");
  return GTOMX_OK;
}

int gtomx_sphere_mask(const float* d_input, float* d_output, gtomx_int3 dims, float radius, float sigma,
                      unsigned batch, void* stream) {
  GX_TRY(check_vol(dims));
  if (!d_input || !d_output || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  size_t n = elements(dims) * batch;
  hipLaunchKernelGGL(k_sphere_mask, dim3(grid_for(n)), dim3(TPB), 0, as_stream(stream), d_input, d_output,
                     make_int3(dims.x, dims.y, dims.z), (double)radius, (double)sigma, batch);
  GX_LAUNCHED("k_sphere_mask");
  return GTOMX_OK;
}

int gtomx_sphere_mask_ex(const void* d_input, void* d_output, gtomx_int3 dims, float radius, float sigma,
                         const gtomx_float3* center, int decentered, int complex_values, unsigned batch,
                         void* stream) {
  GX_TRY(check_vol(dims));
  if (!d_input || !d_output || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer or zero batch");
  const float3 c = center ? make_float3(center->x, center->y, center->z)
                          : make_float3((float)(dims.x / 2), (float)(dims.y / 2), (float)(dims.z / 2));
  const size_t n = elements(dims) * batch;
  const int3 d = make_int3(dims.x, dims.y, dims.z);
  if (complex_values)
    hipLaunchKernelGGL(k_sphere_mask_ex<true>, dim3(grid_for(n)), dim3(TPB), 0, as_stream(stream),
                       static_cast<const float*>(d_input), static_cast<float*>(d_output), d, (double)radius,
                       (double)sigma, c, decentered ? 1 : 0, batch);
  else
    hipLaunchKernelGGL(k_sphere_mask_ex<false>, dim3(grid_for(n)), dim3(TPB), 0, as_stream(stream),
                       static_cast<const float*>(d_input), static_cast<float*>(d_output), d, (double)radius,
                       (double)sigma, c, decentered ? 1 : 0, batch);
  GX_LAUNCHED("k_sphere_mask_ex");
  return GTOMX_OK;
}

int gtomx_local_std(const float* d_map, gtomx_int3 dims, const float* d_fouriermask, float localradius,
                    float* d_std, float* d_mean, void* stream) {
  if (!d_map || !d_std) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  int nd = dims.z > 1 ? 3 : (dims.y > 1 ? 2 : 1);
  GX_TRY(check_fft_dims_any(nd, dims));  // its transforms take every extent up to 2048
  DevBuf sb;
  GX_TRY(sb.alloc(local_std_scratch_bytes(dims)));
  LocalStdScratch sc = local_std_carve(sb.p, dims);
  hipStream_t s = as_stream(stream);
  GX_TRY(local_std_prepare(sc, dims, localradius, s, d_fouriermask));
  GX_TRY(local_std_apply(sc, d_map, dims, d_std, d_mean, nullptr, nullptr, nullptr, s));
  GX_HIP(hipStreamSynchronize(s));
  return GTOMX_OK;
}

}  // extern "C"
