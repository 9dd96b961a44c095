// Per-item reductions and normalisation of a batch:
//  * d_Sum (src/Generics/Sum.cu:99-139) for float, double and int: float and
//    double accumulated in double, int in 64 bits and truncated to 32 (the
//    reference's int sum wraps the same way);
//  * d_NormMonolithic (src/ImageManipulation/Norm.cu:76-91, NormMeanStdDevMonoKernel
//    :242-284): y = (x - mean) / std with sums of x and x^2 in double, std =
//    sqrt(N sum x^2 - (sum x)^2) / N, and y = 0 where that is not positive; the
//    mode argument of the reference is not read there, so there is none here;
//  * d_NormFTMonolithic (:147-154, NormFTMonoKernel :476-509): a complex spectrum
//    divided by max(sqrt(sum |x|^2 / N), 1e-20).
// The reference reduces each volume in one workgroup; here NB workgroups per
// volume write double partial sums, one workgroup per volume finishes them, and
// a wide grid applies the result. Mean and scale are rounded to float before
// they are applied, as the reference stores them.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include <mutex>

#include "runtime.hpp"

namespace gx {
namespace {
constexpr int NB = 256;  // reduction workgroups per volume
}  // namespace
}  // namespace gx

// Scratch of the allocation-free normalisation (gtomx_norm_monolithic_noalloc):
// the double2 partial sums of up to GTOMX_NOALLOC_MAX_BATCH items, then their
// float2 coefficients (two per double2 slot). One copy per GPU, held until the
// process ends. In host code its name is not a device address: every use goes
// through the pointer hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double2 scratch[(size_t)GTOMX_NOALLOC_MAX_BATCH * gx::NB +
                                                 GTOMX_NOALLOC_MAX_BATCH / 2];

namespace gx {
namespace {

__device__ double2 block_sum2(double2 v, double2* red) {
  for (int o = 32; o > 0; o >>= 1) {
    v.x += __shfl_down(v.x, o);
    v.y += __shfl_down(v.y, o);
  }
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  double2 s = make_double2(0, 0);
  if (threadIdx.x == 0)
    for (int w = 0; w < 4; ++w) {
      s.x += red[w].x;
      s.y += red[w].y;
    }
  __syncthreads();
  return s;
}

// part[b][p] = (sum x, sum x^2) for real data, (sum |x|^2, 0) for complex.
template <bool COMPLEX>
__global__ __launch_bounds__(256) void k_mono_partial(const float* __restrict__ x, size_t n,
                                                      double2* __restrict__ part) {
  __shared__ double2 red[4];
  const size_t b = blockIdx.y;
  double2 acc = make_double2(0, 0);
  if constexpr (COMPLEX) {
    const float2* xb = reinterpret_cast<const float2*>(x) + b * n;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
      const float2 v = xb[i];
      acc.x += (double)v.x * v.x + (double)v.y * v.y;
    }
  } else {
    const float* xb = x + b * n;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
      const double v = xb[i];
      acc.x += v;
      acc.y += v * v;
    }
  }
  const double2 t = block_sum2(acc, red);
  if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = t;
}

// coef[b] = (float mean, float 1/std) for real data, (0, float 1/rms) for complex.
template <bool COMPLEX>
__global__ __launch_bounds__(256) void k_mono_finish(const double2* __restrict__ part, int nparts, size_t n,
                                                     float2* __restrict__ coef) {
  __shared__ double2 red[4];
  const size_t b = blockIdx.x;
  double2 acc = make_double2(0, 0);
  for (int i = threadIdx.x; i < nparts; i += 256) {
    acc.x += part[b * nparts + i].x;
    acc.y += part[b * nparts + i].y;
  }
  const double2 t = block_sum2(acc, red);
  if (threadIdx.x != 0) return;
  const double N = (double)n;
  if constexpr (COMPLEX) {
    const float rms = fmaxf((float)sqrt(t.x / N), 1e-20f);
    coef[b] = make_float2(0.f, 1.0f / rms);
  } else {
    const double sd = sqrt(N * t.y - t.x * t.x) / N;
    coef[b] = make_float2((float)(t.x / N), (float)(sd > 0.0 ? 1.0 / sd : 0.0));
  }
}

template <bool COMPLEX>
__global__ __launch_bounds__(256) void k_mono_apply(const float* __restrict__ x, float* __restrict__ y, size_t n,
                                                    const float2* __restrict__ coef) {
  const size_t b = blockIdx.y;
  const float2 c = coef[b];
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    if constexpr (COMPLEX) {
      const float2 v = reinterpret_cast<const float2*>(x)[b * n + i];
      reinterpret_cast<float2*>(y)[b * n + i] = make_float2(v.x * c.y, v.y * c.y);
    } else {
      y[b * n + i] = (x[b * n + i] - c.x) * c.y;
    }
  }
}

template <class T> struct SumAcc { using type = double; };
template <> struct SumAcc<int32_t> { using type = long long; };

// part[b][p]: the sum of this workgroup's share of item b.
template <class T>
__global__ __launch_bounds__(256) void k_sum_partial(const T* __restrict__ x, size_t n,
                                                     typename SumAcc<T>::type* __restrict__ part) {
  using A = typename SumAcc<T>::type;
  __shared__ A red[4];
  const size_t b = blockIdx.y;
  A acc = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) acc += (A)x[b * n + i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

template <class T>
__global__ __launch_bounds__(64) void k_sum_finish(const typename SumAcc<T>::type* __restrict__ part, int nparts,
                                                   T* __restrict__ out) {
  using A = typename SumAcc<T>::type;
  const size_t b = blockIdx.x;
  A acc = 0;
  for (int i = threadIdx.x; i < nparts; i += 64) acc += part[b * nparts + i];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
  if (threadIdx.x == 0) out[b] = (T)acc;
}

template <class T> int sum(const T* in, T* out, size_t n, unsigned batch, void* stream) {
  using A = typename SumAcc<T>::type;
  if (!in || !out) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch must be positive");
  if (batch > 65535) return fail(GTOMX_ERR_UNSUPPORTED, "batch above 65535");
  if (n == 0) return ok();  // the reference returns without writing
  hipStream_t s = as_stream(stream);
  const int nb = (int)std::min<size_t>(NB, (n + 255) / 256);
  DevBuf part;
  GX_TRY(part.alloc((size_t)nb * batch * sizeof(A)));
  hipLaunchKernelGGL(k_sum_partial<T>, dim3(nb, batch), dim3(256), 0, s, in, n, part.as<A>());
  GX_LAUNCHED("k_sum_partial");
  hipLaunchKernelGGL(k_sum_finish<T>, dim3(batch), dim3(64), 0, s, part.as<A>(), nb, out);
  GX_LAUNCHED("k_sum_finish");
  GX_HIP(hipStreamSynchronize(s));  // the scratch is freed on return
  return ok();
}

int mono_check(const float* in, const float* out, size_t n, unsigned batch) {
  if (!in || !out) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (n == 0 || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "elements and batch must be positive");
  if (batch > 65535) return fail(GTOMX_ERR_UNSUPPORTED, "batch above 65535");
  return GTOMX_OK;
}

int mono_parts(size_t n) { return (int)std::min<size_t>(NB, (n + 255) / 256); }

// The three launches of the normalisation with caller-provided scratch: part
// (mono_parts(n) * batch double2) and coef (batch float2).
template <bool COMPLEX>
int mono_launch(const float* in, float* out, size_t n, unsigned batch, double2* part, float2* coef, hipStream_t s) {
  const int nb = mono_parts(n);
  hipLaunchKernelGGL(k_mono_partial<COMPLEX>, dim3(nb, batch), dim3(256), 0, s, in, n, part);
  GX_LAUNCHED("k_mono_partial");
  hipLaunchKernelGGL(k_mono_finish<COMPLEX>, dim3(batch), dim3(256), 0, s, part, nb, n, coef);
  GX_LAUNCHED("k_mono_finish");
  const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_mono_apply<COMPLEX>, dim3(grid, batch), dim3(256), 0, s, in, out, n, coef);
  GX_LAUNCHED("k_mono_apply");
  return GTOMX_OK;
}

template <bool COMPLEX> int mono(const float* in, float* out, size_t n, unsigned batch, void* stream) {
  GX_TRY(mono_check(in, out, n, batch));
  hipStream_t s = as_stream(stream);
  DevBuf part, coef;
  GX_TRY(part.alloc((size_t)mono_parts(n) * batch * sizeof(double2)));
  GX_TRY(coef.alloc((size_t)batch * sizeof(float2)));
  GX_TRY(mono_launch<COMPLEX>(in, out, n, batch, part.as<double2>(), coef.as<float2>(), s));
  GX_HIP(hipStreamSynchronize(s));  // the scratch is freed on return
  return ok();
}

// The device address of `scratch` on the current device, looked up once per
// device with that device current.
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double2* scratch_ptr[MAX_DEVICES] = {};

int scratch_address(double2** out) {
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) return fail(GTOMX_ERR_UNSUPPORTED, "device ordinal above 63");
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    scratch_ptr[dev] = static_cast<double2*>(p);
  }
  *out = scratch_ptr[dev];
  return GTOMX_OK;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" int gtomx_norm_monolithic_noalloc(const float* d_input, float* d_output, size_t elements, unsigned batch,
                                             void* stream) {
  GX_TRY(mono_check(d_input, d_output, elements, batch));
  if (batch > GTOMX_NOALLOC_MAX_BATCH)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch above GTOMX_NOALLOC_MAX_BATCH (256) for the allocation-free call");
  double2* base = nullptr;
  GX_TRY(scratch_address(&base));
  // partials of item b at base + b * mono_parts(n); coefficients after all partials
  float2* coef = reinterpret_cast<float2*>(base + (size_t)GTOMX_NOALLOC_MAX_BATCH * NB);
  GX_TRY(mono_launch<false>(d_input, d_output, elements, batch, base, coef, as_stream(stream)));
  return ok();
}

extern "C" int gtomx_norm_monolithic(const float* d_input, float* d_output, size_t elements, unsigned batch,
                                     void* stream) {
  return mono<false>(d_input, d_output, elements, batch, stream);
}

extern "C" int gtomx_norm_ft_monolithic(const gtomx_complex* d_input, gtomx_complex* d_output, size_t elements,
                                        unsigned batch, void* stream) {
  return mono<true>(reinterpret_cast<const float*>(d_input), reinterpret_cast<float*>(d_output), elements, batch,
                    stream);
}

extern "C" int gtomx_sum_f32(const float* d_input, float* d_output, size_t n, unsigned batch, void* stream) {
  return sum<float>(d_input, d_output, n, batch, stream);
}
extern "C" int gtomx_sum_f64(const double* d_input, double* d_output, size_t n, unsigned batch, void* stream) {
  return sum<double>(d_input, d_output, n, batch, stream);
}
extern "C" int gtomx_sum_i32(const int32_t* d_input, int32_t* d_output, size_t n, unsigned batch, void* stream) {
  return sum<int32_t>(d_input, d_output, n, batch, stream);
}
