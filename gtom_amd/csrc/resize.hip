// Fourier-space resizing of batches, real or complex (src/FFT/FFTResize.cu):
//  * d_FFTCrop (:20-38, FFTCropKernel :121-138): half spectra, rows y < new/2+1
//    keep their index, the others come from y - new + old (likewise z);
//  * d_FFTPad (:68-88, FFTPadEvenKernel :162-205): half spectra, output row y
//    has centred coordinate (y + new/2) % new, old centred coordinate
//    that + (old - new)/2 (C division), read at (c + (old+1)/2) % old when inside
//    [0, old) and x < (old.x+1)/2, else 0;
//  * d_FFTFullCrop (:44-62, FFTFullCropKernel :140-160): full arrays, indices
//    >= new/2 move by old - new on every axis;
//  * d_FFTFullPad (:92-110, FFTFullPadEvenKernel :212-238): full arrays, as
//    d_FFTPad on every axis with the shift (old - new - ((old & (1 - (new &
//    1))) % 2)) / 2 (the kernel's `old & 1 - (new & 1)` parses that way).
// One kernel walks the output row by row and reads the mapped input element or
// writes 0.
// In place (input == output, buffer sized for the larger dims) goes through a
// temporary, as the reference's wrappers do.
#include <algorithm>

#include "runtime.hpp"

namespace gx {

namespace {

__host__ __device__ inline int pad_shift(int o, int n) { return (o - n - ((o & (1 - (n & 1))) % 2)) / 2; }

// Source index along one axis, or -1 for a zero.
template <int OP> __device__ __forceinline__ int src_axis(int i, int o, int n, bool half_x) {
  if (OP == FFT_CROP) {
    if (half_x) return i;
    return i < n / 2 + 1 ? i : i - n + o;
  } else if (OP == FFT_FULL_CROP) {
    return i >= n / 2 ? i + o - n : i;
  } else if (OP == FFT_PAD) {
    if (half_x) return i < (o + 1) / 2 ? i : -1;
    const int c = (i + n / 2) % n + (o - n) / 2;
    return c >= 0 && c < o ? (c + (o + 1) / 2) % o : -1;
  } else {  // FFT_FULL_PAD
    const int c = (i + n / 2) % n + pad_shift(o, n);
    return c >= 0 && c < o ? (c + (o + 1) / 2) % o : -1;
  }
}

// One wave per output row (4 rows per workgroup, blockIdx.y the item): the
// row's source row is decided once, then the wave writes the row with
// coalesced stores. Row arithmetic is 32-bit (n <= 256 per axis).
template <class T, int OP>
__global__ __launch_bounds__(256) void k_fft_resize(const T* __restrict__ in, T* __restrict__ out, int3 od, int3 nd,
                                                    size_t vnew, size_t vold) {
  constexpr bool HALF = OP == FFT_CROP || OP == FFT_PAD;
  const int nxs = HALF ? nd.x / 2 + 1 : nd.x, oxs = HALF ? od.x / 2 + 1 : od.x;
  const int lane = threadIdx.x & 63;
  const unsigned rows = (unsigned)nd.y * (unsigned)nd.z;
  const unsigned row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const int y = (int)(row % (unsigned)nd.y), z = (int)(row / (unsigned)nd.y);
  const size_t b = blockIdx.y;
  const int sy = src_axis<OP>(y, od.y, nd.y, false), sz = src_axis<OP>(z, od.z, nd.z, false);
  T* o = out + b * vnew + (size_t)row * nxs;
  if (sy < 0 || sz < 0) {
    for (int x = lane; x < nxs; x += 64) o[x] = T{};
    return;
  }
  const T* src = in + b * vold + ((size_t)sz * od.y + sy) * oxs;
  for (int x = lane; x < nxs; x += 64) {
    const int sx = src_axis<OP>(x, od.x, nd.x, HALF);
    o[x] = sx >= 0 ? src[sx] : T{};
  }
}

template <class T, int OP> int launch(const void* in, void* out, gtomx_int3 od, gtomx_int3 nd, unsigned batch,
                                      hipStream_t s) {
  constexpr bool HALF = OP == FFT_CROP || OP == FFT_PAD;
  const size_t vnew = HALF ? elements_ft(nd) : elements(nd), vold = HALF ? elements_ft(od) : elements(od);
  const unsigned rows = (unsigned)nd.y * (unsigned)nd.z;
  for (unsigned b0 = 0; b0 < batch; b0 += 65535) {
    const unsigned nb = std::min(batch - b0, 65535u);
    hipLaunchKernelGGL((k_fft_resize<T, OP>), dim3((rows + 3) / 4, nb), dim3(256), 0, s,
                       static_cast<const T*>(in) + (size_t)b0 * vold, static_cast<T*>(out) + (size_t)b0 * vnew,
                       make_int3(od.x, od.y, od.z), make_int3(nd.x, nd.y, nd.z), vnew, vold);
    GX_LAUNCHED("k_fft_resize");
  }
  return GTOMX_OK;
}

template <class T> int dispatch(int op, const void* in, void* out, gtomx_int3 od, gtomx_int3 nd, unsigned batch,
                                hipStream_t s) {
  switch (op) {
    case FFT_CROP: return launch<T, FFT_CROP>(in, out, od, nd, batch, s);
    case FFT_PAD: return launch<T, FFT_PAD>(in, out, od, nd, batch, s);
    case FFT_FULL_CROP: return launch<T, FFT_FULL_CROP>(in, out, od, nd, batch, s);
    default: return launch<T, FFT_FULL_PAD>(in, out, od, nd, batch, s);
  }
}

}  // namespace

int fft_resize(int op, bool complex_values, const void* in, void* out, gtomx_int3 od, gtomx_int3 nd, unsigned batch,
               hipStream_t s) {
  return complex_values ? dispatch<float2>(op, in, out, od, nd, batch, s)
                        : dispatch<float>(op, in, out, od, nd, batch, s);
}

}  // namespace gx

using namespace gx;

namespace {

int resize_entry(int op, bool complex_values, const void* d_input, void* d_output, gtomx_int3 od, gtomx_int3 nd,
                 unsigned batch, void* stream) {
  if (!d_input || !d_output) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (od.x < 1 || od.y < 1 || od.z < 1 || nd.x < 1 || nd.y < 1 || nd.z < 1)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "dims must be positive");
  if (batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch must be positive");
  if ((op == FFT_CROP || op == FFT_FULL_CROP) && (nd.x > od.x || nd.y > od.y || nd.z > od.z))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "a crop needs new dims no larger than the old ones");
  const bool half = op == FFT_CROP || op == FFT_PAD;
  const size_t vnew = half ? elements_ft(nd) : elements(nd);
  if (vnew * batch >= ((size_t)1 << 40) || (size_t)nd.y * nd.z >= ((size_t)1 << 31))
    return fail(GTOMX_ERR_UNSUPPORTED, "output too large");
  hipStream_t s = as_stream(stream);
  if (d_input != d_output) return fft_resize(op, complex_values, d_input, d_output, od, nd, batch, s);
  DevBuf tmp;
  const size_t bytes = vnew * batch * (complex_values ? 8 : 4);
  GX_TRY(tmp.alloc(bytes));
  GX_TRY(fft_resize(op, complex_values, d_input, tmp.p, od, nd, batch, s));
  GX_HIP(hipMemcpyAsync(d_output, tmp.p, bytes, hipMemcpyDeviceToDevice, s));
  GX_HIP(hipStreamSynchronize(s));
  return GTOMX_OK;
}

}  // namespace

extern "C" int gtomx_fft_full_pad(const float* d_input, float* d_output, gtomx_int3 olddims, gtomx_int3 newdims,
                                  unsigned batch, void* stream) {
  return resize_entry(FFT_FULL_PAD, false, d_input, d_output, olddims, newdims, batch, stream);
}

extern "C" int gtomx_fft_resize(int op, int complex_values, const void* d_input, void* d_output, gtomx_int3 olddims,
                                gtomx_int3 newdims, unsigned batch, void* stream) {
  if (op < FFT_CROP || op > FFT_FULL_PAD) return fail(GTOMX_ERR_INVALID_ARGUMENT, "unknown resize operation");
  return resize_entry(op, complex_values != 0, d_input, d_output, olddims, newdims, batch, stream);
}
