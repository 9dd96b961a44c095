// Batched R2C / C2R transforms of 3D, 2D and 1D volumes (axes 2^a or 3*2^a,
// 8..256) for gfx950. Replace the cuFFT plans wrapped by d_FFTR2C and
// d_IFFTC2R (src/FFT/FFT.cu:7-42, src/FFT/IFFT.cu:8-56).
//
// Layout: real (nx, ny, nz) x-fastest; half spectrum (nx/2+1, ny, nz).
// R2C = rows (real -> half, via an nx/2-point complex FFT) then y then z.
// C2R = z then y (in place on a copy) then rows (half -> real). The row C2R
// uses only the real parts of the x = 0 and x = nx/2 bins, which makes the
// result the real part of the Hermitian-completed inverse.
#include "dev_fft.hpp"
#include "runtime.hpp"

namespace gx {

// ---------------------------------------------------------------------------
// Column transforms along a strided axis, in place. Workgroup = TX adjacent
// columns (contiguous in memory) x R2 threads per column. Plane p selects the
// base offset (p / P1) * S1 + (p % P1) * S0.
// ---------------------------------------------------------------------------
// Tile width of the generic column transforms: 32 columns (256-byte row
// segments) at n = 128 and 256, which stream better along the long strides of
// the y and z passes; the search's own column passes keep ColShape's width.
template <int N> constexpr int fft_col_tx() { return (N == 128 || N == 256) ? 32 : ColShape<N>::TX; }

template <int N, int DIR>
__global__ __launch_bounds__(fft_col_tx<N>() * ColShape<N>::R2) void k_cols(
    float2* __restrict__ data, int width, long long axis_stride, long long nplanes, long long P1,
    long long S0, long long S1, const float2* __restrict__ tw) {
  constexpr int R1 = ColShape<N>::R1, R2 = ColShape<N>::R2, TX = fft_col_tx<N>();
  __shared__ float2 xch[R2 > 1 ? TX * N : 1];
  const int t = threadIdx.x, c = t % TX, j = t / TX;
  const int col = blockIdx.x * TX + c;
  const bool valid = col < width;
  for (long long pl = blockIdx.y; pl < nplanes; pl += gridDim.y) {
    const long long base = (pl / P1) * S1 + (pl % P1) * S0 + col;
    float2 v[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m)
      v[m] = valid ? data[base + (long long)(j + R2 * m) * axis_stride] : make_float2(0.f, 0.f);
    __syncthreads();
    fft_two_step<N, R1, R2, DIR>(v, j, tw, [&](int k1, int jj) { return &xch[(k1 * R2 + jj) * TX + c]; });
    if (valid) {
#pragma unroll
      for (int s = 0; s < R1; ++s) data[base + (long long)out_index<R1, R2>(j, s) * axis_stride] = v[s];
    }
  }
}

// ---------------------------------------------------------------------------
// Row C2R: each row holds N/2+1 complex bins and becomes N reals, scaled.
// Z[k] = (X[k] + conj X[N/2-k]) + i w^k (X[k] - conj X[N/2-k]), w = e^{2 pi i/N},
// z = IDFT_{N/2}(Z), out[2m] = Re z[m], out[2m+1] = Im z[m].
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_rows_c2r(const float2* __restrict__ in, float* __restrict__ out,
                                                  long long nrows, float scale,
                                                  const float2* __restrict__ tw) {
  constexpr int NC = N / 2, NH = N / 2 + 1;
  constexpr int R1 = RowShape<NC>::R1, R2 = RowShape<NC>::R2, RW = 256 / R2;
  constexpr int GS = R2 * (R1 + 1);  // exchange floats2 per row (padded)
  __shared__ float2 xch[R2 > 1 ? RW * GS : 1];
  const int t = threadIdx.x, r = t / R2, j = t % R2;
  const long long row = (long long)blockIdx.x * RW + r;
  const bool valid = row < nrows;
  const float2* X = in + row * NH;
  float2 v[R1];
#pragma unroll
  for (int m = 0; m < R1; ++m) {
    const int k = j + R2 * m;
    float2 a = valid ? X[k] : make_float2(0.f, 0.f);
    float2 b = valid ? X[NC - k] : make_float2(0.f, 0.f);
    if (k == 0) {
      a.y = 0.f;
      b.y = 0.f;
    }
    const float2 bc = cconj(b);
    v[m] = cadd(cadd(a, bc), mul_pj<1>(cmul(twiddle<1, N>(tw, k), csub(a, bc))));
  }
  fft_two_step<NC, R1, R2, 1>(v, j, tw, [&](int k1, int jj) { return &xch[r * GS + jj * (R1 + 1) + k1]; });
  if (valid) {
    float2* o = reinterpret_cast<float2*>(out + row * N);
#pragma unroll
    for (int s = 0; s < R1; ++s) o[out_index<R1, R2>(j, s)] = cscale(v[s], scale);
  }
}

// ---------------------------------------------------------------------------
// Row R2C: N reals -> N/2+1 bins. z[m] = x[2m] + i x[2m+1], Z = DFT_{N/2}(z),
// X[k] = ((Z[k] + conj Z[N/2-k]) - i w^k (Z[k] - conj Z[N/2-k])) / 2, w = e^{-2 pi i/N}.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_rows_r2c(const float* __restrict__ in, float2* __restrict__ out,
                                                  long long nrows, const float2* __restrict__ tw) {
  constexpr int NC = N / 2, NH = N / 2 + 1;
  constexpr int R1 = RowShape<NC>::R1, R2 = RowShape<NC>::R2, RW = 256 / R2;
  constexpr int GS = R2 * (R1 + 1);
  constexpr int ZS = NC + 1;  // padded Z row
  __shared__ float2 xch[R2 > 1 ? RW * GS : 1];
  __shared__ float2 zrow[RW * ZS];
  const int t = threadIdx.x, r = t / R2, j = t % R2;
  const long long row = (long long)blockIdx.x * RW + r;
  const bool valid = row < nrows;
  const float2* x = reinterpret_cast<const float2*>(in + row * N);
  float2 v[R1];
#pragma unroll
  for (int m = 0; m < R1; ++m) v[m] = valid ? x[j + R2 * m] : make_float2(0.f, 0.f);
  fft_two_step<NC, R1, R2, -1>(v, j, tw, [&](int k1, int jj) { return &xch[r * GS + jj * (R1 + 1) + k1]; });
#pragma unroll
  for (int s = 0; s < R1; ++s) zrow[r * ZS + out_index<R1, R2>(j, s)] = v[s];
  __syncthreads();
  if (valid) {
    float2* X = out + row * NH;
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const int k = j + R2 * m;
      const float2 zk = zrow[r * ZS + k];
      const float2 zn = cconj(zrow[r * ZS + ((NC - k) % NC)]);
      const float2 s = cadd(zk, zn), d = csub(zk, zn);
      X[k] = cscale(csub(s, mul_pj<1>(cmul(twiddle<-1, N>(tw, k), d))), 0.5f);
    }
    if (j == 0) {
      const float2 z0 = zrow[r * ZS];
      X[NC] = make_float2(z0.x - z0.y, 0.f);
    }
  }
}

// ---------------------------------------------------------------------------
// Host dispatch
// ---------------------------------------------------------------------------
namespace {

template <int DIR>
int launch_cols_n(int n, float2* data, int width, long long axis_stride, long long nplanes,
                  long long P1, long long S0, long long S1, const float2* tw, hipStream_t s) {
  auto go = [&](auto kernel, int TX, int R2) -> int {
    dim3 grid((width + TX - 1) / TX, (unsigned)std::min<long long>(nplanes, 65535));
    hipLaunchKernelGGL(kernel, grid, dim3(TX * R2), 0, s, data, width, axis_stride, nplanes, P1, S0, S1, tw);
    GX_LAUNCHED("k_cols");
    return GTOMX_OK;
  };
  switch (n) {
#define GX_CASE(N) \
  case N:          \
    return go(k_cols<N, DIR>, fft_col_tx<N>(), ColShape<N>::R2);
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
    default: return fail(GTOMX_ERR_UNSUPPORTED, "column FFT length not supported");
  }
}

template <int N> int rows_per_block() { return 256 / RowShape<N / 2>::R2; }

int launch_rows_c2r(int n, const float2* in, float* out, long long nrows, float scale, const float2* tw,
                    hipStream_t s) {
  auto go = [&](auto kernel, int rw) -> int {
    dim3 grid((unsigned)((nrows + rw - 1) / rw));
    hipLaunchKernelGGL(kernel, grid, dim3(256), 0, s, in, out, nrows, scale, tw);
    GX_LAUNCHED("k_rows_c2r");
    return GTOMX_OK;
  };
  switch (n) {
#define GX_CASE(N) \
  case N:          \
    return go(k_rows_c2r<N>, rows_per_block<N>());
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
    default: return fail(GTOMX_ERR_UNSUPPORTED, "row C2R length not supported");
  }
}

int launch_rows_r2c(int n, const float* in, float2* out, long long nrows, const float2* tw, hipStream_t s) {
  auto go = [&](auto kernel, int rw) -> int {
    dim3 grid((unsigned)((nrows + rw - 1) / rw));
    hipLaunchKernelGGL(kernel, grid, dim3(256), 0, s, in, out, nrows, tw);
    GX_LAUNCHED("k_rows_r2c");
    return GTOMX_OK;
  };
  switch (n) {
#define GX_CASE(N) \
  case N:          \
    return go(k_rows_r2c<N>, rows_per_block<N>());
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
    default: return fail(GTOMX_ERR_UNSUPPORTED, "row R2C length not supported");
  }
}

}  // namespace

int fft_r2c(const float* in, float2* out, gtomx_int3 d, unsigned batch, hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  const long long nh = d.x / 2 + 1;
  GX_TRY(launch_rows_r2c(d.x, in, out, (long long)d.y * d.z * batch, tw, s));
  if (d.y > 1)
    GX_TRY(launch_cols_n<-1>(d.y, out, (int)nh, nh, (long long)d.z * batch, 1LL << 62, nh * d.y, 0, tw, s));
  if (d.z > 1)
    GX_TRY(launch_cols_n<-1>(d.z, out, (int)nh, nh * d.y, (long long)d.y * batch, d.y, nh,
                             nh * d.y * d.z, tw, s));
  return GTOMX_OK;
}

int fft_c2r_inplace_yz(float2* data, gtomx_int3 d, unsigned batch, hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  const long long nh = d.x / 2 + 1;
  if (d.z > 1)
    GX_TRY(launch_cols_n<1>(d.z, data, (int)nh, nh * d.y, (long long)d.y * batch, d.y, nh,
                            nh * d.y * d.z, tw, s));
  if (d.y > 1)
    GX_TRY(launch_cols_n<1>(d.y, data, (int)nh, nh, (long long)d.z * batch, 1LL << 62, nh * d.y, 0, tw, s));
  return GTOMX_OK;
}

int fft_c2r(const float2* in, float2* work, float* out, gtomx_int3 d, unsigned batch, float scale,
            hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  const size_t H = elements_ft(d) * batch;
  if (work != in) GX_HIP(hipMemcpyAsync(work, in, H * sizeof(float2), hipMemcpyDeviceToDevice, s));
  GX_TRY(fft_c2r_inplace_yz(work, d, batch, s));
  GX_TRY(launch_rows_c2r(d.x, work, out, (long long)d.y * d.z * batch, scale, tw, s));
  return GTOMX_OK;
}

}  // namespace gx

using namespace gx;

extern "C" int gtomx_fft_r2c(const float* d_input, gtomx_complex* d_output, int ndims, gtomx_int3 dims,
                             unsigned batch, void* stream) {
  GX_TRY(check_fft_dims(ndims, dims));
  if (!d_input || !d_output) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch must be positive");
  return fft_r2c(d_input, reinterpret_cast<float2*>(d_output), dims, batch, as_stream(stream));
}

extern "C" int gtomx_ifft_c2r(const gtomx_complex* d_input, float* d_output, int ndims, gtomx_int3 dims,
                              unsigned batch, int renormalize, void* stream) {
  GX_TRY(check_fft_dims(ndims, dims));
  if (!d_input || !d_output) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch must be positive");
  DevBuf work;
  GX_TRY(work.alloc(elements_ft(dims) * batch * sizeof(float2)));
  hipStream_t s = as_stream(stream);
  float scale = renormalize ? (float)(1.0 / (double)elements(dims)) : 1.0f;
  GX_TRY(fft_c2r(reinterpret_cast<const float2*>(d_input), work.as<float2>(), d_output, dims, batch, scale, s));
  // The workspace is released when this call returns; wait for its last use.
  GX_HIP(hipStreamSynchronize(s));
  return GTOMX_OK;
}
