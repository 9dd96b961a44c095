// d_PickSubTomogramsDiff2 (src/Correlation/SubTomograms.cu:165-209, kernel
// PickSubTomogramsDiff2Kernel :211-313): for every voxel of the relevant box
// (shift d = index - r/2 around the volume origin) and every orientation a,
//   score = sum over the half-spectrum bins k of |R_a(k) ctf(k) e^{-2 pi i k.d/n} - E(k)|^2,
// and the smallest score with its orientation (strict '<', first wins).
//
// The reference evaluates one sincos per (voxel, bin, orientation). Here the
// sum is expanded,
//   score = sum|R_a ctf|^2 + sum|E|^2 - 2 Re X_a(d),  X_a(d) = sum_k Q_a(k) e^{-2 pi i k.d/n},
// Q_a = R_a ctf conj(E), and X_a is separable along the three axes, so it is
// three small batched complex contractions of Q_a with phase matrices (x, then
// y, then z) over the relevant shifts only, each a GEMM on the matrix cores. The two norms are accumulated in
// double, the score is formed in double and stored as float.
#include <cmath>
#include <vector>

#include "dev_project.hpp"
#include "runtime.hpp"

namespace gx {
namespace {

constexpr int NBLK = 128;  // reduction blocks per orientation

__device__ double block_sum_d(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < 4; ++w) s += red[w];
  __syncthreads();
  return s;
}

// Q[a] = R[a] * ctf * conj(E) in place; partial sums of |R ctf|^2 per orientation
// (parts[a*NBLK + block]) and, for a = 0, of |E|^2 (eparts[block]).
__global__ __launch_bounds__(256) void k_diff2_q(float2* __restrict__ R, const float* __restrict__ ctf,
                                                 const float2* __restrict__ E, size_t h, double* __restrict__ parts,
                                                 double* __restrict__ eparts) {
  __shared__ double red[4];
  const unsigned a = blockIdx.y;
  float2* r = R + a * h;
  double acc = 0, eacc = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < h; i += (size_t)gridDim.x * 256) {
    const float c = ctf[i];
    const float2 v = make_float2(r[i].x * c, r[i].y * c), e = E[i];
    acc += (double)v.x * v.x + (double)v.y * v.y;
    if (a == 0) eacc += (double)e.x * e.x + (double)e.y * e.y;
    r[i] = make_float2(v.x * e.x + v.y * e.y, v.y * e.x - v.x * e.y);
  }
  const double t = block_sum_d(acc, red);
  const double te = block_sum_d(eacc, red);
  if (threadIdx.x == 0) {
    parts[a * NBLK + blockIdx.x] = t;
    if (a == 0) eparts[blockIdx.x] = te;
  }
}

// sums[i] = sum of parts[i*np .. (i+1)*np)
__global__ __launch_bounds__(256) void k_diff2_sum(const double* __restrict__ parts, int np,
                                                   double* __restrict__ sums) {
  __shared__ double red[4];
  double acc = 0;
  for (int i = threadIdx.x; i < np; i += 256) acc += parts[(size_t)blockIdx.x * np + i];
  const double t = block_sum_d(acc, red);
  if (threadIdx.x == 0) sums[blockIdx.x] = t;
}

// out[r1*os1 + r2*os2 + j] = sum_k in[r1*is1 + r2*is2 + k*isk] * W[k*nj + j]
// (complex) for rows r = r1 * n2 + r2 < rows, j < nj, k < nk, on the matrix
// cores (v_mfma_f32_16x16x4_f32, fp32 operands and accumulation). The complex
// product is the real product of A' (rows x 2K: re, im of each k interleaved)
// and, per output part, B' (2K x 16): for the real part rows (W_re, -W_im),
// for the imaginary part (W_im, W_re). A wave computes 16 rows x 16 complex columns (two
// accumulator tiles) over K in steps of 2 complex (4 real) k. Operand lanes:
// A row l % 16, B column l % 16, real k = 4 step + l / 16 for both; result
// row 4 (l / 16) + e, column l % 16. Rows, columns and k past the ends
// contribute zeros and store nothing. Grid: x = row tiles / 4 (four waves per
// workgroup), y = column tiles.
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_contract_mfma(const float2* __restrict__ in, float2* __restrict__ out,
                                                       const float2* __restrict__ W, int n2, int nj, int nk,
                                                       long long is1, long long is2, long long isk, long long os1,
                                                       long long os2, long long rows) {
  const int lane = threadIdx.x % 64;
  const long long tile = (long long)blockIdx.x * 4 + threadIdx.x / 64;
  const int c = lane % 16, q = lane / 16;  // operand row / column, real k within the step
  const long long r = tile * 16 + c;        // this lane's A row
  const bool rok = r < rows;
  const float2* src = in + (rok ? (r / n2) * is1 + (r % n2) * is2 : 0);
  const int j = blockIdx.y * 16 + c;  // this lane's B column
  const bool jok = j < nj;
  const int kq = q >> 1;   // complex k offset within the step
  const bool im = q & 1;   // imaginary half of the complex k
  f32x4 acc_re = {0.f, 0.f, 0.f, 0.f}, acc_im = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < nk; k0 += 2) {
    const int k = k0 + kq;
    const bool kok = k < nk;
    const float2 av = (rok && kok) ? src[k * isk] : make_float2(0.f, 0.f);
    const float2 w = (jok && kok) ? W[(size_t)k * nj + j] : make_float2(0.f, 0.f);
    const float a_ = im ? av.y : av.x;
    const float bre = im ? -w.y : w.x;
    const float bim = im ? w.x : w.y;
    acc_re = __builtin_amdgcn_mfma_f32_16x16x4f32(a_, bre, acc_re, 0, 0, 0);
    acc_im = __builtin_amdgcn_mfma_f32_16x16x4f32(a_, bim, acc_im, 0, 0, 0);
  }
  if (!jok) return;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const long long ro = tile * 16 + 4 * q + e;
    if (ro < rows) out[(ro / n2) * os1 + (ro % n2) * os2 + j] = make_float2(acc_re[e], acc_im[e]);
  }
}

// Running minimum over the nb orientations of this batch (strict '<').
// X is [a][dx][dy][dz]; best/angle are [dz][dy][dx].
__global__ __launch_bounds__(256) void k_diff2_update(const float2* __restrict__ X, const double* __restrict__ A,
                                                      const double* __restrict__ Ee, int rx, int ry, int rz, int nb,
                                                      int angle0, float* __restrict__ best,
                                                      float* __restrict__ angle) {
  const int vr = rx * ry * rz;
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= vr) return;
  const int ix = v % rx, iy = (v / rx) % ry, iz = v / (rx * ry);
  const size_t xi = ((size_t)ix * ry + iy) * rz + iz;
  double b = 1e300;
  int k = -1;
  for (int a = 0; a < nb; ++a) {
    const double sc = A[a] + Ee[0] - 2.0 * (double)X[(size_t)a * vr + xi].x;
    if (sc < b) {
      b = sc;
      k = a;
    }
  }
  if ((float)b < best[v]) {
    best[v] = (float)b;
    angle[v] = (float)(angle0 + k);
  }
}

// Phase matrix W[k][j] = exp(-2 pi i k (j - r/2) / n) for k < nk (k is the
// frequency index; signed and unsigned give the same phase for integer shifts).
std::vector<float2> phases(int nk, int r, int n) {
  std::vector<float2> w((size_t)nk * r);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < nk; ++k)
    for (int j = 0; j < r; ++j) {
      const long long m = ((long long)k * (j - r / 2)) % n;
      const long double ph = -two_pi * (long double)m / (long double)n;
      w[(size_t)k * r + j] = make_float2((float)std::cos(ph), (float)std::sin(ph));
    }
  return w;
}

// out[r1*os1 + r2*os2 + j] = sum_k in[r1*is1 + r2*is2 + k*isk] W[k*nj + j] over
// n1 x n2 rows on the matrix cores.
int contract(const float2* in, float2* out, const float2* W, long long n1, int n2, int nj, int nk, long long is1,
             long long is2, long long isk, long long os1, long long os2, hipStream_t s) {
  const long long rows = n1 * n2;
  const long long tiles = (rows + 15) / 16;
  const long long gx_ = (tiles + 3) / 4;
  if (gx_ > 0x7fffffffLL) return fail(GTOMX_ERR_UNSUPPORTED, "contraction too large");
  hipLaunchKernelGGL(k_contract_mfma, dim3((unsigned)gx_, (unsigned)((nj + 15) / 16)), dim3(256), 0, s, in, out, W,
                     n2, nj, nk, is1, is2, isk, os1, os2, rows);
  GX_LAUNCHED("k_contract_mfma");
  return GTOMX_OK;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" int gtomx_pick_subtomograms_diff2(const gtomx_complex* d_projectorft, float oversample,
                                             gtomx_int3 dimsprojector, const gtomx_complex* d_expft,
                                             const float* d_ctf, gtomx_int3 dimsvolume, unsigned nvolumes,
                                             gtomx_int3 dimsrelevant, const gtomx_float3* h_angles, unsigned nangles,
                                             float* d_bestdiff2, float* d_bestangle, void* stream) {
  if (!d_projectorft || !d_expft || !d_ctf || !d_bestdiff2 || !d_bestangle)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL device buffer");
  if (nangles > 0 && !h_angles) return fail(GTOMX_ERR_INVALID_ARGUMENT, "h_angles is NULL");
  const int n = dimsvolume.x, p = dimsprojector.x;
  if (dimsvolume.y != n || dimsvolume.z != n || n < 2 || n % 2 || n > 1024)
    return fail(GTOMX_ERR_UNSUPPORTED, "volumes must be even cubes of side <= 1024");
  if (dimsprojector.y != p || dimsprojector.z != p || p < 2 || p % 2 || p > 512)
    return fail(GTOMX_ERR_UNSUPPORTED, "projector must be an even cube of side <= 512");
  const int rx = dimsrelevant.x, ry = dimsrelevant.y, rz = dimsrelevant.z;
  if (rx < 1 || ry < 1 || rz < 1 || rx > n || ry > n || rz > n)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "dimsrelevant must lie within the volume");
  if (nvolumes == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "nvolumes must be positive");
  if (!(oversample > 0.f) || !std::isfinite(oversample))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "projectoroversample must be positive");
  hipStream_t s = as_stream(stream);
  const int nh = n / 2 + 1;
  const size_t H = (size_t)nh * n * n;
  const size_t Vr = (size_t)rx * ry * rz;
  // d_ValueFill(best, 1e30), angle 0 (src/Correlation/SubTomograms.cu:182-183)
  GX_TRY(fill_f32(d_bestdiff2, Vr * nvolumes, 1e30f, s));
  GX_TRY(fill_f32(d_bestangle, Vr * nvolumes, 0.f, s));
  if (nangles == 0) return GTOMX_OK;

  const unsigned B = std::min<unsigned>(nangles, 32);
  std::vector<float> h_mats(9 * (size_t)nangles);
  GX_TRY(gtomx_projector_matrices(h_angles, nangles, oversample, h_mats.data()));
  const std::vector<float2> wx = phases(nh, rx, n), wy = phases(n, ry, n), wz = phases(n, rz, n);
  DevBuf mats, Wx, Wy, Wz, R, T1, T2, T3, parts, sums, eparts, esum;
  GX_TRY(mats.alloc(h_mats.size() * sizeof(float)));
  GX_TRY(Wx.alloc(wx.size() * sizeof(float2)));
  GX_TRY(Wy.alloc(wy.size() * sizeof(float2)));
  GX_TRY(Wz.alloc(wz.size() * sizeof(float2)));
  GX_TRY(R.alloc((size_t)B * H * sizeof(float2)));
  GX_TRY(T1.alloc((size_t)B * n * n * rx * sizeof(float2)));
  GX_TRY(T2.alloc((size_t)B * n * rx * ry * sizeof(float2)));
  GX_TRY(T3.alloc((size_t)B * Vr * sizeof(float2)));
  GX_TRY(parts.alloc((size_t)B * NBLK * sizeof(double)));
  GX_TRY(sums.alloc((size_t)B * sizeof(double)));
  GX_TRY(eparts.alloc(NBLK * sizeof(double)));
  GX_TRY(esum.alloc(sizeof(double)));
  GX_HIP(hipMemcpyAsync(mats.p, h_mats.data(), h_mats.size() * sizeof(float), hipMemcpyHostToDevice, s));
  GX_HIP(hipMemcpyAsync(Wx.p, wx.data(), wx.size() * sizeof(float2), hipMemcpyHostToDevice, s));
  GX_HIP(hipMemcpyAsync(Wy.p, wy.data(), wy.size() * sizeof(float2), hipMemcpyHostToDevice, s));
  GX_HIP(hipMemcpyAsync(Wz.p, wz.data(), wz.size() * sizeof(float2), hipMemcpyHostToDevice, s));

  for (unsigned v = 0; v < nvolumes; ++v) {
    const float2* Ev = reinterpret_cast<const float2*>(d_expft) + (size_t)v * H;
    const float* cv = d_ctf + (size_t)v * H;  // one CTF per volume (src/Correlation/SubTomograms.cu:199)
    for (unsigned b0 = 0; b0 < nangles; b0 += B) {
      const unsigned nb = std::min(B, nangles - b0);
      // R_a: the projection at the volume size (d_rlnProject, :195), rmax = n/2
      GX_TRY(project_launch(reinterpret_cast<const float2*>(d_projectorft), p, R.as<float2>(), n, (unsigned)(n / 2),
                            mats.as<float>() + 9 * (size_t)b0, nb, s));
      hipLaunchKernelGGL(k_diff2_q, dim3(NBLK, nb), dim3(256), 0, s, R.as<float2>(), cv, Ev, H, parts.as<double>(),
                         eparts.as<double>());
      GX_LAUNCHED("k_diff2_q");
      hipLaunchKernelGGL(k_diff2_sum, dim3(nb), dim3(256), 0, s, parts.as<double>(), NBLK, sums.as<double>());
      GX_LAUNCHED("k_diff2_sum");
      hipLaunchKernelGGL(k_diff2_sum, dim3(1), dim3(256), 0, s, eparts.as<double>(), NBLK, esum.as<double>());
      GX_LAUNCHED("k_diff2_sum");
      // x: Q[a][kz][ky][kx] -> T1[a][kz][ky][dx]
      GX_TRY(contract(R.as<float2>(), T1.as<float2>(), Wx.as<float2>(), (long long)nb * n * n, 1, rx, nh,
                      (long long)nh, 0LL, 1LL, (long long)rx, 0LL, s));
      // y: T1 rows (a*n + kz, dx) over ky -> T2[a][kz][dx][dy]
      GX_TRY(contract(T1.as<float2>(), T2.as<float2>(), Wy.as<float2>(), (long long)nb * n, rx, ry, n,
                      (long long)n * rx, 1LL, (long long)rx, (long long)rx * ry, (long long)ry, s));
      // z: T2 rows (a, dx*ry + dy) over kz -> T3[a][dx][dy][dz]
      GX_TRY(contract(T2.as<float2>(), T3.as<float2>(), Wz.as<float2>(), (long long)nb, rx * ry, rz, n,
                      (long long)n * rx * ry, 1LL, (long long)rx * ry, (long long)Vr, (long long)rz, s));
      hipLaunchKernelGGL(k_diff2_update, dim3((unsigned)((Vr + 255) / 256)), dim3(256), 0, s, T3.as<float2>(),
                         sums.as<double>(), esum.as<double>(), rx, ry, rz, (int)nb, (int)b0,
                         d_bestdiff2 + (size_t)v * Vr, d_bestangle + (size_t)v * Vr);
      GX_LAUNCHED("k_diff2_update");
    }
  }
  GX_HIP(hipStreamSynchronize(s));  // the scratch is freed on return
  return GTOMX_OK;
}
