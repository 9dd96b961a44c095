// The TomoPicker class (include/Correlation.cuh:209-257,
// src/Correlation/TomoPicker.cu) and d_Exact2DWeighting
// (src/Projection/Weighting.cu:21-55, kernel :92-133).
//
// The reference class never runs: its constructor and Initialize throw
// (TomoPicker.cu:23, :28), the projections of the reference and mask at the
// tilt angles are commented out as broken (:223-229, :264-269), and the
// correlation call names a d_CorrelateRealspace overload that does not exist
// (:297). This is the pipeline its code describes, with the projections done
// by the device-array projector as 2D central sections of the padded 3D
// spectra (src/Relion/Project.cu:89-177) and the correlation as the masked,
// locally normalised correlation that d_CorrelateRealspace computes for one
// pair (src/Correlation/Realspace.cu:16-102), evaluated at every image
// position through FFTs:
//   Initialize: P = R2C(ifftshift(pad(ref, 2 nr))), the same for the mask (:80-103);
//   SetImage: filter = Exact2DWeighting(tilts, weights), images <- C2R_n(R2C(images)
//     * filter) (left in the caller's buffer, as the reference does), their spectra
//     and the spectra of their squares (:158-186);
//   PerformCorrelation, per orientation a and tilt image i, with
//   M_i = transpose(Euler(a)) transpose(Euler(tilt_i)) (:208-216):
//     mask (first orientation, or every one if not circular): section ->
//       C2R_n -> FullCrop to nr^2 -> binarise at 0.99 -> M_i = sum -> FullPad to
//       the image -> local sums s1, s2 of image and image^2 under it (:231-263);
//     reference: section -> C2R (unnormalised) -> FullCrop -> mean 0 / std 1 under
//       the mask (d_NormMonolithic with a mask, src/ImageManipulation/Norm.cu:286-333)
//       -> times the mask -> FullPad -> CC_i(x) = C2R_n(imageft * conj(.))(x) /
//       (M_i sigma_i(x)), sigma_i = sqrt(max(0, M_i s2 - s1^2)) / M_i, 0 where
//       sigma_i is not positive;
//     UpdateTomoCorrKernel (:311-358): every voxel of the output volume averages
//       the CC_i (bilinear, clamped, at its projection T_i pos + centre) with the
//       image weights, keeps the maximum with strict '<' and the orientation, and
//       accumulates the average and its square.
// Everything is batched over the tilt images; each orientation is a fixed
// sequence of launches.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dev_fft.hpp"
#include "runtime.hpp"

namespace gx {
namespace {

__device__ __forceinline__ float sincf_ref(float x) {
  // sinc of include/DeviceFunctions.cuh:11-17 (float): 1 for |x| <= 1e-4
  return fabsf(x) <= 1e-4f ? 1.f : sinf(x * 3.14159265358979f) / (x * 3.14159265358979f);
}

// Exact2DWeightingKernel for FFT-order weights: filter b of d_weights (nh x ny)
// in the frame and with the weight of image b, as the reference launches it
// (h_globalB2localB[b] and currentimage = b, src/Projection/Weighting.cu:40-46;
// its h_indices[b] picks a normal the kernel never reads); basis[9 i..] =
// Matrix3Euler(angle_i) row-major.
__global__ __launch_bounds__(256) void k_exact2d_weighting(float* __restrict__ w, int nx, int ny,
                                                           const int* __restrict__ indices,
                                                           const float* __restrict__ basis,
                                                           const float* __restrict__ iw, int nimages, float maxfreq) {
  const int nh = nx / 2 + 1;
  const int b = blockIdx.y;
  (void)indices;
  const float* E = basis + 9 * b;
  const float wc = iw[b];
  float* out = w + (size_t)b * nh * ny;
  for (int id = blockIdx.x * 256 + threadIdx.x; id < nh * ny; id += gridDim.x * 256) {
    const int idx = id % nh, idy = id / nh;
    const int y = ny - 1 - ((idy + ny / 2 - 1) % ny), x = nx / 2 - idx;
    const float ax = (float)idx - (float)(nx / 2), ay = (float)idy - (float)(ny / 2);
    float& dst = out[(size_t)y * nh + x];
    if (sqrtf(ax * ax + ay * ay) <= maxfreq) {
      const float g0 = ax * E[0] + ay * E[3], g1 = ax * E[1] + ay * E[4], g2 = ax * E[2] + ay * E[5];
      float sum = 0.f;
      for (int k = 0; k < nimages; ++k) {
        const float* Nk = basis + 9 * k + 6;
        sum += fabsf(sincf_ref(g0 * Nk[0] + g1 * Nk[1] + g2 * Nk[2])) * iw[k];
      }
      dst *= wc * wc / sum;
    } else {
      dst *= 0.f;
    }
  }
}

__global__ __launch_bounds__(256) void k_mul_real(float2* __restrict__ a, const float* __restrict__ f, size_t n) {
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const float s = f[i];
    a[i] = make_float2(a[i].x * s, a[i].y * s);
  }
}

// out[b][k] = a[b][k] * conj(c[b][k])
__global__ __launch_bounds__(256) void k_times_conj(const float2* __restrict__ a, const float2* __restrict__ c,
                                                    float2* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    out[i] = cmulc(a[i], c[i]);
}

__global__ __launch_bounds__(256) void k_square_f(const float* __restrict__ a, float* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) out[i] = a[i] * a[i];
}

__device__ double block_sum_dd(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < 4; ++w) s += red[w];
  __syncthreads();
  return s;
}

// d_Binarize at 0.99 (src/BinaryManipulation/Binarize.cu:30-36) and the mask sum, one workgroup per image.
__global__ __launch_bounds__(256) void k_binarize_sum(float* __restrict__ m, int n, double* __restrict__ msum) {
  __shared__ double red[4];
  float* mb = m + (size_t)blockIdx.x * n;
  double s = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const float v = mb[i] >= 0.99f ? 1.f : 0.f;
    mb[i] = v;
    s += v;
  }
  const double t = block_sum_dd(s, red);
  if (threadIdx.x == 0) msum[blockIdx.x] = t;
}

// Mean 0 / std 1 of r under the mask m (NormMeanStdDevMonoMaskedKernel,
// src/ImageManipulation/Norm.cu:286-333, sums in double), then times the mask;
// one workgroup per image.
__global__ __launch_bounds__(256) void k_masked_norm(float* __restrict__ r, const float* __restrict__ m, int n) {
  __shared__ double red[4];
  __shared__ double stat[2];
  float* rb = r + (size_t)blockIdx.x * n;
  const float* mb = m + (size_t)blockIdx.x * n;
  double s1 = 0, s2 = 0, sm = 0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double v = rb[i], w = mb[i];
    s1 += v * w;
    s2 += v * v * w;
    sm += w;
  }
  const double t1 = block_sum_dd(s1, red), t2 = block_sum_dd(s2, red), tm = block_sum_dd(sm, red);
  if (threadIdx.x == 0) {
    const double mean = t1 / tm, sd = sqrt(tm * t2 - t1 * t1) / tm;
    stat[0] = mean;
    stat[1] = sd > 0.0 ? 1.0 / sd : 0.0;
  }
  __syncthreads();
  const float mean = (float)stat[0], isd = (float)stat[1];
  for (int i = threadIdx.x; i < n; i += 256) rb[i] = (rb[i] - mean) * isd * mb[i];
}

// CC = num / (M sigma), sigma = sqrt(max(0, M s2 - s1^2)) / M (double), 0 where sigma <= 0.
__global__ __launch_bounds__(256) void k_local_cc(float* __restrict__ num, const float* __restrict__ s1,
                                                  const float* __restrict__ s2, const double* __restrict__ msum,
                                                  size_t vi, unsigned nimg) {
  const size_t total = vi * nimg;
  for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    const double M = msum[i / vi];
    const double a = s1[i], b = s2[i];
    const double sig = sqrt(fmax(0.0, M * b - a * a)) / M;
    num[i] = sig > 0.0 ? (float)((double)num[i] / (M * sig)) : 0.f;
  }
}

// UpdateTomoCorrKernel (src/Correlation/TomoPicker.cu:311-358) with the texture
// fetch as bilinear interpolation of CC_n at (px - 0.5, py - 0.5) with clamped
// indices (unnormalised coordinates clamp); T: nimg x (row 0, row 1) of Matrix3Euler(tilt_n).
// The tilt rows and weights are staged in LDS (at most 128 images); every image
// is sampled without a divergent branch (indices clamped, weight 0 where the
// point falls outside, which adds exactly 0 to both sums), with 32-bit byte
// offsets within an image (images of at most 4096^2). An image whose second
// row is exactly (0, 1, 0) -- a tilt about y alone -- maps every voxel to
// py = y - cy/2 + ny/2 + 0.5, so its sample lies on a pixel row (tv = 0) and
// only the two pixels of that row are read: the general formula's second row
// enters with weight exactly 0 there.
__device__ __forceinline__ bool in_range(float p, float n) { return fminf(p, n - p) >= 0.f; }

__global__ __launch_bounds__(256) void k_tomo_update(const float* __restrict__ cc, int nx, int ny, int nimg,
                                                     const float* __restrict__ T, const float* __restrict__ iw,
                                                     int cx, int cy, int cz, float* __restrict__ corr,
                                                     float3* __restrict__ corrangle, int32_t* __restrict__ corridx,
                                                     float* __restrict__ sum1, float* __restrict__ sum2, float3 angle,
                                                     int aindex) {
  // image n: (T0, T1, T2, T3), (T4, T5, weight, 1 if the second row is (0, 1, 0) else 0)
  __shared__ float4 sT[2 * 128];
  for (int i = threadIdx.x; i < 2 * nimg; i += 256) {
    const int n = i >> 1;
    const float* Tn = T + 6 * n;
    const bool ytilt = Tn[3] == 0.f && Tn[4] == 1.f && Tn[5] == 0.f;
    sT[i] = (i & 1) ? make_float4(Tn[4], Tn[5], iw[n], ytilt ? 1.f : 0.f) : make_float4(Tn[0], Tn[1], Tn[2], Tn[3]);
  }
  __syncthreads();
  const size_t total = (size_t)cx * cy * cz;
  const size_t vi = (size_t)nx * ny;
  const float fnx = (float)nx, fny = (float)ny, fx1 = (float)(nx - 1), fy1 = (float)(ny - 1);
  const float icx = (float)(nx / 2) + 0.5f, icy = (float)(ny / 2) + 0.5f;
  for (size_t o = blockIdx.x * (size_t)256 + threadIdx.x; o < total; o += (size_t)gridDim.x * 256) {
    const int x = (int)(o % cx), y = (int)((o / cx) % cy), z = (int)(o / ((size_t)cx * cy));
    const float px0 = (float)x - (float)(cx / 2), py0 = (float)y - (float)(cy / 2), pz0 = (float)z - (float)(cz / 2);
    // the y-tilt images' row: py = py0 + icy, v = py - 0.5 an integer
    const float pyt = py0 + icy;
    const bool yin = in_range(pyt, fny);
    const float vt = pyt - 0.5f;
    const unsigned rt = __umul24((unsigned)__builtin_amdgcn_fmed3f(vt, 0.f, fy1), (unsigned)nx);
    float sum = 0.f, samples = 0.f;
    for (int n = 0; n < nimg; ++n) {
      const float4 ta = sT[2 * n], tb = sT[2 * n + 1];
      const char* c = reinterpret_cast<const char*>(cc + vi * n);
      const float px = ta.x * px0 + ta.y * py0 + ta.z * pz0 + icx;
      const float u = px - 0.5f;
      const float fu = floorf(u);
      const float tu = u - fu;
      const unsigned xa = (unsigned)__builtin_amdgcn_fmed3f(fu, 0.f, fx1);
      const unsigned xb = (unsigned)__builtin_amdgcn_fmed3f(fu + 1.f, 0.f, fx1);
      float val;
      bool inside;
      if (__builtin_amdgcn_readfirstlane(__float_as_int(tb.w)) != 0) {
        const float a00 = *reinterpret_cast<const float*>(c + 4u * (rt + xa));
        const float a01 = *reinterpret_cast<const float*>(c + 4u * (rt + xb));
        val = (1.f - tu) * a00 + tu * a01;
        inside = static_cast<int>(in_range(px, fnx)) & static_cast<int>(yin);
      } else {
        const float py = ta.w * px0 + tb.x * py0 + tb.y * pz0 + icy;
        const float v = py - 0.5f;
        const float fv = floorf(v);
        const float tv = v - fv;
        const unsigned ra = __umul24((unsigned)__builtin_amdgcn_fmed3f(fv, 0.f, fy1), (unsigned)nx);
        const unsigned rb = __umul24((unsigned)__builtin_amdgcn_fmed3f(fv + 1.f, 0.f, fy1), (unsigned)nx);
        const float a00 = *reinterpret_cast<const float*>(c + 4u * (ra + xa));
        const float a01 = *reinterpret_cast<const float*>(c + 4u * (ra + xb));
        const float a10 = *reinterpret_cast<const float*>(c + 4u * (rb + xa));
        const float a11 = *reinterpret_cast<const float*>(c + 4u * (rb + xb));
        val = (1.f - tv) * ((1.f - tu) * a00 + tu * a01) + tv * ((1.f - tu) * a10 + tu * a11);
        inside = static_cast<int>(in_range(px, fnx)) & static_cast<int>(in_range(py, fny));
      }
      const float w = inside ? tb.z : 0.f;
      sum += val * w;
      samples += w;
    }
    if (samples > 0.f) {
      sum /= samples;
      if (corr[o] < sum) {
        corr[o] = sum;
        if (corrangle) corrangle[o] = angle;
        if (corridx) corridx[o] = aindex;
      }
      if (sum1) sum1[o] += sum;
      if (sum2) sum2[o] += sum * sum;
    }
  }
}

// The same update for a tilt series about one in-plane axis: every image's
// second row is (a, b, 0), so py = a px0 + b py0 + ny/2 + 0.5 is one value per
// voxel for all images (its row pair, fraction and range test are computed
// once), and the voxels are dealt to lanes along the direction that keeps py
// nearly constant: with |a| <= |b| a wave runs along x over the sheared rows
// y = (j + round(-a x / b)) mod cy (else along y over x = (j + round(-b y / a))
// mod cx), so its samples fall on one or two image rows instead of up to 64 |a|.
// Each voxel is still visited once, and the sample and sums are those of
// k_tomo_update.
__global__ __launch_bounds__(256) void k_tomo_update_axis(const float* __restrict__ cc, int nx, int ny, int nimg,
                                                          const float* __restrict__ T, const float* __restrict__ iw,
                                                          int cx, int cy, int cz, float* __restrict__ corr,
                                                          float3* __restrict__ corrangle,
                                                          int32_t* __restrict__ corridx, float* __restrict__ sum1,
                                                          float* __restrict__ sum2, float3 angle, int aindex,
                                                          int xmajor, float slope) {
  __shared__ float4 sT[128];  // image n: (T0, T1, T2, weight)
  for (int n = threadIdx.x; n < nimg; n += 256) sT[n] = make_float4(T[6 * n], T[6 * n + 1], T[6 * n + 2], iw[n]);
  __syncthreads();
  const float a = T[3], b = T[4], c = T[5];
  const size_t total = (size_t)cx * cy * cz;
  const size_t vi = (size_t)nx * ny;
  const int cu = xmajor ? cx : cy, cv = xmajor ? cy : cx;
  const float fnx = (float)nx, fny = (float)ny, fx1 = (float)(nx - 1), fy1 = (float)(ny - 1);
  const float icx = (float)(nx / 2) + 0.5f, icy = (float)(ny / 2) + 0.5f;
  for (size_t q = blockIdx.x * (size_t)256 + threadIdx.x; q < total; q += (size_t)gridDim.x * 256) {
    const int u = (int)(q % cu), j = (int)((q / cu) % cv), z = (int)(q / ((size_t)cu * cv));
    int w = (j + (int)rintf(slope * (float)u)) % cv;
    w += w < 0 ? cv : 0;
    const int x = xmajor ? u : w, y = xmajor ? w : u;
    const size_t o = ((size_t)z * cy + y) * cx + x;
    const float px0 = (float)x - (float)(cx / 2), py0 = (float)y - (float)(cy / 2), pz0 = (float)z - (float)(cz / 2);
    const float py = a * px0 + b * py0 + c * pz0 + icy;
    const bool yin = in_range(py, fny);
    const float v = py - 0.5f;
    const float fv = floorf(v);
    const float tv = v - fv;
    const unsigned ra = __umul24((unsigned)__builtin_amdgcn_fmed3f(fv, 0.f, fy1), (unsigned)nx);
    const unsigned rb = __umul24((unsigned)__builtin_amdgcn_fmed3f(fv + 1.f, 0.f, fy1), (unsigned)nx);
    float sum = 0.f, samples = 0.f;
    for (int n = 0; n < nimg; ++n) {
      const float4 ta = sT[n];
      const char* cb = reinterpret_cast<const char*>(cc + vi * n);
      const float px = ta.x * px0 + ta.y * py0 + ta.z * pz0 + icx;
      const float uu = px - 0.5f;
      const float fu = floorf(uu);
      const float tu = uu - fu;
      const unsigned xa = (unsigned)__builtin_amdgcn_fmed3f(fu, 0.f, fx1);
      const unsigned xb = (unsigned)__builtin_amdgcn_fmed3f(fu + 1.f, 0.f, fx1);
      const float a00 = *reinterpret_cast<const float*>(cb + 4u * (ra + xa));
      const float a01 = *reinterpret_cast<const float*>(cb + 4u * (ra + xb));
      const float a10 = *reinterpret_cast<const float*>(cb + 4u * (rb + xa));
      const float a11 = *reinterpret_cast<const float*>(cb + 4u * (rb + xb));
      const float val = (1.f - tv) * ((1.f - tu) * a00 + tu * a01) + tv * ((1.f - tu) * a10 + tu * a11);
      const float wgt = (static_cast<int>(in_range(px, fnx)) & static_cast<int>(yin)) ? ta.w : 0.f;
      sum += val * wgt;
      samples += wgt;
    }
    if (samples > 0.f) {
      sum /= samples;
      if (corr[o] < sum) {
        corr[o] = sum;
        if (corrangle) corrangle[o] = angle;
        if (corridx) corridx[o] = aindex;
      }
      if (sum1) sum1[o] += sum;
      if (sum2) sum2[o] += sum * sum;
    }
  }
}

unsigned grid_for_n(size_t n) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 16384)); }

// Matrix3Euler (src/Helpers/Euler.cu:39-68), row-major, double.
void euler_rowmajor(const gtomx_float3& a, double* E) {
  const double ca = std::cos((double)a.x), sa = std::sin((double)a.x);
  const double cb = std::cos((double)a.y), sb = std::sin((double)a.y);
  const double cg = std::cos((double)a.z), sg = std::sin((double)a.z);
  const double cc = cb * ca, cs = cb * sa, sc = sb * ca, ss = sb * sa;
  const double A[9] = {cg * cc - sg * sa, cg * cs + sg * ca, -cg * sb, -sg * cc - cg * sa, -sg * cs + cg * ca,
                       sg * sb, sc, ss, cb};
  std::memcpy(E, A, sizeof(A));
}

int exact2d_weighting_launch(float* d_w, int nx, int ny, const int* h_indices, const gtomx_float3* h_angles,
                             const float* d_iw, unsigned nimages, float maxfreq, unsigned batch, hipStream_t s) {
  std::vector<float> basis(9 * (size_t)nimages);
  for (unsigned i = 0; i < nimages; ++i) {
    double E[9];
    euler_rowmajor(h_angles[i], E);
    for (int k = 0; k < 9; ++k) basis[9 * (size_t)i + k] = (float)E[k];
  }
  DevBuf db, di, dw;
  GX_TRY(db.alloc(basis.size() * sizeof(float)));
  GX_TRY(di.alloc(batch * sizeof(int)));
  GX_HIP(hipMemcpyAsync(db.p, basis.data(), basis.size() * sizeof(float), hipMemcpyHostToDevice, s));
  GX_HIP(hipMemcpyAsync(di.p, h_indices, batch * sizeof(int), hipMemcpyHostToDevice, s));
  const float* iw = d_iw;
  if (!iw) {
    std::vector<float> ones(nimages, 1.f);
    GX_TRY(dw.alloc(nimages * sizeof(float)));
    GX_HIP(hipMemcpyAsync(dw.p, ones.data(), nimages * sizeof(float), hipMemcpyHostToDevice, s));
    iw = dw.as<float>();
  }
  const int nh = nx / 2 + 1;
  hipLaunchKernelGGL(k_exact2d_weighting, dim3(grid_for_n((size_t)nh * ny), batch), dim3(256), 0, s, d_w, nx, ny,
                     di.as<int>(), db.as<float>(), iw, (int)nimages, maxfreq);
  GX_LAUNCHED("k_exact2d_weighting");
  GX_HIP(hipStreamSynchronize(s));  // the staged tables are freed on return
  return GTOMX_OK;
}

}  // namespace
}  // namespace gx

struct gtomx_tomopicker {
  int nr = 0, p = 0, nx = 0, ny = 0;
  unsigned nimg = 0;
  bool circular = true;
  int device = 0;
  size_t Hp3 = 0, Hp = 0, Vp = 0, Vr = 0, Hi = 0, Vi = 0;
  float2 *Pref = nullptr, *Pmask = nullptr;
  float2 *imageft = nullptr, *image2ft = nullptr;
  float* iw = nullptr;  // image weights (device)
  std::vector<gtomx_float3> tilt;
  float* T = nullptr;   // nimg x 6 tilt rows (device)
  // every tilt about y alone (second rows (0, 1, 0)), or every second row the same (a, b, 0): a
  // tilt series about one in-plane axis (rot, tilt, psi = 0)
  bool ytilt = false, common_row = false;
  float row_a = 0.f, row_b = 1.f;
  float* mats = nullptr;  // nimg x 9 (device)
  float* h_mats = nullptr;  // pinned staging
  hipEvent_t staged = nullptr;
  bool staged_recorded = false;
  float* mask2d = nullptr;
  double* msum = nullptr;
  float *s1 = nullptr, *s2 = nullptr;
  float2* sec = nullptr;
  float* full = nullptr;
  float* crop = nullptr;
  float* padded = nullptr;
  float2* padft = nullptr;
  float2* prod = nullptr;
  float* cc = nullptr;
  bool have_image = false;
  gx::StreamFence fence;
};

namespace gx {
namespace {

void tomo_release(gtomx_tomopicker* tp) {
  if (!tp) return;
  tp->fence.release();
  for (void* q : {(void*)tp->Pref, (void*)tp->Pmask, (void*)tp->imageft, (void*)tp->image2ft, (void*)tp->iw,
                  (void*)tp->T, (void*)tp->mats, (void*)tp->mask2d, (void*)tp->msum, (void*)tp->s1, (void*)tp->s2,
                  (void*)tp->sec, (void*)tp->full, (void*)tp->crop, (void*)tp->padded, (void*)tp->padft,
                  (void*)tp->prod, (void*)tp->cc})
    dev_free(q);
  if (tp->h_mats) (void)hipHostFree(tp->h_mats);
  if (tp->staged) (void)hipEventDestroy(tp->staged);
  delete tp;
}

int tomo_check(const gtomx_tomopicker* tp) {
  if (!tp) return fail(GTOMX_ERR_INVALID_ARGUMENT, "tomo picker is NULL");
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  if (dev != tp->device) return fail(GTOMX_ERR_INVALID_ARGUMENT, "tomo picker was created on another device");
  return GTOMX_OK;
}

// P = R2C(ifftshift(pad(v, p))) into dst (TomoPicker.cu:80-103); tp->full and tp->sec-sized scratch are not
// large enough for p^3, so a scratch of its own is used.
int prepare_projector3(const float* d_v, int nr, int p, float2* dst, hipStream_t s) {
  const size_t Vp3 = (size_t)p * p * p;
  DevBuf a, b;
  GX_TRY(a.alloc(Vp3 * sizeof(float)));
  GX_TRY(b.alloc(Vp3 * sizeof(float)));
  GX_TRY(gtomx_pad(d_v, a.as<float>(), gtomx_int3{nr, nr, nr}, gtomx_int3{p, p, p}, 0.f, 1, s));
  GX_TRY(gtomx_remap_full2fullfft(a.as<float>(), b.as<float>(), gtomx_int3{p, p, p}, 1, s));
  GX_TRY(fft_r2c(b.as<float>(), dst, gtomx_int3{p, p, p}, 1, s));
  GX_HIP(hipStreamSynchronize(s));  // the scratch is freed on return
  return GTOMX_OK;
}

// One orientation (index a) over all tilt images.
int tomo_angle(gtomx_tomopicker* tp, const gtomx_float3& angle, int a, bool masks, float* corr, float3* corrangle,
               int32_t* corridx, float* sum1, float* sum2, gtomx_int3 dc, hipStream_t s) {
  const unsigned N = tp->nimg;
  const gtomx_int3 dp{tp->p, tp->p, 1}, dr{tp->nr, tp->nr, 1}, di{tp->nx, tp->ny, 1};
  // M_i = transpose(Euler(angle)) * transpose(Euler(tilt_i)) (TomoPicker.cu:208-216)
  if (tp->staged_recorded) GX_HIP(hipEventSynchronize(tp->staged));
  double G[9];
  euler_rowmajor(angle, G);
  for (unsigned i = 0; i < N; ++i) {
    double B[9];
    euler_rowmajor(tp->tilt[i], B);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) {
        double acc = 0;
        for (int k = 0; k < 3; ++k) acc += G[3 * k + r] * B[3 * c + k];
        tp->h_mats[9 * i + 3 * r + c] = (float)acc;
      }
  }
  GX_HIP(hipMemcpyAsync(tp->mats, tp->h_mats, N * 9 * sizeof(float), hipMemcpyHostToDevice, s));
  GX_HIP(hipEventRecord(tp->staged, s));
  tp->staged_recorded = true;
  const size_t nHi = tp->Hi * N, nVi = tp->Vi * N;
  if (masks) {
    GX_TRY(project_launch(tp->Pmask, tp->p, tp->sec, tp->p, (unsigned)(tp->p / 2), tp->mats, N, s, 1));
    GX_TRY(fft_c2r(tp->sec, tp->sec, tp->full, dp, N, 1.0f / (float)tp->Vp, s));
    GX_TRY(fft_resize(FFT_FULL_CROP, false, tp->full, tp->mask2d, dp, dr, N, s));
    hipLaunchKernelGGL(k_binarize_sum, dim3(N), dim3(256), 0, s, tp->mask2d, (int)tp->Vr, tp->msum);
    GX_LAUNCHED("k_binarize_sum");
    GX_TRY(fft_resize(FFT_FULL_PAD, false, tp->mask2d, tp->padded, dr, di, N, s));
    GX_TRY(fft_r2c(tp->padded, tp->padft, di, N, s));
    hipLaunchKernelGGL(k_times_conj, dim3(grid_for_n(nHi)), dim3(256), 0, s, tp->imageft, tp->padft, tp->prod, nHi);
    GX_LAUNCHED("k_times_conj");
    GX_TRY(fft_c2r(tp->prod, tp->prod, tp->s1, di, N, 1.0f / (float)tp->Vi, s));
    hipLaunchKernelGGL(k_times_conj, dim3(grid_for_n(nHi)), dim3(256), 0, s, tp->image2ft, tp->padft, tp->prod,
                       nHi);
    GX_LAUNCHED("k_times_conj");
    GX_TRY(fft_c2r(tp->prod, tp->prod, tp->s2, di, N, 1.0f / (float)tp->Vi, s));
  }
  GX_TRY(project_launch(tp->Pref, tp->p, tp->sec, tp->p, (unsigned)(tp->p / 2), tp->mats, N, s, 1));
  GX_TRY(fft_c2r(tp->sec, tp->sec, tp->full, dp, N, 1.0f, s));  // not renormalised (TomoPicker.cu:272)
  GX_TRY(fft_resize(FFT_FULL_CROP, false, tp->full, tp->crop, dp, dr, N, s));
  hipLaunchKernelGGL(k_masked_norm, dim3(N), dim3(256), 0, s, tp->crop, tp->mask2d, (int)tp->Vr);
  GX_LAUNCHED("k_masked_norm");
  GX_TRY(fft_resize(FFT_FULL_PAD, false, tp->crop, tp->padded, dr, di, N, s));
  GX_TRY(fft_r2c(tp->padded, tp->padft, di, N, s));
  hipLaunchKernelGGL(k_times_conj, dim3(grid_for_n(nHi)), dim3(256), 0, s, tp->imageft, tp->padft, tp->prod, nHi);
  GX_LAUNCHED("k_times_conj");
  GX_TRY(fft_c2r(tp->prod, tp->prod, tp->cc, di, N, 1.0f / (float)tp->Vi, s));
  hipLaunchKernelGGL(k_local_cc, dim3(grid_for_n(nVi)), dim3(256), 0, s, tp->cc, tp->s1, tp->s2, tp->msum, tp->Vi,
                     N);
  GX_LAUNCHED("k_local_cc");
  const size_t V = (size_t)dc.x * dc.y * dc.z;
  if (tp->common_row && !tp->ytilt) {
    const bool xmajor = std::fabs(tp->row_a) <= std::fabs(tp->row_b);
    const float slope = xmajor ? -tp->row_a / tp->row_b : -tp->row_b / tp->row_a;
    hipLaunchKernelGGL(k_tomo_update_axis, dim3(grid_for_n(V)), dim3(256), 0, s, tp->cc, tp->nx, tp->ny, (int)N,
                       tp->T, tp->iw, dc.x, dc.y, dc.z, corr, corrangle, corridx, sum1, sum2,
                       make_float3(angle.x, angle.y, angle.z), a, (int)xmajor, slope);
    GX_LAUNCHED("k_tomo_update_axis");
    return GTOMX_OK;
  }
  hipLaunchKernelGGL(k_tomo_update, dim3(grid_for_n(V)), dim3(256), 0, s, tp->cc, tp->nx, tp->ny, (int)N, tp->T,
                     tp->iw, dc.x, dc.y, dc.z, corr, corrangle, corridx, sum1, sum2,
                     make_float3(angle.x, angle.y, angle.z), a);
  GX_LAUNCHED("k_tomo_update");
  return GTOMX_OK;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int gtomx_exact2d_weighting(float* d_weights, gtomx_int3 dimsimage, const int* h_indices, const gtomx_float3* h_angles,
                            const float* d_imageweights, unsigned nimages, float maxfreq, int iszerocentered,
                            unsigned batch, void* stream) {
  if (!d_weights || !h_indices || !h_angles) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL argument");
  if (iszerocentered) return fail(GTOMX_ERR_UNSUPPORTED, "zero-centred weights are not built (FFT order only)");
  if (nimages == 0 || batch == 0 || dimsimage.x < 2 || dimsimage.y < 1 || dimsimage.z != 1)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "bad dims, nimages or batch");
  if (batch > nimages) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch item b takes image b's frame: batch <= nimages");
  for (unsigned b = 0; b < batch; ++b)
    if (h_indices[b] < 0 || h_indices[b] >= (int)nimages) return fail(GTOMX_ERR_INVALID_ARGUMENT, "index range");
  return exact2d_weighting_launch(d_weights, dimsimage.x, dimsimage.y, h_indices, h_angles, d_imageweights, nimages,
                                  maxfreq, batch, as_stream(stream));
}

int gtomx_tomopicker_create(gtomx_tomopicker** out, const float* d_ref, gtomx_int3 dimsref, const float* d_refmask,
                            int ismaskcircular, gtomx_int3 dimsimage, unsigned nimages, void* stream) {
  if (!out) return fail(GTOMX_ERR_INVALID_ARGUMENT, "tomo picker pointer is NULL");
  *out = nullptr;
  if (!d_ref || !d_refmask) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL reference or mask");
  if (dimsref.x != dimsref.y || dimsref.x != dimsref.z) return fail(GTOMX_ERR_UNSUPPORTED, "the reference is a cube");
  if (nimages == 0 || nimages > 128)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "1 to 128 tilt images (c_tiltmatrices, TomoPicker.cu:19)");
  if (dimsimage.z != 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "dimsimage is 2D");
  const int p = 2 * dimsref.x;
  // every transform here is fft_r2c / fft_c2r, which take any extent up to 2048
  GX_TRY(check_fft_dims_any(3, gtomx_int3{p, p, p}));
  GX_TRY(check_fft_dims_any(2, dimsimage));
  if (dimsref.x > dimsimage.x || dimsref.x > dimsimage.y)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "the reference is larger than the images");
  hipStream_t s = as_stream(stream);
  auto* tp = new gtomx_tomopicker();
  tp->nr = dimsref.x;
  tp->p = p;
  tp->nx = dimsimage.x;
  tp->ny = dimsimage.y;
  tp->nimg = nimages;
  tp->circular = ismaskcircular != 0;
  tp->Hp3 = (size_t)(p / 2 + 1) * p * p;
  tp->Hp = (size_t)(p / 2 + 1) * p;
  tp->Vp = (size_t)p * p;
  tp->Vr = (size_t)tp->nr * tp->nr;
  tp->Hi = (size_t)(tp->nx / 2 + 1) * tp->ny;
  tp->Vi = (size_t)tp->nx * tp->ny;
  int st = GTOMX_OK;
  hipError_t he = hipGetDevice(&tp->device);
  if (he != hipSuccess) {
    tomo_release(tp);
    return hip_fail(he, "hipGetDevice");
  }
  auto alloc = [&](void** dst, size_t bytes) {
    if (st == GTOMX_OK) st = dev_alloc(dst, bytes);
  };
  const size_t N = nimages;
  alloc((void**)&tp->Pref, tp->Hp3 * sizeof(float2));
  alloc((void**)&tp->Pmask, tp->Hp3 * sizeof(float2));
  alloc((void**)&tp->imageft, N * tp->Hi * sizeof(float2));
  alloc((void**)&tp->image2ft, N * tp->Hi * sizeof(float2));
  alloc((void**)&tp->iw, N * sizeof(float));
  alloc((void**)&tp->T, N * 6 * sizeof(float));
  alloc((void**)&tp->mats, N * 9 * sizeof(float));
  alloc((void**)&tp->mask2d, N * tp->Vr * sizeof(float));
  alloc((void**)&tp->msum, N * sizeof(double));
  alloc((void**)&tp->s1, N * tp->Vi * sizeof(float));
  alloc((void**)&tp->s2, N * tp->Vi * sizeof(float));
  alloc((void**)&tp->sec, N * tp->Hp * sizeof(float2));
  alloc((void**)&tp->full, N * tp->Vp * sizeof(float));
  alloc((void**)&tp->crop, N * tp->Vr * sizeof(float));
  alloc((void**)&tp->padded, N * tp->Vi * sizeof(float));
  alloc((void**)&tp->padft, N * tp->Hi * sizeof(float2));
  alloc((void**)&tp->prod, N * tp->Hi * sizeof(float2));
  alloc((void**)&tp->cc, N * tp->Vi * sizeof(float));
  if (st == GTOMX_OK) {
    he = hipHostMalloc((void**)&tp->h_mats, N * 9 * sizeof(float), hipHostMallocDefault);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&tp->staged, hipEventDisableTiming);
    if (he != hipSuccess) st = hip_fail(he, "tomo picker host staging");
  }
  if (st == GTOMX_OK) st = prepare_projector3(d_ref, tp->nr, p, tp->Pref, s);
  if (st == GTOMX_OK) st = prepare_projector3(d_refmask, tp->nr, p, tp->Pmask, s);
  if (st != GTOMX_OK) {
    std::string msg = gtomx_last_error();
    tomo_release(tp);
    return fail(st, msg);
  }
  *out = tp;
  return GTOMX_OK;
}

int gtomx_tomopicker_destroy(gtomx_tomopicker* tp) {
  tomo_release(tp);
  return GTOMX_OK;
}

int gtomx_tomopicker_set_image(gtomx_tomopicker* tp, float* d_image, const gtomx_float3* h_imageangles,
                               const float* h_imageweights, void* stream) {
  GX_TRY(tomo_check(tp));
  if (!d_image || !h_imageangles || !h_imageweights) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL argument");
  hipStream_t s = as_stream(stream);
  GX_TRY(tp->fence.enter(s));
  const unsigned N = tp->nimg;
  const gtomx_int3 di{tp->nx, tp->ny, 1};
  tp->tilt.assign(h_imageangles, h_imageangles + N);
  std::vector<float> T(6 * (size_t)N);
  for (unsigned i = 0; i < N; ++i) {
    double E[9];
    euler_rowmajor(tp->tilt[i], E);
    for (int k = 0; k < 6; ++k) T[6 * (size_t)i + k] = (float)E[k];  // rows 0 and 1 (TomoPicker.cu:193-196)
  }
  tp->common_row = true;
  for (unsigned i = 0; i < N; ++i)
    tp->common_row = tp->common_row && T[6 * (size_t)i + 5] == 0.f && T[6 * (size_t)i + 3] == T[3] &&
                     T[6 * (size_t)i + 4] == T[4];
  tp->ytilt = tp->common_row && T[3] == 0.f && T[4] == 1.f;
  tp->row_a = T[3];
  tp->row_b = T[4];
  GX_HIP(hipMemcpyAsync(tp->T, T.data(), T.size() * sizeof(float), hipMemcpyHostToDevice, s));
  GX_HIP(hipMemcpyAsync(tp->iw, h_imageweights, N * sizeof(float), hipMemcpyHostToDevice, s));
  // filter = Exact2DWeighting(1; tilts, weights, maxfreq = nx) (:168-176); images <- C2R_n(R2C(images) * filter)
  const size_t nHi = tp->Hi * N;
  float* filter = reinterpret_cast<float*>(tp->prod);  // nimg x Hi floats fit in the product buffer
  GX_TRY(fill_f32(filter, nHi, 1.f, s));
  std::vector<int> ind(N);
  for (unsigned i = 0; i < N; ++i) ind[i] = (int)i;
  GX_TRY(exact2d_weighting_launch(filter, tp->nx, tp->ny, ind.data(), h_imageangles, tp->iw, N, (float)tp->nx, N, s));
  GX_TRY(fft_r2c(d_image, tp->imageft, di, N, s));
  hipLaunchKernelGGL(k_mul_real, dim3(grid_for_n(nHi)), dim3(256), 0, s, tp->imageft, filter, nHi);
  GX_LAUNCHED("k_mul_real");
  GX_TRY(fft_c2r(tp->imageft, tp->padft, d_image, di, N, 1.0f / (float)tp->Vi, s));
  // the spectra of the images as stored, and of their squares (:185-186)
  GX_TRY(fft_r2c(d_image, tp->imageft, di, N, s));
  hipLaunchKernelGGL(k_square_f, dim3(grid_for_n(tp->Vi * N)), dim3(256), 0, s, d_image, tp->padded, tp->Vi * N);
  GX_LAUNCHED("k_square_f");
  GX_TRY(fft_r2c(tp->padded, tp->image2ft, di, N, s));
  tp->have_image = true;
  return tp->fence.leave(s);
}

int gtomx_tomopicker_correlate_angles(gtomx_tomopicker* tp, const gtomx_float3* h_angles, unsigned nangles,
                                      float* d_corr, gtomx_float3* d_corrangle, gtomx_int3 dimscorr,
                                      int32_t* d_corrindex, float* d_corrsum1, float* d_corrsum2, void* stream) {
  GX_TRY(tomo_check(tp));
  if (!tp->have_image) return fail(GTOMX_ERR_INVALID_ARGUMENT, "SetImage has not been called");
  if (!d_corr) return fail(GTOMX_ERR_INVALID_ARGUMENT, "d_corr is NULL");
  if (nangles && !h_angles) return fail(GTOMX_ERR_INVALID_ARGUMENT, "h_angles is NULL");
  if (dimscorr.x < 1 || dimscorr.y < 1 || dimscorr.z < 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "bad dimscorr");
  if (nangles > 0x7fffffffu) return fail(GTOMX_ERR_INVALID_ARGUMENT, "orientation index exceeds int32");
  hipStream_t s = as_stream(stream);
  GX_TRY(tp->fence.enter(s));
  const size_t V = (size_t)dimscorr.x * dimscorr.y * dimscorr.z;
  // the sums start at zero (CudaMallocValueFilled, TomoPicker.cu:202-203)
  if (d_corrsum1) GX_TRY(fill_f32(d_corrsum1, V, 0.f, s));
  if (d_corrsum2) GX_TRY(fill_f32(d_corrsum2, V, 0.f, s));
  for (unsigned a = 0; a < nangles; ++a)
    GX_TRY(tomo_angle(tp, h_angles[a], (int)a, a == 0 || !tp->circular, d_corr,
                      reinterpret_cast<float3*>(d_corrangle), d_corrindex, d_corrsum1, d_corrsum2, dimscorr, s));
  return tp->fence.leave(s);
}

int gtomx_tomopicker_correlate(gtomx_tomopicker* tp, float* d_corr, gtomx_float3* d_corrangle, gtomx_int3 dimscorr,
                               float anglestep, void* stream) {
  GX_TRY(tomo_check(tp));
  if (!(anglestep > 0.f) || !std::isfinite(anglestep))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "anglestep must be positive");
  // GetEqualAngularSpacing((0, 2pi), (0, pi), (0, 2pi), anglestep) (TomoPicker.cu:190-191)
  const float PI2 = 6.283185307179586f, PI = 3.14159265358979f;
  int n = 0;
  GX_TRY(gtomx_equal_angular_spacing(0.f, PI2, 0.f, PI, 0.f, PI2, anglestep, nullptr, 0, &n));
  std::vector<gtomx_float3> angles((size_t)n);
  GX_TRY(gtomx_equal_angular_spacing(0.f, PI2, 0.f, PI, 0.f, PI2, anglestep, angles.data(), n, &n));
  return gtomx_tomopicker_correlate_angles(tp, angles.data(), (unsigned)n, d_corr, d_corrangle, dimscorr, nullptr,
                                           nullptr, nullptr, stream);
}

}  // extern "C"
