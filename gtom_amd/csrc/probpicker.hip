// ProbabilisticPicker (include/Correlation.cuh:166-207,
// src/Correlation/ProbabilisticPicker.cu:29-455), 2D and 3D.
//
// Initialize (:33-152), per reference n, on the device:
//   R_n = ifftshift(pad(ref_n, dimsimage, 0)), C_n = FFTCrop(R2C(R_n), 2*lowpass),
//   C_n *= [|k| < lowpass] (d_Bandpass, :136), then d_RemapHalfFFT2Half (:138).
//   That remap (src/FFT/FFTRemap.cu, RemapHalfFFT2HalfKernel) reads row
//   signed(i) + n/2 for output row i, which is n (one past the array) for the
//   zero-frequency row and plane of the centred output; those reads land in the
//   next plane's DC row or past the cropped spectrum, where the reference's buffer
//   still holds R_n (the crop was written over its head). Both are restated here:
//   element j >= ElementsFFT(2*lowpass) is the float pair (R_n[2j], R_n[2j+1]).
//   Masks: M_n = ifftshift(pad(mask_n)), IM_n = R2C(ifftshift(pad(1 - mask_n)));
//   a non-circular mask is kept as the cubic B-spline coefficients of pad(mask_n)
//   (d_CubicBSplinePrefilter3D/2D, :104-110) and rotated into FFT order with each
//   orientation (d_Rotate3D / d_Rotate2D, T_INTERP_CUBIC, :273-279), whose crop
//   then gives that orientation's inverse mask and solvent statistics (:342-356).
// SetImage (:205-224): I = R2C(mean0std1(image)) and I2 = R2C(mean0std1(image)^2),
//   both band-limited to |k| < lowpass; the CTF cropped to 2*lowpass.
// PerformCorrelation (:226-389), orientations B at a time:
//   rot = Rotate{3D,2D}FT(C_n, angle) in FFT order, * ctf_low if doctf,
//   r = C2R(FFTPad(rot, dimsimage)) / N * M_n, corr = C2R_u(I * conj(R2C(r))),
//   sums over the FFTFullCrop of r and M_n to dimsref (in the crop's order):
//   samples = sum m, mean = sum r / samples, mean2 = sum r^2 / samples,
//   ratio = exp((sum_i 5 r_i^2 + 2 r_i G_i) / 5 / (2 samples)) with G_i the sum
//   of the five gaussrand() draws of element i (src/Helpers/Misc.cpp:14-31) in
//   the reference's order. The C runtime's rand() integers after srand(123) are
//   produced on the host (glibc's additive generator, restated and checked
//   against rand() itself; rand() is called directly where they differ) and
//   uploaded per batch while the device works on the previous one; the
//   Box-Muller transform runs on the device inside the reduction;
//   solvent mean/std from I, I2 and IM_n once per call (CalcSolventStatistics,
//   :423-441, ProbSolventStatisticsKernel :443-454);
//   score (ProbPickingUpdateKernel :391-421) in double, kept where it beats the
//   stored float, orientations in increasing order with strict '>'.
// Nothing goes back to the host per orientation: the reference's per-orientation
// device-to-host copies and host sums become one reduction kernel.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <thread>
#include <vector>

#include "dev_bspline.hpp"
#include "dev_fft.hpp"
#include "runtime.hpp"

namespace gx {
namespace {

constexpr int NBS = 64;  // reduction workgroups per orientation for the crop sums

// Per-orientation scalars of the score, as the reference passes them.
struct ProbCoef {
  double ratio;       // expectedpratio
  float normfactor;   // 1 / (N * samples), float arithmetic
  float smr, smr2;    // summaskedref, summaskedref2 rounded to float
  float solventnorm;  // 1 / (N * (Er - samples)), float arithmetic
};

__device__ __forceinline__ int signed_freq(int i, int n) { return i < n / 2 + 1 ? i : i - n; }

// d_Bandpass(low 0, high r, smooth 0) on FFT-order half spectra: keeps |k| < r.
__global__ __launch_bounds__(256) void k_bandpass(float2* __restrict__ x, int3 d, int r, size_t per, unsigned batch) {
  const int nh = d.x / 2 + 1;
  const long long r2 = (long long)r * r;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < per * batch; i += (size_t)gridDim.x * 256) {
    const size_t e = i % per;
    const int kx = (int)(e % nh), ky = signed_freq((int)((e / nh) % d.y), d.y),
              kz = signed_freq((int)(e / ((size_t)nh * d.y)), d.z);
    const long long q = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
    if (!(q < r2)) x[i] = make_float2(0.f, 0.f);
  }
}

// RemapHalfFFT2HalfKernel as the reference runs it on its buffer: C (FFT-order
// crop, `per` complex values) followed by the tail of R viewed as float pairs.
__global__ __launch_bounds__(256) void k_remap_half_quirk(const float2* __restrict__ C, const float* __restrict__ R,
                                                          float2* __restrict__ out, int n, int nz, size_t per,
                                                          size_t vimage, unsigned batch) {
  const int nh = n / 2 + 1;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < per * batch; i += (size_t)gridDim.x * 256) {
    const size_t b = i / per, e = i % per;
    const int x = (int)(e % nh), y = (int)((e / nh) % n), z = (int)(e / ((size_t)nh * n));
    const int rz = (z < nz / 2 + 1 ? z : z - n) + nz / 2;
    const int ry = (y < n / 2 + 1 ? y : y - n) + n / 2;
    const size_t j = ((size_t)rz * n + ry) * nh + x;
    out[i] = j < per ? C[b * per + j] : reinterpret_cast<const float2*>(R + b * vimage)[j];
  }
}

__global__ __launch_bounds__(256) void k_one_minus(const float* __restrict__ in, float* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) out[i] = 1.0f - in[i];
}

__global__ __launch_bounds__(256) void k_times_real(float2* __restrict__ x, const float* __restrict__ v, size_t per,
                                                    unsigned batch) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < per * batch; i += (size_t)gridDim.x * 256) {
    const float f = v[i % per];
    const float2 a = x[i];
    x[i] = make_float2(a.x * f, a.y * f);
  }
}

// x[b][i] *= m[b * mstride + i] (mstride 0: one mask for every item)
__global__ __launch_bounds__(256) void k_times_mask(float* __restrict__ x, const float* __restrict__ m, size_t per,
                                                    unsigned batch, size_t mstride) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < per * batch; i += (size_t)gridDim.x * 256)
    x[i] *= m[(i / per) * mstride + i % per];
}

// The mask product on the rows of a RowBand only (vrows of them, nx values
// each; one mask shared by all volumes): the banded transforms read no other
// rows.
__global__ __launch_bounds__(256) void k_times_mask_band(float* __restrict__ x, const float* __restrict__ m, int nx,
                                                         long long vrows, RowBand rb, long long plane_rows) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < vrows * nx; i += (long long)gridDim.x * 256) {
    const long long row = band_row(i / nx, rb);
    const int c = (int)(i % nx);
    x[row * nx + c] *= m[(row % plane_rows) * nx + c];
  }
}

// out = a * conj(b) (d_ComplexMultiplyByConjVector, src/Generics/Arithmetics.cu:334-344)
__global__ __launch_bounds__(256) void k_mul_conj_into(const float2* __restrict__ a, const float2* __restrict__ b,
                                                       float2* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const float2 x = a[i], y = b[i];
    out[i] = make_float2(x.x * y.x + x.y * y.y, x.y * y.x - x.x * y.y);
  }
}

// d_Rotate3D / d_Rotate2D with T_INTERP_CUBIC into FFT order (Rotate3DKernel,
// Rotate2DKernel<true, false>, src/Transformation/Rotation.cu:386-418, 450-480):
// output voxel i takes x = (i + n/2) % n per axis, p = M (x - n/2) + n/2 + 0.5
// (3D: M = transpose(Euler), 2D: M = R(-angle) from its cos/sin pair), and the
// cubic B-spline of the prefiltered coefficients at p - 0.5 over the 4^ndims
// taps, indices clamped, summed z, y, x as cubicTex3DSimple (include/CubicInterp.cuh:287-312).
__global__ __launch_bounds__(256) void k_rotate_cubic(const float* __restrict__ coef, float* __restrict__ out,
                                                      int3 d, const float* __restrict__ mats, int ndims) {
  const size_t V = (size_t)d.x * d.y * d.z;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= V) return;
  const int b = blockIdx.y;
  const int ix = (int)(i % d.x), iy = (int)((i / d.x) % d.y), iz = (int)(i / ((size_t)d.x * d.y));
  const float cx = (float)((ix + d.x / 2) % d.x - d.x / 2), cy = (float)((iy + d.y / 2) % d.y - d.y / 2),
              cz = (float)((iz + d.z / 2) % d.z - d.z / 2);
  float px, py, pz = 0.f;
  if (ndims == 3) {
    const float* M = mats + 9 * b;
    px = M[0] * cx + M[1] * cy + M[2] * cz;
    py = M[3] * cx + M[4] * cy + M[5] * cz;
    pz = M[6] * cx + M[7] * cy + M[8] * cz;
  } else {
    const float c = mats[2 * b], sn = mats[2 * b + 1];
    px = c * cx - sn * cy;
    py = sn * cx + c * cy;
  }
  // p - 0.5 with p = M(x - n/2) + n/2 + 0.5
  const float gx = px + (float)(d.x / 2), gy = py + (float)(d.y / 2), gz = pz + (float)(d.z / 2);
  const float fx0 = floorf(gx), fy0 = floorf(gy), fz0 = floorf(gz);
  const float fx = gx - fx0, fy = gy - fy0, fz = gz - fz0;
  auto clampi = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  float r = 0.f;
  const int zlo = ndims == 3 ? -1 : 0, zhi = ndims == 3 ? 2 : 0;
  for (int tz = zlo; tz <= zhi; ++tz) {
    const float wz = ndims == 3 ? bspline((float)tz - fz) : 1.0f;
    const int zz = ndims == 3 ? clampi((int)fz0 + tz, d.z - 1) : 0;
    for (int ty = -1; ty <= 2; ++ty) {
      const float wyz = bspline((float)ty - fy) * wz;
      const int yy = clampi((int)fy0 + ty, d.y - 1);
      const float* row = coef + ((size_t)zz * d.y + yy) * d.x;
      for (int tx = -1; tx <= 2; ++tx) r += bspline((float)tx - fx) * wyz * row[clampi((int)fx0 + tx, d.x - 1)];
    }
  }
  out[(size_t)b * V + i] = r;
}

// om[b][i] = 1 - m[b][src(i)] over the FFTFullCrop to dimsref (d_OneMinus of the
// cropped rotated mask, src/Correlation/ProbabilisticPicker.cu:346-348).
__global__ __launch_bounds__(256) void k_crop_one_minus(const float* __restrict__ m, float* __restrict__ om, int3 di,
                                                        int3 dr, size_t vimage, size_t er) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= er) return;
  const unsigned b = blockIdx.y;
  const int x = (int)(i % dr.x), y = (int)((i / dr.x) % dr.y), z = (int)(i / ((size_t)dr.x * dr.y));
  const int sx = x >= dr.x / 2 ? x + di.x - dr.x : x, sy = y >= dr.y / 2 ? y + di.y - dr.y : y,
            sz = z >= dr.z / 2 ? z + di.z - dr.z : z;
  om[(size_t)b * er + i] = 1.0f - m[(size_t)b * vimage + ((size_t)sz * di.y + sy) * di.x + sx];
}

// out[b][i] = a[i] * conj(bb[b][i])
__global__ __launch_bounds__(256) void k_mul_conj_batch(const float2* __restrict__ a, const float2* __restrict__ bb,
                                                        float2* __restrict__ out, size_t n, unsigned batch) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n * batch; i += (size_t)gridDim.x * 256) {
    const float2 x = a[i % n], y = bb[i];
    out[i] = make_float2(x.x * y.x + x.y * y.y, x.y * y.x - x.x * y.y);
  }
}

__global__ __launch_bounds__(256) void k_square_into(const float* __restrict__ in, float* __restrict__ out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) out[i] = in[i] * in[i];
}

// gaussrand draws of one call: draw g (from the call's first) has t = g + p0
// (p0 the pair phase the call starts in); pair q = t / 2 gives sin (t even) or
// cos (t odd) of the Box-Muller pair (U, V). Pair q < p0 is the one a previous
// call left half used (its U, V passed in); pair q >= p0 takes the rand()
// integers raw[q - qlo], U = (a + 1) / (RAND_MAX + 2), V = b / (RAND_MAX + 1),
// rounded to float, Z = sqrt(-2 log U) * {sin, cos}(2 pi V) in float.
struct Draws {
  const int2* raw;
  long long qlo;
  int p0;
  float Us, Vs;
  double rmax;
};

__device__ __forceinline__ float gauss_draw(const Draws& d, long long g) {
  const long long t = g + d.p0;
  const long long q = t >> 1;
  float U = d.Us, V = d.Vs;
  if (q >= d.p0) {
    const int2 w = d.raw[q - d.qlo];
    U = (float)(((double)w.x + 1.) / (d.rmax + 2.));
    V = (float)((double)w.y / (d.rmax + 1.));
  }
  const float m = sqrtf(-2 * logf(U));
  const float a = 2 * 3.1415926535897932384626433832795f * V;
  return (t & 1) ? m * cosf(a) : m * sinf(a);
}

// Sums over the FFTFullCrop (src/FFT/FFTResize.cu:140-160) of r and the mask to
// dimsref, element i of the crop of orientation b weighted by the sum of its
// five draws (draw index (g0 + b * er + i) * 5 + k, g0 = first orientation * er):
// part[b][blk] = (sum m, sum r, sum r^2, sum r*G).
__global__ __launch_bounds__(256) void k_prob_sums(const float* __restrict__ r, const float* __restrict__ mask,
                                                   size_t mstride, Draws draws, long long g0, int3 di, int3 dr,
                                                   size_t vimage, size_t er, double4* __restrict__ part) {
  __shared__ double4 red[4];
  const unsigned b = blockIdx.y;
  double sm = 0, sr = 0, srr = 0, srg = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < er; i += (size_t)gridDim.x * 256) {
    const int x = (int)(i % dr.x), y = (int)((i / dr.x) % dr.y), z = (int)(i / ((size_t)dr.x * dr.y));
    const int sx = x >= dr.x / 2 ? x + di.x - dr.x : x, sy = y >= dr.y / 2 ? y + di.y - dr.y : y,
              sz = z >= dr.z / 2 ? z + di.z - dr.z : z;
    const size_t src = ((size_t)sz * di.y + sy) * di.x + sx;
    const double rv = r[(size_t)b * vimage + src];
    const long long gbase = (g0 + (long long)b * (long long)er + (long long)i) * 5;
    double G = 0;
    for (int k = 0; k < 5; ++k) G += (double)gauss_draw(draws, gbase + k);
    sm += mask[(size_t)b * mstride + src];
    sr += rv;
    srr += rv * rv;
    srg += rv * G;
  }
  for (int o = 32; o > 0; o >>= 1) {
    sm += __shfl_down(sm, o);
    sr += __shfl_down(sr, o);
    srr += __shfl_down(srr, o);
    srg += __shfl_down(srg, o);
  }
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = make_double4(sm, sr, srr, srg);
  __syncthreads();
  if (threadIdx.x == 0) {
    double4 t = red[0];
    for (int w = 1; w < 4; ++w) {
      t.x += red[w].x;
      t.y += red[w].y;
      t.z += red[w].z;
      t.w += red[w].w;
    }
    part[(size_t)b * gridDim.x + blockIdx.x] = t;
  }
}

// The host block of PerformCorrelation (:306-340) and the launch arguments
// (:366-379) from the sums, one thread per orientation.
__global__ void k_prob_coef(const double4* __restrict__ part, int nparts, unsigned batch, size_t vimage, size_t er,
                            ProbCoef* __restrict__ coef) {
  const unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  double sm = 0, sr = 0, srr = 0, srg = 0;
  for (int p = 0; p < nparts; ++p) {
    const double4 t = part[(size_t)b * nparts + p];
    sm += t.x;
    sr += t.y;
    srr += t.z;
    srg += t.w;
  }
  ProbCoef c;
  const double pr = (5.0 * srr + 2.0 * srg) / 5.0;
  c.ratio = exp(pr / (2.0 * sm));
  c.smr = (float)(sr / sm);
  c.smr2 = (float)(srr / sm);
  const float masksum = (float)sm;                   // h_masksum[n] is a tfloat
  const float invmasksum = (float)((double)er - sm);  // h_invmasksum[n] likewise
  c.normfactor = 1.0f / ((float)vimage * masksum);
  c.solventnorm = 1.0f / ((float)vimage * invmasksum);
  coef[b] = c;
}

// ProbSolventStatisticsKernel: mean = m * f, std = sqrt(max(0, s * f - mean^2)), float.
// (item b of `batch` uses orientation b's factor)
__global__ __launch_bounds__(256) void k_solvent_stats(float* __restrict__ mean, float* __restrict__ sd, size_t n,
                                                       const ProbCoef* __restrict__ coef, unsigned batch) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n * batch; i += (size_t)gridDim.x * 256) {
    const float f = coef[i / n].solventnorm;
    const float m = mean[i] * f;
    mean[i] = m;
    const float v = sd[i] * f - m * m;
    sd[i] = sqrtf(v < 0.f ? 0.f : v);
  }
}

// ProbPickingUpdateKernel over B maps in orientation order.
// (solvent statistics of item b at b * sstride: 0 when one set serves every orientation)
__global__ __launch_bounds__(256) void k_prob_update(const float* __restrict__ corr, const float* __restrict__ smean,
                                                     const float* __restrict__ sstd, size_t sstride,
                                                     const ProbCoef* __restrict__ coef,
                                                     const float3* __restrict__ angles, unsigned batch, size_t n,
                                                     int refid, float* __restrict__ bestccf,
                                                     float3* __restrict__ bestangle, int* __restrict__ bestref) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    float best = bestccf[i];
    int k = -1;
    for (unsigned b = 0; b < batch; ++b) {
      const ProbCoef c = coef[b];
      const double m = smean[b * sstride + i];
      const double sd = fmax(1e-7, (double)sstd[b * sstride + i]);
      double d = -2.0 * (double)c.normfactor * (double)corr[(size_t)b * n + i];
      d += 2.0 * m * (double)c.smr;
      d /= sd;
      d += (double)c.smr2;
      d = exp(-d / 2.0);
      d = (d - 1.0) / (c.ratio - 1.0);
      if (d > (double)best) {
        best = (float)d;
        k = (int)b;
      }
    }
    if (k >= 0) {
      bestccf[i] = best;
      bestangle[i] = angles[k];
      bestref[i] = refid;
    }
  }
}

inline unsigned grid_of(size_t n, size_t cap = 16384) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, cap)); }

// The pair state of gaussrand (src/Helpers/Misc.cpp:14-31): its function-scope
// statics U, V (tfloat) and phase live for the process and are shared by every
// picker, as there.
struct GaussState {
  float U = 0, V = 0;
  int phase = 0;
};
GaussState g_gauss;
std::mutex g_gauss_mutex;

// glibc's rand(): the additive generator of random_r with a 31-word state
// seeded by srandom_r (word 0 = seed, word i = 16807 * word i-1 mod 2^31 - 1 by
// Schrage's method) and 310 outputs discarded. With r_i its words (r_0..r_30 from
// the seeding, r_31..r_33 = r_0..r_2, r_i = r_(i-31) + r_(i-3) from i = 34), output
// k is r_(k+344) >> 1, and every r_(3+e) is the combination of the words r_3..r_33
// whose coefficients are those of x^e mod (x^31 - x^28 - 1) over Z/2^32, so any
// stretch of the sequence can be produced on its own.
struct AdditiveJump {
  uint32_t base[31];  // r_3 .. r_33
  explicit AdditiveJump(unsigned seed) {
    uint32_t r[34];
    int32_t word = seed == 0 ? 1 : (int32_t)seed;
    r[0] = (uint32_t)word;
    for (int i = 1; i < 31; ++i) {
      const long long hi = word / 127773, lo = word % 127773;
      long long w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      word = (int32_t)w;
      r[i] = (uint32_t)word;
    }
    for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
    for (int i = 0; i < 31; ++i) base[i] = r[3 + i];
  }
  static void mulmod(const uint32_t* a, const uint32_t* b, uint32_t* out) {
    uint32_t c[61] = {0};
    for (int i = 0; i < 31; ++i)
      for (int j = 0; j < 31; ++j) c[i + j] += a[i] * b[j];
    for (int d = 60; d >= 31; --d) {  // x^31 = x^28 + 1
      c[d - 3] += c[d];
      c[d - 31] += c[d];
    }
    for (int i = 0; i < 31; ++i) out[i] = c[i];
  }
  // Fills n outputs starting at output index k (k >= 0).
  void outputs(long long k, long long n, int* out) const {
    const long long m = k + 344 - 31;  // the first word of the window before output k
    uint32_t res[31] = {1}, sq[31] = {0};
    sq[1] = 1;  // x
    for (long long e = m - 3; e > 0; e >>= 1) {
      uint32_t t[31];
      if (e & 1) {
        mulmod(res, sq, t);
        std::copy(t, t + 31, res);
      }
      mulmod(sq, sq, t);
      std::copy(t, t + 31, sq);
    }
    uint32_t w[31];  // r_m .. r_(m+30)
    uint32_t q[31];
    std::copy(res, res + 31, q);
    for (int j = 0; j < 31; ++j) {
      uint32_t v = 0;
      for (int i = 0; i < 31; ++i) v += q[i] * base[i];
      w[j] = v;
      const uint32_t top = q[30];  // q <- x q mod P
      for (int i = 30; i > 0; --i) q[i] = q[i - 1];
      q[0] = top;
      q[28] += top;
    }
    int f = 0;  // w is a ring: the word 31 back sits at f, the word 3 back at f + 28
    for (long long i = 0; i < n; ++i) {
      const uint32_t v = w[f] + w[(f + 28) % 31];
      w[f] = v;
      out[i] = (int)(v >> 1);
      if (++f == 31) f = 0;
    }
  }
};

// Whether the C runtime's rand() is that generator: compared over 1000
// outputs after srand(123) (the only seed used), once per process.
bool runtime_rand_is_additive() {
  static int known = -1;
  if (known < 0) {
    std::vector<int> mine(1000);
    AdditiveJump(123).outputs(0, 1000, mine.data());
    std::srand(123);
    bool same = RAND_MAX == 2147483647;
    for (int i = 0; i < 1000 && same; ++i) same = std::rand() == mine[(size_t)i];
    known = same ? 1 : 0;
  }
  return known == 1;
}

}  // namespace
}  // namespace gx

struct gtomx_prob_picker {
  gtomx_int3 dr{}, di{}, dl{};
  int ndims = 3;
  unsigned nrefs = 0;
  bool circular = true, doctf = false;
  int device = 0;
  unsigned B = 0;
  size_t Er = 0, Vi = 0, Hi = 0, Hl = 0;
  float2* refs_low = nullptr;    // nrefs x Hl, zero-centred half spectra
  float* masks = nullptr;        // nrefs x Vi: FFT order (circular), or centred B-spline coefficients
  float2* invmasksft = nullptr;  // nrefs x Hi (circular)
  float* mrot = nullptr;         // B x Vi rotated masks in FFT order (non-circular)
  float2* imft = nullptr;        // B x Hi inverse-mask spectra (non-circular)
  float* om = nullptr;           // B x Er cropped inverse masks (non-circular)
  float2* imageft = nullptr;
  float2* image2ft = nullptr;
  float* ctf_low = nullptr;      // Hl
  float2* rot = nullptr;         // B x Hl
  float2* padft = nullptr;       // B x Hi: padded spectra, C2R workspace, then R2C of r
  float* real = nullptr;         // B x Vi: r, then the correlation maps
  float* smean = nullptr;        // Vi (circular) or B x Vi
  float* sstd = nullptr;         // Vi (circular) or B x Vi
  int2* raw = nullptr;           // rand() pairs of one batch
  double4* part = nullptr;       // B x NBS
  gx::ProbCoef* coef = nullptr;  // B
  float* mats = nullptr;         // B x 9 (3D) or B x 2 (2D)
  float3* dang = nullptr;        // B
  double* nscr = nullptr;
  int2* h_raw[2] = {nullptr, nullptr};     // pinned rand() pairs of one batch each
  size_t raw_cap = 0;                      // pairs per batch buffer
  float* h_stage[2] = {nullptr, nullptr};  // pinned, B x 12 floats each: matrices then angles
  hipEvent_t uploaded[2] = {nullptr, nullptr};
  bool have_image = false;
};

namespace gx {
namespace {

void prob_release(gtomx_prob_picker* pp) {
  if (!pp) return;
  for (void* p : {(void*)pp->refs_low, (void*)pp->masks, (void*)pp->invmasksft, (void*)pp->imageft,
                  (void*)pp->image2ft, (void*)pp->ctf_low, (void*)pp->rot, (void*)pp->padft, (void*)pp->real,
                  (void*)pp->smean, (void*)pp->sstd, (void*)pp->raw, (void*)pp->part, (void*)pp->coef,
                  (void*)pp->mats, (void*)pp->dang, (void*)pp->nscr, (void*)pp->mrot, (void*)pp->imft,
                  (void*)pp->om})
    dev_free(p);
  for (int k = 0; k < 2; ++k) {
    if (pp->h_raw[k]) (void)hipHostFree(pp->h_raw[k]);
    if (pp->h_stage[k]) (void)hipHostFree(pp->h_stage[k]);
    if (pp->uploaded[k]) (void)hipEventDestroy(pp->uploaded[k]);
  }
  delete pp;
}

int prob_check(const gtomx_prob_picker* pp) {
  if (!pp) return fail(GTOMX_ERR_INVALID_ARGUMENT, "picker is NULL");
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  if (dev != pp->device) return fail(GTOMX_ERR_INVALID_ARGUMENT, "picker was created on another device");
  return GTOMX_OK;
}

// Highest complex index the reference's half remap reads for a 2*lowpass square
// or cube (the zero-frequency plane reads plane n, the zero-frequency row row n).
size_t remap_quirk_max_index(int n, int nz) {
  const size_t nh = (size_t)(n / 2 + 1);
  const int rzmax = nz > 1 ? n : 0;
  return ((size_t)rzmax * n + n) * nh + (nh - 1);
}

// The orientation list of PerformCorrelation (:234-244).
int prob_angles(int ndims, float anglestep, std::vector<gtomx_float3>& out) {
  const float PIf = 3.1415926535897932384626433832795f, PI2f = 6.283185307179586476925286766559f;
  if (!(anglestep > 0.0f)) return fail(GTOMX_ERR_INVALID_ARGUMENT, "anglestep must be positive");
  if (ndims == 3) {
    int count = 0;
    GX_TRY(gtomx_equal_angular_spacing(0, PI2f, 0, PIf, 0, PI2f, anglestep, nullptr, 0, &count));
    out.resize((size_t)count);
    GX_TRY(gtomx_equal_angular_spacing(0, PI2f, 0, PIf, 0, PI2f, anglestep, out.data(), count, &count));
  } else {
    const int steps = (int)std::ceil(PI2f / anglestep);
    const float step = PI2f / (float)steps;
    out.resize((size_t)steps);
    for (int a = 0; a < steps; ++a) out[(size_t)a] = gtomx_float3{0.f, 0.f, (float)a * step};
  }
  return GTOMX_OK;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int gtomx_prob_picker_create(gtomx_prob_picker** out, const float* d_refs, gtomx_int3 dimsref, unsigned nrefs,
                             const float* d_refmasks, int ismaskcircular, int doctf, gtomx_int3 dimsimage,
                             unsigned lowpassfreq, void* stream) {
  if (!out) return fail(GTOMX_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  if (!d_refs || !d_refmasks) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (nrefs == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "nrefs must be positive");
  const int nd = dimsimage.z > 1 ? 3 : 2;
  if ((dimsref.z > 1 ? 3 : 2) != nd) return fail(GTOMX_ERR_INVALID_ARGUMENT, "reference and image dimensionality differ");
  GX_TRY(check_fft_dims_any(nd, dimsimage));  // non-native extents: chirp-z, full transforms
  if (dimsref.x < 1 || dimsref.y < 1 || dimsref.z < 1 || dimsref.x > dimsimage.x || dimsref.y > dimsimage.y ||
      dimsref.z > dimsimage.z)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "reference dims must be positive and fit the image");
  const int n = 2 * (int)lowpassfreq;
  if (lowpassfreq == 0 || n > dimsimage.x || n > dimsimage.y || (nd == 3 && n > dimsimage.z))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "2*lowpassfreq must be positive and fit the image");
  const gtomx_int3 dl{n, n, nd == 3 ? n : 1};
  const size_t Vi = elements(dimsimage), Hl = elements_ft(dl);
  if (2 * remap_quirk_max_index(n, dl.z) + 2 > Vi)
    return fail(GTOMX_ERR_UNSUPPORTED,
                "the image is too small next to 2*lowpassfreq: the reference's half-spectrum remap would read "
                "memory it never wrote");
  if (nrefs > 65535) return fail(GTOMX_ERR_UNSUPPORTED, "more than 65535 references");

  auto* pp = new gtomx_prob_picker();
  pp->dr = dimsref;
  pp->di = dimsimage;
  pp->dl = dl;
  pp->ndims = nd;
  pp->nrefs = nrefs;
  pp->circular = ismaskcircular != 0;
  pp->doctf = doctf != 0;
  pp->Er = elements(dimsref);
  pp->Vi = Vi;
  pp->Hi = elements_ft(dimsimage);
  pp->Hl = Hl;
  hipStream_t s = as_stream(stream);
  int st = GTOMX_OK;
  auto alloc = [&](void** p, size_t bytes) {
    if (st == GTOMX_OK) st = dev_alloc(p, bytes);
  };
  st = hipGetDevice(&pp->device) == hipSuccess ? GTOMX_OK : fail(GTOMX_ERR_DEVICE, "hipGetDevice");
  // orientations per batch: about 2 GiB of per-orientation buffers, at most 32
  const size_t per = Hl * 8 + pp->Hi * 8 + Vi * 4 + pp->Er * 20 +
                     (pp->circular ? 0 : Vi * 12 + pp->Hi * 8 + pp->Er * 4);
  pp->B = (unsigned)std::max<size_t>(1, std::min<size_t>(32, ((size_t)2 << 30) / per));
  const unsigned B = pp->B;
  alloc((void**)&pp->refs_low, (size_t)nrefs * Hl * sizeof(float2));
  alloc((void**)&pp->masks, (size_t)nrefs * Vi * sizeof(float));
  if (pp->circular) {
    alloc((void**)&pp->invmasksft, (size_t)nrefs * pp->Hi * sizeof(float2));
  } else {
    alloc((void**)&pp->mrot, (size_t)B * Vi * sizeof(float));
    alloc((void**)&pp->imft, (size_t)B * pp->Hi * sizeof(float2));
    alloc((void**)&pp->om, (size_t)B * pp->Er * sizeof(float));
  }
  alloc((void**)&pp->imageft, pp->Hi * sizeof(float2));
  alloc((void**)&pp->image2ft, pp->Hi * sizeof(float2));
  alloc((void**)&pp->ctf_low, Hl * sizeof(float));
  alloc((void**)&pp->rot, (size_t)B * Hl * sizeof(float2));
  alloc((void**)&pp->padft, (size_t)B * pp->Hi * sizeof(float2));
  alloc((void**)&pp->real, (size_t)B * Vi * sizeof(float));
  alloc((void**)&pp->smean, (pp->circular ? 1 : (size_t)B) * Vi * sizeof(float));
  alloc((void**)&pp->sstd, (pp->circular ? 1 : (size_t)B) * Vi * sizeof(float));
  pp->raw_cap = ((size_t)B * pp->Er * 5 + 3) / 2 + 1;  // pairs touched by B orientations' draws
  alloc((void**)&pp->raw, pp->raw_cap * sizeof(int2));
  alloc((void**)&pp->part, (size_t)B * NBS * sizeof(double4));
  alloc((void**)&pp->coef, (size_t)B * sizeof(ProbCoef));
  alloc((void**)&pp->mats, (size_t)B * 9 * sizeof(float));
  alloc((void**)&pp->dang, (size_t)B * sizeof(float3));
  alloc((void**)&pp->nscr, norm_scratch_doubles(1) * sizeof(double));
  for (int k = 0; k < 2 && st == GTOMX_OK; ++k) {
    if (hipHostMalloc((void**)&pp->h_raw[k], pp->raw_cap * sizeof(int2), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&pp->h_stage[k], (size_t)B * 12 * sizeof(float), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&pp->uploaded[k], hipEventDisableTiming) != hipSuccess)
      st = fail(GTOMX_ERR_OUT_OF_MEMORY, "pinned staging buffers");
  }
  if (st != GTOMX_OK) {
    prob_release(pp);
    return st;
  }
  // The banded transforms of micrographs leave rows of `real` outside the
  // mask's band as they were, to be multiplied by the mask's zeros: they must
  // hold finite values from the start.
  if (hipMemsetAsync(pp->real, 0, (size_t)B * Vi * sizeof(float), s) != hipSuccess) {
    prob_release(pp);
    return fail(GTOMX_ERR_DEVICE, "clearing the batch buffer");
  }

  // Reference spectra and masks, all references at once. Workspace: the padded
  // references R (nrefs x Vi) in `real`-sized chunks would not fit every case, so
  // one temporary of nrefs x (Vi floats + Hi complex) is taken here.
  DevBuf tR, tF, tC;
  auto fin = [&](int status) {
    if (status != GTOMX_OK) prob_release(pp);
    return status;
  };
  if ((st = tR.alloc((size_t)nrefs * Vi * sizeof(float))) != GTOMX_OK) return fin(st);
  if ((st = tF.alloc((size_t)nrefs * pp->Hi * sizeof(float2))) != GTOMX_OK) return fin(st);
  if ((st = tC.alloc((size_t)nrefs * std::max(Hl * sizeof(float2), pp->Er * sizeof(float)))) != GTOMX_OK)
    return fin(st);
  float* R = tR.as<float>();
  float2* F = tF.as<float2>();
  // R = ifftshift(pad(ref)); F = R2C(R); C = crop(F) band-limited; quirky remap.
  if ((st = gtomx_pad(d_refs, reinterpret_cast<float*>(F), dimsref, dimsimage, 0.0f, nrefs, stream)) != GTOMX_OK)
    return fin(st);
  if ((st = gtomx_remap_full2fullfft(reinterpret_cast<float*>(F), R, dimsimage, nrefs, stream)) != GTOMX_OK)
    return fin(st);
  if ((st = fft_r2c(R, F, dimsimage, nrefs, s)) != GTOMX_OK) return fin(st);
  float2* C = tC.as<float2>();
  if ((st = fft_resize(FFT_CROP, true, F, C, dimsimage, dl, nrefs, s)) != GTOMX_OK) return fin(st);
  hipLaunchKernelGGL(k_bandpass, dim3(grid_of(Hl * nrefs)), dim3(256), 0, s, C, make_int3(dl.x, dl.y, dl.z),
                     (int)lowpassfreq, Hl, nrefs);
  if (hipGetLastError() != hipSuccess) return fin(fail(GTOMX_ERR_DEVICE, "k_bandpass"));
  hipLaunchKernelGGL(k_remap_half_quirk, dim3(grid_of(Hl * nrefs)), dim3(256), 0, s, C, R, pp->refs_low, n, dl.z,
                     Hl, Vi, nrefs);
  if (hipGetLastError() != hipSuccess) return fin(fail(GTOMX_ERR_DEVICE, "k_remap_half_quirk"));
  if (!pp->circular) {
    // the padded masks, centred, as cubic B-spline coefficients (:102-110)
    if ((st = gtomx_pad(d_refmasks, pp->masks, dimsref, dimsimage, 0.0f, nrefs, stream)) != GTOMX_OK) return fin(st);
    if ((st = bspline_prefilter(pp->masks, dimsimage, nd, nrefs, 1, s)) != GTOMX_OK) return fin(st);
    if (hipStreamSynchronize(s) != hipSuccess) return fin(fail(GTOMX_ERR_DEVICE, "hipStreamSynchronize"));
    *out = pp;
    return GTOMX_OK;
  }
  // M = ifftshift(pad(mask)); IM = R2C(ifftshift(pad(1 - mask)))
  if ((st = gtomx_pad(d_refmasks, R, dimsref, dimsimage, 0.0f, nrefs, stream)) != GTOMX_OK) return fin(st);
  if ((st = gtomx_remap_full2fullfft(R, pp->masks, dimsimage, nrefs, stream)) != GTOMX_OK) return fin(st);
  float* om = tC.as<float>();
  hipLaunchKernelGGL(k_one_minus, dim3(grid_of(pp->Er * nrefs)), dim3(256), 0, s, d_refmasks, om, pp->Er * nrefs);
  if (hipGetLastError() != hipSuccess) return fin(fail(GTOMX_ERR_DEVICE, "k_one_minus"));
  if ((st = gtomx_pad(om, reinterpret_cast<float*>(F), dimsref, dimsimage, 0.0f, nrefs, stream)) != GTOMX_OK)
    return fin(st);
  if ((st = gtomx_remap_full2fullfft(reinterpret_cast<float*>(F), R, dimsimage, nrefs, stream)) != GTOMX_OK)
    return fin(st);
  if ((st = fft_r2c(R, pp->invmasksft, dimsimage, nrefs, s)) != GTOMX_OK) return fin(st);
  if (hipStreamSynchronize(s) != hipSuccess) return fin(fail(GTOMX_ERR_DEVICE, "hipStreamSynchronize"));
  *out = pp;
  return GTOMX_OK;
}

int gtomx_prob_picker_destroy(gtomx_prob_picker* picker) {
  prob_release(picker);
  return GTOMX_OK;
}

int gtomx_prob_picker_set_image(gtomx_prob_picker* pp, const float* d_image, const float* d_ctf, void* stream) {
  GX_TRY(prob_check(pp));
  if (!d_image) return fail(GTOMX_ERR_INVALID_ARGUMENT, "d_image is NULL");
  if (pp->doctf && !d_ctf) return fail(GTOMX_ERR_INVALID_ARGUMENT, "d_ctf is NULL and the picker applies a CTF");
  hipStream_t s = as_stream(stream);
  const int lp = pp->dl.x / 2;
  const int3 di = make_int3(pp->di.x, pp->di.y, pp->di.z);
  // the mean0/std1 image goes to the first map of `real`, its square to a temporary
  DevBuf sq;
  GX_TRY(sq.alloc(pp->Vi * sizeof(float)));
  GX_TRY(norm_mean01std(d_image, pp->real, pp->Vi, 1, pp->nscr, s));
  GX_TRY(fft_r2c(pp->real, pp->imageft, pp->di, 1, s));
  hipLaunchKernelGGL(k_bandpass, dim3(grid_of(pp->Hi)), dim3(256), 0, s, pp->imageft, di, lp, pp->Hi, 1u);
  GX_LAUNCHED("k_bandpass");
  hipLaunchKernelGGL(k_square_into, dim3(grid_of(pp->Vi)), dim3(256), 0, s, pp->real, sq.as<float>(), pp->Vi);
  GX_LAUNCHED("k_square_into");
  GX_TRY(fft_r2c(sq.as<float>(), pp->image2ft, pp->di, 1, s));
  hipLaunchKernelGGL(k_bandpass, dim3(grid_of(pp->Hi)), dim3(256), 0, s, pp->image2ft, di, lp, pp->Hi, 1u);
  GX_LAUNCHED("k_bandpass");
  if (d_ctf) GX_TRY(fft_resize(FFT_CROP, false, d_ctf, pp->ctf_low, pp->di, pp->dl, 1, s));
  GX_HIP(hipStreamSynchronize(s));
  pp->have_image = true;
  return GTOMX_OK;
}

void gtomx_debug_rand_outputs(unsigned seed, long long first, int n, int* out) {
  if (!out || n <= 0 || first < 0) return;
  AdditiveJump(seed).outputs(first, n, out);
}

int gtomx_prob_picker_count_angles(int ndims, float anglestep, int* count) {
  if (!count) return fail(GTOMX_ERR_INVALID_ARGUMENT, "count is NULL");
  if (ndims != 2 && ndims != 3) return fail(GTOMX_ERR_INVALID_ARGUMENT, "ndims must be 2 or 3");
  std::vector<gtomx_float3> a;
  GX_TRY(prob_angles(ndims, anglestep, a));
  *count = (int)a.size();
  return GTOMX_OK;
}

int gtomx_prob_picker_correlate(gtomx_prob_picker* pp, unsigned n, float anglestep, float* d_bestccf,
                                gtomx_float3* d_bestangle, int* d_bestref, void* stream) {
  GX_TRY(prob_check(pp));
  if (!pp->have_image) return fail(GTOMX_ERR_INVALID_ARGUMENT, "set_image was not called");
  if (n >= pp->nrefs) return fail(GTOMX_ERR_INVALID_ARGUMENT, "reference index out of range");
  if (!d_bestccf || !d_bestangle || !d_bestref) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL output");
  std::vector<gtomx_float3> angles;
  GX_TRY(prob_angles(pp->ndims, anglestep, angles));
  hipStream_t s = as_stream(stream);
  const unsigned B = pp->B;
  const size_t Er = pp->Er, Vi = pp->Vi, Hi = pp->Hi, Hl = pp->Hl;
  const int3 di = make_int3(pp->di.x, pp->di.y, pp->di.z), dr = make_int3(pp->dr.x, pp->dr.y, pp->dr.z);
  const float2* refft = pp->refs_low + (size_t)n * Hl;
  const float* mask = pp->masks + (size_t)n * Vi;
  const float2* invmaskft = pp->circular ? pp->invmasksft + (size_t)n * Hi : nullptr;
  const float maxfreq = (float)(pp->dl.x / 2);
  std::lock_guard<std::mutex> lock(g_gauss_mutex);
  // rand() integers in draw order: the restated generator where it is the C
  // runtime's, else rand() itself (reseeded as the reference does, :229)
  const bool additive = runtime_rand_is_additive();
  const AdditiveJump jump(123);
  const int fill_threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::srand(123);
  auto next_int = [&]() { return std::rand(); };
  const int p0 = g_gauss.phase;
  const float Us = g_gauss.U, Vs = g_gauss.V;
  long long pairs_made = 0;  // rand() pairs drawn so far (pair q uses pair index q - p0)
  int2 last_pair = make_int2(0, 0);
  bool used[2] = {false, false};
  for (size_t a0 = 0, k = 0; a0 < angles.size(); a0 += B, ++k) {
    const unsigned nb = (unsigned)std::min<size_t>(B, angles.size() - a0);
    const int slot = (int)(k % 2);
    if (used[slot]) GX_HIP(hipEventSynchronize(pp->uploaded[slot]));
    // the pairs these orientations' draws touch: q in [qlo, qhi]
    const long long gs = (long long)a0 * (long long)Er * 5, ge = (long long)(a0 + nb) * (long long)Er * 5;
    const long long qlo = (gs + p0) / 2, qhi = (ge - 1 + p0) / 2;
    int2* hr = pp->h_raw[slot];
    if (qlo < p0) hr[0] = make_int2(0, 0);  // the half-used pair of the previous call; U, V passed instead
    const long long q0 = std::max<long long>(qlo, p0);
    if (additive && qhi >= q0) {
      // pair q holds outputs 2(q - p0) and 2(q - p0) + 1: threads fill slices from their own positions
      const long long npairs = qhi - q0 + 1;
      const int nt = (int)std::max<long long>(1, std::min<long long>(fill_threads, npairs / 4096));
      std::vector<std::thread> pool;
      for (int t = 0; t < nt; ++t) {
        const long long a = q0 + npairs * t / nt, e = q0 + npairs * (t + 1) / nt;
        pool.emplace_back([&, a, e]() {
          jump.outputs(2 * (a - p0), 2 * (e - a), reinterpret_cast<int*>(hr + (a - qlo)));
        });
      }
      for (auto& th : pool) th.join();
      last_pair = hr[qhi - qlo];
    } else {
      for (long long q = q0; q <= qhi; ++q) {
        if (q - p0 >= pairs_made) {
          const int a = next_int();
          const int b2 = next_int();
          last_pair = make_int2(a, b2);
          ++pairs_made;
        }
        hr[q - qlo] = last_pair;  // a pair straddling two batches is repeated
      }
    }
    float* hs = pp->h_stage[slot];
    if (pp->ndims == 3) {
      GX_TRY(gtomx_projector_matrices(angles.data() + a0, nb, 1.0f, hs));
    } else {
      for (unsigned b = 0; b < nb; ++b) rotate2d_coefficients(angles[a0 + b].z, reinterpret_cast<float2*>(hs) + b);
    }
    float* ha = hs + (size_t)B * 9;
    for (unsigned b = 0; b < nb; ++b) {
      ha[3 * b] = angles[a0 + b].x;
      ha[3 * b + 1] = angles[a0 + b].y;
      ha[3 * b + 2] = angles[a0 + b].z;
    }
    GX_HIP(hipMemcpyAsync(pp->raw, hr, (size_t)(qhi - qlo + 1) * sizeof(int2), hipMemcpyHostToDevice, s));
    const size_t nmat = (size_t)nb * (pp->ndims == 3 ? 9 : 2);
    GX_HIP(hipMemcpyAsync(pp->mats, hs, nmat * sizeof(float), hipMemcpyHostToDevice, s));
    GX_HIP(hipMemcpyAsync(pp->dang, ha, (size_t)nb * 3 * sizeof(float), hipMemcpyHostToDevice, s));
    GX_HIP(hipEventRecord(pp->uploaded[slot], s));
    used[slot] = true;

    // rotate, CTF, pad, r = C2R / N, r *= mask
    if (pp->ndims == 3)
      GX_TRY(rotate3d_ft_batch(refft, pp->rot, pp->dl.x, pp->mats, nb, false, s));
    else
      GX_TRY(rotate2d_ft_batch(refft, 0, pp->rot, pp->dl.x, pp->dl.y, reinterpret_cast<const float2*>(pp->mats), nb,
                               maxfreq, false, s));
    if (pp->doctf) {
      hipLaunchKernelGGL(k_times_real, dim3(grid_of(Hl * nb)), dim3(256), 0, s, pp->rot, pp->ctf_low, Hl, nb);
      GX_LAUNCHED("k_times_real");
    }
    GX_TRY(fft_resize(FFT_PAD, true, pp->rot, pp->padft, pp->dl, pp->di, nb, s));
    // Square micrographs with a circular mask: only the padded reference's
    // low-pass columns are nonzero, and only the rows of the mask's band
    // survive the mask, so the transforms skip the rest (the same values
    // where they are used).
    const int n2d = pp->di.x;
    const bool native = dims_native(pp->di);
    const bool banded = native && pp->ndims == 2 && pp->circular && n2d > 256 && pp->di.y == n2d;
    const int band = std::min(n2d / 2, (pp->dr.y / 2 + 1 + 7) / 8 * 8);
    // the same for volumes and images up to 256 per axis: bands in y (whole
    // row blocks of the row kernels) and, for volumes, in z
    const int rw = native ? rows_per_workgroup(pp->di.x) : 0;
    const int vby = rw > 0 ? (pp->dr.y / 2 + 1 + rw - 1) / rw * rw : 0;
    const int vbz = pp->ndims == 3 && 2 * (pp->dr.z / 2 + 1) <= pp->di.z ? pp->dr.z / 2 + 1 : 0;
    const bool vbanded = !banded && pp->circular && pp->di.x <= 256 && pp->di.y <= 256 && pp->di.z <= 256 &&
                         vby > 0 && 2 * vby <= pp->di.y;
    if (banded)
      GX_TRY(image_c2r_banded(pp->padft, pp->real, n2d, nb, pp->dl.x / 2 + 1, band, 1.0f / (float)Vi, s));
    else if (vbanded)
      GX_TRY(volume_c2r_banded(pp->padft, pp->real, pp->di, nb, pp->dl.x / 2 + 1, vby, vbz, 1.0f / (float)Vi, s));
    else  // the padded spectrum is zero beyond the low-pass half width
      GX_TRY(fft_c2r(pp->padft, pp->padft, pp->real, pp->di, nb, 1.0f / (float)Vi, s, pp->dl.x / 2 + 1));
    // the mask of each orientation: the same one (circular), or the B-spline
    // coefficients rotated into FFT order (d_Rotate3D / d_Rotate2D, T_INTERP_CUBIC, :273-279)
    const float* mk = mask;
    size_t mstride = 0;
    if (!pp->circular) {
      hipLaunchKernelGGL(k_rotate_cubic, dim3((unsigned)((Vi + 255) / 256), nb), dim3(256), 0, s, mask, pp->mrot,
                         di, pp->mats, pp->ndims);
      GX_LAUNCHED("k_rotate_cubic");
      mk = pp->mrot;
      mstride = Vi;
    }
    if (banded || vbanded) {  // circular masks: one mask, rows of the band only
      RowBand rb;
      rb.ny = pp->di.y, rb.by = banded ? band : vby, rb.nz = pp->di.z, rb.bz = banded ? 0 : vbz;
      const long long vrows = (long long)nb * (rb.bz ? 2 * rb.bz : rb.nz) * 2 * rb.by;
      hipLaunchKernelGGL(k_times_mask_band, dim3((unsigned)std::min<long long>((vrows * pp->di.x + 255) / 256, 8192)),
                         dim3(256), 0, s, pp->real, mk, pp->di.x, vrows, rb, (long long)pp->di.y * pp->di.z);
      GX_LAUNCHED("k_times_mask_band");
    } else {
      hipLaunchKernelGGL(k_times_mask, dim3(grid_of(Vi * nb)), dim3(256), 0, s, pp->real, mk, Vi, nb, mstride);
      GX_LAUNCHED("k_times_mask");
    }
    // statistics of the cropped, masked reference
    const int nparts = (int)std::min<size_t>(NBS, (Er + 255) / 256);
    const Draws draws{pp->raw, qlo, p0, Us, Vs, (double)RAND_MAX};
    hipLaunchKernelGGL(k_prob_sums, dim3(nparts, nb), dim3(256), 0, s, pp->real, mk, mstride, draws,
                       (long long)a0 * (long long)Er, di, dr, Vi, Er, pp->part);
    GX_LAUNCHED("k_prob_sums");
    hipLaunchKernelGGL(k_prob_coef, dim3((nb + 63) / 64), dim3(64), 0, s, pp->part, nparts, nb, Vi, Er, pp->coef);
    GX_LAUNCHED("k_prob_coef");
    // correlation with the image
    if (banded)
      GX_TRY(image_r2c_banded(pp->real, pp->padft, n2d, nb, band, s));
    else if (vbanded)
      GX_TRY(volume_r2c_banded(pp->real, pp->padft, pp->di, nb, vby, vbz, s));
    else
      GX_TRY(fft_r2c(pp->real, pp->padft, pp->di, nb, s));
    GX_TRY(image_times_conj(pp->imageft, pp->padft, Hi, nb, s));
    GX_TRY(fft_c2r(pp->padft, pp->padft, pp->real, pp->di, nb, 1.0f, s));
    if (pp->circular && a0 == 0) {
      // solvent statistics with the first orientation's mask sum (the mask is
      // circular, so every orientation has the same); padft[0] is free again.
      float2* tmp = pp->padft;
      hipLaunchKernelGGL(k_mul_conj_into, dim3(grid_of(Hi)), dim3(256), 0, s, pp->imageft, invmaskft, tmp, Hi);
      GX_LAUNCHED("k_mul_conj_into");
      GX_TRY(fft_c2r(tmp, tmp, pp->smean, pp->di, 1, 1.0f, s));
      hipLaunchKernelGGL(k_mul_conj_into, dim3(grid_of(Hi)), dim3(256), 0, s, pp->image2ft, invmaskft, tmp, Hi);
      GX_LAUNCHED("k_mul_conj_into");
      GX_TRY(fft_c2r(tmp, tmp, pp->sstd, pp->di, 1, 1.0f, s));
      hipLaunchKernelGGL(k_solvent_stats, dim3(grid_of(Vi)), dim3(256), 0, s, pp->smean, pp->sstd, Vi, pp->coef, 1u);
      GX_LAUNCHED("k_solvent_stats");
    } else if (!pp->circular) {
      // every orientation's own inverse mask and solvent statistics (:342-356):
      // 1 - crop(mask), full pad, R2C; then I and I2 against it, as above
      hipLaunchKernelGGL(k_crop_one_minus, dim3((unsigned)((Er + 255) / 256), nb), dim3(256), 0, s, pp->mrot, pp->om,
                         di, dr, Vi, Er);
      GX_LAUNCHED("k_crop_one_minus");
      GX_TRY(fft_resize(FFT_FULL_PAD, false, pp->om, pp->smean, pp->dr, pp->di, nb, s));
      GX_TRY(fft_r2c(pp->smean, pp->imft, pp->di, nb, s));
      hipLaunchKernelGGL(k_mul_conj_batch, dim3(grid_of(Hi * nb)), dim3(256), 0, s, pp->imageft, pp->imft, pp->padft,
                         Hi, nb);
      GX_LAUNCHED("k_mul_conj_batch");
      GX_TRY(fft_c2r(pp->padft, pp->padft, pp->smean, pp->di, nb, 1.0f, s));
      hipLaunchKernelGGL(k_mul_conj_batch, dim3(grid_of(Hi * nb)), dim3(256), 0, s, pp->image2ft, pp->imft,
                         pp->padft, Hi, nb);
      GX_LAUNCHED("k_mul_conj_batch");
      GX_TRY(fft_c2r(pp->padft, pp->padft, pp->sstd, pp->di, nb, 1.0f, s));
      hipLaunchKernelGGL(k_solvent_stats, dim3(grid_of(Vi * nb)), dim3(256), 0, s, pp->smean, pp->sstd, Vi, pp->coef,
                         nb);
      GX_LAUNCHED("k_solvent_stats");
    }
    hipLaunchKernelGGL(k_prob_update, dim3(grid_of(Vi)), dim3(256), 0, s, pp->real, pp->smean, pp->sstd,
                       pp->circular ? (size_t)0 : Vi, pp->coef, pp->dang, nb, Vi, (int)n, d_bestccf,
                       reinterpret_cast<float3*>(d_bestangle), d_bestref);
    GX_LAUNCHED("k_prob_update");
  }
  GX_HIP(hipStreamSynchronize(s));
  // the pair state left for the next call, as gaussrand's statics
  const long long tend = (long long)angles.size() * (long long)Er * 5 + p0;
  if (tend % 2 == 1) {
    const long long q = (tend - 1) / 2;
    if (q >= p0) {
      g_gauss.U = (float)(((double)last_pair.x + 1.) / (RAND_MAX + 2.));
      g_gauss.V = (float)((double)last_pair.y / (RAND_MAX + 1.));
    }
  }
  g_gauss.phase = (int)(tend % 2);
  return GTOMX_OK;
}

}  // extern "C"
