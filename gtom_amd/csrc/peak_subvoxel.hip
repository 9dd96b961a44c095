// Sub-voxel peak location on correlation maps:
//  * d_Peak (src/Correlation/Peak.cu:26-216): the maximum of each map, as an
//    integer voxel, refined along lines through it by sinc interpolation
//    (sub-coarse), or by Fourier up-scaling of the region around it (sub-fine);
//  * d_PeakOne2D (:218-225, PeakOne2DKernel :380-454): the maximum of a
//    centred region of each image plus the 3x3 value-weighted centroid;
//  * d_SubpixelMax (src/Correlation/LocalPeaks.cu:64-91, SubpixelMaxKernel
//    :145-178): the running maximum of a map sinc-shifted by every sub-sample
//    offset.
//
// The reference moves integer positions, line extracts and interpolated
// regions between host and device between its kernels and up-scales the
// sub-fine region with two FFTs into a materialised (samples*subdivisions)^3
// volume. Here every stage reads the previous one's result from device memory:
//  k_argmax_partial / k_argmax_final  two-stage max with the first raster index
//                                     on ties, written straight as positions;
//  k_peak_subcoarse<ND>               one workgroup per map: ND line searches
//                                     over 1024 positions in double and the
//                                     sinc evaluation at the result;
//  k_subfine_spectrum / k_subfine_rows / k_subfine_argmax / k_subfine_finish
//                                     the up-scaled region is never stored: the
//                                     few non-zero bins of the padded spectrum
//                                     are transformed along y and z once
//                                     (G(kx, y', z')), and the C2R along x is
//                                     evaluated on the fly inside the argmax;
//  k_subpixel_max<ND>                 LDS tiles with a 6-voxel halo; the
//                                     13^ND-point sinc sum separated into x, y
//                                     and z passes, each shared by the shifts of
//                                     the later dimensions.
#include <climits>
#include <cmath>
#include <vector>

#include "runtime.hpp"

namespace gx {
namespace {

constexpr double PI_D = 3.14159265358979323846;

// (ov, orank) ranks before (v, rank): larger value, then smaller rank.
__device__ __forceinline__ bool beats(float ov, long long orank, float v, long long rank) {
  return ov > v || (ov == v && orank < rank);
}

__device__ __forceinline__ long long shfl_xor_ll(long long x, int m) {
  const int lo = __shfl_xor((int)(unsigned)(x & 0xffffffffll), m);
  const int hi = __shfl_xor((int)(x >> 32), m);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Best (value, rank) over a 256-thread block; every thread receives it.
__device__ void block_best256(float& v, long long& r) {
  __shared__ float sv[4];
  __shared__ long long sr[4];
  for (int off = 32; off >= 1; off >>= 1) {
    const float ov = __shfl_xor(v, off);
    const long long orr = shfl_xor_ll(r, off);
    if (beats(ov, orr, v, r)) {
      v = ov;
      r = orr;
    }
  }
  if ((threadIdx.x & 63) == 0) {
    sv[threadIdx.x >> 6] = v;
    sr[threadIdx.x >> 6] = r;
  }
  __syncthreads();
  v = sv[0];
  r = sr[0];
  for (int k = 1; k < 4; ++k)
    if (beats(sv[k], sr[k], v, r)) {
      v = sv[k];
      r = sr[k];
    }
  __syncthreads();
}

__device__ float block_sum256(float v) {
  __shared__ float s[4];
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
  __syncthreads();
  const float r = (s[0] + s[1]) + (s[2] + s[3]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ int wrapi(long long i, int n) {
  const long long r = i % n;
  return (int)(r < 0 ? r + n : r);
}

// sinc of include/DeviceFunctions.cuh:11-25: 1 within the threshold, else
// sin(pi x)/(pi x).
__device__ __forceinline__ double sinc_d(double x) {
  return fabs(x) <= 1e-8 ? 1.0 : sin(x * PI_D) / (x * PI_D);
}
__device__ __forceinline__ float sinc_f(float x) {
  return fabsf(x) <= 1e-4f ? 1.0f : sinpif(x) / (x * (float)PI_D);
}

// ---- integer maximum -------------------------------------------------------

// grid (P, batch): block p reduces elements [p*chunk, (p+1)*chunk) of map b.
__global__ __launch_bounds__(256) void k_argmax_partial(const float* __restrict__ in, long long V, long long chunk,
                                                        float* __restrict__ pv, long long* __restrict__ pi) {
  const float* m = in + (size_t)blockIdx.y * V;
  const long long lo = (long long)blockIdx.x * chunk, hi = min(V, lo + chunk);
  float v = -INFINITY;
  long long r = LLONG_MAX;
  for (long long i = lo + threadIdx.x; i < hi; i += 256) {
    const float x = m[i];
    if (x > v) {
      v = x;
      r = i;
    }
  }
  block_best256(v, r);
  if (threadIdx.x == 0) {
    pv[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = v;
    pi[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = r;
  }
}

// One block per map: the best partial, stored as the integer result.
__global__ __launch_bounds__(256) void k_argmax_final(const float* __restrict__ in, long long V, int P,
                                                      const float* __restrict__ pv, const long long* __restrict__ pi,
                                                      int nx, int ny, long long* __restrict__ best_i,
                                                      float* __restrict__ best_v, gtomx_float3* __restrict__ pos,
                                                      float* __restrict__ val) {
  const int b = blockIdx.x;
  float v = -INFINITY;
  long long r = LLONG_MAX;
  for (int k = threadIdx.x; k < P; k += 256)
    if (beats(pv[(size_t)b * P + k], pi[(size_t)b * P + k], v, r)) {
      v = pv[(size_t)b * P + k];
      r = pi[(size_t)b * P + k];
    }
  block_best256(v, r);
  if (threadIdx.x == 0) {
    if (r == LLONG_MAX) {  // no value above -inf (or NaN only): the first voxel
      r = 0;
      v = in[(size_t)b * V];
    }
    best_i[b] = r;
    best_v[b] = v;
    const long long plane = (long long)nx * ny;
    pos[b] = gtomx_float3{(float)(r % nx), (float)((r / nx) % ny), (float)(r / plane)};
    val[b] = v;
  }
}

// ---- sub-coarse --------------------------------------------------------------

__device__ __forceinline__ int bitrev8(int t) { return (int)(__brev((unsigned)t) >> 24); }

// One block per map. n samples per line (even, <= 20) through the integer
// maximum along each of the ND dimensions, starting at peak - n/2, cyclic.
// Line search (SincMax1DKernel, Peak.cu:230-320): y(x) = sum over |k| <=
// n/2 - 1 of sinc(k - x) s[n/2 + k] at x = id/512 - 1, outermost pair first,
// in double; thread t visits id = t + 256 j keeping a float (x, y) that a later
// id replaces when (double)y_kept < y; among equal float maxima the thread with
// the smallest bit-reversed index wins (the reference's tree order). Value
// (SincEvalMaxKernel :322-378): the float sinc interpolation over the n^ND
// region starting at peak - n/2, at (n/2 + shift) per dimension.
template <int ND>
__global__ __launch_bounds__(256) void k_peak_subcoarse(const float* __restrict__ in, int nx, int ny, int nz, int n,
                                                        const long long* __restrict__ best_i,
                                                        gtomx_float3* __restrict__ pos, float* __restrict__ val) {
  const int b = blockIdx.x, t = threadIdx.x;
  const long long V = (long long)nx * ny * nz;
  const float* m = in + (size_t)b * V;
  const long long i0 = best_i[b];
  const int dims[3] = {nx, ny, nz};
  const int ip[3] = {(int)(i0 % nx), (int)((i0 / nx) % ny), (int)(i0 / ((long long)nx * ny))};
  __shared__ double line[3][20];
  __shared__ float s_x[256];
  __shared__ float s_shift[3];
  if (t < ND * n) {
    const int d = t / n, i = t % n;
    int q[3] = {ip[0], ip[1], ip[2]};
    q[d] = wrapi((long long)q[d] + i - n / 2, dims[d]);
    line[d][i] = (double)m[((size_t)q[2] * ny + q[1]) * nx + q[0]];
  }
  __syncthreads();
  for (int d = 0; d < ND; ++d) {
    float bx = 0.f, by = -1e30f;
    for (int id = t; id < 1024; id += 256) {
      const double x = (double)id / 512.0 - 1.0;
      double y = 0;
      for (int k = 9; k >= 1; --k) {
        if (n <= 2 * k) continue;
        y += sinc_d(-(double)k - x) * line[d][n / 2 - k];
        y += sinc_d((double)k - x) * line[d][n / 2 + k];
      }
      y += sinc_d(x) * line[d][n / 2];
      if ((double)by < y) {
        bx = (float)x;
        by = (float)y;
      }
    }
    s_x[t] = bx;
    float v = by;
    long long r = bitrev8(t);
    block_best256(v, r);  // synchronises, so s_x is complete
    if (t == 0) s_shift[d] = s_x[bitrev8((int)r)];
    __syncthreads();
  }
  float e[3] = {0.f, 0.f, 0.f};
  for (int d = 0; d < ND; ++d) e[d] = (float)(n / 2) + s_shift[d];
  const int ry = ND > 1 ? n : 1, rz = ND > 2 ? n : 1;
  float sum = 0.f;
  for (int idx = t; idx < n * ry * rz; idx += 256) {
    const int x = idx % n, y = (idx / n) % ry, z = idx / (n * ry);
    float w = sinc_f((float)x - e[0]);
    if (ND > 1) w *= sinc_f((float)y - e[1]);
    if (ND > 2) w *= sinc_f((float)z - e[2]);
    const int gx = wrapi((long long)ip[0] - n / 2 + x, nx), gy = wrapi((long long)ip[1] - ry / 2 + y, ny),
              gz = wrapi((long long)ip[2] - rz / 2 + z, nz);
    sum += m[((size_t)gz * ny + gy) * nx + gx] * w;
  }
  sum = block_sum256(sum);
  if (t == 0) {
    float p[3] = {(float)ip[0], (float)ip[1], (float)ip[2]};
    for (int d = 0; d < ND; ++d) p[d] += s_shift[d];
    pos[b] = gtomx_float3{p[0], p[1], p[2]};
    val[b] = sum;
  }
}

// ---- sub-fine ----------------------------------------------------------------
// Region (r0, r1, r2) of samples centred on the integer maximum, up-scaled to
// (u0, u1, u2) = (r0, r1, r2) * sub (1 stays 1): d_Scale T_INTERP_FOURIER
// (src/Transformation/Scale.cu:109-139) = R2C, zero pad of the half spectrum
// (FFTPadEvenKernel, src/FFT/FFTResize.cu:162-184: kx < (r0+1)/2, old bin k of
// an axis of r samples lands on signed frequency ((k + r/2) % r) - r/2 of the
// new axis), C2R_u, times 1/N_old.

struct SubfineShape {
  int r[3];   // region
  int u[3];   // up-scaled region
  int kxn;    // x bins kept by the pad
  int sub, centre;
};

// One block per map: the kept bins S(kx, ky, kz) / N_old, doubled for kx > 0
// (the Hermitian mate of the C2R), in double. Layout [kz][ky][kx].
__global__ __launch_bounds__(256) void k_subfine_spectrum(const float* __restrict__ in, int nx, int ny, int nz,
                                                          SubfineShape sh, const long long* __restrict__ best_i,
                                                          double2* __restrict__ spec) {
  const int b = blockIdx.x, t = threadIdx.x;
  const float* m = in + (size_t)b * nx * ny * nz;
  const long long i0 = best_i[b];
  const int dims[3] = {nx, ny, nz};
  const int ip[3] = {(int)(i0 % nx), (int)((i0 / nx) % ny), (int)(i0 / ((long long)nx * ny))};
  __shared__ float reg[128];
  const int R = sh.r[0] * sh.r[1] * sh.r[2];
  if (t < R) {
    const int x = t % sh.r[0], y = (t / sh.r[0]) % sh.r[1], z = t / (sh.r[0] * sh.r[1]);
    const int q[3] = {x, y, z};
    int g[3];
    for (int d = 0; d < 3; ++d) g[d] = wrapi((long long)wrapi((long long)ip[d] - sh.r[d] / 2, dims[d]) + q[d], dims[d]);
    reg[t] = m[((size_t)g[2] * ny + g[1]) * nx + g[0]];
  }
  __syncthreads();
  const int nbins = sh.kxn * sh.r[1] * sh.r[2];
  const double norm = 1.0 / R;
  for (int bin = t; bin < nbins; bin += 256) {
    const int kx = bin % sh.kxn, ky = (bin / sh.kxn) % sh.r[1], kz = bin / (sh.kxn * sh.r[1]);
    double re = 0, im = 0;
    for (int i = 0; i < R; ++i) {
      const int x = i % sh.r[0], y = (i / sh.r[0]) % sh.r[1], z = i / (sh.r[0] * sh.r[1]);
      const double f = (double)((kx * x) % sh.r[0]) / sh.r[0] + (double)((ky * y) % sh.r[1]) / sh.r[1] +
                       (double)((kz * z) % sh.r[2]) / sh.r[2];
      double sn, cs;
      sincospi(-2.0 * f, &sn, &cs);
      re += reg[i] * cs;
      im += reg[i] * sn;
    }
    const double w = (kx > 0 ? 2.0 : 1.0) * norm;
    spec[(size_t)b * nbins + bin] = make_double2(re * w, im * w);
  }
}

__device__ __forceinline__ int pad_freq(int k, int r) { return (k + r / 2) % r - r / 2; }

// G(kx, y', z') = sum over (ky, kz) of S e^{2 pi i (fy y'/u1 + fz z'/u2)}; one
// thread per row (y', z'), all kx. grid (rows/256, maps).
__global__ __launch_bounds__(256) void k_subfine_rows(const double2* __restrict__ spec, SubfineShape sh,
                                                      float2* __restrict__ G) {
  const int b = blockIdx.y;
  const int rows = sh.u[1] * sh.u[2];
  const int row = blockIdx.x * 256 + threadIdx.x;
  if (row >= rows) return;
  const int yy = row % sh.u[1], zz = row / sh.u[1];
  const int nbins = sh.kxn * sh.r[1] * sh.r[2];
  const double2* S = spec + (size_t)b * nbins;
  double acc[5][2] = {};
  for (int kz = 0; kz < sh.r[2]; ++kz) {
    const long long fz = pad_freq(kz, sh.r[2]);
    double sz, cz;
    sincospi(2.0 * (double)(((fz * zz) % sh.u[2] + sh.u[2]) % sh.u[2]) / sh.u[2], &sz, &cz);
    for (int ky = 0; ky < sh.r[1]; ++ky) {
      const long long fy = pad_freq(ky, sh.r[1]);
      double sy, cy;
      sincospi(2.0 * (double)(((fy * yy) % sh.u[1] + sh.u[1]) % sh.u[1]) / sh.u[1], &sy, &cy);
      const double pr = cz * cy - sz * sy, pim = cz * sy + sz * cy;
#pragma unroll
      for (int kx = 0; kx < 5; ++kx) {
        if (kx >= sh.kxn) break;
        const double2 s = S[((size_t)kz * sh.r[1] + ky) * sh.kxn + kx];
        acc[kx][0] += s.x * pr - s.y * pim;
        acc[kx][1] += s.x * pim + s.y * pr;
      }
    }
  }
  float2* g = G + ((size_t)b * rows + row) * sh.kxn;
  for (int kx = 0; kx < sh.kxn; ++kx) g[kx] = make_float2((float)acc[kx][0], (float)acc[kx][1]);
}

// The up-scaled region value at (x', row) = sum over kx of Re(G(kx, row)
// e^{2 pi i kx x'/u0}); grid (P, maps), block p takes rows [p*rpb, (p+1)*rpb).
// Best per block with the first raster index on ties.
constexpr int SUBFINE_MAX_U0 = 945;
__global__ __launch_bounds__(256) void k_subfine_argmax(const float2* __restrict__ G, SubfineShape sh, int rpb,
                                                        float* __restrict__ pv, long long* __restrict__ pi) {
  const int b = blockIdx.y;
  const int rows = sh.u[1] * sh.u[2], u0 = sh.u[0];
  __shared__ float2 tw[5 * SUBFINE_MAX_U0];
  for (int j = threadIdx.x; j < sh.kxn * u0; j += 256) {
    const int kx = j / u0, x = j % u0;
    float s, c;
    sincospif(2.0f * (float)((kx * x) % u0) / (float)u0, &s, &c);
    tw[j] = make_float2(c, s);
  }
  __syncthreads();
  const int r0 = blockIdx.x * rpb, r1 = min(rows, r0 + rpb);
  float v = -INFINITY;
  long long r = LLONG_MAX;
  for (int row = r0; row < r1; ++row) {
    const float2* g = G + ((size_t)b * rows + row) * sh.kxn;
    float2 gk[5];
#pragma unroll
    for (int kx = 0; kx < 5; ++kx) gk[kx] = kx < sh.kxn ? g[kx] : make_float2(0.f, 0.f);
    for (int x = threadIdx.x; x < u0; x += 256) {
      float a = 0.f;
#pragma unroll
      for (int kx = 0; kx < 5; ++kx) {
        if (kx >= sh.kxn) break;
        const float2 w = tw[kx * u0 + x];
        a += gk[kx].x * w.x - gk[kx].y * w.y;
      }
      if (a > v) {
        v = a;
        r = (long long)row * u0 + x;
      }
    }
  }
  block_best256(v, r);
  if (threadIdx.x == 0) {
    pv[(size_t)b * gridDim.x + blockIdx.x] = v;
    pi[(size_t)b * gridDim.x + blockIdx.x] = r;
  }
}

// One block per map: the best partial moves the integer position by
// (index - centre)/sub per up-scaled dimension (Peak.cu:190-200); the value is
// the larger of the integer and the up-scaled maxima.
__global__ __launch_bounds__(256) void k_subfine_finish(const float* __restrict__ pv,
                                                        const long long* __restrict__ pi, int P, SubfineShape sh,
                                                        int ny, int nz, const float* __restrict__ best_v,
                                                        gtomx_float3* __restrict__ pos, float* __restrict__ val) {
  const int b = blockIdx.x;
  float v = -INFINITY;
  long long r = LLONG_MAX;
  for (int k = threadIdx.x; k < P; k += 256)
    if (beats(pv[(size_t)b * P + k], pi[(size_t)b * P + k], v, r)) {
      v = pv[(size_t)b * P + k];
      r = pi[(size_t)b * P + k];
    }
  block_best256(v, r);
  if (threadIdx.x == 0) {
    if (r == LLONG_MAX) r = 0;
    const long long S = sh.r[0], plane = S * S * sh.sub * sh.sub, rowlen = S * sh.sub;
    const long long iz = r / plane, iy = (r - iz * plane) / rowlen, ix = r - iz * plane - iy * rowlen;
    gtomx_float3 p = pos[b];
    p.x += (float)((int)ix - sh.centre) / (float)sh.sub;
    if (ny > 1) p.y += (float)((int)iy - sh.centre) / (float)sh.sub;
    if (nz > 1) p.z += (float)((int)iz - sh.centre) / (float)sh.sub;
    pos[b] = p;
    val[b] = fmaxf(best_v[b], v);
  }
}

// ---- d_PeakOne2D ---------------------------------------------------------------

// blockDim = min(128, rx*ry) threads; thread t visits id = t + T j of the centred
// region with strict '>' from (-1e20, (0, 0)); thread 0 takes the others in
// order with strict '>', then adds the 3x3 value-weighted centroid in float.
__global__ __launch_bounds__(128) void k_peak_one_2d(const float* __restrict__ in, int nx, int ny, int rx, int ry,
                                                     int subtract, gtomx_float3* __restrict__ pos,
                                                     float* __restrict__ val) {
  const int T = blockDim.x;
  const float* m = in + (size_t)blockIdx.x * nx * ny;
  const int ox = (nx - rx) / 2, oy = (ny - ry) / 2;
  float tv = -1e20f;
  int tx = 0, ty = 0;
  for (int id = threadIdx.x; id < rx * ry; id += T) {
    const int y = id / rx, x = id - y * rx;
    const float v = m[(size_t)(y + oy) * nx + x + ox];
    if (v > tv) {
      tv = v;
      tx = x + ox;
      ty = y + oy;
    }
  }
  __shared__ float sv[128];
  __shared__ int sx[128], sy[128];
  sv[threadIdx.x] = tv;
  sx[threadIdx.x] = tx;
  sy[threadIdx.x] = ty;
  __syncthreads();
  if (threadIdx.x != 0) return;
  float mv = sv[0];
  int mx = sx[0], my = sy[0];
  for (int t = 1; t < T; ++t)
    if (sv[t] > mv) {
      mv = sv[t];
      mx = sx[t];
      my = sy[t];
    }
  float cx = 0.f, cy = 0.f, sw = 0.f;
  for (int y = -1; y <= 1; ++y) {
    if (my + y < 0 || my + y >= ny) continue;
    for (int x = -1; x <= 1; ++x) {
      if (mx + x < 0 || mx + x >= nx) continue;
      const float v = m[(size_t)(my + y) * nx + mx + x];
      cx += v * x;
      cy += v * y;
      sw += v;
    }
  }
  if (sw != 0.f) {
    cx /= sw;
    cy /= sw;
  }
  float fx = (float)mx + cx, fy = (float)my + cy;
  if (subtract) {
    fx -= (float)(nx / 2);
    fy -= (float)(ny / 2);
  }
  pos[blockIdx.x] = gtomx_float3{fx, fy, 0.f};
  val[blockIdx.x] = mv;
}

// ---- d_SubpixelMax ---------------------------------------------------------------

template <int ND> struct SpxTile;
template <> struct SpxTile<3> { static constexpr int TX = 16, TY = 8, TZ = 4; };
template <> struct SpxTile<2> { static constexpr int TX = 64, TY = 16, TZ = 1; };
constexpr int SPX_R = 6, SPX_K = 2 * SPX_R + 1, SPX_MAX_STEPS = 64;

// One block per output tile; wt holds steps x 13 float weights sinc(s - k).
// x pass into s_x for each x shift, y pass into s_y for each y shift (3D), z
// pass (3D) or the y pass (2D) into the running maximum per output voxel.
template <int ND>
__global__ __launch_bounds__(256) void k_subpixel_max(const float* __restrict__ in, float* __restrict__ out, int nx,
                                                      int ny, int nz, int xlo, int xhi, int ylo, int yhi, int zlo,
                                                      int zhi, int steps, const float* __restrict__ wt) {
  constexpr int TX = SpxTile<ND>::TX, TY = SpxTile<ND>::TY, TZ = SpxTile<ND>::TZ;
  constexpr int IX = TX + 2 * SPX_R, IY = TY + 2 * SPX_R, IZ = ND == 3 ? TZ + 2 * SPX_R : 1;
  constexpr int NOUT = TX * TY * TZ / 256;
  __shared__ float s_in[IZ * IY * IX];
  __shared__ float s_x[IZ * IY * TX];
  __shared__ float s_y[ND == 3 ? IZ * TY * TX : 1];
  __shared__ float s_w[SPX_MAX_STEPS * SPX_K];
  const int ox = xlo + blockIdx.x * TX, oy = ylo + blockIdx.y * TY, oz = ND == 3 ? zlo + blockIdx.z * TZ : 0;
  for (int i = threadIdx.x; i < steps * SPX_K; i += 256) s_w[i] = wt[i];
  for (int i = threadIdx.x; i < IZ * IY * IX; i += 256) {
    const int lx = i % IX, ly = (i / IX) % IY, lz = i / (IX * IY);
    const int gx = ox - SPX_R + lx, gy = oy - SPX_R + ly, gz = ND == 3 ? oz - SPX_R + lz : 0;
    float v = 0.f;
    if ((unsigned)gx < (unsigned)nx && (unsigned)gy < (unsigned)ny && (unsigned)gz < (unsigned)nz)
      v = in[((size_t)gz * ny + gy) * nx + gx];
    s_in[i] = v;
  }
  float best[NOUT];
#pragma unroll
  for (int j = 0; j < NOUT; ++j) best[j] = -INFINITY;
  __syncthreads();
  for (int sx = 0; sx < steps; ++sx) {
    const float* wx = s_w + sx * SPX_K;
    for (int i = threadIdx.x; i < IZ * IY * TX; i += 256) {
      const int x = i % TX, row = i / TX;
      const float* src = s_in + row * IX + x;
      float a = 0.f;
#pragma unroll
      for (int k = 0; k < SPX_K; ++k) a += wx[k] * src[k];
      s_x[i] = a;
    }
    __syncthreads();
    for (int sy = 0; sy < steps; ++sy) {
      const float* wy = s_w + sy * SPX_K;
      if constexpr (ND == 2) {
#pragma unroll
        for (int j = 0; j < NOUT; ++j) {
          const int i = threadIdx.x + 256 * j, x = i % TX, y = i / TX;
          float a = 0.f;
#pragma unroll
          for (int k = 0; k < SPX_K; ++k) a += wy[k] * s_x[(y + k) * TX + x];
          best[j] = fmaxf(best[j], a);
        }
      } else {
        for (int i = threadIdx.x; i < IZ * TY * TX; i += 256) {
          const int x = i % TX, y = (i / TX) % TY, z = i / (TX * TY);
          float a = 0.f;
#pragma unroll
          for (int k = 0; k < SPX_K; ++k) a += wy[k] * s_x[(z * IY + y + k) * TX + x];
          s_y[i] = a;
        }
        __syncthreads();
        for (int sz = 0; sz < steps; ++sz) {
          const float* wz = s_w + sz * SPX_K;
#pragma unroll
          for (int j = 0; j < NOUT; ++j) {
            const int i = threadIdx.x + 256 * j, x = i % TX, y = (i / TX) % TY, z = i / (TX * TY);
            float a = 0.f;
#pragma unroll
            for (int k = 0; k < SPX_K; ++k) a += wz[k] * s_y[((z + k) * TY + y) * TX + x];
            best[j] = fmaxf(best[j], a);
          }
        }
        __syncthreads();
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < NOUT; ++j) {
    const int i = threadIdx.x + 256 * j, x = i % TX, y = (i / TX) % TY, z = i / (TX * TY);
    const int gx = ox + x, gy = oy + y, gz = oz + z;
    if (gx <= xhi && gy <= yhi && gz <= zhi) {
      float& o = out[((size_t)gz * ny + gy) * nx + gx];
      o = fmaxf(o, best[j]);
    }
  }
}

inline int dimension_count(gtomx_int3 d) { return d.z > 1 ? 3 : (d.y > 1 ? 2 : 1); }

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" int gtomx_peak(const float* d_input, gtomx_float3* d_positions, float* d_values, gtomx_int3 dims, int mode,
                          unsigned batch, void* stream) {
  if (!d_input || !d_positions || !d_values) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (dims.x < 1 || dims.y < 1 || dims.z < 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "dims must be positive");
  if (dims.y == 1 && dims.z > 1) return fail(GTOMX_ERR_UNSUPPORTED, "dims with y = 1 and z > 1");
  if (batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch must be positive");
  if (mode < GTOMX_PEAK_INTEGER || mode > GTOMX_PEAK_SUBFINE) return fail(GTOMX_ERR_INVALID_ARGUMENT, "unknown mode");
  if (batch > 65535) return fail(GTOMX_ERR_UNSUPPORTED, "batch above 65535");
  hipStream_t s = as_stream(stream);
  const long long V = (long long)elements(dims);
  const int nd = dimension_count(dims);

  const int P = (int)std::min<long long>(1024, std::max<long long>(1, (V + 16383) / 16384));
  const long long chunk = ((V + P - 1) / P + 255) / 256 * 256;
  DevBuf d_pv, d_pi, d_bi, d_bv;
  GX_TRY(d_pv.alloc((size_t)P * batch * sizeof(float)));
  GX_TRY(d_pi.alloc((size_t)P * batch * sizeof(long long)));
  GX_TRY(d_bi.alloc((size_t)batch * sizeof(long long)));
  GX_TRY(d_bv.alloc((size_t)batch * sizeof(float)));
  hipLaunchKernelGGL(k_argmax_partial, dim3(P, batch), dim3(256), 0, s, d_input, V, chunk, d_pv.as<float>(),
                     d_pi.as<long long>());
  GX_LAUNCHED("k_argmax_partial");
  hipLaunchKernelGGL(k_argmax_final, dim3(batch), dim3(256), 0, s, d_input, V, P, d_pv.as<float>(),
                     d_pi.as<long long>(), dims.x, dims.y, d_bi.as<long long>(), d_bv.as<float>(), d_positions,
                     d_values);
  GX_LAUNCHED("k_argmax_final");

  if (mode == GTOMX_PEAK_SUBCOARSE) {
    const int n = std::min(nd == 3 ? 10 : 20, (dims.x + 1) / 2 * 2);
    if (nd == 1)
      hipLaunchKernelGGL((k_peak_subcoarse<1>), dim3(batch), dim3(256), 0, s, d_input, dims.x, dims.y, dims.z, n,
                         d_bi.as<long long>(), d_positions, d_values);
    else if (nd == 2)
      hipLaunchKernelGGL((k_peak_subcoarse<2>), dim3(batch), dim3(256), 0, s, d_input, dims.x, dims.y, dims.z, n,
                         d_bi.as<long long>(), d_positions, d_values);
    else
      hipLaunchKernelGGL((k_peak_subcoarse<3>), dim3(batch), dim3(256), 0, s, d_input, dims.x, dims.y, dims.z, n,
                         d_bi.as<long long>(), d_positions, d_values);
    GX_LAUNCHED("k_peak_subcoarse");
  } else if (mode == GTOMX_PEAK_SUBFINE) {
    SubfineShape sh;
    int S = nd < 3 ? 9 : 5;
    const int d3[3] = {dims.x, dims.y, dims.z};
    for (int d = 0; d < nd; ++d) S = std::min(S, d3[d]);
    sh.sub = nd < 3 ? 105 : 63;
    sh.centre = S / 2 * sh.sub;
    sh.r[0] = S;
    sh.r[1] = std::min(dims.y, S);
    sh.r[2] = std::min(dims.z, S);
    sh.u[0] = S * sh.sub;
    sh.u[1] = dims.y == 1 ? 1 : S * sh.sub;
    sh.u[2] = dims.z == 1 ? 1 : S * sh.sub;
    sh.kxn = (S + 1) / 2;
    const int rows = sh.u[1] * sh.u[2];
    const int nbins = sh.kxn * sh.r[1] * sh.r[2];
    const int P2 = std::min(1024, rows);
    const int rpb = (rows + P2 - 1) / P2;
    const int P2u = (rows + rpb - 1) / rpb;
    const unsigned chunk_maps = std::min(batch, 64u);
    DevBuf d_spec, d_G, d_fv, d_fi;
    GX_TRY(d_spec.alloc((size_t)chunk_maps * nbins * sizeof(double2)));
    GX_TRY(d_G.alloc((size_t)chunk_maps * rows * sh.kxn * sizeof(float2)));
    GX_TRY(d_fv.alloc((size_t)chunk_maps * P2u * sizeof(float)));
    GX_TRY(d_fi.alloc((size_t)chunk_maps * P2u * sizeof(long long)));
    for (unsigned b0 = 0; b0 < batch; b0 += chunk_maps) {
      const unsigned nb = std::min(chunk_maps, batch - b0);
      hipLaunchKernelGGL(k_subfine_spectrum, dim3(nb), dim3(256), 0, s, d_input + (size_t)b0 * V, dims.x, dims.y,
                         dims.z, sh, d_bi.as<long long>() + b0, d_spec.as<double2>());
      GX_LAUNCHED("k_subfine_spectrum");
      hipLaunchKernelGGL(k_subfine_rows, dim3((rows + 255) / 256, nb), dim3(256), 0, s, d_spec.as<double2>(), sh,
                         d_G.as<float2>());
      GX_LAUNCHED("k_subfine_rows");
      hipLaunchKernelGGL(k_subfine_argmax, dim3(P2u, nb), dim3(256), 0, s, d_G.as<float2>(), sh, rpb,
                         d_fv.as<float>(), d_fi.as<long long>());
      GX_LAUNCHED("k_subfine_argmax");
      hipLaunchKernelGGL(k_subfine_finish, dim3(nb), dim3(256), 0, s, d_fv.as<float>(), d_fi.as<long long>(), P2u,
                         sh, dims.y, dims.z, d_bv.as<float>() + b0, d_positions + b0, d_values + b0);
      GX_LAUNCHED("k_subfine_finish");
    }
    GX_HIP(hipStreamSynchronize(s));
  }
  GX_HIP(hipStreamSynchronize(s));
  return ok();
}

extern "C" int gtomx_peak_one_2d(const float* d_input, gtomx_float3* d_positions, float* d_values, gtomx_int3 dims,
                                 gtomx_int3 dimsregion, int subtractcenter, unsigned batch, void* stream) {
  if (!d_input || !d_positions || !d_values) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (dims.x < 1 || dims.y < 1 || dims.z != 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "dims must be 2D");
  if (dimsregion.x < 1 || dimsregion.y < 1 || dimsregion.x > dims.x || dimsregion.y > dims.y)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "region must be positive and fit the image");
  if ((long long)dimsregion.x * dimsregion.y >= (1ll << 31)) return fail(GTOMX_ERR_UNSUPPORTED, "region too large");
  if (batch == 0 || batch > 2147483647u) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch out of range");
  const int T = (int)std::min<long long>(128, (long long)dimsregion.x * dimsregion.y);
  hipLaunchKernelGGL(k_peak_one_2d, dim3(batch), dim3(T), 0, as_stream(stream), d_input, dims.x, dims.y,
                     dimsregion.x, dimsregion.y, subtractcenter ? 1 : 0, d_positions, d_values);
  GX_LAUNCHED("k_peak_one_2d");
  return ok();
}

extern "C" int gtomx_subpixel_max(const float* d_input, float* d_output, gtomx_int3 dims, int subpixsteps,
                                  void* stream) {
  if (!d_input || !d_output) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (dims.x < 1 || dims.y < 1 || dims.z < 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "dims must be positive");
  const int nd = dimension_count(dims);
  if (nd < 2) return fail(GTOMX_ERR_UNSUPPORTED, "1D maps (the reference throws)");
  if (dims.y == 1) return fail(GTOMX_ERR_UNSUPPORTED, "dims with y = 1 and z > 1");
  if (subpixsteps < 1 || subpixsteps > SPX_MAX_STEPS)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "subpixsteps must be in [1, 64]");
  // Voxels the reference's launch covers: grid ((nx-15)/128 rounded up, ny-15,
  // nz-15) from (6, 6, 6), guard x < nx-8, y < ny-8, z < nz-8.
  const int gx = (dims.x - 15 + 127) / 128, gy = dims.y - 15, gz = nd == 3 ? dims.z - 15 : 1;
  if (gx <= 0 || gy <= 0 || gz <= 0) return ok();
  const int xlo = SPX_R, ylo = SPX_R, zlo = nd == 3 ? SPX_R : 0;
  const int xhi = std::min(dims.x - SPX_R - 3, gx * 128 - 1 + SPX_R);
  const int yhi = std::min(dims.y - SPX_R - 3, gy - 1 + SPX_R);
  const int zhi = nd == 3 ? std::min(dims.z - SPX_R - 3, gz - 1 + SPX_R) : 0;
  // Weights sinc(s - k), s = i/steps - 1/2 + 1/(2 steps) in float as at
  // LocalPeaks.cu:70-74, evaluated in double and rounded.
  std::vector<float> w((size_t)subpixsteps * SPX_K);
  const float step = 1.0f / subpixsteps;
  for (int i = 0; i < subpixsteps; ++i) {
    const float a = (float)i * step;
    const float b = a - 0.5f;
    const float c = b + step / 2;
    for (int k = -SPX_R; k <= SPX_R; ++k) {
      const float d = c - (float)k;
      w[(size_t)i * SPX_K + k + SPX_R] =
          std::fabs(d) <= 1e-4f ? 1.0f : (float)(std::sin(PI_D * (double)d) / (PI_D * (double)d));
    }
  }
  hipStream_t s = as_stream(stream);
  DevBuf d_w;
  GX_TRY(d_w.alloc(w.size() * sizeof(float)));
  GX_HIP(hipMemcpyAsync(d_w.p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, s));
  if (nd == 3) {
    using T = SpxTile<3>;
    const dim3 grid((xhi - xlo + T::TX) / T::TX, (yhi - ylo + T::TY) / T::TY, (zhi - zlo + T::TZ) / T::TZ);
    hipLaunchKernelGGL((k_subpixel_max<3>), grid, dim3(256), 0, s, d_input, d_output, dims.x, dims.y, dims.z, xlo,
                       xhi, ylo, yhi, zlo, zhi, subpixsteps, d_w.as<float>());
  } else {
    using T = SpxTile<2>;
    const dim3 grid((xhi - xlo + T::TX) / T::TX, (yhi - ylo + T::TY) / T::TY, 1);
    hipLaunchKernelGGL((k_subpixel_max<2>), grid, dim3(256), 0, s, d_input, d_output, dims.x, dims.y, dims.z, xlo,
                       xhi, ylo, yhi, zlo, zhi, subpixsteps, d_w.as<float>());
  }
  GX_LAUNCHED("k_subpixel_max");
  GX_HIP(hipStreamSynchronize(s));
  return ok();
}
