// The learned orientation-proposal network: a small 3D CNN over 32^3 tiles of
// a tomogram block that scores 500 coarse directions (the Fibonacci-sphere
// directions of the search's orientation set) and one "no peak" class, trained
// on the exhaustive search's own argmax labels. The search then runs only the
// shortlisted directions' orientations on each tile (pick.hip, tile masks).
// There is no counterpart in the reference; its labels are the argmax
// orientation of UpdateCorrelationKernel (src/Correlation/SubTomograms.cu:138-158)
// divided by the 18 in-plane steps of a direction.
//
// Network (channels-last bf16 activations, bf16 weights, fp32 accumulation):
//   x (32^3 x 1) -> conv1 3^3 (1->16) + ReLU -> avgpool 2 -> conv2 (16->16) + ReLU
//   -> avgpool 2 -> conv3 (16->16) + ReLU -> conv4 (16->16) + ReLU -> avgpool 2
//   -> 1024 features (4^3 x 16) -> linear 1024 -> 501 logits.
// Every convolution, its data gradient and its weight gradient is an implicit
// GEMM on the matrix cores (v_mfma_f32_16x16x32_bf16): forward and data
// gradient with M = 16 voxels, N = 16 output channels, K = 2 taps x 16 input
// channels per instruction (14 instructions for the 27 taps, the last half
// zero); weight gradient with M = 16 output channels, N = 16 input channels (or
// 16 taps for the one-channel input), K = 32 voxels. The head is a generic
// strided bf16 GEMM on the same instruction. Parameters are fp32 masters with
// bf16 working copies refreshed after every Adam step.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "runtime.hpp"

namespace gx {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

constexpr int PT = 32;    // tile side
constexpr int PC = 16;    // channels
constexpr int NDIR = 500;
constexpr int NCLS = 501; // + "no peak"
constexpr int CPAD = 512; // padded class count
constexpr int NFEAT = 1024;
constexpr int NTAP = 27;

// Parameter layout of the fp32 master array (and of the gradients).
constexpr int OFF_W1 = 0;                        // [16 co][27 taps]
constexpr int OFF_B1 = OFF_W1 + PC * NTAP;       // [16]
constexpr int OFF_W2 = OFF_B1 + PC;              // [27 taps][16 co][16 ci]
constexpr int OFF_B2 = OFF_W2 + NTAP * PC * PC;
constexpr int OFF_W3 = OFF_B2 + PC;
constexpr int OFF_B3 = OFF_W3 + NTAP * PC * PC;
constexpr int OFF_W4 = OFF_B3 + PC;
constexpr int OFF_B4 = OFF_W4 + NTAP * PC * PC;
constexpr int OFF_WF = OFF_B4 + PC;              // [501][1024]
constexpr int OFF_BF = OFF_WF + NCLS * NFEAT;    // [501]
constexpr int NPARAM = OFF_BF + NCLS;

__device__ __forceinline__ u16 f2bf(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (u16)(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }
__device__ __forceinline__ bool bf_pos(u16 h) { return (h & 0x8000u) == 0 && (h & 0x7fffu) != 0; }

__device__ __forceinline__ f32x4 mfma(u16x8 a, u16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}
__device__ __forceinline__ u16x8 load8(const u16* p) { return *reinterpret_cast<const u16x8*>(p); }

// Tap t in [0, 27): offset (t/9 - 1, (t/3)%3 - 1, t%3 - 1) in (z, y, x); tap 26 - t is its negation.
__device__ __forceinline__ bool tap_src(int t, int S, int z, int y, int x, int& v) {
  const int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
  v = (zz * S + yy) * S + xx;
  return (unsigned)zz < (unsigned)S && (unsigned)yy < (unsigned)S && (unsigned)xx < (unsigned)S;
}

// ---------------------------------------------------------------------------
// Convolution (3^3, zero padding), forward or data gradient. One wave: 16
// consecutive voxels of one tile x 16 output channels. out = (acc + bias),
// ReLU if relu, zeroed where mask (a post-ReLU activation) is not positive.
// wb: CIN = 16: [27][16 out][16 in] bf16; CIN = 1: [16 out][32 taps] bf16.
// ---------------------------------------------------------------------------
template <int S, int CIN>
__global__ __launch_bounds__(256) void k_conv(const u16* __restrict__ in, const u16* __restrict__ wb,
                                              const float* __restrict__ bias, const u16* __restrict__ mask,
                                              u16* __restrict__ out, int ntiles, int relu) {
  constexpr int V = S * S * S, WPT = V / 16;
  const int lane = threadIdx.x & 63;
  const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wave >= (long)ntiles * WPT) return;
  const int tile = (int)(wave / WPT), v0 = (int)(wave % WPT) * 16;
  const int i = lane & 15, g = lane >> 4;
  const int v = v0 + i, x = v % S, y = (v / S) % S, z = v / (S * S);
  const u16* tin = in + (size_t)tile * V * CIN;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  if constexpr (CIN == 1) {
    u16x8 a;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int t = 8 * g + e;
      int src;
      a[e] = (t < NTAP && tap_src(t, S, z, y, x, src)) ? tin[src] : (u16)0;
    }
    acc = mfma(a, load8(wb + i * 32 + 8 * g), acc);
  } else {
    const int ci0 = 8 * (g & 1);
#pragma unroll
    for (int p = 0; p < 14; ++p) {
      const int t = 2 * p + (g >> 1);
      u16x8 a = {0, 0, 0, 0, 0, 0, 0, 0}, b = {0, 0, 0, 0, 0, 0, 0, 0};
      int src;
      if (t < NTAP) {
        b = load8(wb + (t * PC + i) * PC + ci0);
        if (tap_src(t, S, z, y, x, src)) a = load8(tin + (size_t)src * PC + ci0);
      }
      acc = mfma(a, b, acc);
    }
  }
  const int co = i;
  const float bb = bias ? bias[co] : 0.f;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const size_t o = ((size_t)tile * V + v0 + 4 * g + e) * PC + co;
    float val = acc[e] + bb;
    if (relu) val = fmaxf(val, 0.f);
    if (mask && !bf_pos(mask[o])) val = 0.f;
    out[o] = f2bf(val);
  }
}

// ---------------------------------------------------------------------------
// Weight gradient partial sums over a chunk of voxels (all tiles' voxels in
// one linear range): dW[t][co][ci] += dz[v][co] * in[v + off_t][ci], and the
// bias gradient dz[v][co]. One wave per chunk of `steps` x 32 voxels; its
// partials go to part[wave][...] (CIN = 16: 27*256 weights then 16 biases;
// CIN = 1: [16 co][27] then 16 biases).
// ---------------------------------------------------------------------------
template <int CIN> constexpr int dw_part_len() { return CIN == 1 ? PC * NTAP + PC : NTAP * PC * PC + PC; }

template <int S, int CIN>
__global__ __launch_bounds__(256) void k_conv_dw(const u16* __restrict__ in, const u16* __restrict__ dz,
                                                 float* __restrict__ part, long nvox, int steps, int nwaves) {
  constexpr int V = S * S * S;
  constexpr int NB = CIN == 1 ? 2 : NTAP;  // accumulators
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wave >= nwaves) return;
  const int i = lane & 15, g = lane >> 4;
  f32x4 acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = {0.f, 0.f, 0.f, 0.f};
  float bsum = 0.f;
  const long first = (long)wave * steps * 32;
  for (int st = 0; st < steps; ++st) {
    const long vb = first + (long)st * 32 + 8 * g;
    if (first + (long)st * 32 >= nvox) break;
    u16x8 a;
    int tl[8], zz[8], yy[8], xx[8];
    bool ok[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const long vv = vb + e;
      ok[e] = vv < nvox;
      const long vc = ok[e] ? vv : 0;
      tl[e] = (int)(vc / V);
      const int v = (int)(vc % V);
      xx[e] = v % S;
      yy[e] = (v / S) % S;
      zz[e] = v / (S * S);
      a[e] = ok[e] ? dz[vc * PC + i] : (u16)0;
      bsum += bf2f(a[e]);
    }
    if constexpr (CIN == 1) {
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        const int t = 16 * nb + i;
        u16x8 b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          int src;
          b[e] = (ok[e] && t < NTAP && tap_src(t, S, zz[e], yy[e], xx[e], src)) ? in[(size_t)tl[e] * V + src]
                                                                                 : (u16)0;
        }
        acc[nb] = mfma(a, b, acc[nb]);
      }
    } else {
#pragma unroll
      for (int t = 0; t < NTAP; ++t) {
        u16x8 b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          int src;
          b[e] = (ok[e] && tap_src(t, S, zz[e], yy[e], xx[e], src)) ? in[((size_t)tl[e] * V + src) * PC + i]
                                                                   : (u16)0;
        }
        acc[t] = mfma(a, b, acc[t]);
      }
    }
  }
  // D: column = lane & 15 (ci, or tap), rows co = 4g + e
  float* pw = part + (size_t)wave * dw_part_len<CIN>();
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int co = 4 * g + e;
      if constexpr (CIN == 1) {
        const int t = 16 * b + i;
        if (t < NTAP) pw[co * NTAP + t] = acc[b][e];
      } else {
        pw[(b * PC + co) * PC + i] = acc[b][e];
      }
    }
  // bias: lanes i, i+16, i+32, i+48 hold the four voxel groups of channel i
  bsum += __shfl_xor(bsum, 16);
  bsum += __shfl_xor(bsum, 32);
  if (g == 0) pw[dw_part_len<CIN>() - PC + i] = bsum;
}

// grad[off + k] = sum over waves of part[w][k], k < len; one workgroup per k.
__global__ __launch_bounds__(256) void k_reduce_parts(const float* __restrict__ part, int len, int nwaves,
                                                      float* __restrict__ grad) {
  __shared__ float red[256];
  const int k = blockIdx.x;
  float s = 0.f;
  for (int w = threadIdx.x; w < nwaves; w += 256) s += part[(size_t)w * len + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) grad[k] = red[0];
}

// ---------------------------------------------------------------------------
// Average pooling by 2 (channels-last) and its gradient with the ReLU mask of
// the pooled activation. pool: out rows [ntiles, nrows) are zeroed (they pad
// the head's GEMM). unpool: din = dout(parent) / 8 where a > 0.
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void k_pool(const u16* __restrict__ in, u16* __restrict__ out, int ntiles,
                                              int nrows) {
  constexpr int H = S / 2, VO = H * H * H;
  const long total = (long)nrows * VO * PC;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    const int c = (int)(id % PC);
    const long vo = id / PC;
    const int tile = (int)(vo / VO), r = (int)(vo % VO);
    if (tile >= ntiles) {
      out[id] = 0;
      continue;
    }
    const int x = r % H, y = (r / H) % H, z = r / (H * H);
    float s = 0.f;
#pragma unroll
    for (int d = 0; d < 8; ++d) {
      const int vi = ((2 * z + (d >> 2)) * S + 2 * y + ((d >> 1) & 1)) * S + 2 * x + (d & 1);
      s += bf2f(in[((size_t)tile * S * S * S + vi) * PC + c]);
    }
    out[id] = f2bf(s * 0.125f);
  }
}

template <int S, class T>
__global__ __launch_bounds__(256) void k_unpool(const T* __restrict__ dout, const u16* __restrict__ act,
                                                u16* __restrict__ din, int ntiles) {
  constexpr int H = S / 2, V = S * S * S;
  const long total = (long)ntiles * V * PC;
  for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < total; id += (long)gridDim.x * blockDim.x) {
    const int c = (int)(id % PC);
    const long vi = id / PC;
    const int tile = (int)(vi / V), r = (int)(vi % V);
    const int x = r % S, y = (r / S) % S, z = r / (S * S);
    const size_t po = ((size_t)tile * H * H * H + ((z / 2) * H + y / 2) * H + x / 2) * PC + c;
    float d;
    if constexpr (sizeof(T) == 4)
      d = dout[po];
    else
      d = bf2f(dout[po]);
    din[id] = bf_pos(act[id]) ? f2bf(d * 0.125f) : (u16)0;
  }
}

// ---------------------------------------------------------------------------
// C[m][n] = sum_k A(m, k) B(k, n) (+ bias[n]) with A(m, k) = A[m*sam + k*sak],
// B(k, n) = B[k*sbk + n*sbn] (bf16) and C fp32 at C[m*scm + n]; one wave per
// 16 x 16 block of C; rows m >= M and columns n >= N are not stored, operand
// entries past M, N or K read as zero.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gemm(const u16* __restrict__ A, long sam, long sak, const u16* __restrict__ B,
                                              long sbk, long sbn, float* __restrict__ C, long scm, int M, int N,
                                              int K, const float* __restrict__ bias) {
  const int lane = threadIdx.x & 63;
  const int nbn = (N + 15) / 16;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int mb = wave / nbn, nb = wave % nbn;
  if (mb * 16 >= M) return;
  const int i = lane & 15, g = lane >> 4;
  const int m = mb * 16 + i, n = nb * 16 + i;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 32) {
    u16x8 a, b;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = k0 + 8 * g + e;
      a[e] = (m < M && k < K) ? A[m * sam + k * sak] : (u16)0;
      b[e] = (n < N && k < K) ? B[k * sbk + n * sbn] : (u16)0;
    }
    acc = mfma(a, b, acc);
  }
  const int col = nb * 16 + i;
  if (col >= N) return;
  const float bb = bias ? bias[col] : 0.f;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int row = mb * 16 + 4 * g + e;
    if (row < M) C[row * scm + col] = acc[e] + bb;
  }
}

// ---------------------------------------------------------------------------
// Softmax cross-entropy over the 501 classes, one wave per row. Rows < n:
// dlog = (softmax - onehot(label)) / n (fp32 and bf16), loss[r]; rows in
// [n, nrows): zeros.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_softmax_ce(const float* __restrict__ logits, const int* __restrict__ labels,
                                                    int n, int nrows, float* __restrict__ dlog,
                                                    u16* __restrict__ dlogb, float* __restrict__ loss) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= nrows) return;
  float* dr = dlog + (size_t)r * CPAD;
  u16* db = dlogb + (size_t)r * CPAD;
  if (r >= n) {
    for (int k = lane; k < CPAD; k += 64) {
      dr[k] = 0.f;
      db[k] = 0;
    }
    return;
  }
  const float* lr = logits + (size_t)r * CPAD;
  float mx = -INFINITY;
  for (int k = lane; k < NCLS; k += 64) mx = fmaxf(mx, lr[k]);
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
  float s = 0.f;
  for (int k = lane; k < NCLS; k += 64) s += expf(lr[k] - mx);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  const float lse = mx + logf(s);
  const int lab = min(max(labels[r], 0), NCLS - 1);  // labels outside [0, 501) are clamped
  const float inv = 1.f / (float)n;
  for (int k = lane; k < CPAD; k += 64) {
    float d = 0.f;
    if (k < NCLS) d = (expf(lr[k] - lse) - (k == lab ? 1.f : 0.f)) * inv;
    dr[k] = d;
    db[k] = f2bf(d);
  }
  if (lane == 0) loss[r] = lse - lr[lab];
}

// out[k] = sum over rows of in[r*ld + k], k < len (one workgroup per k).
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ in, int rows, int ld, float* __restrict__ out) {
  __shared__ float red[256];
  const int k = blockIdx.x;
  float s = 0.f;
  for (int r = threadIdx.x; r < rows; r += 256) s += in[(size_t)r * ld + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

// ---------------------------------------------------------------------------
// Proposals: softmax top-1 probability and class over the 501 classes, the
// top-k of the 500 direction logits (larger logit first, ties to the lower
// index), the summed probability of those k directions (mass) and the
// probability of "no peak" (pnp). One wave per row.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_propose(const float* __restrict__ logits, int n, int topk,
                                                 int* __restrict__ dirs, float* __restrict__ top1,
                                                 int* __restrict__ cls, float* __restrict__ mass,
                                                 float* __restrict__ pnp) {
  const int lane = threadIdx.x & 63;
  const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= n) return;
  const float* lr = logits + (size_t)r * CPAD;
  float v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = lane + 64 * q;
    v[q] = k < NCLS ? lr[k] : -INFINITY;
  }
  float mx = -INFINITY;
  int mk = 0x7fffffff;
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (v[q] > mx) {
      mx = v[q];
      mk = lane + 64 * q;
    }
  for (int o = 32; o > 0; o >>= 1) {
    const float om = __shfl_xor(mx, o);
    const int ok = __shfl_xor(mk, o);
    if (om > mx || (om == mx && ok < mk)) {
      mx = om;
      mk = ok;
    }
  }
  float s = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) s += expf(v[q] - mx);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) {
    top1[r] = 1.f / s;
    cls[r] = mk;
    if (pnp) pnp[r] = expf(lr[NDIR] - mx) / s;
  }
  // the background class takes no part in the direction ranking
  for (int q = 0; q < 8; ++q)
    if (lane + 64 * q >= NDIR) v[q] = -INFINITY;
  float msum = 0.f;
  for (int t = 0; t < topk; ++t) {
    float bm = -INFINITY;
    int bk = 0x7fffffff;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (v[q] > bm) {
        bm = v[q];
        bk = lane + 64 * q;
      }
    for (int o = 32; o > 0; o >>= 1) {
      const float om = __shfl_xor(bm, o);
      const int ok = __shfl_xor(bk, o);
      if (om > bm || (om == bm && ok < bk)) {
        bm = om;
        bk = ok;
      }
    }
    if (lane == 0) dirs[(size_t)r * topk + t] = bk < NDIR ? bk : -1;
    if (bk < NDIR) msum += expf(bm - mx) / s;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (lane + 64 * q == bk) v[q] = -INFINITY;
  }
  if (lane == 0 && mass) mass[r] = msum;
}

// ---------------------------------------------------------------------------
// Adam on the fp32 masters, then the bf16 working copies.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ gr,
                                              float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                              float b1, float b2, float eps, float c1, float c2, float wd) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = gr[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
}

// Working copies: w1b [16][32]; wNb [27][16 co][16 ci]; wNt [27][16 ci][16 co]
// with wNt[t][a][b] = W[26 - t][b][a] (the data-gradient convolution); wfb
// [512][1024] with rows past 501 zero.
__global__ __launch_bounds__(256) void k_refresh(const float* __restrict__ p, u16* __restrict__ w1b,
                                                 u16* __restrict__ wcb, u16* __restrict__ wct,
                                                 u16* __restrict__ wfb) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  const int nconv = 3 * NTAP * PC * PC;
  if (id < PC * 32) {
    const int co = id / 32, t = id % 32;
    w1b[id] = t < NTAP ? f2bf(p[OFF_W1 + co * NTAP + t]) : (u16)0;
  }
  if (id < nconv) {
    const int l = id / (NTAP * PC * PC), r = id % (NTAP * PC * PC);
    const int t = r / (PC * PC), a = (r / PC) % PC, b = r % PC;
    const float* W = p + (l == 0 ? OFF_W2 : l == 1 ? OFF_W3 : OFF_W4);
    wcb[id] = f2bf(W[r]);
    wct[id] = f2bf(W[((NTAP - 1 - t) * PC + b) * PC + a]);
  }
  if (id < CPAD * NFEAT) {
    const int c = id / NFEAT;
    wfb[id] = c < NCLS ? f2bf(p[OFF_WF + id]) : (u16)0;
  }
}

// ---------------------------------------------------------------------------
// Tiles of an n^3 volume (n a multiple of 32), standardised by the volume's
// mean and standard deviation (double sums) and rounded to bf16:
// tiles[(tz*nt + ty)*nt + tx][z][y][x].
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_moments(const float* __restrict__ vol, size_t n, double* __restrict__ part) {
  __shared__ double r1[256], r2[256];
  double s1 = 0, s2 = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double x = vol[i];
    s1 += x;
    s2 += x * x;
  }
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      r1[threadIdx.x] += r1[threadIdx.x + h];
      r2[threadIdx.x] += r2[threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
  }
}

__global__ __launch_bounds__(256) void k_tiles(const float* __restrict__ vol, int n, const double* __restrict__ part,
                                               int nparts, u16* __restrict__ tiles) {
  __shared__ float stat[2];
  if (threadIdx.x == 0) {
    double s1 = 0, s2 = 0;
    for (int k = 0; k < nparts; ++k) {
      s1 += part[2 * k];
      s2 += part[2 * k + 1];
    }
    const double N = (double)n * n * n, mean = s1 / N;
    const double var = fmax(s2 / N - mean * mean, 0.0);
    stat[0] = (float)mean;
    stat[1] = var > 0 ? (float)(1.0 / sqrt(var)) : 0.f;
  }
  __syncthreads();
  const int nt = n / PT;
  const size_t total = (size_t)n * n * n;
  for (size_t id = blockIdx.x * (size_t)blockDim.x + threadIdx.x; id < total; id += (size_t)gridDim.x * blockDim.x) {
    const int x = (int)(id % n), y = (int)((id / n) % n), z = (int)(id / ((size_t)n * n));
    const size_t tile = ((size_t)(z / PT) * nt + y / PT) * nt + x / PT;
    const size_t o = tile * PT * PT * PT + ((size_t)(z % PT) * PT + y % PT) * PT + x % PT;
    tiles[o] = f2bf((vol[id] - stat[0]) * stat[1]);
  }
}

__global__ __launch_bounds__(256) void k_f32_to_bf16(const float* __restrict__ in, u16* __restrict__ out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = f2bf(in[i]);
}

unsigned grid_of(size_t n, unsigned cap = 65535) { return (unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, cap)); }

}  // namespace
}  // namespace gx

using namespace gx;

// ---------------------------------------------------------------------------
// Network state
// ---------------------------------------------------------------------------
struct gtomx_pnet {
  int device = 0;
  unsigned cap = 0;   // tiles per pass
  unsigned rows = 0;  // cap rounded up to 16 (head GEMM rows)
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // NPARAM each
  long step = 0;
  u16 *w1b = nullptr, *wcb = nullptr, *wct = nullptr, *wfb = nullptr;
  u16 *x = nullptr, *a1 = nullptr, *p1 = nullptr, *a2 = nullptr, *p2 = nullptr, *a3 = nullptr, *a4 = nullptr,
      *feat = nullptr;
  float* logits = nullptr;
  float* dlog = nullptr;
  u16* dlogb = nullptr;
  float* dfeat = nullptr;
  float* loss = nullptr;
  u16 *dz4 = nullptr, *dz3 = nullptr, *dp2 = nullptr, *dz2 = nullptr, *dp1 = nullptr, *dz1 = nullptr;
  float* part = nullptr;
  size_t part_floats = 0;
  double* mom = nullptr;  // moments partials of k_moments
  unsigned nlast = 0;     // tiles of the last forward pass
  size_t device_bytes = 0;
  StreamFence fence;
};

namespace gx {
namespace {

constexpr int MOM_BLOCKS = 512;

void pnet_release(gtomx_pnet* nt) {
  if (!nt) return;
  nt->fence.release();
  for (void* q : {(void*)nt->p, (void*)nt->g, (void*)nt->m, (void*)nt->v, (void*)nt->w1b, (void*)nt->wcb,
                  (void*)nt->wct, (void*)nt->wfb, (void*)nt->x, (void*)nt->a1, (void*)nt->p1, (void*)nt->a2,
                  (void*)nt->p2, (void*)nt->a3, (void*)nt->a4, (void*)nt->feat, (void*)nt->logits, (void*)nt->dlog,
                  (void*)nt->dlogb, (void*)nt->dfeat, (void*)nt->loss, (void*)nt->dz4, (void*)nt->dz3,
                  (void*)nt->dp2, (void*)nt->dz2, (void*)nt->dp1, (void*)nt->dz1, (void*)nt->part, (void*)nt->mom})
    dev_free(q);
  delete nt;
}

int pnet_check(const gtomx_pnet* nt) {
  if (!nt) return fail(GTOMX_ERR_INVALID_ARGUMENT, "network is NULL");
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  if (dev != nt->device) return fail(GTOMX_ERR_INVALID_ARGUMENT, "network was created on another device");
  return GTOMX_OK;
}

int refresh(gtomx_pnet* nt, hipStream_t s) {
  hipLaunchKernelGGL(k_refresh, dim3(grid_of(CPAD * NFEAT)), dim3(256), 0, s, nt->p, nt->w1b, nt->wcb, nt->wct,
                     nt->wfb);
  GX_LAUNCHED("k_refresh");
  return GTOMX_OK;
}

template <int S, int CIN>
int conv(const u16* in, const u16* wb, const float* bias, const u16* mask, u16* out, unsigned n, int relu,
         hipStream_t s) {
  const long waves = (long)n * (S * S * S / 16);
  hipLaunchKernelGGL((k_conv<S, CIN>), dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, s, in, wb, bias, mask, out,
                     (int)n, relu);
  GX_LAUNCHED("k_conv");
  return GTOMX_OK;
}

// Weight and bias gradient of one convolution into grad + woff (weights) and grad + boff (bias).
template <int S, int CIN>
int conv_dw(gtomx_pnet* nt, const u16* in, const u16* dz, unsigned n, float* wgrad, float* bgrad, hipStream_t s) {
  const long nvox = (long)n * S * S * S;
  const int len = dw_part_len<CIN>();
  // about 2048 waves, each at least 4 steps of 32 voxels
  int steps = (int)std::max<long>(4, (nvox + 2048L * 32 - 1) / (2048L * 32));
  int nw = (int)((nvox + (long)steps * 32 - 1) / ((long)steps * 32));
  if ((size_t)nw * len > nt->part_floats) return fail(GTOMX_ERR_INTERNAL, "gradient partials exceed the scratch");
  hipLaunchKernelGGL((k_conv_dw<S, CIN>), dim3((nw + 3) / 4), dim3(256), 0, s, in, dz, nt->part, nvox, steps, nw);
  GX_LAUNCHED("k_conv_dw");
  const int wl = len - PC;
  hipLaunchKernelGGL(k_reduce_parts, dim3(wl), dim3(256), 0, s, nt->part, len, nw, wgrad);
  GX_LAUNCHED("k_reduce_parts");
  hipLaunchKernelGGL(k_reduce_parts, dim3(PC), dim3(256), 0, s, nt->part + wl, len, nw, bgrad);
  GX_LAUNCHED("k_reduce_parts");
  return GTOMX_OK;
}

int gemm(const u16* A, long sam, long sak, const u16* B, long sbk, long sbn, float* C, long scm, int M, int N, int K,
         const float* bias, hipStream_t s) {
  const int waves = ((M + 15) / 16) * ((N + 15) / 16);
  hipLaunchKernelGGL(k_gemm, dim3((waves + 3) / 4), dim3(256), 0, s, A, sam, sak, B, sbk, sbn, C, scm, M, N, K, bias);
  GX_LAUNCHED("k_gemm");
  return GTOMX_OK;
}

template <int S, class T> int unpool(const T* dout, const u16* act, u16* din, unsigned n, hipStream_t s) {
  hipLaunchKernelGGL((k_unpool<S, T>), dim3(grid_of((size_t)n * S * S * S * PC)), dim3(256), 0, s, dout, act, din,
                     (int)n);
  GX_LAUNCHED("k_unpool");
  return GTOMX_OK;
}

template <int S> int pool(const u16* in, u16* out, unsigned n, unsigned rows, hipStream_t s) {
  hipLaunchKernelGGL((k_pool<S>), dim3(grid_of((size_t)rows * (S / 2) * (S / 2) * (S / 2) * PC)), dim3(256), 0, s,
                     in, out, (int)n, (int)rows);
  GX_LAUNCHED("k_pool");
  return GTOMX_OK;
}

unsigned rows_of(unsigned n) { return (n + 15) / 16 * 16; }

// Forward pass over n tiles already in nt->x.
int forward(gtomx_pnet* nt, unsigned n, hipStream_t s) {
  const unsigned rows = rows_of(n);
  GX_TRY((conv<32, 1>(nt->x, nt->w1b, nt->p + OFF_B1, nullptr, nt->a1, n, 1, s)));
  GX_TRY(pool<32>(nt->a1, nt->p1, n, n, s));
  GX_TRY((conv<16, 16>(nt->p1, nt->wcb, nt->p + OFF_B2, nullptr, nt->a2, n, 1, s)));
  GX_TRY(pool<16>(nt->a2, nt->p2, n, n, s));
  GX_TRY((conv<8, 16>(nt->p2, nt->wcb + NTAP * PC * PC, nt->p + OFF_B3, nullptr, nt->a3, n, 1, s)));
  GX_TRY((conv<8, 16>(nt->a3, nt->wcb + 2 * NTAP * PC * PC, nt->p + OFF_B4, nullptr, nt->a4, n, 1, s)));
  GX_TRY(pool<8>(nt->a4, nt->feat, n, rows, s));
  // logits[r][c] = feat[r] . Wf[c] + bf[c] for the 501 classes (columns 501..511 stay zero)
  GX_TRY(gemm(nt->feat, NFEAT, 1, nt->wfb, 1, NFEAT, nt->logits, CPAD, (int)rows, NCLS, NFEAT, nt->p + OFF_BF, s));
  nt->nlast = n;
  return GTOMX_OK;
}

// Backward pass from the labels (after forward over the same n tiles); the
// gradients land in nt->g, the per-tile losses in nt->loss.
int backward(gtomx_pnet* nt, const int* d_labels, unsigned n, hipStream_t s) {
  const unsigned rows = rows_of(n);
  hipLaunchKernelGGL(k_softmax_ce, dim3((rows + 3) / 4), dim3(256), 0, s, nt->logits, d_labels, (int)n, (int)rows,
                     nt->dlog, nt->dlogb, nt->loss);
  GX_LAUNCHED("k_softmax_ce");
  hipLaunchKernelGGL(k_colsum, dim3(NCLS), dim3(256), 0, s, nt->dlog, (int)rows, CPAD, nt->g + OFF_BF);
  GX_LAUNCHED("k_colsum");
  // dWf[c][f] = sum_r dlog[r][c] feat[r][f]
  GX_TRY(gemm(nt->dlogb, 1, CPAD, nt->feat, NFEAT, 1, nt->g + OFF_WF, NFEAT, NCLS, NFEAT, (int)rows, nullptr, s));
  // dfeat[r][f] = sum_c dlog[r][c] Wf[c][f]
  GX_TRY(gemm(nt->dlogb, CPAD, 1, nt->wfb, NFEAT, 1, nt->dfeat, NFEAT, (int)rows, NFEAT, CPAD, nullptr, s));
  GX_TRY((unpool<8, float>(nt->dfeat, nt->a4, nt->dz4, n, s)));
  GX_TRY((conv_dw<8, 16>(nt, nt->a3, nt->dz4, n, nt->g + OFF_W4, nt->g + OFF_B4, s)));
  GX_TRY((conv<8, 16>(nt->dz4, nt->wct + 2 * NTAP * PC * PC, nullptr, nt->a3, nt->dz3, n, 0, s)));
  GX_TRY((conv_dw<8, 16>(nt, nt->p2, nt->dz3, n, nt->g + OFF_W3, nt->g + OFF_B3, s)));
  GX_TRY((conv<8, 16>(nt->dz3, nt->wct + NTAP * PC * PC, nullptr, nullptr, nt->dp2, n, 0, s)));
  GX_TRY((unpool<16, u16>(nt->dp2, nt->a2, nt->dz2, n, s)));
  GX_TRY((conv_dw<16, 16>(nt, nt->p1, nt->dz2, n, nt->g + OFF_W2, nt->g + OFF_B2, s)));
  GX_TRY((conv<16, 16>(nt->dz2, nt->wct, nullptr, nullptr, nt->dp1, n, 0, s)));
  GX_TRY((unpool<32, u16>(nt->dp1, nt->a1, nt->dz1, n, s)));
  GX_TRY((conv_dw<32, 1>(nt, nt->x, nt->dz1, n, nt->g + OFF_W1, nt->g + OFF_B1, s)));
  return GTOMX_OK;
}

}  // namespace
}  // namespace gx

extern "C" {

int gtomx_pnet_param_count(void) { return NPARAM; }

int gtomx_pnet_create(gtomx_pnet** out, unsigned max_tiles) {
  if (!out) return fail(GTOMX_ERR_INVALID_ARGUMENT, "network pointer is NULL");
  *out = nullptr;
  if (max_tiles == 0 || max_tiles > 65536) return fail(GTOMX_ERR_INVALID_ARGUMENT, "max_tiles must be in [1, 65536]");
  auto* nt = new gtomx_pnet();
  nt->cap = max_tiles;
  nt->rows = rows_of(max_tiles);
  hipError_t he = hipGetDevice(&nt->device);
  if (he != hipSuccess) {
    pnet_release(nt);
    return hip_fail(he, "hipGetDevice");
  }
  int st = GTOMX_OK;
  auto alloc = [&](auto** dst, size_t bytes) {
    if (st != GTOMX_OK) return;
    void* d = nullptr;
    st = dev_alloc(&d, bytes);
    if (st == GTOMX_OK) {
      *dst = static_cast<std::remove_reference_t<decltype(*dst)>>(d);
      nt->device_bytes += bytes;
    }
  };
  const size_t R = nt->rows, C = nt->cap;
  const size_t V1 = 32 * 32 * 32, V2 = 16 * 16 * 16, V3 = 8 * 8 * 8;
  for (float** q : {&nt->p, &nt->g, &nt->m, &nt->v}) alloc(q, NPARAM * sizeof(float));
  alloc(&nt->w1b, PC * 32 * 2);
  alloc(&nt->wcb, 3 * NTAP * PC * PC * 2);
  alloc(&nt->wct, 3 * NTAP * PC * PC * 2);
  alloc(&nt->wfb, (size_t)CPAD * NFEAT * 2);
  alloc(&nt->x, C * V1 * 2);
  alloc(&nt->a1, C * V1 * PC * 2);
  alloc(&nt->p1, C * V2 * PC * 2);
  alloc(&nt->a2, C * V2 * PC * 2);
  alloc(&nt->p2, C * V3 * PC * 2);
  alloc(&nt->a3, C * V3 * PC * 2);
  alloc(&nt->a4, C * V3 * PC * 2);
  alloc(&nt->feat, R * NFEAT * 2);
  alloc(&nt->logits, R * CPAD * 4);
  alloc(&nt->dlog, R * CPAD * 4);
  alloc(&nt->dlogb, R * CPAD * 2);
  alloc(&nt->dfeat, R * NFEAT * 4);
  alloc(&nt->loss, R * 4);
  alloc(&nt->dz4, C * V3 * PC * 2);
  alloc(&nt->dz3, C * V3 * PC * 2);
  alloc(&nt->dp2, C * V3 * PC * 2);
  alloc(&nt->dz2, C * V2 * PC * 2);
  alloc(&nt->dp1, C * V2 * PC * 2);
  alloc(&nt->dz1, C * V1 * PC * 2);
  nt->part_floats = (size_t)2100 * (NTAP * PC * PC + PC);
  alloc(&nt->part, nt->part_floats * 4);
  alloc(&nt->mom, (size_t)MOM_BLOCKS * 2 * sizeof(double));
  if (st != GTOMX_OK) {
    std::string msg = gtomx_last_error();
    pnet_release(nt);
    return fail(st, msg);
  }
  // zero parameters and optimiser state until gtomx_pnet_set_params
  for (float* q : {nt->p, nt->g, nt->m, nt->v}) he = he == hipSuccess ? hipMemset(q, 0, NPARAM * sizeof(float)) : he;
  if (he == hipSuccess) he = hipMemset(nt->logits, 0, R * CPAD * 4);
  if (he == hipSuccess) he = hipStreamSynchronize(nullptr);  // the clears ran on the null stream
  if (he != hipSuccess) {
    pnet_release(nt);
    return hip_fail(he, "network set-up");
  }
  *out = nt;
  return GTOMX_OK;
}

int gtomx_pnet_destroy(gtomx_pnet* net) {
  pnet_release(net);
  return GTOMX_OK;
}

int gtomx_pnet_set_params(gtomx_pnet* nt, const float* h_params, size_t count, void* stream) {
  GX_TRY(pnet_check(nt));
  if (!h_params || count != (size_t)NPARAM) return fail(GTOMX_ERR_INVALID_ARGUMENT, "expected all parameters");
  hipStream_t s = as_stream(stream);
  GX_TRY(nt->fence.enter(s));
  GX_HIP(hipMemcpyAsync(nt->p, h_params, NPARAM * sizeof(float), hipMemcpyHostToDevice, s));
  GX_HIP(hipMemsetAsync(nt->m, 0, NPARAM * sizeof(float), s));
  GX_HIP(hipMemsetAsync(nt->v, 0, NPARAM * sizeof(float), s));
  nt->step = 0;
  GX_TRY(refresh(nt, s));
  GX_TRY(nt->fence.leave(s));
  GX_HIP(hipStreamSynchronize(s));  // the host array may go away on return
  return GTOMX_OK;
}

int gtomx_pnet_get_params(gtomx_pnet* nt, float* h_params, size_t count, void* stream) {
  GX_TRY(pnet_check(nt));
  if (!h_params || count != (size_t)NPARAM) return fail(GTOMX_ERR_INVALID_ARGUMENT, "expected all parameters");
  hipStream_t s = as_stream(stream);
  GX_TRY(nt->fence.enter(s));
  GX_HIP(hipMemcpyAsync(h_params, nt->p, NPARAM * sizeof(float), hipMemcpyDeviceToHost, s));
  GX_TRY(nt->fence.leave(s));
  GX_HIP(hipStreamSynchronize(s));
  return GTOMX_OK;
}

// The network's input from n tiles of 32^3 floats (already standardised), rounded to bf16.
int gtomx_pnet_load_tiles(gtomx_pnet* nt, const float* d_tiles, unsigned n, void* stream) {
  GX_TRY(pnet_check(nt));
  if (!d_tiles || n == 0 || n > nt->cap) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL tiles or n out of range");
  hipStream_t s = as_stream(stream);
  GX_TRY(nt->fence.enter(s));
  const size_t cnt = (size_t)n * PT * PT * PT;
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_of(cnt)), dim3(256), 0, s, d_tiles, nt->x, cnt);
  GX_LAUNCHED("k_f32_to_bf16");
  return nt->fence.leave(s);
}

// The network's input from every 32^3 tile of an n^3 volume (n a multiple of 32, (n/32)^3 <= max_tiles),
// standardised by the volume's mean and standard deviation; tile (tz*nt + ty)*nt + tx.
int gtomx_pnet_load_volume(gtomx_pnet* nt, const float* d_volume, gtomx_int3 dims, void* stream) {
  GX_TRY(pnet_check(nt));
  if (!d_volume) return fail(GTOMX_ERR_INVALID_ARGUMENT, "d_volume is NULL");
  if (dims.x != dims.y || dims.x != dims.z || dims.x % PT || dims.x < PT)
    return fail(GTOMX_ERR_UNSUPPORTED, "the volume must be a cube whose side is a multiple of 32");
  const size_t ntl = (size_t)(dims.x / PT) * (dims.x / PT) * (dims.x / PT);
  if (ntl > nt->cap) return fail(GTOMX_ERR_INVALID_ARGUMENT, "more tiles than the network's max_tiles");
  hipStream_t s = as_stream(stream);
  GX_TRY(nt->fence.enter(s));
  const size_t V = elements(dims);
  hipLaunchKernelGGL(k_moments, dim3(MOM_BLOCKS), dim3(256), 0, s, d_volume, V, nt->mom);
  GX_LAUNCHED("k_moments");
  hipLaunchKernelGGL(k_tiles, dim3(grid_of(V)), dim3(256), 0, s, d_volume, dims.x, nt->mom, MOM_BLOCKS, nt->x);
  GX_LAUNCHED("k_tiles");
  return nt->fence.leave(s);
}

// Forward pass over the first n loaded tiles: d_logits (optional) receives n x 512 floats (501 used).
int gtomx_pnet_forward(gtomx_pnet* nt, unsigned n, float* d_logits, void* stream) {
  GX_TRY(pnet_check(nt));
  if (n == 0 || n > nt->cap) return fail(GTOMX_ERR_INVALID_ARGUMENT, "n out of range");
  hipStream_t s = as_stream(stream);
  GX_TRY(nt->fence.enter(s));
  GX_TRY(forward(nt, n, s));
  if (d_logits) GX_HIP(hipMemcpyAsync(d_logits, nt->logits, (size_t)n * CPAD * 4, hipMemcpyDeviceToDevice, s));
  return nt->fence.leave(s);
}

// Loss and gradients of the loaded tiles against labels (class per tile): d_grads (optional) receives
// the NPARAM gradients, d_loss (optional) the n per-tile losses.
int gtomx_pnet_gradients(gtomx_pnet* nt, const int32_t* d_labels, unsigned n, float* d_grads, float* d_loss,
                         void* stream) {
  GX_TRY(pnet_check(nt));
  if (!d_labels || n == 0 || n > nt->cap) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL labels or n out of range");
  hipStream_t s = as_stream(stream);
  GX_TRY(nt->fence.enter(s));
  GX_TRY(forward(nt, n, s));
  GX_TRY(backward(nt, d_labels, n, s));
  if (d_grads) GX_HIP(hipMemcpyAsync(d_grads, nt->g, NPARAM * 4, hipMemcpyDeviceToDevice, s));
  if (d_loss) GX_HIP(hipMemcpyAsync(d_loss, nt->loss, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  return nt->fence.leave(s);
}

// One Adam step (beta 0.9 / 0.999, eps 1e-8, decoupled-free L2 weight decay wd) on the loaded tiles.
int gtomx_pnet_train_step(gtomx_pnet* nt, const int32_t* d_labels, unsigned n, float lr, float wd, float* d_loss,
                          void* stream) {
  GX_TRY(pnet_check(nt));
  if (!d_labels || n == 0 || n > nt->cap) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL labels or n out of range");
  if (!(lr >= 0.f) || !std::isfinite(lr)) return fail(GTOMX_ERR_INVALID_ARGUMENT, "lr must be finite, >= 0");
  hipStream_t s = as_stream(stream);
  GX_TRY(nt->fence.enter(s));
  GX_TRY(forward(nt, n, s));
  GX_TRY(backward(nt, d_labels, n, s));
  nt->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)nt->step)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)nt->step)));
  hipLaunchKernelGGL(k_adam, dim3(grid_of(NPARAM, 2048)), dim3(256), 0, s, nt->p, nt->g, nt->m, nt->v, NPARAM, lr, b1,
                     b2, 1e-8f, c1, c2, wd);
  GX_LAUNCHED("k_adam");
  GX_TRY(refresh(nt, s));
  if (d_loss) GX_HIP(hipMemcpyAsync(d_loss, nt->loss, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
  return nt->fence.leave(s);
}

// Proposals from the last forward pass: per tile the top-k directions (d_dirs, n x topk), the softmax
// probability of the top class (d_top1) and that class (d_class; 500 = no peak), and optionally the
// summed probability of the k directions (d_mass) and the probability of "no peak" (d_nopeak).
int gtomx_pnet_propose(gtomx_pnet* nt, unsigned topk, int32_t* d_dirs, float* d_top1, int32_t* d_class,
                       float* d_mass, float* d_nopeak, void* stream) {
  GX_TRY(pnet_check(nt));
  if (nt->nlast == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "no forward pass has run");
  if (topk == 0 || topk > (unsigned)NDIR) return fail(GTOMX_ERR_INVALID_ARGUMENT, "topk must be in [1, 500]");
  if (!d_dirs || !d_top1 || !d_class) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL output buffer");
  hipStream_t s = as_stream(stream);
  GX_TRY(nt->fence.enter(s));
  hipLaunchKernelGGL(k_propose, dim3((nt->nlast + 3) / 4), dim3(256), 0, s, nt->logits, (int)nt->nlast, (int)topk,
                     d_dirs, d_top1, d_class, d_mass, d_nopeak);
  GX_LAUNCHED("k_propose");
  return nt->fence.leave(s);
}

// The head's GEMM on its own (bf16 operands from float inputs, fp32 result), for tests of the
// matrix-core operand layout: C (M x N, row stride N) = A (M x K) . B (K x N), both row-major.
int gtomx_pnet_debug_gemm(const float* d_a, const float* d_b, float* d_c, int M, int N, int K, void* stream) {
  if (!d_a || !d_b || !d_c || M <= 0 || N <= 0 || K <= 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "bad GEMM");
  hipStream_t s = as_stream(stream);
  DevBuf ab, bb;
  GX_TRY(ab.alloc((size_t)M * K * 2));
  GX_TRY(bb.alloc((size_t)K * N * 2));
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_of((size_t)M * K)), dim3(256), 0, s, d_a, ab.as<u16>(), (size_t)M * K);
  GX_LAUNCHED("k_f32_to_bf16");
  hipLaunchKernelGGL(k_f32_to_bf16, dim3(grid_of((size_t)K * N)), dim3(256), 0, s, d_b, bb.as<u16>(), (size_t)K * N);
  GX_LAUNCHED("k_f32_to_bf16");
  GX_TRY(gemm(ab.as<u16>(), K, 1, bb.as<u16>(), N, 1, d_c, N, M, N, K, nullptr, s));
  GX_HIP(hipStreamSynchronize(s));
  return GTOMX_OK;
}

}  // extern "C"
