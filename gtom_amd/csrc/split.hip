// Split-precision form of the exhaustive search on 64^3 volumes: the same
// dataflow as d_PickSubTomograms (src/Correlation/SubTomograms.cu:19-158) with
// every per-orientation inverse transform (the C2R of SubTomograms.cu:81) run
// as 64 = 16 x 4 with its radix-16 stage on the matrix cores.
//
// A 16-point DFT of 16 columns is one real 32 x 32 by 32 x 16 product,
//   [Yr; Yi] = [[Fr, -Fi], [Fi, Fr]] [Xr; Xi],  F[k][n] = exp(+2 pi i k n / 16),
// two v_mfma_f32_16x16x32_{f16,bf16} tiles (rows Yr, rows Yi) with K = 32. Each
// operand is split as a = hi + lo with hi = round(a), lo = round(a - hi) in the
// 16-bit type, and the product is hi*hi + hi*lo + lo*hi accumulated in fp32
// (the lo*lo term, about 2^-22 of the product in fp16 and 2^-16 in bf16, is
// dropped). Each column (a 16-point subsequence) is scaled by a power of two
// to a maximum magnitude in [1/2, 1) before the split, so fp16 neither
// overflows nor loses lo to underflow, and scaled back exactly afterwards.
// The twiddles exp(+2 pi i q k1 / 64) and the radix-4 stage across the four
// subsequences (lanes of one DPP quad) stay in fp32. With SPLIT = false only
// hi*hi runs, which shows what the compensation buys.
//
// Per batch of orientations:
//   k_split_project: R = project(M_o k) * ctf (src/Relion/Project.cu:89-177,
//     src/Generics/Arithmetics.cu:275-286), Q = E * conj(R) with DC = 0
//     (SubTomograms.cu:127-136) into the batch buffer ([z][y][x], x < 33), and
//     partial sums of |Herm R|^2 (the Fourier form of the reference's
//     C2R -> mean0/std1 -> R2C -> /RMS, src/ImageManipulation/Norm.cu:476-509);
//   k_split_scale: 1 / max(RMS, 1e-20) per orientation;
//   k_split_lines (z, then y): inverse 64-point DFTs in place;
//   k_split_x_max: Hermitian completion of each x row, inverse 64-point DFT,
//     real part times the scale, running max/argmax with strict '>' in
//     increasing orientation (SubTomograms.cu:138-158).
// The search's begin and finish (local-std division) are the plain path's.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "dev_fft.hpp"
#include "dev_project.hpp"
#include "pick_plan.hpp"
#include "runtime.hpp"

namespace gx {
namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int SN = 64;            // transform length along every axis
constexpr int SNH = SN / 2 + 1;   // 33 stored x bins
constexpr int SPLANE = SNH * SN;  // 2112 bins per z plane
constexpr int SH = SPLANE * SN;   // 135168 bins per orientation
constexpr int LP = SN + 1;        // LDS pitch of a line (complex), odd against bank conflicts
constexpr int PROJ_BLOCKS = 64;   // projection workgroups per orientation

struct F16 {
  static __device__ __forceinline__ unsigned short bits(float f) {
    return __builtin_bit_cast(unsigned short, (_Float16)f);
  }
  static __device__ __forceinline__ float value(unsigned short b) { return (float)__builtin_bit_cast(_Float16, b); }
  static __device__ __forceinline__ f32x4 mfma(u16x8 a, u16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                  0, 0);
  }
};

struct BF16 {
  static __device__ __forceinline__ unsigned short bits(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (unsigned short)(u >> 16);
  }
  static __device__ __forceinline__ float value(unsigned short b) { return __uint_as_float((unsigned)b << 16); }
  static __device__ __forceinline__ f32x4 mfma(u16x8 a, u16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
  }
};

// cos(pi m / 8), m in [0, 16); sin(pi m / 8) = COS16[(m + 12) & 15].
__constant__ double COS16[16] = {1.0,
                                         0.92387953251128675613,
                                         0.70710678118654752440,
                                         0.38268343236508977173,
                                         0.0,
                                         -0.38268343236508977173,
                                         -0.70710678118654752440,
                                         -0.92387953251128675613,
                                         -1.0,
                                         -0.92387953251128675613,
                                         -0.70710678118654752440,
                                         -0.38268343236508977173,
                                         0.0,
                                         0.38268343236508977173,
                                         0.70710678118654752440,
                                         0.92387953251128675613};

// The per-lane operands of the inverse 16-point DFT matrix and the stage
// twiddles. Lane (i = lane % 16, g = lane / 16) holds row i of each tile at
// K = 8g .. 8g + 7: K < 16 multiplies Xr[K], K >= 16 multiplies Xi[K - 16].
template <class T> struct Dft16Operands {
  u16x8 a0h, a0l, a1h, a1l;  // tile 0 (rows Yr) and tile 1 (rows Yi), hi and lo
  float2 tw[4];              // exp(+2 pi i q (4g + e) / 64), q = (lane % 16) % 4

  __device__ void init(int lane, const float2* __restrict__ table) {
    const int i = lane & 15, g = lane >> 4;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int n1 = (8 * g + e) & 15;
      const int m = (i * n1) & 15;
      const double c = COS16[m], s = COS16[(m + 12) & 15];
      const double v0 = g < 2 ? c : -s;  // Yr = Fr Xr - Fi Xi
      const double v1 = g < 2 ? s : c;   // Yi = Fi Xr + Fr Xi
      a0h[e] = T::bits((float)v0);
      a0l[e] = T::bits((float)(v0 - (double)T::value(a0h[e])));
      a1h[e] = T::bits((float)v1);
      a1l[e] = T::bits((float)(v1 - (double)T::value(a1h[e])));
    }
    const int q = i & 3;
#pragma unroll
    for (int e = 0; e < 4; ++e) tw[e] = twiddle<1, SN>(table, q * (4 * g + e));
  }
};

__device__ __forceinline__ float quad_bcast(float v, int j) {
  // DPP quad_perm(j, j, j, j): every lane of a quad reads quad lane j
  const int ctrl = j * 0x55;
  int r;
  switch (ctrl) {
    case 0x00: r = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x00, 0xF, 0xF, false); break;
    case 0x55: r = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x55, 0xF, 0xF, false); break;
    case 0xAA: r = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xAA, 0xF, 0xF, false); break;
    default: r = __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xFF, 0xF, 0xF, false); break;
  }
  return __int_as_float(r);
}

// a * i^m
__device__ __forceinline__ float2 rot_i(float2 a, int m) {
  switch (m & 3) {
    case 0: return a;
    case 1: return make_float2(-a.y, a.x);
    case 2: return make_float2(-a.x, -a.y);
    default: return make_float2(a.y, -a.x);
  }
}

// Inverse 64-point DFTs of the 16 lines buf[L * LP + k] (one wave, L < 16, k <
// 64), unnormalised; out(L, k, X) receives X[k] of line L. Lines 4grp .. 4grp+3
// form one group: column c = 4 (L % 4) + q holds the subsequence x[4 n1 + q].
// The caller synchronises before (buf written) and after (out's target read).
template <class T, bool SPLIT, class Out>
__device__ __forceinline__ void dft64_lines(const float2* buf, int lane, const Dft16Operands<T>& A, Out out) {
  const int c = lane & 15, g = lane >> 4, q = c & 3;
  float x[4][8];
#pragma unroll
  for (int grp = 0; grp < 4; ++grp) {
    const int L = 4 * grp + (c >> 2);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float2 z = buf[L * LP + 4 * (8 * (g & 1) + e) + q];
      x[grp][e] = g < 2 ? z.x : z.y;
    }
  }
  __syncthreads();  // every input is in registers before any output lands in buf
#pragma unroll
  for (int grp = 0; grp < 4; ++grp) {
    const int L = 4 * grp + (c >> 2);
    float m = 0.f;
#pragma unroll
    for (int e = 0; e < 8; ++e) m = fmaxf(m, fabsf(x[grp][e]));
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 32));
    int ex = 0;
    if (m > 0.f) (void)frexpf(m, &ex);  // m = f 2^ex, f in [1/2, 1)
    u16x8 bh, bl;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float v = ldexpf(x[grp][e], -ex);
      bh[e] = T::bits(v);
      bl[e] = T::bits(v - T::value(bh[e]));
    }
    f32x4 re = {0.f, 0.f, 0.f, 0.f}, im = {0.f, 0.f, 0.f, 0.f};
    if constexpr (SPLIT) {
      re = T::mfma(A.a0l, bh, re);
      re = T::mfma(A.a0h, bl, re);
      im = T::mfma(A.a1l, bh, im);
      im = T::mfma(A.a1h, bl, im);
    }
    re = T::mfma(A.a0h, bh, re);
    im = T::mfma(A.a1h, bh, im);
    // lane (c, g) holds Y_q[k1] for k1 = 4g + e: twiddle, then the radix-4 stage over q
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float2 y = cmul(make_float2(ldexpf(re[e], ex), ldexpf(im[e], ex)), A.tw[e]);
      float2 acc = make_float2(0.f, 0.f);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float2 zj = make_float2(quad_bcast(y.x, j), quad_bcast(y.y, j));
        acc = cadd(acc, rot_i(zj, j * q));
      }
      out(L, 4 * g + e + 16 * q, acc);  // X[k1 + 16 k2] with k2 = q
    }
  }
}

// Inverse 64-point DFTs in place along one axis of each orientation's batch
// spectrum (SH bins at Q + o * SH): line l starts at (l / lpp) * pstride +
// l % lpp, its elements estride apart. One wave per 16 lines.
template <class T, bool SPLIT>
__global__ __launch_bounds__(64) void k_split_lines(float2* __restrict__ Q, int nlines, int lpp, int pstride,
                                                    int estride, size_t ostride, const float2* __restrict__ tw) {
  __shared__ float2 buf[16 * LP];
  const int lane = threadIdx.x;
  Dft16Operands<T> A;
  A.init(lane, tw);
  float2* Qo = Q + (size_t)blockIdx.y * ostride;
  const int line0 = blockIdx.x * 16;
  auto addr = [&](int L, int k) -> long {
    const int line = line0 + L;
    return (long)(line / lpp) * pstride + line % lpp + (long)k * estride;
  };
  for (int i = lane; i < 16 * SN; i += 64) {
    const int L = i & 15, k = i >> 4;
    buf[L * LP + k] = line0 + L < nlines ? Qo[addr(L, k)] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2* wb = buf;
  dft64_lines<T, SPLIT>(buf, lane, A, [&](int L, int k, float2 v) { wb[L * LP + k] = v; });
  __syncthreads();
  for (int i = lane; i < 16 * SN; i += 64) {
    const int L = i & 15, k = i >> 4;
    if (line0 + L < nlines) Qo[addr(L, k)] = buf[L * LP + k];
  }
}

// The x rows (z, y) of nb orientations: Hermitian completion, inverse DFT,
// real part times scale[o], running max/argmax of the 16 rows r0 .. r0 + 15
// held in registers (lane: row lane / 4, x in [16 (lane % 4), + 16)).
template <class T, bool SPLIT>
__global__ __launch_bounds__(64) void k_split_x_max(const float2* __restrict__ Q, int nb, int angle0,
                                                    const float* __restrict__ scale, float* __restrict__ best,
                                                    int32_t* __restrict__ bidx, const float2* __restrict__ tw) {
  __shared__ float2 buf[16 * LP];
  __shared__ float obuf[16 * LP];
  const int lane = threadIdx.x;
  Dft16Operands<T> A;
  A.init(lane, tw);
  const int r0 = blockIdx.x * 16;
  const int R = lane >> 2, x0 = (lane & 3) * 16;
  const size_t vox = (size_t)(r0 + R) * SN + x0;
  float bv[16];
  int bi[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    bv[k] = best[vox + k];
    bi[k] = bidx[vox + k];
  }
  float* wo = obuf;
  for (int o = 0; o < nb; ++o) {
    const float2* Qo = Q + (size_t)o * SH + (size_t)r0 * SNH;
    for (int i = lane; i < 16 * SNH; i += 64) {
      const int L = i / SNH, k = i % SNH;
      const float2 v = Qo[i];
      buf[L * LP + k] = v;
      if (k >= 1 && k < SN / 2) buf[L * LP + SN - k] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    dft64_lines<T, SPLIT>(buf, lane, A, [&](int L, int k, float2 v) { wo[L * LP + k] = v.x; });
    __syncthreads();
    const float sc = scale[o];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const float cc = obuf[R * LP + x0 + k] * sc;
      if (cc > bv[k]) {
        bv[k] = cc;
        bi[k] = angle0 + o;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    best[vox + k] = bv[k];
    bidx[vox + k] = bi[k];
  }
}

// Q = E * conj(R) (DC 0) for every bin of nb orientations, and per-workgroup
// partial sums of |Herm R|^2 over the stored half: on the x = 0 plane Herm R(k)
// = (R(k) + conj R(-k)) / 2, on the x = 32 plane (inside the sphere only at its
// own mirror (32, 0, 0)) Re R, elsewhere R.
__global__ __launch_bounds__(256) void k_split_project(Projector P, const float* __restrict__ mats,
                                                       const float2* __restrict__ E, const float* __restrict__ ctf,
                                                       float2* __restrict__ Q, float* __restrict__ partial) {
  __shared__ float red[4];
  const int o = blockIdx.y;
  float M[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) M[k] = mats[9 * o + k];
  const __amdgpu_buffer_rsrc_t rs = projector_rsrc(P);
  float2* Qo = Q + (size_t)o * SH;
  float acc = 0.f;
  for (int b = blockIdx.x * 256 + threadIdx.x; b < SH; b += PROJ_BLOCKS * 256) {
    const int x = b % SNH, yi = (b / SNH) % SN, zi = b / SPLANE;
    const int y = signed_freq(yi, SN), z = signed_freq(zi, SN);
    float2 q = make_float2(0.f, 0.f);
    if (x * x + y * y + z * z <= (SN / 2) * (SN / 2) && b != 0) {
      const float c = ctf ? ctf[b] : 1.f;
      const float2 R = cscale(project_sample(rs, P.p, P.ph, M, x, y, z), c);
      q = cmulc(E[b], R);
      float2 h = R;
      if (x == 0) {
        const int bm = (((SN - zi) % SN) * SN + (SN - yi) % SN) * SNH;
        const float cm = ctf ? ctf[bm] : 1.f;
        const float2 Rm = cscale(project_sample(rs, P.p, P.ph, M, 0, -y, -z), cm);
        h = make_float2(0.5f * (R.x + Rm.x), 0.5f * (R.y - Rm.y));
      } else if (x == SN / 2) {
        h = make_float2(R.x, 0.f);
      }
      acc += h.x * h.x + h.y * h.y;
    }
    Qo[b] = q;
  }
  const float tot = block_sum(acc, red);
  if (threadIdx.x == 0) partial[(size_t)o * PROJ_BLOCKS + blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_split_scale(const float* __restrict__ partial, float* __restrict__ scale) {
  __shared__ double red[256];
  const int o = blockIdx.x;
  double s = 0;
  for (int i = threadIdx.x; i < PROJ_BLOCKS; i += 256) s += (double)partial[(size_t)o * PROJ_BLOCKS + i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) scale[o] = (float)(1.0 / fmax(sqrt(red[0] / (double)SH), 1e-20));
}

template <class T, bool SPLIT>
int run_split_batch(gtomx_pick_plan* pl, const Projector& P, const float* mats, int nb, const float2* E,
                    const float* ctf, float* best, int32_t* idx, int angle0, const float2* tw, hipStream_t s) {
  hipLaunchKernelGGL(k_split_project, dim3(PROJ_BLOCKS, nb), dim3(256), 0, s, P, mats, E, ctf, pl->G, pl->partial);
  GX_LAUNCHED("k_split_project");
  hipLaunchKernelGGL(k_split_scale, dim3(nb), dim3(256), 0, s, pl->partial, pl->scale);
  GX_LAUNCHED("k_split_scale");
  // z lines: (x, y) -> x + 33 y, elements SPLANE apart; y lines: (x, z) -> x + 33 z
  hipLaunchKernelGGL((k_split_lines<T, SPLIT>), dim3(SPLANE / 16, nb), dim3(64), 0, s, pl->G, SPLANE, SPLANE, 0,
                     SPLANE, (size_t)SH, tw);
  GX_LAUNCHED("k_split_lines");
  hipLaunchKernelGGL((k_split_lines<T, SPLIT>), dim3(SPLANE / 16, nb), dim3(64), 0, s, pl->G, SPLANE, SNH, SPLANE,
                     SNH, (size_t)SH, tw);
  GX_LAUNCHED("k_split_lines");
  hipLaunchKernelGGL((k_split_x_max<T, SPLIT>), dim3(SN * SN / 16), dim3(64), 0, s, pl->G, nb, angle0, pl->scale,
                     best, idx, tw);
  GX_LAUNCHED("k_split_x_max");
  return GTOMX_OK;
}

int run_split_batch_p(int precision, gtomx_pick_plan* pl, const Projector& P, const float* mats, int nb,
                      const float2* E, const float* ctf, float* best, int32_t* idx, int angle0, const float2* tw,
                      hipStream_t s) {
  switch (precision) {
    case GTOMX_SPLIT_FP16: return run_split_batch<F16, true>(pl, P, mats, nb, E, ctf, best, idx, angle0, tw, s);
    case GTOMX_SPLIT_BF16: return run_split_batch<BF16, true>(pl, P, mats, nb, E, ctf, best, idx, angle0, tw, s);
    case GTOMX_SINGLE_FP16: return run_split_batch<F16, false>(pl, P, mats, nb, E, ctf, best, idx, angle0, tw, s);
    default: return run_split_batch<BF16, false>(pl, P, mats, nb, E, ctf, best, idx, angle0, tw, s);
  }
}

template <class T, bool SPLIT>
int split_lines(const float2* in, float2* out, unsigned nlines, const float2* tw, hipStream_t s) {
  const size_t bytes = (size_t)nlines * SN * sizeof(float2);
  if (in != out) GX_HIP(hipMemcpyAsync(out, in, bytes, hipMemcpyDeviceToDevice, s));
  // contiguous lines: line l at l * 64 (lpp = 1, pstride = 64), elements adjacent
  hipLaunchKernelGGL((k_split_lines<T, SPLIT>), dim3((nlines + 15) / 16, 1), dim3(64), 0, s, out, (int)nlines, 1, SN,
                     1, (size_t)0, tw);
  GX_LAUNCHED("k_split_lines");
  return GTOMX_OK;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int gtomx_pick_update_split(gtomx_pick_plan* plan, int precision, const gtomx_complex* d_projectorft,
                            float oversample, gtomx_int3 dimsprojector, const gtomx_complex* d_expft,
                            const float* d_ctf, size_t ctf_volume_stride, gtomx_int3 dimsvolume, unsigned nvolumes,
                            const gtomx_float3* h_angles, unsigned nangles, unsigned angle_offset, float* d_best,
                            int32_t* d_idx, void* stream) {
  GX_TRY(check_update_args(plan, d_projectorft, oversample, dimsprojector, d_expft, dimsvolume, nvolumes, d_best,
                           d_idx));
  if (plan->n != SN || plan->nz != SN) return fail(GTOMX_ERR_UNSUPPORTED, "the split-precision search runs on 64^3");
  if (precision < GTOMX_SPLIT_FP16 || precision > GTOMX_SINGLE_BF16)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "unknown precision");
  if (nangles == 0) return GTOMX_OK;
  if (!h_angles) return fail(GTOMX_ERR_INVALID_ARGUMENT, "h_angles is NULL");
  if ((unsigned long long)angle_offset + nangles > 0x7fffffffULL)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "orientation index exceeds int32");
  if ((size_t)plan->ntiles * 4 < (size_t)PROJ_BLOCKS)
    return fail(GTOMX_ERR_INTERNAL, "plan's partial-sum buffer is too small");
  hipStream_t s = as_stream(stream);
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  GX_TRY(plan->fence.enter(s));
  GX_TRY(upload_matrices(plan, h_angles, nangles, oversample, s));
  const Projector P{reinterpret_cast<const float2*>(d_projectorft), plan->p, plan->p / 2 + 1,
                    (plan->p / 2 + 1) * plan->p * plan->p * 8};
  for (unsigned v = 0; v < nvolumes; ++v) {
    const float2* Ev = reinterpret_cast<const float2*>(d_expft) + (size_t)v * plan->H;
    const float* cv = d_ctf ? d_ctf + (size_t)v * ctf_volume_stride : nullptr;
    float* bv = d_best + (size_t)v * plan->V;
    int32_t* iv = d_idx + (size_t)v * plan->V;
    for (unsigned b0 = 0; b0 < nangles; b0 += plan->B) {
      const int nb = (int)std::min<unsigned>(plan->B, nangles - b0);
      GX_TRY(run_split_batch_p(precision, plan, P, plan->mats + 9 * (size_t)b0, nb, Ev, cv, bv, iv,
                               (int)(angle_offset + b0), tw, s));
    }
  }
  return plan->fence.leave(s);
}

int gtomx_split_dft64(const gtomx_complex* d_in, gtomx_complex* d_out, unsigned nlines, int precision,
                      void* stream) {
  if (!d_in || !d_out) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (nlines == 0) return GTOMX_OK;
  if (nlines > (1u << 24)) return fail(GTOMX_ERR_UNSUPPORTED, "more than 2^24 lines");
  hipStream_t s = as_stream(stream);
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  const float2* in = reinterpret_cast<const float2*>(d_in);
  float2* out = reinterpret_cast<float2*>(d_out);
  switch (precision) {
    case GTOMX_SPLIT_FP16: return split_lines<F16, true>(in, out, nlines, tw, s);
    case GTOMX_SPLIT_BF16: return split_lines<BF16, true>(in, out, nlines, tw, s);
    case GTOMX_SINGLE_FP16: return split_lines<F16, false>(in, out, nlines, tw, s);
    case GTOMX_SINGLE_BF16: return split_lines<BF16, false>(in, out, nlines, tw, s);
    default: return fail(GTOMX_ERR_INVALID_ARGUMENT, "unknown precision");
  }
}

}  // extern "C"
