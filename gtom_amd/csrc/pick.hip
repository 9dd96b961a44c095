// The exhaustive template-matching search, d_PickSubTomograms
// (src/Correlation/SubTomograms.cu:19-125), as three fused passes per batch
// of B orientations over one n^3 volume:
//
//  pass 1 (k_pick_project_z): for a tile of TX x-columns at one y and all z:
//     R = project(projector, M_o) * ctf        (src/Relion/Project.cu:89-177,
//                                               src/Generics/Arithmetics.cu:275-286)
//     Q = E_v * conj(R) with DC = 0            (src/Correlation/SubTomograms.cu:127-136)
//     and the partial sums of |R|^2; inverse DFT along z, written to the
//     batch buffer G[o].
//  k_pick_herm_norm: the correction of those sums on the x = 0 plane (the
//     x = n/2 plane meets the sphere only at its centre bin), so that together
//     they are sum |R'|^2 for R' = Herm(R) with DC
//     zeroed: the Fourier form of C2R -> mean0/std1 -> R2C -> /RMS
//     (src/Correlation/SubTomograms.cu:66-70, src/ImageManipulation/Norm.cu:242-284, 476-509).
//  pass 2 (k_pick_y): inverse DFT along y in place; one workgroup per
//     orientation also reduces the sums to the scale 1/RMS.
//  pass 3 (k_pick_x_max): C2R along x (real part of the Hermitian-completed
//     inverse), times the scale, then the running max/argmax with strict '>'
//     in increasing orientation index (src/Correlation/SubTomograms.cu:138-158)
//     for all B orientations, best/argmax held in registers.
// After the last orientation, pick_finish divides by the local standard
// deviation of C2R(E_v) (src/Correlation/SubTomograms.cu:103-118).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "dev_fft.hpp"
#include "dev_project.hpp"
#include "pick_plan.hpp"
#include "runtime.hpp"

namespace gx {

// Layout of the batch buffer G (one block of STRIDE complex values per
// orientation). When n/2 is a multiple of the column-tile width TX ("split"),
// the x < n/2 columns form rows of exactly PITCH = n/2 values, so every tile
// row is a whole number of 128-byte lines, and the x = n/2 (Nyquist) column is
// stored after them as a compact [z][y] array. Otherwise rows hold all n/2+1
// columns padded to whole tiles (PITCH), and the padding is never read.
template <int N> struct GLayout {
  static constexpr int TX = ColShape<N>::TX;
  static constexpr bool SPLIT = (N / 2) % TX == 0;
  static constexpr int PITCH = SPLIT ? N / 2 : (N / 2 + 1 + TX - 1) / TX * TX;
  static constexpr size_t MAIN = (size_t)PITCH * N * N;
  static constexpr size_t STRIDE = MAIN + (SPLIT ? (size_t)N * N : 0);
  static constexpr int ROWS = N * N;  // rows (z, y) of the x transform
};

// The 2D branch (n x n images): per orientation, n rows (y) of all n/2+1
// columns, padded to whole column tiles.
template <int N> struct GLayout2 {
  static constexpr int TX = ColShape<N>::TX;
  static constexpr bool SPLIT = false;
  static constexpr int FT = (N / 2 + 1 + TX - 1) / TX;  // column tiles
  static constexpr int PITCH = FT * TX;
  static constexpr size_t MAIN = (size_t)PITCH * N;
  static constexpr size_t STRIDE = MAIN;
  static constexpr int ROWS = N;
};

// Workgroup layout of the column passes (1 and 2). In the split layout the
// first NQ workgroups take the Nyquist column of TX rows at once, one row per
// lane (their strided accesses are slower, so they start first), and the
// remaining FT*N workgroups take TX full columns of one row each. Otherwise
// every row has FT tiles and lanes past x = n/2 idle.
template <int N, int TXW = ColShape<N>::TX> struct ColMap {
  static constexpr int TX = TXW;
  static_assert(TXW == ColShape<N>::TX || (GLayout<N>::SPLIT && (N / 2) % TXW == 0),
                "a tile width other than the layout's needs the split layout");
  static constexpr bool SPLIT = GLayout<N>::SPLIT;
  static constexpr int FT = SPLIT ? (N / 2) / TX : (N / 2 + TX) / TX;
  static constexpr int NQ = SPLIT ? N / TX : 0;
  static constexpr int NWG = FT * N + NQ;
  // (workgroup w, lane column c) -> (x, row); row is y in pass 1 and z in pass 2.
  __device__ static void map(int w, int c, int& x, int& row) {
    if (w < NQ) {
      x = N / 2;
      row = w * TX + c;
    } else {
      x = ((w - NQ) % FT) * TX + c;
      row = (w - NQ) / FT;
    }
  }
};

// Cache-policy bits of buffer accesses: non-temporal (streaming) data should
// not displace the projector from L2.
constexpr int CPOL_NT = 2;
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// ---------------------------------------------------------------------------
// pass 1
//
// The Hermitian projection of the x = 0 and x = n/2 planes does not change
// C2R(E * conj R) when E is the R2C of a real volume (C2R keeps the real part
// of the completed inverse, and Herm(E conj R) = E conj Herm(R)), so the
// samples here are not projected; only the unit-RMS sum needs Herm(R), and
// k_pick_herm_norm supplies its correction on the two planes. The DC bin is
// zeroed (the reference subtracts the mean).
// ---------------------------------------------------------------------------
__device__ __forceinline__ __amdgpu_buffer_rsrc_t sampler_rsrc(const Projector& P) { return projector_rsrc(P); }
__device__ __forceinline__ __amdgpu_buffer_rsrc_t sampler_rsrc(const CentredProjector& C) {
  return make_rsrc(C.data, (unsigned)C.bytes);
}
__device__ __forceinline__ float2 sample(__amdgpu_buffer_rsrc_t rs, const Projector& P, float px, float py, float pz,
                                         bool in) {
  return project_at(rs, P.p, P.ph, px, py, pz, in);
}
__device__ __forceinline__ float2 sample(__amdgpu_buffer_rsrc_t rs, const CentredProjector& C, float px, float py,
                                         float pz, bool in) {
  return project_centred(rs, C, px, py, pz, in);
}

// Tile masks (the learned search): tmask[tz * nt + ty] holds one bit per
// 32-voxel tile tx of the output row band (z in tile tz, y in tile ty), nt =
// N / 32; nullptr searches everything. Pass 2 transforms only the planes z
// whose tile layer has a tile set and writes only the rows y of set (tz, ty)
// bands; pass 3 takes only those rows and updates only the voxels of set
// tiles. The unit-RMS scale is always computed.
template <int N> __device__ __forceinline__ bool layer_any(const uint8_t* tmask, int z) {
  constexpr int NT = N / 32;
  const uint8_t* m = tmask + (z / 32) * NT;
  uint8_t a = 0;
#pragma unroll
  for (int k = 0; k < NT; ++k) a |= m[k];
  return a != 0;
}

// Columns (x, y) with x^2 + y^2 > (n/2)^2 hold no bin of the sphere, so Q is
// zero along all of z there and stays zero through the z transform: pass 1
// writes nothing for them and pass 2 reads zeros in their place.
template <int N> __device__ __forceinline__ bool column_in_disc(int x, int y) {
  return x * x + y * y <= (N / 2) * (N / 2);
}

// Pass 1 runs as one wave of resident workgroups over the work items
// (tile, orientation), tile = a column tile with at least one column inside
// the disc (tiles[] lists their workgroup indices w of ColMap), in tile-major
// order: workgroup g takes the contiguous items [g*I/G, (g+1)*I/G) of the
// I = ntiles*nb items, so every workgroup does the same share of the work
// (one launch of one workgroup per tile leaves the last round mostly idle).
// E and the CTF are re-read when a workgroup moves to its next tile.
// Resident pass-1 workgroups per CU asked of the compiler: 3 at n = 256 with
// the centred sampler (at most 168 VGPRs, 12 waves per CU); no bound otherwise.
template <int N, class PJ> struct Pass1Occupancy {
  static constexpr int MIN_BLOCKS = (N == 256 && std::is_same<PJ, CentredProjector>::value) ? 3 : 1;
};

// Visiting order of a thread's samples m (z = j + R2 m in FFT order): 0, R1-1,
// 1, R1-2, ... is increasing |z|, so each group of samples is a shell of |z| and
// whole groups fall outside the sphere for the outer columns.
template <int R1> __device__ constexpr int zorder(int k) { return k % 2 == 0 ? k / 2 : R1 - 1 - k / 2; }

// Waves per pass-1 workgroup; each stores its own partial sum of |R|^2 per
// (orientation, tile), so that no barrier is spent on a block reduction.
template <int N> constexpr int pass1_waves() { return (ColShape<N>::TX * ColShape<N>::R2 + 63) / 64; }

template <int N, class PJ, bool MASKED = false>
__global__ __launch_bounds__(ColShape<N>::TX * ColShape<N>::R2, (Pass1Occupancy<N, PJ>::MIN_BLOCKS)) void
k_pick_project_z(
    PJ P, const float* __restrict__ mats, int nb, const float2* __restrict__ E,
    const float* __restrict__ ctf, float2* __restrict__ G, float* __restrict__ partial,
    const float2* __restrict__ tw, const int* __restrict__ tiles, int ntiles,
    const uint8_t* __restrict__ tmask) {
  using L = GLayout<N>;
  constexpr int R1 = ColShape<N>::R1, R2 = ColShape<N>::R2, TX = ColShape<N>::TX;
  constexpr int NH = N / 2 + 1, HALF = N / 2, RMAX2 = HALF * HALF;
  constexpr int NT = TX * R2;
  constexpr int GRP = 2;  // projector samples per group of in-flight gathers
  static_assert(R1 % GRP == 0, "R1 must be a multiple of the sample group");
  static_assert(R1 % 2 == 0, "zorder is a permutation of 0..R1-1 for even R1");
  __shared__ float2 xch[R2 > 1 ? TX * N : 1];
  // The N twiddles of the z transform, staged once per workgroup in LDS (the
  // first barrier before their use is the one ahead of the first transform).
  __shared__ float2 twl[N];
  const int t = threadIdx.x, c = t % TX, j = t / TX;
  for (int i = t; i < N; i += NT) twl[i] = twiddle<-1, N>(tw, i);
  // With tile masks, bit tz of `layers` says whether z layer tz (planes 32tz ..
  // 32tz + 31) holds a searched tile; pass 2 neither reads nor transforms the
  // other planes, so their outputs are not stored.
  unsigned layers = 0xffffffffu;
  if constexpr (MASKED) {
    layers = 0;
    for (int tz = 0; tz < N / 32; ++tz) layers |= layer_any<N>(tmask, 32 * tz) ? 1u << tz : 0u;
  }

  const __amdgpu_buffer_rsrc_t rs = sampler_rsrc(P);
  constexpr unsigned HB = (unsigned)(NH * N * N);
  const __amdgpu_buffer_rsrc_t rE = make_rsrc(E, HB * 8u);
  const __amdgpu_buffer_rsrc_t rC = make_rsrc(ctf, ctf ? HB * 4u : 0u);
  const long long items = (long long)ntiles * nb;
  const long long i1 = items * (blockIdx.x + 1) / gridDim.x;
  for (long long it = items * blockIdx.x / gridDim.x; it < i1;) {
    const int ti = (int)(it / nb);
    const int o0 = (int)(it % nb);
    const int o1 = (int)min((long long)nb, o0 + (i1 - it));
    it += o1 - o0;
    const int w = tiles[ti];
    int ix, iy;
    ColMap<N>::map(w, c, ix, iy);
    const bool nyq_wg = w < ColMap<N>::NQ;
    const int y = signed_freq(iy, N);
    const bool valid = ix < NH && column_in_disc<N>(ix, y);

    // E_v and the CTF at this thread's bins of the tile (element (x, y, z) at
    // (z*N + y)*NH + x; the z part is the constant m*R2*N*NH).
    const int ebase = valid ? (j * N + iy) * NH + ix : (int)HB;  // invalid lanes read nothing (0)
    float2 e[R1];
    float cf[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      auto ve = __builtin_amdgcn_raw_buffer_load_b64(rE, ebase * 8, m * R2 * N * NH * 8, CPOL_NT);
      __builtin_memcpy(&e[m], &ve, 8);
      cf[m] = ctf ? buf_load1(rC, ebase * 4, m * R2 * N * NH * 4) : (valid ? 1.f : 0.f);
    }
    const bool dc_thread = ix == 0 && iy == 0 && j == 0;  // holds the DC bin as m = 0
    // Output of this thread's first bin (z = j*(R1/R2)) and the step between its outputs.
    const int zrow0 = j * (R1 / R2);
    // Lanes of columns outside the disc store nothing (out-of-range offset).
    const int gbase8 = !valid ? OUT_OF_RANGE
                              : (nyq_wg ? (int)L::MAIN + zrow0 * N + iy : (zrow0 * N + iy) * L::PITCH + ix) * 8;
    const int gstep = nyq_wg ? N : N * L::PITCH;  // per unit of output z

    for (int o = o0; o < o1; ++o) {
      float M[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) M[k] = mats[9 * o + k];
      // q = M (x, y, z) = q0 + z * M[:, 2]
      const float q0x = M[0] * ix + M[1] * y, q0y = M[3] * ix + M[4] * y, q0z = M[6] * ix + M[7] * y;
      float acc = 0.f;
      float2 v[R1];
      // Samples in groups of GRP, in increasing |z|: the group's gathers are in
      // flight together, the scheduling barrier keeps the next group's loads from
      // being hoisted (which would exhaust the registers), and a group whose bins
      // all lie outside the sphere in this wave is skipped.
#pragma unroll
      for (int g = 0; g < R1 / GRP; ++g) {
        bool ins[GRP];
        bool any = false;
#pragma unroll
        for (int u = 0; u < GRP; ++u) {
          const int z = signed_freq(j + R2 * zorder<R1>(g * GRP + u), N);
          ins[u] = valid && ix * ix + y * y + z * z <= RMAX2;
          any |= ins[u];
        }
        if (__any(any)) {
#pragma unroll
          for (int u = 0; u < GRP; ++u) {
            const int m = zorder<R1>(g * GRP + u);
            const float z = (float)signed_freq(j + R2 * m, N);
            const float2 r = cscale(
                sample(rs, P, fmaf(M[2], z, q0x), fmaf(M[5], z, q0y), fmaf(M[8], z, q0z), ins[u]), cf[m]);
            acc += r.x * r.x + r.y * r.y;
            v[m] = cmulc(e[m], r);
          }
        } else {
#pragma unroll
          for (int u = 0; u < GRP; ++u) v[zorder<R1>(g * GRP + u)] = make_float2(0.f, 0.f);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      // DC of Q is zero (the reference's mean subtraction); its |R|^2 term is
      // removed from the sum by k_pick_herm_norm.
      if (dc_thread) v[0] = make_float2(0.f, 0.f);
      __syncthreads();
      fft_two_step<N, R1, R2, 1>(v, j, twl, [&](int k1, int jj) { return &xch[(k1 * R2 + jj) * TX + c]; },
                                 TwiddleLDS{});
      const __amdgpu_buffer_rsrc_t rG = make_rsrc(G + (size_t)o * L::STRIDE, (unsigned)(L::STRIDE * 8));
#pragma unroll
      for (int s = 0; s < R1; ++s) {
        const int u = s / R2, k2 = s % R2;  // output z = j*(R1/R2) + u + R1*k2
        if (MASKED && !((layers >> ((zrow0 + u + R1 * k2) >> 5)) & 1u)) continue;
        const u32x2 vs = __builtin_bit_cast(u32x2, v[s]);
        __builtin_amdgcn_raw_buffer_store_b64(vs, rG, gbase8, (u + R1 * k2) * gstep * 8, CPOL_NT);
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
      if ((t & 63) == 0) partial[((size_t)o * ntiles + ti) * pass1_waves<N>() + (t >> 6)] = acc;
    }
  }
}

// Workgroup indices w of the pass-1 tiles with at least one column inside
// the disc, in increasing order (the same test as column_in_disc).
template <int N> std::vector<int> disc_tiles() {
  using CM = ColMap<N>;
  std::vector<int> out;
  for (int w = 0; w < CM::NWG; ++w) {
    bool in;
    if (w < CM::NQ) {
      in = w == 0;  // the Nyquist column (n/2, y) meets the disc only at y = 0
    } else {
      const int x0 = ((w - CM::NQ) % CM::FT) * CM::TX, row = (w - CM::NQ) / CM::FT;
      const int y = row <= N / 2 ? row : row - N;
      in = x0 * x0 + y * y <= (N / 2) * (N / 2);
    }
    if (in) out.push_back(w);
  }
  return out;
}

// ---------------------------------------------------------------------------
// Unit-RMS correction on the x = 0 and x = n/2 planes: for every bin k of the
// two planes, |Herm R(k)|^2 - |R(k)|^2 with Herm R(k) = (R(k) + conj R(-k))/2
// (R(-k) sampled at the mirrored stored bin) and Herm R(0) = 0. Added to the
// pass-1 sums it gives sum |R'|^2 for R' = Herm(R) with DC zeroed, the
// spectrum that C2R -> mean0/std1 -> R2C produces in the reference.
// ---------------------------------------------------------------------------
constexpr int HERM_BLOCKS = 32;

// NT = 0: the side is the run-time n.
template <int NT>
__global__ __launch_bounds__(256) void k_pick_herm_norm(Projector P, const float* __restrict__ mats,
                                                        const float* __restrict__ ctf, float* __restrict__ corr,
                                                        int n) {
  const int N = NT ? NT : n;
  const int NH = N / 2 + 1, HALF = N / 2, RMAX2 = HALF * HALF;
  __shared__ float red[4];
  const int o = blockIdx.y;
  float M[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) M[k] = mats[9 * o + k];
  const __amdgpu_buffer_rsrc_t rs = projector_rsrc(P);
  float acc = 0.f;
  // the x = 0 plane: bins of the disc y^2 + z^2 <= (n/2)^2 (R is 0 elsewhere), paired with -k
  for (int b = blockIdx.x * 256 + threadIdx.x; b < N * N; b += gridDim.x * 256) {
    const int iy = b % N, iz = b / N;
    const int y = signed_freq(iy, N), z = signed_freq(iz, N);
    if (y * y + z * z > RMAX2) continue;
    const int iyp = (N - iy) % N, izp = (N - iz) % N;
    const int yp = signed_freq(iyp, N), zp = signed_freq(izp, N);
    const float c = ctf ? ctf[((size_t)iz * N + iy) * NH] : 1.f;
    const float cp = ctf ? ctf[((size_t)izp * N + iyp) * NH] : 1.f;
    const float2 R = cscale(project_at(rs, P.p, P.ph, M[1] * y + M[2] * z, M[4] * y + M[5] * z, M[7] * y + M[8] * z),
                            c);
    const float2 Rp = cscale(
        project_at(rs, P.p, P.ph, M[1] * yp + M[2] * zp, M[4] * yp + M[5] * zp, M[7] * yp + M[8] * zp), cp);
    float2 h = make_float2(0.5f * (R.x + Rp.x), 0.5f * (R.y - Rp.y));
    if (b == 0) h = make_float2(0.f, 0.f);  // DC
    acc += (h.x * h.x + h.y * h.y) - (R.x * R.x + R.y * R.y);
  }
  // the x = n/2 plane meets the sphere only at (n/2, 0, 0), its own mirror: Herm R = Re R (odd n have
  // no such plane)
  if ((N & 1) == 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    const float c = ctf ? ctf[HALF] : 1.f;
    const float2 R = cscale(project_at(rs, P.p, P.ph, M[0] * HALF, M[3] * HALF, M[6] * HALF), c);
    acc -= R.y * R.y;
  }
  const float tot = block_sum(acc, red);
  if (threadIdx.x == 0) corr[(size_t)o * gridDim.x + blockIdx.x] = tot;
}

// ---------------------------------------------------------------------------
// pass 2
// ---------------------------------------------------------------------------
// Column tile width of pass 2 (its only tie to pass 1 is the batch buffer's
// layout): 32 columns (256-byte row segments) where the split layout allows,
// which streams measurably better than 16.
template <int N> struct Pass2Tile {
  static constexpr int TX = (GLayout<N>::SPLIT && (N / 2) % 32 == 0) ? 32 : ColShape<N>::TX;
};

template <int N>
__global__ __launch_bounds__(Pass2Tile<N>::TX * ColShape<N>::R2) void k_pick_y(
    float2* __restrict__ G, const float* __restrict__ partial, int ntiles, const float* __restrict__ corr,
    float* __restrict__ scale, const float2* __restrict__ tw, const uint8_t* __restrict__ tmask) {
  using L = GLayout<N>;
  constexpr int R1 = ColShape<N>::R1, R2 = ColShape<N>::R2, TX = Pass2Tile<N>::TX;
  using CM = ColMap<N, TX>;
  constexpr int NH = N / 2 + 1;
  constexpr size_t H = (size_t)NH * N * N;
  constexpr int NT = TX * R2;
  // The unit-RMS reduction reuses the exchange area after the transform.
  static_assert(R2 == 1 || TX * N * sizeof(float2) >= NT * sizeof(double), "exchange area too small");
  __shared__ float2 xch[R2 > 1 ? TX * N : NT];
  double* dred = reinterpret_cast<double*>(xch);

  const int t = threadIdx.x, c = t % TX, j = t / TX;
  int ix, z;
  CM::map(blockIdx.x, c, ix, z);
  const int o = blockIdx.y;
  const bool nyq_wg = (int)blockIdx.x < CM::NQ;
  // Column (x, z) along y: main rows (stride PITCH) or the compact Nyquist column (stride 1).
  const int base = nyq_wg ? (int)L::MAIN + z * N : z * N * L::PITCH + ix;
  const int ystep = nyq_wg ? 1 : L::PITCH;
  const __amdgpu_buffer_rsrc_t rG = make_rsrc(G + (size_t)o * L::STRIDE, (unsigned)(L::STRIDE * 8));
  // a whole workgroup's plane is skipped when its tile layer is empty (the Nyquist
  // workgroups hold several planes and always run)
  const bool skip = tmask && !nyq_wg && !layer_any<N>(tmask, z);
  if (!skip) {
  float2 v[R1];
#pragma unroll
  for (int m = 0; m < R1; ++m) {
    // bins of columns outside the disc were not written by pass 1 and are zero
    const int voff = column_in_disc<N>(ix, signed_freq(j + R2 * m, N)) ? (base + j * ystep) * 8 : OUT_OF_RANGE;
    auto vl = __builtin_amdgcn_raw_buffer_load_b64(rG, voff, m * R2 * ystep * 8, CPOL_NT);
    __builtin_memcpy(&v[m], &vl, 8);
  }
  // The N twiddles of this length, staged once per workgroup in LDS (one
  // global load per thread instead of R1 - 1).
  __shared__ float2 twl[N];
  for (int i = t; i < N; i += NT) twl[i] = twiddle<-1, N>(tw, i);
  __syncthreads();
  fft_two_step<N, R1, R2, 1>(v, j, twl, [&](int k1, int jj) { return &xch[(k1 * R2 + jj) * TX + c]; },
                             TwiddleLDS{});
  const int obase = base + j * (R1 / R2) * ystep;
#pragma unroll
  for (int s = 0; s < R1; ++s) {
    const int u = s / R2, k2 = s % R2;  // output y = j*(R1/R2) + u + R1*k2
    const int y = j * (R1 / R2) + u + R1 * k2;
    if (tmask && !nyq_wg && tmask[(z / 32) * (N / 32) + y / 32] == 0) continue;
    const u32x2 vs = __builtin_bit_cast(u32x2, v[s]);
    __builtin_amdgcn_raw_buffer_store_b64(vs, rG, obase * 8, (u + R1 * k2) * ystep * 8, CPOL_NT);
  }
  }  // !skip
  if (blockIdx.x == 0) {
    // unit-RMS scale over the H half-spectrum bins (src/ImageManipulation/Norm.cu:476-509)
    __syncthreads();  // the exchange area is free again
    double sum = 0;
    const int nparts = ntiles * pass1_waves<N>();
    for (int i = t; i < nparts; i += NT) sum += (double)partial[(size_t)o * nparts + i];
    for (int i = t; i < HERM_BLOCKS; i += NT) sum += (double)corr[(size_t)o * HERM_BLOCKS + i];
    dred[t] = sum;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
      if (t < w) dred[t] += dred[t + w];
      __syncthreads();
    }
    if (t == 0) {
      double rms = sqrt(dred[0] / (double)H);
      scale[o] = (float)(1.0 / fmax(rms, 1e-20));
    }
  }
}

// ---------------------------------------------------------------------------
// pass 3
// ---------------------------------------------------------------------------
// L is the batch buffer's layout (GLayout for volumes, GLayout2 for images).
// With FRESH the running maximum starts at -1e30 / 0 instead of the values in
// best/bidx: orientation group blockIdx.y (og orientations from og*blockIdx.y)
// then writes its own maximum to best/bidx + blockIdx.y * ROWS * N, and
// k_pick_merge combines the groups in order.
template <int N, class L = GLayout<N>, bool FRESH = false>
__global__ __launch_bounds__(256) void k_pick_x_max(const float2* __restrict__ G, int nb, int angle0,
                                                    const float* __restrict__ scale, float* __restrict__ best,
                                                    int32_t* __restrict__ bidx, const float2* __restrict__ tw,
                                                    int og = 0, const uint8_t* __restrict__ tmask = nullptr) {
  constexpr int NC = N / 2;
  constexpr int R1 = RowShape<NC>::R1, R2 = RowShape<NC>::R2, RW = 256 / R2;
  constexpr int GS = R2 * (R1 + 2);  // two float2 of padding per thread slot (fewer bank conflicts than one)
  __shared__ float2 xch[R2 > 1 ? RW * GS : 1];
  const int t = threadIdx.x, r = t / R2, j = t % R2;
  const int row = blockIdx.x * RW + r;  // row = z*N + y (volumes) or y (images)
  // tile bits of this row (volumes with a mask): bit k covers x in [32k, 32k + 32)
  unsigned xbits = 0xffffffffu;
  if (!FRESH && tmask) {
    const int rz = min(row, L::ROWS - 1) / N, ry = min(row, L::ROWS - 1) % N;
    xbits = tmask[(rz / 32) * (N / 32) + ry / 32];
    // the workgroup leaves when none of its rows is searched
    if (__syncthreads_or(row < L::ROWS && xbits != 0) == 0) return;
  }
  const bool valid = row < L::ROWS && xbits != 0;
  if (FRESH) {
    const int o0 = og * (int)blockIdx.y;
    G += (size_t)o0 * L::STRIDE;
    scale += o0;
    angle0 += o0;
    nb = min(og, nb - o0);
    best += (size_t)blockIdx.y * L::ROWS * N;
    bidx += (size_t)blockIdx.y * L::ROWS * N;
  }

  float bv[2 * R1];
  int32_t bi[2 * R1];
  float2* bp = reinterpret_cast<float2*>(best + (size_t)row * N);
  int2* ip = reinterpret_cast<int2*>(bidx + (size_t)row * N);
#pragma unroll
  for (int s = 0; s < R1; ++s) {
    const int m = out_index<R1, R2>(j, s);
    float2 b = valid && !FRESH ? bp[m] : make_float2(-1e30f, -1e30f);
    int2 a = valid && !FRESH ? ip[m] : make_int2(0, 0);
    bv[2 * s] = b.x;
    bv[2 * s + 1] = b.y;
    bi[2 * s] = a.x;
    bi[2 * s + 1] = a.y;
  }
  // Row bins X[0..n/2-1] in the main rows, X[n/2] in the Nyquist column
  // (split layout) or at the end of the row.
  // This thread's twiddles, the same for every orientation, held in
  // registers: exp(2 pi i k/N) for its pre-processing bins k = j + R2*m and
  // the stage twiddles of the half-length transform.
  float2 twp[R1], twh[R1];
#pragma unroll
  for (int m = 0; m < R1; ++m) twp[m] = twiddle<1, N>(tw, j + R2 * m);
  load_two_step_twiddles<NC, R1, 1>(twh, j, tw);
  const int xrow = valid ? row * L::PITCH : (int)L::STRIDE;  // invalid rows read nothing (0)
  const int xnyq = valid ? (L::SPLIT ? (int)L::MAIN + row : row * L::PITCH + NC) : (int)L::STRIDE;

  for (int o = 0; o < nb; ++o) {
    const float sc = scale[o];
    const __amdgpu_buffer_rsrc_t rG = make_rsrc(G + (size_t)o * L::STRIDE, (unsigned)(L::STRIDE * 8));
    float2 v[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const int k = j + R2 * m;
      // a = X[k]; b = X[n/2 - k], which is X[n/2] for k = 0
      const float2 a = buf_load2(rG, (xrow + j) * 8, R2 * m * 8);
      const float2 b = buf_load2(rG, (m == 0 && j == 0 ? xnyq : xrow + NC - k) * 8, 0);
      // the imaginary parts of X[0] and X[n/2] do not enter the real inverse
      const float2 a0 = k == 0 ? make_float2(a.x, 0.f) : a;
      const float2 bc = k == 0 ? make_float2(b.x, 0.f) : cconj(b);
      v[m] = cadd(cadd(a0, bc), mul_pj<1>(cmul(twp[m], csub(a0, bc))));
    }
    __syncthreads();
    fft_two_step<NC, R1, R2, 1>(v, j, tw, [&](int k1, int jj) { return &xch[r * GS + jj * (R1 + 2) + k1]; },
                                TwiddleRegs<R1>{twh});
    const int32_t oi = angle0 + o;
#pragma unroll
    for (int s = 0; s < R1; ++s) {
      const float c0 = v[s].x * sc, c1 = v[s].y * sc;
      if (tmask && !FRESH && !((xbits >> (out_index<R1, R2>(j, s) >> 4)) & 1u)) continue;  // x = 2m, 2m + 1
      if (c0 > bv[2 * s]) {
        bv[2 * s] = c0;
        bi[2 * s] = oi;
      }
      if (c1 > bv[2 * s + 1]) {
        bv[2 * s + 1] = c1;
        bi[2 * s + 1] = oi;
      }
    }
  }
  if (valid) {
#pragma unroll
    for (int s = 0; s < R1; ++s) {
      const int m = out_index<R1, R2>(j, s);
      bp[m] = make_float2(bv[2 * s], bv[2 * s + 1]);
      ip[m] = make_int2(bi[2 * s], bi[2 * s + 1]);
    }
  }
}

// ---------------------------------------------------------------------------
// The centred projector layout (CentredProjector): element (cx, cy, cz) of a
// PX x PY x PY array, PX = p/2 + 2, PY = p + 3, is the reference's projector
// at linear index (iz*p + iy)*(p/2+1) + cx with iy = (cy - OFF) mod p,
// iz = (cz - OFF) mod p, OFF = p/2 + 1; cx = p/2 + 1 continues into the next
// row at that linear address, and indices past the array read 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_centre_projector(const float2* __restrict__ P, int p,
                                                          float2* __restrict__ out) {
  const int ph = p / 2 + 1, PX = ph + 1, PY = p + 3, OFF = p / 2 + 1;
  const int total = PX * PY * PY;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const int cx = i % PX, cy = (i / PX) % PY, cz = i / (PX * PY);
  const int iy = ((cy - OFF) % p + p) % p, iz = ((cz - OFF) % p + p) % p;
  const int lin = (iz * p + iy) * ph + cx;
  out[i] = lin < ph * p * p ? P[lin] : make_float2(0.f, 0.f);
}

inline int centred_bytes(int p) { return (p / 2 + 2) * (p + 3) * (p + 3) * 8; }

// Combines the per-group maxima of the 2D pass 3 in increasing group order with
// strict '>', so the first maximum over all orientations wins.
__global__ __launch_bounds__(256) void k_pick_merge(const float* __restrict__ gbest, const int32_t* __restrict__ gidx,
                                                   int groups, size_t v, float* __restrict__ best,
                                                   int32_t* __restrict__ bidx) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < v; i += (size_t)gridDim.x * 256) {
    float b = best[i];
    int32_t k = bidx[i];
    for (int g = 0; g < groups; ++g) {
      const float c = gbest[(size_t)g * v + i];
      if (c > b) {
        b = c;
        k = gidx[(size_t)g * v + i];
      }
    }
    best[i] = b;
    bidx[i] = k;
  }
}

// ---------------------------------------------------------------------------
// The 2D branch, pass 1: one workgroup per (column tile, orientation) projects
// the central section M (x, y, 0) (Project3DArraytoNDKernel<2>) at its TX
// columns and all y, applies the CTF, forms E * conj(R) with the DC zeroed and
// inverse-transforms along y; per-workgroup sums of |R|^2 go to partial.
// ---------------------------------------------------------------------------
template <int N, class PJ>
__global__ __launch_bounds__(ColShape<N>::TX * ColShape<N>::R2) void k_pick2d_project_y(
    PJ P, const float* __restrict__ mats, const float2* __restrict__ E, const float* __restrict__ ctf,
    float2* __restrict__ G, float* __restrict__ partial, const float2* __restrict__ tw) {
  using L = GLayout2<N>;
  constexpr int R1 = ColShape<N>::R1, R2 = ColShape<N>::R2, TX = ColShape<N>::TX;
  constexpr int NH = N / 2 + 1, HALF = N / 2, RMAX2 = HALF * HALF;
  constexpr int NT = TX * R2;
  __shared__ float2 xch[R2 > 1 ? TX * N : 1];
  __shared__ float red[(NT + 63) / 64];
  __shared__ float2 twl[N];
  const int t = threadIdx.x, c = t % TX, j = t / TX;
  for (int i = t; i < N; i += NT) twl[i] = twiddle<-1, N>(tw, i);
  const int tile = blockIdx.x, o = blockIdx.y;
  const int ix = tile * TX + c;
  const bool valid = ix < NH;
  float M[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) M[k] = mats[9 * o + k];
  const __amdgpu_buffer_rsrc_t rs = sampler_rsrc(P);
  float acc = 0.f;
  float2 v[R1];
#pragma unroll
  for (int m = 0; m < R1; ++m) {
    const int iy = j + R2 * m, y = signed_freq(iy, N);
    const bool ins = valid && ix * ix + y * y <= RMAX2;
    const float2 e = valid ? E[iy * NH + ix] : make_float2(0.f, 0.f);
    const float cf = valid ? (ctf ? ctf[iy * NH + ix] : 1.f) : 0.f;
    const float2 r = cscale(sample(rs, P, M[0] * ix + M[1] * y, M[3] * ix + M[4] * y, M[6] * ix + M[7] * y, ins),
                            cf);
    acc += r.x * r.x + r.y * r.y;
    v[m] = cmulc(e, r);
  }
  if (ix == 0 && j == 0) v[0] = make_float2(0.f, 0.f);  // DC
  __syncthreads();
  fft_two_step<N, R1, R2, 1>(v, j, twl, [&](int k1, int jj) { return &xch[(k1 * R2 + jj) * TX + c]; },
                             TwiddleLDS{});
  float2* g = G + (size_t)o * L::STRIDE;
#pragma unroll
  for (int s = 0; s < R1; ++s) g[(size_t)out_index<R1, R2>(j, s) * L::PITCH + ix] = v[s];
  const float tot = block_sum(acc, red);
  if (t == 0) partial[(size_t)o * L::FT + tile] = tot;
}

// The 2D unit-RMS correction: |Herm R|^2 - |R|^2 on the lines x = 0 and
// x = n/2, Herm R(0, y) = (R(0, y) + conj R(0, -y))/2, DC removed.
template <int N>
__global__ __launch_bounds__(256) void k_pick2d_herm_norm(Projector P, const float* __restrict__ mats,
                                                          const float* __restrict__ ctf, float* __restrict__ corr) {
  constexpr int NH = N / 2 + 1, HALF = N / 2, RMAX2 = HALF * HALF;
  __shared__ float red[4];
  const int o = blockIdx.x;
  float M[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) M[k] = mats[9 * o + k];
  const __amdgpu_buffer_rsrc_t rs = projector_rsrc(P);
  float acc = 0.f;
  for (int b = threadIdx.x; b < 2 * N; b += 256) {
    const int ix = b < N ? 0 : HALF;
    const int iy = b % N, iyp = (N - iy) % N;
    const int y = signed_freq(iy, N), yp = signed_freq(iyp, N);
    const bool in = ix * ix + y * y <= RMAX2;
    const float c = ctf ? ctf[iy * NH + ix] : 1.f;
    const float cp = ctf ? ctf[iyp * NH + ix] : 1.f;
    const float2 R = cscale(project_at(rs, P.p, P.ph, M[0] * ix + M[1] * y, M[3] * ix + M[4] * y,
                                       M[6] * ix + M[7] * y, in), c);
    const float2 Rp = cscale(project_at(rs, P.p, P.ph, M[0] * ix + M[1] * yp, M[3] * ix + M[4] * yp,
                                        M[6] * ix + M[7] * yp, in), cp);
    float2 h = make_float2(0.5f * (R.x + Rp.x), 0.5f * (R.y - Rp.y));
    if (b == 0) h = make_float2(0.f, 0.f);
    acc += (h.x * h.x + h.y * h.y) - (R.x * R.x + R.y * R.y);
  }
  const float tot = block_sum(acc, red);
  if (threadIdx.x == 0) corr[o] = tot;
}

// scale[o] = 1/max(sqrt((sum of partial sums + corrections)/H), 1e-20) in double
// (src/ImageManipulation/Norm.cu:476-509).
__global__ __launch_bounds__(256) void k_pick_scale(const float* __restrict__ partial, int nparts,
                                                   const float* __restrict__ corr, int ncorr, double h,
                                                   float* __restrict__ scale) {
  __shared__ double red[256];
  const int o = blockIdx.x, t = threadIdx.x;
  double sum = 0;
  for (int i = t; i < nparts; i += 256) sum += (double)partial[(size_t)o * nparts + i];
  for (int i = t; i < ncorr; i += 256) sum += (double)corr[(size_t)o * ncorr + i];
  red[t] = sum;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) scale[o] = (float)(1.0 / fmax(sqrt(red[0] / h), 1e-20));
}

// ---------------------------------------------------------------------------
// The 2D branch for images above 256 (micrographs up to 4096^2), on the
// generic transforms: pass 1 forms E * conj(CTF * R) for every bin of the
// half spectrum (DC zeroed) with per-workgroup sums of |R|^2, the unit-RMS
// scale comes from k_pick2d_herm_norm and k_pick_scale as in the fused form,
// the batch goes through the generic column inverse, and the row inverse of
// every orientation is folded into best / index in orientation order
// (strict >, first index on ties) by one row kernel (pick_rows_max_long).
// ---------------------------------------------------------------------------
// Pass 1 above 256, with n = L1 * L2: the projection fused into the first
// pass of the inverse y transform. Workgroup (column tile, (o, n2)) projects
// its TX columns at the rows iy = n2 + L2 n1 (E * conj(CTF R), DC zeroed),
// transforms over n1, multiplies by W_n^{n2 k1} and stores at row n2 + L2 k1
// of the scratch spectra (image_cols_pass2_inverse finishes the transform);
// one sum of |R|^2 per workgroup.
template <class PJ, int L1, int TX>
__global__ __launch_bounds__(TX * LineShape<L1>::R2) void k_pick2d_project_cols(
    PJ P, const float* __restrict__ mats, const float2* __restrict__ E, const float* __restrict__ ctf,
    float2* __restrict__ S, float* __restrict__ partial, int n, int L2, const float2* __restrict__ tw) {
  constexpr int R1 = LineShape<L1>::R1, R2 = LineShape<L1>::R2, NT = TX * R2;
  __shared__ float2 xch[R2 > 1 ? TX * L1 : 1];
  __shared__ float red[(NT + 63) / 64];
  const int NH = n / 2 + 1, HALF = n / 2;
  const int t = threadIdx.x, c = t % TX, j = t / TX;
  const int o = blockIdx.y / L2, n2 = blockIdx.y % L2;
  const int ix = blockIdx.x * TX + c;
  const bool valid = ix < NH;
  float M[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) M[k] = mats[9 * o + k];
  const __amdgpu_buffer_rsrc_t rs = sampler_rsrc(P);
  float acc = 0.f;
  float2 v[R1];
#pragma unroll
  for (int m = 0; m < R1; ++m) {
    const int iy = n2 + L2 * (j + R2 * m), y = signed_freq(iy, n);
    const bool ins = valid && ix * ix + y * y <= HALF * HALF;
    const size_t b = (size_t)iy * NH + ix;
    const float2 e = valid ? E[b] : make_float2(0.f, 0.f);
    const float cf = valid ? (ctf ? ctf[b] : 1.f) : 0.f;
    const float2 r =
        cscale(sample(rs, P, M[0] * ix + M[1] * y, M[3] * ix + M[4] * y, M[6] * ix + M[7] * y, ins), cf);
    acc += r.x * r.x + r.y * r.y;
    v[m] = ix == 0 && iy == 0 ? make_float2(0.f, 0.f) : cmulc(e, r);
  }
  fft_two_step<L1, R1, R2, 1>(v, j, tw, [&](int k1, int jj) { return &xch[(k1 * R2 + jj) * TX + c]; });
  if (valid) {
    float2* So = S + (size_t)o * NH * n;
#pragma unroll
    for (int s = 0; s < R1; ++s) {
      const int k1 = out_index<R1, R2>(j, s);
      // streaming stores: the spectra are read once, by the second pass
      const float2 w = cmul(v[s], twiddle_rt<1>(tw, n, n2 * k1));
      float* dst = reinterpret_cast<float*>(So + (size_t)(n2 + L2 * k1) * NH + ix);
      __builtin_nontemporal_store(w.x, dst);
      __builtin_nontemporal_store(w.y, dst + 1);
    }
  }
  const float tot = block_sum(acc, red);
  if (t == 0) partial[((size_t)o * gridDim.x + blockIdx.x) * L2 + n2] = tot;
}

// ---------------------------------------------------------------------------
// Volumes above 256 per side (n^3, n in {512, 768, 1024}), on the generic
// transforms: pass 1 writes Q_o = E * conj(CTF R_o) over the whole half
// spectrum of every orientation of the batch (zero outside the sphere and at
// DC) with one sum of |R|^2 per workgroup and orientation; the x = 0 / n/2
// plane correction and the scale as in the fused form; then the generic
// column inverse along z and y in place and the long-row pass (x C2R, scale,
// in-order max/argmax, pick_rows_max_long).
// ---------------------------------------------------------------------------
constexpr int P3_BINS = 4;  // consecutive-in-x bins per thread (stride 256)

template <class PJ>
__global__ __launch_bounds__(256) void k_pick3d_project_any(PJ P, const float* __restrict__ mats, int nb,
                                                            const float2* __restrict__ E,
                                                            const float* __restrict__ ctf, float2* __restrict__ G,
                                                            float* __restrict__ partial, int n, long long H) {
  __shared__ float red[4];
  const int NH = n / 2 + 1, HALF = n / 2;
  const int t = threadIdx.x;
  const long long base = (long long)blockIdx.x * 256 * P3_BINS + t;
  const __amdgpu_buffer_rsrc_t rs = sampler_rsrc(P);
  float2 e[P3_BINS];
  float cf[P3_BINS];
  bool live[P3_BINS], ins[P3_BINS];
  float fx[P3_BINS], fy[P3_BINS], fz[P3_BINS];
#pragma unroll
  for (int k = 0; k < P3_BINS; ++k) {
    const long long i = base + 256LL * k;
    live[k] = i < H;
    const long long ii = live[k] ? i : 0;
    const int ix = (int)(ii % NH), iy = (int)((ii / NH) % n), iz = (int)(ii / ((long long)NH * n));
    const int y = signed_freq(iy, n), z = signed_freq(iz, n);
    ins[k] = live[k] && (long long)ix * ix + (long long)y * y + (long long)z * z <= (long long)HALF * HALF;
    fx[k] = (float)ix;
    fy[k] = (float)y;
    fz[k] = (float)z;
    e[k] = ins[k] && ii != 0 ? E[ii] : make_float2(0.f, 0.f);  // DC of Q zeroed
    cf[k] = ins[k] ? (ctf ? ctf[ii] : 1.f) : 0.f;
  }
  for (int o = 0; o < nb; ++o) {
    float M[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) M[k] = mats[9 * o + k];
    float acc = 0.f;
    float2* Go = G + (size_t)o * H;
#pragma unroll
    for (int k = 0; k < P3_BINS; ++k) {
      const float2 r = cscale(sample(rs, P, fmaf(M[2], fz[k], fmaf(M[1], fy[k], M[0] * fx[k])),
                                     fmaf(M[5], fz[k], fmaf(M[4], fy[k], M[3] * fx[k])),
                                     fmaf(M[8], fz[k], fmaf(M[7], fy[k], M[6] * fx[k])), ins[k]),
                              cf[k]);
      acc += r.x * r.x + r.y * r.y;
      if (live[k]) {  // streaming stores: each spectrum is read once, by the column inverse
        const float2 q = cmulc(e[k], r);
        float* dst = reinterpret_cast<float*>(Go + base + 256LL * k);
        __builtin_nontemporal_store(q.x, dst);
        __builtin_nontemporal_store(q.y, dst + 1);
      }
    }
    const float tot = block_sum(acc, red);
    if (t == 0) partial[(size_t)o * gridDim.x + blockIdx.x] = tot;
  }
}

}  // namespace gx

// The plan (struct gtomx_pick_plan) is declared in pick_plan.hpp.

namespace gx {
namespace {

void plan_release(gtomx_pick_plan* pl) {
  if (!pl) return;
  pl->fence.release();
  dev_free(pl->G);
  dev_free(pl->partial);
  dev_free(pl->tiles);
  dev_free(pl->corr);
  dev_free(pl->scale);
  dev_free(pl->mats);
  dev_free(pl->ls_mem);
  dev_free(pl->ebuf);
  dev_free(pl->cproj);
  dev_free(pl->gbest);
  dev_free(pl->gidx);
  dev_free(pl->scr);
  dev_free(pl->tmasks);
  dev_free(pl->ldirs);
  dev_free(pl->ltop1);
  dev_free(pl->lcls);
  dev_free(pl->lmass);
  dev_free(pl->lpnp);
  if (pl->h_tmasks) (void)hipHostFree(pl->h_tmasks);
  if (pl->h_mats) (void)hipHostFree(pl->h_mats);
  if (pl->mats_ev) (void)hipEventDestroy(pl->mats_ev);
  for (hipEvent_t e : pl->ev_pool) (void)hipEventDestroy(e);
  delete pl;
}

size_t g_stride_for(int n) {
  switch (n) {
#define GX_CASE(N) \
  case N:          \
    return GLayout<N>::STRIDE;
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
  }
  return 0;
}

size_t g2_stride_for(int n) {
  switch (n) {
#define GX_CASE(N) \
  case N:          \
    return GLayout2<N>::STRIDE;
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
  }
  return 0;
}

int g2_tiles_for(int n) {
  switch (n) {
#define GX_CASE(N) \
  case N:          \
    return GLayout2<N>::FT;
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
  }
  return 0;
}

std::vector<int> tiles_for(int n) {
  switch (n) {
#define GX_CASE(N) \
  case N:          \
    return disc_tiles<N>();
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
  }
  return {};
}

// Workgroups of pass 1 that are resident at once on the whole device.
template <int N, class PJ> int resident_grid(int cus) {
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pick_project_z<N, PJ>,
                                                              ColShape<N>::TX * ColShape<N>::R2, 0);
  if (e != hipSuccess || per_cu < 1) per_cu = 1;
  return per_cu * cus;
}
int resident_grid_for(int n, bool centred, int cus) {
  switch (n) {
#define GX_CASE(N) \
  case N:          \
    return centred ? resident_grid<N, CentredProjector>(cus) : resident_grid<N, Projector>(cus);
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
  }
  return cus;
}

int event_at(gtomx_pick_plan* pl, size_t i, hipEvent_t* ev) {
  while (pl->ev_pool.size() <= i) {
    hipEvent_t e;
    GX_HIP(hipEventCreate(&e));
    pl->ev_pool.push_back(e);
  }
  *ev = pl->ev_pool[i];
  return GTOMX_OK;
}

// Runs `launch` bracketed by timing events when the plan's timing is enabled.
template <class F> int timed(gtomx_pick_plan* pl, int stage, hipStream_t s, F launch) {
  if (!pl->timing) return launch();
  hipEvent_t a, b;
  size_t ia = pl->ev_used, ib = pl->ev_used + 1;
  GX_TRY(event_at(pl, ia, &a));
  GX_TRY(event_at(pl, ib, &b));
  pl->ev_used += 2;
  GX_HIP(hipEventRecord(a, s));
  GX_TRY(launch());
  GX_HIP(hipEventRecord(b, s));
  pl->pending.push_back({stage, ia, ib});
  return GTOMX_OK;
}

template <int N, class PJ>
int run_batch(gtomx_pick_plan* pl, const PJ& PS, const Projector& P, const float* mats, int nb, const float2* E,
              const float* ctf, float* best, int32_t* idx, int angle0, const float2* tw, hipStream_t s,
              const uint8_t* tmask) {
  constexpr int TX = ColShape<N>::TX, R2 = ColShape<N>::R2;
  constexpr int RW = 256 / RowShape<N / 2>::R2;
  GX_TRY(timed(pl, 0, s, [&]() -> int {
    const int g = pl->grid1[std::is_same<PJ, CentredProjector>::value ? 1 : 0];
    if (tmask)
      hipLaunchKernelGGL((k_pick_project_z<N, PJ, true>), dim3(g), dim3(TX * R2), 0, s, PS, mats, nb, E, ctf, pl->G,
                         pl->partial, tw, pl->tiles, pl->ntiles, tmask);
    else
      hipLaunchKernelGGL((k_pick_project_z<N, PJ>), dim3(g), dim3(TX * R2), 0, s, PS, mats, nb, E, ctf, pl->G,
                         pl->partial, tw, pl->tiles, pl->ntiles, nullptr);
    GX_LAUNCHED("k_pick_project_z");
    return GTOMX_OK;
  }));
  // stage 1 of the timings: the plane correction and pass 2
  GX_TRY(timed(pl, 1, s, [&]() -> int {
    hipLaunchKernelGGL(k_pick_herm_norm<N>, dim3(HERM_BLOCKS, nb), dim3(256), 0, s, P, mats, ctf, pl->corr, N);
    GX_LAUNCHED("k_pick_herm_norm");
    hipLaunchKernelGGL(k_pick_y<N>, dim3(ColMap<N, Pass2Tile<N>::TX>::NWG, nb), dim3(Pass2Tile<N>::TX * R2), 0, s,
                       pl->G, pl->partial, pl->ntiles, pl->corr, pl->scale, tw, tmask);
    GX_LAUNCHED("k_pick_y");
    return GTOMX_OK;
  }));
  GX_TRY(timed(pl, 2, s, [&]() -> int {
    hipLaunchKernelGGL(k_pick_x_max<N>, dim3((N * N + RW - 1) / RW), dim3(256), 0, s, pl->G, nb, angle0, pl->scale,
                       best, idx, tw, 0, tmask);
    GX_LAUNCHED("k_pick_x_max");
    return GTOMX_OK;
  }));
  return GTOMX_OK;
}

// Volumes above 256: pass 1 over the whole half spectrum, the plane correction
// and the scale, the generic column inverse along z and y in place, then the
// long-row pass folds every orientation into best / index in order.
int herm_norm3_any(int n, const Projector& P, const float* mats, const float* ctf, float* corr, int nb,
                   hipStream_t s) {
  switch (n) {
#define GX_CASE(N)                                                                                    \
  case N:                                                                                             \
    hipLaunchKernelGGL(k_pick_herm_norm<N>, dim3(HERM_BLOCKS, nb), dim3(256), 0, s, P, mats, ctf, corr, N); \
    GX_LAUNCHED("k_pick_herm_norm");                                                                  \
    return GTOMX_OK;
    GX_CASE(512) GX_CASE(768) GX_CASE(1024)
#undef GX_CASE
  }
  hipLaunchKernelGGL(k_pick_herm_norm<0>, dim3(HERM_BLOCKS, nb), dim3(256), 0, s, P, mats, ctf, corr, n);
  GX_LAUNCHED("k_pick_herm_norm");
  return GTOMX_OK;
}

// The in-order update of the generic form: the real maps of nb orientations
// (`stride` floats apart), each times scale[o], folded into best / index
// (strict >, the first orientation winning ties), as k_pick_rows_max_long does.
__global__ __launch_bounds__(256) void k_pick_max_maps(const float* __restrict__ maps, long long stride, int nb,
                                                       const float* __restrict__ scale, long long V, int angle0,
                                                       float* __restrict__ best, int32_t* __restrict__ idx) {
  for (long long v = blockIdx.x * 256LL + threadIdx.x; v < V; v += (long long)gridDim.x * 256) {
    float b = best[v];
    int32_t i = idx[v];
    for (int o = 0; o < nb; ++o) {
      const float c = maps[o * stride + v] * scale[o];
      if (c > b) {
        b = c;
        i = angle0 + o;
      }
    }
    best[v] = b;
    idx[v] = i;
  }
}

template <class PJ>
int run_batch3d_any(gtomx_pick_plan* pl, const PJ& PS, const Projector& P, const float* mats, int nb,
                    const float2* E, const float* ctf, float* best, int32_t* idx, int angle0, hipStream_t s) {
  const int n = pl->n;
  const long long H = (long long)pl->H;
  GX_TRY(timed(pl, 0, s, [&]() -> int {
    hipLaunchKernelGGL(k_pick3d_project_any<PJ>, dim3((unsigned)pl->nparts), dim3(256), 0, s, PS, mats, nb, E, ctf,
                       pl->G, pl->partial, n, H);
    GX_LAUNCHED("k_pick3d_project_any");
    return GTOMX_OK;
  }));
  GX_TRY(herm_norm3_any(n, P, mats, ctf, pl->corr, nb, s));
  GX_TRY(timed(pl, 1, s, [&]() -> int {
    hipLaunchKernelGGL(k_pick_scale, dim3(nb), dim3(256), 0, s, pl->partial, pl->nparts, pl->corr, HERM_BLOCKS,
                       (double)pl->H, pl->scale);
    GX_LAUNCHED("k_pick_scale");
    if (pl->gen3d) return yz_axes_any(pl->G, gtomx_int3{n, n, n}, (unsigned)nb, 1, s);
    return fft_c2r_cols_inplace(pl->G, gtomx_int3{n, n, n}, (unsigned)nb, s);
  }));
  return timed(pl, 2, s, [&]() -> int {
    if (pl->gen3d) {
      // the rows' C2R into the first n^3 floats of each spectrum's slot, then the in-order update
      GX_TRY(rows_c2r_any_inplace(pl->G, n, nb, H, s));
      const long long V = (long long)n * n * n;
      hipLaunchKernelGGL(k_pick_max_maps, dim3((unsigned)std::min<long long>((V + 255) / 256, 16384)), dim3(256), 0,
                         s, reinterpret_cast<const float*>(pl->G), 2 * H, nb, pl->scale, V, angle0, best, idx);
      GX_LAUNCHED("k_pick_max_maps");
      return GTOMX_OK;
    }
    return pick_rows_max_long(n, pl->G, nb, H, n * n, angle0, pl->scale, best, idx, s, pl->rows_variant);
  });
}

// Pass 1 samples the centred layout when every sampled q stays within
// |q| < p/2 + 1/2 (oversample * n/2 <= p/2 + 1/4, which leaves room for the
// rounding of q), and the reference's layout otherwise.
int run_batch_n(gtomx_pick_plan* pl, const CentredProjector* C, const Projector& P, const float* mats, int nb,
                const float2* E, const float* ctf, float* best, int32_t* idx, int angle0, const float2* tw,
                hipStream_t s, const uint8_t* tmask = nullptr) {
  if (pl->n > 256 || pl->gen3d) {
    if (tmask) return fail(GTOMX_ERR_UNSUPPORTED, "tile masks need volumes of side 2^a or 3*2^a, at most 256");
    return C ? run_batch3d_any(pl, *C, P, mats, nb, E, ctf, best, idx, angle0, s)
             : run_batch3d_any(pl, P, P, mats, nb, E, ctf, best, idx, angle0, s);
  }
  switch (pl->n) {
#define GX_CASE(N)                                                                            \
  case N:                                                                                     \
    return C ? run_batch<N>(pl, *C, P, mats, nb, E, ctf, best, idx, angle0, tw, s, tmask)     \
             : run_batch<N>(pl, P, P, mats, nb, E, ctf, best, idx, angle0, tw, s, tmask);
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
  }
  return fail(GTOMX_ERR_UNSUPPORTED, "volume side not supported");
}

// The 2D branch: pass 1 (projection + y transform), the line correction and
// the scale, then pass 3 per orientation group and the in-order merge.
template <int N, class PJ>
int run_batch2d(gtomx_pick_plan* pl, const PJ& PS, const Projector& P, const float* mats, int nb, const float2* E,
                const float* ctf, float* best, int32_t* idx, int angle0, const float2* tw, hipStream_t s) {
  using L = GLayout2<N>;
  constexpr int TX = ColShape<N>::TX, R2 = ColShape<N>::R2;
  constexpr int RW = 256 / RowShape<N / 2>::R2;
  GX_TRY(timed(pl, 0, s, [&]() -> int {
    hipLaunchKernelGGL((k_pick2d_project_y<N, PJ>), dim3(L::FT, nb), dim3(TX * R2), 0, s, PS, mats, E, ctf, pl->G,
                       pl->partial, tw);
    GX_LAUNCHED("k_pick2d_project_y");
    return GTOMX_OK;
  }));
  hipLaunchKernelGGL(k_pick2d_herm_norm<N>, dim3(nb), dim3(256), 0, s, P, mats, ctf, pl->corr);
  GX_LAUNCHED("k_pick2d_herm_norm");
  GX_TRY(timed(pl, 1, s, [&]() -> int {
    hipLaunchKernelGGL(k_pick_scale, dim3(nb), dim3(256), 0, s, pl->partial, L::FT, pl->corr, 1, (double)pl->H,
                       pl->scale);
    GX_LAUNCHED("k_pick_scale");
    return GTOMX_OK;
  }));
  const int groups = (nb + pl->og - 1) / pl->og;
  GX_TRY(timed(pl, 2, s, [&]() -> int {
    hipLaunchKernelGGL((k_pick_x_max<N, L, true>), dim3((N + RW - 1) / RW, groups), dim3(256), 0, s, pl->G, nb,
                       angle0, pl->scale, pl->gbest, pl->gidx, tw, pl->og);
    GX_LAUNCHED("k_pick_x_max");
    const unsigned grid = (unsigned)std::min<size_t>((pl->V + 255) / 256, 4096);
    hipLaunchKernelGGL(k_pick_merge, dim3(grid), dim3(256), 0, s, pl->gbest, pl->gidx, groups, pl->V, best, idx);
    GX_LAUNCHED("k_pick_merge");
    return GTOMX_OK;
  }));
  return GTOMX_OK;
}

// The 2D branch above 256: pass 1 fused with the first half of the y
// inverse, the same scale, the second half, then the fused row inverse +
// in-order update.
int herm_norm_any(int n, const Projector& P, const float* mats, const float* ctf, float* corr, int nb,
                  hipStream_t s) {
  switch (n) {
#define GX_CASE(N)                                                                                            \
  case N:                                                                                                     \
    hipLaunchKernelGGL(k_pick2d_herm_norm<N>, dim3(nb), dim3(256), 0, s, P, mats, ctf, corr);                 \
    GX_LAUNCHED("k_pick2d_herm_norm");                                                                        \
    return GTOMX_OK;
    GX_CASE(384) GX_CASE(512) GX_CASE(768) GX_CASE(1024) GX_CASE(1536) GX_CASE(2048) GX_CASE(3072) GX_CASE(4096)
#undef GX_CASE
  }
  return fail(GTOMX_ERR_UNSUPPORTED, "image side not supported");
}

template <class PJ>
int run_batch2d_any(gtomx_pick_plan* pl, const PJ& PS, const Projector& P, const float* mats, int nb,
                    const float2* E, const float* ctf, float* best, int32_t* idx, int angle0, const float2* tw,
                    hipStream_t s) {
  const int n = pl->n;
  int L1, L2;
  GX_TRY(fft_len_split(n, &L1, &L2));
  const int tiles = (n / 2 + 1 + 63) / 64;
  GX_TRY(timed(pl, 0, s, [&]() -> int {
    auto go = [&](auto kernel, int threads) -> int {
      hipLaunchKernelGGL(kernel, dim3(tiles, nb * L2), dim3(threads), 0, s, PS, mats, E, ctf, pl->scr, pl->partial, n,
                         L2, tw);
      GX_LAUNCHED("k_pick2d_project_cols");
      return GTOMX_OK;
    };
    switch (L1) {
      case 16: return go(k_pick2d_project_cols<PJ, 16, 64>, 64 * LineShape<16>::R2);
      case 32: return go(k_pick2d_project_cols<PJ, 32, 64>, 64 * LineShape<32>::R2);
      case 48: return go(k_pick2d_project_cols<PJ, 48, 64>, 64 * LineShape<48>::R2);
      case 64: return go(k_pick2d_project_cols<PJ, 64, 64>, 64 * LineShape<64>::R2);
      default: return fail(GTOMX_ERR_INTERNAL, "image split not instantiated");
    }
  }));
  GX_TRY(herm_norm_any(n, P, mats, ctf, pl->corr, nb, s));
  GX_TRY(timed(pl, 1, s, [&]() -> int {
    hipLaunchKernelGGL(k_pick_scale, dim3(nb), dim3(256), 0, s, pl->partial, tiles * L2, pl->corr, 1, (double)pl->H,
                       pl->scale);
    GX_LAUNCHED("k_pick_scale");
    return GTOMX_OK;
  }));
  return timed(pl, 2, s, [&]() -> int {
    GX_TRY(image_cols_pass2_inverse(n, pl->scr, pl->G, (unsigned)nb, s));
    return pick_rows_max_long(n, pl->G, nb, (long long)pl->H, n, angle0, pl->scale, best, idx, s, pl->rows_variant);
  });
}

int run_batch2d_n(gtomx_pick_plan* pl, const CentredProjector* C, const Projector& P, const float* mats, int nb,
                  const float2* E, const float* ctf, float* best, int32_t* idx, int angle0, const float2* tw,
                  hipStream_t s) {
  if (pl->n > 256)
    return C ? run_batch2d_any(pl, *C, P, mats, nb, E, ctf, best, idx, angle0, tw, s)
             : run_batch2d_any(pl, P, P, mats, nb, E, ctf, best, idx, angle0, tw, s);
  switch (pl->n) {
#define GX_CASE(N)                                                                         \
  case N:                                                                                  \
    return C ? run_batch2d<N>(pl, *C, P, mats, nb, E, ctf, best, idx, angle0, tw, s)       \
             : run_batch2d<N>(pl, P, P, mats, nb, E, ctf, best, idx, angle0, tw, s);
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
  }
  return fail(GTOMX_ERR_UNSUPPORTED, "image side not supported");
}

}  // namespace

int check_plan_dims(const gtomx_pick_plan* pl, gtomx_int3 dimsvolume, unsigned nvolumes) {
  if (!pl) return fail(GTOMX_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (dimsvolume.x != pl->n || dimsvolume.y != pl->n || dimsvolume.z != pl->nz)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "dimsvolume does not match the plan");
  if (nvolumes == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "nvolumes must be positive");
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  if (dev != pl->device) return fail(GTOMX_ERR_INVALID_ARGUMENT, "plan was created on another device");
  return GTOMX_OK;
}

int upload_matrices(gtomx_pick_plan* pl, const gtomx_float3* h_angles, unsigned nangles, float oversample,
                    hipStream_t s) {
  if (pl->mats_ev_recorded) {
    GX_HIP(hipEventSynchronize(pl->mats_ev));
    pl->mats_ev_recorded = false;
  }
  if (pl->mats_cap < nangles) {
    dev_free(pl->mats);
    pl->mats = nullptr;
    pl->device_bytes -= pl->mats_cap * 9 * sizeof(float);
    pl->mats_cap = 0;
    void* d = nullptr;
    GX_TRY(dev_alloc(&d, (size_t)nangles * 9 * sizeof(float)));
    pl->mats = static_cast<float*>(d);
    pl->mats_cap = nangles;
    pl->device_bytes += pl->mats_cap * 9 * sizeof(float);
  }
  if (pl->h_mats_cap < nangles) {
    if (pl->h_mats) (void)hipHostFree(pl->h_mats);
    pl->h_mats = nullptr;
    pl->h_mats_cap = 0;
    void* h = nullptr;
    hipError_t e = hipHostMalloc(&h, (size_t)nangles * 9 * sizeof(float), hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(matrices)");
    pl->h_mats = static_cast<float*>(h);
    pl->h_mats_cap = nangles;
  }
  GX_TRY(gtomx_projector_matrices(h_angles, nangles, oversample, pl->h_mats));
  GX_HIP(hipMemcpyAsync(pl->mats, pl->h_mats, (size_t)nangles * 9 * sizeof(float), hipMemcpyHostToDevice, s));
  GX_HIP(hipEventRecord(pl->mats_ev, s));
  pl->mats_ev_recorded = true;
  return GTOMX_OK;
}

// Copies `bytes` of tile masks to the device through pinned staging; call after
// upload_matrices in the same call (which waited for the staging's last copy).
int upload_masks(gtomx_pick_plan* pl, const uint8_t* h, size_t bytes, hipStream_t s) {
  if (pl->tmasks_cap < bytes) {
    dev_free(pl->tmasks);
    pl->tmasks = nullptr;
    pl->device_bytes -= pl->tmasks_cap;
    pl->tmasks_cap = 0;
    void* d = nullptr;
    GX_TRY(dev_alloc(&d, bytes));
    pl->tmasks = static_cast<uint8_t*>(d);
    pl->tmasks_cap = bytes;
    pl->device_bytes += bytes;
  }
  if (pl->h_tmasks_cap < bytes) {
    if (pl->h_tmasks) (void)hipHostFree(pl->h_tmasks);
    pl->h_tmasks = nullptr;
    pl->h_tmasks_cap = 0;
    void* h2 = nullptr;
    hipError_t e = hipHostMalloc(&h2, bytes, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(tile masks)");
    pl->h_tmasks = static_cast<uint8_t*>(h2);
    pl->h_tmasks_cap = bytes;
  }
  std::memcpy(pl->h_tmasks, h, bytes);
  GX_HIP(hipMemcpyAsync(pl->tmasks, pl->h_tmasks, bytes, hipMemcpyHostToDevice, s));
  GX_HIP(hipEventRecord(pl->mats_ev, s));
  pl->mats_ev_recorded = true;
  return GTOMX_OK;
}

// The projector the sampler reads: the centred copy when every sampled q stays
// within |q| < p/2 + 1/2, else the reference's layout. Returns whether centred.
bool setup_projector(gtomx_pick_plan* pl, const gtomx_complex* d_projectorft, float oversample, hipStream_t s,
                     Projector* P, CentredProjector* C, int* status) {
  *P = Projector{reinterpret_cast<const float2*>(d_projectorft), pl->p, pl->p / 2 + 1,
                 (pl->p / 2 + 1) * pl->p * pl->p * 8};
  *status = GTOMX_OK;
  const bool centred = (double)oversample * (pl->n / 2) <= pl->p / 2 + 0.25;
  if (centred) {
    const int p = pl->p, PX = p / 2 + 2, PY = p + 3, OFF = p / 2 + 1;
    hipLaunchKernelGGL(k_centre_projector, dim3((PX * PY * PY + 255) / 256), dim3(256), 0, s, P->data, p, pl->cproj);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) *status = hip_fail(e, "k_centre_projector");
    *C = CentredProjector{pl->cproj, centred_bytes(p), PX * 8, PX * PY * 8, (float)((OFF * PY + OFF) * PX * 8)};
  }
  return centred;
}

int check_update_args(const gtomx_pick_plan* plan, const gtomx_complex* d_projectorft, float oversample,
                      gtomx_int3 dimsprojector, const gtomx_complex* d_expft, gtomx_int3 dimsvolume,
                      unsigned nvolumes, const float* d_best, const int32_t* d_idx) {
  GX_TRY(check_plan_dims(plan, dimsvolume, nvolumes));
  if (dimsprojector.x != plan->p || dimsprojector.y != plan->p || dimsprojector.z != plan->p)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "dimsprojector does not match the plan");
  if (!d_projectorft || !d_expft || !d_best || !d_idx)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL device buffer");
  if (!(oversample > 0.f) || !std::isfinite(oversample))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "projectoroversample must be positive");
  return GTOMX_OK;
}

int launch_update_batches(gtomx_pick_plan* pl, const Projector& P, const CentredProjector* C, const float* mats,
                          unsigned nangles, const float2* E, const float* ctf, float* best, int32_t* idx,
                          int angle0, hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  for (unsigned b0 = 0; b0 < nangles; b0 += pl->B) {
    const int nb = (int)std::min<unsigned>(pl->B, nangles - b0);
    if (pl->nz == 1)
      GX_TRY(run_batch2d_n(pl, C, P, mats + 9 * (size_t)b0, nb, E, ctf, best, idx, angle0 + (int)b0, tw, s));
    else
      GX_TRY(run_batch_n(pl, C, P, mats + 9 * (size_t)b0, nb, E, ctf, best, idx, angle0 + (int)b0, tw, s));
  }
  return GTOMX_OK;
}

namespace {

// The grouped search: group g (orientations g*group .. g*group + group - 1 of
// h_angles, which are also their indices) runs as one batch over the tiles set
// in its mask (volume v's masks at h_masks + (v*ngroups + g)*nt*nt); groups with
// empty masks are skipped.
int update_tiles(gtomx_pick_plan* plan, const gtomx_complex* d_projectorft, float oversample,
                 const gtomx_complex* d_expft, const float* d_ctf, size_t ctf_volume_stride, unsigned nvolumes,
                 const gtomx_float3* h_angles, unsigned ngroups, unsigned group, const uint8_t* h_masks,
                 float* d_best, int32_t* d_idx, hipStream_t s, unsigned* groups_run) {
  const int nt = plan->n / 32;
  const size_t mbytes = (size_t)nt * nt;
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  GX_TRY(upload_matrices(plan, h_angles, ngroups * group, oversample, s));
  GX_TRY(upload_masks(plan, h_masks, nvolumes * (size_t)ngroups * mbytes, s));
  Projector P;
  CentredProjector C{};
  int st = GTOMX_OK;
  const bool centred = setup_projector(plan, d_projectorft, oversample, s, &P, &C, &st);
  GX_TRY(st);
  // Consecutive groups with identical masks are one contiguous range of orientations and run
  // in batches of the plan's size (the in-order update is the same); a mask with every tile
  // set runs the unmasked kernels.
  const uint8_t full = (uint8_t)((1u << nt) - 1u);
  unsigned run = 0;
  for (unsigned v = 0; v < nvolumes; ++v) {
    const float2* Ev = reinterpret_cast<const float2*>(d_expft) + (size_t)v * plan->H;
    const float* cv = d_ctf ? d_ctf + (size_t)v * ctf_volume_stride : nullptr;
    float* bv = d_best + (size_t)v * plan->V;
    int32_t* iv = d_idx + (size_t)v * plan->V;
    for (unsigned g = 0; g < ngroups;) {
      const uint8_t* hm = h_masks + ((size_t)v * ngroups + g) * mbytes;
      bool any = false, all = true;
      for (size_t k = 0; k < mbytes; ++k) {
        any |= hm[k] != 0;
        all &= hm[k] == full;
      }
      unsigned g1 = g + 1;
      while (g1 < ngroups && std::memcmp(hm, h_masks + ((size_t)v * ngroups + g1) * mbytes, mbytes) == 0) ++g1;
      if (any) {
        run += g1 - g;
        const uint8_t* dm = all ? nullptr : plan->tmasks + ((size_t)v * ngroups + g) * mbytes;
        for (unsigned o0 = g * group, o1 = g1 * group; o0 < o1; o0 += plan->B) {
          const int nb = (int)std::min<unsigned>(plan->B, o1 - o0);
          GX_TRY(run_batch_n(plan, centred ? &C : nullptr, P, plan->mats + 9 * (size_t)o0, nb, Ev, cv, bv, iv,
                             (int)o0, tw, s, dm));
        }
      }
      g = g1;
    }
  }
  if (groups_run) *groups_run = run;
  return GTOMX_OK;
}

int check_grouped(const gtomx_pick_plan* plan, unsigned ngroups, unsigned group) {
  if (plan->nz == 1 || plan->n % 32 || plan->n / 32 > 8)
    return fail(GTOMX_ERR_UNSUPPORTED, "tile masks need n^3 volumes with n a multiple of 32, at most 256");
  if (group == 0 || group > plan->B) return fail(GTOMX_ERR_INVALID_ARGUMENT, "group must be in [1, plan batch]");
  if ((unsigned long long)ngroups * group > 0x7fffffffULL)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "orientation index exceeds int32");
  return GTOMX_OK;
}

// Voxels of tiles that were not searched: best = 0, index = -1 (and the float angle -1).
__global__ __launch_bounds__(256) void k_fill_unsearched(const uint8_t* __restrict__ searched, int n, float* best,
                                                         int32_t* idx, float* angle, size_t total) {
  const int nt = n / 32;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t V = (size_t)n * n * n;
    const size_t v = i / V, r = i % V;
    const int x = (int)(r % n), y = (int)((r / n) % n), z = (int)(r / ((size_t)n * n));
    if (searched[v * nt * nt * nt + ((size_t)(z / 32) * nt + y / 32) * nt + x / 32]) continue;
    best[i] = 0.f;
    idx[i] = -1;
    if (angle) angle[i] = -1.f;
  }
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int gtomx_pick_plan_create(gtomx_pick_plan** out, gtomx_int3 dp, gtomx_int3 dv, unsigned batch) {
  if (!out) return fail(GTOMX_ERR_INVALID_ARGUMENT, "plan pointer is NULL");
  *out = nullptr;
  const bool is2d = dv.z == 1;
  if (dv.x != dv.y || (!is2d && dv.x != dv.z))
    return fail(GTOMX_ERR_UNSUPPORTED, "volumes must be cubes (or images squares)");
  const bool big2d = is2d && dv.x > 256;  // micrographs: the generic-transform form of the 2D branch
  // volumes above 256: the generic-transform form (sides with a long-row pass)
  const bool big3d = !is2d && (dv.x == 512 || dv.x == 768 || dv.x == 1024);
  // every other volume side from 8 to 1024: the same form on the chirp-z transforms
  const bool gen3d = !is2d && !big3d && !fft_len_ok(dv.x) && dv.x >= 8 && dv.x <= 1024;
  const bool any3d = big3d || gen3d;
  if (big2d ? !row_len_native(dv.x) : (!any3d && !fft_len_ok(dv.x)))
    return fail(GTOMX_ERR_UNSUPPORTED,
                "volume side must be from 8 to 1024; image side 2^a or 3*2^a from 8 to 256, or 512, 768, "
                "1024, 1536, 2048, 3072 or 4096");
  if (dp.x != dp.y || dp.x != dp.z || dp.x < 2 || dp.x % 2 || dp.x > 512)
    return fail(GTOMX_ERR_UNSUPPORTED, "projector must be an even cube of side <= 512");
  if (batch > 256) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch must be <= 256");
  auto* pl = new gtomx_pick_plan();
  pl->n = dv.x;
  pl->nz = dv.z;
  pl->p = dp.x;
  // 2D: the reference's 240 orientations per pass (src/Correlation/SubTomograms.cu:35-36)
  pl->H = elements_ft(dv);
  pl->V = elements(dv);
  // above 256 the default batch keeps the spectra near 1 GiB (16 orientations at 4096^2)
  const unsigned big_b = (unsigned)std::max<size_t>(1, std::min<size_t>(240, (1ull << 30) / (pl->H * sizeof(float2))));
  // volumes above 256: spectra of about 8 GiB per batch (15 orientations at 512^3, 1 at 1024^3)
  const unsigned big3_b = (unsigned)std::max<size_t>(1, std::min<size_t>(32, (8ull << 30) / (pl->H * sizeof(float2))));
  // volumes of 128 and more: 128 orientations per batch (the fixed costs of a batch -- E and the CTF
  // in pass 1, the running maxima in pass 3 -- spread over more orientations; 8.7 GB of spectra at 256^3);
  // smaller volumes: 256 (their launches are short, so their number matters most: 1.89 ms for 588
  // orientations of 64^3 at 256 against 2.68 ms at 16)
  pl->B = batch ? batch : (is2d ? (big2d ? big_b : 240u) : (any3d ? big3_b : (dv.x >= 128 ? 128u : 256u)));
  pl->gen3d = gen3d;
  pl->og = dv.x >= 128 ? 2 : 1;
  int L1 = 0, L2 = 0;
  if (big2d && fft_len_split(dv.x, &L1, &L2) != GTOMX_OK) {
    delete pl;
    return fail(GTOMX_ERR_UNSUPPORTED, "image side has no two-pass split");
  }
  pl->nparts = big2d ? (dv.x / 2 + 1 + 63) / 64 * L2
                     : (any3d ? (int)((pl->H + 256 * P3_BINS - 1) / (256 * P3_BINS)) : 0);
  const std::vector<int> tiles = (is2d || any3d) ? std::vector<int>{0} : tiles_for(pl->n);
  pl->ntiles = (big2d || any3d) ? 1 : (is2d ? g2_tiles_for(pl->n) : (int)tiles.size());
  int s = GTOMX_OK;
  hipError_t he = hipGetDevice(&pl->device);
  if (he != hipSuccess) {
    plan_release(pl);
    return hip_fail(he, "hipGetDevice");
  }
  void* d = nullptr;
  auto alloc = [&](void** dst, size_t bytes) {
    if (s != GTOMX_OK) return;
    s = dev_alloc(dst, bytes);
    if (s == GTOMX_OK) pl->device_bytes += bytes;
  };
  // The batch buffer also serves as the C2R workspace in pick_finish (H elements).
  const size_t gstride = (big2d || any3d) ? pl->H : (is2d ? g2_stride_for(pl->n) : g_stride_for(pl->n));
  alloc(&d, std::max((size_t)pl->B * gstride, pl->H) * sizeof(float2));
  pl->G = static_cast<float2*>(d);
  d = nullptr;
  // up to 4 pass-1 waves per tile; above 256, one sum per pass-1 workgroup
  alloc(&d, (size_t)pl->B * ((big2d || any3d) ? (size_t)pl->nparts : (size_t)pl->ntiles * 4) * sizeof(float));
  pl->partial = static_cast<float*>(d);
  d = nullptr;
  alloc(&d, (size_t)pl->ntiles * sizeof(int));
  pl->tiles = static_cast<int*>(d);
  d = nullptr;
  alloc(&d, (size_t)pl->B * HERM_BLOCKS * sizeof(float));
  pl->corr = static_cast<float*>(d);
  d = nullptr;
  alloc(&d, (size_t)pl->B * sizeof(float));
  pl->scale = static_cast<float*>(d);
  d = nullptr;
  alloc(&pl->ls_mem, local_std_scratch_bytes(dv));
  alloc(&d, pl->V * sizeof(float));
  pl->ebuf = static_cast<float*>(d);
  d = nullptr;
  alloc(&d, (size_t)centred_bytes(pl->p));
  pl->cproj = static_cast<float2*>(d);
  if (big2d) {
    d = nullptr;
    alloc(&d, (size_t)pl->B * pl->H * sizeof(float2));
    pl->scr = static_cast<float2*>(d);
  } else if (is2d) {
    const size_t groups = (pl->B + pl->og - 1) / pl->og;
    d = nullptr;
    alloc(&d, groups * pl->V * sizeof(float));
    pl->gbest = static_cast<float*>(d);
    d = nullptr;
    alloc(&d, groups * pl->V * sizeof(int32_t));
    pl->gidx = static_cast<int32_t*>(d);
  }
  if (s != GTOMX_OK) {
    std::string msg = gtomx_last_error();
    plan_release(pl);
    return fail(s, msg);
  }
  pl->ls = local_std_carve(pl->ls_mem, dv);
  he = hipMemcpy(pl->tiles, tiles.data(), std::min(tiles.size(), (size_t)pl->ntiles) * sizeof(int),
                 hipMemcpyHostToDevice);
  int cus = 0;
  if (he == hipSuccess) he = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device);
  if (he != hipSuccess) {
    plan_release(pl);
    return hip_fail(he, "plan set-up");
  }
  pl->grid1[0] = resident_grid_for(pl->n, false, cus);
  pl->grid1[1] = resident_grid_for(pl->n, true, cus);
  he = hipEventCreateWithFlags(&pl->mats_ev, hipEventDisableTiming);
  if (he != hipSuccess) {
    plan_release(pl);
    return hip_fail(he, "hipEventCreate");
  }
  const float2* tw = nullptr;
  s = twiddles(&tw);
  if (s != GTOMX_OK) {
    std::string msg = gtomx_last_error();
    plan_release(pl);
    return fail(s, msg);
  }
  *out = pl;
  return GTOMX_OK;
}

int gtomx_pick_plan_destroy(gtomx_pick_plan* plan) {
  if (!plan) return GTOMX_OK;
  // Buffers may still be in use by the last call's work (plan_release waits for it).
  plan_release(plan);
  return GTOMX_OK;
}

int gtomx_pick_plan_set_rows_variant(gtomx_pick_plan* plan, int variant) {
  if (!plan) return fail(GTOMX_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (variant != GTOMX_ROWS_PLAIN && variant != GTOMX_ROWS_PREFETCH && variant != GTOMX_ROWS_ONE)
    return fail(GTOMX_ERR_INVALID_ARGUMENT,
                "rows variant must be GTOMX_ROWS_PLAIN, GTOMX_ROWS_PREFETCH or GTOMX_ROWS_ONE");
  if (variant == GTOMX_ROWS_ONE && !(plan->nz == 1 && (plan->n == 2048 || plan->n == 3072 || plan->n == 4096)))
    return fail(GTOMX_ERR_UNSUPPORTED, "GTOMX_ROWS_ONE takes images of side 2048, 3072 or 4096");
  plan->rows_variant = variant;
  return GTOMX_OK;
}

int gtomx_pick_plan_info(const gtomx_pick_plan* plan, unsigned* batch, size_t* device_bytes) {
  if (!plan) return fail(GTOMX_ERR_INVALID_ARGUMENT, "plan is NULL");
  if (batch) *batch = plan->B;
  if (device_bytes) *device_bytes = plan->device_bytes;
  return GTOMX_OK;
}

int gtomx_pick_plan_enable_timing(gtomx_pick_plan* plan, int enable) {
  if (!plan) return fail(GTOMX_ERR_INVALID_ARGUMENT, "plan is NULL");
  plan->timing = enable != 0;
  return GTOMX_OK;
}

int gtomx_pick_plan_timings(gtomx_pick_plan* plan, double* ms3, unsigned long long* launches3, int reset) {
  if (!plan) return fail(GTOMX_ERR_INVALID_ARGUMENT, "plan is NULL");
  for (const auto& pd : plan->pending) {
    GX_HIP(hipEventSynchronize(plan->ev_pool[pd.b]));
    float ms = 0.f;
    GX_HIP(hipEventElapsedTime(&ms, plan->ev_pool[pd.a], plan->ev_pool[pd.b]));
    plan->ms[pd.stage] += ms;
    plan->launches[pd.stage] += 1;
  }
  plan->pending.clear();
  plan->ev_used = 0;
  for (int k = 0; k < 3; ++k) {
    if (ms3) ms3[k] = plan->ms[k];
    if (launches3) launches3[k] = plan->launches[k];
    if (reset) {
      plan->ms[k] = 0;
      plan->launches[k] = 0;
    }
  }
  return GTOMX_OK;
}

int gtomx_pick_begin(gtomx_pick_plan* plan, gtomx_int3 dimsvolume, unsigned nvolumes, float* d_best,
                     int32_t* d_idx, void* stream) {
  GX_TRY(check_plan_dims(plan, dimsvolume, nvolumes));
  if (!d_best || !d_idx) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL output buffer");
  hipStream_t s = as_stream(stream);
  GX_TRY(plan->fence.enter(s));
  // d_ValueFill best = -1e30, angle = 0 (src/Correlation/SubTomograms.cu:41-42)
  GX_TRY(fill_f32(d_best, plan->V * nvolumes, -1e30f, s));
  GX_TRY(fill_i32(d_idx, plan->V * nvolumes, 0, s));
  return plan->fence.leave(s);
}

int gtomx_pick_update(gtomx_pick_plan* plan, const gtomx_complex* d_projectorft, float oversample,
                      gtomx_int3 dimsprojector, const gtomx_complex* d_expft, const float* d_ctf,
                      size_t ctf_volume_stride, gtomx_int3 dimsvolume, unsigned nvolumes,
                      const gtomx_float3* h_angles, unsigned nangles, unsigned angle_offset, float* d_best,
                      int32_t* d_idx, void* stream) {
  GX_TRY(check_plan_dims(plan, dimsvolume, nvolumes));
  if (dimsprojector.x != plan->p || dimsprojector.y != plan->p || dimsprojector.z != plan->p)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "dimsprojector does not match the plan");
  if (!d_projectorft || !d_expft || !d_best || !d_idx)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL device buffer");
  if (nangles == 0) return GTOMX_OK;
  if (!h_angles) return fail(GTOMX_ERR_INVALID_ARGUMENT, "h_angles is NULL");
  if ((unsigned long long)angle_offset + nangles > 0x7fffffffULL)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "orientation index exceeds int32");
  if (!(oversample > 0.f) || !std::isfinite(oversample))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "projectoroversample must be positive");
  hipStream_t s = as_stream(stream);
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  GX_TRY(plan->fence.enter(s));
  GX_TRY(upload_matrices(plan, h_angles, nangles, oversample, s));
  Projector P;
  CentredProjector C{};
  int st = GTOMX_OK;
  const bool centred = setup_projector(plan, d_projectorft, oversample, s, &P, &C, &st);
  GX_TRY(st);
  for (unsigned v = 0; v < nvolumes; ++v) {
    const float2* Ev = reinterpret_cast<const float2*>(d_expft) + (size_t)v * plan->H;
    const float* cv = d_ctf ? d_ctf + (size_t)v * ctf_volume_stride : nullptr;
    float* bv = d_best + (size_t)v * plan->V;
    int32_t* iv = d_idx + (size_t)v * plan->V;
    for (unsigned b0 = 0; b0 < nangles; b0 += plan->B) {
      int nb = (int)std::min<unsigned>(plan->B, nangles - b0);
      if (plan->nz == 1)
        GX_TRY(run_batch2d_n(plan, centred ? &C : nullptr, P, plan->mats + 9 * (size_t)b0, nb, Ev, cv, bv, iv,
                             (int)(angle_offset + b0), tw, s));
      else
        GX_TRY(run_batch_n(plan, centred ? &C : nullptr, P, plan->mats + 9 * (size_t)b0, nb, Ev, cv, bv, iv,
                           (int)(angle_offset + b0), tw, s));
    }
  }
  return plan->fence.leave(s);
}

int gtomx_pick_finish(gtomx_pick_plan* plan, const gtomx_complex* d_expft, gtomx_int3 dimsvolume,
                      unsigned nvolumes, float maskradius, float* d_best, int32_t* d_idx, float* d_bestangle,
                      void* stream) {
  GX_TRY(check_plan_dims(plan, dimsvolume, nvolumes));
  if (!d_expft || !d_best || !d_idx) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL device buffer");
  if (!(maskradius >= 0.f) || !std::isfinite(maskradius))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "maskradius must be finite and non-negative");
  hipStream_t s = as_stream(stream);
  GX_TRY(plan->fence.enter(s));
  GX_TRY(local_std_prepare(plan->ls, dimsvolume, maskradius, s));
  for (unsigned v = 0; v < nvolumes; ++v) {
    const float2* Ev = reinterpret_cast<const float2*>(d_expft) + (size_t)v * plan->H;
    // e = C2R_u(E_v) (src/Correlation/SubTomograms.cu:105); the batch buffer is the workspace.
    GX_TRY(fft_c2r(Ev, plan->G, plan->ebuf, dimsvolume, 1, 1.0f, s));
    GX_TRY(local_std_apply(plan->ls, plan->ebuf, dimsvolume, nullptr, nullptr, d_best + (size_t)v * plan->V,
                           d_idx + (size_t)v * plan->V, d_bestangle ? d_bestangle + (size_t)v * plan->V : nullptr,
                           s));
  }
  return plan->fence.leave(s);
}

int gtomx_pick_subtomograms(gtomx_pick_plan* plan, const gtomx_complex* d_projectorft, float oversample,
                            gtomx_int3 dimsprojector, const gtomx_complex* d_expft, const float* d_ctf,
                            size_t ctf_volume_stride, gtomx_int3 dimsvolume, unsigned nvolumes,
                            const gtomx_float3* h_angles, unsigned nangles, float maskradius, float* d_best,
                            int32_t* d_idx, float* d_bestangle, void* stream) {
  if (nangles > 0 && !h_angles) return fail(GTOMX_ERR_INVALID_ARGUMENT, "h_angles is NULL");
  GX_TRY(gtomx_pick_begin(plan, dimsvolume, nvolumes, d_best, d_idx, stream));
  GX_TRY(gtomx_pick_update(plan, d_projectorft, oversample, dimsprojector, d_expft, d_ctf, ctf_volume_stride,
                           dimsvolume, nvolumes, h_angles, nangles, 0, d_best, d_idx, stream));
  GX_TRY(gtomx_pick_finish(plan, d_expft, dimsvolume, nvolumes, maskradius, d_best, d_idx, d_bestangle, stream));
  return GTOMX_OK;
}

int gtomx_pick_update_tiles(gtomx_pick_plan* plan, const gtomx_complex* d_projectorft, float oversample,
                            gtomx_int3 dimsprojector, const gtomx_complex* d_expft, const float* d_ctf,
                            size_t ctf_volume_stride, gtomx_int3 dimsvolume, unsigned nvolumes,
                            const gtomx_float3* h_angles, unsigned ngroups, unsigned group, const uint8_t* h_tilemasks,
                            float* d_best, int32_t* d_idx, void* stream) {
  GX_TRY(check_update_args(plan, d_projectorft, oversample, dimsprojector, d_expft, dimsvolume, nvolumes, d_best,
                           d_idx));
  GX_TRY(check_grouped(plan, ngroups, group));
  if (ngroups == 0) return GTOMX_OK;
  if (!h_angles || !h_tilemasks) return fail(GTOMX_ERR_INVALID_ARGUMENT, "h_angles or h_tilemasks is NULL");
  hipStream_t s = as_stream(stream);
  GX_TRY(plan->fence.enter(s));
  GX_TRY(update_tiles(plan, d_projectorft, oversample, d_expft, d_ctf, ctf_volume_stride, nvolumes, h_angles, ngroups,
                      group, h_tilemasks, d_best, d_idx, s, nullptr));
  return plan->fence.leave(s);
}

int gtomx_pick_subtomograms_learned(gtomx_pick_plan* plan, gtomx_pnet* net, const gtomx_complex* d_projectorft,
                                    float oversample, gtomx_int3 dimsprojector, const gtomx_complex* d_expft,
                                    const float* d_ctf, size_t ctf_volume_stride, gtomx_int3 dimsvolume,
                                    unsigned nvolumes, const gtomx_float3* h_angles, unsigned ndirections,
                                    unsigned npsi, unsigned topk, float confidence, float maskradius, float* d_best,
                                    int32_t* d_idx, float* d_bestangle, gtomx_learned_stats* h_stats, void* stream) {
  GX_TRY(check_update_args(plan, d_projectorft, oversample, dimsprojector, d_expft, dimsvolume, nvolumes, d_best,
                           d_idx));
  GX_TRY(check_grouped(plan, ndirections, npsi));
  if (!net) return fail(GTOMX_ERR_INVALID_ARGUMENT, "network is NULL");
  if (ndirections != 500) return fail(GTOMX_ERR_UNSUPPORTED, "the network scores 500 directions");
  if (!h_angles) return fail(GTOMX_ERR_INVALID_ARGUMENT, "h_angles is NULL");
  if (topk == 0 || topk > ndirections) return fail(GTOMX_ERR_INVALID_ARGUMENT, "topk must be in [1, 500]");
  if (!(confidence >= 0.f && confidence <= 1.f)) return fail(GTOMX_ERR_INVALID_ARGUMENT, "confidence in [0, 1]");
  hipStream_t s = as_stream(stream);
  const int n = plan->n, nt = n / 32;
  const size_t ntl = (size_t)nt * nt * nt, mbytes = (size_t)nt * nt;
  GX_TRY(plan->fence.enter(s));
  if (plan->lcap < ntl * topk) {
    for (void* q : {(void*)plan->ldirs, (void*)plan->ltop1, (void*)plan->lcls, (void*)plan->lmass, (void*)plan->lpnp})
      dev_free(q);
    plan->ldirs = nullptr;
    plan->ltop1 = nullptr;
    plan->lcls = nullptr;
    plan->lmass = nullptr;
    plan->lpnp = nullptr;
    plan->lcap = 0;
    void *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr, *e = nullptr;
    GX_TRY(dev_alloc(&a, ntl * topk * sizeof(int32_t)));
    plan->ldirs = static_cast<int32_t*>(a);
    GX_TRY(dev_alloc(&b, ntl * sizeof(float)));
    plan->ltop1 = static_cast<float*>(b);
    GX_TRY(dev_alloc(&c, ntl * sizeof(int32_t)));
    plan->lcls = static_cast<int32_t*>(c);
    GX_TRY(dev_alloc(&d, ntl * sizeof(float)));
    plan->lmass = static_cast<float*>(d);
    GX_TRY(dev_alloc(&e, ntl * sizeof(float)));
    plan->lpnp = static_cast<float*>(e);
    plan->lcap = ntl * topk;
  }
  std::vector<uint8_t> masks((size_t)nvolumes * ndirections * mbytes, 0), searched((size_t)nvolumes * ntl, 0);
  std::vector<int32_t> dirs(ntl * topk);
  std::vector<float> mass(ntl), pnp(ntl);
  gtomx_learned_stats st{};
  for (unsigned v = 0; v < nvolumes; ++v) {
    // the network reads the block in real space: C2R_u(E_v) (its standardisation removes the scale)
    const float2* Ev = reinterpret_cast<const float2*>(d_expft) + (size_t)v * plan->H;
    GX_TRY(fft_c2r(Ev, plan->G, plan->ebuf, dimsvolume, 1, 1.0f, s));
    GX_TRY(gtomx_pnet_load_volume(net, plan->ebuf, dimsvolume, stream));
    GX_TRY(gtomx_pnet_forward(net, (unsigned)ntl, nullptr, stream));
    GX_TRY(gtomx_pnet_propose(net, topk, plan->ldirs, plan->ltop1, plan->lcls, plan->lmass, plan->lpnp, stream));
    GX_HIP(hipMemcpyAsync(dirs.data(), plan->ldirs, ntl * topk * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(mass.data(), plan->lmass, ntl * sizeof(float), hipMemcpyDeviceToHost, s));
    GX_HIP(hipMemcpyAsync(pnp.data(), plan->lpnp, ntl * sizeof(float), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    uint8_t* mv = masks.data() + (size_t)v * ndirections * mbytes;
    for (size_t t = 0; t < ntl; ++t) {
      const int tz = (int)(t / mbytes), ty = (int)((t / nt) % nt), tx = (int)(t % nt);
      const size_t mi = (size_t)tz * nt + ty;
      const uint8_t bit = (uint8_t)(1u << tx);
      // the tile's proposal is "no peak" or its top-k directions, whichever the network gives the
      // larger probability; below `confidence` the tile falls back to every direction
      const bool none = pnp[t] >= mass[t];
      const float conf = none ? pnp[t] : mass[t];
      if (conf < confidence) {  // not confident: every direction
        for (unsigned g = 0; g < ndirections; ++g) mv[g * mbytes + mi] |= bit;
        searched[v * ntl + t] = 1;
        st.exhaustive++;
      } else if (none) {  // confidently no peak: not searched
        st.skipped++;
      } else {
        for (unsigned k = 0; k < topk; ++k) {
          const int d = dirs[t * topk + k];
          if (d >= 0 && d < (int)ndirections) mv[(size_t)d * mbytes + mi] |= bit;
        }
        searched[v * ntl + t] = 1;
        st.shortlisted++;
      }
    }
  }
  st.tiles = (unsigned)(ntl * nvolumes);
  GX_TRY(gtomx_pick_begin(plan, dimsvolume, nvolumes, d_best, d_idx, stream));
  GX_TRY(update_tiles(plan, d_projectorft, oversample, d_expft, d_ctf, ctf_volume_stride, nvolumes, h_angles,
                      ndirections, npsi, masks.data(), d_best, d_idx, s, &st.groups_searched));
  GX_TRY(gtomx_pick_finish(plan, d_expft, dimsvolume, nvolumes, maskradius, d_best, d_idx, d_bestangle, stream));
  // the searched flags follow the masks in the staging: one more staged copy
  DevBuf sflag;
  GX_TRY(sflag.alloc(searched.size()));
  GX_HIP(hipMemcpyAsync(sflag.p, searched.data(), searched.size(), hipMemcpyHostToDevice, s));
  const size_t total = (size_t)nvolumes * plan->V;
  hipLaunchKernelGGL(k_fill_unsearched, dim3((unsigned)std::min<size_t>((total + 255) / 256, 8192)), dim3(256), 0, s,
                     sflag.as<uint8_t>(), n, d_best, d_idx, d_bestangle, total);
  GX_LAUNCHED("k_fill_unsearched");
  GX_HIP(hipStreamSynchronize(s));  // the flags' host copy and scratch go away on return
  if (h_stats) *h_stats = st;
  return plan->fence.leave(s);
}

}  // extern "C"
