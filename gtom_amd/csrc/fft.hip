// Batched R2C / C2R transforms of 3D, 2D and 1D volumes (axes 2^a up to 4096
// or 3*2^a up to 3072, from 8; the standalone entry points also take any other
// extent up to 2048 by the chirp-z form at the end of this file) for gfx950.
// Replace the cuFFT plans wrapped by d_FFTR2C and d_IFFTC2R (src/FFT/FFT.cu:7-42,
// src/FFT/IFFT.cu:8-56). Axes up to 256 (rows up to 256 reals) take one pass
// each; longer ones two.
//
// Layout: real (nx, ny, nz) x-fastest; half spectrum (nx/2+1, ny, nz).
// R2C = rows (real -> half, via an nx/2-point complex FFT) then y then z.
// C2R = z then y (in place on a copy) then rows (half -> real). The row C2R
// uses only the real parts of the x = 0 and x = nx/2 bins, which makes the
// result the real part of the Hermitian-completed inverse.
#include <algorithm>
#include <complex>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>
#include <type_traits>

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
__global__ __launch_bounds__(fft_col_tx<N>() * ColShape<N>::R2, (N == 512 ? 2 : 1)) void k_cols(
    float2* __restrict__ data, int width, long long axis_stride, long long nplanes, long long P1,
    long long S0, long long S1, const float2* __restrict__ tw) {
  constexpr int R1 = ColShape<N>::R1, R2 = ColShape<N>::R2, TX = fft_col_tx<N>();
  __shared__ float2 xch[R2 > 1 ? TX * N : 1];
  const int t = threadIdx.x, c = t % TX, j = t / TX;
  const int col = blockIdx.x * TX + c;
  const bool valid = col < width;
  // at 512 the stage twiddles come from an LDS copy (31 per thread would otherwise be held in registers)
  constexpr bool LDS_TW = N == 512;
  __shared__ float2 twl[LDS_TW ? N : 1];
  if constexpr (LDS_TW)
    for (int i = t; i < N; i += TX * R2) twl[i] = twiddle<-1, N>(tw, i);
  for (long long pl = blockIdx.y; pl < nplanes; pl += gridDim.y) {
    const long long base = (pl / P1) * S1 + (pl % P1) * S0 + col;
    float2 v[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m)
      v[m] = valid ? data[base + (long long)(j + R2 * m) * axis_stride] : make_float2(0.f, 0.f);
    __syncthreads();
    if constexpr (LDS_TW)
      fft_two_step<N, R1, R2, DIR>(v, j, twl, [&](int k1, int jj) { return &xch[(k1 * R2 + jj) * TX + c]; },
                                   TwiddleLDS{});
    else
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
// The workgroup's RW rows are one contiguous block in and out, so both go
// through LDS (the exchange area) with coalesced accesses; barriers separate
// the area's uses.
template <int N>
__global__ __launch_bounds__(256) void k_rows_c2r(const float2* __restrict__ in, float* __restrict__ out,
                                                  long long nrows, float scale,
                                                  const float2* __restrict__ tw, RowBand rb) {
  constexpr int NC = N / 2, NH = N / 2 + 1;
  constexpr int R1 = RowShape<NC>::R1, R2 = RowShape<NC>::R2, RW = 256 / R2;
  constexpr int GS = R2 * (R1 + 1);  // exchange floats2 per row (padded)
  constexpr int AREA = RW * (GS > NH ? GS : NH);
  __shared__ alignas(16) float2 xch[AREA];
  const int t = threadIdx.x, r = t / R2, j = t % R2;
  const long long v0 = (long long)blockIdx.x * RW, row0 = band_row(v0, rb);
  const int nr = (int)(nrows - v0 < RW ? nrows - v0 : RW);
  const float2* src = in + row0 * NH;
  for (int i = t; i < nr * NH; i += 256) xch[i] = src[i];
  __syncthreads();
  const bool valid = r < nr;
  float2 v[R1];
#pragma unroll
  for (int m = 0; m < R1; ++m) {
    const int k = j + R2 * m;
    float2 a = valid ? xch[r * NH + k] : make_float2(0.f, 0.f);
    float2 b = valid ? xch[r * NH + NC - k] : make_float2(0.f, 0.f);
    if (k == 0) {
      a.y = 0.f;
      b.y = 0.f;
    }
    const float2 bc = cconj(b);
    v[m] = cadd(cadd(a, bc), mul_pj<1>(cmul(twiddle<1, N>(tw, k), csub(a, bc))));
  }
  __syncthreads();  // the rows are in registers: the area becomes the exchange
  fft_two_step<NC, R1, R2, 1>(v, j, tw, [&](int k1, int jj) { return &xch[r * GS + jj * (R1 + 1) + k1]; });
  __syncthreads();  // exchange reads done: the area becomes the output rows
#pragma unroll
  for (int s = 0; s < R1; ++s) xch[r * NC + out_index<R1, R2>(j, s)] = cscale(v[s], scale);
  __syncthreads();
  float* dst = out + row0 * N;
  if ((reinterpret_cast<uintptr_t>(dst) & 15) == 0) {
    float4* d4 = reinterpret_cast<float4*>(dst);
    const float4* b4 = reinterpret_cast<const float4*>(xch);
    for (int i = t; i < nr * N / 4; i += 256) d4[i] = b4[i];
  } else {
    float2* d2 = reinterpret_cast<float2*>(dst);
    for (int i = t; i < nr * NC; i += 256) d2[i] = xch[i];
  }
}

// ---------------------------------------------------------------------------
// Row R2C: N reals -> N/2+1 bins. z[m] = x[2m] + i x[2m+1], Z = DFT_{N/2}(z),
// X[k] = ((Z[k] + conj Z[N/2-k]) - i w^k (Z[k] - conj Z[N/2-k])) / 2, w = e^{-2 pi i/N}.
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_rows_r2c(const float* __restrict__ in, float2* __restrict__ out,
                                                  long long nrows, const float2* __restrict__ tw, RowBand rb) {
  constexpr int NC = N / 2, NH = N / 2 + 1;
  constexpr int R1 = RowShape<NC>::R1, R2 = RowShape<NC>::R2, RW = 256 / R2;
  constexpr int GS = R2 * (R1 + 1);
  constexpr int ZS = NC + 1;  // padded Z row
  __shared__ float2 xch[R2 > 1 ? RW * GS : 1];
  __shared__ float2 zrow[RW * ZS];
  const int t = threadIdx.x, r = t / R2, j = t % R2;
  const bool valid = (long long)blockIdx.x * RW + r < nrows;
  const long long row = band_row((long long)blockIdx.x * RW, rb) + r;
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
// Strided axes above 256: the four-step decomposition L = L1 * L2. With n = n2 + L2 n1 and k = k1 + L1 k2,
//   X[k1 + L1 k2] = sum_n2 W_L2^{n2 k2} W_L^{n2 k1} sum_n1 x[n2 + L2 n1] W_L1^{n1 k1}.
// Pass 1 runs an L1-point transform for each n2, multiplies by W_L^{n2 k1} and
// stores at n2 + L2 k1 (the positions it read) in a scratch copy. Pass 2 runs
// an L2-point transform for each k1 over the entries L2 k1 + n2 and stores at
// k1 + L1 k2 of the destination: natural order without a transpose pass (any
// order of whole lines along a strided axis is coalesced). Both passes are
// batches of strided lines, addressed by LineMap. Rows above 256 reals run the
// same decomposition inside LDS (k_rows_*_long below).
// ---------------------------------------------------------------------------
struct LineMap {
  long long P1, S0s, S1s, S0d, S1d;  // plane q: base = (q / P1) * S1 + (q % P1) * S0
  long long Cs, Cd;                  // column c (adjacent threads) offset multipliers
  long long Gs, Gd;                  // sub-line g offset multipliers, g in [0, G)
  long long Ms, Md;                  // element strides in and out
  long long nplanes;                 // planes * G
  int width, G;
  int L, TG, TC;                     // output times W_L^{(g TG + c TC) k} when L > 0
  float scale;
  int band = 0;                      // > 0: entries at axis positions g + G i outside [0, band) and
                                     // [L - band, L) read as zero (the first pass of a banded input)
};

template <int N, int DIR, int TX>
__global__ __launch_bounds__(TX * LineShape<N>::R2) void k_lines(const float2* __restrict__ src,
                                                                 float2* __restrict__ dst, LineMap p,
                                                                 const float2* __restrict__ tw) {
  constexpr int R1 = LineShape<N>::R1, R2 = LineShape<N>::R2;
  __shared__ float2 xch[R2 > 1 ? TX * N : 1];
  const int t = threadIdx.x, c = t % TX, j = t / TX;
  const int col = blockIdx.x * TX + c;
  const bool valid = col < p.width;
  for (long long pl = blockIdx.y; pl < p.nplanes; pl += gridDim.y) {
    const long long q = pl / p.G;
    const int g = (int)(pl % p.G);
    // 64-bit line bases; offsets along a line are 32-bit (the host keeps every
    // buffer below 2^31 elements)
    const float2* sl = src + ((q / p.P1) * p.S1s + (q % p.P1) * p.S0s + col * p.Cs + g * p.Gs);
    float2* dl = dst + ((q / p.P1) * p.S1d + (q % p.P1) * p.S0d + col * p.Cd + g * p.Gd);
    const int ms = (int)p.Ms, md = (int)p.Md;
    float2 v[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m) {
      const int y = g + p.G * (j + R2 * m);
      const bool in = p.band == 0 || y < p.band || y >= p.L - p.band;
      v[m] = valid && in ? sl[(j + R2 * m) * ms] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_two_step<N, R1, R2, DIR>(v, j, tw, [&](int k1, int jj) { return &xch[(k1 * R2 + jj) * TX + c]; });
    if (valid) {
      const int e = g * p.TG + col * p.TC;
#pragma unroll
      for (int s = 0; s < R1; ++s) {
        const int k = out_index<R1, R2>(j, s);
        float2 o = v[s];
        if (p.L > 0) o = cmul(o, twiddle_rt<DIR>(tw, p.L, e * k));
        dl[k * md] = cscale(o, p.scale);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Rows of N = 2 NC reals with NC = L1 * L2 in [256, 2048]: the whole four-step
// in LDS, one HBM round trip. RWG rows per workgroup of 256 threads. Entry i
// of a row sits at i + 4 (i / L2) in LDS (4 float2 of padding per group of L2,
// which spreads both passes' strided reads over the banks). Pass A transforms
// the L2 lines {n2 + L2 n1} in place, times W_NC^{n2 k1}; pass B transforms
// the L1 contiguous groups in place, leaving Z[k1 + L1 k2] at entry L2 k1 + k2.
// Lines are dealt to threads in rounds of 256 items (R2 threads per line);
// threads past the last line compute on zeros, exchange through a sink and
// store nothing, so every thread meets every barrier.
// ---------------------------------------------------------------------------
template <int L1, int L2> struct LongRow {
  static constexpr int NC = L1 * L2, N = 2 * NC, NH = NC + 1;
  static constexpr int PITCH = NC + 4 * L1 + 1;
  __device__ static constexpr int phys(int i) { return i + 4 * (i / L2); }
  __device__ static constexpr int zpos(int k) { return phys(L2 * (k % L1) + k / L1); }  // Z[k] after pass B
};

// One pass over all lines of the workgroup's rows: line ln of row r holds
// entries pos(ln, i), i in [0, LN). The two-step exchange of a line goes
// through the line's own entries (thread j writes exchange slots (k1, j) to
// the entries j + R2 k1 it read itself), so a pass needs no other LDS.
template <int LN, int NL, int RWG, int DIR, int PITCH, class Pos, class Out>
__device__ __forceinline__ void long_row_pass(float2* buf, float2* sink, const float2* __restrict__ tw, Pos pos,
                                              Out store, int t) {
  constexpr int R1 = ColShape<LN>::R1, R2 = ColShape<LN>::R2, ITEMS = RWG * NL * R2,
                ROUNDS = (ITEMS + 255) / 256;
  const int j = t % R2;
  for (int rd = 0; rd < ROUNDS; ++rd) {
    const int item = rd * 256 + t;
    const bool active = item < ITEMS;
    const int lg = active ? item / R2 : 0, r = lg / NL, ln = lg % NL;
    float2* row = buf + r * PITCH;
    float2 v[R1];
#pragma unroll
    for (int m = 0; m < R1; ++m) v[m] = active ? row[pos(ln, j + R2 * m)] : make_float2(0.f, 0.f);
    fft_two_step<LN, R1, R2, DIR>(v, j, tw,
                                  [&](int k1, int jj) { return active ? &row[pos(ln, k1 * R2 + jj)] : sink; });
    if constexpr (R2 > 1) __syncthreads();  // the line's exchange reads are done before its entries change
    if (active) {
#pragma unroll
      for (int s = 0; s < R1; ++s) {
        const int k = out_index<R1, R2>(j, s);
        row[pos(ln, k)] = store(ln, k, v[s]);
      }
    }
  }
}

template <int L1, int L2, int RWG, int DIR>
__device__ __forceinline__ void long_row_passes(float2* buf, float2* sink, const float2* __restrict__ tw, int t) {
  using LR = LongRow<L1, L2>;
  // pass A: the L2 lines {n2 + L2 n1} of L1 points, times W_NC^{n2 k1}
  long_row_pass<L1, L2, RWG, DIR, LR::PITCH>(
      buf, sink, tw, [](int n2, int i) { return LR::phys(n2 + L2 * i); },
      [&](int n2, int k1, float2 x) { return cmul(x, twiddle_rt<DIR>(tw, LR::NC, n2 * k1)); }, t);
  __syncthreads();
  // pass B: the L1 contiguous groups of L2 points
  long_row_pass<L2, L1, RWG, DIR, LR::PITCH>(
      buf, sink, tw, [](int k1, int i) { return LR::phys(L2 * k1 + i); }, [](int, int, float2 x) { return x; }, t);
  __syncthreads();
}

// First row of a workgroup: rows in order, or with band > 0 only the rows
// [0, band) and [per - band, per) of every image of `per` rows (band a
// multiple of RWG, so a workgroup's rows stay contiguous).
template <int RWG>
__device__ __forceinline__ long long long_row0(long long nrows, int per, int band, int* nr) {
  const long long v = (long long)blockIdx.x * RWG;
  *nr = (int)(nrows - v < RWG ? nrows - v : RWG);
  if (band == 0) return v;
  const long long b = v / (2 * band);
  const int r = (int)(v % (2 * band));
  return b * per + (r < band ? r : per - 2 * band + r);
}

template <int L1, int L2, int RWG>
__global__ __launch_bounds__(256) void k_rows_r2c_long(const float* __restrict__ in, float2* __restrict__ out,
                                                       long long nrows, int per, int band,
                                                       const float2* __restrict__ tw) {
  using LR = LongRow<L1, L2>;
  constexpr int NC = LR::NC, NH = LR::NH;
  __shared__ float2 buf[RWG * LR::PITCH];
  __shared__ float2 sink[1];  // exchange target of threads without a line
  const int t = threadIdx.x;
  int nr;
  const long long row0 = long_row0<RWG>(nrows, per, band, &nr);
  const float2* src = reinterpret_cast<const float2*>(in) + row0 * NC;
  for (int i = t; i < nr * NC; i += 256) buf[(i / NC) * LR::PITCH + LR::phys(i % NC)] = src[i];
  __syncthreads();
  long_row_passes<L1, L2, RWG, -1>(buf, sink, tw, t);
  float2* dst = out + row0 * NH;
  for (int i = t; i < nr * NH; i += 256) {
    const int r = i / NH, k = i % NH;
    const float2* row = buf + r * LR::PITCH;
    float2 x;
    if (k == NC) {
      const float2 z0 = row[LR::zpos(0)];
      x = make_float2(z0.x - z0.y, 0.f);
    } else {
      const float2 zk = row[LR::zpos(k)], zn = cconj(row[LR::zpos((NC - k) % NC)]);
      x = cscale(csub(cadd(zk, zn), mul_pj<1>(cmul(twiddle_rt<-1>(tw, LR::N, k), csub(zk, zn)))), 0.5f);
    }
    dst[i] = x;
  }
}

// Half spectrum -> Z for the pair (k, NC - k) of a row: a = X[k], b = X[NC - k]
// (the DC and Nyquist bins at k = 0 taken as real); both results into LDS.
template <int L1, int L2>
__device__ __forceinline__ void long_row_z_pair(float2* row, int k, float2 a, float2 b,
                                                const float2* __restrict__ tw) {
  using LR = LongRow<L1, L2>;
  constexpr int NC = LR::NC;
  if (k == 0) {
    a.y = 0.f;
    b.y = 0.f;
  }
  const float2 ac = cconj(a), bc = cconj(b);
  row[LR::phys(k)] = cadd(cadd(a, bc), mul_pj<1>(cmul(twiddle_rt<1>(tw, LR::N, k), csub(a, bc))));
  if (k > 0)
    row[LR::phys(NC - k)] = cadd(cadd(b, ac), mul_pj<1>(cmul(twiddle_rt<1>(tw, LR::N, NC - k), csub(b, ac))));
}

// Loads nr rows of half spectra (pitch NH in `src`) into LDS and leaves their
// inverse Z (z[m] = x[2m] + i x[2m+1], unscaled) at zpos(m) of each row. Each
// thread reads its pairs (k, NC - k) from memory and writes Z directly (both
// reads coalesce: k rises and NC - k falls along the wavefront).
template <int L1, int L2, int RWG>
__device__ __forceinline__ void long_row_c2r_core(float2* buf, float2* sink, const float2* __restrict__ src, int nr,
                                                  const float2* __restrict__ tw, int t) {
  using LR = LongRow<L1, L2>;
  constexpr int NC = LR::NC, NH = LR::NH, NP = NC / 2 + 1;
  for (int i = t; i < nr * NP; i += 256) {
    const int r = i / NP, k = i % NP;
    long_row_z_pair<L1, L2>(buf + r * LR::PITCH, k, src[r * NH + k], src[r * NH + NC - k], tw);
  }
  __syncthreads();
  long_row_passes<L1, L2, RWG, 1>(buf, sink, tw, t);
}

template <int L1, int L2, int RWG>
__global__ __launch_bounds__(256) void k_rows_c2r_long(const float2* __restrict__ in, float* __restrict__ out,
                                                       long long nrows, int per, int band, float scale,
                                                       const float2* __restrict__ tw) {
  using LR = LongRow<L1, L2>;
  constexpr int NC = LR::NC, NH = LR::NH;
  __shared__ float2 buf[RWG * LR::PITCH];
  __shared__ float2 sink[1];  // exchange target of threads without a line
  const int t = threadIdx.x;
  int nr;
  const long long row0 = long_row0<RWG>(nrows, per, band, &nr);
  long_row_c2r_core<L1, L2, RWG>(buf, sink, in + row0 * NH, nr, tw, t);
  float2* dst = reinterpret_cast<float2*>(out) + row0 * NC;
  for (int i = t; i < nr * NC; i += 256) dst[i] = cscale(buf[(i / NC) * LR::PITCH + LR::zpos(i % NC)], scale);
}

// The search's row pass for images above 256: the row C2R of every
// orientation of a batch (spectra H apart in G), each value times scale[o],
// folded into best / index in orientation order (strict >, first index on
// ties). The workgroup's rows keep their running maxima in registers. With PF
// the pairs of orientation o + 1 are loaded into registers while orientation
// o is transformed (the loads are in flight across the passes; more VGPRs, so
// fewer resident waves: the autotuner chooses per image side).
template <int L1, int L2, int RWG, bool PF>
__global__ __launch_bounds__(256) void k_pick_rows_max_long(const float2* __restrict__ G, int nb, long long H,
                                                            int nrows, int angle0, const float* __restrict__ scale,
                                                            float* __restrict__ best, int32_t* __restrict__ bidx,
                                                            const float2* __restrict__ tw) {
  using LR = LongRow<L1, L2>;
  constexpr int NC = LR::NC, NH = LR::NH, NP = NC / 2 + 1, Q = (RWG * NC + 255) / 256,
                QP = (RWG * NP + 255) / 256;
  __shared__ float2 buf[RWG * LR::PITCH];
  __shared__ float2 sink[1];
  const int t = threadIdx.x;
  const int row0 = blockIdx.x * RWG;
  const int nr = nrows - row0 < RWG ? nrows - row0 : RWG;
  float2* bp = reinterpret_cast<float2*>(best) + (size_t)row0 * NC;  // values 2m, 2m+1 of row r at r * NC + m
  int2* ip = reinterpret_cast<int2*>(bidx) + (size_t)row0 * NC;
  float bv[2 * Q];
  int32_t bi[2 * Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int i = t + 256 * q;
    const float2 b = i < nr * NC ? bp[i] : make_float2(-1e30f, -1e30f);
    const int2 a = i < nr * NC ? ip[i] : make_int2(0, 0);
    bv[2 * q] = b.x;
    bv[2 * q + 1] = b.y;
    bi[2 * q] = a.x;
    bi[2 * q + 1] = a.y;
  }
  float2 pa[PF ? QP : 1], pb[PF ? QP : 1];  // prefetched pairs (k, NC - k) of the next orientation
  auto fetch = [&](int o) {
    const float2* src = G + (size_t)o * H + (size_t)row0 * NH;
#pragma unroll
    for (int q = 0; q < QP; ++q) {
      const int i = t + 256 * q;
      if (i < nr * NP) {
        const int r = i / NP, k = i % NP;
        pa[q] = src[r * NH + k];
        pb[q] = src[r * NH + NC - k];
      }
    }
  };
  if constexpr (PF) fetch(0);
  for (int o = 0; o < nb; ++o) {
    const float sc = scale[o];
    // The transform's LDS addresses and twiddles are recomputed from an opaque
    // copy of the thread index in each iteration, rather than hoisted out of
    // the loop into registers (which costs occupancy).
    int tt = t;
    asm volatile("" : "+v"(tt));
    __syncthreads();  // the previous orientation's reads of buf are done
    if constexpr (PF) {
#pragma unroll
      for (int q = 0; q < QP; ++q) {
        const int i = tt + 256 * q;
        if (i < nr * NP) long_row_z_pair<L1, L2>(buf + (i / NP) * LR::PITCH, i % NP, pa[q], pb[q], tw);
      }
      if (o + 1 < nb) fetch(o + 1);
      __syncthreads();
      long_row_passes<L1, L2, RWG, 1>(buf, sink, tw, tt);
    } else {
      long_row_c2r_core<L1, L2, RWG>(buf, sink, G + (size_t)o * H + (size_t)row0 * NH, nr, tw, tt);
    }
    const int32_t oi = angle0 + o;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = tt + 256 * q;
      if (i < nr * NC) {
        const float2 z = buf[(i / NC) * LR::PITCH + LR::zpos(i % NC)];
        const float c0 = z.x * sc, c1 = z.y * sc;
        if (c0 > bv[2 * q]) {
          bv[2 * q] = c0;
          bi[2 * q] = oi;
        }
        if (c1 > bv[2 * q + 1]) {
          bv[2 * q + 1] = c1;
          bi[2 * q + 1] = oi;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int i = t + 256 * q;
    if (i < nr * NC) {
      bp[i] = make_float2(bv[2 * q], bv[2 * q + 1]);
      ip[i] = make_int2(bi[2 * q], bi[2 * q + 1]);
    }
  }
}

// The row pass with one row per workgroup of NC / 8 threads (NC = L1 * L2
// complex points, L2 = 32; rows of 2048, 3072 and 4096 reals): the transform
// and in-order update of k_pick_rows_max_long, with every LDS address a
// per-thread base plus compile-time offsets (the row buffer keeps LongRow's
// layout: element 32 a + b at 36 a + b), the twiddles of the Z pairs
// (exp(-2 pi i m / 2 NC), m < NC) and of the two-step stages in LDS tables
// staged once per workgroup, and pass A's output twiddles in registers.
// Thread t: pass A line la = t / R2a (points la + 32 i) with index ja; pass B
// line lb = t / R2b (points 32 lb + i) with index jb; the running maxima of
// pairs t + T q, q < 8.
template <int L1, int L2>
__global__ __launch_bounds__(L1 * L2 / 8) void k_pick_rows_max_one(const float2* __restrict__ G, int nb, long long H,
                                                                   int angle0, const float* __restrict__ scale,
                                                                   float* __restrict__ best,
                                                                   int32_t* __restrict__ bidx,
                                                                   const float2* __restrict__ tw) {
  using LR = LongRow<L1, L2>;
  constexpr int NC = LR::NC, NH = LR::NH, N = LR::N, T = NC / 8, Q = 8;
  using RA = ColShape<L1>;
  using RB = ColShape<L2>;
  constexpr int R1a = RA::R1, R2a = RA::R2, R1b = RB::R1, R2b = RB::R2;
  constexpr int ITEMS_A = L2 * R2a, ITEMS_B = L1 * R2b;
  static_assert(L2 == 32 && LR::phys(32 * 5 + 7) == 36 * 5 + 7, "row layout: element 32 a + b at 36 a + b");
  static_assert(ITEMS_B == T && ITEMS_A <= T && ITEMS_A % 64 == 0, "pass B one line item per thread");
  static_assert(T % L1 == 0 && T / L1 * 8 <= 32, "update addresses");
  __shared__ float2 buf[LR::PITCH];
  __shared__ float2 tz[NC];  // exp(-2 pi i m / N), m < NC
  __shared__ float2 ta[L1];  // exp(-2 pi i m / L1)
  __shared__ float2 tb[L2];  // exp(-2 pi i m / L2)
  const int t = threadIdx.x;
  const long long row = blockIdx.x;
  for (int m = t; m < NC; m += T) tz[m] = twiddle<-1, N>(tw, m);
  if (t < L1) ta[t] = twiddle<-1, L1>(tw, t);
  if (t < L2) tb[t] = twiddle<-1, L2>(tw, t);
  // pass A's output twiddles conj(W_NC^{la k1}), the same for every orientation
  float2 twa[R1a];
  {
    const int la = t / R2a, ja = t % R2a;
#pragma unroll
    for (int sl = 0; sl < R1a; ++sl) twa[sl] = twiddle<1, N>(tw, 2 * la * out_index<R1a, R2a>(ja, sl));
  }
  float2* bp = reinterpret_cast<float2*>(best) + row * NC;  // values 2m, 2m+1 of the row at m
  int2* ip = reinterpret_cast<int2*>(bidx) + row * NC;
  float bv[2 * Q];
  int32_t bi[2 * Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const float2 b = bp[t + T * q];
    const int2 a = ip[t + T * q];
    bv[2 * q] = b.x;
    bv[2 * q + 1] = b.y;
    bi[2 * q] = a.x;
    bi[2 * q + 1] = a.y;
  }
  for (int o = 0; o < nb; ++o) {
    const float sc = scale[o];
    const float2* src = G + (size_t)o * H + row * NH;
    // bases from an opaque copy of the thread index in each iteration (not
    // hoisted into registers across the loop); the modulo keeps its range known
    int tt = t;
    asm volatile("" : "+v"(tt));
    tt %= T;
    __syncthreads();  // the previous orientation's reads of buf are done (and the tables are written)
    {
      // Z[k] and Z[NC - k] from X[k], X[NC - k], k = tt + T q (X[0], X[NC] real); thread 0 also k = NC / 2
      const int pz = tt + 4 * (tt >> 5);                // phys(tt)
      const int pm = (NC - tt) + 4 * ((NC - tt) >> 5);  // phys(NC - tt)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int k = tt + T * q;
        float2 a = src[k], b = src[NC - k];
        if (q == 0 && tt == 0) {
          a.y = 0.f;
          b.y = 0.f;
        }
        const float2 ac = cconj(a), bc = cconj(b);
        buf[pz + (T + T / 8) * q] = cadd(cadd(a, bc), mul_pj<1>(cmul(cconj(tz[k]), csub(a, bc))));
        if (q > 0 || tt > 0)
          buf[pm - (T + T / 8) * q] = cadd(cadd(b, ac), mul_pj<1>(cmul(cconj(tz[NC - k]), csub(b, ac))));
      }
      if (tt == 0) {
        const float2 a = src[NC / 2], ac = cconj(a);
        buf[LR::phys(NC / 2)] = cadd(cadd(a, ac), mul_pj<1>(cmul(cconj(tz[NC / 2]), csub(a, ac))));
      }
    }
    __syncthreads();
    if (tt < ITEMS_A) {  // whole waves
      // pass A: line la of L1 points as R1a x R2a, then times conj(W_NC^{la k1})
      const int la = tt / R2a, ja = tt % R2a;
      const int bA = 36 * ja + la, bA2 = 36 * R1a * ja + la;
      float2 v[R1a];
#pragma unroll
      for (int m = 0; m < R1a; ++m) v[m] = buf[bA + 36 * R2a * m];
      dft<R1a, 1>(v);
#pragma unroll
      for (int k1 = 1; k1 < R1a; ++k1) v[k1] = cmul(v[k1], cconj(ta[ja * k1]));
#pragma unroll
      for (int k1 = 0; k1 < R1a; ++k1) buf[bA + 36 * R2a * k1] = v[k1];
      __syncthreads();
#pragma unroll
      for (int u = 0; u < R1a / R2a; ++u) {
        float2 z[R2a];
#pragma unroll
        for (int jj = 0; jj < R2a; ++jj) z[jj] = buf[bA2 + 36 * R2a * u + 36 * jj];
        dft<R2a, 1>(z);
#pragma unroll
        for (int k2 = 0; k2 < R2a; ++k2) v[u * R2a + k2] = z[k2];
      }
      __syncthreads();  // the exchange reads are done before the line's entries change
#pragma unroll
      for (int sl = 0; sl < R1a; ++sl) buf[36 * out_index<R1a, R2a>(ja, sl) + la] = cmul(v[sl], twa[sl]);
    } else {
      __syncthreads();  // the two barriers of the waves with a line
      __syncthreads();
    }
    __syncthreads();
    {
      // pass B: line lb (32 contiguous points) as R1b x R2b
      const int lb = tt / R2b, jb = tt % R2b;
      const int bB = 36 * lb + jb, bB2 = 36 * lb + R1b * jb, bB3 = 36 * lb + (R1b / R2b) * jb;
      float2 v[R1b];
#pragma unroll
      for (int m = 0; m < R1b; ++m) v[m] = buf[bB + R2b * m];
      dft<R1b, 1>(v);
#pragma unroll
      for (int k1 = 1; k1 < R1b; ++k1) v[k1] = cmul(v[k1], cconj(tb[jb * k1]));
#pragma unroll
      for (int k1 = 0; k1 < R1b; ++k1) buf[bB + R2b * k1] = v[k1];
      __syncthreads();
#pragma unroll
      for (int u = 0; u < R1b / R2b; ++u) {
        float2 z[R2b];
#pragma unroll
        for (int jj = 0; jj < R2b; ++jj) z[jj] = buf[bB2 + R2b * u + jj];
        dft<R2b, 1>(z);
#pragma unroll
        for (int k2 = 0; k2 < R2b; ++k2) v[u * R2b + k2] = z[k2];
      }
      __syncthreads();
#pragma unroll
      for (int u = 0; u < R1b / R2b; ++u)
#pragma unroll
        for (int k2 = 0; k2 < R2b; ++k2) buf[bB3 + u + R1b * k2] = v[u * R2b + k2];
    }
    __syncthreads();
    const int32_t oi = angle0 + o;
    const int bM = 36 * (tt % L1) + tt / L1;  // Z[tt + T q] at bM + (T / L1) q
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const float2 z = buf[bM + (T / L1) * q];
      const float c0 = z.x * sc, c1 = z.y * sc;
      if (c0 > bv[2 * q]) {
        bv[2 * q] = c0;
        bi[2 * q] = oi;
      }
      if (c1 > bv[2 * q + 1]) {
        bv[2 * q + 1] = c1;
        bi[2 * q + 1] = oi;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    bp[t + T * q] = make_float2(bv[2 * q], bv[2 * q + 1]);
    ip[t + T * q] = make_int2(bi[2 * q], bi[2 * q + 1]);
  }
}

// Zeroes the rows of half spectra outside a RowBand (rows of NH bins, nrows
// in all: every plane of every volume).
__global__ __launch_bounds__(256) void k_zero_outside_band(float2* __restrict__ data, long long nrows, int NH,
                                                           RowBand b) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nrows * NH; i += (long long)gridDim.x * 256) {
    const long long row = i / NH;
    const int y = (int)(row % b.ny), z = (int)((row / b.ny) % b.nz);
    const bool in = (y < b.by || y >= b.ny - b.by) && (b.bz == 0 || z < b.bz || z >= b.nz - b.bz);
    if (!in) data[i] = make_float2(0.f, 0.f);
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
    GX_CASE(512)
#undef GX_CASE
    default: return fail(GTOMX_ERR_UNSUPPORTED, "column FFT length not supported");
  }
}

template <int N> int rows_per_block() { return 256 / RowShape<N / 2>::R2; }

int launch_rows_c2r(int n, const float2* in, float* out, long long nrows, float scale, const float2* tw,
                    hipStream_t s, RowBand rb = RowBand{}) {
  auto go = [&](auto kernel, int rw) -> int {
    if (rb.by % rw) return fail(GTOMX_ERR_INTERNAL, "row band not a multiple of the rows per workgroup");
    dim3 grid((unsigned)((nrows + rw - 1) / rw));
    hipLaunchKernelGGL(kernel, grid, dim3(256), 0, s, in, out, nrows, scale, tw, rb);
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

int launch_rows_r2c(int n, const float* in, float2* out, long long nrows, const float2* tw, hipStream_t s,
                    RowBand rb = RowBand{}) {
  auto go = [&](auto kernel, int rw) -> int {
    if (rb.by % rw) return fail(GTOMX_ERR_INTERNAL, "row band not a multiple of the rows per workgroup");
    dim3 grid((unsigned)((nrows + rw - 1) / rw));
    hipLaunchKernelGGL(kernel, grid, dim3(256), 0, s, in, out, nrows, tw, rb);
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

// Factorisation L = L1 * L2 of the lengths above the one-pass kernels.
bool split_len(int L, int* L1, int* L2) {
  switch (L) {
    case 256: *L1 = 16, *L2 = 16; return true;
    case 384: *L1 = 16, *L2 = 24; return true;
    case 512: *L1 = 16, *L2 = 32; return true;
    case 768: *L1 = 32, *L2 = 24; return true;
    case 1024: *L1 = 32, *L2 = 32; return true;
    case 1536: *L1 = 48, *L2 = 32; return true;
    case 2048: *L1 = 64, *L2 = 32; return true;
    case 3072: *L1 = 64, *L2 = 48; return true;
    case 4096: *L1 = 64, *L2 = 64; return true;
    default: return false;
  }
}

// Column tile of a pass: up to 64 adjacent lines (512-byte segments) within
// 256 threads, narrower when the pass has fewer lines.
template <int N, int DIR>
int launch_lines_n(const float2* src, float2* dst, const LineMap& p, const float2* tw, hipStream_t s) {
  constexpr int R2 = LineShape<N>::R2;
  auto go = [&](auto kernel, int TX) -> int {
    dim3 grid((p.width + TX - 1) / TX, (unsigned)std::min<long long>(p.nplanes, 65535));
    hipLaunchKernelGGL(kernel, grid, dim3(TX * R2), 0, s, src, dst, p, tw);
    GX_LAUNCHED("k_lines");
    return GTOMX_OK;
  };
  if (p.width >= 64 && 64 * R2 <= 256) return go(k_lines<N, DIR, 64>, 64);
  if (p.width >= 32 && 32 * R2 <= 256) return go(k_lines<N, DIR, 32>, 32);
  return go(k_lines<N, DIR, 16>, 16);
}

template <int DIR>
int launch_lines(int n, const float2* src, float2* dst, const LineMap& p, const float2* tw, hipStream_t s) {
  switch (n) {
    case 16: return launch_lines_n<16, DIR>(src, dst, p, tw, s);
    case 24: return launch_lines_n<24, DIR>(src, dst, p, tw, s);
    case 32: return launch_lines_n<32, DIR>(src, dst, p, tw, s);
    case 48: return launch_lines_n<48, DIR>(src, dst, p, tw, s);
    case 64: return launch_lines_n<64, DIR>(src, dst, p, tw, s);
    default: return fail(GTOMX_ERR_INTERNAL, "four-step factor not instantiated");
  }
}

// One strided axis of length L in place; above 256 through `scratch`, which
// mirrors data's extent.
template <int DIR>
int cols_axis(int L, float2* data, float2* scratch, int width, long long as, long long nplanes, long long P1,
              long long S0, long long S1, const float2* tw, hipStream_t s, int band = 0) {
  // one pass up to 256 and at 512 (the exchange of 16 columns fits LDS), two otherwise or
  // with a band
  if (L <= 256 || (L == 512 && band == 0))
    return launch_cols_n<DIR>(L, data, width, as, nplanes, P1, S0, S1, tw, s);
  int L1, L2;
  if (!split_len(L, &L1, &L2) || !scratch) return fail(GTOMX_ERR_INTERNAL, "axis length not split");
  if ((long long)L * as >= (1LL << 31)) return fail(GTOMX_ERR_UNSUPPORTED, "axis span of 2^31 elements or more");
  LineMap a{};
  a.P1 = P1, a.S0s = a.S0d = S0, a.S1s = a.S1d = S1, a.Cs = a.Cd = 1;
  a.G = L2, a.Gs = a.Gd = as, a.Ms = a.Md = L2 * as, a.nplanes = nplanes * L2, a.width = width;
  a.L = L, a.TG = 1, a.TC = 0, a.scale = 1.f, a.band = band;
  GX_TRY(launch_lines<DIR>(L1, data, scratch, a, tw, s));
  LineMap b = a;
  b.G = L1, b.Gs = L2 * as, b.Gd = as, b.Ms = as, b.Md = L1 * as, b.nplanes = nplanes * L1, b.L = 0, b.band = 0;
  return launch_lines<DIR>(L2, scratch, data, b, tw, s);
}

// Long rows in one pass (NC = N/2 in [256, 2048]).
// nrows rows in order, or (band > 0) the 2 band rows of nrows / (2 band)
// images of `per` rows each.
template <bool R2C>
int launch_rows_long(int n, const void* in, void* out, long long nrows, float scale, const float2* tw,
                     hipStream_t s, int per = 0, int band = 0) {
  auto go = [&](auto r2c_kernel, auto c2r_kernel, int rwg) -> int {
    if (band % rwg) return fail(GTOMX_ERR_INTERNAL, "row band not a multiple of the rows per workgroup");
    dim3 grid((unsigned)((nrows + rwg - 1) / rwg));
    if constexpr (R2C) {
      hipLaunchKernelGGL(r2c_kernel, grid, dim3(256), 0, s, static_cast<const float*>(in), static_cast<float2*>(out),
                         nrows, per, band, tw);
      GX_LAUNCHED("k_rows_r2c_long");
    } else {
      hipLaunchKernelGGL(c2r_kernel, grid, dim3(256), 0, s, static_cast<const float2*>(in), static_cast<float*>(out),
                         nrows, per, band, scale, tw);
      GX_LAUNCHED("k_rows_c2r_long");
    }
    return GTOMX_OK;
  };
  switch (n / 2) {
#define GX_LONG(NC, A, B, W) \
  case NC: return go(k_rows_r2c_long<A, B, W>, k_rows_c2r_long<A, B, W>, W);
    GX_LONG(256, 16, 16, 8)
    GX_LONG(384, 16, 24, 8)
    GX_LONG(512, 16, 32, 4)
    GX_LONG(768, 32, 24, 2)
    GX_LONG(1024, 32, 32, 2)
    GX_LONG(1536, 48, 32, 1)
    GX_LONG(2048, 64, 32, 1)
#undef GX_LONG
    default: return fail(GTOMX_ERR_INTERNAL, "long row length not instantiated");
  }
}


// Stream-ordered scratch for the two-pass axes: taken from the stream's pool
// and given back on the same stream, so no synchronisation is needed.
struct StreamScratch {
  float2* p = nullptr;
  hipStream_t s = nullptr;
  int take(size_t n, hipStream_t st) {
    s = st;
    void* q = nullptr;
    hipError_t e = hipMallocAsync(&q, n * sizeof(float2), st);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync(FFT scratch)");
    p = static_cast<float2*>(q);
    return GTOMX_OK;
  }
  ~StreamScratch() {
    if (p) (void)hipFreeAsync(p, s);
  }
};

// The y and z transforms of half spectra over their first `cols` columns
// (cols = 0: all n/2 + 1).
int yz_axes(float2* data, float2* scr, gtomx_int3 d, unsigned batch, int dir, const float2* tw, hipStream_t s,
            int cols = 0) {
  const long long nh = d.x / 2 + 1;
  const int w = cols > 0 ? cols : (int)nh;
  auto zax = [&](auto DIRC) {
    return cols_axis<decltype(DIRC)::value>(d.z, data, scr, w, nh * d.y, (long long)d.y * batch, d.y, nh,
                                            nh * d.y * d.z, tw, s);
  };
  auto yax = [&](auto DIRC) {
    return cols_axis<decltype(DIRC)::value>(d.y, data, scr, w, nh, (long long)d.z * batch, 1LL << 62, nh * d.y, 0,
                                            tw, s);
  };
  using F = std::integral_constant<int, -1>;
  using B = std::integral_constant<int, 1>;
  if (dir < 0) {
    if (d.y > 1) GX_TRY(yax(F{}));
    if (d.z > 1) GX_TRY(zax(F{}));
  } else {
    if (d.z > 1) GX_TRY(zax(B{}));
    if (d.y > 1) GX_TRY(yax(B{}));
  }
  return GTOMX_OK;
}

int fft_r2c_any(const float* in, float2* out, gtomx_int3 d, unsigned batch, hipStream_t s);
int fft_c2r_any(float2* work, float* out, gtomx_int3 d, unsigned batch, float scale, hipStream_t s);

int fft_r2c(const float* in, float2* out, gtomx_int3 d, unsigned batch, hipStream_t s) {
  if (!dims_native(d)) return fft_r2c_any(in, out, d, batch, s);
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  const long long rows = (long long)d.y * d.z * batch;
  StreamScratch scr;
  if (d.y > 256 || d.z > 256) GX_TRY(scr.take(elements_ft(d) * batch, s));
  if (d.x > 256)
    GX_TRY(launch_rows_long<true>(d.x, in, out, rows, 1.f, tw, s));
  else
    GX_TRY(launch_rows_r2c(d.x, in, out, rows, tw, s));
  return yz_axes(out, scr.p, d, batch, -1, tw, s);
}

int fft_c2r(const float2* in, float2* work, float* out, gtomx_int3 d, unsigned batch, float scale, hipStream_t s,
            int cols) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  const size_t H = elements_ft(d) * batch;
  if (cols < 0 || cols > d.x / 2 + 1) return fail(GTOMX_ERR_INTERNAL, "column count outside the half spectrum");
  if (work != in) GX_HIP(hipMemcpyAsync(work, in, H * sizeof(float2), hipMemcpyDeviceToDevice, s));
  if (!dims_native(d)) return fft_c2r_any(work, out, d, batch, scale, s);
  StreamScratch scr;
  if (d.y > 256 || d.z > 256) GX_TRY(scr.take(H, s));
  GX_TRY(yz_axes(work, scr.p, d, batch, 1, tw, s, cols));
  const long long rows = (long long)d.y * d.z * batch;
  if (d.x > 256) return launch_rows_long<false>(d.x, work, out, rows, scale, tw, s);
  return launch_rows_c2r(d.x, work, out, rows, scale, tw, s);
}


// The search's row pass above 256 (k_pick_rows_max_long): nrows rows per
// spectrum, nb spectra H apart.
int pick_rows_max_long(int n, const float2* G, int nb, long long H, int nrows, int angle0, const float* scale,
                       float* best, int32_t* idx, hipStream_t s, int variant) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  auto go = [&](auto kernel, int rwg) -> int {
    hipLaunchKernelGGL(kernel, dim3((unsigned)((nrows + rwg - 1) / rwg)), dim3(256), 0, s, G, nb, H, nrows, angle0,
                       scale, best, idx, tw);
    GX_LAUNCHED("k_pick_rows_max_long");
    return GTOMX_OK;
  };
  if (variant == GTOMX_ROWS_ONE) {  // one row per workgroup of NC / 8 threads, LDS twiddles
    auto one = [&](auto kernel, int threads) -> int {
      hipLaunchKernelGGL(kernel, dim3((unsigned)nrows), dim3(threads), 0, s, G, nb, H, angle0, scale, best, idx, tw);
      GX_LAUNCHED("k_pick_rows_max_one");
      return GTOMX_OK;
    };
    switch (n) {
      case 2048: return one(k_pick_rows_max_one<32, 32>, 128);
      case 3072: return one(k_pick_rows_max_one<48, 32>, 192);
      case 4096: return one(k_pick_rows_max_one<64, 32>, 256);
      default: return fail(GTOMX_ERR_UNSUPPORTED, "the one-row pass takes rows of 2048, 3072 or 4096 reals");
    }
  }
  switch (n / 2) {
#define GX_LONG(NC, A, B, W) \
  case NC:                   \
    return variant ? go(k_pick_rows_max_long<A, B, W, true>, W) : go(k_pick_rows_max_long<A, B, W, false>, W);
    GX_LONG(256, 16, 16, 8)
    GX_LONG(384, 16, 24, 8)
    GX_LONG(512, 16, 32, 4)
    GX_LONG(768, 32, 24, 2)
    GX_LONG(1024, 32, 32, 2)
    GX_LONG(1536, 48, 32, 1)
    GX_LONG(2048, 64, 32, 1)
#undef GX_LONG
    default: return fail(GTOMX_ERR_UNSUPPORTED, "row length not supported by the search's long-row pass");
  }
}

int fft_len_split(int L, int* L1, int* L2) {
  if (L <= 256 || !split_len(L, L1, L2)) return fail(GTOMX_ERR_UNSUPPORTED, "length has no two-pass split");
  return GTOMX_OK;
}

int image_cols_pass2_inverse(int n, const float2* scratch, float2* out, unsigned nb, hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  int L1, L2;
  GX_TRY(fft_len_split(n, &L1, &L2));
  const long long nh = n / 2 + 1, as = nh;
  LineMap b{};
  b.P1 = 1LL << 62, b.S0s = b.S0d = nh * n, b.Cs = b.Cd = 1;
  b.G = L1, b.Gs = L2 * as, b.Gd = as, b.Ms = as, b.Md = L1 * as, b.nplanes = (long long)nb * L1;
  b.width = (int)nh, b.L = 0, b.scale = 1.f;
  return launch_lines<1>(L2, scratch, out, b, tw, s);
}

int image_c2r_banded(float2* work, float* out, int n, unsigned nb, int cols, int band, float scale, hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  if (n <= 256 || band <= 0 || 2 * band > n || cols < 1 || cols > n / 2 + 1)
    return fail(GTOMX_ERR_INTERNAL, "banded C2R outside its range");
  const long long nh = n / 2 + 1;
  StreamScratch scr;
  GX_TRY(scr.take((size_t)nh * n * nb, s));
  GX_TRY(cols_axis<1>(n, work, scr.p, cols, nh, (long long)nb, 1LL << 62, nh * n, 0, tw, s));
  return launch_rows_long<false>(n, work, out, (long long)nb * 2 * band, scale, tw, s, n, band);
}

int image_r2c_banded(const float* in, float2* out, int n, unsigned nb, int band, hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  if (n <= 256 || band <= 0 || 2 * band > n) return fail(GTOMX_ERR_INTERNAL, "banded R2C outside its range");
  const long long nh = n / 2 + 1;
  // the rows between the bands are never written: the y transform's first
  // pass reads them as zero and its second pass writes every row
  GX_TRY(launch_rows_long<true>(n, in, out, (long long)nb * 2 * band, 1.f, tw, s, n, band));
  StreamScratch scr;
  GX_TRY(scr.take((size_t)nh * n * nb, s));
  return cols_axis<-1>(n, out, scr.p, (int)nh, nh, (long long)nb, 1LL << 62, nh * n, 0, tw, s, band);
}

int rows_per_workgroup(int n) {
  switch (n) {
#define GX_CASE(N) \
  case N:          \
    return rows_per_block<N>();
    GX_FOR_SIZES(GX_CASE)
#undef GX_CASE
  }
  return 0;
}

int volume_c2r_banded(float2* work, float* out, gtomx_int3 d, unsigned nb, int cols, int by, int bz, float scale,
                      hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  if (d.x > 256 || d.y > 256 || d.z > 256 || by <= 0 || 2 * by > d.y || bz < 0 || 2 * bz > d.z)
    return fail(GTOMX_ERR_INTERNAL, "banded C2R outside its range");
  GX_TRY(yz_axes(work, nullptr, d, nb, 1, tw, s, cols));
  RowBand rb;
  rb.ny = d.y, rb.by = by, rb.nz = d.z, rb.bz = bz;
  const long long rows = (long long)nb * (bz ? 2 * bz : d.z) * 2 * by;
  return launch_rows_c2r(d.x, work, out, rows, scale, tw, s, rb);
}

int volume_r2c_banded(const float* in, float2* out, gtomx_int3 d, unsigned nb, int by, int bz, hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  if (d.x > 256 || d.y > 256 || d.z > 256 || by <= 0 || 2 * by > d.y || bz < 0 || 2 * bz > d.z)
    return fail(GTOMX_ERR_INTERNAL, "banded R2C outside its range");
  const long long nh = d.x / 2 + 1, all = (long long)nb * d.z * d.y;
  RowBand rb;
  rb.ny = d.y, rb.by = by, rb.nz = d.z, rb.bz = bz;
  hipLaunchKernelGGL(k_zero_outside_band, dim3((unsigned)std::min<long long>((all * nh + 255) / 256, 8192)),
                     dim3(256), 0, s, out, all, (int)nh, rb);
  GX_LAUNCHED("k_zero_outside_band");
  GX_TRY(launch_rows_r2c(d.x, in, out, (long long)nb * (bz ? 2 * bz : d.z) * 2 * by, tw, s, rb));
  if (d.y > 1) {  // planes outside the z band are zero: their y transform is skipped
    if (bz > 0) {
      const long long pl = nh * d.y;
      GX_TRY(launch_cols_n<-1>(d.y, out, (int)nh, nh, (long long)nb * bz, bz, pl, pl * d.z, tw, s));
      GX_TRY(launch_cols_n<-1>(d.y, out + (d.z - bz) * pl, (int)nh, nh, (long long)nb * bz, bz, pl, pl * d.z, tw,
                               s));
    } else {
      GX_TRY(launch_cols_n<-1>(d.y, out, (int)nh, nh, (long long)d.z * nb, 1LL << 62, nh * d.y, 0, tw, s));
    }
  }
  if (d.z > 1)
    GX_TRY(launch_cols_n<-1>(d.z, out, (int)nh, nh * d.y, (long long)d.y * nb, d.y, nh, nh * d.y * d.z, tw, s));
  return GTOMX_OK;
}

int fft_c2r_cols_inplace(float2* data, gtomx_int3 d, unsigned batch, hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  StreamScratch scr;
  if (d.y > 256 || d.z > 256) GX_TRY(scr.take(elements_ft(d) * batch, s));
  return yz_axes(data, scr.p, d, batch, 1, tw, s);
}

// ---------------------------------------------------------------------------
// Axes of any length L <= 2048 that the kernels above do not take (odd
// lengths, other prime factors), by Bluestein's chirp-z algorithm: with
// c[j] = exp(-i pi j^2 / L), the forward DFT is
//   X[k] = c[k] sum_j (x[j] c[j]) conj(c[k - j]),
// a circular convolution of length M (the shortest native length >= 2L - 1)
// done with the native M-point column transforms: load a[j] = x[j] c[j]
// (zero for j >= L), forward M-point DFT, times B[k] / M with B the DFT of
// b[m] = conj(c[m]) (m in (-L, L), wrapped), inverse M-point DFT, times c[k].
// The inverse DFT uses conj(c) for c and c for b. The lines of one axis are
// gathered into an M x (lines) block (line index fastest, the layout the
// column kernels take), in chunks that keep the block within 2^31 elements
// and 256 MB.

// Line l of an axis: its first point at (l / P) * S1 + (l % P) * S0, its
// points es apart.
struct AxisLines {
  long long P, S0, S1, es;
};
// (line indices below 2^32: 32-bit division; P above 2^31 means no period)
__device__ __forceinline__ long long line_base(const AxisLines& a, long long l) {
  if (a.P > 0x7fffffffLL) return l * a.S0;
  const unsigned u = (unsigned)l, p = (unsigned)a.P;
  return (long long)(u / p) * a.S1 + (long long)(u % p) * a.S0;
}

// MODE 0: complex points; 1: real points (the x rows of an R2C); 2: a half
// spectrum of nh points completed as X[k] = conj(X[L - k]) for k >= nh (the x
// rows of a C2R). DIR -1 multiplies by c, +1 by conj(c).
template <int MODE, int DIR>
__global__ __launch_bounds__(256) void k_blue_load(const void* __restrict__ src, AxisLines a, long long l0, int nl,
                                                   int L, int nh, int M, const float2* __restrict__ chirp,
                                                   float2* __restrict__ W) {
  const long long total = (long long)M * nl;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int j = (int)(i / nl), l = (int)(i % nl);
    float2 v = make_float2(0.f, 0.f);
    if (j < L) {
      const long long b = line_base(a, l0 + l);
      if constexpr (MODE == 0) {
        v = static_cast<const float2*>(src)[b + j * a.es];
      } else if constexpr (MODE == 1) {
        v = make_float2(static_cast<const float*>(src)[b + j * a.es], 0.f);
      } else {
        const float2* X = static_cast<const float2*>(src);
        v = j < nh ? X[b + j * a.es] : cconj(X[b + (L - j) * a.es]);
      }
      const float2 c = chirp[j];
      v = cmul(v, DIR < 0 ? c : cconj(c));
    }
    W[i] = v;
  }
}

__global__ __launch_bounds__(256) void k_blue_mul(float2* __restrict__ W, int nl, int M,
                                                  const float2* __restrict__ B) {
  const long long total = (long long)M * nl;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < total; i += (long long)gridDim.x * 256)
    W[i] = cmul(W[i], B[i / nl]);
}

// MODE 0: complex points k < L; 1: the half spectrum k < nh; 2: the real part
// times scale (k < L).
template <int MODE, int DIR>
__global__ __launch_bounds__(256) void k_blue_store(const float2* __restrict__ W, AxisLines a, long long l0, int nl,
                                                    int nout, const float2* __restrict__ chirp, void* __restrict__ dst,
                                                    float scale) {
  const long long total = (long long)nout * nl;
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
    const int k = (int)(i / nl), l = (int)(i % nl);
    const float2 c = chirp[k];
    const float2 v = cmul(W[i], DIR < 0 ? c : cconj(c));
    const long long o = line_base(a, l0 + l) + k * a.es;
    if constexpr (MODE == 2)
      static_cast<float*>(dst)[o] = v.x * scale;
    else
      static_cast<float2*>(dst)[o] = v;
  }
}

// The row modes (1, 2: points contiguous, es = 1) through a 64 x 64 LDS tile:
// the rows are read along the row and the block written along the line index,
// both coalesced. Grid (ceil(M / 64), ceil(nl / 64)) for the load,
// (ceil(nout / 64), ceil(nl / 64)) for the store.
template <int MODE, int DIR>
__global__ __launch_bounds__(256) void k_blue_load_rows(const void* __restrict__ src, AxisLines a, long long l0,
                                                        int nl, int L, int nh, int M,
                                                        const float2* __restrict__ chirp, float2* __restrict__ W) {
  __shared__ float tx[64][65], ty[64][65];  // split: a column read is conflict-free
  __shared__ long long lbase[64];
  const int jb = blockIdx.x * 64, lb = blockIdx.y * 64, t = threadIdx.x;
  if (t < 64) lbase[t] = line_base(a, l0 + lb + t);
  __syncthreads();
  for (int r = 0; r < 16; ++r) {
    const int ll = (t >> 6) + 4 * r, jj = t & 63;
    const int l = lb + ll, j = jb + jj;
    float2 v = make_float2(0.f, 0.f);
    if (l < nl && j < L) {
      const long long b = lbase[ll];
      if constexpr (MODE == 1) {
        v = make_float2(static_cast<const float*>(src)[b + j], 0.f);
      } else {
        const float2* X = static_cast<const float2*>(src);
        v = j < nh ? X[b + j] : cconj(X[b + (L - j)]);
      }
      const float2 c = chirp[j];
      v = cmul(v, DIR < 0 ? c : cconj(c));
    }
    tx[ll][jj] = v.x;
    ty[ll][jj] = v.y;
  }
  __syncthreads();
  for (int r = 0; r < 16; ++r) {
    const int jj = (t >> 6) + 4 * r, ll = t & 63;
    const int l = lb + ll, j = jb + jj;
    if (l < nl && j < M) W[(long long)j * nl + l] = make_float2(tx[ll][jj], ty[ll][jj]);
  }
}

template <int MODE, int DIR>
__global__ __launch_bounds__(256) void k_blue_store_rows(const float2* __restrict__ W, AxisLines a, long long l0,
                                                         int nl, int nout, const float2* __restrict__ chirp,
                                                         void* __restrict__ dst, float scale) {
  __shared__ float tx[64][65], ty[64][65];  // split: a column read is conflict-free
  __shared__ long long lbase[64];
  const int kb = blockIdx.x * 64, lb = blockIdx.y * 64, t = threadIdx.x;
  if (t < 64) lbase[t] = line_base(a, l0 + lb + t);
  for (int r = 0; r < 16; ++r) {
    const int kk = (t >> 6) + 4 * r, ll = t & 63;
    const int l = lb + ll, k = kb + kk;
    if (l < nl && k < nout) {
      const float2 c = chirp[k];
      const float2 v = cmul(W[(long long)k * nl + l], DIR < 0 ? c : cconj(c));
      tx[ll][kk] = v.x;
      ty[ll][kk] = v.y;
    }
  }
  __syncthreads();
  for (int r = 0; r < 16; ++r) {
    const int ll = (t >> 6) + 4 * r, kk = t & 63;
    const int l = lb + ll, k = kb + kk;
    if (l < nl && k < nout) {
      const long long o = lbase[ll] + k;
      if constexpr (MODE == 2)
        static_cast<float*>(dst)[o] = tx[ll][kk] * scale;
      else
        static_cast<float2*>(dst)[o] = make_float2(tx[ll][kk], ty[ll][kk]);
    }
  }
}

// Chirp and kernel spectra of one length, per device, computed once in long
// double: chirp[j] = exp(-i pi (j^2 mod 2L) / L), bf / bi = the M-point DFTs
// of the forward / inverse convolution kernels, divided by M.
struct BlueTables {
  int M = 0;
  float2* chirp = nullptr;
  float2* bf = nullptr;
  float2* bi = nullptr;
};
std::mutex g_blue_mu;
std::map<std::pair<int, int>, BlueTables> g_blue;

int blue_tables(int L, BlueTables* out) {
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_blue_mu);
  auto it = g_blue.find({dev, L});
  if (it != g_blue.end()) {
    *out = it->second;
    return GTOMX_OK;
  }
  // the shortest native length (2^a or 3 * 2^a, at least 8) that holds the linear convolution
  if (L > 2048) return fail(GTOMX_ERR_UNSUPPORTED, "lengths without a native transform must be at most 2048");
  int M = 8;
  while (M < 2 * L - 1 || !fft_len_ok_any(M)) ++M;
  const long double pi = 3.141592653589793238462643383279502884L;
  std::vector<std::complex<long double>> c(L), bf(M), bi(M);
  for (int j = 0; j < L; ++j) {
    const long long r = ((long long)j * j) % (2LL * L);
    c[j] = std::polar(1.0L, -pi * (long double)r / (long double)L);
  }
  std::vector<std::complex<long double>> kf(M, 0.0L), ki(M, 0.0L);
  for (int m = 0; m < L; ++m) {
    kf[m] = std::conj(c[m]);
    ki[m] = c[m];
    if (m > 0) {
      kf[M - m] = std::conj(c[m]);
      ki[M - m] = c[m];
    }
  }
  std::vector<std::complex<long double>> w(M);
  for (int t = 0; t < M; ++t) w[t] = std::polar(1.0L, -2.0L * pi * (long double)t / (long double)M);
  for (int k = 0; k < M; ++k) {
    std::complex<long double> sf = 0.0L, si = 0.0L;
    for (int m = 0; m < M; ++m) {
      const std::complex<long double>& e = w[(int)(((long long)m * k) % M)];
      sf += kf[m] * e;
      si += ki[m] * e;
    }
    bf[k] = sf / (long double)M;
    bi[k] = si / (long double)M;
  }
  std::vector<float2> h((size_t)L + 2 * (size_t)M);
  for (int j = 0; j < L; ++j) h[j] = make_float2((float)c[j].real(), (float)c[j].imag());
  for (int k = 0; k < M; ++k) {
    h[L + k] = make_float2((float)bf[k].real(), (float)bf[k].imag());
    h[L + M + k] = make_float2((float)bi[k].real(), (float)bi[k].imag());
  }
  void* d = nullptr;
  hipError_t e = hipMalloc(&d, h.size() * sizeof(float2));
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(chirp tables)");
  e = hipMemcpy(d, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return hip_fail(e, "hipMemcpy(chirp tables)");
  }
  BlueTables t;
  t.M = M;
  t.chirp = static_cast<float2*>(d);
  t.bf = t.chirp + L;
  t.bi = t.bf + M;
  g_blue[{dev, L}] = t;
  *out = t;
  return GTOMX_OK;
}

unsigned blue_grid(long long n) { return (unsigned)std::max<long long>(1, std::min<long long>((n + 255) / 256, 65536)); }

// One axis of nlines lines of length L by the chirp-z form: IN / OUT are the
// load and store modes above, src / dst laid out by as / ad (they may be the
// same array: every chunk is loaded before it is stored).
template <int IN, int OUT, int DIR>
int blue_axis(const void* src, AxisLines as, void* dst, AxisLines ad, long long nlines, int L, int nh, float scale,
              hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  BlueTables t;
  GX_TRY(blue_tables(L, &t));
  const int M = t.M;
  // chunks within 2^31 elements, 256 MB and 65535 tiles of 64 lines
  const long long cap =
      std::max<long long>(1, std::min({((1LL << 31) - 1) / M, (32LL << 20) / M, 65535LL * 64}));
  if (nlines >= (1LL << 32)) return fail(GTOMX_ERR_UNSUPPORTED, "2^32 or more lines on one axis");
  const int chunk = (int)std::min<long long>(nlines, cap);
  StreamScratch W, W2;
  GX_TRY(W.take((size_t)M * chunk, s));
  if (M > 256 && M != 512) GX_TRY(W2.take((size_t)M * chunk, s));  // the two-pass column lengths
  const int nout = OUT == 1 ? nh : L;
  for (long long l0 = 0; l0 < nlines; l0 += chunk) {
    const int nl = (int)std::min<long long>(chunk, nlines - l0);
    if constexpr (IN == 0) {
      hipLaunchKernelGGL((k_blue_load<IN, DIR>), dim3(blue_grid((long long)M * nl)), dim3(256), 0, s, src, as, l0,
                         nl, L, nh, M, t.chirp, W.p);
      GX_LAUNCHED("k_blue_load");
    } else {
      if (as.es != 1) return fail(GTOMX_ERR_INTERNAL, "row loads take contiguous lines");
      hipLaunchKernelGGL((k_blue_load_rows<IN, DIR>), dim3((M + 63) / 64, (nl + 63) / 64), dim3(256), 0, s, src, as,
                         l0, nl, L, nh, M, t.chirp, W.p);
      GX_LAUNCHED("k_blue_load_rows");
    }
    GX_TRY(cols_axis<-1>(M, W.p, W2.p, nl, nl, 1, 1LL << 62, (long long)M * nl, 0, tw, s));
    hipLaunchKernelGGL(k_blue_mul, dim3(blue_grid((long long)M * nl)), dim3(256), 0, s, W.p, nl, M,
                       DIR < 0 ? t.bf : t.bi);
    GX_LAUNCHED("k_blue_mul");
    GX_TRY(cols_axis<1>(M, W.p, W2.p, nl, nl, 1, 1LL << 62, (long long)M * nl, 0, tw, s));
    if constexpr (OUT == 0) {
      hipLaunchKernelGGL((k_blue_store<OUT, DIR>), dim3(blue_grid((long long)nout * nl)), dim3(256), 0, s, W.p, ad,
                         l0, nl, nout, t.chirp, dst, scale);
      GX_LAUNCHED("k_blue_store");
    } else {
      if (ad.es != 1) return fail(GTOMX_ERR_INTERNAL, "row stores take contiguous lines");
      hipLaunchKernelGGL((k_blue_store_rows<OUT, DIR>), dim3((nout + 63) / 64, (nl + 63) / 64), dim3(256), 0, s, W.p,
                         ad, l0, nl, nout, t.chirp, dst, scale);
      GX_LAUNCHED("k_blue_store_rows");
    }
  }
  return GTOMX_OK;
}

// The y and z axes of half spectra, each native or by the chirp-z form.
int yz_axes_any(float2* data, gtomx_int3 d, unsigned batch, int dir, hipStream_t s) {
  const float2* tw = nullptr;
  GX_TRY(twiddles(&tw));
  const long long nh = d.x / 2 + 1;
  StreamScratch scr;
  if ((d.y > 256 && fft_len_ok_any(d.y)) || (d.z > 256 && fft_len_ok_any(d.z)))
    GX_TRY(scr.take(elements_ft(d) * batch, s));
  auto yax = [&](auto DIRC) -> int {
    constexpr int D = decltype(DIRC)::value;
    if (fft_len_ok_any(d.y))
      return cols_axis<D>(d.y, data, scr.p, (int)nh, nh, (long long)d.z * batch, 1LL << 62, nh * d.y, 0, tw, s);
    const AxisLines a{nh, 1, nh * d.y, nh};
    return blue_axis<0, 0, D>(data, a, data, a, nh * d.z * batch, d.y, 0, 1.f, s);
  };
  auto zax = [&](auto DIRC) -> int {
    constexpr int D = decltype(DIRC)::value;
    if (fft_len_ok_any(d.z))
      return cols_axis<D>(d.z, data, scr.p, (int)nh, nh * d.y, (long long)d.y * batch, d.y, nh, nh * d.y * d.z, tw,
                          s);
    const AxisLines a{nh * d.y, 1, nh * d.y * d.z, nh * d.y};
    return blue_axis<0, 0, D>(data, a, data, a, nh * d.y * batch, d.z, 0, 1.f, s);
  };
  using F = std::integral_constant<int, -1>;
  using B = std::integral_constant<int, 1>;
  if (dir < 0) {
    if (d.y > 1) GX_TRY(yax(F{}));
    if (d.z > 1) GX_TRY(zax(F{}));
  } else {
    if (d.z > 1) GX_TRY(zax(B{}));
    if (d.y > 1) GX_TRY(yax(B{}));
  }
  return GTOMX_OK;
}

bool dims_native(gtomx_int3 d) {
  return row_len_native(d.x) && (d.y == 1 || fft_len_ok_any(d.y)) && (d.z == 1 || fft_len_ok_any(d.z));
}

int fft_r2c_any(const float* in, float2* out, gtomx_int3 d, unsigned batch, hipStream_t s) {
  if (dims_native(d)) return fft_r2c(in, out, d, batch, s);
  const long long rows = (long long)d.y * d.z * batch, nh = d.x / 2 + 1;
  if (row_len_native(d.x)) {
    const float2* tw = nullptr;
    GX_TRY(twiddles(&tw));
    if (d.x > 256)
      GX_TRY(launch_rows_long<true>(d.x, in, out, rows, 1.f, tw, s));
    else
      GX_TRY(launch_rows_r2c(d.x, in, out, rows, tw, s));
  } else {
    GX_TRY((blue_axis<1, 1, -1>(in, AxisLines{1LL << 62, d.x, 0, 1}, out, AxisLines{1LL << 62, nh, 0, 1}, rows, d.x,
                                (int)nh, 1.f, s)));
  }
  return yz_axes_any(out, d, batch, -1, s);
}

// work holds a copy of the half spectra (it is transformed in place).
int fft_c2r_any(float2* work, float* out, gtomx_int3 d, unsigned batch, float scale, hipStream_t s) {
  GX_TRY(yz_axes_any(work, d, batch, 1, s));
  const long long rows = (long long)d.y * d.z * batch, nh = d.x / 2 + 1;
  if (row_len_native(d.x)) {
    const float2* tw = nullptr;
    GX_TRY(twiddles(&tw));
    if (d.x > 256) return launch_rows_long<false>(d.x, work, out, rows, scale, tw, s);
    return launch_rows_c2r(d.x, work, out, rows, scale, tw, s);
  }
  return blue_axis<2, 2, 1>(work, AxisLines{1LL << 62, nh, 0, 1}, out, AxisLines{1LL << 62, d.x, 0, 1}, rows, d.x,
                            (int)nh, scale, s);
}

int rows_c2r_any_inplace(float2* G, int n, int nb, long long H, hipStream_t s) {
  const long long nh = n / 2 + 1, nn = (long long)n * n;
  if (H != nh * nn) return fail(GTOMX_ERR_INTERNAL, "rows_c2r_any_inplace takes n^3 half spectra");
  return blue_axis<2, 2, 1>(G, AxisLines{1LL << 62, nh, 0, 1}, G, AxisLines{nn, n, 2 * H, 1}, (long long)nb * nn, n,
                            (int)nh, 1.f, s);
}

}  // namespace gx

using namespace gx;

extern "C" int gtomx_fft_r2c(const float* d_input, gtomx_complex* d_output, int ndims, gtomx_int3 dims,
                             unsigned batch, void* stream) {
  GX_TRY(check_fft_dims_any(ndims, dims));
  if (!d_input || !d_output) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch must be positive");
  return fft_r2c(d_input, reinterpret_cast<float2*>(d_output), dims, batch, as_stream(stream));
}

extern "C" int gtomx_ifft_c2r(const gtomx_complex* d_input, float* d_output, int ndims, gtomx_int3 dims,
                              unsigned batch, int renormalize, void* stream) {
  GX_TRY(check_fft_dims_any(ndims, dims));
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
