// Device-side FFT building blocks for gfx950 (wave64).
//
// A length-N transform is split as N = R1 * R2. A thread holds R1 points
// (stride R2 apart), runs a register DFT-R1, applies the W_N^{j*k1} twiddles,
// exchanges through LDS, and finishes with R1/R2 register DFT-R2s. With R2 = 1
// the whole transform stays in one thread's registers. This is the Stockham
// two-step (four-step) factorisation; no reference source exists for the FFT
// arithmetic (the reference calls closed cuFFT, src/FFT/FFT.cu:17-42).
#pragma once
#include <hip/hip_runtime.h>

namespace gx {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * (DIR * i): DIR = -1 forward (multiply by -i), +1 inverse (multiply by +i)
template <int DIR> __device__ __forceinline__ float2 mul_pj(float2 a) {
  return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// Twiddle tables (one allocation): tw[m] = exp(-2 pi i m / 4096) for
// m in [0, 4096) serves power-of-two lengths; tw[4096 + m] = exp(-2 pi i m / 3072)
// for m in [0, 3072) serves lengths 3 * 2^k. Both are evaluated in extended
// precision on the host and rounded to float.
constexpr int TW_LEN = 4096;
constexpr int TW3_LEN = 3072;
constexpr int TW3_OFF = 4096;
constexpr bool pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
template <int DIR, int N> __device__ __forceinline__ float2 twiddle(const float2* __restrict__ tw, int m) {
  float2 w;
  if constexpr (pow2(N)) {
    static_assert(TW_LEN % N == 0, "twiddle table too short");
    w = tw[(m & (N - 1)) * (TW_LEN / N)];
  } else {
    static_assert(TW3_LEN % N == 0, "length must be 3 * 2^k <= 3072");
    w = tw[TW3_OFF + (m % N) * (TW3_LEN / N)];
  }
  return DIR > 0 ? make_float2(w.x, -w.y) : w;
}

// exp(DIR * 2 pi i m / R) for R dividing 24, from the 24th roots of unity.
template <int DIR, int R> __device__ __forceinline__ float2 root24(int m) {
  static_assert(24 % R == 0, "R must divide 24");
  constexpr float c15 = 0.96592582628906828675f, s15 = 0.25881904510252076235f;
  constexpr float c30 = 0.86602540378443864676f, h = 0.70710678118654752440f;
  constexpr float C[24] = {1.f, c15, c30, h, 0.5f, s15, 0.f, -s15, -0.5f, -h, -c30, -c15,
                           -1.f, -c15, -c30, -h, -0.5f, -s15, 0.f, s15, 0.5f, h, c30, c15};
  constexpr float S[24] = {0.f, s15, 0.5f, h, c30, c15, 1.f, c15, c30, h, 0.5f, s15,
                           0.f, -s15, -0.5f, -h, -c30, -c15, -1.f, -c15, -c30, -h, -0.5f, -s15};
  const int k = (m * (24 / R)) % 24;
  return make_float2(C[k], DIR * S[k]);
}

// ---- register DFTs, natural order in and out, unnormalised ----
template <int DIR> __device__ __forceinline__ void dft2(float2& a0, float2& a1) {
  float2 t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}
template <int DIR> __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_pj<DIR>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}
template <int DIR> __device__ __forceinline__ void dft8(float2* a) {
  float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<DIR>(e0, e1, e2, e3);
  dft4<DIR>(o0, o1, o2, o3);
  constexpr float h = 0.70710678118654752440f;  // sqrt(1/2)
  o1 = cmul(o1, make_float2(h, DIR * h));
  o2 = mul_pj<DIR>(o2);
  o3 = cmul(o3, make_float2(-h, DIR * h));
  a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
  a[1] = cadd(e1, o1); a[5] = csub(e1, o1);
  a[2] = cadd(e2, o2); a[6] = csub(e2, o2);
  a[3] = cadd(e3, o3); a[7] = csub(e3, o3);
}
template <int DIR> __device__ __forceinline__ void dft16(float2* a) {
  float2 e[8], o[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    e[i] = a[2 * i];
    o[i] = a[2 * i + 1];
  }
  dft8<DIR>(e);
  dft8<DIR>(o);
  // W16^k = cos(pi k/8) + DIR i sin(pi k/8)
  constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
  constexpr float h = 0.70710678118654752440f;
  o[1] = cmul(o[1], make_float2(c1, DIR * s1));
  o[2] = cmul(o[2], make_float2(h, DIR * h));
  o[3] = cmul(o[3], make_float2(s1, DIR * c1));
  o[4] = mul_pj<DIR>(o[4]);
  o[5] = cmul(o[5], make_float2(-s1, DIR * c1));
  o[6] = cmul(o[6], make_float2(-h, DIR * h));
  o[7] = cmul(o[7], make_float2(-c1, DIR * s1));
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    a[k] = cadd(e[k], o[k]);
    a[k + 8] = csub(e[k], o[k]);
  }
}
// DFT-32 as two DFT-16 of the even and odd points and one radix-2 stage with W32^k.
template <int DIR> __device__ __forceinline__ void dft32(float2* a) {
  float2 e[16], o[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    e[i] = a[2 * i];
    o[i] = a[2 * i + 1];
  }
  dft16<DIR>(e);
  dft16<DIR>(o);
  // cos(pi k / 16), sin(pi k / 16), k = 0..15
  constexpr float C[16] = {1.f, 0.98078528040323044913f, 0.92387953251128675613f, 0.83146961230254523708f,
                           0.70710678118654752440f, 0.55557023301960222474f, 0.38268343236508977173f,
                           0.19509032201612826785f, 0.f, -0.19509032201612826785f, -0.38268343236508977173f,
                           -0.55557023301960222474f, -0.70710678118654752440f, -0.83146961230254523708f,
                           -0.92387953251128675613f, -0.98078528040323044913f};
  constexpr float S[16] = {0.f, 0.19509032201612826785f, 0.38268343236508977173f, 0.55557023301960222474f,
                           0.70710678118654752440f, 0.83146961230254523708f, 0.92387953251128675613f,
                           0.98078528040323044913f, 1.f, 0.98078528040323044913f, 0.92387953251128675613f,
                           0.83146961230254523708f, 0.70710678118654752440f, 0.55557023301960222474f,
                           0.38268343236508977173f, 0.19509032201612826785f};
#pragma unroll
  for (int k = 1; k < 16; ++k) o[k] = k == 8 ? mul_pj<DIR>(o[k]) : cmul(o[k], make_float2(C[k], DIR * S[k]));
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    a[k] = cadd(e[k], o[k]);
    a[k + 16] = csub(e[k], o[k]);
  }
}
template <int DIR> __device__ __forceinline__ void dft3(float2& a0, float2& a1, float2& a2) {
  constexpr float s3 = 0.86602540378443864676f;  // sin(2 pi / 3)
  const float2 t = cadd(a1, a2);
  const float2 m = make_float2(a0.x - 0.5f * t.x, a0.y - 0.5f * t.y);
  const float2 u = mul_pj<DIR>(cscale(csub(a1, a2), s3));
  a0 = cadd(a0, t);
  a1 = cadd(m, u);
  a2 = csub(m, u);
}

template <int R, int DIR> __device__ __forceinline__ void dft(float2* a);

// Register DFT of length A*B: A-point DFTs over x[b + B*a], twiddles
// W_{AB}^{b*k1}, then B-point DFTs; X[k1 + A*k2].
template <int A, int B, int DIR> __device__ __forceinline__ void dft_mixed(float2* x) {
  float2 y[B][A];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    float2 t[A];
#pragma unroll
    for (int a = 0; a < A; ++a) t[a] = x[b + B * a];
    dft<A, DIR>(t);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) y[b][k1] = (b * k1 == 0) ? t[k1] : cmul(t[k1], root24<DIR, A * B>(b * k1));
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    float2 t[B];
#pragma unroll
    for (int b = 0; b < B; ++b) t[b] = y[b][k1];
    dft<B, DIR>(t);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) x[k1 + A * k2] = t[k2];
  }
}

template <int R, int DIR> __device__ __forceinline__ void dft(float2* a) {
  if constexpr (R == 1) {
  } else if constexpr (R == 2) {
    dft2<DIR>(a[0], a[1]);
  } else if constexpr (R == 3) {
    dft3<DIR>(a[0], a[1], a[2]);
  } else if constexpr (R == 4) {
    dft4<DIR>(a[0], a[1], a[2], a[3]);
  } else if constexpr (R == 6) {
    dft_mixed<2, 3, DIR>(a);
  } else if constexpr (R == 8) {
    dft8<DIR>(a);
  } else if constexpr (R == 12) {
    dft_mixed<4, 3, DIR>(a);
  } else if constexpr (R == 16) {
    dft16<DIR>(a);
  } else if constexpr (R == 24) {
    dft_mixed<8, 3, DIR>(a);
  } else if constexpr (R == 32) {
    dft32<DIR>(a);
  } else {
    static_assert(R == 16, "unsupported radix");
  }
}

// ---- length factorisations ----
// Column transforms (strided axes): TX adjacent columns per workgroup, R2
// threads per column, R1 points per thread.
template <int N> struct ColShape;
template <> struct ColShape<8>   { static constexpr int R1 = 8,  R2 = 1,  TX = 64; };
template <> struct ColShape<12>  { static constexpr int R1 = 12, R2 = 1,  TX = 64; };
template <> struct ColShape<16>  { static constexpr int R1 = 16, R2 = 1,  TX = 64; };
template <> struct ColShape<24>  { static constexpr int R1 = 24, R2 = 1,  TX = 64; };
template <> struct ColShape<32>  { static constexpr int R1 = 8,  R2 = 4,  TX = 16; };
template <> struct ColShape<48>  { static constexpr int R1 = 12, R2 = 4,  TX = 16; };
template <> struct ColShape<64>  { static constexpr int R1 = 8,  R2 = 8,  TX = 16; };
template <> struct ColShape<96>  { static constexpr int R1 = 24, R2 = 4,  TX = 16; };
template <> struct ColShape<128> { static constexpr int R1 = 16, R2 = 8,  TX = 16; };
template <> struct ColShape<192> { static constexpr int R1 = 24, R2 = 8,  TX = 16; };
template <> struct ColShape<256> { static constexpr int R1 = 16, R2 = 16, TX = 16; };
// one-pass columns of the generic transforms at 512 (64 KB of exchange)
template <> struct ColShape<512> { static constexpr int R1 = 32, R2 = 16, TX = 16; };

// Row transforms (contiguous axis, complex length NC): R2 threads per row,
// R1 points per thread.
template <int NC> struct RowShape;
template <> struct RowShape<4>   { static constexpr int R1 = 4,  R2 = 1; };
template <> struct RowShape<6>   { static constexpr int R1 = 6,  R2 = 1; };
template <> struct RowShape<8>   { static constexpr int R1 = 8,  R2 = 1; };
template <> struct RowShape<12>  { static constexpr int R1 = 12, R2 = 1; };
template <> struct RowShape<16>  { static constexpr int R1 = 16, R2 = 1; };
template <> struct RowShape<24>  { static constexpr int R1 = 12, R2 = 2; };
template <> struct RowShape<32>  { static constexpr int R1 = 8,  R2 = 4; };
template <> struct RowShape<48>  { static constexpr int R1 = 12, R2 = 4; };
template <> struct RowShape<64>  { static constexpr int R1 = 8,  R2 = 8; };
template <> struct RowShape<96>  { static constexpr int R1 = 24, R2 = 4; };
template <> struct RowShape<128> { static constexpr int R1 = 16, R2 = 8; };

// Index of the k-th output held by thread j after fft_two_step:
// k = (j*(R1/R2) + u) + R1*k2 for slot u*R2 + k2.
template <int R1, int R2> __device__ __forceinline__ int out_index(int j, int slot) {
  if constexpr (R2 == 1) {
    return slot;
  } else {
    int u = slot / R2, k2 = slot % R2;
    return (j * (R1 / R2) + u) + R1 * k2;
  }
}

// Two-step transform of length N = R1*R2 for one thread of a group of R2
// threads (index j) sharing an LDS exchange area. On entry v[m] holds input
// point j + R2*m; on exit v[slot] holds output point out_index(j, slot).
// xch(k1, jj) returns the LDS address used for the exchange of stage-1
// output k1 of thread jj; it must be private to the group. The caller
// synchronises the workgroup before the call if xch's memory was in use.
// Twiddle source tags for fft_two_step: the global table above, or an LDS
// copy holding exp(-2 pi i m / N) for m in [0, N).
struct TwiddleGlobal {};
struct TwiddleLDS {};
template <int DIR, int N>
__device__ __forceinline__ float2 twiddle_from(const float2* tw, int m, TwiddleGlobal) {
  return twiddle<DIR, N>(tw, m);
}
template <int DIR, int N>
__device__ __forceinline__ float2 twiddle_from(const float2* tw, int m, TwiddleLDS) {
  const float2 w = tw[m % N];
  return DIR > 0 ? make_float2(w.x, -w.y) : w;
}

// Or twiddles the thread holds in registers: w[k1] = exp(DIR 2 pi i j k1 / N)
// for its j (w[0] unused), loaded once by load_two_step_twiddles.
template <int R1> struct TwiddleRegs {
  const float2 (&w)[R1];
};
template <int N, int R1, int DIR>
__device__ __forceinline__ void load_two_step_twiddles(float2 (&w)[R1], int j, const float2* __restrict__ tw) {
  w[0] = make_float2(1.f, 0.f);
#pragma unroll
  for (int k1 = 1; k1 < R1; ++k1) w[k1] = twiddle<DIR, N>(tw, j * k1);
}
template <int DIR, int N, class Src>
__device__ __forceinline__ float2 stage_twiddle(const float2* tw, int j, int k1, Src src) {
  return twiddle_from<DIR, N>(tw, j * k1, src);
}
template <int DIR, int N, int R1>
__device__ __forceinline__ float2 stage_twiddle(const float2*, int, int k1, TwiddleRegs<R1> src) {
  return src.w[k1];
}

template <int N, int R1, int R2, int DIR, class Xch, class Src = TwiddleGlobal>
__device__ __forceinline__ void fft_two_step(float2 (&v)[R1], int j, const float2* __restrict__ tw, Xch xch,
                                             Src src = Src{}) {
  static_assert(R1 * R2 == N, "bad factorisation");
  dft<R1, DIR>(v);
  if constexpr (R2 > 1) {
    static_assert(R1 % R2 == 0, "R1 must be a multiple of R2");
#pragma unroll
    for (int k1 = 1; k1 < R1; ++k1) v[k1] = cmul(v[k1], stage_twiddle<DIR, N>(tw, j, k1, src));
#pragma unroll
    for (int k1 = 0; k1 < R1; ++k1) *xch(k1, j) = v[k1];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < R1 / R2; ++u) {
      float2 z[R2];
      int k1 = j * (R1 / R2) + u;
#pragma unroll
      for (int jj = 0; jj < R2; ++jj) z[jj] = *xch(k1, jj);
      dft<R2, DIR>(z);
#pragma unroll
      for (int k2 = 0; k2 < R2; ++k2) v[u * R2 + k2] = z[k2];
    }
  }
}

// exp(DIR 2 pi i e / L) from the twiddle table, L = 2^a <= 4096 or 3 * 2^a <= 3072.
template <int DIR> __device__ __forceinline__ float2 twiddle_rt(const float2* __restrict__ tw, int L, int e) {
  const float2 w = (L & (L - 1)) == 0 ? tw[(e & (L - 1)) * (TW_LEN / L)] : tw[TW3_OFF + (e % L) * (TW3_LEN / L)];
  return DIR > 0 ? make_float2(w.x, -w.y) : w;
}

// Register split of the passes' line lengths: the column shapes, except 64 =
// 16 x 4, which lets a 256-thread workgroup span 64 lines (512-byte segments).
template <int N> struct LineShape {
  static constexpr int R1 = ColShape<N>::R1, R2 = ColShape<N>::R2;
};
template <> struct LineShape<64> {
  static constexpr int R1 = 16, R2 = 4;
};

// Row selection of the row kernels (and of work on the same rows): by == 0
// takes rows in order;
// otherwise only rows y in [0, by) and [ny - by, ny) of each plane, and (bz >
// 0) only planes z in [0, bz) and [nz - bz, nz) of each volume, counted in
// that order. by must be a multiple of the rows per workgroup.
struct RowBand {
  int ny = 0, by = 0, nz = 1, bz = 0;
};
__device__ __forceinline__ long long band_row(long long v, const RowBand& b) {
  if (b.by == 0) return v;
  const long long py = 2LL * b.by, pz = b.bz ? 2LL * b.bz : b.nz;
  const int yy = (int)(v % py);
  const long long q = v / py;
  const int zz = (int)(q % pz);
  const long long img = q / pz;
  const int y = yy < b.by ? yy : b.ny - (int)py + yy;
  const int z = b.bz == 0 ? zz : (zz < b.bz ? zz : b.nz - (int)pz + zz);
  return (img * b.nz + z) * b.ny + y;
}

// Workgroup sum of one float per thread (all threads must call). Returns the
// total in thread 0. red must hold one float per wave.
__device__ __forceinline__ float block_sum(float v, float* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  float s = 0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) s += red[w];
  }
  __syncthreads();
  return s;
}

}  // namespace gx
