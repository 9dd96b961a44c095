// CPU oracle for GTOM's FFT template-matching path (d_PickSubTomograms).
//
// TEST INFRASTRUCTURE ONLY. This file is a plain restatement of the reference
// algorithm in double precision (the correctness oracle) and, instantiated at
// float, the host-core baseline that is timed beside the GPU path. Only the
// test suite, the smoke check and the CPU-baseline leg of the benchmark load
// it; the product library never links or calls it.
//
// Every function cites the reference file:line it restates. Conventions:
//  * real volumes are x-fastest, index (z*ny + y)*nx + x;
//  * half spectra hold x in [0, nx/2], index (z*ny + y)*(nx/2+1) + x, y and z
//    in FFT order (signed frequency = i <= n/2 ? i : i - n);
//  * R2C is the unnormalised forward DFT (exp(-2 pi i k.r/n));
//  * C2R_u is the unnormalised inverse, defined as the real part of the inverse
//    DFT of the Hermitian-completed spectrum (the imaginary parts of the x=0
//    and x=nx/2 bins are ignored after the y/z transforms); C2R_n = C2R_u / N.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

template <class T> using cpx = std::complex<T>;

// ---------------------------------------------------------------------------
// 1D complex FFT: iterative radix-2 for powers of two, direct DFT otherwise.
// Twiddles are evaluated in long double and rounded to T.
// Replaces the closed cuFFT plans behind src/FFT/FFT.cu:17-42 and
// src/FFT/IFFT.cu:18-56 (no reference source exists for the arithmetic).
// ---------------------------------------------------------------------------
template <class T> struct Fft1D {
  int n = 0;
  bool pow2 = true;
  std::vector<cpx<T>> tw;  // exp(-2 pi i k / n), k in [0, n)
  std::vector<int> rev;

  explicit Fft1D(int n_) : n(n_) {
    pow2 = n > 0 && (n & (n - 1)) == 0;
    tw.resize(std::max(n, 1));
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < n; ++k) {
      long double a = -twopi * (long double)k / (long double)n;
      tw[k] = cpx<T>((T)std::cos(a), (T)std::sin(a));
    }
    if (pow2) {
      rev.resize(n);
      int lg = 0;
      while ((1 << lg) < n) ++lg;
      for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
        rev[i] = r;
      }
    }
  }

  // sign = -1: forward (exp(-i...)), sign = +1: inverse (exp(+i...)), unnormalised.
  void run(cpx<T>* a, int sign, cpx<T>* scratch) const {
    if (n <= 1) return;
    if (!pow2) {
      for (int k = 0; k < n; ++k) {
        cpx<T> acc(0, 0);
        for (int j = 0; j < n; ++j) {
          cpx<T> w = tw[(int)(((long long)j * k) % n)];
          if (sign > 0) w = std::conj(w);
          acc += a[j] * w;
        }
        scratch[k] = acc;
      }
      for (int k = 0; k < n; ++k) a[k] = scratch[k];
      return;
    }
    for (int i = 0; i < n; ++i)
      if (i < rev[i]) std::swap(a[i], a[rev[i]]);
    for (int len = 2; len <= n; len <<= 1) {
      int half = len >> 1, step = n / len;
      for (int i = 0; i < n; i += len) {
        for (int j = 0; j < half; ++j) {
          cpx<T> w = tw[j * step];
          if (sign > 0) w = std::conj(w);
          cpx<T> u = a[i + j], v = a[i + j + half] * w;
          a[i + j] = u + v;
          a[i + j + half] = u - v;
        }
      }
    }
  }
};

inline bool in_parallel() {
#ifdef _OPENMP
  return omp_in_parallel();
#else
  return false;
#endif
}

// Transform every line along one axis of a complex array with dims (dx, dy, dz).
// axis 0: x (contiguous), 1: y (stride dx), 2: z (stride dx*dy).
template <class T>
void fft_axis(cpx<T>* data, int dx, int dy, int dz, int axis, int sign) {
  int len = axis == 0 ? dx : (axis == 1 ? dy : dz);
  if (len <= 1) return;
  Fft1D<T> plan(len);
  long long stride = axis == 0 ? 1 : (axis == 1 ? dx : (long long)dx * dy);
  long long nlines = (long long)dx * dy * dz / len;
#pragma omp parallel if (!in_parallel())
  {
    std::vector<cpx<T>> line(len), scratch(len);
#pragma omp for schedule(static)
    for (long long l = 0; l < nlines; ++l) {
      long long base;
      if (axis == 0) base = l * dx;
      else if (axis == 1) base = (l / dx) * (long long)dx * dy + (l % dx);
      else base = l;  // l indexes (x, y) in [0, dx*dy)
      for (int i = 0; i < len; ++i) line[i] = data[base + i * stride];
      plan.run(line.data(), sign, scratch.data());
      for (int i = 0; i < len; ++i) data[base + i * stride] = line[i];
    }
  }
}

// R2C: unnormalised forward transform of a real (nx, ny, nz) volume into the
// (nx/2+1, ny, nz) half spectrum. Follows cufftExecR2C as called at
// src/FFT/FFT.cu:36-42 (dims reversed to {z,y,x} at src/FFT/FFT.cu:21).
template <class T>
void r2c3(const T* in, cpx<T>* out, int nx, int ny, int nz) {
  int nh = nx / 2 + 1;
  Fft1D<T> plan(nx);
  long long rows = (long long)ny * nz;
#pragma omp parallel if (!in_parallel())
  {
    std::vector<cpx<T>> line(nx), scratch(nx);
#pragma omp for schedule(static)
    for (long long r = 0; r < rows; ++r) {
      for (int x = 0; x < nx; ++x) line[x] = cpx<T>(in[r * nx + x], 0);
      plan.run(line.data(), -1, scratch.data());
      for (int x = 0; x < nh; ++x) out[r * nh + x] = line[x];
    }
  }
  fft_axis(out, nh, ny, nz, 1, -1);
  fft_axis(out, nh, ny, nz, 2, -1);
}

// C2R_u: unnormalised inverse of a half spectrum, real-part semantics (see the
// file header). Follows cufftExecC2R at src/FFT/IFFT.cu:48-56 (no 1/N).
template <class T>
void c2r3(const cpx<T>* in, T* out, int nx, int ny, int nz) {
  int nh = nx / 2 + 1;
  long long H = (long long)nh * ny * nz;
  std::vector<cpx<T>> tmp(in, in + H);
  fft_axis(tmp.data(), nh, ny, nz, 2, +1);
  fft_axis(tmp.data(), nh, ny, nz, 1, +1);
  Fft1D<T> plan(nx);
  long long rows = (long long)ny * nz;
#pragma omp parallel if (!in_parallel())
  {
    std::vector<cpx<T>> line(nx), scratch(nx);
#pragma omp for schedule(static)
    for (long long r = 0; r < rows; ++r) {
      const cpx<T>* h = &tmp[r * nh];
      line[0] = cpx<T>(h[0].real(), 0);
      for (int k = 1; k < nh; ++k) {
        if (2 * k == nx) {
          line[k] = cpx<T>(h[k].real(), 0);
        } else {
          line[k] = h[k];
          line[nx - k] = std::conj(h[k]);
        }
      }
      plan.run(line.data(), +1, scratch.data());
      for (int x = 0; x < nx; ++x) out[r * nx + x] = line[x].real();
    }
  }
}

// ---------------------------------------------------------------------------
// Fourier-space projector: restates Project3DArraytoNDKernel<3>
// (src/Relion/Project.cu:89-177) for one orientation. M is the row-major 3x3
// matrix applied to the integer frequency (x, y, z); the reference builds it as
// transpose(Matrix3Euler(angles)) * Matrix3Scale(oversample)
// (src/Relion/Project.cu:12-17). The reference reads the projector at the
// linear address ((z*p + y)*(p/2+1) + x) without a bounds check; x1 = p/2+1
// therefore reaches into the next row. The same linear address is used here;
// an address outside the array reads 0 (the reference's behaviour there is
// undefined), which changes nothing for in-range reads.
// ---------------------------------------------------------------------------
// nz = n gives the 3D projection (Project3DArraytoNDKernel<3>), nz = 1 the 2D
// central section (Project3DArraytoNDKernel<2>: z = 0, sphere test x^2 + y^2).
// pz is the volume's z extent: p for a 3D volume, 1 for a 2D image's spectrum
// (the Picker's 2D branch, which rotates a 2D reference in plane); reads past
// the volume are 0 (the reference reads them with weight 0 when pz = 1).
template <class T>
void project1(const cpx<T>* vol, int p, cpx<T>* out, int n, const double* M,
              int rmax, int nz = -1, int pz = -1) {
  if (nz < 0) nz = n;
  if (pz < 0) pz = p;
  int ph = p / 2 + 1, nh = n / 2 + 1;
  long long total = (long long)ph * p * pz;
  int rmax2 = rmax * rmax;
  long long H = (long long)nh * n * nz;
  auto at = [&](long long xx, long long yy, long long zz) -> cpx<T> {
    long long i = (zz * p + yy) * ph + xx;
    if (i < 0 || i >= total) return cpx<T>(0, 0);
    return vol[i];
  };
  auto lerp = [](cpx<T> a, cpx<T> b, T t) -> cpx<T> {
    return cpx<T>(std::fma(b.real() - a.real(), t, a.real()),
                  std::fma(b.imag() - a.imag(), t, a.imag()));
  };
#pragma omp parallel for schedule(static) if (!in_parallel())
  for (long long id = 0; id < H; ++id) {
    int ix = (int)(id % nh);
    int iy = (int)((id / nh) % n);
    int iz = (int)(id / ((long long)nh * n));
    int x = ix;
    int y = iy <= n / 2 ? iy : iy - n;
    int z = iz <= n / 2 ? iz : iz - n;
    int r2 = z * z + y * y + x * x;
    if (r2 > rmax2) {
      out[id] = cpx<T>(0, 0);
      continue;
    }
    T px = (T)(M[0] * x + M[1] * y + M[2] * z);
    T py = (T)(M[3] * x + M[4] * y + M[5] * z);
    T pz = (T)(M[6] * x + M[7] * y + M[8] * z);
    T neg = 1;
    if (px < (T)-1e-5) {
      px = std::fabs(px);
      py = -py;
      pz = -pz;
      neg = -1;
    }
    T fx = std::floor(px + (T)1e-5);
    long long x0 = (long long)fx, x1 = x0 + 1;
    T tx = px - fx;
    T fy = std::floor(py);
    long long y0 = (long long)fy, y1 = y0 + 1;
    if (y0 < 0) y0 += p;
    if (y1 < 0) y1 += p;
    T ty = py - fy;
    T fz = std::floor(pz);
    long long z0 = (long long)fz, z1 = z0 + 1;
    if (z0 < 0) z0 += p;
    if (z1 < 0) z1 += p;
    T tz = pz - fz;
    cpx<T> d000 = at(x0, y0, z0), d001 = at(x1, y0, z0);
    cpx<T> d010 = at(x0, y1, z0), d011 = at(x1, y1, z0);
    cpx<T> d100 = at(x0, y0, z1), d101 = at(x1, y0, z1);
    cpx<T> d110 = at(x0, y1, z1), d111 = at(x1, y1, z1);
    cpx<T> dx00 = lerp(d000, d001, tx), dx01 = lerp(d010, d011, tx);
    cpx<T> dx10 = lerp(d100, d101, tx), dx11 = lerp(d110, d111, tx);
    cpx<T> dxy0 = lerp(dx00, dx01, ty), dxy1 = lerp(dx10, dx11, ty);
    cpx<T> val = lerp(dxy0, dxy1, tz);
    out[id] = cpx<T>(val.real(), val.imag() * neg);
  }
}

// Matrix3Euler (src/Helpers/Euler.cu:39-68): ZYZ, RELION Euler_angles2matrix
// layout, returned row-major (A[r][c] = A[3r + c]).
void euler_matrix(double rot, double tilt, double psi, double* A) {
  double ca = std::cos(rot), sa = std::sin(rot);
  double cb = std::cos(tilt), sb = std::sin(tilt);
  double cg = std::cos(psi), sg = std::sin(psi);
  double cc = cb * ca, cs = cb * sa, sc = sb * ca, ss = sb * sa;
  A[0] = cg * cc - sg * sa;  A[1] = cg * cs + sg * ca;  A[2] = -cg * sb;
  A[3] = -sg * cc - cg * sa; A[4] = -sg * cs + cg * ca; A[5] = sg * sb;
  A[6] = sc;                 A[7] = ss;                 A[8] = cb;
}

// The projector matrix transpose(A) * s (src/Relion/Project.cu:16).
void projector_matrix(double rot, double tilt, double psi, double s, double* M) {
  double A[9];
  euler_matrix(rot, tilt, psi, A);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) M[3 * r + c] = s * A[3 * c + r];
}

// Soft sphere mask value, SphereMaskKernel with decentered = false
// (src/Masking/SphereMask.cu:49-106), centre (n/2, n/2, n/2), sigma 1.
double soft_sphere(int x, int y, int z, int nx, int ny, int nz, double radius,
                   double sigma) {
  double xx = x - nx / 2, yy = y - ny / 2, zz = nz > 1 ? z - nz / 2 : 0;
  double len = std::sqrt(xx * xx + yy * yy + zz * zz);
  if (len < radius) return 1.0;
  if (sigma > 0) {
    double v = (std::cos(std::min(1.0, (len - radius) / sigma) * M_PI) + 1.0) * 0.5;
    return std::max(0.0, v);
  }
  return 0.0;
}

// IFFTShift(x, d) = (x + (d+1)/2) % d (include/Prerequisites.cuh:227);
// RemapFull2FullFFTKernel writes out[ifftshift(r)] = in[r]
// (src/FFT/FFTRemap.cu:227-240).
template <class T>
void remap_full2fullfft(const T* in, T* out, int nx, int ny, int nz) {
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        int rx = (x + (nx + 1) / 2) % nx, ry = (y + (ny + 1) / 2) % ny,
            rz = (z + (nz + 1) / 2) % nz;
        out[((long long)rz * ny + ry) * nx + rx] = in[((long long)z * ny + y) * nx + x];
      }
}

// d_LocalStd (src/Generics/LocalStd.cu:10-118), optionally under a Fourier mask:
// sigma = sqrt(max(0, M*conv(map^2) - conv(map)^2)) / M, with conv = C2R_n of
// the product with the conjugate FT of the ifftshifted soft sphere and
// M = sum of the mask. Optional local mean = conv(map) / M.
template <class T>
void local_std(const T* map, int nx, int ny, int nz, double radius, T* std_out,
               T* mean_out, const T* fouriermask = nullptr) {
  long long V = (long long)nx * ny * nz;
  int nh = nx / 2 + 1;
  long long H = (long long)nh * ny * nz;
  std::vector<T> mask(V), maskr(V);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x)
        mask[((long long)z * ny + y) * nx + x] =
            (T)soft_sphere(x, y, z, nx, ny, nz, radius, 1.0);
  remap_full2fullfft(mask.data(), maskr.data(), nx, ny, nz);
  double msum = 0;
  for (long long i = 0; i < V; ++i) msum += maskr[i];
  std::vector<cpx<T>> mft(H), aft(H), bft(H);
  r2c3(maskr.data(), mft.data(), nx, ny, nz);
  if (fouriermask) {
    // Filtered sphere (LocalStd.cu:38-52): spectrum times the mask, sum of the
    // renormalised inverse transform.
    for (long long i = 0; i < H; ++i) mft[i] *= fouriermask[i];
    std::vector<cpx<T>> tmp(mft);
    std::vector<T> conv(V);
    c2r3(tmp.data(), conv.data(), nx, ny, nz);
    msum = 0;
    for (long long i = 0; i < V; ++i) msum += conv[i] / (T)V;
  }
  std::vector<T> sq(V);
  for (long long i = 0; i < V; ++i) sq[i] = map[i] * map[i];
  r2c3(map, aft.data(), nx, ny, nz);
  r2c3(sq.data(), bft.data(), nx, ny, nz);
  for (long long i = 0; i < H; ++i) {
    aft[i] *= std::conj(mft[i]);
    bft[i] *= std::conj(mft[i]);
  }
  std::vector<T> c1(V), c2(V);
  c2r3(aft.data(), c1.data(), nx, ny, nz);
  c2r3(bft.data(), c2.data(), nx, ny, nz);
  T invN = (T)(1.0 / (double)V);
  T M = (T)msum;
  for (long long i = 0; i < V; ++i) {
    T a = c1[i] * invN, b = c2[i] * invN;
    if (mean_out) mean_out[i] = a / M;
    T v = b * M - a * a;
    v = std::max(v, (T)0);
    std_out[i] = std::sqrt(v) / M;
  }
}

// 3D CTF on the half spectrum: CTFParamsLean (include/CTF.cuh:100-134) with
// d_GetCTF<false,false> (include/CTF.cuh:195-223) and the 3D frequency mapping
// k = |r| * ny, ny = 1/(n * pixelsize) (src/CTF/Wiener.cu:99-110).
// Lengths in SI units as in CTFParams (include/CTF.cuh:11-43).
template <class T>
void ctf3(T* out, int n, double pixelsize_m, double voltage_V, double Cs_m,
          double defocus_m, double amplitude, double phaseshift,
          double bfactor_m2, double scale) {
  int nh = n / 2 + 1;
  double ny = 1.0 / ((double)n * (pixelsize_m * 1e10));
  double lambda = 12.2643247 / std::sqrt(voltage_V * (1.0 + voltage_V * 0.978466e-6));
  double defocus = defocus_m * 1e10;
  double Cs = Cs_m * 1e10;
  double K1 = M_PI * lambda;
  double K2 = M_PI_2 * Cs * lambda * lambda * lambda;
  double K3 = std::atan(amplitude / std::sqrt(1 - amplitude * amplitude));
  double B = bfactor_m2 * 0.25e20;
  for (int iz = 0; iz < n; ++iz)
    for (int iy = 0; iy < n; ++iy)
      for (int ix = 0; ix < nh; ++ix) {
        int y = ((iy + n / 2) % n) - n / 2, z = ((iz + n / 2) % n) - n / 2;
        double r = std::sqrt((double)ix * ix + (double)y * y + (double)z * z);
        double k = r * ny, k2 = k * k, k4 = k2 * k2;
        double gamma = K1 * defocus * k2 + K2 * k4 - phaseshift - K3;
        double v = -std::sin(gamma);
        if (B != 0) v *= std::exp(B * k2);
        v *= scale;
        out[((long long)iz * n + iy) * nh + ix] = (T)v;
      }
}

// ---------------------------------------------------------------------------
// Full path: restates d_PickSubTomograms (src/Correlation/SubTomograms.cu:19-125)
// with the device-array projector, for one volume at a time. Per orientation:
//  R = project(P, M_o) (src/Relion/Project.cu:89-177)
//  R *= ctf_v                      (src/Generics/Arithmetics.cu:275-286)
//  r = C2R_u(R); r = (r - mean)/std (src/ImageManipulation/Norm.cu:242-284)
//  Y = R2C(r); Y /= max(sqrt(sum|Y|^2 / H), 1e-20)
//                                  (src/ImageManipulation/Norm.cu:476-509)
//  Q = E_v * conj(Y)               (src/Correlation/SubTomograms.cu:127-136)
//  C = C2R_u(Q)                    (src/Correlation/SubTomograms.cu:81)
//  best = max over o, strict '>' in increasing o (SubTomograms.cu:138-158)
// then sigma = local_std(C2R_u(E_v), maskradius) and
// out = |sigma| > 1e-15 ? best / sigma : 0 (src/Generics/Arithmetics.cu:478-493).
// The oracle also keeps the second-best CC per voxel.
// ---------------------------------------------------------------------------
// nz = n: volumes of n^3; nz = 1: images of n^2 (the reference's ndims == 2 branch,
// src/Correlation/SubTomograms.cu:33-36, same sequence with 2D transforms).
template <class T>
double pick(const T* projft, int p, double oversample, const T* expft,
            const T* ctf, int ctf_per_volume, int n, int nvol,
            const double* angles, int nangles, double maskradius, T* best_out,
            int32_t* idx_out, T* top1_out, T* top2_out, T* sigma_out, int nz = -1) {
  if (nz < 0) nz = n;
  const cpx<T>* P = reinterpret_cast<const cpx<T>*>(projft);
  const cpx<T>* E = reinterpret_cast<const cpx<T>*>(expft);
  int nh = n / 2 + 1;
  long long V = (long long)n * n * nz, H = (long long)nh * n * nz;
  std::vector<cpx<T>> R(H), Y(H), Q(H);
  std::vector<T> r(V), C(V);
  std::vector<T> top1(V * nvol, (T)-1e30), top2(V * nvol, (T)-1e30);
  std::vector<int32_t> idx(V * nvol, 0);
  double loop_seconds = 0;
  for (int v = 0; v < nvol; ++v) {
    const T* c = ctf + (ctf_per_volume ? (long long)v * H : 0);
    const cpx<T>* Ev = E + (long long)v * H;
    T* t1 = &top1[v * V];
    T* t2 = &top2[v * V];
    int32_t* ix = &idx[v * V];
    auto t0 = std::chrono::steady_clock::now();
    for (int o = 0; o < nangles; ++o) {
      double M[9];
      projector_matrix(angles[3 * o], angles[3 * o + 1], angles[3 * o + 2], oversample, M);
      project1(P, p, R.data(), n, M, n / 2, nz);
      for (long long i = 0; i < H; ++i) R[i] *= c[i];
      c2r3(R.data(), r.data(), n, n, nz);
      double s1 = 0, s2 = 0;
      for (long long i = 0; i < V; ++i) {
        double val = r[i];
        s1 += val;
        s2 += val * val;
      }
      double mean = s1 / (double)V;
      double sd = std::sqrt((double)V * s2 - s1 * s1) / (double)V;
      T tmean = (T)mean, tinv = (T)(sd > 0 ? 1.0 / sd : 0.0);
      for (long long i = 0; i < V; ++i) r[i] = (r[i] - tmean) * tinv;
      r2c3(r.data(), Y.data(), n, n, nz);
      double ss = 0;
      for (long long i = 0; i < H; ++i) ss += (double)std::norm(Y[i]);
      double rms = std::sqrt(ss / (double)H);
      T inv = (T)(1.0 / std::max(rms, 1e-20));
      for (long long i = 0; i < H; ++i) {
        Y[i] *= inv;
        Q[i] = Ev[i] * std::conj(Y[i]);
      }
      c2r3(Q.data(), C.data(), n, n, nz);
      for (long long i = 0; i < V; ++i) {
        T cc = C[i];
        if (cc > t1[i]) {
          t2[i] = t1[i];
          t1[i] = cc;
          ix[i] = o;
        } else if (cc > t2[i]) {
          t2[i] = cc;
        }
      }
    }
    loop_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // Local-std normalisation (SubTomograms.cu:103-118)
    std::vector<T> e(V), sig(V);
    c2r3(Ev, e.data(), n, n, nz);
    local_std(e.data(), n, n, nz, maskradius, sig.data(), (T*)nullptr);
    for (long long i = 0; i < V; ++i) {
      T s = sig[i];
      T out = std::fabs(s) > (T)1e-15 ? t1[i] / s : (T)0;
      if (best_out) best_out[v * V + i] = out;
      if (sigma_out) sigma_out[v * V + i] = s;
    }
  }
  if (idx_out) std::memcpy(idx_out, idx.data(), sizeof(int32_t) * V * nvol);
  if (top1_out) std::memcpy(top1_out, top1.data(), sizeof(T) * V * nvol);
  if (top2_out) std::memcpy(top2_out, top2.data(), sizeof(T) * V * nvol);
  return loop_seconds;
}

// d_FFTFullPad on real volumes (src/FFT/FFTResize.cu:92-110, FFTFullPadEvenKernel
// :212-238): both volumes in FFT order (origin at index 0). Output element x
// (per axis) has centred coordinate nr = (x + new/2) % new, which maps to
// or = nr + (old - new - ((old & (1 - (new & 1))) % 2)) / 2 in the input's centred
// coordinates (the kernel's `old & 1 - (new & 1)` parses as old & (1 - (new & 1)));
// inside [0, old) the value is read at (or + (old+1)/2) % old, outside it is 0.
template <class T>
void full_pad(const T* in, T* out, int ox, int oy, int oz, int nx, int ny, int nz) {
  auto shift = [](int o, int n) { return (o - n - ((o & (1 - (n & 1))) % 2)) / 2; };
  const int sx = shift(ox, nx), sy = shift(oy, ny), sz = shift(oz, nz);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const int rx = (x + nx / 2) % nx + sx, ry = (y + ny / 2) % ny + sy, rz = (z + nz / 2) % nz + sz;
        T v = 0;
        if (rx >= 0 && rx < ox && ry >= 0 && ry < oy && rz >= 0 && rz < oz) {
          const int ix = (rx + (ox + 1) / 2) % ox, iy = (ry + (oy + 1) / 2) % oy, iz = (rz + (oz + 1) / 2) % oz;
          v = in[((long long)iz * oy + iy) * ox + ix];
        }
        out[((long long)z * ny + y) * nx + x] = v;
      }
}

// d_Norm with T_NORM_MEAN01STD and no mask (src/ImageManipulation/Norm.cu:36-70,
// NormMeanStdDevKernel :192-212, statistics from d_Dev src/Generics/Dev.cu:13-72):
// mean = sum/N, variance = sum((x - mean)^2)/N (population), out = (x - mean)/std,
// or x - mean when std = 0. Accumulated in double.
template <class T> void norm_mean01std(const T* in, T* out, long long n) {
  double s = 0;
  for (long long i = 0; i < n; ++i) s += in[i];
  const double mean = s / (double)n;
  double v = 0;
  for (long long i = 0; i < n; ++i) v += ((double)in[i] - mean) * ((double)in[i] - mean);
  const double sd = std::sqrt(v / (double)n);
  for (long long i = 0; i < n; ++i) out[i] = (T)(sd == 0 ? (double)in[i] - mean : ((double)in[i] - mean) / sd);
}

// The Picker class (include/Correlation.cuh:136-164, src/Correlation/Picker.cu:25-276)
// as its code intends, with the rotation its author left commented out
// (src/Correlation/Picker.cu:181-187) done by the device-array projector:
//  Initialize: refft = R2C(ifftshift(ref)) at the reference size (:92-94, :112-114);
//  SetImage: image -> mean0/std1 (:148), imageft = R2C(image) (:152);
//  PerformCorrelation, for every orientation (:172-245):
//    R = project(refft, transpose(Euler(angle)), rmax = nref/2) (:176, :185),
//    R *= ctf (:189), r = C2R_n(R) (:190), padded = FullPad(r) (:225),
//    padded -> mean0/std1 (:226), P = R2C(padded) (:227),
//    C = C2R_u(imageft * conj(P)) (:228-229), best/angle updated where
//    C > best (PickingUpdateKernel :249-276).
// One reference of side nr (a cube, or with nrz = 1 a square image: the 2D
// branch, whose orientations are in-plane steps, src/Correlation/Picker.cu:163-171,
// and whose rotation is the 2D section of the rotated 2D spectrum), one image
// nx x ny x nz (nz = 1 with a 2D reference), ctf of (nr/2+1)*nr*nrz.
// best/idx/top2 are updated in place (the caller fills them), idx receives the
// index into `angles`.
template <class T>
void picker(const T* ref, int nr, int nrz, const T* image, int nx, int ny, int nz, const T* ctf,
            const double* angles, int nangles, T* best, int32_t* idx, T* top2) {
  const int nrh = nr / 2 + 1;
  const long long Vr = (long long)nr * nr * nrz, Hr = (long long)nrh * nr * nrz;
  const long long V = (long long)nx * ny * nz, H = (long long)(nx / 2 + 1) * ny * nz;
  std::vector<T> shifted(Vr), r(Vr), padded(V), C(V), img(V);
  std::vector<cpx<T>> refft(Hr), R(Hr), imageft(H), P(H);
  remap_full2fullfft(ref, shifted.data(), nr, nr, nrz);
  r2c3(shifted.data(), refft.data(), nr, nr, nrz);
  norm_mean01std(image, img.data(), V);
  r2c3(img.data(), imageft.data(), nx, ny, nz);
  for (int o = 0; o < nangles; ++o) {
    double M[9];
    projector_matrix(angles[3 * o], angles[3 * o + 1], angles[3 * o + 2], 1.0, M);
    project1(refft.data(), nr, R.data(), nr, M, nr / 2, nrz, nrz);
    for (long long i = 0; i < Hr; ++i) R[i] *= ctf[i];
    c2r3(R.data(), r.data(), nr, nr, nrz);
    for (long long i = 0; i < Vr; ++i) r[i] /= (T)Vr;
    full_pad(r.data(), padded.data(), nr, nr, nrz, nx, ny, nz);
    norm_mean01std(padded.data(), padded.data(), V);
    r2c3(padded.data(), P.data(), nx, ny, nz);
    for (long long i = 0; i < H; ++i) P[i] = imageft[i] * std::conj(P[i]);
    c2r3(P.data(), C.data(), nx, ny, nz);
    for (long long i = 0; i < V; ++i) {
      if (C[i] > best[i]) {
        if (top2) top2[i] = best[i];
        best[i] = C[i];
        idx[i] = o;
      } else if (top2 && C[i] > top2[i]) {
        top2[i] = C[i];
      }
    }
  }
}

void fft_resize(int op, const double* in, double* out, const int* od, const int* nd, int ncomp);

// ---------------------------------------------------------------------------
// TomoPicker (include/Correlation.cuh:209-257, src/Correlation/TomoPicker.cu)
// as its code describes, with the projections its author left broken done by
// the device-array projector (src/Relion/Project.cu:89-177, 2D sections):
// ---------------------------------------------------------------------------
// d_Exact2DWeighting (src/Projection/Weighting.cu:21-55, kernel :92-133) for
// images in FFT order (iszerocentered false), maxfreq in pixels: filter b
// (n/2+1 x n half spectrum, multiplied in place) at frequency f is
// w_b^2 / sum_k |sinc(dot(A, normal_k))| w_k, A = the frequency -f of the
// kernel's loop written in the frame of image b (rows 0 and 1 of
// Matrix3Euler(angle_b) as the local axes), normal_k = row 2 of
// Matrix3Euler(angle_k); 0 where |f| > maxfreq. As in the reference, batch item
// b takes the frame (h_globalB2localB[b], :40) and weight (currentimage = b) of
// image b; indices[b] only selects a normal the kernel never uses. sinc(x) =
// sin(pi x)/(pi x), 1 for |x| <= 1e-8 (include/DeviceFunctions.cuh:19-25).
void exact2d_weighting(double* w, int nx, int ny, const double* angles, const double* weights, int nimages,
                       const int* indices, int batch, double maxfreq) {
  std::vector<double> normals(3 * (size_t)nimages), B(9 * (size_t)nimages);
  for (int i = 0; i < nimages; ++i) {
    euler_matrix(angles[3 * i], angles[3 * i + 1], angles[3 * i + 2], &B[9 * (size_t)i]);
    for (int c = 0; c < 3; ++c) normals[3 * (size_t)i + c] = B[9 * (size_t)i + 6 + c];
  }
  const int nh = nx / 2 + 1;
  for (int b = 0; b < batch; ++b) {
    (void)indices;
    const double* E = &B[9 * (size_t)b];
    double* out = w + (size_t)b * nh * ny;
    for (int idy = 0; idy < ny; ++idy) {
      const int y = ny - 1 - ((idy + ny / 2 - 1) % ny);
      for (int idx = 0; idx < nh; ++idx) {
        const int x = nx / 2 - idx;
        const double ax = (double)idx - nx / 2, ay = (double)idy - ny / 2;
        double& dst = out[(size_t)y * nh + x];
        if (std::sqrt(ax * ax + ay * ay) <= maxfreq) {
          // localA * globalB2localB: the vector ax * row0 + ay * row1 of the image's Euler matrix
          const double g[3] = {ax * E[0] + ay * E[3], ax * E[1] + ay * E[4], ax * E[2] + ay * E[5]};
          double sum = 0;
          for (int k = 0; k < nimages; ++k) {
            const double d = g[0] * normals[3 * k] + g[1] * normals[3 * k + 1] + g[2] * normals[3 * k + 2];
            const double sn = std::fabs(d) <= 1e-8 ? 1.0 : std::sin(d * M_PI) / (d * M_PI);
            sum += std::fabs(sn) * weights[k];
          }
          dst *= weights[b] * weights[b] / sum;
        } else {
          dst *= 0.0;
        }
      }
    }
  }
}

// The whole TomoPicker in double. ref and mask: nr^3 (centred); images: nimages
// x ny x nx, filtered in place as SetImage leaves them; tilt: nimages (rot,
// tilt, psi); weights: nimages; angles: the search orientations. corr (dimscorr,
// caller-initialised) keeps the maximum over orientations of the weighted mean
// of the tilt images' correlation maps sampled at each voxel's projection,
// with strict '<' (UpdateTomoCorrKernel, :311-358), corridx the orientation
// index, corrsum1/2 the sums of that mean and its square over orientations,
// corr2 (optional) the second-best value.
// The correlation of one tilt image with one projected reference (the call the
// reference marks BROKEN, :297) is the masked, locally normalised correlation
// of d_CorrelateRealspace (src/Correlation/Realspace.cu:16-102) at every image
// position x:  sum_m mask(m) ref(m) img(x + m) / (M sigma(x)), with ref
// normalised to mean 0, std 1 under the binarised mask (d_NormMonolithic with a
// mask, src/ImageManipulation/Norm.cu:286-333), M = sum mask, sigma(x) =
// sqrt(max(0, M s2 - s1^2)) / M from the local sums s1, s2 of img and img^2
// under the mask (:255-263), and 0 where sigma is not positive.
void tomo_picker(const double* ref, const double* mask, int nr, bool circular, double* images, int nx, int ny,
                 int nimages, const double* tilt, const double* weights, const double* angles, int nangles,
                 int cx, int cy, int cz, double* corr, int32_t* corridx, double* corrsum1, double* corrsum2,
                 double* corr2) {
  const int p = 2 * nr, ph = p / 2 + 1;
  const long long Vp3 = (long long)p * p * p, Hp3 = (long long)ph * p * p;
  const long long Vp = (long long)p * p, Hp = (long long)ph * p, Vr = (long long)nr * nr;
  const long long Vi = (long long)nx * ny, Hi = (long long)(nx / 2 + 1) * ny;
  // Initialize: pad (centre p/2 <- nr/2), ifftshift, R2C (TomoPicker.cu:80-103)
  auto prepare = [&](const double* v, std::vector<cpx<double>>& ft) {
    std::vector<double> padded(Vp3, 0.0), shifted(Vp3);
    const int off = p / 2 - nr / 2;
    for (int z = 0; z < nr; ++z)
      for (int y = 0; y < nr; ++y)
        for (int x = 0; x < nr; ++x)
          padded[((size_t)(z + off) * p + y + off) * p + x + off] = v[((size_t)z * nr + y) * nr + x];
    remap_full2fullfft(padded.data(), shifted.data(), p, p, p);
    ft.resize(Hp3);
    r2c3(shifted.data(), ft.data(), p, p, p);
  };
  std::vector<cpx<double>> Pref, Pmask;
  prepare(ref, Pref);
  prepare(mask, Pmask);
  // SetImage (:158-186): filter, filtered images, their spectra and the spectra of their squares
  std::vector<double> filter(Hi * nimages, 1.0);
  std::vector<int> ind(nimages);
  for (int i = 0; i < nimages; ++i) ind[i] = i;
  exact2d_weighting(filter.data(), nx, ny, tilt, weights, nimages, ind.data(), nimages, (double)nx);
  std::vector<cpx<double>> imageft(Hi * nimages), image2ft(Hi * nimages);
  std::vector<double> sq(Vi);
  for (int i = 0; i < nimages; ++i) {
    cpx<double>* F = &imageft[Hi * i];
    double* im = images + Vi * i;
    r2c3(im, F, nx, ny, 1);
    for (long long k = 0; k < Hi; ++k) F[k] *= filter[Hi * i + k];
    c2r3(F, im, nx, ny, 1);
    for (long long k = 0; k < Vi; ++k) im[k] /= (double)Vi;
    r2c3(im, F, nx, ny, 1);  // the spectrum of the image as stored (the product, Hermitian-projected)
    for (long long k = 0; k < Vi; ++k) sq[k] = im[k] * im[k];
    r2c3(sq.data(), &image2ft[Hi * i], nx, ny, 1);
  }
  // PerformCorrelation (:188-305)
  std::vector<double> mask2d(Vr * nimages), masksum(nimages), s1(Vi * nimages), s2(Vi * nimages);
  std::vector<double> imagecorr(Vi * nimages);
  std::vector<cpx<double>> sec(Hp), tmpft(Hi);
  std::vector<double> full(Vp), padded(Vi), num(Vi);
  const int od[3] = {p, p, 1}, nd[3] = {nr, nr, 1}, id3[3] = {nr, nr, 1}, pd3[3] = {nx, ny, 1};
  std::vector<double> M(9 * (size_t)nimages), T(6 * (size_t)nimages);
  for (int i = 0; i < nimages; ++i) {
    double A[9];
    euler_matrix(tilt[3 * i], tilt[3 * i + 1], tilt[3 * i + 2], A);
    for (int c = 0; c < 3; ++c) {
      T[6 * (size_t)i + c] = A[c];          // row 0
      T[6 * (size_t)i + 3 + c] = A[3 + c];  // row 1
    }
  }
  for (long long v = 0; v < (long long)cx * cy * cz; ++v) corrsum1[v] = corrsum2[v] = 0.0;
  for (int a = 0; a < nangles; ++a) {
    double G[9];
    euler_matrix(angles[3 * a], angles[3 * a + 1], angles[3 * a + 2], G);
    for (int i = 0; i < nimages; ++i) {
      double B[9];
      euler_matrix(tilt[3 * i], tilt[3 * i + 1], tilt[3 * i + 2], B);
      // transpose(Euler(angle)) * transpose(Euler(tilt_i))
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double acc = 0;
          for (int k = 0; k < 3; ++k) acc += G[3 * k + r] * B[3 * c + k];
          M[9 * (size_t)i + 3 * r + c] = acc;
        }
    }
    if (a == 0 || !circular) {
      for (int i = 0; i < nimages; ++i) {
        project1(Pmask.data(), p, sec.data(), p, &M[9 * (size_t)i], p / 2, 1);
        c2r3(sec.data(), full.data(), p, p, 1);
        for (long long k = 0; k < Vp; ++k) full[k] /= (double)Vp;
        double* m2 = &mask2d[Vr * i];
        fft_resize(3, full.data(), m2, od, nd, 1);
        double ms = 0;
        for (long long k = 0; k < Vr; ++k) {
          m2[k] = m2[k] >= 0.99 ? 1.0 : 0.0;  // d_Binarize (src/BinaryManipulation/Binarize.cu:30-36)
          ms += m2[k];
        }
        masksum[i] = ms;
        fft_resize(4, m2, padded.data(), id3, pd3, 1);
        r2c3(padded.data(), tmpft.data(), nx, ny, 1);
        std::vector<cpx<double>> prod(Hi);
        for (long long k = 0; k < Hi; ++k) prod[k] = imageft[Hi * i + k] * std::conj(tmpft[k]);
        c2r3(prod.data(), &s1[Vi * i], nx, ny, 1);
        for (long long k = 0; k < Hi; ++k) prod[k] = image2ft[Hi * i + k] * std::conj(tmpft[k]);
        c2r3(prod.data(), &s2[Vi * i], nx, ny, 1);
        for (long long k = 0; k < Vi; ++k) {
          s1[Vi * i + k] /= (double)Vi;
          s2[Vi * i + k] /= (double)Vi;
        }
      }
    }
    for (int i = 0; i < nimages; ++i) {
      project1(Pref.data(), p, sec.data(), p, &M[9 * (size_t)i], p / 2, 1);
      c2r3(sec.data(), full.data(), p, p, 1);  // not renormalised (the 3-argument d_IFFTC2R, :272)
      std::vector<double> r(Vr);
      fft_resize(3, full.data(), r.data(), od, nd, 1);
      const double* m2 = &mask2d[Vr * i];
      double a1 = 0, a2 = 0, sm = 0;
      for (long long k = 0; k < Vr; ++k) {
        a1 += r[k] * m2[k];
        a2 += r[k] * r[k] * m2[k];
        sm += m2[k];
      }
      const double mean = a1 / sm, sd = std::sqrt(sm * a2 - a1 * a1) / sm;
      const double isd = sd > 0 ? 1.0 / sd : 0.0;
      for (long long k = 0; k < Vr; ++k) r[k] = (r[k] - mean) * isd * m2[k];
      fft_resize(4, r.data(), padded.data(), id3, pd3, 1);
      r2c3(padded.data(), tmpft.data(), nx, ny, 1);
      std::vector<cpx<double>> prod(Hi);
      for (long long k = 0; k < Hi; ++k) prod[k] = imageft[Hi * i + k] * std::conj(tmpft[k]);
      c2r3(prod.data(), num.data(), nx, ny, 1);
      const double Mi = masksum[i];
      for (long long k = 0; k < Vi; ++k) {
        const double a = s1[Vi * i + k], b = s2[Vi * i + k];
        const double sig = std::sqrt(std::max(0.0, Mi * b - a * a)) / Mi;
        imagecorr[Vi * i + k] = sig > 0 ? num[k] / (double)Vi / (Mi * sig) : 0.0;
      }
    }
    // back-projection (UpdateTomoCorrKernel, :311-358): bilinear interpolation of
    // the linear-filtered texture with clamped addressing, texel centres at +0.5
    for (int z = 0; z < cz; ++z)
      for (int y = 0; y < cy; ++y)
        for (int x = 0; x < cx; ++x) {
          const double pos[3] = {(double)x - cx / 2, (double)y - cy / 2, (double)z - cz / 2};
          double sum = 0, samples = 0;
          for (int n = 0; n < nimages; ++n) {
            const double* Tn = &T[6 * (size_t)n];
            const double px = Tn[0] * pos[0] + Tn[1] * pos[1] + Tn[2] * pos[2] + (nx / 2 + 0.5);
            const double py = Tn[3] * pos[0] + Tn[4] * pos[1] + Tn[5] * pos[2] + (ny / 2 + 0.5);
            if (px >= 0 && px <= nx && py >= 0 && py <= ny) {
              const double u = px - 0.5, v = py - 0.5;
              const double fu = std::floor(u), fv = std::floor(v);
              const double tu = u - fu, tv = v - fv;
              auto at = [&](long long xi, long long yi) {
                xi = std::min<long long>(std::max<long long>(xi, 0), nx - 1);
                yi = std::min<long long>(std::max<long long>(yi, 0), ny - 1);
                return imagecorr[Vi * n + yi * nx + xi];
              };
              const long long x0 = (long long)fu, y0 = (long long)fv;
              const double val = (1 - tv) * ((1 - tu) * at(x0, y0) + tu * at(x0 + 1, y0)) +
                                 tv * ((1 - tu) * at(x0, y0 + 1) + tu * at(x0 + 1, y0 + 1));
              sum += val * weights[n];
              samples += weights[n];
            }
          }
          if (samples > 0) {
            sum /= samples;
            const size_t o = ((size_t)z * cy + y) * cx + x;
            if (corr[o] < sum) {
              if (corr2) corr2[o] = corr[o];
              corr[o] = sum;
              corridx[o] = a;
            } else if (corr2 && corr2[o] < sum) {
              corr2[o] = sum;
            }
            corrsum1[o] += sum;
            corrsum2[o] += sum * sum;
          }
        }
  }
}

// d_PickSubTomogramsDiff2 (src/Correlation/SubTomograms.cu:165-209, kernel
// PickSubTomogramsDiff2Kernel :211-313), literally in double: for every voxel
// (ix, iy, iz) of dimsrelevant the shift
//   s = -((ix - rx/2)/nx, (iy - ry/2)/ny, (iz - rz/2)/nz) * 2 pi   (rx/2 etc. integer)
// and for every orientation a the score
//   sum over all half-spectrum bins k of |R_a(k) ctf(k) exp(i k.s) - E(k)|^2,
// with k = (x, signed y, signed z) (:248-250); best = the smallest score with
// strict '<' in increasing orientation (:285-299, :303-311), started at 1e30.
// One cubic volume per call (the reference loops volumes with its own CTF).
void diff2(const cpx<double>* P, int p, double oversample, const cpx<double>* E, const double* ctf, int n,
           int rx, int ry, int rz, const double* angles, int nangles, double* best, int32_t* idx,
           double* second) {
  const int nh = n / 2 + 1;
  const long long H = (long long)nh * n * n;
  const long long Vr = (long long)rx * ry * rz;
  std::vector<cpx<double>> R(H);
  for (long long i = 0; i < Vr; ++i) {
    best[i] = 1e30;
    idx[i] = 0;
    if (second) second[i] = 1e30;
  }
  for (int a = 0; a < nangles; ++a) {
    double M[9];
    projector_matrix(angles[3 * a], angles[3 * a + 1], angles[3 * a + 2], oversample, M);
    project1(P, p, R.data(), n, M, n / 2);
    for (long long i = 0; i < H; ++i) R[i] *= ctf[i];
#pragma omp parallel for schedule(static)
    for (long long v = 0; v < Vr; ++v) {
      const int ix = (int)(v % rx), iy = (int)((v / rx) % ry), iz = (int)(v / ((long long)rx * ry));
      const double two_pi = 6.283185307179586476925286766559;
      const double sx = -((double)ix - rx / 2) / n * two_pi, sy = -((double)iy - ry / 2) / n * two_pi,
                   sz = -((double)iz - rz / 2) / n * two_pi;
      double score = 0;
      for (long long k = 0; k < H; ++k) {
        const int kx = (int)(k % nh), jy = (int)((k / nh) % n), jz = (int)(k / ((long long)nh * n));
        const double y = jy > n / 2 ? jy - n : jy, z = jz > n / 2 ? jz - n : jz;
        const double ph = kx * sx + y * sy + z * sz;
        const cpx<double> d = R[k] * cpx<double>(std::cos(ph), std::sin(ph)) - E[k];
        score += std::norm(d);
      }
      if (score < best[v]) {
        if (second) second[v] = best[v];
        best[v] = score;
        idx[v] = a;
      } else if (second && score < second[v]) {
        second[v] = score;
      }
    }
  }
}

// CTFParams (include/CTF.cuh:11-43) in SI units, as the reference stores them.
struct CtfFull {
  double pixelsize, pixeldelta, pixelangle, Cs, voltage, defocus, astigmatismangle, defocusdelta, amplitude,
      Bfactor, Bfactordelta, Bfactorangle, scale, phaseshift;
};

// CTFSimulateKernel (src/CTF/Simulate.cu:117-139) with d_GetCTF
// (include/CTF.cuh:195-223) and CTFParamsLean's unit conversions
// (include/CTF.cuh:100-134), in double: address (k, angle) with k in cycles per
// pixel; pixel size pixelsize + pixeldelta/2 cos 2(angle - pixelangle) [A];
// defocus + defocusdelta/2 cos 2(angle - astigmatismangle) [A];
// gamma = K1 deltaf k^2 + K2 k^4 - phaseshift - K3 + gammacorrection;
// -sin(gamma) (1 when ignorefirstpeak and |gamma| < pi/2), times
// exp((B + dB cos 2(angle - Bangle)) k^2) with B in A^2/4, |.| when
// amplitudesquared, times scale.
double ctf_at(const CtfFull& q, double k, double angle, double gammacorr, bool ampsq, bool ignorefirst) {
  const double pi = 3.14159265358979323846;
  const double lambda = 12.2643247 / std::sqrt(q.voltage * (1.0 + q.voltage * 0.978466e-6));
  const double px = q.pixelsize * 1e10 + q.pixeldelta * 0.5e10 * std::cos(2.0 * (angle - q.pixelangle));
  k /= px;
  const double K1 = pi * lambda, K2 = pi / 2 * (q.Cs * 1e10) * lambda * lambda * lambda;
  const double K3 = std::atan(q.amplitude / std::sqrt(1.0 - q.amplitude * q.amplitude));
  const double k2 = k * k, k4 = k2 * k2;
  const double deltaf = q.defocus * 1e10 + q.defocusdelta * 0.5e10 * std::cos(2.0 * (angle - q.astigmatismangle));
  const double gamma = K1 * deltaf * k2 + K2 * k4 - q.phaseshift - K3 + gammacorr;
  double v = (ignorefirst && std::fabs(gamma) < pi / 2) ? 1.0 : -std::sin(gamma);
  if (q.Bfactor != 0 || q.Bfactordelta != 0) {
    double b = q.Bfactor * 0.25e20;
    if (q.Bfactordelta != 0) b += q.Bfactordelta * 0.25e20 * std::cos(2.0 * (angle - q.Bfactorangle));
    v *= std::exp(b * k2);
  }
  if (ampsq) v = std::fabs(v);
  return v * q.scale;
}

// SphereMaskKernel in full (src/Masking/SphereMask.cu:49-106): mask values of a
// (nx, ny, nz) volume with float centre c (the difference i - c truncated to int)
// or, decentered, FFT-order distances from the origin; z only when nz > 1.
void sphere_mask_ex(double* out, int nx, int ny, int nz, double radius, double sigma, const float* c,
                    bool decentered) {
  const double pi = 3.14159265358979323846;
  auto axis = [&](int v, int n, float cv) { return decentered ? (v < n / 2 + 1 ? v : v - n) : (int)((float)v - cv); };
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const long long xx = axis(x, nx, c[0]), yy = axis(y, ny, c[1]), zz = nz > 1 ? axis(z, nz, c[2]) : 0;
        const double len = std::sqrt((double)(xx * xx + yy * yy + zz * zz));
        double m = 1.0;
        if (!(len < radius))
          m = sigma > 0 ? std::max(0.0, (std::cos(std::min(1.0, (len - radius) / sigma) * pi) + 1.0) * 0.5) : 0.0;
        out[((size_t)z * ny + y) * nx + x] = m;
      }
}

// Fourier-space resizing (src/FFT/FFTResize.cu), one array of ncomp doubles per
// element: op 1 FFTCropKernel (:121-138), 2 FFTPadEvenKernel (:162-205), 3
// FFTFullCropKernel (:140-160), 4 FFTFullPadEvenKernel (:212-238). Half
// spectra (ops 1, 2) have (n/2+1) x n x n elements, full arrays n^3.
void fft_resize(int op, const double* in, double* out, const int* od, const int* nd, int ncomp) {
  const bool half = op <= 2;
  const int oxs = half ? od[0] / 2 + 1 : od[0], nxs = half ? nd[0] / 2 + 1 : nd[0];
  for (int z = 0; z < nd[2]; ++z)
    for (int y = 0; y < nd[1]; ++y)
      for (int x = 0; x < nxs; ++x) {
        const int p[3] = {x, y, z};
        int q[3];
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
          const int o = od[a], n = nd[a], i = p[a];
          if (op == 1) {
            q[a] = (a == 0) ? i : (i < n / 2 + 1 ? i : i - n + o);
          } else if (op == 3) {
            q[a] = i >= n / 2 ? i + o - n : i;
          } else if (op == 2 && a == 0) {
            ok = ok && i < (o + 1) / 2;
            q[a] = i;
          } else {
            const int shift = op == 2 ? (o - n) / 2 : (o - n - ((o & (1 - (n & 1))) % 2)) / 2;
            const int c = (i + n / 2) % n + shift;
            ok = ok && c >= 0 && c < o;
            q[a] = (c + (o + 1) / 2) % o;
          }
        }
        double* dst = out + (((size_t)z * nd[1] + y) * nxs + x) * ncomp;
        for (int k = 0; k < ncomp; ++k)
          dst[k] = ok ? in[(((size_t)q[2] * od[1] + q[1]) * oxs + q[0]) * ncomp + k] : 0.0;
      }
}

// d_NormMonolithic (src/ImageManipulation/Norm.cu:76-91, NormMeanStdDevMonoKernel
// :242-284): (x - mean) / std, std = sqrt(N sum x^2 - (sum x)^2) / N in double, 0
// where std is not positive. In double throughout (the reference rounds mean
// and 1/std to float).
void norm_monolithic(const float* in, double* out, long long n) {
  double s1 = 0, s2 = 0;
  for (long long i = 0; i < n; ++i) {
    s1 += in[i];
    s2 += (double)in[i] * in[i];
  }
  const double mean = s1 / n, sd = std::sqrt(n * s2 - s1 * s1) / n;
  const double inv = sd > 0 ? 1.0 / sd : 0.0;
  for (long long i = 0; i < n; ++i) out[i] = (in[i] - mean) * inv;
}

// d_NormFTMonolithic (:147-154, NormFTMonoKernel :476-509): x / max(sqrt(sum |x|^2 / N), 1e-20).
void norm_ft_monolithic(const float* in, double* out, long long n) {
  double s = 0;
  for (long long i = 0; i < 2 * n; ++i) s += (double)in[i] * in[i];
  const double inv = 1.0 / std::max(std::sqrt(s / n), 1e-20);
  for (long long i = 0; i < 2 * n; ++i) out[i] = in[i] * inv;
}

// d_Rotate3DFT, T_INTERP_LINEAR (src/Transformation/Rotation.cu:263-374,
// Rotate3DFTKernel :541-654), one orientation; M = transpose(Euler) row-major.
// in: centred half layout (n/2+1, n, n), element (tx, ty, tz) at frequency
// (tx - n/2, ty - n/2, tz - n/2); ncomp = 2 for interleaved complex, 1 for real.
// Trilinear interpolation with exact weights and clamped texel indices (the
// reference's texture: 8-bit weights; wrap mode on unnormalised coordinates
// Cubic B-spline prefilter of one line (ConvertToInterpolationCoefficients,
// src/CubicSplines/internal/cubicPrefilter_kernel.cu): pole z = sqrt(3) - 2, gain
// (1 - z)(1 - 1/z), causal start c0 + sum_{n < min(12, len)} z^(n+1) c_n (the
// reference's truncated clamped-boundary sum), anticausal start z/(z - 1) c_(len-1).
void bspline_line(double* c, int len, size_t stride) {
  const double z = std::sqrt(3.0) - 2.0, lambda = (1.0 - z) * (1.0 - 1.0 / z);
  const int horizon = std::min(12, len);
  double zn = z, sum = c[0];
  for (int n = 0; n < horizon; ++n) {
    sum += zn * c[(size_t)n * stride];
    zn *= z;
  }
  double prev = lambda * sum;
  c[0] = prev;
  for (int n = 1; n < len; ++n) c[(size_t)n * stride] = prev = lambda * c[(size_t)n * stride] + z * prev;
  c[(size_t)(len - 1) * stride] = prev = (z / (z - 1.0)) * prev;
  for (int n = len - 2; n >= 0; --n) c[(size_t)n * stride] = prev = z * (prev - c[(size_t)n * stride]);
}

// d_CubicBSplinePrefilter3D / 2D (src/CubicSplines/cubicPrefilter3D.cu:114-127,
// cubicPrefilter2D.cu): lines along x, then y, then z.
void bspline_prefilter(double* v, int nx, int ny, int nz) {
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y) bspline_line(v + ((size_t)z * ny + y) * nx, nx, 1);
  for (int z = 0; z < nz; ++z)
    for (int x = 0; x < nx; ++x) bspline_line(v + (size_t)z * ny * nx + x, ny, nx);
  if (nz > 1)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) bspline_line(v + (size_t)y * nx + x, nz, (size_t)nx * ny);
}

double bspline_w(double t) {
  t = std::fabs(t);
  const double a = 2.0 - t;
  if (t < 1.0) return 2.0 / 3.0 - 0.5 * t * t * a;
  if (t < 2.0) return a * a * a / 6.0;
  return 0.0;
}

// Cubic B-spline of coef (nx x ny x nz, nz = 1 for images) at texel
// coordinate t: taps floor(t) - 1 .. floor(t) + 2 per axis, clamped.
double cubic_at(const double* coef, int nx, int ny, int nz, double tx, double ty, double tz) {
  auto cl = [](long v, long hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  const long fx = (long)std::floor(tx), fy = (long)std::floor(ty), fz = (long)std::floor(tz);
  const bool three = nz > 1;
  double r = 0;
  for (int kz = three ? -1 : 0; kz <= (three ? 2 : 0); ++kz) {
    const double wz = three ? bspline_w(kz - (tz - fz)) : 1.0;
    const long zz = three ? cl(fz + kz, nz - 1) : 0;
    for (int ky = -1; ky <= 2; ++ky) {
      const double wyz = bspline_w(ky - (ty - fy)) * wz;
      const long yy = cl(fy + ky, ny - 1);
      for (int kx = -1; kx <= 2; ++kx) r += bspline_w(kx - (tx - fx)) * wyz * coef[((size_t)zz * ny + yy) * nx + cl(fx + kx, nx - 1)];
    }
  }
  return r;
}

// Per-component cubic B-spline coefficients of an interleaved array of ncomp
// components (d_CubicBSplinePrefilter3D on the reference's Re and Im copies).
std::vector<std::vector<double>> spline_components(const double* in, int nx, int ny, int nz, int ncomp) {
  const size_t V = (size_t)nx * ny * nz;
  std::vector<std::vector<double>> out((size_t)ncomp, std::vector<double>(V));
  for (int c = 0; c < ncomp; ++c) {
    for (size_t i = 0; i < V; ++i) out[(size_t)c][i] = in[i * ncomp + c];
    bspline_prefilter(out[(size_t)c].data(), nx, ny, nz);
  }
  return out;
}

// acts as clamp).
void rotate3d_ft(const double* in, double* out, int n, const double* M, bool zerocentred, int ncomp,
                 bool cubic = false) {
  const int nh = n / 2 + 1;
  std::vector<std::vector<double>> coef;
  if (cubic) coef = spline_components(in, nh, n, n, ncomp);
  auto clampi = [](long v, long hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  for (int idz = 0; idz < n; ++idz)
    for (int idy = 0; idy < n; ++idy)
      for (int idx = 0; idx < nh; ++idx) {
        int x = idx, y = idy, z = idz;
        if (!zerocentred) {
          x = n / 2 - idx;
          y = n - 1 - ((idy + n / 2 - 1) % n);
          z = n - 1 - ((idz + n / 2 - 1) % n);
        }
        const double cx = x - n / 2, cy = y - n / 2, cz = z - n / 2;
        double p[3];
        for (int r = 0; r < 3; ++r) p[r] = M[3 * r] * cx + M[3 * r + 1] * cy + M[3 * r + 2] * cz;
        // the radius cut and the flip are decided in float, as the reference
        // does, with the device's fixed evaluation order; the value is
        // interpolated at the double position
        float pf[3];
        for (int r = 0; r < 3; ++r)
          pf[r] = std::fma((float)M[3 * r], (float)cx, std::fma((float)M[3 * r + 1], (float)cy,
                                                                (float)M[3 * r + 2] * (float)cz));
        double* o = out + (((size_t)idz * n + idy) * nh + idx) * ncomp;
        if (std::fma(pf[0], pf[0], std::fma(pf[1], pf[1], pf[2] * pf[2])) >= (float)(n * n / 4)) {
          for (int c = 0; c < ncomp; ++c) o[c] = 0;
          continue;
        }
        const bool neg = pf[0] > 1e-6f;
        double t[3];
        for (int r = 0; r < 3; ++r) t[r] = (neg ? -p[r] : p[r]) + n / 2;
        if (cubic) {
          for (int c = 0; c < ncomp; ++c) {
            const double v = cubic_at(coef[(size_t)c].data(), nh, n, n, t[0], t[1], t[2]);
            o[c] = (c == 1 && neg) ? -v : v;
          }
          continue;
        }
        const long f[3] = {(long)std::floor(t[0]), (long)std::floor(t[1]), (long)std::floor(t[2])};
        const double a[3] = {t[0] - f[0], t[1] - f[1], t[2] - f[2]};
        const long hi[3] = {nh - 1, n - 1, n - 1};
        for (int c = 0; c < ncomp; ++c) {
          double v = 0;
          for (int k = 0; k < 8; ++k) {
            const int bx = k & 1, by = (k >> 1) & 1, bz = k >> 2;
            const double w = (bx ? a[0] : 1 - a[0]) * (by ? a[1] : 1 - a[1]) * (bz ? a[2] : 1 - a[2]);
            const long ix = clampi(f[0] + bx, hi[0]), iy = clampi(f[1] + by, hi[1]), iz = clampi(f[2] + bz, hi[2]);
            v += w * in[(((size_t)iz * n + iy) * nh + ix) * ncomp + c];
          }
          o[c] = (c == 1 && neg) ? -v : v;
        }
      }
}

// Tilted specimen plane: CTFTiltParams::GetZGrid2D (include/CTF.cuh:160-180).
// Matrix3RotationCustom (src/Helpers/Euler.cu:312-331) as written there, glm
// column-major (m[col][row]); its element [0][0] is c + a.x + a.x (1 - c), not
// the Rodrigues c + a.x^2 (1 - c), and is kept. Matrix3RotationInPlaneAxis
// (:333-336) uses the axis (cos a, sin a, 0); Matrix2Rotation (:343-349) the
// columns (c, s), (-s, c).
void rot_custom(double ax, double ay, double az, double angle, double m[3][3]) {
  const double c = std::cos(angle), c1 = 1.0 - c, s = std::sin(angle);
  m[0][0] = c + ax + ax * c1;
  m[0][1] = ay * ax * c1 + az * s;
  m[0][2] = az * ax * c1 - ay * s;
  m[1][0] = ax * ay * c1 - az * s;
  m[1][1] = c + ay * ay * c1;
  m[1][2] = az * ay * c1 + ax * s;
  m[2][0] = ax * az * c1 + ay * s;
  m[2][1] = ay * az * c1 - ax * s;
  m[2][2] = c + az * az * c1;
}

// angles = (imageangle, stage axis, stage rotation, specimen axis, specimen
// rotation); z[n] = defocus + the plane's height under the region centre.
void tilt_zgrid(const double* angles, double defocus, double pixelsize, int imx, int imy, int rx, int ry,
                const int* pos, int n, double* z) {
  double A[3][3], B[3][3];
  rot_custom(std::cos(angles[1]), std::sin(angles[1]), 0.0, angles[2], A);
  rot_custom(std::cos(angles[3]), std::sin(angles[3]), 0.0, angles[4], B);
  double nrm[3];
  for (int r = 0; r < 3; ++r) nrm[r] = A[0][r] * B[2][0] + A[1][r] * B[2][1] + A[2][r] * B[2][2];
  const double c = std::cos(angles[0]), s = std::sin(angles[0]);
  for (int i = 0; i < n; ++i) {
    const double fx = (pos[3 * i] + rx / 2.0 - imx / 2.0) * pixelsize;
    const double fy = (pos[3 * i + 1] + ry / 2.0 - imy / 2.0) * pixelsize;
    const double gx = c * fx - s * fy, gy = s * fx + c * fy;
    z[i] = defocus + (-gx * nrm[0] - gy * nrm[1]) / nrm[2];
  }
}

// ---------------------------------------------------------------------------
// Sub-voxel peak location: d_Peak (src/Correlation/Peak.cu:26-216), d_PeakOne2D
// (:218-225, PeakOne2DKernel :380-454) and d_SubpixelMax
// (src/Correlation/LocalPeaks.cu:64-91, SubpixelMaxKernel :145-178), per map,
// in double from the float map.
// ---------------------------------------------------------------------------
// DimensionCount (include/Prerequisites.cuh): 3 if nz > 1, 2 if ny > 1, else 1.
inline int dimension_count(int ny, int nz) { return nz > 1 ? 3 : (ny > 1 ? 2 : 1); }
inline long long wrap(long long i, long long n) { const long long r = i % n; return r < 0 ? r + n : r; }
// sinc of include/DeviceFunctions.cuh:11-25 with the given zero threshold
// (1e-4 for the float overload, 1e-8 for the double one), evaluated in double.
inline double sinc_t(double x, double zero) {
  const double pi = 3.14159265358979323846;
  return std::fabs(x) <= zero ? 1.0 : std::sin(x * pi) / (x * pi);
}

struct PeakResult { double pos[3]; double value; };

// T_PEAK_INTEGER (:28-53): the maximum and its voxel. d_Max breaks ties in the
// order of its block reduction (src/Generics/MinMax.cu:382-425); the first
// maximum in raster order is this restatement's rule.
PeakResult peak_integer(const float* in, int nx, int ny, int nz) {
  const long long V = (long long)nx * ny * nz;
  long long arg = 0;
  for (long long i = 1; i < V; ++i)
    if (in[i] > in[arg]) arg = i;
  PeakResult r;
  r.pos[0] = (double)(arg % nx);
  r.pos[1] = (double)((arg / nx) % ny);
  r.pos[2] = (double)(arg / ((long long)nx * ny));
  r.value = in[arg];
  return r;
}

// SincMax1DKernel<n> (:230-320): y(x) = sum over |k| <= n/2 - 1 of
// sinc(k - x) s[n/2 + k] at x = id/512 - 1, id = 0..1023, terms added from the
// outermost pair inwards (-k before +k), the centre last, in double. 256
// threads: thread t visits id = t + 256 j in order and keeps a float pair
// (x, y), replaced when (double)best_y < y; the tree then keeps the lower slot
// unless the upper one's float y is strictly larger, so among equal float
// maxima the thread with the smallest bit-reversed index wins. Returns x.
double sinc_max_1d(const double* s, int n) {
  float bx[256], by[256];
  for (int t = 0; t < 256; ++t) {
    bx[t] = 0.f;
    by[t] = -1e30f;
    for (int id = t; id < 1024; id += 256) {
      const double x = (double)id / 512.0 - 1.0;
      double y = 0;
      for (int k = 9; k >= 1; --k) {
        if (!(n > 2 * k)) continue;
        y += sinc_t(-(double)k - x, 1e-8) * s[n / 2 - k];
        y += sinc_t((double)k - x, 1e-8) * s[n / 2 + k];
      }
      y += sinc_t(x, 1e-8) * s[n / 2];
      if ((double)by[t] < y) {
        bx[t] = (float)x;
        by[t] = (float)y;
      }
    }
  }
  for (int st = 128; st >= 1; st /= 2)
    for (int t = 0; t < st; ++t)
      if (by[t + st] > by[t]) {
        by[t] = by[t + st];
        bx[t] = bx[t + st];
      }
  return bx[0];
}

// T_PEAK_SUBCOARSE (:54-158): n = min(3D ? 10 : 20, nx rounded up to even)
// samples on the line through the integer peak along each dimension (cyclic,
// ExtractKernel src/Generics/Extraction.cu:139-153), starting at peak - n/2;
// each line gives the shift in [-1, 1) of its sinc maximum; the position is
// peak + shift. The value is the sinc interpolation at that position over the
// n^ndims region starting at peak - n/2 (SincEvalMaxKernel :322-378, float
// sinc of the sample-to-position distance per dimension).
PeakResult peak_subcoarse(const float* in, int nx, int ny, int nz) {
  PeakResult r = peak_integer(in, nx, ny, nz);
  const int nd = dimension_count(ny, nz);
  const int n = std::min(nd == 3 ? 10 : 20, (nx + 1) / 2 * 2);
  const long long dims[3] = {nx, ny, nz};
  const long long ip[3] = {(long long)r.pos[0], (long long)r.pos[1], (long long)r.pos[2]};
  auto at = [&](long long x, long long y, long long z) {
    return (double)in[(wrap(z, dims[2]) * dims[1] + wrap(y, dims[1])) * dims[0] + wrap(x, dims[0])];
  };
  double shift[3] = {0, 0, 0};
  std::vector<double> line(n);
  for (int d = 0; d < nd; ++d) {
    for (int i = 0; i < n; ++i) {
      long long q[3] = {ip[0], ip[1], ip[2]};
      q[d] += i - n / 2;
      line[i] = at(q[0], q[1], q[2]);
    }
    shift[d] = sinc_max_1d(line.data(), n);
  }
  double e[3] = {0, 0, 0};
  for (int d = 0; d < nd; ++d) {
    r.pos[d] = (double)ip[d] + shift[d];
    e[d] = (double)(n / 2) + shift[d];
  }
  const int ry = nd > 1 ? n : 1, rz = nd > 2 ? n : 1;
  double sum = 0;
  for (int z = 0; z < rz; ++z)
    for (int y = 0; y < ry; ++y)
      for (int x = 0; x < n; ++x) {
        double w = sinc_t(x - e[0], 1e-4);
        if (nd > 1) w *= sinc_t(y - e[1], 1e-4);
        if (nd > 2) w *= sinc_t(z - e[2], 1e-4);
        sum += at(ip[0] - n / 2 + x, ip[1] - ry / 2 + y, ip[2] - rz / 2 + z) * w;
      }
  r.value = sum;
  return r;
}

// The (new index, old index) pairs of one axis that FFTPadEvenKernel
// (src/FFT/FFTResize.cu:162-184) copies: new index i has centred coordinate
// nr = (i + nn/2) % nn, old centred coordinate or = nr + (no - nn)/2, kept
// when 0 <= or < no and read at (or + (no+1)/2) % no.
std::vector<std::pair<int, int>> pad_pairs(int no, int nn) {
  std::vector<std::pair<int, int>> v;
  for (int i = 0; i < nn; ++i) {
    const int nr = (i + nn / 2) % nn, orr = nr + (no - nn) / 2;
    if (orr >= 0 && orr < no) v.push_back({i, (orr + (no + 1) / 2) % no});
  }
  return v;
}

// T_PEAK_SUBFINE (:159-210): samples = 3D ? 5 : 9 (at most each dimension),
// subdivisions = 3D ? 63 : 105; the samples^ndims region centred on the integer
// peak (d_Extract src/Generics/Extraction.cu:25-35, cyclic) is scaled by
// d_Scale T_INTERP_FOURIER (src/Transformation/Scale.cu:109-139): R2C, zero
// pad of the half spectrum (FFTPadEvenKernel), C2R_u, times 1/N_old. The
// maximum of the scaled region (first in raster order on ties) moves the peak
// by (index - samples/2 * subdivisions) / subdivisions per dimension; the value
// is the larger of the integer maximum and the scaled one. The C2R is
// evaluated directly at every output point from the few non-zero bins.
PeakResult peak_subfine(const float* in, int nx, int ny, int nz) {
  PeakResult r = peak_integer(in, nx, ny, nz);
  const int nd = dimension_count(ny, nz);
  const long long dims[3] = {nx, ny, nz};
  int S = nd < 3 ? 9 : 5;
  for (int d = 0; d < nd; ++d) S = std::min<long long>(S, dims[d]);
  const int sub = nd < 3 ? 105 : 63, centre = S / 2 * sub;
  const int ro[3] = {S, (int)std::min<long long>(ny, S), (int)std::min<long long>(nz, S)};
  const int rn[3] = {S * sub, ny == 1 ? 1 : S * sub, nz == 1 ? 1 : S * sub};
  std::vector<double> reg((size_t)ro[0] * ro[1] * ro[2]);
  long long org[3];
  for (int d = 0; d < 3; ++d) org[d] = wrap((long long)r.pos[d] - ro[d] / 2, dims[d]);
  for (int z = 0; z < ro[2]; ++z)
    for (int y = 0; y < ro[1]; ++y)
      for (int x = 0; x < ro[0]; ++x)
        reg[((size_t)z * ro[1] + y) * ro[0] + x] =
            in[(wrap(org[2] + z, nz) * ny + wrap(org[1] + y, ny)) * nx + wrap(org[0] + x, nx)];
  const int hx = ro[0] / 2 + 1;
  std::vector<cpx<double>> spec((size_t)hx * ro[1] * ro[2]);
  r2c3<double>(reg.data(), spec.data(), ro[0], ro[1], ro[2]);
  const int kxn = (ro[0] + 1) / 2;  // x bins the pad keeps
  const auto py = pad_pairs(ro[1], rn[1]), pz = pad_pairs(ro[2], rn[2]);
  const double tau = 6.283185307179586476925286766559;
  auto phase = [&](long long k, long long j, long long n) { return std::polar(1.0, tau * (double)((k * j) % n) / n); };
  // G(kx, y', z') = sum over the kept (y, z) bins of S e^{2 pi i (ky y'/Ny' + kz z'/Nz')}
  std::vector<cpx<double>> ey(py.size() * rn[1]), ez(pz.size() * rn[2]);
  for (size_t b = 0; b < py.size(); ++b)
    for (int yy = 0; yy < rn[1]; ++yy) ey[b * rn[1] + yy] = phase(py[b].first, yy, rn[1]);
  for (size_t a = 0; a < pz.size(); ++a)
    for (int zz = 0; zz < rn[2]; ++zz) ez[a * rn[2] + zz] = phase(pz[a].first, zz, rn[2]);
  std::vector<cpx<double>> G((size_t)kxn * rn[1] * rn[2]);
  for (int zz = 0; zz < rn[2]; ++zz)
    for (int yy = 0; yy < rn[1]; ++yy)
      for (int kx = 0; kx < kxn; ++kx) {
        cpx<double> acc = 0;
        for (size_t a = 0; a < pz.size(); ++a)
          for (size_t b = 0; b < py.size(); ++b)
            acc += spec[((size_t)pz[a].second * ro[1] + py[b].second) * hx + kx] * ez[a * rn[2] + zz] *
                   ey[b * rn[1] + yy];
        G[((size_t)zz * rn[1] + yy) * kxn + kx] = acc;
      }
  std::vector<cpx<double>> tw((size_t)kxn * rn[0]);
  for (int kx = 0; kx < kxn; ++kx)
    for (int xx = 0; xx < rn[0]; ++xx) tw[(size_t)kx * rn[0] + xx] = (kx == 0 ? 1.0 : 2.0) * phase(kx, xx, rn[0]);
  const double norm = 1.0 / ((double)ro[0] * ro[1] * ro[2]);
  double best = -HUGE_VAL;
  long long arg = 0;
  for (long long row = 0; row < (long long)rn[1] * rn[2]; ++row)
    for (int xx = 0; xx < rn[0]; ++xx) {
      double v = 0;
      for (int kx = 0; kx < kxn; ++kx) v += (G[(size_t)row * kxn + kx] * tw[(size_t)kx * rn[0] + xx]).real();
      v *= norm;
      if (v > best) {
        best = v;
        arg = row * rn[0] + xx;
      }
    }
  const long long plane = (long long)S * S * sub * sub, rowlen = (long long)S * sub;
  const long long iz = arg / plane, iy = (arg - iz * plane) / rowlen, ix = arg - iz * plane - iy * rowlen;
  r.value = std::max(r.value, best);
  r.pos[0] += (double)(ix - centre) / sub;
  if (ny > 1) r.pos[1] += (double)(iy - centre) / sub;
  if (nz > 1) r.pos[2] += (double)(iz - centre) / sub;
  return r;
}

// PeakOne2DKernel (:380-454): the maximum of the centred region (rx, ry) of an
// nx x ny image. min(128, rx*ry) threads: thread t visits id = t + T j with
// strict '>' from (-1e20, (0, 0)); thread 0 then takes the others in order with
// strict '>'. The position moves by the value-weighted centroid of the 3x3
// neighbourhood (clipped at the borders) and loses (nx/2, ny/2) on request.
void peak_one_2d(const float* in, int nx, int ny, int rx, int ry, bool subtractcenter, double* pos, double* value) {
  const int T = std::min(128, rx * ry), ox = (nx - rx) / 2, oy = (ny - ry) / 2;
  float mv = -1e20f;
  int mx = 0, my = 0;
  for (int t = 0; t < T; ++t) {
    float tv = -1e20f;
    int tx = 0, ty = 0;
    for (int id = t; id < rx * ry; id += T) {
      const int y = id / rx + oy, x = id % rx + ox;
      const float v = in[(size_t)y * nx + x];
      if (v > tv) { tv = v; tx = x; ty = y; }
    }
    if (t == 0 || tv > mv) { mv = tv; mx = tx; my = ty; }
  }
  double cx = 0, cy = 0, sw = 0;
  for (int y = -1; y <= 1; ++y) {
    if (my + y < 0 || my + y >= ny) continue;
    for (int x = -1; x <= 1; ++x) {
      if (mx + x < 0 || mx + x >= nx) continue;
      const double v = in[(size_t)(my + y) * nx + mx + x];
      cx += v * x;
      cy += v * y;
      sw += v;
    }
  }
  if (sw != 0) { cx /= sw; cy /= sw; }
  pos[0] = mx + cx - (subtractcenter ? nx / 2 : 0);
  pos[1] = my + cy - (subtractcenter ? ny / 2 : 0);
  pos[2] = 0;
  *value = mv;
}

// d_SubpixelMax: for every sub-sample shift s = (i + 1/2)/steps - 1/2 per
// dimension (float arithmetic as at :70-74; s.z = 0 in 2D), out = max(out,
// sum over the 13^ndims window of in * sinc(s - offset) per dimension) at the
// voxels the launch covers: x in [6, min(nx - 9, 128 ceil((nx - 15)/128) + 5)],
// y in [6, ny - 10], z in [6, nz - 10] (3D; z = 0 in 2D). Nothing is written when
// a grid dimension (nx - 15, ny - 15, nz - 15) is not positive. 1D is rejected by
// the reference (throw) and here returns without writing.
void subpixel_max(const float* in, double* out, int nx, int ny, int nz, int steps) {
  const int nd = dimension_count(ny, nz), rad = 6;
  if (nd < 2 || steps < 1) return;
  const int gx = (nx - 15 + 127) / 128, gy = ny - 15, gz = nd > 2 ? nz - 15 : 1;
  if (gx <= 0 || gy <= 0 || gz <= 0) return;
  const int xhi = std::min(nx - rad - 3, gx * 128 - 1 + rad), yhi = std::min(ny - rad - 3, gy - 1 + rad);
  const int zlo = nd > 2 ? rad : 0, zhi = nd > 2 ? std::min(nz - rad - 3, gz - 1 + rad) : 0;
  volatile float step = 1.0f / steps;
  std::vector<double> wx(13), wy(13), wz(13);
  auto weights = [&](int i, std::vector<double>& w) {
    volatile float a = (float)i * step;
    volatile float b = a - 0.5f;
    volatile float c = b + step / 2;
    for (int k = -rad; k <= rad; ++k) {
      volatile float d = c - (float)k;
      w[k + rad] = sinc_t(d, 1e-4);
    }
  };
  for (int sz = 0; sz < (nd == 3 ? steps : 1); ++sz) {
    if (nd == 3) weights(sz, wz);
    for (int sy = 0; sy < steps; ++sy) {
      weights(sy, wy);
      for (int sx = 0; sx < steps; ++sx) {
        weights(sx, wx);
        for (int cz = zlo; cz <= zhi; ++cz)
          for (int cy = rad; cy <= yhi; ++cy)
            for (int cx = rad; cx <= xhi; ++cx) {
              double sum = 0;
              for (int z = (nd == 3 ? -rad : 0); z <= (nd == 3 ? rad : 0); ++z) {
                const double fz = nd == 3 ? wz[z + rad] : 1.0;
                for (int y = -rad; y <= rad; ++y)
                  for (int x = -rad; x <= rad; ++x)
                    sum += in[((size_t)(cz + z) * ny + cy + y) * nx + cx + x] * wx[x + rad] * wy[y + rad] * fz;
              }
              double& o = out[((size_t)cz * ny + cy) * nx + cx];
              o = std::max(o, sum);
            }
      }
    }
  }
}

}  // namespace

extern "C" int orc_equal_angular_spacing(float phimin, float phimax, float thetamin, float thetamax, float psimin,
                                         float psimax, float increment, float* out, int maxcount);

namespace {

// gaussrand (src/Helpers/Misc.cpp:14-31): Box-Muller pairs drawn from the C
// runtime's rand(); U, V and the result are tfloat (float) and the C++ float
// overloads of log, sqrt, sin and cos apply to them. The state persists for the
// process, as the reference's function-scope statics do.
struct OrcGaussRand {
  float U = 0, V = 0;
  int phase = 0;
  float next() {
    float Z;
    if (phase == 0) {
      U = (float)((std::rand() + 1.) / (RAND_MAX + 2.));
      V = (float)(std::rand() / (RAND_MAX + 1.));
      Z = std::sqrt(-2 * std::log(U)) * std::sin(2 * 3.1415926535897932384626433832795f * V);
    } else {
      Z = std::sqrt(-2 * std::log(U)) * std::cos(2 * 3.1415926535897932384626433832795f * V);
    }
    phase = 1 - phase;
    return Z;
  }
};
OrcGaussRand g_orc_gauss;

// d_Rotate2DFT with linear interpolation (src/Transformation/Rotation.cu:235-256,
// Rotate2DFTKernel :485-539) on a zero-centred complex half layout
// ((nx/2+1) x ny): the position comes from the output thread's own (idx, idy),
// pos = R(-angle) (idx - nx/2, idy - ny/2) with the matrix entries rounded to
// float as glm::mat2 stores them (Matrix2Rotation, src/Helpers/Euler.cu:343-349);
// 0 when |pos| > maxfreq; the Hermitian mate (imaginary part negated) when
// pos.x > 1e-5; bilinear at texel pos + n/2 with indices clamped; written at
// (idx, idy) or at the FFT-order bin (nx/2 - idx, ny-1 - ((idy + ny/2 - 1) % ny)).
void rotate2d_ft(const double* in, double* out, int nx, int ny, double angle, double maxfreq, bool zerocentred,
                 bool cubic = false) {
  const int nh = nx / 2 + 1;
  std::vector<std::vector<double>> coef;
  if (cubic) coef = spline_components(in, nh, ny, 1, 2);
  const double c = (float)std::cos(-angle), s = (float)std::sin(-angle);
  auto clampi = [](long v, long hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  for (int idy = 0; idy < ny; ++idy)
    for (int idx = 0; idx < nh; ++idx) {
      int x = idx, y = idy;
      if (!zerocentred) {
        x = nx / 2 - idx;
        y = ny - 1 - ((idy + ny / 2 - 1) % ny);
      }
      const double vx = idx - nx / 2, vy = idy - ny / 2;
      double px = c * vx - s * vy, py = s * vx + c * vy;
      // the radius cut and the flip in float (the reference's glm::vec2), in
      // the device's fixed evaluation order; the value at the double position
      const float qx = std::fma((float)c, (float)vx, -((float)s * (float)vy));
      const float qy = std::fma((float)s, (float)vx, (float)c * (float)vy);
      double* o = out + ((size_t)y * nh + x) * 2;
      if (std::sqrt(std::fma(qx, qx, qy * qy)) > (float)maxfreq) {
        o[0] = o[1] = 0;
        continue;
      }
      const bool neg = qx > 1e-5f;
      if (neg) {
        px = -px;
        py = -py;
      }
      const double tx = px + nx / 2, ty = py + ny / 2;
      if (cubic) {
        for (int k = 0; k < 2; ++k) {
          const double v = cubic_at(coef[(size_t)k].data(), nh, ny, 1, tx, ty, 0.0);
          o[k] = (k == 1 && neg) ? -v : v;
        }
        continue;
      }
      const long fx = (long)std::floor(tx), fy = (long)std::floor(ty);
      const double ax = tx - fx, ay = ty - fy;
      for (int k = 0; k < 2; ++k) {
        double v = 0;
        for (int q = 0; q < 4; ++q) {
          const int bx = q & 1, by = q >> 1;
          const double w = (bx ? ax : 1 - ax) * (by ? ay : 1 - ay);
          v += w * in[((size_t)clampi(fy + by, ny - 1) * nh + clampi(fx + bx, nh - 1)) * 2 + k];
        }
        o[k] = (k == 1 && neg) ? -v : v;
      }
    }
}

// d_Rotate3D / d_Rotate2D with T_INTERP_CUBIC into FFT order
// (Rotate3DKernel/Rotate2DKernel<true, false>, src/Transformation/Rotation.cu):
// output voxel i reads x = (i + n/2) % n per axis at p = M (x - n/2) + n/2 + 0.5,
// the cubic B-spline of coef at p - 0.5 over 4^ndims taps with clamped indices
// (cubicTex3DSimple, include/CubicInterp.cuh:287-312). M: row-major 3x3 (3D,
// transpose(Euler)), or the 2x2 R(-angle) of Matrix2Rotation in M[0], M[1],
// M[3], M[4] (2D, nz = 1).
void rotate_cubic(const double* coef, double* out, int nx, int ny, int nz, const double* M) {
  const bool three = nz > 1;
  auto cl = [](long v, long hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const double cx = (ix + nx / 2) % nx - nx / 2, cy = (iy + ny / 2) % ny - ny / 2,
                     cz = three ? (iz + nz / 2) % nz - nz / 2 : 0;
        const double gx = M[0] * cx + M[1] * cy + M[2] * cz + nx / 2;
        const double gy = M[3] * cx + M[4] * cy + M[5] * cz + ny / 2;
        const double gz = three ? M[6] * cx + M[7] * cy + M[8] * cz + nz / 2 : 0;
        const long fx = (long)std::floor(gx), fy = (long)std::floor(gy), fz = (long)std::floor(gz);
        double r = 0;
        for (int tz = three ? -1 : 0; tz <= (three ? 2 : 0); ++tz) {
          const double wz = three ? bspline_w(tz - (gz - fz)) : 1.0;
          const long zz = three ? cl(fz + tz, nz - 1) : 0;
          for (int ty = -1; ty <= 2; ++ty) {
            const double wyz = bspline_w(ty - (gy - fy)) * wz;
            const long yy = cl(fy + ty, ny - 1);
            for (int tx = -1; tx <= 2; ++tx)
              r += bspline_w(tx - (gx - fx)) * wyz * coef[((size_t)zz * ny + yy) * nx + cl(fx + tx, nx - 1)];
          }
        }
        out[((size_t)iz * ny + iy) * nx + ix] = r;
      }
}

// ProbabilisticPicker (src/Correlation/ProbabilisticPicker.cu), for one
// reference (index refid), in the reference's order:
//  Initialize (:33-152): R = ifftshift(pad(ref, dimsimage, 0)) (:128-129),
//   C = FFTCrop(R2C(R), 2*lowpass) (:131-132), C *= [|k| < lowpass] (d_Bandpass
//   :136 with a hard sphere, src/ImageManipulation/Bandpass.cu:37-87),
//   Hc = d_RemapHalfFFT2Half(C) as RemapHalfFFT2HalfKernel (src/FFT/FFTRemap.cu)
//   reads it: row signed(i) + n/2 (n for the zero-frequency row and plane),
//   complex index j >= ElementsFFT(2*lowpass) reading the float pair
//   (R[2j], R[2j+1]) that the reference's buffer still holds there;
//   M = ifftshift(pad(mask)) (:104-106), IM = R2C(ifftshift(pad(1 - mask))) (:95-100);
//  SetImage (:205-224): image to mean 0 / std 1 (d_Norm), I = R2C(image) and
//   I2 = R2C(image^2) band-limited, ctf cropped to 2*lowpass;
//  PerformCorrelation (:226-389): srand(123); per orientation rot =
//   Rotate{3D,2D}FT(Hc) in FFT order, *ctf, r = C2R(FFTPad(rot)) / N * M,
//   corr = C2R_u(I * conj(R2C(r))); over the FFTFullCrop of r and M to dimsref,
//   the host block (:306-340) with five gaussrand() draws per element;
//   solvent mean/std at the first orientation (CalcSolventStatistics :423-454);
//   the score of ProbPickingUpdateKernel (:391-421), kept where it exceeds best.
// top2 (optional) receives the second-highest score over the orientations.
// With a non-circular mask (circular = false): the padded mask, centred, as
// cubic B-spline coefficients (:102-110); per orientation it is rotated into FFT
// order (:273-279) and replaces M, and the inverse mask, its spectrum and the
// solvent statistics are recomputed from its crop (:342-356).
void prob_picker(const double* ref, const double* mask, const int* dr, const int* di, int lowpass, bool doctf,
                 const double* image, const double* ctf, float anglestep, int refid, double* bestccf,
                 double* bestangle, int* bestref, double* top2, bool circular) {
  const bool three = di[2] > 1;
  const int n = 2 * lowpass;
  const int dl[3] = {n, n, three ? n : 1};
  const size_t Vi = (size_t)di[0] * di[1] * di[2], Hi = (size_t)(di[0] / 2 + 1) * di[1] * di[2];
  const size_t Hl = (size_t)(n / 2 + 1) * n * dl[2], Er = (size_t)dr[0] * dr[1] * dr[2];
  auto pad = [&](const double* in) {
    std::vector<double> out(Vi, 0.0);
    const int off[3] = {dr[0] / 2 - di[0] / 2, dr[1] / 2 - di[1] / 2, dr[2] / 2 - di[2] / 2};
    for (int z = 0; z < di[2]; ++z)
      for (int y = 0; y < di[1]; ++y)
        for (int x = 0; x < di[0]; ++x) {
          const int ix = x + off[0], iy = y + off[1], iz = z + off[2];
          if (ix >= 0 && ix < dr[0] && iy >= 0 && iy < dr[1] && iz >= 0 && iz < dr[2])
            out[((size_t)z * di[1] + y) * di[0] + x] = in[((size_t)iz * dr[1] + iy) * dr[0] + ix];
        }
    return out;
  };
  auto shifted = [&](const std::vector<double>& a) {
    std::vector<double> out(Vi);
    remap_full2fullfft(a.data(), out.data(), di[0], di[1], di[2]);
    return out;
  };
  auto r2c = [&](const std::vector<double>& r) {
    std::vector<cpx<double>> f(Hi);
    r2c3(r.data(), f.data(), di[0], di[1], di[2]);
    return f;
  };
  auto c2r = [&](std::vector<cpx<double>> f) {
    std::vector<double> r(Vi);
    c2r3(f.data(), r.data(), di[0], di[1], di[2]);
    return r;
  };
  auto sfreq = [](int i, int len) { return i < len / 2 + 1 ? i : i - len; };
  auto bandpass = [&](cpx<double>* x, const int* d) {
    const int nh = d[0] / 2 + 1;
    for (int z = 0; z < d[2]; ++z)
      for (int y = 0; y < d[1]; ++y)
        for (int xx = 0; xx < nh; ++xx) {
          const long long kz = sfreq(z, d[2]), ky = sfreq(y, d[1]);
          if (!((long long)xx * xx + ky * ky + kz * kz < (long long)lowpass * lowpass))
            x[((size_t)z * d[1] + y) * nh + xx] = 0;
        }
  };

  // Initialize
  const std::vector<double> R = shifted(pad(ref));
  std::vector<cpx<double>> F = r2c(R);
  std::vector<cpx<double>> C(Hl);
  fft_resize(1, reinterpret_cast<const double*>(F.data()), reinterpret_cast<double*>(C.data()), di, dl, 2);
  bandpass(C.data(), dl);
  std::vector<double> Hc(2 * Hl);
  {
    const int nh = n / 2 + 1, nz = dl[2];
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < n; ++y)
        for (int x = 0; x < nh; ++x) {
          const int rz = (z < nz / 2 + 1 ? z : z - n) + nz / 2, ry = (y < n / 2 + 1 ? y : y - n) + n / 2;
          const size_t j = ((size_t)rz * n + ry) * nh + x, o = ((size_t)z * n + y) * nh + x;
          if (j < Hl) {
            Hc[2 * o] = C[j].real();
            Hc[2 * o + 1] = C[j].imag();
          } else {
            Hc[2 * o] = R[2 * j];
            Hc[2 * o + 1] = R[2 * j + 1];
          }
        }
  }
  std::vector<double> M = circular ? shifted(pad(mask)) : pad(mask);
  if (!circular) bspline_prefilter(M.data(), di[0], di[1], di[2]);
  std::vector<double> om(Er);
  for (size_t i = 0; i < Er; ++i) om[i] = 1.0 - mask[i];
  std::vector<cpx<double>> IM = r2c(shifted(pad(om.data())));
  std::vector<double> Mrot(Vi);

  // SetImage
  std::vector<double> imn(Vi), sq(Vi);
  norm_mean01std(image, imn.data(), (long long)Vi);
  std::vector<cpx<double>> I = r2c(imn);
  bandpass(I.data(), di);
  for (size_t i = 0; i < Vi; ++i) sq[i] = imn[i] * imn[i];
  std::vector<cpx<double>> I2 = r2c(sq);
  bandpass(I2.data(), di);
  std::vector<double> ctfl(Hl, 1.0);
  if (doctf) fft_resize(1, ctf, ctfl.data(), di, dl, 1);

  // PerformCorrelation
  std::srand(123);
  std::vector<float> ang;
  if (three) {
    const float PIf = 3.1415926535897932384626433832795f, PI2f = 6.283185307179586476925286766559f;
    const int cnt = orc_equal_angular_spacing(0, PI2f, 0, PIf, 0, PI2f, anglestep, nullptr, 0);
    ang.resize(3 * (size_t)cnt);
    orc_equal_angular_spacing(0, PI2f, 0, PIf, 0, PI2f, anglestep, ang.data(), cnt);
  } else {
    const float PI2f = 6.283185307179586476925286766559f;
    const int steps = (int)std::ceil(PI2f / anglestep);
    const float step = PI2f / (float)steps;
    for (int a = 0; a < steps; ++a) {
      ang.push_back(0.f);
      ang.push_back(0.f);
      ang.push_back((float)a * step);
    }
  }
  std::vector<double> smean(Vi), sstd(Vi);
  std::vector<double> rot(2 * Hl);
  std::vector<cpx<double>> P(Hi);
  for (size_t a = 0; a < ang.size() / 3; ++a) {
    double Mr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (three) {
      projector_matrix(ang[3 * a], ang[3 * a + 1], ang[3 * a + 2], 1.0, Mr);
      rotate3d_ft(Hc.data(), rot.data(), n, Mr, false, 2);
    } else {
      rotate2d_ft(Hc.data(), rot.data(), n, n, ang[3 * a + 2], (double)(float)(n / 2), false);
      const double c = (float)std::cos(-(double)ang[3 * a + 2]), sn = (float)std::sin(-(double)ang[3 * a + 2]);
      Mr[0] = c;
      Mr[1] = -sn;
      Mr[3] = sn;
      Mr[4] = c;
    }
    const double* Mk = M.data();
    if (!circular) {
      rotate_cubic(M.data(), Mrot.data(), di[0], di[1], di[2], Mr);
      Mk = Mrot.data();
    }
    if (doctf)
      for (size_t i = 0; i < Hl; ++i) {
        rot[2 * i] *= ctfl[i];
        rot[2 * i + 1] *= ctfl[i];
      }
    fft_resize(2, rot.data(), reinterpret_cast<double*>(P.data()), dl, di, 2);
    std::vector<double> r = c2r(P);
    for (size_t i = 0; i < Vi; ++i) r[i] = r[i] / (double)Vi * Mk[i];
    std::vector<cpx<double>> B1 = r2c(r);
    for (size_t i = 0; i < Hi; ++i) B1[i] = I[i] * std::conj(B1[i]);
    const std::vector<double> corr = c2r(B1);
    std::vector<double> refc(Er), maskc(Er);
    fft_resize(3, r.data(), refc.data(), di, dr, 1);
    fft_resize(3, Mk, maskc.data(), di, dr, 1);
    double smr = 0, smr2 = 0, samples = 0, epr[5] = {0, 0, 0, 0, 0};
    for (size_t i = 0; i < Er; ++i) {
      samples += maskc[i];
      const double rs = refc[i], rs2 = rs * rs;
      for (int k = 0; k < 5; ++k) epr[k] += rs2 + 2.0 * rs * (double)g_orc_gauss.next();
      smr += rs;
      smr2 += rs2;
    }
    smr /= samples;
    smr2 /= samples;
    double ep = 0;
    for (int k = 0; k < 5; ++k) ep += epr[k];
    ep /= 5.0;
    ep = std::exp(ep / (2.0 * samples));
    if (!circular) {
      std::vector<double> omc(Er), padded(Vi);
      for (size_t i = 0; i < Er; ++i) omc[i] = 1.0 - maskc[i];
      fft_resize(4, omc.data(), padded.data(), dr, di, 1);
      IM = r2c(padded);
    }
    if (a == 0 || !circular) {
      const double f = 1.0 / ((double)Vi * ((double)Er - samples));
      std::vector<cpx<double>> t(Hi);
      for (size_t i = 0; i < Hi; ++i) t[i] = I[i] * std::conj(IM[i]);
      smean = c2r(t);
      for (size_t i = 0; i < Hi; ++i) t[i] = I2[i] * std::conj(IM[i]);
      sstd = c2r(t);
      for (size_t i = 0; i < Vi; ++i) {
        smean[i] *= f;
        sstd[i] = std::sqrt(std::max(0.0, sstd[i] * f - smean[i] * smean[i]));
      }
    }
    const double nf = 1.0 / ((double)Vi * samples);
    for (size_t i = 0; i < Vi; ++i) {
      double d = -2.0 * nf * corr[i];
      d += 2.0 * smean[i] * smr;
      d /= std::max(1e-7, sstd[i]);
      d += smr2;
      d = std::exp(-d / 2.0);
      d = (d - 1.0) / (ep - 1.0);
      if (d > bestccf[i]) {
        if (top2) top2[i] = std::max(top2[i], bestccf[i]);
        bestccf[i] = d;
        for (int k = 0; k < 3; ++k) bestangle[3 * i + k] = ang[3 * a + k];
        bestref[i] = refid;
      } else if (top2) {
        top2[i] = std::max(top2[i], d);
      }
    }
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI for ctypes. Complex arrays are interleaved (re, im).
// ---------------------------------------------------------------------------
extern "C" {

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

void orc_fft_c2c_1d_f64(double* data, int n, int sign) {
  Fft1D<double> plan(n);
  std::vector<cpx<double>> s(n);
  plan.run(reinterpret_cast<cpx<double>*>(data), sign, s.data());
}

void orc_r2c_3d_f64(const double* in, double* out, int nx, int ny, int nz, int batch) {
  long long V = (long long)nx * ny * nz, H = (long long)(nx / 2 + 1) * ny * nz;
  for (int b = 0; b < batch; ++b)
    r2c3(in + b * V, reinterpret_cast<cpx<double>*>(out) + b * H, nx, ny, nz);
}
void orc_c2r_3d_f64(const double* in, double* out, int nx, int ny, int nz, int batch) {
  long long V = (long long)nx * ny * nz, H = (long long)(nx / 2 + 1) * ny * nz;
  for (int b = 0; b < batch; ++b)
    c2r3(reinterpret_cast<const cpx<double>*>(in) + b * H, out + b * V, nx, ny, nz);
}
void orc_r2c_3d_f32(const float* in, float* out, int nx, int ny, int nz, int batch) {
  long long V = (long long)nx * ny * nz, H = (long long)(nx / 2 + 1) * ny * nz;
  for (int b = 0; b < batch; ++b)
    r2c3(in + b * V, reinterpret_cast<cpx<float>*>(out) + b * H, nx, ny, nz);
}
void orc_c2r_3d_f32(const float* in, float* out, int nx, int ny, int nz, int batch) {
  long long V = (long long)nx * ny * nz, H = (long long)(nx / 2 + 1) * ny * nz;
  for (int b = 0; b < batch; ++b)
    c2r3(reinterpret_cast<const cpx<float>*>(in) + b * H, out + b * V, nx, ny, nz);
}

void orc_euler_matrix(double rot, double tilt, double psi, double* A) {
  euler_matrix(rot, tilt, psi, A);
}
void orc_projector_matrix(double rot, double tilt, double psi, double s, double* M) {
  projector_matrix(rot, tilt, psi, s, M);
}

// One orientation per call; M row-major (includes the oversampling factor).
void orc_project_f64(const double* vol, int p, double* out, int n, const double* M, int rmax) {
  project1(reinterpret_cast<const cpx<double>*>(vol), p, reinterpret_cast<cpx<double>*>(out), n, M, rmax);
}

// 2D central sections ((n/2+1) x n output), Project3DArraytoNDKernel<2>.
void orc_project2d_f64(const double* vol, int p, double* out, int n, const double* M, int rmax) {
  project1(reinterpret_cast<const cpx<double>*>(vol), p, reinterpret_cast<cpx<double>*>(out), n, M, rmax, 1);
}

// GetEqualAngularSpacing (src/Helpers/AngularSpacing.cu:46-67), restated with
// the reference's float32 loop counters. Writes up to maxcount (rot, tilt, psi)
// triples and returns the total count.
int orc_equal_angular_spacing(float phimin, float phimax, float thetamin, float thetamax,
                              float psimin, float psimax, float increment, float* out,
                              int maxcount) {
  const float PI2f = 6.283185307179586476925286766559f;
  int count = 0;
  for (float tilt = thetamin; tilt <= thetamax + 1e-5f; tilt += increment) {
    int rot_nstep = tilt > 0 ? (int)std::ceil(PI2f * std::sin(tilt) / increment) : 1;
    float rot_sam = PI2f / (float)rot_nstep;
    for (float rot = phimin; rot <= phimax - rot_sam + 1e-5f; rot += rot_sam)
      for (float psi = psimin; psi <= psimax - increment + 1e-5f; psi += increment) {
        if (out && count < maxcount) {
          out[3 * count] = rot;
          out[3 * count + 1] = tilt;
          out[3 * count + 2] = psi;
        }
        count++;
      }
  }
  return count;
}

void orc_ctf3_f64(double* out, int n, double pixelsize_m, double voltage_V, double Cs_m,
                  double defocus_m, double amplitude, double phaseshift, double bfactor_m2,
                  double scale) {
  ctf3(out, n, pixelsize_m, voltage_V, Cs_m, defocus_m, amplitude, phaseshift, bfactor_m2, scale);
}

void orc_sphere_mask_f64(double* out, int nx, int ny, int nz, double radius, double sigma) {
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x)
        out[((long long)z * ny + y) * nx + x] = soft_sphere(x, y, z, nx, ny, nz, radius, sigma);
}

void orc_remap_full2fullfft_f64(const double* in, double* out, int nx, int ny, int nz) {
  remap_full2fullfft(in, out, nx, ny, nz);
}

void orc_local_std_f64(const double* map, int nx, int ny, int nz, double radius,
                       double* std_out, double* mean_out, const double* fouriermask) {
  local_std(map, nx, ny, nz, radius, std_out, mean_out, fouriermask);
}

// Running max/argmax over a batch of maps (UpdateCorrelationKernel,
// src/Correlation/SubTomograms.cu:138-158), int32 index.
void orc_update_max_argmax_f64(const double* corr, long long vectorlength, int batch,
                               int batchoffset, double* best, int32_t* bestidx) {
  for (long long i = 0; i < vectorlength; ++i) {
    double b = best[i];
    int32_t a = bestidx[i];
    for (int k = 0; k < batch; ++k) {
      double c = corr[k * vectorlength + i];
      if (c > b) {
        b = c;
        a = k + batchoffset;
      }
    }
    best[i] = b;
    bestidx[i] = a;
  }
}

// d_LocalPeaks (src/Correlation/LocalPeaks.cu:24-62) with LocalPeaksKernel
// (src/Correlation/LocalPeaks.cu:98-143), one volume: voxel (x, y, z) is a
// peak unless its value is below the threshold or some other voxel within the
// box clamped to the volume, at squared distance <= localextent^2 (the
// kernel's sqdist > r^2 + 1e-5 test on integers), holds a strictly larger
// value. Writes up to `capacity` linear indices in raster order (z, then y,
// then x, the reference's host scan at :47-51) and returns the peak count.
long long orc_local_peaks_f32(const float* in, int nx, int ny, int nz, int localextent, float threshold,
                              long long* out, long long capacity) {
  const long long r2 = (long long)localextent * localextent;
  long long count = 0;
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const float v = in[((long long)z * ny + y) * nx + x];
        if (v < threshold) continue;
        bool peak = true;
        for (int zz = std::max(0, z - localextent); peak && zz <= std::min(nz - 1, z + localextent); ++zz)
          for (int yy = std::max(0, y - localextent); peak && yy <= std::min(ny - 1, y + localextent); ++yy)
            for (int xx = std::max(0, x - localextent); xx <= std::min(nx - 1, x + localextent); ++xx) {
              const long long d2 = (long long)(zz - z) * (zz - z) + (long long)(yy - y) * (yy - y) +
                                   (long long)(xx - x) * (xx - x);
              if (d2 > r2 || d2 == 0) continue;
              if (v < in[((long long)zz * ny + yy) * nx + xx]) {
                peak = false;
                break;
              }
            }
        if (peak) {
          if (count < capacity) out[count] = ((long long)z * ny + y) * nx + x;
          ++count;
        }
      }
  return count;
}

void orc_diff2_f64(const double* projft, int p, double oversample, const double* expft, const double* ctf, int n,
                   int rx, int ry, int rz, const double* angles, int nangles, double* best, int32_t* idx,
                   double* second) {
  diff2(reinterpret_cast<const cpx<double>*>(projft), p, oversample, reinterpret_cast<const cpx<double>*>(expft), ctf,
        n, rx, ry, rz, angles, nangles, best, idx, second);
}

// d_CTFSimulate at n addresses (k [cycles/pixel], angle) per parameter set;
// params: 14 doubles per set in CtfFull order.
void orc_ctf_simulate_f64(const double* params, const double* addresses, const double* gammacorr, double* out,
                          int n, int ampsq, int ignorefirst, int batch) {
  for (int b = 0; b < batch; ++b) {
    CtfFull q;
    std::memcpy(&q, params + 14 * b, sizeof(q));
    for (int i = 0; i < n; ++i)
      out[(long long)b * n + i] = ctf_at(q, addresses[2 * i], addresses[2 * i + 1], gammacorr ? gammacorr[i] : 0.0,
                                         ampsq != 0, ignorefirst != 0);
  }
}

void orc_full_pad_f64(const double* in, double* out, int ox, int oy, int oz, int nx, int ny, int nz) {
  full_pad(in, out, ox, oy, oz, nx, ny, nz);
}

void orc_norm_mean01std_f64(const double* in, double* out, long long n) { norm_mean01std(in, out, n); }

void orc_picker_f64(const double* ref, int nr, int nrz, const double* image, int nx, int ny, int nz,
                    const double* ctf, const double* angles, int nangles, double* best, int32_t* idx,
                    double* top2) {
  picker(ref, nr, nrz, image, nx, ny, nz, ctf, angles, nangles, best, idx, top2);
}

// Returns the wall time of the orientation loop in seconds.
void orc_exact2d_weighting_f64(double* w, int nx, int ny, const double* angles, const double* weights, int nimages,
                               const int* indices, int batch, double maxfreq) {
  exact2d_weighting(w, nx, ny, angles, weights, nimages, indices, batch, maxfreq);
}

void orc_tomo_picker_f64(const double* ref, const double* mask, int nr, int circular, double* images, int nx, int ny,
                         int nimages, const double* tilt, const double* weights, const double* angles, int nangles,
                         int cx, int cy, int cz, double* corr, int32_t* corridx, double* corrsum1, double* corrsum2,
                         double* corr2) {
  tomo_picker(ref, mask, nr, circular != 0, images, nx, ny, nimages, tilt, weights, angles, nangles, cx, cy, cz, corr,
              corridx, corrsum1, corrsum2, corr2);
}

double orc_pick_f64(const double* projft, int p, double oversample, const double* expft,
                    const double* ctf, int ctf_per_volume, int n, int nvol,
                    const double* angles, int nangles, double maskradius, double* best_out,
                    int32_t* idx_out, double* top1_out, double* top2_out, double* sigma_out) {
  return pick<double>(projft, p, oversample, expft, ctf, ctf_per_volume, n, nvol, angles,
                      nangles, maskradius, best_out, idx_out, top1_out, top2_out, sigma_out);
}

// The same search on n x n images (nvol images, CTF of (n/2+1)*n per image or shared).
double orc_pick2d_f64(const double* projft, int p, double oversample, const double* expft,
                      const double* ctf, int ctf_per_volume, int n, int nvol,
                      const double* angles, int nangles, double maskradius, double* best_out,
                      int32_t* idx_out, double* top1_out, double* top2_out, double* sigma_out) {
  return pick<double>(projft, p, oversample, expft, ctf, ctf_per_volume, n, nvol, angles,
                      nangles, maskradius, best_out, idx_out, top1_out, top2_out, sigma_out, 1);
}

double orc_pick_f32(const float* projft, int p, double oversample, const float* expft,
                    const float* ctf, int ctf_per_volume, int n, int nvol,
                    const double* angles, int nangles, double maskradius, float* best_out,
                    int32_t* idx_out, float* top1_out, float* top2_out, float* sigma_out) {
  return pick<float>(projft, p, oversample, expft, ctf, ctf_per_volume, n, nvol, angles,
                     nangles, maskradius, best_out, idx_out, top1_out, top2_out, sigma_out);
}

void orc_sphere_mask_ex_f64(double* out, int nx, int ny, int nz, double radius, double sigma, const float* c,
                            int decentered) {
  sphere_mask_ex(out, nx, ny, nz, radius, sigma, c, decentered != 0);
}

void orc_fft_resize_f64(int op, const double* in, double* out, const int* od, const int* nd, int ncomp) {
  fft_resize(op, in, out, od, nd, ncomp);
}

void orc_norm_monolithic_f64(const float* in, double* out, long long n, int batch, int complex_values) {
  for (int b = 0; b < batch; ++b) {
    if (complex_values)
      norm_ft_monolithic(in + (size_t)b * 2 * n, out + (size_t)b * 2 * n, n);
    else
      norm_monolithic(in + (size_t)b * n, out + (size_t)b * n, n);
  }
}

void orc_rotate3d_ft_f64(const double* in, double* out, int n, const double* M, int zerocentred, int ncomp,
                         int cubic) {
  rotate3d_ft(in, out, n, M, zerocentred != 0, ncomp, cubic != 0);
}

void orc_ctf_tilt_zgrid_f64(const double* angles, double defocus, double pixelsize, int imx, int imy, int rx,
                            int ry, const int* pos, int n, double* z) {
  tilt_zgrid(angles, defocus, pixelsize, imx, imy, rx, ry, pos, n, z);
}

// d_Peak on `batch` float maps of (nx, ny, nz): mode 1 integer, 2 sub-coarse,
// 3 sub-fine. pos receives (x, y, z) per map, value one double per map.
void orc_peak_f64(const float* in, int nx, int ny, int nz, int mode, int batch, double* pos, double* value) {
  const long long V = (long long)nx * ny * nz;
  for (int b = 0; b < batch; ++b) {
    const float* m = in + (size_t)b * V;
    PeakResult r = mode == 1 ? peak_integer(m, nx, ny, nz)
                 : mode == 2 ? peak_subcoarse(m, nx, ny, nz)
                             : peak_subfine(m, nx, ny, nz);
    for (int d = 0; d < 3; ++d) pos[3 * b + d] = r.pos[d];
    value[b] = r.value;
  }
}

void orc_peak_one_2d_f64(const float* in, int nx, int ny, int rx, int ry, int subtractcenter, int batch,
                         double* pos, double* value) {
  for (int b = 0; b < batch; ++b)
    peak_one_2d(in + (size_t)b * nx * ny, nx, ny, rx, ry, subtractcenter != 0, pos + 3 * b, value + b);
}

void orc_subpixel_max_f64(const float* in, double* out, int nx, int ny, int nz, int steps) {
  subpixel_max(in, out, nx, ny, nz, steps);
}

void orc_rotate2d_ft_f64(const double* in, double* out, int nx, int ny, double angle, double maxfreq,
                         int zerocentred, int cubic) {
  rotate2d_ft(in, out, nx, ny, angle, maxfreq, zerocentred != 0, cubic != 0);
}

void orc_prob_picker_f64(const double* ref, const double* mask, const int* dimsref, const int* dimsimage,
                         int lowpass, int doctf, const double* image, const double* ctf, float anglestep, int refid,
                         double* bestccf, double* bestangle, int* bestref, double* top2, int circular) {
  prob_picker(ref, mask, dimsref, dimsimage, lowpass, doctf != 0, image, ctf, anglestep, refid, bestccf, bestangle,
              bestref, top2, circular != 0);
}

void orc_bspline_prefilter_f64(double* v, int nx, int ny, int nz) { bspline_prefilter(v, nx, ny, nz); }

void orc_rotate_cubic_f64(const double* coef, double* out, int nx, int ny, int nz, const double* M) {
  rotate_cubic(coef, out, nx, ny, nz, M);
}

// The oracle's own gaussrand stream: srand(seed), reset of its phase, n draws.
void orc_gaussrand_f32(unsigned seed, int reset, float* out, int n) {
  std::srand(seed);
  if (reset) g_orc_gauss = OrcGaussRand();
  for (int i = 0; i < n; ++i) out[i] = g_orc_gauss.next();
}

}  // extern "C"
