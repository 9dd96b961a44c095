// Exercises the header-only gtom:: drop-in (include/gtom_pick.hpp) from C++,
// the way a host program written against the reference API would call it.
// Prints one line per case and exits non-zero on the first failure.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#include "gtom_pick.hpp"

using namespace gtom;

#define EXPECT(cond, ...)                          \
  do {                                             \
    if (!(cond)) {                                 \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                    \
      std::printf("\n");                           \
      std::exit(1);                                \
    }                                              \
  } while (0)

template <class T> T* dev(size_t n) {
  T* p = nullptr;
  EXPECT(hipMalloc((void**)&p, n * sizeof(T)) == hipSuccess, "hipMalloc");
  return p;
}

static void fft_roundtrip() {
  int3 dims = make_int3(32, 16, 8);
  size_t V = 32 * 16 * 8, H = 17 * 16 * 8;
  std::vector<float> h(V);
  std::mt19937 rng(1);
  std::normal_distribution<float> nd;
  for (auto& v : h) v = nd(rng);
  float* d_x = dev<float>(V);
  float* d_y = dev<float>(V);
  tcomplex* d_X = dev<tcomplex>(H);
  hipMemcpy(d_x, h.data(), V * sizeof(float), hipMemcpyHostToDevice);
  d_FFTR2C(d_x, d_X, 3, dims);
  d_IFFTC2R(d_X, d_y, 3, dims);  // renormalised
  std::vector<float> back(V);
  hipMemcpy(back.data(), d_y, V * sizeof(float), hipMemcpyDeviceToHost);
  double err = 0;
  for (size_t i = 0; i < V; ++i) err = std::max(err, (double)std::fabs(back[i] - h[i]));
  EXPECT(err < 1e-5, "round trip error %g", err);
  std::printf("ok fft_roundtrip (max error %.2e)\n", err);
  hipFree(d_x); hipFree(d_y); hipFree(d_X);
}

// The centred template moved by a known circular shift s is found at its
// centre n/2 + s with the identity orientation (index 0 of the 30-degree list).
static void pick_finds_shift() {
  const int n = 32, box = 16, p = 32;
  std::vector<float> tmpl((size_t)box * box * box);
  for (int z = 0; z < box; ++z)
    for (int y = 0; y < box; ++y)
      for (int x = 0; x < box; ++x) {
        double dx = x - box / 2 - 1.5, dy = y - box / 2 + 1.0, dz = z - box / 2;
        tmpl[((size_t)z * box + y) * box + x] =
            (float)(std::exp(-(dx * dx / 4 + dy * dy / 9 + dz * dz / 2.5) / 2) +
                    0.5 * std::exp(-((dx - 3) * (dx - 3) + dy * dy + (dz + 2) * (dz + 2)) / 3));
      }
  float* d_t = dev<float>(tmpl.size());
  hipMemcpy(d_t, tmpl.data(), tmpl.size() * sizeof(float), hipMemcpyHostToDevice);
  float* d_pad = dev<float>((size_t)p * p * p);
  float* d_shift = dev<float>((size_t)p * p * p);
  tcomplex* d_proj = dev<tcomplex>((size_t)(p / 2 + 1) * p * p);
  d_Pad(d_t, d_pad, make_int3(box, box, box), make_int3(p, p, p), 0.f);
  d_RemapFull2FullFFT(d_pad, d_shift, make_int3(p, p, p));
  d_FFTR2C(d_shift, d_proj, 3, make_int3(p, p, p));

  // tomogram: the padded (centred) template moved by (sx, sy, sz)
  std::vector<float> centred((size_t)p * p * p), tomo((size_t)n * n * n);
  hipMemcpy(centred.data(), d_pad, centred.size() * sizeof(float), hipMemcpyDeviceToHost);
  const int sx = 5, sy = -3, sz = 7;
  // Noise keeps the local standard deviation away from zero everywhere, so the
  // normalised map peaks where the correlation does.
  std::mt19937 rng(7);
  std::normal_distribution<float> noise(0.f, 0.05f);
  for (int z = 0; z < n; ++z)
    for (int y = 0; y < n; ++y)
      for (int x = 0; x < n; ++x)
        tomo[((size_t)((z + sz + n) % n) * n + (y + sy + n) % n) * n + (x + sx + n) % n] =
            centred[((size_t)z * n + y) * n + x] + noise(rng);
  float* d_tomo = dev<float>(tomo.size());
  hipMemcpy(d_tomo, tomo.data(), tomo.size() * sizeof(float), hipMemcpyHostToDevice);
  tcomplex* d_E = dev<tcomplex>((size_t)(n / 2 + 1) * n * n);
  d_FFTR2C(d_tomo, d_E, 3, make_int3(n, n, n));
  std::vector<float> ctf((size_t)(n / 2 + 1) * n * n, 1.f);
  float* d_ctf = dev<float>(ctf.size());
  hipMemcpy(d_ctf, ctf.data(), ctf.size() * sizeof(float), hipMemcpyHostToDevice);

  const float pi = 3.14159265358979f;
  std::vector<float3> angles = GetEqualAngularSpacing(make_float2(0, 2 * pi), make_float2(0, pi),
                                                      make_float2(0, 2 * pi), 30.f * pi / 180.f);
  EXPECT(angles.size() == 588, "angle count %zu", angles.size());
  float* d_best = dev<float>((size_t)n * n * n);
  float* d_angle = dev<float>((size_t)n * n * n);
  d_PickSubTomograms(d_proj, (float)p / n, make_int3(p, p, p), d_E, d_ctf, make_int3(n, n, n), 1, angles.data(),
                     (uint)angles.size(), 6.f, d_best, d_angle);
  std::vector<float> best((size_t)n * n * n), ang((size_t)n * n * n);
  hipMemcpy(best.data(), d_best, best.size() * sizeof(float), hipMemcpyDeviceToHost);
  hipMemcpy(ang.data(), d_angle, ang.size() * sizeof(float), hipMemcpyDeviceToHost);
  size_t arg = 0;
  for (size_t i = 1; i < best.size(); ++i)
    if (best[i] > best[arg]) arg = i;
  int ax = (int)(arg % n), ay = (int)((arg / n) % n), az = (int)(arg / ((size_t)n * n));
  EXPECT(ax == (n / 2 + sx + n) % n && ay == (n / 2 + sy + n) % n && az == (n / 2 + sz + n) % n,
         "peak at (%d,%d,%d)", ax, ay, az);
  EXPECT(ang[arg] == 0.f, "orientation index %g", ang[arg]);
  std::printf("ok pick_finds_shift (peak %.3f at (%d,%d,%d), orientation %g)\n", best[arg], ax, ay, az, ang[arg]);

  // the picked-particle list: the planted position is a local peak above half the maximum
  int3* peaks[1] = {nullptr};
  int npeaks[1] = {0};
  d_LocalPeaks(d_best, peaks, npeaks, make_int3(n, n, n), 4, 0.5f * best[arg]);
  bool found = false;
  for (int k = 0; k < npeaks[0]; ++k)
    found |= peaks[0][k].x == ax && peaks[0][k].y == ay && peaks[0][k].z == az;
  EXPECT(npeaks[0] >= 1 && found, "%d peaks, planted position not among them", npeaks[0]);
  std::printf("ok local_peaks (%d peaks above half the maximum)\n", npeaks[0]);

  // the same maximum through d_Peak: integer and sub-voxel positions agree to a voxel
  tfloat3* d_pos = dev<tfloat3>(1);
  float* d_val = dev<float>(1);
  for (T_PEAK_MODE mode : {T_PEAK_INTEGER, T_PEAK_SUBCOARSE, T_PEAK_SUBFINE}) {
    d_Peak(d_best, d_pos, d_val, make_int3(n, n, n), mode);
    tfloat3 hp;
    float hv;
    hipMemcpy(&hp, d_pos, sizeof(hp), hipMemcpyDeviceToHost);
    hipMemcpy(&hv, d_val, sizeof(hv), hipMemcpyDeviceToHost);
    EXPECT(std::fabs(hp.x - ax) <= 1.f && std::fabs(hp.y - ay) <= 1.f && std::fabs(hp.z - az) <= 1.f,
           "d_Peak mode %d at (%g,%g,%g)", (int)mode, hp.x, hp.y, hp.z);
    EXPECT(mode != T_PEAK_INTEGER || (hp.x == ax && hp.y == ay && hp.z == az && hv == best[arg]),
           "integer peak (%g,%g,%g) %g", hp.x, hp.y, hp.z, hv);
  }
  std::printf("ok d_Peak (integer, sub-coarse, sub-fine)\n");
  hipFree(d_pos); hipFree(d_val);
  std::free(peaks[0]);
  hipFree(d_t); hipFree(d_pad); hipFree(d_shift); hipFree(d_proj); hipFree(d_tomo); hipFree(d_E);
  hipFree(d_ctf); hipFree(d_best); hipFree(d_angle);
}

// The Picker class: a copy of the reference, shifted by a known amount, is
// found at that shift with the identity orientation (the first of the list).
static void picker_finds_copy() {
  const int nr = 16, n = 32;
  std::vector<float> ref((size_t)nr * nr * nr);
  for (int z = 0; z < nr; ++z)
    for (int y = 0; y < nr; ++y)
      for (int x = 0; x < nr; ++x) {
        double dx = x - nr / 2 - 1.0, dy = y - nr / 2 + 2.0, dz = z - nr / 2;
        ref[((size_t)z * nr + y) * nr + x] = (float)(std::exp(-(dx * dx / 3 + dy * dy / 6 + dz * dz / 2) / 2) +
                                                     0.7 * std::exp(-((dx - 3) * (dx - 3) + dy * dy + dz * dz) / 2));
      }
  float* d_ref = dev<float>(ref.size());
  hipMemcpy(d_ref, ref.data(), ref.size() * sizeof(float), hipMemcpyHostToDevice);
  // image: the reference padded to n around its centre, moved by (sx, sy, sz), plus noise
  std::vector<float> image((size_t)n * n * n, 0.f);
  const int sx = 3, sy = -4, sz = 5;
  std::mt19937 rng(9);
  std::normal_distribution<float> noise(0.f, 0.02f);
  for (auto& v : image) v = noise(rng);
  for (int z = 0; z < nr; ++z)
    for (int y = 0; y < nr; ++y)
      for (int x = 0; x < nr; ++x) {
        int X = (x - nr / 2 + sx + n) % n, Y = (y - nr / 2 + sy + n) % n, Z = (z - nr / 2 + sz + n) % n;
        image[((size_t)Z * n + Y) * n + X] += ref[((size_t)z * nr + y) * nr + x];
      }
  float* d_image = dev<float>(image.size());
  hipMemcpy(d_image, image.data(), image.size() * sizeof(float), hipMemcpyHostToDevice);
  std::vector<float> ctf((size_t)(nr / 2 + 1) * nr * nr, 1.f);
  float* d_ctf = dev<float>(ctf.size());
  hipMemcpy(d_ctf, ctf.data(), ctf.size() * sizeof(float), hipMemcpyHostToDevice);
  std::vector<float> best0(image.size(), -1e30f);
  float* d_best = dev<float>(image.size());
  hipMemcpy(d_best, best0.data(), best0.size() * sizeof(float), hipMemcpyHostToDevice);
  tfloat3* d_angle = dev<tfloat3>(image.size());
  hipMemset(d_angle, 0, image.size() * sizeof(tfloat3));

  Picker picker;
  picker.Initialize(d_ref, make_int3(nr, nr, nr), nullptr, make_int3(n, n, n));
  picker.SetImage(d_image, d_ctf);
  picker.PerformCorrelation(90.f * 3.14159265f / 180.f, d_best, d_angle);
  std::vector<float> best(image.size());
  std::vector<tfloat3> ang(image.size());
  hipMemcpy(best.data(), d_best, best.size() * sizeof(float), hipMemcpyDeviceToHost);
  hipMemcpy(ang.data(), d_angle, ang.size() * sizeof(tfloat3), hipMemcpyDeviceToHost);
  size_t arg = 0;
  for (size_t i = 1; i < best.size(); ++i)
    if (best[i] > best[arg]) arg = i;
  int ax = (int)(arg % n), ay = (int)((arg / n) % n), az = (int)(arg / ((size_t)n * n));
  EXPECT(ax == (sx + n) % n && ay == (sy + n) % n && az == (sz + n) % n, "Picker peak at (%d,%d,%d)", ax, ay, az);
  EXPECT(ang[arg].x == 0.f && ang[arg].y == 0.f && ang[arg].z == 0.f, "Picker angle (%g,%g,%g)", ang[arg].x,
         ang[arg].y, ang[arg].z);
  std::printf("ok picker_finds_copy (peak %.1f at (%d,%d,%d))\n", best[arg], ax, ay, az);
  hipFree(d_ref); hipFree(d_image); hipFree(d_ctf); hipFree(d_best); hipFree(d_angle);
}

// ProbabilisticPicker: the class passes its arguments through unchanged, so two
// references correlated through it equal the C entry points called directly
// (bit for bit: rand() is reseeded by every call).
static void prob_picker_matches_c_abi() {
  const int n = 48, nr = 12, nrefs = 2, lp = 6;
  const size_t V = (size_t)n * n, E = (size_t)nr * nr;
  std::mt19937 rng(31);
  std::normal_distribution<float> nd(0.f, 1.f);
  std::vector<float> refs(E * nrefs), masks(E * nrefs), image(V);
  for (auto& v : refs) v = nd(rng);
  for (int r = 0; r < nrefs; ++r)
    for (int y = 0; y < nr; ++y)
      for (int x = 0; x < nr; ++x)
        masks[r * E + y * nr + x] = std::hypot(x - nr / 2.0, y - nr / 2.0) < 5.0 ? 1.f : 0.f;
  for (auto& v : image) v = nd(rng);
  float* d_refs = dev<float>(refs.size());
  float* d_masks = dev<float>(masks.size());
  float* d_image = dev<float>(V);
  hipMemcpy(d_refs, refs.data(), refs.size() * sizeof(float), hipMemcpyHostToDevice);
  hipMemcpy(d_masks, masks.data(), masks.size() * sizeof(float), hipMemcpyHostToDevice);
  hipMemcpy(d_image, image.data(), V * sizeof(float), hipMemcpyHostToDevice);
  std::vector<float> init(V, -1e30f);
  float* best[2];
  tfloat3* ang[2];
  int* ref[2];
  for (int k = 0; k < 2; ++k) {
    best[k] = dev<float>(V);
    ang[k] = dev<tfloat3>(V);
    ref[k] = dev<int>(V);
    hipMemcpy(best[k], init.data(), V * sizeof(float), hipMemcpyHostToDevice);
    hipMemset(ang[k], 0, V * sizeof(tfloat3));
    hipMemset(ref[k], 0xff, V * sizeof(int));
  }
  {
    ProbabilisticPicker pp;
    pp.Initialize(d_refs, make_int3(nr, nr, 1), nrefs, d_masks, true, false, make_int3(n, n, 1), lp);
    EXPECT(pp.dimslowpass.x == 2 * lp && pp.dimslowpass.z == 1 && pp.ndims == 2, "ProbabilisticPicker dims");
    pp.SetImage(d_image, nullptr);
    for (uint r = 0; r < (uint)nrefs; ++r) pp.PerformCorrelation(r, 0.6f, best[0], ang[0], ref[0]);
  }
  gtomx_prob_picker* h = nullptr;
  EXPECT(gtomx_prob_picker_create(&h, d_refs, gtomx_int3{nr, nr, 1}, nrefs, d_masks, 1, 0, gtomx_int3{n, n, 1}, lp,
                                  nullptr) == GTOMX_OK, "gtomx_prob_picker_create: %s", gtomx_last_error());
  EXPECT(gtomx_prob_picker_set_image(h, d_image, nullptr, nullptr) == GTOMX_OK, "set_image");
  for (unsigned r = 0; r < (unsigned)nrefs; ++r)
    EXPECT(gtomx_prob_picker_correlate(h, r, 0.6f, best[1], reinterpret_cast<gtomx_float3*>(ang[1]), ref[1],
                                       nullptr) == GTOMX_OK, "correlate");
  gtomx_prob_picker_destroy(h);
  std::vector<float> b0(V), b1(V);
  std::vector<int> r0(V), r1(V);
  std::vector<tfloat3> a0(V), a1(V);
  hipMemcpy(b0.data(), best[0], V * sizeof(float), hipMemcpyDeviceToHost);
  hipMemcpy(b1.data(), best[1], V * sizeof(float), hipMemcpyDeviceToHost);
  hipMemcpy(r0.data(), ref[0], V * sizeof(int), hipMemcpyDeviceToHost);
  hipMemcpy(r1.data(), ref[1], V * sizeof(int), hipMemcpyDeviceToHost);
  hipMemcpy(a0.data(), ang[0], V * sizeof(tfloat3), hipMemcpyDeviceToHost);
  hipMemcpy(a1.data(), ang[1], V * sizeof(tfloat3), hipMemcpyDeviceToHost);
  EXPECT(std::memcmp(b0.data(), b1.data(), V * sizeof(float)) == 0, "ProbabilisticPicker scores differ");
  EXPECT(r0 == r1, "ProbabilisticPicker reference indices differ");
  EXPECT(std::memcmp(a0.data(), a1.data(), V * sizeof(tfloat3)) == 0, "ProbabilisticPicker angles differ");
  int seen[2] = {0, 0};
  for (int v : r0) {
    EXPECT(v == 0 || v == 1, "reference index %d", v);
    seen[v]++;
  }
  EXPECT(seen[0] > 0 && seen[1] > 0, "both references win somewhere (%d, %d)", seen[0], seen[1]);
  {
    // a non-circular mask: rotated per orientation; every voxel is scored
    ProbabilisticPicker nc;
    nc.Initialize(d_refs, make_int3(nr, nr, 1), 1, d_masks, false, false, make_int3(n, n, 1), lp);
    nc.SetImage(d_image, nullptr);
    hipMemcpy(best[0], init.data(), V * sizeof(float), hipMemcpyHostToDevice);
    hipMemset(ref[0], 0xff, V * sizeof(int));
    nc.PerformCorrelation(0, 0.6f, best[0], ang[0], ref[0]);
    hipMemcpy(b0.data(), best[0], V * sizeof(float), hipMemcpyDeviceToHost);
    hipMemcpy(r0.data(), ref[0], V * sizeof(int), hipMemcpyDeviceToHost);
    for (size_t i = 0; i < V; ++i) EXPECT(std::isfinite(b0[i]) && b0[i] > -1e30f && r0[i] == 0, "non-circular voxel %zu", i);
  }
  std::printf("ok ProbabilisticPicker (class = C ABI; references won %d / %d; non-circular mask runs)\n", seen[0],
              seen[1]);
  for (int k = 0; k < 2; ++k) {
    hipFree(best[k]); hipFree(ang[k]); hipFree(ref[k]);
  }
  hipFree(d_refs); hipFree(d_masks); hipFree(d_image);
}

// MRC: a device volume written with d_WriteMRC reads back with ReadMRC, and
// MixedToDeviceTfloat brings it to the device unchanged.
static void mrc_round_trip() {
  const int3 dims = make_int3(12, 10, 8);
  const size_t n = 12 * 10 * 8;
  std::vector<float> h(n);
  for (size_t i = 0; i < n; ++i) h[i] = (float)std::sin(0.37 * i) * 3.f;
  float* d = dev<float>(n);
  hipMemcpy(d, h.data(), n * sizeof(float), hipMemcpyHostToDevice);
  const std::string path = "gtomx_shim_test.mrc";
  d_WriteMRC(d, dims, path);
  HeaderMRC hdr = ReadMRCHeader(path);
  EXPECT(hdr.dimensions.x == 12 && hdr.dimensions.y == 10 && hdr.dimensions.z == 8 && hdr.mode == MRC_FLOAT,
         "header %d %d %d mode %d", hdr.dimensions.x, hdr.dimensions.y, hdr.dimensions.z, (int)hdr.mode);
  void* raw = nullptr;
  ReadMRC(path, &raw);
  float* d2 = dev<float>(n);
  MixedToDeviceTfloat(raw, d2, hdr.mode, n);
  std::vector<float> back(n);
  hipMemcpy(back.data(), d2, n * sizeof(float), hipMemcpyDeviceToHost);
  for (size_t i = 0; i < n; ++i) EXPECT(back[i] == h[i], "voxel %zu: %g vs %g", i, back[i], h[i]);
  EXPECT(hdr.minvalue <= -2.9f && hdr.maxvalue >= 2.9f, "min/max %g %g", hdr.minvalue, hdr.maxvalue);
  hipHostFree(raw);
  std::remove(path.c_str());
  hipFree(d); hipFree(d2);
  std::printf("ok mrc_round_trip\n");
}

// d_CTFSimulate: CTF(0) equals the amplitude contrast, and along the
// astigmatism angle an astigmatic CTF equals the symmetric one at
// defocus + defocusdelta/2 (two parameter sets, one call, shared addresses).
static void ctf_simulate_known_answers() {
  CTFParams p[2];
  p[0].defocusdelta = 0.4e-6f;
  p[0].astigmatismangle = 0.3f;
  p[1].defocus = p[0].defocus + 0.2e-6f;
  const int n = 6;
  std::vector<float2> addr(n);
  for (int i = 0; i < n; ++i) addr[i] = make_float2(0.08f * i, 0.3f);
  float2* d_addr = dev<float2>(n);
  float* d_out = dev<float>(2 * n);
  hipMemcpy(d_addr, addr.data(), n * sizeof(float2), hipMemcpyHostToDevice);
  d_CTFSimulate(p, d_addr, nullptr, d_out, n, false, false, 2);
  std::vector<float> out(2 * n);
  hipMemcpy(out.data(), d_out, out.size() * sizeof(float), hipMemcpyDeviceToHost);
  EXPECT(std::fabs(out[0] - 0.07f) < 1e-6f && std::fabs(out[n] - 0.07f) < 1e-6f, "CTF(0) = %g, %g", out[0], out[n]);
  for (int i = 0; i < n; ++i) EXPECT(std::fabs(out[i] - out[n + i]) < 1e-4f, "k %d: %g vs %g", i, out[i], out[n + i]);
  std::printf("ok ctf_simulate_known_answers\n");

  // a stage tilt of 0.5 rad about x: defocus changes by tan(0.5) * y offset
  CTFTiltParams tilt(make_float3(0.f, 0.5f, 0.f), CTFParams());
  int3 origins[2] = {make_int3(448, 448, 0), make_int3(448, 0, 0)};
  float z[2];
  tilt.GetZGrid2D(make_int2(1024, 1024), make_int2(128, 128), origins, 2, z);
  EXPECT(std::fabs(z[0] - p[0].defocus) < 1e-12f, "centre defocus %g", z[0]);
  EXPECT(std::fabs(z[1] - (p[0].defocus - 448e-10f * std::tan(0.5f))) < 1e-12f, "tilted defocus %g", z[1]);
  std::printf("ok ctf_tilt_grid\n");
  hipFree(d_addr); hipFree(d_out);
}

static void errors_throw() {
  bool threw = false;
  try {
    d_FFTR2C(nullptr, nullptr, 3, make_int3(100, 100, 100));
  } catch (const std::runtime_error& e) {
    threw = true;
  }
  EXPECT(threw, "unsupported size did not throw");
  std::printf("ok errors_throw\n");
}

// The TomoPicker class: one untilted tilt image holding the reference's
// projection along z at a known pixel; the voxels above that pixel hold the
// best correlation (near 1) with the identity orientation.
static void tomopicker_finds_projection() {
  const int nr = 16, n = 64;
  std::vector<float> ref((size_t)nr * nr * nr), mask(ref.size(), 1.f);
  for (int z = 0; z < nr; ++z)
    for (int y = 0; y < nr; ++y)
      for (int x = 0; x < nr; ++x) {
        double dx = x - nr / 2 - 1.0, dy = y - nr / 2 + 2.0, dz = z - nr / 2;
        ref[((size_t)z * nr + y) * nr + x] = (float)(std::exp(-(dx * dx / 3 + dy * dy / 6 + dz * dz / 2) / 2) +
                                                     0.7 * std::exp(-((dx - 3) * (dx - 3) + dy * dy + dz * dz) / 2));
      }
  std::vector<float> image((size_t)n * n, 0.f);
  const int px = 37, py = 22;
  std::mt19937 rng(11);
  std::normal_distribution<float> noise(0.f, 0.05f);
  for (auto& v : image) v = noise(rng);
  for (int z = 0; z < nr; ++z)
    for (int y = 0; y < nr; ++y)
      for (int x = 0; x < nr; ++x)
        image[(size_t)(y - nr / 2 + py) * n + (x - nr / 2 + px)] += ref[((size_t)z * nr + y) * nr + x];
  float* d_ref = dev<float>(ref.size());
  float* d_mask = dev<float>(mask.size());
  float* d_image = dev<float>(image.size());
  hipMemcpy(d_ref, ref.data(), ref.size() * sizeof(float), hipMemcpyHostToDevice);
  hipMemcpy(d_mask, mask.data(), mask.size() * sizeof(float), hipMemcpyHostToDevice);
  hipMemcpy(d_image, image.data(), image.size() * sizeof(float), hipMemcpyHostToDevice);
  const int cx = 32, cy = 32, cz = 2;
  std::vector<float> c0((size_t)cx * cy * cz, -1e30f);
  float* d_corr = dev<float>(c0.size());
  hipMemcpy(d_corr, c0.data(), c0.size() * sizeof(float), hipMemcpyHostToDevice);
  tfloat3* d_angle = dev<tfloat3>(c0.size());
  hipMemset(d_angle, 0xff, c0.size() * sizeof(tfloat3));
  tfloat3 tilt{0.f, 0.f, 0.f};
  tfloat weight = 1.f;

  TomoPicker tp;
  tp.Initialize(d_ref, make_int3(nr, nr, nr), d_mask, true, int2{n, n}, 1);
  EXPECT(tp.dimsrefpadded.x == 2 * nr && tp.nimages == 1, "TomoPicker dims");
  tp.SetImage(d_image, &tilt, &weight);
  tp.PerformCorrelation(d_corr, d_angle, make_int3(cx, cy, cz), 90.f * 3.14159265f / 180.f);
  std::vector<float> corr(c0.size());
  std::vector<tfloat3> ang(c0.size());
  hipMemcpy(corr.data(), d_corr, corr.size() * sizeof(float), hipMemcpyDeviceToHost);
  hipMemcpy(ang.data(), d_angle, ang.size() * sizeof(tfloat3), hipMemcpyDeviceToHost);
  const int vx = px + cx / 2 - n / 2, vy = py + cy / 2 - n / 2;
  for (int z = 0; z < cz; ++z) {
    size_t arg = (size_t)z * cx * cy;
    for (size_t i = arg; i < (size_t)(z + 1) * cx * cy; ++i)
      if (corr[i] > corr[arg]) arg = i;
    const int ax = (int)(arg % cx), ay = (int)((arg / cx) % cy);
    EXPECT(ax == vx && ay == vy && corr[arg] > 0.97f, "TomoPicker peak %.3f at (%d,%d) in z %d", corr[arg], ax, ay, z);
    EXPECT(ang[arg].x == 0.f && ang[arg].y == 0.f && ang[arg].z == 0.f, "TomoPicker angle");
  }
  std::printf("ok tomopicker_finds_projection\n");
  hipFree(d_ref); hipFree(d_mask); hipFree(d_image); hipFree(d_corr); hipFree(d_angle);
}

int main() {
  fft_roundtrip();
  pick_finds_shift();
  picker_finds_copy();
  tomopicker_finds_projection();
  prob_picker_matches_c_abi();
  mrc_round_trip();
  ctf_simulate_known_answers();
  errors_throw();
  std::printf("all passed\n");
  return 0;
}
