// Peak extraction from a correlation map: d_LocalPeaks
// (src/Correlation/LocalPeaks.cu:24-62, LocalPeaksKernel :98-143). A voxel is
// a peak when its value is not below the threshold and no other voxel within
// the Euclidean ball of radius localextent (clamped at the volume faces, no
// wrap) holds a strictly larger value.
//
// Two kernels per volume instead of the reference's one thread per voxel per
// z slice (which scans the whole ball from every voxel above threshold):
//  k_peak_candidates: one thread per voxel applies the threshold and the
//     neighbours of the 3x3x3 block that lie in the ball; survivors are
//     appended to a list with one atomic per wave;
//  k_peak_verify: one wave per candidate scans the rest of the ball, 64
//     offsets per step, nearest shells first, and stops at the first strictly
//     larger neighbour.
// The host sorts the surviving linear indices, which is the reference's
// raster order (its host scan over z, y, x at :47-51).
#include <algorithm>
#include <vector>

#include "runtime.hpp"

namespace gx {
namespace {

__global__ __launch_bounds__(256) void k_peak_candidates(const float* __restrict__ in, int nx, int ny, int nz,
                                                         int r2, float thr, int* __restrict__ cand,
                                                         unsigned* __restrict__ count) {
  const int V = nx * ny * nz;
  const int i = blockIdx.x * 256 + threadIdx.x;
  bool is = false;
  if (i < V) {
    const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    const float v = in[i];
    is = !(v < thr);
    if (is) {
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const int d2 = dx * dx + dy * dy + dz * dz;
            const int xx = x + dx, yy = y + dy, zz = z + dz;
            if (d2 == 0 || d2 > r2 || (unsigned)xx >= (unsigned)nx || (unsigned)yy >= (unsigned)ny ||
                (unsigned)zz >= (unsigned)nz)
              continue;
            if (v < in[(zz * ny + yy) * nx + xx]) is = false;
          }
    }
  }
  const unsigned long long mask = __ballot(is);
  if (mask) {
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(mask));
    base = __shfl(base, leader);
    if (is) cand[base + __popcll(mask & ((1ull << lane) - 1ull))] = i;
  }
}

// Offsets (dx, dy, dz) of the ball outside the 3x3x3 block, nearest first.
__global__ __launch_bounds__(256) void k_peak_verify(const float* __restrict__ in, int nx, int ny, int nz,
                                                     const short4* __restrict__ offs, int noffs,
                                                     int* __restrict__ cand, const unsigned* __restrict__ count) {
  const int lane = threadIdx.x % 64;
  const unsigned ncand = *count;
  const unsigned waves = gridDim.x * (256 / 64);
  for (unsigned c = blockIdx.x * (256 / 64) + threadIdx.x / 64; c < ncand; c += waves) {
    const int i = cand[c];
    const int x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    const float v = in[i];
    bool rejected = false;
    for (int k0 = 0; k0 < noffs; k0 += 64) {
      const int k = k0 + lane;
      bool bad = false;
      if (k < noffs) {
        const short4 o = offs[k];
        const int xx = x + o.x, yy = y + o.y, zz = z + o.z;
        if ((unsigned)xx < (unsigned)nx && (unsigned)yy < (unsigned)ny && (unsigned)zz < (unsigned)nz)
          bad = v < in[(zz * ny + yy) * nx + xx];
      }
      if (__any(bad)) {
        rejected = true;
        break;
      }
    }
    if (rejected && lane == 0) cand[c] = -1;
  }
}

constexpr size_t MAX_OFFSETS = (size_t)1 << 26;

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" int gtomx_local_peaks(const float* d_input, gtomx_int3 dims, int localextent, float threshold,
                                 unsigned batch, gtomx_int3* h_peaks, size_t capacity, size_t* h_counts,
                                 void* stream) {
  if (!d_input || !h_counts) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (capacity > 0 && !h_peaks) return fail(GTOMX_ERR_INVALID_ARGUMENT, "h_peaks is NULL");
  if (dims.x < 1 || dims.y < 1 || dims.z < 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "dims must be positive");
  if (batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch must be positive");
  if (localextent < 0 || localextent > 46340) return fail(GTOMX_ERR_INVALID_ARGUMENT, "localextent out of range");
  const size_t V = elements(dims);
  if (V >= ((size_t)1 << 31)) return fail(GTOMX_ERR_UNSUPPORTED, "volume exceeds 2^31 voxels");
  hipStream_t s = as_stream(stream);

  // Offsets of the ball beyond the 3x3x3 block, limited to what fits the volume.
  const long long r2 = (long long)localextent * localextent;
  const int ex = std::min(localextent, dims.x - 1), ey = std::min(localextent, dims.y - 1),
            ez = std::min(localextent, dims.z - 1);
  std::vector<std::pair<long long, short4>> ball;
  for (int dz = -ez; dz <= ez; ++dz)
    for (int dy = -ey; dy <= ey; ++dy)
      for (int dx = -ex; dx <= ex; ++dx) {
        const long long d2 = (long long)dx * dx + (long long)dy * dy + (long long)dz * dz;
        if (d2 > r2 || (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1)) continue;
        if (ball.size() >= MAX_OFFSETS) return fail(GTOMX_ERR_UNSUPPORTED, "localextent too large");
        ball.push_back({d2, make_short4((short)dx, (short)dy, (short)dz, 0)});
      }
  std::stable_sort(ball.begin(), ball.end(),
                   [](const auto& a, const auto& b) { return a.first < b.first; });
  std::vector<short4> offs(ball.size());
  for (size_t k = 0; k < ball.size(); ++k) offs[k] = ball[k].second;

  DevBuf d_offs, d_cand, d_count;
  GX_TRY(d_offs.alloc(std::max<size_t>(1, offs.size()) * sizeof(short4)));
  GX_TRY(d_cand.alloc(V * sizeof(int)));
  GX_TRY(d_count.alloc(sizeof(unsigned)));
  if (!offs.empty())
    GX_HIP(hipMemcpyAsync(d_offs.p, offs.data(), offs.size() * sizeof(short4), hipMemcpyHostToDevice, s));
  const int r2_block = (int)std::min<long long>(r2, 3);  // the 3x3x3 block holds squared distances <= 3
  std::vector<int> lin;
  for (unsigned b = 0; b < batch; ++b) {
    const float* in = d_input + (size_t)b * V;
    GX_HIP(hipMemsetAsync(d_count.p, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_peak_candidates, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, in, dims.x, dims.y,
                       dims.z, r2_block, threshold, static_cast<int*>(d_cand.p), static_cast<unsigned*>(d_count.p));
    GX_LAUNCHED("k_peak_candidates");
    if (!offs.empty()) {
      hipLaunchKernelGGL(k_peak_verify, dim3(1024), dim3(256), 0, s, in, dims.x, dims.y, dims.z,
                         static_cast<const short4*>(d_offs.p), (int)offs.size(), static_cast<int*>(d_cand.p),
                         static_cast<const unsigned*>(d_count.p));
      GX_LAUNCHED("k_peak_verify");
    }
    unsigned ncand = 0;
    GX_HIP(hipMemcpyAsync(&ncand, d_count.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    GX_HIP(hipStreamSynchronize(s));
    lin.resize(ncand);
    if (ncand) {
      GX_HIP(hipMemcpyAsync(lin.data(), d_cand.p, (size_t)ncand * sizeof(int), hipMemcpyDeviceToHost, s));
      GX_HIP(hipStreamSynchronize(s));
    }
    lin.erase(std::remove(lin.begin(), lin.end(), -1), lin.end());
    std::sort(lin.begin(), lin.end());
    h_counts[b] = lin.size();
    const size_t k = std::min(capacity, lin.size());
    for (size_t q = 0; q < k; ++q) {
      const int i = lin[q];
      h_peaks[(size_t)b * capacity + q] = gtomx_int3{i % dims.x, (i / dims.x) % dims.y, i / (dims.x * dims.y)};
    }
  }
  return GTOMX_OK;
}
