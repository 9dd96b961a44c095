// Host-side runtime shared by every entry point: status codes, the calling
// thread's last error message, checked device allocation (with a test hook
// that forces allocation failure), and per-device constant tables.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/gtomx.h"

namespace gx {

// Device twiddle tables (must match TW_LEN / TW3_LEN / TW3_OFF in dev_fft.hpp):
// exp(-2 pi i m/4096) for m < 4096, then exp(-2 pi i m/3072) for m < 3072.
constexpr int TW_LEN_HOST = 4096;
constexpr int TW3_LEN_HOST = 3072;

// Records a message for gtomx_last_error() and returns the status.
int fail(int status, const std::string& msg);
int ok();

// Maps a HIP error to GTOMX_ERR_DEVICE / GTOMX_ERR_OUT_OF_MEMORY with context.
int hip_fail(hipError_t e, const char* what);

#define GX_HIP(call)                                       \
  do {                                                     \
    hipError_t gx_e_ = (call);                             \
    if (gx_e_ != hipSuccess) return ::gx::hip_fail(gx_e_, #call); \
  } while (0)

// Checks hipGetLastError after a kernel launch.
#define GX_LAUNCHED(name)                                  \
  do {                                                     \
    hipError_t gx_e_ = hipGetLastError();                  \
    if (gx_e_ != hipSuccess) return ::gx::hip_fail(gx_e_, name); \
  } while (0)

#define GX_TRY(expr)              \
  do {                            \
    int gx_s_ = (expr);           \
    if (gx_s_ != GTOMX_OK) return gx_s_; \
  } while (0)

// Device allocation through the library's allocator (honours the test hook).
int dev_alloc(void** p, size_t bytes);
void dev_free(void* p);

// Twiddle table exp(-2 pi i m / 4096), m in [0, 4096), for the current device.
int twiddles(const float2** table);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
// Axis extents of the one-pass FFT kernels (and of the fused search kernels
// built on them): 2^a or 3 * 2^a in [8, 256].
inline bool fft_len_ok(int n) {
  return n >= 8 && n <= 256 && (is_pow2(n) || (n % 3 == 0 && is_pow2(n / 3)));
}
// Axis extents of the generic transforms: the above, plus 2^a up to 4096 and
// 3 * 2^a up to 3072 through the two-pass (four-step) decomposition.
inline bool fft_len_ok_any(int n) {
  return fft_len_ok(n) || (n > 256 && ((is_pow2(n) && n <= 4096) || (n % 3 == 0 && is_pow2(n / 3) && n <= 3072)));
}
// Row (x) extents of the native transforms: one pass up to 256, the long-row
// kernels at 512, 768, 1024, 1536, 2048, 3072 and 4096 (other x extents,
// 384 among them, take the chirp-z form).
inline bool row_len_native(int n) {
  return fft_len_ok(n) || n == 512 || n == 768 || n == 1024 || n == 1536 || n == 2048 || n == 3072 || n == 4096;
}
// Expands X(n) for every supported transform length.
#define GX_FOR_SIZES(X) X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(192) X(256)

// Validates dims for an ndims-dimensional transform; returns GTOMX_OK or an error.
int check_fft_dims(int ndims, gtomx_int3 dims);
// The same for the standalone transforms, which also take any extent up to
// 2048 that has no native transform (by the chirp-z form).
int check_fft_dims_any(int ndims, gtomx_int3 dims);

inline size_t elements(gtomx_int3 d) { return (size_t)d.x * d.y * d.z; }
inline size_t elements_ft(gtomx_int3 d) { return (size_t)(d.x / 2 + 1) * d.y * d.z; }

// Orders a stateful object's calls across streams. Every call that enqueues
// work on the object's buffers begins with enter(s): when the previous call
// ran on another stream, s waits for that call's work (an event, no host
// synchronisation). leave(s) records the event after the call's work, and
// drain() blocks the host until it has completed (before buffers are freed).
struct StreamFence {
  hipEvent_t ev = nullptr;
  hipStream_t last = nullptr;
  bool recorded = false;
  int enter(hipStream_t s) {
    if (recorded && s != last) {
      hipError_t e = hipStreamWaitEvent(s, ev, 0);
      if (e != hipSuccess) return hip_fail(e, "hipStreamWaitEvent");
    }
    return GTOMX_OK;
  }
  int leave(hipStream_t s) {
    if (!ev) {
      hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    }
    hipError_t e = hipEventRecord(ev, s);
    if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
    last = s;
    recorded = true;
    return GTOMX_OK;
  }
  void drain() {
    if (recorded) (void)hipEventSynchronize(ev);
    recorded = false;
  }
  void release() {
    drain();
    if (ev) (void)hipEventDestroy(ev);
    ev = nullptr;
  }
};

// Scratch device buffer that frees itself.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { dev_free(p); }
  int alloc(size_t b) {
    dev_free(p);
    p = nullptr;
    bytes = 0;
    int s = dev_alloc(&p, b);
    if (s == GTOMX_OK) bytes = b;
    return s;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// ---- internal host API shared between translation units ----
// Batched 3D/2D transforms (dims with dims.z == 1 for 2D). `work` must hold
// elements_ft(dims) * batch complex values for c2r (the input is copied
// there first so it is left unmodified); r2c needs no workspace.
// Extents without a native transform (at most 2048) go through the chirp-z form.
int fft_r2c(const float* in, float2* out, gtomx_int3 dims, unsigned batch, hipStream_t s);
// cols > 0 declares the columns x >= cols of every half spectrum zero: the y
// and z inverses skip them (the result is the same).
int fft_c2r(const float2* in, float2* work, float* out, gtomx_int3 dims, unsigned batch,
            float scale, hipStream_t s, int cols = 0);
// Axes above 256 go through two passes with a scratch copy of the data, taken
// from the stream's memory pool (hipMallocAsync) and returned on the stream.
// L = L1 * L2 as the two-pass transforms split it (L above 256).
int fft_len_split(int L, int* L1, int* L2);
// Pass 2 of the inverse y transform of nb n x n half spectra (n above 256),
// from the pass-1 layout in `scratch` (the column position n2 + L2 k1 holds
// pass 1's output k1 of line n2, twiddled) into natural order in `out`.
int image_cols_pass2_inverse(int n, const float2* scratch, float2* out, unsigned nb, hipStream_t s);
// Transforms of nb n x n images (n above 256) whose data is known to matter
// only in part, with results equal to the full transforms where they are used:
// C2R with the y inverse over the first `cols` columns only (the rest must be
// zero) and the row inverse only over the rows [0, band) and [n - band, n)
// (other rows of `out` keep their contents); R2C of input that is zero
// outside those rows (its rows between the bands are not read). band is a
// multiple of 8 and at most n / 2.
int image_c2r_banded(float2* work, float* out, int n, unsigned nb, int cols, int band, float scale, hipStream_t s);
int image_r2c_banded(const float* in, float2* out, int n, unsigned nb, int band, hipStream_t s);
// The same for volumes and images with every axis up to 256: rows y in
// [0, by) and [ny - by, ny) of the planes z in [0, bz) and [nz - bz, nz)
// (bz = 0: every plane); by a multiple of rows_per_workgroup(nx).
int rows_per_workgroup(int nx);
int volume_c2r_banded(float2* work, float* out, gtomx_int3 dims, unsigned nb, int cols, int by, int bz, float scale,
                      hipStream_t s);
int volume_r2c_banded(const float* in, float2* out, gtomx_int3 dims, unsigned nb, int by, int bz, hipStream_t s);
// The inverse along z then y only, in place (the rows stay half spectra).
int fft_c2r_cols_inplace(float2* data, gtomx_int3 dims, unsigned batch, hipStream_t s);
// For n x n images with n > 256: the row C2R of nb spectra (H apart in G, in
// place after fft_c2r_cols_inplace), times scale[o], folded into best / idx in
// orientation order from angle0 (strict >). variant 1 prefetches the next
// orientation's spectrum rows into registers during each transform.
int pick_rows_max_long(int n, const float2* G, int nb, long long H, int nrows, int angle0, const float* scale,
                       float* best, int32_t* idx, hipStream_t s, int variant = 0);
// True when every axis of dims has a native transform.
bool dims_native(gtomx_int3 dims);
// The inverse along z then y of nb half spectra in place, each axis native or
// by the chirp-z form (dir -1 forward, +1 inverse).
int yz_axes_any(float2* data, gtomx_int3 dims, unsigned batch, int dir, hipStream_t s);
// The x C2R (real part, unscaled) of the nb * n * n rows of nb n^3 half spectra
// H apart in G, by the chirp-z form, written in place: the real map of
// spectrum o fills the first n^3 floats of its slot (every row is read before
// the rows it overwrites are written). n at most 2048.
int rows_c2r_any_inplace(float2* G, int n, int nb, long long H, hipStream_t s);

// Fourier-space resizing (resize.hip): op is one of the GTOMX_FFT_* values.
constexpr int FFT_CROP = GTOMX_FFT_CROP, FFT_PAD = GTOMX_FFT_PAD, FFT_FULL_CROP = GTOMX_FFT_FULL_CROP,
              FFT_FULL_PAD = GTOMX_FFT_FULL_PAD;
int fft_resize(int op, bool complex_values, const void* in, void* out, gtomx_int3 od, gtomx_int3 nd, unsigned batch,
               hipStream_t s);

// Fourier-space rotation (rotate.hip). rotate3d_ft_batch: one zero-centred half
// spectrum of an n^3 cube rotated by `batch` matrices (9 floats each, from
// gtomx_projector_matrices at scale 1). rotate2d_ft_batch: item b of the input
// (stride in_stride complex values, 0 for one shared image) rotated with the
// coefficients cs[b] from rotate2d_coefficients(angle_b).
int rotate3d_ft_batch(const float2* in, float2* out, int n, const float* d_mats, unsigned batch, bool zerocentred,
                      hipStream_t s);
int rotate2d_ft_batch(const float2* in, size_t in_stride, float2* out, int nx, int ny, const float2* d_cs,
                      unsigned batch, float maxfreq, bool zerocentred, hipStream_t s);
void rotate2d_coefficients(float angle, float2* cs);

// d_CubicBSplinePrefilter3D / 2D (rotate.hip) in place on `batch` arrays of
// dims, ndims axes (x, then y, then z), `channels` interleaved components (1
// real, 2 complex), each line by ConvertToInterpolationCoefficients.
int bspline_prefilter(float* v, gtomx_int3 dims, int ndims, unsigned batch, int channels, hipStream_t s);

// d_Norm with T_NORM_MEAN01STD per item (picker.hip); scratch holds
// norm_scratch_doubles(batch) doubles.
int norm_mean01std(const float* in, float* out, size_t n, unsigned batch, double* scratch, hipStream_t s);
size_t norm_scratch_doubles(unsigned batch);
// P[b][i] = E[i] * conj(P[b][i]) for b < batch (picker.hip).
int image_times_conj(const float2* E, float2* P, size_t h, unsigned batch, hipStream_t s);

// Element-wise helpers (ops.hip).
int fill_f32(float* d, size_t n, float v, hipStream_t s);
int fill_i32(int32_t* d, size_t n, int32_t v, hipStream_t s);
int sphere_mask_ifftshift(float* out, gtomx_int3 dims, float radius, float sigma, hipStream_t s);
int sum_f64(const float* d, size_t n, double* d_result, hipStream_t s);
int mul_conj_inplace(float2* a, const float2* b, size_t n, hipStream_t s);
int square_f32(const float* in, float* out, size_t n, hipStream_t s);
// std = sqrt(max(0, M*c2/N - (c1/N)^2)) / M (double), optional mean = c1/(N*M);
// with best != nullptr also best = |std| > 1e-15 ? best/std : 0 and the float index copy.
// d_mapsum (or nullptr): the map's sum; its mean was subtracted before the convolutions and is added back to the mean.
int local_std_combine(const float* c1, const float* c2, const double* d_masksum, size_t n,
                      float* std_out, float* mean_out, float* best, const int32_t* idx,
                      float* angle_out, hipStream_t s, const double* d_mapsum = nullptr);

// Local standard deviation chain (ops.hip). The scratch holds the mask FT,
// two spectra, three real volumes and a few doubles.
struct LocalStdScratch {
  float2* maskft;
  float2* aft;
  float2* bft;
  float* rbuf;
  float* c1;
  float* c2;
  double* dsum;  // LS_DOUBLES doubles; the mask sum is dsum[256], the map sum dsum[LS_MAPSUM]
};
constexpr int LS_MAPSUM = 520;  // its 256 partials precede it
constexpr int LS_DOUBLES = LS_MAPSUM + 1;
size_t local_std_scratch_bytes(gtomx_int3 d);
LocalStdScratch local_std_carve(void* base, gtomx_int3 d);
// fouriermask: NULL, or a real half spectrum (FFT order) that scales the mask spectrum.
int local_std_prepare(const LocalStdScratch& sc, gtomx_int3 d, float radius, hipStream_t s,
                      const float* fouriermask = nullptr);
int local_std_apply(const LocalStdScratch& sc, const float* map, gtomx_int3 d, float* std_out, float* mean_out,
                    float* best, const int32_t* idx, float* angle_out, hipStream_t s);

// Standalone projector launch (ops.hip).
// nz = n: n^3 projections; nz = 1: n x n central sections (2D). volz: the
// volume's z extent (p for a 3D spectrum; 1 for a 2D image's spectrum, whose
// in-plane rotations read the missing planes as 0).
int project_launch(const float2* vol, int p, float2* out, int n, unsigned rmax, const float* d_mats,
                   unsigned batch, hipStream_t s, int nz = -1, int volz = -1);

}  // namespace gx
