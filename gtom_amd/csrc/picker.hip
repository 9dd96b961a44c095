// The Picker class (include/Correlation.cuh:136-164, src/Correlation/Picker.cu)
// and the two standalone pieces only it uses: d_FFTFullPad on real volumes
// (src/FFT/FFTResize.cu:92-110, kernel :212-238) and d_Norm with
// T_NORM_MEAN01STD (src/ImageManipulation/Norm.cu:36-70, statistics of d_Dev,
// src/Generics/Dev.cu:13-72).
//
// The reference class never runs (its constructor throws, Picker.cu:27, and the
// rotation in PerformCorrelation is commented out, Picker.cu:181-187). This is
// the pipeline its code describes, with the rotation done by the device-array
// projector (src/Relion/Project.cu:89-177) at the reference size:
//   R = project(refft, transpose(Euler(angle)), rmax = nref/2) * ctf,
//   r = C2R_n(R), padded = FullPad(r) -> mean0/std1, P = R2C(padded),
//   C = C2R_u(imageft * conj(P)), best/angle updated where C > best.
// Orientations are processed B at a time: the projection, C2R, padding,
// normalisation, R2C, product and C2R are each one batched launch (or a few),
// and the running maximum takes the B maps in increasing orientation order,
// which keeps the reference's per-orientation strict '>' rule.
#include <algorithm>
#include <cmath>
#include <vector>

#include "runtime.hpp"

namespace gx {
namespace {

// ---------------------------------------------------------------------------
// mean0/std1 per volume: two passes over the data for the mean and the
// population variance (as d_Dev), each reduced in double by NB blocks per
// volume and finished by one block per volume; then the normalisation.
// ---------------------------------------------------------------------------
constexpr int NB = 256;  // reduction blocks per volume

__device__ double block_sum_d(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < 4; ++w) s += red[w];
  __syncthreads();
  return s;
}

// pass 0: sum x; pass 1: sum (x - mean)^2 with mean = stats[2b]
__global__ __launch_bounds__(256) void k_stats_partial(const float* __restrict__ x, size_t n, int pass,
                                                       const double* __restrict__ stats, double* __restrict__ part) {
  __shared__ double red[4];
  const size_t b = blockIdx.y;
  const float* xb = x + b * n;
  const double mean = pass ? stats[2 * b] : 0.0;
  double acc = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const double d = (double)xb[i] - mean;
    acc += pass ? d * d : d;
  }
  const double t = block_sum_d(acc, red);
  if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = t;
}

// stats[2b] = mean (pass 0) or stats[2b+1] = std (pass 1)
__global__ __launch_bounds__(256) void k_stats_finish(const double* __restrict__ part, int nparts, size_t n, int pass,
                                                      double* __restrict__ stats) {
  __shared__ double red[4];
  const size_t b = blockIdx.x;
  double acc = 0;
  for (int i = threadIdx.x; i < nparts; i += 256) acc += part[b * nparts + i];
  const double t = block_sum_d(acc, red);
  if (threadIdx.x == 0) {
    if (pass == 0)
      stats[2 * b] = t / (double)n;
    else
      stats[2 * b + 1] = sqrt(t / (double)n);
  }
}

__global__ __launch_bounds__(256) void k_norm_apply(const float* __restrict__ x, float* __restrict__ y, size_t n,
                                                    const double* __restrict__ stats) {
  const size_t b = blockIdx.y;
  const double mean = stats[2 * b], sd = stats[2 * b + 1];
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const double d = (double)x[b * n + i] - mean;
    y[b * n + i] = (float)(sd == 0.0 ? d : d / sd);
  }
}

}  // namespace

// The (mean, sd) of each item, in double, at scratch + NB * batch.
static int norm_stats(const float* in, size_t n, unsigned batch, double* scratch, hipStream_t s) {
  double* part = scratch;
  double* stats = scratch + (size_t)NB * batch;
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL(k_stats_partial, dim3(NB, batch), dim3(256), 0, s, in, n, pass, stats, part);
    GX_LAUNCHED("k_stats_partial");
    hipLaunchKernelGGL(k_stats_finish, dim3(batch), dim3(256), 0, s, part, NB, n, pass, stats);
    GX_LAUNCHED("k_stats_finish");
  }
  return GTOMX_OK;
}

// scratch: NB * batch + 2 * batch doubles
int norm_mean01std(const float* in, float* out, size_t n, unsigned batch, double* scratch, hipStream_t s) {
  double* stats = scratch + (size_t)NB * batch;
  GX_TRY(norm_stats(in, n, batch, scratch, s));
  const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_norm_apply, dim3(grid, batch), dim3(256), 0, s, in, out, n, stats);
  GX_LAUNCHED("k_norm_apply");
  return GTOMX_OK;
}
size_t norm_scratch_doubles(unsigned batch) { return (size_t)NB * batch + 2 * (size_t)batch; }

namespace {

// P[b][i] = E[i] * conj(P[b][i])  (d_ComplexMultiplyByConjVector(d_imageft, P, P),
// src/Correlation/Picker.cu:228, src/Generics/Arithmetics.cu:334-344)
__global__ __launch_bounds__(256) void k_image_times_conj(const float2* __restrict__ E, float2* __restrict__ P,
                                                          size_t h, unsigned batch) {
  const size_t total = h * batch;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    const float2 e = E[i % h], p = P[i];
    P[i] = make_float2(e.x * p.x + e.y * p.y, e.y * p.x - e.x * p.y);
  }
}

// The same product with P the R2C of the un-normalised padded reference: the
// normalisation (x - mean) / sd of k_norm_apply applied in Fourier space,
// P' = (P - mean * n at the DC bin) / sd (sd = 0: no division), since the
// transform is linear. Lets the R2C skip the rows the padding leaves zero.
__global__ __launch_bounds__(256) void k_image_times_conj_norm(const float2* __restrict__ E, float2* __restrict__ P,
                                                               size_t h, unsigned batch,
                                                               const double* __restrict__ stats, size_t n) {
  const size_t total = h * batch;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < total; i += (size_t)gridDim.x * 256) {
    const size_t b = i / h;
    const double mean = stats[2 * b], sd = stats[2 * b + 1];
    const float2 e = E[i % h], q = P[i];
    double px = q.x, py = q.y;
    if (i % h == 0) px -= mean * (double)n;
    if (sd != 0.0) {
      px /= sd;
      py /= sd;
    }
    const float2 p = make_float2((float)px, (float)py);
    P[i] = make_float2(e.x * p.x + e.y * p.y, e.y * p.x - e.x * p.y);
  }
}

// PickingUpdateKernel (src/Correlation/Picker.cu:249-276) over B maps in order.
__global__ __launch_bounds__(256) void k_update_max_angle(const float* __restrict__ C, size_t v, unsigned batch,
                                                          const float3* __restrict__ angles, float* __restrict__ best,
                                                          float3* __restrict__ bestangle) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < v; i += (size_t)gridDim.x * 256) {
    float b = best[i];
    int k = -1;
    for (unsigned o = 0; o < batch; ++o) {
      const float c = C[o * v + i];
      if (c > b) {
        b = c;
        k = (int)o;
      }
    }
    if (k >= 0) {
      best[i] = b;
      bestangle[i] = angles[k];
    }
  }
}

}  // namespace

int image_times_conj(const float2* E, float2* P, size_t h, unsigned batch, hipStream_t s) {
  const unsigned grid = (unsigned)std::min<size_t>((h * batch + 255) / 256, 16384);
  hipLaunchKernelGGL(k_image_times_conj, dim3(grid), dim3(256), 0, s, E, P, h, batch);
  GX_LAUNCHED("k_image_times_conj");
  return GTOMX_OK;
}

}  // namespace gx

// ---------------------------------------------------------------------------
// Picker state
// ---------------------------------------------------------------------------
struct gtomx_picker {
  int nr = 0;           // reference side
  int nrz = 0;          // reference z extent: nr (3D) or 1 (2D)
  gtomx_int3 di{};      // image dims
  int device = 0;
  unsigned B = 0;
  size_t Hr = 0, Vr = 0, H = 0, V = 0;
  float2* refft = nullptr;
  float2* imageft = nullptr;
  const float* ctf = nullptr;  // caller's, as in the reference (SetImage keeps the pointer)
  float2* proj = nullptr;      // B x Hr, also the C2R workspace at the reference size
  float* realref = nullptr;    // B x Vr
  float* padded = nullptr;     // B x V, then the correlation maps
  float2* pft = nullptr;       // B x H, also the C2R workspace at the image size
  double* nscr = nullptr;      // normalisation scratch
  float* mats = nullptr;       // B x 9
  float3* dang = nullptr;      // B angles
  float* h_mats = nullptr;     // pinned staging: B x 9 matrices, then B x 3 angles
  hipEvent_t staged = nullptr;
  bool staged_recorded = false;
  bool have_image = false;
  gx::StreamFence fence;  // buffers are reused by calls on any stream
};

namespace gx {
namespace {

void picker_release(gtomx_picker* pk) {
  if (!pk) return;
  pk->fence.release();
  for (void* p : {(void*)pk->refft, (void*)pk->imageft, (void*)pk->proj, (void*)pk->realref, (void*)pk->padded,
                  (void*)pk->pft, (void*)pk->nscr, (void*)pk->mats, (void*)pk->dang})
    dev_free(p);
  if (pk->h_mats) (void)hipHostFree(pk->h_mats);
  if (pk->staged) (void)hipEventDestroy(pk->staged);
  delete pk;
}

int picker_check(const gtomx_picker* pk) {
  if (!pk) return fail(GTOMX_ERR_INVALID_ARGUMENT, "picker is NULL");
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  if (dev != pk->device) return fail(GTOMX_ERR_INVALID_ARGUMENT, "picker was created on another device");
  return GTOMX_OK;
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

int gtomx_norm_mean01std(const float* d_input, float* d_output, size_t elements_, unsigned batch, void* stream) {
  if (!d_input || !d_output) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  if (elements_ == 0 || batch == 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "elements and batch must be positive");
  hipStream_t s = as_stream(stream);
  DevBuf scr;
  GX_TRY(scr.alloc(norm_scratch_doubles(batch) * sizeof(double)));
  GX_TRY(norm_mean01std(d_input, d_output, elements_, batch, static_cast<double*>(scr.p), s));
  GX_HIP(hipStreamSynchronize(s));  // the scratch is freed on return
  return GTOMX_OK;
}

int gtomx_picker_create(gtomx_picker** out, const float* d_ref, gtomx_int3 dimsref, const float* d_refmask,
                        gtomx_int3 dimsimage, unsigned batch, void* stream) {
  (void)d_refmask;  // prepared but never used by the reference (src/Correlation/Picker.cu:56-88, 193-221)
  if (!out) return fail(GTOMX_ERR_INVALID_ARGUMENT, "picker pointer is NULL");
  *out = nullptr;
  if (!d_ref) return fail(GTOMX_ERR_INVALID_ARGUMENT, "d_ref is NULL");
  // 3D: a cubic reference and a volume; 2D (the reference's ndims == 2 branch,
  // src/Correlation/Picker.cu:35, 163-171): a square reference and an image
  const bool is2d = dimsref.z == 1;
  if (dimsref.x != dimsref.y || (!is2d && dimsref.x != dimsref.z))
    return fail(GTOMX_ERR_UNSUPPORTED, "the reference must be a cube (3D) or a square (2D)");
  if (is2d != (dimsimage.z == 1))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "reference and image must have the same dimensionality");
  // extents without a native transform take the chirp-z form (and full, not banded, transforms)
  GX_TRY(check_fft_dims_any(is2d ? 2 : 3, dimsref));
  GX_TRY(check_fft_dims_any(is2d ? 2 : 3, dimsimage));
  if (dimsref.x > dimsimage.x || dimsref.y > dimsimage.y || dimsref.z > dimsimage.z)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "the reference is larger than the image");
  if (batch > 64) return fail(GTOMX_ERR_INVALID_ARGUMENT, "batch must be <= 64");
  hipStream_t s = as_stream(stream);
  auto* pk = new gtomx_picker();
  pk->nr = dimsref.x;
  pk->nrz = dimsref.z;
  pk->di = dimsimage;
  pk->B = batch ? batch : 8u;
  pk->Hr = elements_ft(dimsref);
  pk->Vr = elements(dimsref);
  pk->H = elements_ft(dimsimage);
  pk->V = elements(dimsimage);
  int st = GTOMX_OK;
  hipError_t he = hipGetDevice(&pk->device);
  if (he != hipSuccess) {
    picker_release(pk);
    return hip_fail(he, "hipGetDevice");
  }
  auto alloc = [&](void** dst, size_t bytes) {
    if (st == GTOMX_OK) st = dev_alloc(dst, bytes);
  };
  const size_t B = pk->B;
  alloc((void**)&pk->refft, pk->Hr * sizeof(float2));
  alloc((void**)&pk->imageft, pk->H * sizeof(float2));
  alloc((void**)&pk->proj, B * pk->Hr * sizeof(float2));
  alloc((void**)&pk->realref, B * pk->Vr * sizeof(float));
  alloc((void**)&pk->padded, B * pk->V * sizeof(float));
  alloc((void**)&pk->pft, B * pk->H * sizeof(float2));
  alloc((void**)&pk->nscr, norm_scratch_doubles((unsigned)B) * sizeof(double));
  alloc((void**)&pk->mats, B * 9 * sizeof(float));
  alloc((void**)&pk->dang, B * sizeof(float3));
  if (st == GTOMX_OK) {
    he = hipHostMalloc((void**)&pk->h_mats, B * 12 * sizeof(float), hipHostMallocDefault);
    if (he == hipSuccess) he = hipEventCreateWithFlags(&pk->staged, hipEventDisableTiming);
    if (he != hipSuccess) st = hip_fail(he, "picker host staging");
  }
  if (st == GTOMX_OK) {
    // Initialize: refft = R2C(ifftshift(ref)) at the reference size (src/Correlation/Picker.cu:92-94, 112-114);
    // realref serves as the shifted copy.
    st = gtomx_remap_full2fullfft(d_ref, pk->realref, dimsref, 1, stream);
    if (st == GTOMX_OK) st = fft_r2c(pk->realref, pk->refft, dimsref, 1, s);
    if (st == GTOMX_OK) st = pk->fence.leave(s);
  }
  if (st != GTOMX_OK) {
    std::string msg = gtomx_last_error();
    picker_release(pk);
    return fail(st, msg);
  }
  *out = pk;
  return GTOMX_OK;
}

int gtomx_picker_destroy(gtomx_picker* picker) {
  if (!picker) return GTOMX_OK;
  picker_release(picker);  // waits for the last call's work on the picker's buffers
  return GTOMX_OK;
}

int gtomx_picker_set_image(gtomx_picker* pk, float* d_image, const float* d_ctf, void* stream) {
  GX_TRY(picker_check(pk));
  if (!d_image || !d_ctf) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL buffer");
  hipStream_t s = as_stream(stream);
  GX_TRY(pk->fence.enter(s));
  pk->ctf = d_ctf;
  // image -> mean0/std1 in place (src/Correlation/Picker.cu:148), imageft = R2C(image) (:152),
  // then the image is left holding its square (:154), as in the reference.
  GX_TRY(norm_mean01std(d_image, d_image, pk->V, 1, pk->nscr, s));
  GX_TRY(fft_r2c(d_image, pk->imageft, pk->di, 1, s));
  GX_TRY(square_f32(d_image, d_image, pk->V, s));
  pk->have_image = true;
  return pk->fence.leave(s);
}

int gtomx_picker_correlate_angles(gtomx_picker* pk, const gtomx_float3* h_angles, unsigned nangles,
                                  float* d_bestccf, gtomx_float3* d_bestangle, void* stream) {
  GX_TRY(picker_check(pk));
  if (!pk->have_image) return fail(GTOMX_ERR_INVALID_ARGUMENT, "SetImage has not been called");
  if (!d_bestccf || !d_bestangle) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL output buffer");
  if (nangles > 0 && !h_angles) return fail(GTOMX_ERR_INVALID_ARGUMENT, "h_angles is NULL");
  hipStream_t s = as_stream(stream);
  const gtomx_int3 dr{pk->nr, pk->nr, pk->nrz};
  const unsigned grid_i = (unsigned)std::min<size_t>((pk->H * pk->B + 255) / 256, 65536);
  const unsigned grid_v = (unsigned)std::min<size_t>((pk->V + 255) / 256, 65536);
  GX_TRY(pk->fence.enter(s));
  for (unsigned b0 = 0; b0 < nangles; b0 += pk->B) {
    const unsigned nb = std::min(pk->B, nangles - b0);
    // stage the matrices transpose(Euler) (no scale) and the angles for this batch
    if (pk->staged_recorded) GX_HIP(hipEventSynchronize(pk->staged));
    GX_TRY(gtomx_projector_matrices(h_angles + b0, nb, 1.0f, pk->h_mats));
    for (unsigned k = 0; k < nb; ++k) {
      pk->h_mats[9 * pk->B + 3 * k] = h_angles[b0 + k].x;
      pk->h_mats[9 * pk->B + 3 * k + 1] = h_angles[b0 + k].y;
      pk->h_mats[9 * pk->B + 3 * k + 2] = h_angles[b0 + k].z;
    }
    GX_HIP(hipMemcpyAsync(pk->mats, pk->h_mats, nb * 9 * sizeof(float), hipMemcpyHostToDevice, s));
    GX_HIP(hipMemcpyAsync(pk->dang, pk->h_mats + 9 * pk->B, nb * 3 * sizeof(float), hipMemcpyHostToDevice, s));
    GX_HIP(hipEventRecord(pk->staged, s));
    pk->staged_recorded = true;
    // R = project(refft) * ctf; r = C2R_n(R)   (src/Correlation/Picker.cu:185-190)
    GX_TRY(project_launch(pk->refft, pk->nr, pk->proj, pk->nr, (unsigned)(pk->nr / 2), pk->mats, nb, s, pk->nrz,
                          pk->nrz));
    GX_TRY(gtomx_complex_multiply_by_vector(reinterpret_cast<const gtomx_complex*>(pk->proj), pk->ctf,
                                            reinterpret_cast<gtomx_complex*>(pk->proj), pk->Hr, nb, stream));
    GX_TRY(fft_c2r(pk->proj, pk->proj, pk->realref, dr, nb, 1.0f / (float)pk->Vr, s));
    // FullPad, mean0/std1, R2C   (:225-227)
    GX_TRY(fft_resize(FFT_FULL_PAD, false, pk->realref, pk->padded, dr, pk->di, nb, s));
    // The padded reference is zero outside the z x y box of the reference
    // (FFT order); where the row kernels allow, the R2C runs over that band
    // of rows and the normalisation moves into the product (same values up
    // to rounding).
    const int rw = dims_native(pk->di) ? rows_per_workgroup(pk->di.x) : 0;
    const int vby = rw > 0 ? (dr.y / 2 + 1 + rw - 1) / rw * rw : 0;
    const int vbz = 2 * (dr.z / 2 + 1) <= pk->di.z ? dr.z / 2 + 1 : 0;
    const bool vbanded = pk->nrz > 1 && pk->di.x <= 256 && pk->di.y <= 256 && pk->di.z <= 256 && vby > 0 &&
                         2 * vby <= pk->di.y;
    if (vbanded) {
      GX_TRY(norm_stats(pk->padded, pk->V, nb, pk->nscr, s));
      GX_TRY(volume_r2c_banded(pk->padded, pk->pft, pk->di, nb, vby, vbz, s));
      // C = C2R_u(imageft * conj(P))   (:228-229)
      hipLaunchKernelGGL(k_image_times_conj_norm, dim3(grid_i), dim3(256), 0, s, pk->imageft, pk->pft, pk->H, nb,
                         pk->nscr + (size_t)NB * nb, pk->V);
      GX_LAUNCHED("k_image_times_conj_norm");
    } else {
      GX_TRY(norm_mean01std(pk->padded, pk->padded, pk->V, nb, pk->nscr, s));
      GX_TRY(fft_r2c(pk->padded, pk->pft, pk->di, nb, s));
      // C = C2R_u(imageft * conj(P))   (:228-229)
      hipLaunchKernelGGL(k_image_times_conj, dim3(grid_i), dim3(256), 0, s, pk->imageft, pk->pft, pk->H, nb);
      GX_LAUNCHED("k_image_times_conj");
    }
    GX_TRY(fft_c2r(pk->pft, pk->pft, pk->padded, pk->di, nb, 1.0f, s));
    // best/angle where C > best   (:233-241)
    hipLaunchKernelGGL(k_update_max_angle, dim3(grid_v), dim3(256), 0, s, pk->padded, pk->V, nb, pk->dang,
                       d_bestccf, reinterpret_cast<float3*>(d_bestangle));
    GX_LAUNCHED("k_update_max_angle");
  }
  return pk->fence.leave(s);
}

int gtomx_picker_correlate(gtomx_picker* pk, float anglestep, float* d_bestccf, gtomx_float3* d_bestangle,
                           void* stream) {
  GX_TRY(picker_check(pk));
  if (!(anglestep > 0.f) || !std::isfinite(anglestep))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "anglestep must be positive");
  const float PI2 = 6.283185307179586f, PI = 3.14159265358979f;
  std::vector<gtomx_float3> angles;
  if (pk->nrz == 1) {
    // 2D: ceil(2 pi / anglestep) in-plane steps of 2 pi / steps (src/Correlation/Picker.cu:165-171), in float
    const int steps = (int)std::ceil(PI2 / anglestep);
    const float step = PI2 / (float)steps;
    for (int a = 0; a < steps; ++a) angles.push_back(gtomx_float3{0.f, 0.f, (float)a * step});
  } else {
    // GetEqualAngularSpacing((0, 2pi), (0, pi), (0, 2pi), anglestep) (src/Correlation/Picker.cu:163-164)
    int n = 0;
    GX_TRY(gtomx_equal_angular_spacing(0.f, PI2, 0.f, PI, 0.f, PI2, anglestep, nullptr, 0, &n));
    angles.resize((size_t)n);
    GX_TRY(gtomx_equal_angular_spacing(0.f, PI2, 0.f, PI, 0.f, PI2, anglestep, angles.data(), n, &n));
  }
  const int n = (int)angles.size();
  GX_TRY(gtomx_picker_correlate_angles(pk, angles.data(), (unsigned)n, d_bestccf, d_bestangle, stream));
  // the angle list lives on this stack frame; the uploads have been staged through pinned memory
  return GTOMX_OK;
}

}  // extern "C"
