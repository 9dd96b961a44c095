// Host runtime: diagnostics, allocation, per-device tables and the host-only
// angle utilities of the C ABI.
#include "runtime.hpp"

#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace gx {

namespace {
thread_local std::string g_last_error;
std::atomic<int> g_fail_alloc_after{0};

struct DeviceTables {
  float2* tw = nullptr;
};
std::mutex g_tables_mu;
std::vector<DeviceTables> g_tables;
}  // namespace

int fail(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}

int ok() { return GTOMX_OK; }

int hip_fail(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    return fail(GTOMX_ERR_OUT_OF_MEMORY, m);
  return fail(GTOMX_ERR_DEVICE, m);
}

int dev_alloc(void** p, size_t bytes) {
  *p = nullptr;
  int n = g_fail_alloc_after.load();
  if (n > 0) {
    if (n == 1) {
      g_fail_alloc_after.store(0);
      return fail(GTOMX_ERR_OUT_OF_MEMORY, "device allocation failed (injected)");
    }
    g_fail_alloc_after.store(n - 1);
  }
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return hip_fail(e, "hipMalloc");
  }
  return GTOMX_OK;
}

void dev_free(void* p) {
  if (p) (void)hipFree(p);
}

int twiddles(const float2** table) {
  int dev = 0;
  GX_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_tables_mu);
  if ((int)g_tables.size() <= dev) g_tables.resize(dev + 1);
  if (!g_tables[dev].tw) {
    const int total = TW_LEN_HOST + TW3_LEN_HOST;
    std::vector<float2> h(total);
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int m = 0; m < TW_LEN_HOST; ++m) {
      long double a = -twopi * (long double)m / (long double)TW_LEN_HOST;
      h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    for (int m = 0; m < TW3_LEN_HOST; ++m) {
      long double a = -twopi * (long double)m / (long double)TW3_LEN_HOST;
      h[TW_LEN_HOST + m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(float2) * total);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(twiddles)");
    e = hipMemcpy(d, h.data(), sizeof(float2) * total, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(d);
      return hip_fail(e, "hipMemcpy(twiddles)");
    }
    g_tables[dev].tw = static_cast<float2*>(d);
  }
  *table = g_tables[dev].tw;
  return GTOMX_OK;
}

int check_fft_dims(int ndims, gtomx_int3 d) {
  if (ndims < 1 || ndims > 3) return fail(GTOMX_ERR_INVALID_ARGUMENT, "ndims must be 1, 2 or 3");
  if (d.x < 1 || d.y < 1 || d.z < 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "dims must be positive");
  int want = 3 - (d.z == 1) - (d.z == 1 && d.y == 1);
  if (ndims != want)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "ndims does not match dims (trailing extents must be 1)");
  if (!fft_len_ok_any(d.x))
    return fail(GTOMX_ERR_UNSUPPORTED, "dims.x must be 2^a in [8, 4096] or 3*2^a in [12, 3072]");
  if (d.y != 1 && !fft_len_ok_any(d.y))
    return fail(GTOMX_ERR_UNSUPPORTED, "dims.y must be 1, 2^a in [8, 4096] or 3*2^a in [12, 3072]");
  if (d.z != 1 && !fft_len_ok_any(d.z))
    return fail(GTOMX_ERR_UNSUPPORTED, "dims.z must be 1, 2^a in [8, 4096] or 3*2^a in [12, 3072]");
  return GTOMX_OK;
}

int check_fft_dims_any(int ndims, gtomx_int3 d) {
  if (ndims < 1 || ndims > 3) return fail(GTOMX_ERR_INVALID_ARGUMENT, "ndims must be 1, 2 or 3");
  if (d.x < 1 || d.y < 1 || d.z < 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "dims must be positive");
  int want = 3 - (d.z == 1) - (d.z == 1 && d.y == 1);
  if (ndims != want)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "ndims does not match dims (trailing extents must be 1)");
  auto ok = [](int n) { return fft_len_ok_any(n) || n <= 2048; };
  if (!ok(d.x) || !ok(d.y) || !ok(d.z))
    return fail(GTOMX_ERR_UNSUPPORTED,
                "extents must be at most 2048, or 2^a up to 4096, or 3*2^a up to 3072");
  return GTOMX_OK;
}

}  // namespace gx

using namespace gx;

extern "C" {

const char* gtomx_version(void) { return "gtomx 0.1 (gfx950)"; }

const char* gtomx_status_string(int s) {
  switch (s) {
    case GTOMX_OK: return "ok";
    case GTOMX_ERR_INVALID_ARGUMENT: return "invalid argument";
    case GTOMX_ERR_UNSUPPORTED: return "unsupported";
    case GTOMX_ERR_OUT_OF_MEMORY: return "out of memory";
    case GTOMX_ERR_DEVICE: return "device error";
    case GTOMX_ERR_INTERNAL: return "internal error";
    default: return "unknown status";
  }
}

const char* gtomx_last_error(void) { return g_last_error.c_str(); }

void gtomx_debug_fail_alloc_after(int n) { g_fail_alloc_after.store(n < 0 ? 0 : n); }

// GetEqualAngularSpacing (src/Helpers/AngularSpacing.cu:46-67) with float32
// loop counters so that the orientation count and order match the reference.
int gtomx_equal_angular_spacing(float phimin, float phimax, float thetamin, float thetamax,
                                float psimin, float psimax, float increment,
                                gtomx_float3* h_angles, int capacity, int* count) {
  if (!count) return fail(GTOMX_ERR_INVALID_ARGUMENT, "count is NULL");
  if (!(increment > 0.0f)) return fail(GTOMX_ERR_INVALID_ARGUMENT, "increment must be positive");
  if (capacity < 0 || (capacity > 0 && !h_angles))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "bad output buffer");
  const float PI2f = 6.283185307179586476925286766559f;
  int n = 0;
  for (float tilt = thetamin; tilt <= thetamax + 1e-5f; tilt += increment) {
    int rot_nstep = tilt > 0 ? (int)std::ceil(PI2f * std::sin(tilt) / increment) : 1;
    float rot_sam = PI2f / (float)rot_nstep;
    for (float rot = phimin; rot <= phimax - rot_sam + 1e-5f; rot += rot_sam)
      for (float psi = psimin; psi <= psimax - increment + 1e-5f; psi += increment) {
        if (n < capacity) h_angles[n] = gtomx_float3{rot, tilt, psi};
        ++n;
      }
  }
  *count = n;
  return GTOMX_OK;
}

int gtomx_fibonacci_angles(int ndirs, int npsi, gtomx_float3* h_angles) {
  if (ndirs <= 0 || npsi <= 0 || !h_angles)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "ndirs and npsi must be positive, h_angles non-NULL");
  const double pi = 3.14159265358979323846;
  const double golden = pi * (3.0 - std::sqrt(5.0));
  for (int i = 0; i < ndirs; ++i) {
    double z = 1.0 - (2.0 * i + 1.0) / ndirs;
    double tilt = std::acos(std::max(-1.0, std::min(1.0, z)));
    double rot = std::fmod(golden * i, 2.0 * pi);
    for (int j = 0; j < npsi; ++j) {
      double psi = 2.0 * pi * j / npsi;
      h_angles[(size_t)i * npsi + j] = gtomx_float3{(float)rot, (float)tilt, (float)psi};
    }
  }
  return GTOMX_OK;
}

// transpose(Matrix3Euler(a)) * s, row-major (src/Relion/Project.cu:16,
// src/Helpers/Euler.cu:39-68), evaluated in double.
int gtomx_projector_matrices(const gtomx_float3* h_angles, unsigned n, float oversample,
                             float* h_matrices) {
  if ((n > 0 && (!h_angles || !h_matrices)))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL angle or matrix buffer");
  for (unsigned i = 0; i < n; ++i) {
    double ca = std::cos((double)h_angles[i].x), sa = std::sin((double)h_angles[i].x);
    double cb = std::cos((double)h_angles[i].y), sb = std::sin((double)h_angles[i].y);
    double cg = std::cos((double)h_angles[i].z), sg = std::sin((double)h_angles[i].z);
    double cc = cb * ca, cs = cb * sa, sc = sb * ca, ss = sb * sa;
    double A[9] = {cg * cc - sg * sa,  cg * cs + sg * ca,  -cg * sb,
                   -sg * cc - cg * sa, -sg * cs + cg * ca, sg * sb,
                   sc,                 ss,                 cb};
    double s = oversample;
    float* M = h_matrices + 9 * (size_t)i;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) M[3 * r + c] = (float)(s * A[3 * c + r]);
  }
  return GTOMX_OK;
}

}  // extern "C"
