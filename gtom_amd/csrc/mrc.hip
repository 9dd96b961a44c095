// MRC file I/O (src/IO/mrc.cpp:7-115, HeaderMRC include/IO.cuh:26-96) and the
// conversion of MRC data modes to float on the device (MixedToDeviceTfloat,
// src/Helpers/Memory.cu:424-436): the raw file bytes are copied to the GPU as
// they are (1 or 2 bytes per voxel for modes 0, 1 and 6) and widened there.
#include <cstdio>
#include <cstring>
#include <vector>

#include "runtime.hpp"

static_assert(sizeof(gtomx_mrc_header) == 1024, "the MRC header is 1024 bytes");

namespace gx {
namespace {

// Bytes per voxel by mode (the reference's MRC_DATATYPE_SIZE, include/IO.cuh:24,
// for the modes it defines); 0 for modes it does not.
size_t mode_size(int mode) {
  switch (mode) {
    case 0: return 1;   // MRC_BYTE
    case 1: return 2;   // MRC_SHORT
    case 2: return 4;   // MRC_FLOAT
    case 3: return 4;   // MRC_SHORTCOMPLEX
    case 4: return 8;   // MRC_FLOATCOMPLEX
    case 6: return 2;   // MRC_UNSIGNEDSHORT
    case 16: return 3;  // MRC_RGB
    default: return 0;
  }
}

struct File {
  FILE* f = nullptr;
  explicit File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

int check_header(const gtomx_mrc_header& h) {
  if (h.nx < 1 || h.ny < 1 || h.nz < 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "MRC dimensions must be positive");
  if (mode_size(h.mode) == 0) return fail(GTOMX_ERR_UNSUPPORTED, "MRC mode not supported");
  if (h.nsymbt < 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "negative extended header size");
  return GTOMX_OK;
}

template <class T> __global__ __launch_bounds__(256) void k_widen(const T* __restrict__ in, float* __restrict__ out,
                                                                   size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) out[i] = (float)in[i];
}

}  // namespace
}  // namespace gx

using namespace gx;

extern "C" {

void gtomx_mrc_header_default(gtomx_mrc_header* h) {
  if (!h) return;
  std::memset(h, 0, sizeof(*h));
  // HeaderMRC() (include/IO.cuh:72-95)
  h->nx = h->ny = h->nz = 1;
  h->mode = 2;
  h->mx = h->my = h->mz = 1;
  h->cellx = h->celly = h->cellz = 1.f;
  h->mapc = 1;
  h->mapr = 2;
  h->maps = 3;
  h->dmin = -1.f;
  h->dmax = 1.f;
  // left unset by the reference; the MRC format's map identifier and a
  // little-endian machine stamp
  std::memcpy(h->cmap, "MAP ", 4);
  h->stamp[0] = 0x44;
  h->stamp[1] = 0x44;
}

int gtomx_mrc_read_header(const char* path, gtomx_mrc_header* header) {
  if (!path || !header) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL argument");
  File in(path, "rb");
  if (!in.f) return fail(GTOMX_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  if (std::fread(header, 1, sizeof(*header), in.f) != sizeof(*header))
    return fail(GTOMX_ERR_INVALID_ARGUMENT, std::string("short MRC header in ") + path);
  return GTOMX_OK;
}

int gtomx_mrc_read(const char* path, int nframe, void* h_data, size_t capacity, gtomx_mrc_header* header) {
  if (!path || !h_data) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL argument");
  gtomx_mrc_header h;
  GX_TRY(gtomx_mrc_read_header(path, &h));
  GX_TRY(check_header(h));
  if (header) *header = h;
  // ReadMRC (src/IO/mrc.cpp:7-36): one z section (nframe >= 0) or the whole volume,
  // after the 1024-byte header and the extended header
  const size_t section = (size_t)h.nx * h.ny * mode_size(h.mode);
  const size_t bytes = nframe >= 0 ? section : section * (size_t)h.nz;
  if (nframe >= h.nz) return fail(GTOMX_ERR_INVALID_ARGUMENT, "frame index beyond nz");
  if (capacity < bytes) return fail(GTOMX_ERR_INVALID_ARGUMENT, "buffer smaller than the MRC data");
  File in(path, "rb");
  if (!in.f) return fail(GTOMX_ERR_INVALID_ARGUMENT, std::string("cannot open ") + path);
  const long long off = 1024LL + h.nsymbt + (nframe >= 0 ? (long long)section * nframe : 0);
  if (fseeko(in.f, (off_t)off, SEEK_SET) != 0) return fail(GTOMX_ERR_INVALID_ARGUMENT, "seek failed");
  if (std::fread(h_data, 1, bytes, in.f) != bytes)
    return fail(GTOMX_ERR_INVALID_ARGUMENT, std::string("short MRC data in ") + path);
  return GTOMX_OK;
}

int gtomx_mrc_write(const void* h_data, const gtomx_mrc_header* header, const char* path) {
  if (!h_data || !header || !path) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL argument");
  gtomx_mrc_header h = *header;
  GX_TRY(check_header(h));
  // WriteMRC (src/IO/mrc.cpp:68-83) writes no extended header, so none is declared.
  h.nsymbt = 0;
  File out(path, "wb");
  if (!out.f) return fail(GTOMX_ERR_INVALID_ARGUMENT, std::string("cannot create ") + path);
  const size_t bytes = (size_t)h.nx * h.ny * h.nz * mode_size(h.mode);
  if (std::fwrite(&h, 1, sizeof(h), out.f) != sizeof(h) || std::fwrite(h_data, 1, bytes, out.f) != bytes)
    return fail(GTOMX_ERR_DEVICE, std::string("write failed: ") + path);
  return GTOMX_OK;
}

int gtomx_mrc_write_f32(const float* h_data, gtomx_int3 dims, const char* path) {
  if (!h_data) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL argument");
  // WriteMRC(tfloat*, int3, path) (src/IO/mrc.cpp:85-106): default header,
  // float mode, min and max of the data
  gtomx_mrc_header h;
  gtomx_mrc_header_default(&h);
  h.nx = dims.x;
  h.ny = dims.y;
  h.nz = dims.z;
  float mn = 1e30f, mx = -1e30f;
  const size_t n = elements(dims);
  for (size_t i = 0; i < n; ++i) {
    mn = h_data[i] < mn ? h_data[i] : mn;
    mx = h_data[i] > mx ? h_data[i] : mx;
  }
  h.dmin = mn;
  h.dmax = mx;
  return gtomx_mrc_write(h_data, &h, path);
}

int gtomx_mrc_write_device_f32(const float* d_data, gtomx_int3 dims, const char* path, void* stream) {
  if (!d_data || !path) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL argument");
  if (dims.x < 1 || dims.y < 1 || dims.z < 1) return fail(GTOMX_ERR_INVALID_ARGUMENT, "dims must be positive");
  // d_WriteMRC (src/IO/mrc.cpp:108-115)
  std::vector<float> h(elements(dims));
  hipStream_t s = as_stream(stream);
  GX_HIP(hipMemcpyAsync(h.data(), d_data, h.size() * sizeof(float), hipMemcpyDeviceToHost, s));
  GX_HIP(hipStreamSynchronize(s));
  return gtomx_mrc_write_f32(h.data(), dims, path);
}

int gtomx_mrc_to_device_f32(const void* h_data, int mode, size_t elements_, float* d_output, void* stream) {
  if (!h_data || !d_output) return fail(GTOMX_ERR_INVALID_ARGUMENT, "NULL argument");
  if (elements_ == 0) return GTOMX_OK;
  hipStream_t s = as_stream(stream);
  if (mode == 2) {
    GX_HIP(hipMemcpyAsync(d_output, h_data, elements_ * sizeof(float), hipMemcpyHostToDevice, s));
    GX_HIP(hipStreamSynchronize(s));
    return GTOMX_OK;
  }
  if (mode != 0 && mode != 1 && mode != 6)
    return fail(GTOMX_ERR_UNSUPPORTED, "only MRC modes 0, 1, 2 and 6 convert to float");
  const size_t esz = mode_size(mode);
  DevBuf raw;
  GX_TRY(raw.alloc(elements_ * esz));
  GX_HIP(hipMemcpyAsync(raw.p, h_data, elements_ * esz, hipMemcpyHostToDevice, s));
  const unsigned grid = (unsigned)std::min<size_t>((elements_ + 255) / 256, 65536);
  if (mode == 0)  // unsigned bytes, as the reference converts MRC_BYTE
    hipLaunchKernelGGL(k_widen<unsigned char>, dim3(grid), dim3(256), 0, s,
                       static_cast<const unsigned char*>(raw.p), d_output, elements_);
  else if (mode == 1)
    hipLaunchKernelGGL(k_widen<short>, dim3(grid), dim3(256), 0, s, static_cast<const short*>(raw.p), d_output,
                       elements_);
  else
    hipLaunchKernelGGL(k_widen<unsigned short>, dim3(grid), dim3(256), 0, s,
                       static_cast<const unsigned short*>(raw.p), d_output, elements_);
  GX_LAUNCHED("k_widen");
  GX_HIP(hipStreamSynchronize(s));  // the staging buffer is freed on return
  return GTOMX_OK;
}

}  // extern "C"
