// gtom_pick.hpp - header-only C++ drop-in for GTOM's template-matching path.
//
// Re-exposes the reference's d_-prefixed host API (namespace gtom, device
// pointers, int3 dims, batch counts) on top of the C ABI in gtomx.h. Like the
// reference these calls are host-synchronous on the default stream and throw
// std::runtime_error on failure (the reference throws on cuFFT errors,
// include/FFT.cuh:45-52, and calls std::terminate on bad dimensions).
//
// Differences forced by the MI355X build, each also noted at its function:
//  * the projector is a device tcomplex* half spectrum instead of the
//    cudaTex Re/Im pair (the author's stated intent, src/Correlation/Picker.cu:27);
//  * d_PickSubTomograms takes an optional per-volume CTF stride;
//  * d_experimentalft is left unmodified.
#ifndef GTOM_PICK_HPP
#define GTOM_PICK_HPP

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "gtomx.h"

namespace gtom {

typedef float tfloat;     // include/Prerequisites.cuh:41
typedef float2 tcomplex;  // include/Prerequisites.cuh:42
typedef float3 tfloat3;   // angles (rot, tilt, psi)
typedef unsigned int uint;

static_assert(sizeof(tcomplex) == sizeof(gtomx_complex), "complex layout");
static_assert(sizeof(int3) == sizeof(gtomx_int3), "int3 layout");
static_assert(sizeof(tfloat3) == sizeof(gtomx_float3), "float3 layout");

namespace detail {
inline void check(int status, const char* where) {
  if (status != GTOMX_OK)
    throw std::runtime_error(std::string(where) + ": " + gtomx_status_string(status) + ": " + gtomx_last_error());
}
inline gtomx_int3 i3(int3 d) { return gtomx_int3{d.x, d.y, d.z}; }
inline int ndims(int3 d) { return d.z > 1 ? 3 : (d.y > 1 ? 2 : 1); }
inline void sync() { check(hipStreamSynchronize(nullptr) == hipSuccess ? GTOMX_OK : GTOMX_ERR_DEVICE, "hipStreamSynchronize"); }
}  // namespace detail

// d_FFTR2C, include/FFT.cuh:70: unnormalised forward transform.
inline void d_FFTR2C(tfloat* const d_input, tcomplex* const d_output, int const ndimensions, int3 const dimensions,
                     int batch = 1) {
  detail::check(gtomx_fft_r2c(d_input, reinterpret_cast<gtomx_complex*>(d_output), ndimensions,
                              detail::i3(dimensions), (unsigned)batch, nullptr), "d_FFTR2C");
  detail::sync();
}

// d_IFFTC2R, include/FFT.cuh:150: inverse transform, renormalised by 1/N unless asked not to.
inline void d_IFFTC2R(tcomplex* const d_input, tfloat* const d_output, int const ndimensions, int3 const dimensions,
                      int batch = 1, bool renormalize = true) {
  detail::check(gtomx_ifft_c2r(reinterpret_cast<const gtomx_complex*>(d_input), d_output, ndimensions,
                               detail::i3(dimensions), (unsigned)batch, renormalize ? 1 : 0, nullptr),
                "d_IFFTC2R");
}

// d_Pad (T_PAD_VALUE), include/Generics.cuh:813.
inline void d_Pad(tfloat* d_input, tfloat* d_output, int3 inputdims, int3 outputdims, tfloat value, int batch = 1) {
  detail::check(gtomx_pad(d_input, d_output, detail::i3(inputdims), detail::i3(outputdims), value, (unsigned)batch,
                          nullptr), "d_Pad");
  detail::sync();
}

// d_RemapFull2FullFFT, include/FFT.cuh:287 (out of place).
inline void d_RemapFull2FullFFT(tfloat* d_input, tfloat* d_output, int3 dims, int batch = 1) {
  detail::check(gtomx_remap_full2fullfft(d_input, d_output, detail::i3(dims), (unsigned)batch, nullptr),
                "d_RemapFull2FullFFT");
  detail::sync();
}

// d_rlnProject (device-array form with host angles), include/Relion.cuh:34.
inline void d_rlnProject(tcomplex* d_volumeft, int3 dimsvolume, tcomplex* d_proj, int3 dimsproj, tfloat3* h_angles,
                         float supersample, uint batch) {
  detail::check(gtomx_project_angles(reinterpret_cast<const gtomx_complex*>(d_volumeft), detail::i3(dimsvolume),
                                     reinterpret_cast<gtomx_complex*>(d_proj), detail::i3(dimsproj),
                                     reinterpret_cast<const gtomx_float3*>(h_angles), supersample, batch, nullptr),
                "d_rlnProject");
}

// d_rlnProject (device matrices), include/Relion.cuh:35; d_matrices holds `batch`
// row-major 3x3 float matrices (the reference passes glm::mat3, column-major of
// the transpose: the same 9 floats).
inline void d_rlnProject(tcomplex* d_volumeft, int3 dimsvolume, tcomplex* d_proj, int3 dimsproj, uint rmax,
                         float* d_matrices, uint batch) {
  detail::check(gtomx_project(reinterpret_cast<const gtomx_complex*>(d_volumeft), detail::i3(dimsvolume),
                              reinterpret_cast<gtomx_complex*>(d_proj), detail::i3(dimsproj), rmax, d_matrices,
                              batch, nullptr), "d_rlnProject");
  detail::sync();
}

// d_ComplexMultiplyByVector (real multiplicators), include/Generics.cuh:71.
inline void d_ComplexMultiplyByVector(tcomplex* d_input, tfloat* d_multiplicators, tcomplex* d_output, size_t elements,
                                      int batch = 1) {
  detail::check(gtomx_complex_multiply_by_vector(reinterpret_cast<const gtomx_complex*>(d_input), d_multiplicators,
                                                 reinterpret_cast<gtomx_complex*>(d_output), elements,
                                                 (unsigned)batch, nullptr), "d_ComplexMultiplyByVector");
  detail::sync();
}

// d_ComplexMultiplyByConjVector, include/Generics.cuh:150.
inline void d_ComplexMultiplyByConjVector(tcomplex* d_input, tcomplex* d_multiplicators, tcomplex* d_output,
                                          size_t elements, int batch = 1) {
  detail::check(gtomx_complex_multiply_by_conj_vector(reinterpret_cast<const gtomx_complex*>(d_input),
                                                      reinterpret_cast<const gtomx_complex*>(d_multiplicators),
                                                      reinterpret_cast<gtomx_complex*>(d_output), elements,
                                                      (unsigned)batch, nullptr), "d_ComplexMultiplyByConjVector");
  detail::sync();
}

// d_DivideSafeByVector, include/Generics.cuh:190.
inline void d_DivideSafeByVector(tfloat* d_input, tfloat* d_divisors, tfloat* d_output, size_t elements,
                                 int batch = 1) {
  detail::check(gtomx_divide_safe_by_vector(d_input, d_divisors, d_output, elements, (unsigned)batch, nullptr),
                "d_DivideSafeByVector");
  detail::sync();
}

// d_SphereMask, include/Masking.cuh:20, for tfloat and tcomplex: radius NULL
// uses min(dims)/2 (z only in 3D), center NULL the volume centre (dims/2).
template <class T>
inline void d_SphereMask(T* d_input, T* d_output, int3 size, tfloat* radius, tfloat sigma, tfloat3* center,
                         bool decentered, int batch = 1) {
  static_assert(std::is_same<T, tfloat>::value || std::is_same<T, tcomplex>::value, "tfloat or tcomplex");
  tfloat r = radius ? *radius : (tfloat)(std::min(std::min(size.x, size.y), size.z > 1 ? size.z : size.x) / 2);
  detail::check(gtomx_sphere_mask_ex(d_input, d_output, detail::i3(size), r, sigma,
                                     reinterpret_cast<const gtomx_float3*>(center), decentered ? 1 : 0,
                                     std::is_same<T, tcomplex>::value ? 1 : 0, (unsigned)batch, nullptr),
                "d_SphereMask");
  detail::sync();
}

// d_LocalStd, include/Generics.cuh:782; d_fouriermask NULL or a real half spectrum.
inline void d_LocalStd(tfloat* d_map, int3 dimsmap, tfloat* d_fouriermask, tfloat localradius, tfloat* d_std,
                       tfloat* d_mean = nullptr) {
  detail::check(gtomx_local_std(d_map, detail::i3(dimsmap), d_fouriermask, localradius, d_std, d_mean, nullptr),
                "d_LocalStd");
}

// d_PickSubTomogramsDiff2, include/Correlation.cuh:111-123, with the
// device-array projector in place of the texture pair.
inline void d_PickSubTomogramsDiff2(tcomplex* d_projectordata, tfloat projectoroversample, int3 dimsprojector,
                                    tcomplex* d_experimentalft, tfloat* d_ctf, int3 dimsvolume, uint nvolumes,
                                    int3 dimsrelevant, tfloat3* h_angles, uint nangles, tfloat* d_bestdiff2,
                                    float* d_bestangle) {
  detail::check(gtomx_pick_subtomograms_diff2(reinterpret_cast<const gtomx_complex*>(d_projectordata),
                                              projectoroversample, detail::i3(dimsprojector),
                                              reinterpret_cast<const gtomx_complex*>(d_experimentalft), d_ctf,
                                              detail::i3(dimsvolume), nvolumes, detail::i3(dimsrelevant),
                                              reinterpret_cast<const gtomx_float3*>(h_angles), nangles, d_bestdiff2,
                                              d_bestangle, nullptr),
                "d_PickSubTomogramsDiff2");
}

// d_LocalPeaks, include/Correlation.cuh:81 / src/Correlation/LocalPeaks.cu:24-62:
// h_peaks[b] is malloc'ed (free() it) when volume b has peaks, as in the
// reference; h_peaksnum[b] receives the count.
inline void d_LocalPeaks(tfloat* d_input, int3** h_peaks, int* h_peaksnum, int3 dims, int localextent,
                         tfloat threshold, int batch = 1) {
  if (batch < 1) throw std::runtime_error("d_LocalPeaks: batch must be positive");
  size_t cap = 4096;
  std::vector<gtomx_int3> buf;
  std::vector<size_t> counts((size_t)batch);
  for (;;) {
    buf.resize(cap * (size_t)batch);
    detail::check(gtomx_local_peaks(d_input, detail::i3(dims), localextent, threshold, (unsigned)batch, buf.data(),
                                    cap, counts.data(), nullptr),
                  "d_LocalPeaks");
    size_t need = 0;
    for (size_t c : counts) need = c > need ? c : need;
    if (need <= cap) break;
    cap = need;
  }
  for (int b = 0; b < batch; ++b) {
    h_peaksnum[b] = (int)counts[(size_t)b];
    if (counts[(size_t)b] == 0) continue;
    h_peaks[b] = static_cast<int3*>(std::malloc(counts[(size_t)b] * sizeof(int3)));
    if (!h_peaks[b]) throw std::runtime_error("d_LocalPeaks: out of host memory");
    for (size_t k = 0; k < counts[(size_t)b]; ++k) {
      const gtomx_int3 q = buf[(size_t)b * cap + k];
      h_peaks[b][k] = make_int3(q.x, q.y, q.z);
    }
  }
}

// T_INTERP_MODE, include/Prerequisites.cuh:239-245.
enum T_INTERP_MODE { T_INTERP_LINEAR = 1, T_INTERP_CUBIC = 2, T_INTERP_FOURIER = 3, T_INTERP_SINC = 4 };

// d_Rotate3DFT, include/Transformation.cuh:54 (complex) and the real-valued
// overload (src/Transformation/Rotation.cu:325): T_INTERP_LINEAR or T_INTERP_CUBIC.
inline void d_Rotate3DFT(tcomplex* d_volume, tcomplex* d_output, int3 dims, tfloat3* h_angles, int nangles,
                         T_INTERP_MODE mode, bool outputzerocentered) {
  detail::check(gtomx_rotate3d_ft(reinterpret_cast<const gtomx_complex*>(d_volume),
                                  reinterpret_cast<gtomx_complex*>(d_output), detail::i3(dims),
                                  reinterpret_cast<const gtomx_float3*>(h_angles), (unsigned)nangles, (int)mode,
                                  outputzerocentered ? 1 : 0, nullptr),
                "d_Rotate3DFT");
}
inline void d_Rotate3DFT(tfloat* d_volume, tfloat* d_output, int3 dims, tfloat3* h_angles, int nangles,
                         T_INTERP_MODE mode, bool outputzerocentered) {
  detail::check(gtomx_rotate3d_ft_real(d_volume, d_output, detail::i3(dims),
                                       reinterpret_cast<const gtomx_float3*>(h_angles), (unsigned)nangles, (int)mode,
                                       outputzerocentered ? 1 : 0, nullptr),
                "d_Rotate3DFT");
}

// d_Rotate2DFT, include/Transformation.cuh:52: `batch` images, image b rotated
// by angles[b] (host array, radians); T_INTERP_LINEAR or T_INTERP_CUBIC.
inline void d_Rotate2DFT(tcomplex* d_input, tcomplex* d_output, int3 dims, tfloat* angles, tfloat maxfreq,
                         T_INTERP_MODE mode, bool isoutputzerocentered, int batch = 1) {
  detail::check(gtomx_rotate2d_ft(reinterpret_cast<const gtomx_complex*>(d_input),
                                  reinterpret_cast<gtomx_complex*>(d_output), detail::i3(dims), angles, maxfreq,
                                  (int)mode, isoutputzerocentered ? 1 : 0, (unsigned)batch, nullptr),
                "d_Rotate2DFT");
}

// struct Peak, include/Correlation.cuh:127-134: a picked particle, filled by
// the caller from the search output (position from d_LocalPeaks / d_Peak,
// angles from the orientation index, fom from the score map).
struct Peak {
  tfloat3 position;
  uint ref;
  tfloat3 angles;
  tfloat fom;
  tfloat relativefom;
};

// T_PEAK_MODE, include/Correlation.cuh:44-52.
enum T_PEAK_MODE { T_PEAK_INTEGER = 1, T_PEAK_SUBCOARSE = 2, T_PEAK_SUBFINE = 3 };

// d_Peak, include/Correlation.cuh:65. The optional cuFFT plans of the
// reference have no counterpart; positions are tfloat3 (x, y, z) per map.
inline void d_Peak(tfloat* d_input, tfloat3* d_positions, tfloat* d_values, int3 dims, T_PEAK_MODE mode,
                   int batch = 1) {
  detail::check(gtomx_peak(d_input, reinterpret_cast<gtomx_float3*>(d_positions), d_values, detail::i3(dims),
                           (int)mode, (unsigned)batch, nullptr),
                "d_Peak");
}

// d_PeakOne2D, include/Correlation.cuh:67.
inline void d_PeakOne2D(tfloat* d_input, float3* d_positions, tfloat* d_values, int2 dims, int2 dimsregion,
                        bool subtractcenter, int batch = 1) {
  detail::check(gtomx_peak_one_2d(d_input, reinterpret_cast<gtomx_float3*>(d_positions), d_values,
                                  gtomx_int3{dims.x, dims.y, 1}, gtomx_int3{dimsregion.x, dimsregion.y, 1},
                                  subtractcenter ? 1 : 0, (unsigned)batch, nullptr),
                "d_PeakOne2D");
  detail::sync();
}

// d_SubpixelMax, include/Correlation.cuh:83: d_output must hold the starting values.
inline void d_SubpixelMax(tfloat* d_input, tfloat* d_output, int3 dims, int subpixsteps) {
  detail::check(gtomx_subpixel_max(d_input, d_output, detail::i3(dims), subpixsteps, nullptr), "d_SubpixelMax");
}

// GetEqualAngularSpacing, src/Helpers/AngularSpacing.cu:46-67.
inline std::vector<float3> GetEqualAngularSpacing(float2 phirange, float2 thetarange, float2 psirange,
                                                  float increment) {
  int n = 0;
  detail::check(gtomx_equal_angular_spacing(phirange.x, phirange.y, thetarange.x, thetarange.y, psirange.x,
                                            psirange.y, increment, nullptr, 0, &n), "GetEqualAngularSpacing");
  std::vector<float3> out((size_t)n);
  detail::check(gtomx_equal_angular_spacing(phirange.x, phirange.y, thetarange.x, thetarange.y, psirange.x,
                                            psirange.y, increment, reinterpret_cast<gtomx_float3*>(out.data()), n,
                                            &n), "GetEqualAngularSpacing");
  return out;
}

// d_PickSubTomograms, include/Correlation.cuh:97-109, with the device-array
// projector d_projectordata ((p/2+1)*p*p complex, FFT order) in place of the
// texture pair, and an optional element stride between per-volume CTFs
// (0: one CTF for all volumes, as in the reference). d_bestangle receives the
// orientation index as float, like the reference.
inline void d_PickSubTomograms(tcomplex* d_projectordata, tfloat projectoroversample, int3 dimsprojector,
                               tcomplex* d_experimentalft, tfloat* d_ctf, int3 dimsvolume, uint nvolumes,
                               tfloat3* h_angles, uint nangles, tfloat maskradius, tfloat* d_bestcorrelation,
                               float* d_bestangle, size_t ctf_volume_stride = 0) {
  gtomx_pick_plan* plan = nullptr;
  detail::check(gtomx_pick_plan_create(&plan, detail::i3(dimsprojector), detail::i3(dimsvolume), 0),
                "d_PickSubTomograms");
  int32_t* d_index = nullptr;
  size_t nvox = (size_t)dimsvolume.x * dimsvolume.y * dimsvolume.z * nvolumes;
  if (hipMalloc((void**)&d_index, nvox * sizeof(int32_t)) != hipSuccess) {
    gtomx_pick_plan_destroy(plan);
    throw std::runtime_error("d_PickSubTomograms: out of device memory");
  }
  int s = gtomx_pick_subtomograms(plan, reinterpret_cast<const gtomx_complex*>(d_projectordata),
                                  projectoroversample, detail::i3(dimsprojector),
                                  reinterpret_cast<const gtomx_complex*>(d_experimentalft), d_ctf,
                                  ctf_volume_stride, detail::i3(dimsvolume), nvolumes,
                                  reinterpret_cast<const gtomx_float3*>(h_angles), nangles, maskradius,
                                  d_bestcorrelation, d_index, d_bestangle, nullptr);
  (void)hipStreamSynchronize(nullptr);
  (void)hipFree(d_index);
  gtomx_pick_plan_destroy(plan);
  detail::check(s, "d_PickSubTomograms");
}

// MRC I/O, include/IO.cuh:13-96, 165-171. HeaderMRC keeps the reference's
// member names over the 1024-byte MRC header (the file format fixes the layout).
enum MRC_DATATYPE { MRC_BYTE = 0, MRC_SHORT = 1, MRC_FLOAT = 2, MRC_SHORTCOMPLEX = 3, MRC_FLOATCOMPLEX = 4,
                    MRC_UNSIGNEDSHORT = 6, MRC_RGB = 16 };
struct HeaderMRC {
  int3 dimensions;
  MRC_DATATYPE mode;
  int3 startsubimage;
  int3 griddimensions;
  float3 pixelsize;
  float3 angles;
  int3 maporder;
  float minvalue, maxvalue, meanvalue;
  int spacegroup;
  int extendedbytes;
  short creatid;
  unsigned char extradata1[30];
  short nint, nreal;
  unsigned char extradata2[28];
  short idtype, lens, nd1, nd2, vd1, vd2;
  float3 tiltoriginal, tiltcurrent, origin;
  unsigned char cmap[4];
  unsigned char stamp[4];
  float stddevvalue;
  int numlabels;
  unsigned char labels[10][80];
  HeaderMRC() { gtomx_mrc_header_default(reinterpret_cast<gtomx_mrc_header*>(this)); }
};
static_assert(sizeof(HeaderMRC) == sizeof(gtomx_mrc_header), "HeaderMRC is the 1024-byte MRC header");

inline HeaderMRC ReadMRCHeader(std::string path) {
  HeaderMRC h;
  detail::check(gtomx_mrc_read_header(path.c_str(), reinterpret_cast<gtomx_mrc_header*>(&h)), "ReadMRCHeader");
  return h;
}
// *data receives pinned host memory (hipHostFree it), as the reference's cudaMallocHost.
inline void ReadMRC(std::string path, void** data, int nframe = -1) {
  HeaderMRC h = ReadMRCHeader(path);
  static const size_t sizes[17] = {1, 2, 4, 4, 8, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 3};
  if ((int)h.mode < 0 || (int)h.mode > 16) throw std::runtime_error("ReadMRC: unknown mode");
  size_t bytes = (size_t)h.dimensions.x * h.dimensions.y * sizes[(int)h.mode] * (nframe >= 0 ? 1 : h.dimensions.z);
  if (hipHostMalloc(data, bytes, hipHostMallocDefault) != hipSuccess) throw std::runtime_error("ReadMRC: out of memory");
  detail::check(gtomx_mrc_read(path.c_str(), nframe, *data, bytes, nullptr), "ReadMRC");
}
inline void WriteMRC(void* data, HeaderMRC header, std::string path) {
  detail::check(gtomx_mrc_write(data, reinterpret_cast<const gtomx_mrc_header*>(&header), path.c_str()), "WriteMRC");
}
inline void WriteMRC(tfloat* data, int3 dims, std::string path) {
  detail::check(gtomx_mrc_write_f32(data, detail::i3(dims), path.c_str()), "WriteMRC");
}
inline void d_WriteMRC(tfloat* d_data, int3 dims, std::string path) {
  detail::check(gtomx_mrc_write_device_f32(d_data, detail::i3(dims), path.c_str(), nullptr), "d_WriteMRC");
}
// MixedToDeviceTfloat(void*, tfloat*, MRC_DATATYPE, size_t), src/Helpers/Memory.cu:424-436.
inline void MixedToDeviceTfloat(void* h_input, tfloat* d_output, MRC_DATATYPE datatype, size_t elements) {
  detail::check(gtomx_mrc_to_device_f32(h_input, (int)datatype, elements, d_output, nullptr), "MixedToDeviceTfloat");
}

// d_FFTFullPad on real volumes, include/FFT.cuh:297.
inline void d_FFTFullPad(tfloat* d_input, tfloat* d_output, int3 olddims, int3 newdims, int batch = 1) {
  detail::check(gtomx_fft_full_pad(d_input, d_output, detail::i3(olddims), detail::i3(newdims), (unsigned)batch,
                                   nullptr),
                "d_FFTFullPad");
  detail::sync();
}

// d_FFTCrop / d_FFTPad (half spectra) and d_FFTFullCrop / d_FFTFullPad (full
// arrays), include/FFT.cuh:294-297, for tfloat and tcomplex.
namespace detail {
template <class T> inline void fft_resize(int op, T* in, T* out, int3 od, int3 nd, int batch, const char* name) {
  static_assert(std::is_same<T, tfloat>::value || std::is_same<T, tcomplex>::value, "tfloat or tcomplex");
  check(gtomx_fft_resize(op, std::is_same<T, tcomplex>::value ? 1 : 0, in, out, i3(od), i3(nd), (unsigned)batch,
                         nullptr),
        name);
  sync();
}
}  // namespace detail
template <class T> inline void d_FFTCrop(T* d_input, T* d_output, int3 olddims, int3 newdims, int batch = 1) {
  detail::fft_resize(GTOMX_FFT_CROP, d_input, d_output, olddims, newdims, batch, "d_FFTCrop");
}
template <class T> inline void d_FFTPad(T* d_input, T* d_output, int3 olddims, int3 newdims, int batch = 1) {
  detail::fft_resize(GTOMX_FFT_PAD, d_input, d_output, olddims, newdims, batch, "d_FFTPad");
}
template <class T> inline void d_FFTFullCrop(T* d_input, T* d_output, int3 olddims, int3 newdims, int batch = 1) {
  detail::fft_resize(GTOMX_FFT_FULL_CROP, d_input, d_output, olddims, newdims, batch, "d_FFTFullCrop");
}
inline void d_FFTFullPad(tcomplex* d_input, tcomplex* d_output, int3 olddims, int3 newdims, int batch = 1) {
  detail::fft_resize(GTOMX_FFT_FULL_PAD, d_input, d_output, olddims, newdims, batch, "d_FFTFullPad");
}

// d_Sum, include/Generics.cuh:585, for float, double and int.
inline int sum_dispatch(const float* in, float* out, size_t n, unsigned b) { return gtomx_sum_f32(in, out, n, b, nullptr); }
inline int sum_dispatch(const double* in, double* out, size_t n, unsigned b) {
  return gtomx_sum_f64(in, out, n, b, nullptr);
}
inline int sum_dispatch(const int* in, int* out, size_t n, unsigned b) {
  return gtomx_sum_i32(reinterpret_cast<const int32_t*>(in), reinterpret_cast<int32_t*>(out), n, b, nullptr);
}
template <class T> inline void d_Sum(T* d_input, T* d_output, size_t n, int batch = 1) {
  detail::check(sum_dispatch(d_input, d_output, n, (unsigned)batch), "d_Sum");
}

// T_NORM_MODE, include/ImageManipulation.cuh:37-47.
enum T_NORM_MODE {
  T_NORM_NONE = 0, T_NORM_MEAN01STD = 1, T_NORM_PHASE = 2, T_NORM_STD1 = 3, T_NORM_STD2 = 4, T_NORM_STD3 = 5,
  T_NORM_OSCAR = 6, T_NORM_CUSTOM = 7
};

// d_NormMonolithic, include/ImageManipulation.cuh:49: mean0/std1 per item
// whatever the mode, as the reference's kernel does.
inline void d_NormMonolithic(tfloat* d_input, tfloat* d_output, size_t elements, T_NORM_MODE mode, int batch) {
  (void)mode;
  detail::check(gtomx_norm_monolithic(d_input, d_output, elements, (unsigned)batch, nullptr), "d_NormMonolithic");
}

// d_NormFTMonolithic, include/ImageManipulation.cuh:55.
inline void d_NormFTMonolithic(tcomplex* d_input, tcomplex* d_output, size_t elements, int batch) {
  detail::check(gtomx_norm_ft_monolithic(reinterpret_cast<const gtomx_complex*>(d_input),
                                         reinterpret_cast<gtomx_complex*>(d_output), elements, (unsigned)batch,
                                         nullptr),
                "d_NormFTMonolithic");
}

// d_Norm with T_NORM_MEAN01STD and no mask, include/ImageManipulation.cuh:48
// (other modes and masks are not built).
inline void d_NormMean01Std(tfloat* d_input, tfloat* d_output, size_t elements, int batch = 1) {
  detail::check(gtomx_norm_mean01std(d_input, d_output, elements, (unsigned)batch, nullptr), "d_Norm");
}

// CTFParams, include/CTF.cuh:11-43: same fields, order and defaults (SI units,
// angles in radians), laid out as gtomx_ctf_params_full.
struct CTFParams {
  tfloat pixelsize = 1e-10f, pixeldelta = 0, pixelangle = 0;
  tfloat Cs = 2e-3f, voltage = 300e3f;
  tfloat defocus = -3e-6f, astigmatismangle = 0, defocusdelta = 0;
  tfloat amplitude = 0.07f;
  tfloat Bfactor = 0, Bfactordelta = 0, Bfactorangle = 0;
  tfloat scale = 1.0f, phaseshift = 0;
};
static_assert(sizeof(CTFParams) == sizeof(gtomx_ctf_params_full), "CTFParams layout");

// d_CTFSimulate, include/CTF.cuh:317: batch CTFs at n (k, angle) addresses,
// d_gammacorrection optional (nullptr), output batch x n.
inline void d_CTFSimulate(CTFParams* h_params, float2* d_addresses, float* d_gammacorrection, tfloat* d_output,
                          uint n, bool amplitudesquared = false, bool ignorefirstpeak = false, int batch = 1) {
  detail::check(gtomx_ctf_simulate(reinterpret_cast<const gtomx_ctf_params_full*>(h_params),
                                   reinterpret_cast<const float*>(d_addresses), d_gammacorrection, d_output, n,
                                   amplitudesquared ? 1 : 0, ignorefirstpeak ? 1 : 0, (unsigned)batch, nullptr),
                "d_CTFSimulate");
}

// CTFTiltParams, include/CTF.cuh:136-194: the plane geometry runs in the
// library (double precision, the reference's rotation matrix as written).
class CTFTiltParams {
 public:
  tfloat imageangle;
  float2 stageangle;
  float2 specimenangle;
  CTFParams centerparams;

  CTFTiltParams(tfloat3 _angles, CTFParams _centerparams)
      : imageangle(_angles.z), stageangle(make_float2(_angles.x, _angles.y)), specimenangle(make_float2(0, 0)),
        centerparams(_centerparams) {}
  CTFTiltParams(tfloat _imageangle, float2 _stageangle, float2 _specimenangle, CTFParams _centerparams)
      : imageangle(_imageangle), stageangle(_stageangle), specimenangle(_specimenangle),
        centerparams(_centerparams) {}

  void GetParamsGrid2D(int2 dimsimage, int2 dimsregion, int3* h_positions, uint npoints, CTFParams* h_params) {
    gtomx_ctf_tilt_params t;
    t.imageangle = imageangle;
    t.stageangle[0] = stageangle.x;
    t.stageangle[1] = stageangle.y;
    t.specimenangle[0] = specimenangle.x;
    t.specimenangle[1] = specimenangle.y;
    static_assert(sizeof(CTFParams) == sizeof(gtomx_ctf_params_full), "CTFParams layout");
    std::memcpy(&t.centerparams, &centerparams, sizeof(CTFParams));
    detail::check(gtomx_ctf_tilt_params_grid(&t, gtomx_int3{dimsimage.x, dimsimage.y, 1},
                                             gtomx_int3{dimsregion.x, dimsregion.y, 1},
                                             reinterpret_cast<const gtomx_int3*>(h_positions), npoints,
                                             reinterpret_cast<gtomx_ctf_params_full*>(h_params)),
                  "CTFTiltParams::GetParamsGrid2D");
  }

  void GetZGrid2D(int2 dimsimage, int2 dimsregion, int3* h_positions, uint npoints, float* h_zvalues) {
    std::vector<CTFParams> p(npoints);
    GetParamsGrid2D(dimsimage, dimsregion, h_positions, npoints, p.data());
    for (uint n = 0; n < npoints; ++n) h_zvalues[n] = p[n].defocus;
  }
};

// class Picker, include/Correlation.cuh:136-164. The reference's members
// (cuFFT plans, textures, scratch) live inside the library object; the public
// dimensions are kept. 3D volumes, or 2D images with a square 2D reference
// (in-plane orientations).
class Picker {
 public:
  int3 dimsref{}, dimsimage{};
  uint ndims = 0;

  Picker() = default;
  Picker(const Picker&) = delete;
  Picker& operator=(const Picker&) = delete;
  ~Picker() { gtomx_picker_destroy(handle_); }

  void Initialize(tfloat* _d_ref, int3 _dimsref, tfloat* _d_refmask, int3 _dimsimage) {
    gtomx_picker_destroy(handle_);
    handle_ = nullptr;
    dimsref = _dimsref;
    dimsimage = _dimsimage;
    ndims = (uint)detail::ndims(_dimsref);
    detail::check(gtomx_picker_create(&handle_, _d_ref, detail::i3(_dimsref), _d_refmask, detail::i3(_dimsimage), 0,
                                      nullptr),
                  "Picker::Initialize");
    detail::sync();
  }
  // Normalises _d_image and leaves it holding its square, as the reference does;
  // _d_ctf is kept by pointer.
  void SetImage(tfloat* _d_image, tfloat* _d_ctf) {
    detail::check(gtomx_picker_set_image(handle_, _d_image, _d_ctf, nullptr), "Picker::SetImage");
    detail::sync();
  }
  void PerformCorrelation(tfloat anglestep, tfloat* d_bestccf, tfloat3* d_bestangle) {
    detail::check(gtomx_picker_correlate(handle_, anglestep, d_bestccf, reinterpret_cast<gtomx_float3*>(d_bestangle),
                                         nullptr),
                  "Picker::PerformCorrelation");
    detail::sync();
  }

 private:
  gtomx_picker* handle_ = nullptr;
};

// class TomoPicker, include/Correlation.cuh:209-257: the tilt-series picker
// (reference, mask and the images' tilt geometry; correlation maps of every
// tilt image back-projected into a volume). The reference's public buffers
// (textures, spectra, plans) live inside the library object; the dimensions
// and flags are kept.
class TomoPicker {
 public:
  int3 dimsref{}, dimsrefpadded{};
  int2 dimsimage{};
  uint nimages = 0;
  bool ismaskcircular = true;

  TomoPicker() = default;
  TomoPicker(const TomoPicker&) = delete;
  TomoPicker& operator=(const TomoPicker&) = delete;
  ~TomoPicker() { gtomx_tomopicker_destroy(handle_); }

  void Initialize(tfloat* _d_ref, int3 _dimsref, tfloat* _d_refmask, bool _ismaskcircular, int2 _dimsimage,
                  uint _nimages) {
    gtomx_tomopicker_destroy(handle_);
    handle_ = nullptr;
    dimsref = _dimsref;
    dimsrefpadded = int3{2 * _dimsref.x, 2 * _dimsref.y, 2 * _dimsref.z};
    dimsimage = _dimsimage;
    ismaskcircular = _ismaskcircular;
    nimages = _nimages;
    detail::check(gtomx_tomopicker_create(&handle_, _d_ref, detail::i3(_dimsref), _d_refmask, _ismaskcircular ? 1 : 0,
                                          gtomx_int3{_dimsimage.x, _dimsimage.y, 1}, _nimages, nullptr),
                  "TomoPicker::Initialize");
    detail::sync();
  }
  // Weights _d_image in place (d_Exact2DWeighting of the tilt geometry), as the reference does.
  void SetImage(tfloat* _d_image, tfloat3* _h_imageangles, tfloat* _h_imageweights) {
    detail::check(gtomx_tomopicker_set_image(handle_, _d_image, reinterpret_cast<const gtomx_float3*>(_h_imageangles),
                                             _h_imageweights, nullptr),
                  "TomoPicker::SetImage");
    detail::sync();
  }
  void PerformCorrelation(tfloat* d_corr, tfloat3* d_corrangle, int3 dimscorr, tfloat anglestep) {
    detail::check(gtomx_tomopicker_correlate(handle_, d_corr, reinterpret_cast<gtomx_float3*>(d_corrangle),
                                             detail::i3(dimscorr), anglestep, nullptr),
                  "TomoPicker::PerformCorrelation");
    detail::sync();
  }

 private:
  gtomx_tomopicker* handle_ = nullptr;
};

// class ProbabilisticPicker, include/Correlation.cuh:166-207, circular and
// non-circular masks. The reference keeps its low-pass references and masks in pinned host
// memory and uploads one per PerformCorrelation; here they stay on the device.
// PickImage is declared but never defined in the reference and is absent here.
class ProbabilisticPicker {
 public:
  int3 dimsref{}, dimsimage{}, dimslowpass{};
  uint ndims = 0;
  uint nrefs = 0;
  bool ismaskcircular = true;
  bool doctf = false;

  ProbabilisticPicker() = default;
  ProbabilisticPicker(const ProbabilisticPicker&) = delete;
  ProbabilisticPicker& operator=(const ProbabilisticPicker&) = delete;
  ~ProbabilisticPicker() { gtomx_prob_picker_destroy(handle_); }

  void Initialize(tfloat* _d_refs, int3 _dimsref, uint _nrefs, tfloat* _d_refmasks, bool _ismaskcircular, bool _doctf,
                  int3 _dimsimage, uint _lowpassfreq) {
    gtomx_prob_picker_destroy(handle_);
    handle_ = nullptr;
    dimsref = _dimsref;
    nrefs = _nrefs;
    ndims = (uint)detail::ndims(_dimsref);
    dimsimage = _dimsimage;
    const int lp2 = (int)_lowpassfreq * 2;
    dimslowpass = ndims == 3 ? make_int3(lp2, lp2, lp2) : make_int3(lp2, lp2, 1);
    ismaskcircular = _ismaskcircular;
    doctf = _doctf;
    detail::check(gtomx_prob_picker_create(&handle_, _d_refs, detail::i3(_dimsref), _nrefs, _d_refmasks,
                                           _ismaskcircular ? 1 : 0, _doctf ? 1 : 0, detail::i3(_dimsimage),
                                           _lowpassfreq, nullptr),
                  "ProbabilisticPicker::Initialize");
    detail::sync();
  }
  void SetImage(tfloat* _d_image, tfloat* _d_ctf) {
    detail::check(gtomx_prob_picker_set_image(handle_, _d_image, _d_ctf, nullptr), "ProbabilisticPicker::SetImage");
    detail::sync();
  }
  void PerformCorrelation(uint n, tfloat anglestep, tfloat* d_bestccf, tfloat3* d_bestangle, int* d_bestref) {
    detail::check(gtomx_prob_picker_correlate(handle_, n, anglestep, d_bestccf,
                                              reinterpret_cast<gtomx_float3*>(d_bestangle), d_bestref, nullptr),
                  "ProbabilisticPicker::PerformCorrelation");
    detail::sync();
  }

 private:
  gtomx_prob_picker* handle_ = nullptr;
};

}  // namespace gtom

#endif  // GTOM_PICK_HPP
