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

#include <cstdint>
#include <stdexcept>
#include <string>
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

// d_SphereMask (real, centred), include/Masking.cuh:20; radius NULL uses min(dims)/2.
inline void d_SphereMask(tfloat* d_input, tfloat* d_output, int3 size, tfloat* radius, tfloat sigma,
                         tfloat3* center, bool decentered, int batch = 1) {
  if (center != nullptr || decentered)
    throw std::runtime_error("d_SphereMask: only the centred, non-decentered form is built");
  tfloat r = radius ? *radius : (tfloat)(std::min(std::min(size.x, size.y), size.z > 1 ? size.z : size.x) / 2);
  detail::check(gtomx_sphere_mask(d_input, d_output, detail::i3(size), r, sigma, (unsigned)batch, nullptr),
                "d_SphereMask");
  detail::sync();
}

// d_LocalStd without a Fourier mask, include/Generics.cuh:782.
inline void d_LocalStd(tfloat* d_map, int3 dimsmap, tfloat* d_fouriermask, tfloat localradius, tfloat* d_std,
                       tfloat* d_mean = nullptr) {
  if (d_fouriermask != nullptr) throw std::runtime_error("d_LocalStd: a Fourier mask is not supported");
  detail::check(gtomx_local_std(d_map, detail::i3(dimsmap), localradius, d_std, d_mean, nullptr), "d_LocalStd");
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

}  // namespace gtom

#endif  // GTOM_PICK_HPP
