/*
 * gtomx.h - C ABI of the MI355X (gfx950) implementation of GTOM's FFT
 * template-matching path.
 *
 * Conventions shared by every entry point:
 *  - All d_ pointers are device pointers (HIP), h_ pointers are host pointers.
 *    The caller owns every buffer passed in.
 *  - Real volumes are x-fastest, element (x, y, z) at (z*dims.y + y)*dims.x + x.
 *  - Half spectra ("ft") hold x in [0, dims.x/2]: element (x, y, z) at
 *    (z*dims.y + y)*(dims.x/2 + 1) + x, with y and z in FFT order. Complex
 *    values are interleaved float pairs (gtomx_complex, the layout of float2 /
 *    tcomplex in the reference, include/Prerequisites.cuh:42).
 *  - `stream` is a hipStream_t passed as void*; NULL selects the default
 *    stream. Entry points enqueue work on that stream and return without
 *    waiting for it unless documented otherwise.
 *  - Every function returns a gtomx_status; no C++ exception crosses the ABI.
 *    gtomx_last_error() gives a human-readable message for the calling thread.
 *  - Supported transform sizes: every axis 2^a in [8, 4096] or 3*2^a in
 *    [12, 3072] (an axis of extent 1 is allowed for 2D and 1D data). The
 *    search (gtomx_pick_*) takes volume sides up to 256 (fused passes) and
 *    512, 768 or 1024 (through the generic transforms), and image sides up
 *    to 4096 (above 256 through the generic transforms); the transforms and
 *    the pickers built on them take the full range, strided axes above 256 in
 *    two passes with stream-ordered scratch (hipMallocAsync).
 *
 * The reference interface each function replaces is cited as file:line of
 * the reference tree (dtegunov/gtom).
 */
#ifndef GTOMX_H
#define GTOMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int x, y, z; } gtomx_int3;       /* int3 */
typedef struct { float x, y; } gtomx_complex;     /* tcomplex, include/Prerequisites.cuh:42 */
typedef struct { float x, y, z; } gtomx_float3;   /* tfloat3, include/Prerequisites.cuh:62-73 */

typedef enum {
  GTOMX_OK = 0,
  GTOMX_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad dims, zero batch, ... */
  GTOMX_ERR_UNSUPPORTED = 2,      /* valid request outside the supported sizes */
  GTOMX_ERR_OUT_OF_MEMORY = 3,    /* device or pinned-host allocation failed */
  GTOMX_ERR_DEVICE = 4,           /* a HIP runtime call or kernel launch failed */
  GTOMX_ERR_INTERNAL = 5
} gtomx_status;

/* Library version and diagnostics (no reference counterpart). */
const char* gtomx_version(void);
const char* gtomx_status_string(int status);
const char* gtomx_last_error(void);

/* Fault-injection hook for host-side tests of the error paths (no reference
 * counterpart): makes the n-th following device allocation made by this
 * library fail with GTOMX_ERR_OUT_OF_MEMORY (n = 0 disables). */
void gtomx_debug_fail_alloc_after(int n);

/* ------------------------------------------------------------------------ */
/* Angles and rotation matrices (host)                                       */
/* ------------------------------------------------------------------------ */

/* GetEqualAngularSpacing, src/Helpers/AngularSpacing.cu:46-67 (restated with
 * the reference's float32 loop counters). Writes min(count, capacity)
 * (rot, tilt, psi) triples in radians to h_angles (may be NULL) and the total
 * count to *count. */
int gtomx_equal_angular_spacing(float phimin, float phimax, float thetamin, float thetamax,
                                float psimin, float psimax, float increment,
                                gtomx_float3* h_angles, int capacity, int* count);

/* Fibonacci-sphere directions x in-plane steps (no reference counterpart):
 * direction i of ndirs has z_i = 1 - (2i+1)/ndirs, phi_i = i*pi*(3 - sqrt 5);
 * psi_j = j * 2*pi/npsi. Output index i*npsi + j, as ZYZ (rot, tilt, psi).
 * Computed in double. h_angles must hold ndirs*npsi triples. */
int gtomx_fibonacci_angles(int ndirs, int npsi, gtomx_float3* h_angles);

/* transpose(Matrix3Euler(angles)) * oversample for each angle, row-major
 * 3x3 per orientation, computed in double and rounded to float
 * (src/Relion/Project.cu:12-17, src/Helpers/Euler.cu:39-68). */
int gtomx_projector_matrices(const gtomx_float3* h_angles, unsigned n, float oversample,
                             float* h_matrices);

/* ------------------------------------------------------------------------ */
/* FFT (replaces the cuFFT wrappers)                                         */
/* ------------------------------------------------------------------------ */

/* d_FFTR2C, include/FFT.cuh:70 / src/FFT/FFT.cu:7-42: unnormalised forward
 * transform of `batch` real volumes into half spectra. Extents: 2^a up to 4096
 * and 3*2^a up to 3072 natively, any other extent up to 2048 by the chirp-z
 * (Bluestein) form over the native power-of-two transforms. */
int gtomx_fft_r2c(const float* d_input, gtomx_complex* d_output, int ndims, gtomx_int3 dims,
                  unsigned batch, void* stream);

/* d_IFFTC2R, include/FFT.cuh:150 / src/FFT/IFFT.cu:8-56: inverse transform of
 * half spectra; renormalize != 0 multiplies by 1/N. The result is the real
 * part of the inverse DFT of the Hermitian-completed spectrum. The input is
 * not modified (the reference's cuFFT call may clobber it). Extents as for
 * gtomx_fft_r2c. */
int gtomx_ifft_c2r(const gtomx_complex* d_input, float* d_output, int ndims, gtomx_int3 dims,
                   unsigned batch, int renormalize, void* stream);

/* ------------------------------------------------------------------------ */
/* Projector preparation and projection                                      */
/* ------------------------------------------------------------------------ */

/* d_Pad with T_PAD_VALUE, include/Generics.cuh:813 / src/Generics/Padding.cu:22-45:
 * centred copy of `batch` real volumes into outputdims, filling with value. */
int gtomx_pad(const float* d_input, float* d_output, gtomx_int3 inputdims, gtomx_int3 outputdims,
              float value, unsigned batch, void* stream);

/* d_RemapFull2FullFFT, include/FFT.cuh:287 / src/FFT/FFTRemap.cu:89-106:
 * ifftshift of real volumes (d_input != d_output required). */
int gtomx_remap_full2fullfft(const float* d_input, float* d_output, gtomx_int3 dims,
                             unsigned batch, void* stream);

/* d_rlnProject (device-array form), include/Relion.cuh:35 /
 * src/Relion/Project.cu:25-43, kernel src/Relion/Project.cu:89-177: samples
 * the (p/2+1, p, p) projector half spectrum at M_b * k for every half-spectrum
 * bin k of a dimsproj volume with |k| <= rmax, for `batch` row-major 3x3
 * matrices (device memory, 9 floats each). dimsproj = (n, n, 1) gives 2D
 * central sections (k = (x, y, 0), Project3DArraytoNDKernel<2>). */
int gtomx_project(const gtomx_complex* d_volumeft, gtomx_int3 dimsvolume, gtomx_complex* d_proj,
                  gtomx_int3 dimsproj, unsigned rmax, const float* d_matrices, unsigned batch,
                  void* stream);

/* d_rlnProject (angles form), include/Relion.cuh:34 / src/Relion/Project.cu:12-23:
 * matrices from host angles (transpose(Euler) * supersample), rmax = dimsproj.x/2. */
int gtomx_project_angles(const gtomx_complex* d_volumeft, gtomx_int3 dimsvolume,
                         gtomx_complex* d_proj, gtomx_int3 dimsproj,
                         const gtomx_float3* h_angles, float supersample, unsigned batch,
                         void* stream);

/* ------------------------------------------------------------------------ */
/* CTF                                                                       */
/* ------------------------------------------------------------------------ */

/* CTF parameters in SI units, the subset of CTFParams (include/CTF.cuh:11-43)
 * used by the radially symmetric 3D form. */
typedef struct {
  float pixelsize;   /* m (default 1e-10) */
  float Cs;          /* m (2e-3) */
  float voltage;     /* V (300e3) */
  float defocus;     /* m, negative = underfocus (-3e-6) */
  float amplitude;   /* amplitude contrast (0.07) */
  float phaseshift;  /* rad (0) */
  float Bfactor;     /* m^2 (0) */
  float scale;       /* (1) */
} gtomx_ctf_params;

/* Radially symmetric CTF on the half spectrum of a cube of side n (or of an
 * n x n image, dims.z = 1), one array of (n/2+1)*n*dims.z floats per parameter set: d_GetCTF<false,false>
 * (include/CTF.cuh:195-223) with CTFParamsLean (include/CTF.cuh:100-134) and
 * the 3D frequency mapping of src/CTF/Wiener.cu:99-110. Evaluated in double. */
int gtomx_ctf_3d(const gtomx_ctf_params* h_params, unsigned batch, gtomx_int3 dims, float* d_ctf,
                 void* stream);

/* CTFParams, include/CTF.cuh:11-43, all fields (SI units, angles in radians). */
typedef struct {
  float pixelsize, pixeldelta, pixelangle;
  float Cs, voltage;
  float defocus, astigmatismangle, defocusdelta;
  float amplitude;
  float Bfactor, Bfactordelta, Bfactorangle;
  float scale, phaseshift;
} gtomx_ctf_params_full;

/* d_CTFSimulate, include/CTF.cuh:317 / src/CTF/Simulate.cu:22-47 (kernel :117-139):
 * the CTF of each of `batch` parameter sets at n addresses (d_addresses: n
 * (k, angle) float pairs, k in cycles per pixel) with astigmatic defocus
 * defocus + defocusdelta/2 cos 2(angle - astigmatismangle), anisotropic pixel
 * size and B-factor, optional per-address gamma correction (may be NULL);
 * output batch x n floats. Evaluated in double. Synchronous. */
int gtomx_ctf_simulate(const gtomx_ctf_params_full* h_params, const float* d_addresses,
                       const float* d_gammacorrection, float* d_output, unsigned n, int amplitudesquared,
                       int ignorefirstpeak, unsigned batch, void* stream);

/* CTFTiltParams, include/CTF.cuh:136-158: the specimen plane tilted by the
 * stage and the specimen, each an (in-plane axis angle, rotation about it)
 * pair; imageangle rotates image coordinates. Radians. */
typedef struct {
  float imageangle;
  float stageangle[2];
  float specimenangle[2];
  gtomx_ctf_params_full centerparams;
} gtomx_ctf_tilt_params;

/* CTFTiltParams::GetParamsGrid2D, include/CTF.cuh:182-194 (GetZGrid2D
 * :160-180): for npoints regions of dimsregion at origins h_positions (x, y) of
 * an image of dimsimage, centerparams with the defocus of the tilted plane
 * under the region centre. The plane normal is built from the reference's
 * rotation matrix as written (its element [0][0] is c + a.x + a.x(1 - c),
 * src/Helpers/Euler.cu:319; 1 + a.x for an untilted stage), which changes the
 * plane whenever the specimen tilt gives the normal an x component.
 * Host computation in double; no device work. */
int gtomx_ctf_tilt_params_grid(const gtomx_ctf_tilt_params* tilt, gtomx_int3 dimsimage, gtomx_int3 dimsregion,
                               const gtomx_int3* h_positions, unsigned npoints, gtomx_ctf_params_full* h_params);

/* The same on the (n/2+1) x n half spectrum of n x n images (dims.z = 1):
 * bin (x, y) at k = |(x, y)|/n, angle = atan2(y, x) (y in [-n/2, n/2)), one array per parameter
 * set: per-micrograph astigmatic CTFs for the 2D branch (no reference
 * counterpart beyond d_CTFSimulate itself). Synchronous. */
int gtomx_ctf_2d(const gtomx_ctf_params_full* h_params, unsigned batch, gtomx_int3 dims, float* d_ctf, void* stream);

/* ------------------------------------------------------------------------ */
/* Element-wise pieces of the path (drop-ins for the unfused reference)      */
/* ------------------------------------------------------------------------ */

/* d_ComplexMultiplyByVector (real multiplicators), include/Generics.cuh:71 /
 * src/Generics/Arithmetics.cu:192-197: out[b][i] = in[b][i] * v[i]. */
int gtomx_complex_multiply_by_vector(const gtomx_complex* d_input, const float* d_multiplicators,
                                     gtomx_complex* d_output, size_t elements, unsigned batch,
                                     void* stream);

/* d_ComplexMultiplyByConjVector, include/Generics.cuh:150 /
 * src/Generics/Arithmetics.cu:220-225: out[b][i] = in[b][i] * conj(v[i]). */
int gtomx_complex_multiply_by_conj_vector(const gtomx_complex* d_input,
                                          const gtomx_complex* d_multiplicators,
                                          gtomx_complex* d_output, size_t elements,
                                          unsigned batch, void* stream);

/* UpdateCorrelationKernel, src/Correlation/SubTomograms.cu:138-158: running
 * max over `batch` maps of `vectorlength` floats with strict '>', the first
 * maximum winning; d_bestindex receives batchoffset + b (int32). */
int gtomx_update_max_argmax(const float* d_correlation, size_t vectorlength, unsigned batch,
                            int batchoffset, float* d_best, int32_t* d_bestindex, void* stream);

/* d_DivideSafeByVector, include/Generics.cuh:190 / src/Generics/Arithmetics.cu:478-493:
 * out = |div| > 1e-15 ? in / div : 0. */
int gtomx_divide_safe_by_vector(const float* d_input, const float* d_divisors, float* d_output,
                                size_t elements, unsigned batch, void* stream);

/* d_SphereMask (real, not decentered), include/Masking.cuh:20 /
 * src/Masking/SphereMask.cu:18-33, 49-106: soft sphere with raised-cosine
 * edge of width sigma around (dims/2). */
int gtomx_sphere_mask(const float* d_input, float* d_output, gtomx_int3 dims, float radius,
                      float sigma, unsigned batch, void* stream);

/* d_SphereMask in full, include/Masking.cuh:20 / src/Masking/SphereMask.cu:18-33
 * (kernel :49-106): center NULL = (dims/2), a float centre is subtracted from
 * each integer index and the difference truncated to int as in the kernel;
 * decentered measures FFT-order distances from the origin (index i < n/2+1 ?
 * i : i - n) and ignores center; z contributes only when dims.z > 1. Real
 * (complex_values = 0) or interleaved complex data, both parts scaled. */
int gtomx_sphere_mask_ex(const void* d_input, void* d_output, gtomx_int3 dims, float radius, float sigma,
                         const gtomx_float3* center, int decentered, int complex_values, unsigned batch,
                         void* stream);

/* d_LocalStd, include/Generics.cuh:782 / src/Generics/LocalStd.cu:10-118:
 * local standard deviation (and, if d_mean is not NULL, mean) of d_map under a
 * soft sphere of localradius. d_fouriermask is NULL or a real half spectrum
 * (FFT order, dims.x/2+1 fastest) that multiplies the sphere's spectrum; the
 * mask sum is then that of the filtered sphere. The variance is combined in
 * double. Allocates and frees its own scratch. */
int gtomx_local_std(const float* d_map, gtomx_int3 dims, const float* d_fouriermask, float localradius,
                    float* d_std, float* d_mean, void* stream);

/* ------------------------------------------------------------------------ */
/* The squared-difference scorer                                             */
/* ------------------------------------------------------------------------ */

/* d_PickSubTomogramsDiff2, include/Correlation.cuh:111-123 /
 * src/Correlation/SubTomograms.cu:165-313 (device-array projector instead of
 * the texture pair): for each of nvolumes cubes (E and CTF of (n/2+1)*n*n
 * each, one CTF per volume), every voxel of dimsrelevant (the shift
 * d = index - dimsrelevant/2 around the volume origin) and every orientation,
 * score = sum over the half spectrum of |R ctf e^{-2 pi i k.d/n} - E|^2;
 * d_bestdiff2 receives the smallest score (started at 1e30, strict '<') and
 * d_bestangle its orientation index as float (Elements(dimsrelevant) per
 * volume). Synchronous (its scratch is freed on return). */
int gtomx_pick_subtomograms_diff2(const gtomx_complex* d_projectorft, float projectoroversample,
                                  gtomx_int3 dimsprojector, const gtomx_complex* d_experimentalft,
                                  const float* d_ctf, gtomx_int3 dimsvolume, unsigned nvolumes,
                                  gtomx_int3 dimsrelevant, const gtomx_float3* h_angles, unsigned nangles,
                                  float* d_bestdiff2, float* d_bestangle, void* stream);

/* ------------------------------------------------------------------------ */
/* The Picker class (whole-volume matcher)                                   */
/* ------------------------------------------------------------------------ */

/* Fourier-space resizing, include/FFT.cuh:294-297 (src/FFT/FFTResize.cu):
 * GTOMX_FFT_CROP = d_FFTCrop and GTOMX_FFT_PAD = d_FFTPad on half spectra
 * ((n/2+1) x n x n layout), GTOMX_FFT_FULL_CROP = d_FFTFullCrop and
 * GTOMX_FFT_FULL_PAD = d_FFTFullPad on full arrays in FFT order. */
enum { GTOMX_FFT_CROP = 1, GTOMX_FFT_PAD = 2, GTOMX_FFT_FULL_CROP = 3, GTOMX_FFT_FULL_PAD = 4 };

/* One of the four resizes on `batch` real (complex_values = 0) or interleaved
 * complex arrays, index maps as the reference's kernels (FFTCropKernel :121-138,
 * FFTPadEvenKernel :162-205, FFTFullCropKernel :140-160, FFTFullPadEvenKernel
 * :212-238); crops need newdims <= olddims. In place (d_input == d_output,
 * buffer sized for the larger dims) is allowed and then synchronises the stream. */
int gtomx_fft_resize(int op, int complex_values, const void* d_input, void* d_output, gtomx_int3 olddims,
                     gtomx_int3 newdims, unsigned batch, void* stream);

/* d_FFTFullPad on real volumes, include/FFT.cuh:297 / src/FFT/FFTResize.cu:92-110
 * (kernel :212-238): pads `batch` volumes in FFT order (origin at index 0)
 * from olddims to newdims around the origin, zero elsewhere. In place
 * (d_input == d_output, buffer sized for newdims) is allowed and then
 * synchronises the stream. */
int gtomx_fft_full_pad(const float* d_input, float* d_output, gtomx_int3 olddims, gtomx_int3 newdims,
                       unsigned batch, void* stream);

/* d_Sum, include/Generics.cuh:585 / src/Generics/Sum.cu:99-139: the sum of each
 * of `batch` arrays of n elements into d_output[b]; float and double summed in
 * double, int in 64 bits and truncated to 32 (as the reference's int sum wraps).
 * n = 0 writes nothing. Synchronises the stream. */
int gtomx_sum_f32(const float* d_input, float* d_output, size_t n, unsigned batch, void* stream);
int gtomx_sum_f64(const double* d_input, double* d_output, size_t n, unsigned batch, void* stream);
int gtomx_sum_i32(const int32_t* d_input, int32_t* d_output, size_t n, unsigned batch, void* stream);

/* d_NormMonolithic, include/ImageManipulation.cuh:49 / src/ImageManipulation/Norm.cu:76-91
 * (kernel :242-284): each of `batch` arrays of `elements` floats becomes
 * (x - mean) / std with sums of x and x^2 in double, std = sqrt(N sum x^2 -
 * (sum x)^2) / N, and 0 where std is not positive; mean and 1/std are rounded to
 * float before use. (The reference's mode argument is not read by it.)
 * Synchronises the stream (its scratch is freed on return). */
int gtomx_norm_monolithic(const float* d_input, float* d_output, size_t elements, unsigned batch, void* stream);

/* No counterpart in the reference. gtomx_norm_monolithic for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops): the same kernels and result, with the partial sums and coefficients
 * in a file-scope device array of about 1 MiB per GPU (held until the process
 * ends) instead of per-call allocations, and no synchronisation of the stream.
 * batch may be at most GTOMX_NOALLOC_MAX_BATCH (else GTOMX_ERR_INVALID_ARGUMENT).
 * Every call on one GPU shares that array: calls that overlap on one GPU (on
 * different streams, or from several threads) race. */
#define GTOMX_NOALLOC_MAX_BATCH 256
int gtomx_norm_monolithic_noalloc(const float* d_input, float* d_output, size_t elements, unsigned batch,
                                  void* stream);

/* d_NormFTMonolithic, include/ImageManipulation.cuh:55 / src/ImageManipulation/Norm.cu:147-154
 * (kernel :476-509): each of `batch` complex arrays divided by
 * max(sqrt(sum |x|^2 / elements), 1e-20) (sum in double, the root rounded to
 * float). Synchronises the stream. */
int gtomx_norm_ft_monolithic(const gtomx_complex* d_input, gtomx_complex* d_output, size_t elements,
                             unsigned batch, void* stream);

/* d_Norm with T_NORM_MEAN01STD and no mask, include/ImageManipulation.cuh:48 /
 * src/ImageManipulation/Norm.cu:36-70 (statistics of d_Dev, src/Generics/Dev.cu:13-72):
 * each of `batch` arrays of `elements` floats becomes (x - mean)/std with the
 * population std (x - mean when std = 0), statistics accumulated in double.
 * Synchronises the stream (its scratch is freed on return). */
int gtomx_norm_mean01std(const float* d_input, float* d_output, size_t elements, unsigned batch,
                         void* stream);

typedef struct gtomx_picker gtomx_picker;

/* Picker::Initialize, include/Correlation.cuh:161 / src/Correlation/Picker.cu:30-116:
 * a cubic 3D reference d_ref (centred, real) for volumes of dimsimage, or a
 * square 2D reference (dimsref.z = 1) for images (dimsimage.z = 1; the
 * reference's ndims == 2 branch, whose orientations are in-plane steps). The
 * reference is transformed once (R2C of its ifftshift). d_refmask is accepted
 * and ignored: the reference prepares it but never uses it
 * (src/Correlation/Picker.cu:56-88, 193-221). batch = orientations per pass
 * (0 = default). The reference's constructor throws (src/Correlation/Picker.cu:27);
 * this object runs the pipeline its code describes. */
int gtomx_picker_create(gtomx_picker** picker, const float* d_ref, gtomx_int3 dimsref,
                        const float* d_refmask, gtomx_int3 dimsimage, unsigned batch, void* stream);
int gtomx_picker_destroy(gtomx_picker* picker);

/* Picker::SetImage, include/Correlation.cuh:162 / src/Correlation/Picker.cu:143-156:
 * normalises d_image to mean 0 / std 1 and transforms it; as in the reference
 * the caller's d_image is left holding the square of the normalised image.
 * d_ctf ((nref/2+1)*nref*nref floats) is kept by pointer and must stay valid. */
int gtomx_picker_set_image(gtomx_picker* picker, float* d_image, const float* d_ctf, void* stream);

/* Picker::PerformCorrelation, include/Correlation.cuh:163 /
 * src/Correlation/Picker.cu:158-247: for every orientation of
 * GetEqualAngularSpacing((0, 2pi), (0, pi), (0, 2pi), anglestep) (2D: the
 * ceil(2 pi / anglestep) in-plane steps (0, 0, k 2 pi / steps), :165-171), with the
 * rotation the reference left commented out done by the device-array
 * projector at the reference size:
 *   C = C2R(imageft * conj(R2C(norm(FullPad(C2R_n(project(refft) * ctf)))))),
 * and where C > d_bestccf: d_bestccf = C, d_bestangle = the (rot, tilt, psi)
 * triple (PickingUpdateKernel, src/Correlation/Picker.cu:249-276). The caller
 * initialises d_bestccf and d_bestangle, as in the reference. */
int gtomx_picker_correlate(gtomx_picker* picker, float anglestep, float* d_bestccf,
                           gtomx_float3* d_bestangle, void* stream);
/* The same over an explicit orientation list (no reference counterpart). */
int gtomx_picker_correlate_angles(gtomx_picker* picker, const gtomx_float3* h_angles, unsigned nangles,
                                  float* d_bestccf, gtomx_float3* d_bestangle, void* stream);

/* d_Exact2DWeighting, include/Projection.cuh / src/Projection/Weighting.cu:21-55
 * (kernel :92-133): multiplies batch half-spectrum filters (n/2+1 x n each, FFT
 * order; zero-centred weights are refused) by w_b^2 / sum_k |sinc(A . normal_k)| w_k
 * inside |f| <= maxfreq and 0 outside, A the frequency in the frame of image b of
 * the tilt series with (rot, tilt, psi) h_angles (d_imageweights NULL: all 1).
 * As in the reference, filter b takes image b's frame and weight (batch <=
 * nimages); h_indices is validated but, as there, selects nothing used. Synchronous. */
int gtomx_exact2d_weighting(float* d_weights, gtomx_int3 dimsimage, const int* h_indices,
                            const gtomx_float3* h_angles, const float* d_imageweights, unsigned nimages,
                            float maxfreq, int iszerocentered, unsigned batch, void* stream);

/* The TomoPicker class, include/Correlation.cuh:209-257 / src/Correlation/TomoPicker.cu.
 * The reference's constructor and Initialize throw and its tilt projections and
 * correlation call are marked broken (TomoPicker.cu:23, 28, 223-229, 264-269, 297);
 * this object runs the pipeline its code describes, with 2D central sections of
 * the padded (2 nr)^3 spectra from the device-array projector and, per tilt
 * image, the masked locally normalised correlation of d_CorrelateRealspace
 * (src/Correlation/Realspace.cu:16-102) at every image position.
 * create = Initialize (:26-104): reference and mask nr^3 (centred), circular mask
 * (projected once) or not (every orientation), images nx x ny, 1..128 of them. */
typedef struct gtomx_tomopicker gtomx_tomopicker;
int gtomx_tomopicker_create(gtomx_tomopicker** picker, const float* d_ref, gtomx_int3 dimsref,
                            const float* d_refmask, int ismaskcircular, gtomx_int3 dimsimage,
                            unsigned nimages, void* stream);
int gtomx_tomopicker_destroy(gtomx_tomopicker* picker);
/* SetImage (:158-186): d_image (nimages x ny x nx) is weighted by
 * d_Exact2DWeighting(h_imageangles, h_imageweights, maxfreq = nx) in place, as in
 * the reference; its spectrum and the spectrum of its square are kept. */
int gtomx_tomopicker_set_image(gtomx_tomopicker* picker, float* d_image, const gtomx_float3* h_imageangles,
                               const float* h_imageweights, void* stream);
/* PerformCorrelation (:188-305) over GetEqualAngularSpacing(anglestep):
 * d_corr (dimscorr, caller-initialised) keeps, with strict '<', the maximum over
 * orientations of each voxel's weighted mean of the tilt correlation maps at
 * its projections (UpdateTomoCorrKernel, :311-358), d_corrangle the triple. */
int gtomx_tomopicker_correlate(gtomx_tomopicker* picker, float* d_corr, gtomx_float3* d_corrangle,
                               gtomx_int3 dimscorr, float anglestep, void* stream);
/* The same over an explicit orientation list, with the optional orientation
 * index and the sums of the mean and of its square over orientations (the
 * reference's d_corrsum1/2, :202-203); any output but d_corr may be NULL. */
int gtomx_tomopicker_correlate_angles(gtomx_tomopicker* picker, const gtomx_float3* h_angles,
                                      unsigned nangles, float* d_corr, gtomx_float3* d_corrangle,
                                      gtomx_int3 dimscorr, int32_t* d_corrindex, float* d_corrsum1,
                                      float* d_corrsum2, void* stream);

typedef struct gtomx_prob_picker gtomx_prob_picker;

/* ProbabilisticPicker::Initialize, include/Correlation.cuh:202 /
 * src/Correlation/ProbabilisticPicker.cu:33-152: nrefs centred references
 * (d_refs) and masks (d_refmasks) of dimsref for images of dimsimage (2D when
 * dims.z == 1), band-limited to |k| < lowpassfreq and rotated at 2*lowpassfreq.
 * A non-circular mask (ismaskcircular == 0) is kept as cubic B-spline
 * coefficients and rotated with each orientation (d_Rotate3D / d_Rotate2D,
 * T_INTERP_CUBIC, exact weights), with its own solvent statistics. The low-pass spectra
 * are stored as the reference stores them, including what its half remap
 * (src/FFT/FFTRemap.cu, RemapHalfFFT2HalfKernel) reads past the cropped spectrum
 * for the zero-frequency row and plane; dimsimage must be large enough that
 * those reads stay inside the padded reference (GTOMX_ERR_UNSUPPORTED
 * otherwise). Everything stays in device memory. */
int gtomx_prob_picker_create(gtomx_prob_picker** picker, const float* d_refs, gtomx_int3 dimsref, unsigned nrefs,
                             const float* d_refmasks, int ismaskcircular, int doctf, gtomx_int3 dimsimage,
                             unsigned lowpassfreq, void* stream);
int gtomx_prob_picker_destroy(gtomx_prob_picker* picker);

/* ProbabilisticPicker::SetImage, include/Correlation.cuh:204 /
 * src/Correlation/ProbabilisticPicker.cu:205-224: the image normalised to mean 0
 * / std 1 and its square, transformed and band-limited; d_ctf (half spectrum
 * of dimsimage) cropped to the low-pass size. d_ctf may be NULL when the picker
 * was created without a CTF. The caller's image is left unchanged. */
int gtomx_prob_picker_set_image(gtomx_prob_picker* picker, const float* d_image, const float* d_ctf, void* stream);

/* ProbabilisticPicker::PerformCorrelation, include/Correlation.cuh:205 /
 * src/Correlation/ProbabilisticPicker.cu:226-389, for reference n: for every
 * orientation of GetEqualAngularSpacing((0, 2pi), (0, pi), (0, 2pi), anglestep)
 * (3D) or of ceil(2pi/anglestep) in-plane steps (2D), the probability-ratio
 * score of ProbPickingUpdateKernel (:391-421); where it exceeds d_bestccf,
 * d_bestccf, d_bestangle and d_bestref (= n) are replaced. Calls srand(123) and
 * draws the reference's gaussrand() noise from the C runtime's rand(), whose
 * state, like gaussrand's own, is shared with the rest of the process. The
 * caller initialises the three outputs (dimsimage elements each). */
int gtomx_prob_picker_correlate(gtomx_prob_picker* picker, unsigned n, float anglestep, float* d_bestccf,
                                gtomx_float3* d_bestangle, int* d_bestref, void* stream);
/* Test hook (no reference counterpart): outputs first .. first+n-1 of glibc's
 * rand() after srand(seed), from the library's jump-ahead restatement of its
 * additive generator, which PerformCorrelation uses once it has matched the
 * C runtime's rand(). */
void gtomx_debug_rand_outputs(unsigned seed, long long first, int n, int* out);
/* Number of orientations PerformCorrelation visits for ndims (2 or 3) and
 * anglestep (no reference counterpart; the list is the reference's). */
int gtomx_prob_picker_count_angles(int ndims, float anglestep, int* count);

/* Interpolation modes of the rotations, the values of T_INTERP_MODE
 * (include/Prerequisites.cuh:239-245). */
enum { GTOMX_INTERP_LINEAR = 1, GTOMX_INTERP_CUBIC = 2 };

/* d_Rotate3DFT, include/Transformation.cuh:54 / src/Transformation/Rotation.cu:263-374
 * (Rotate3DFTKernel :541-654): each of nangles rotations of the half spectrum of
 * a cube of side n. Input in the zero-centred half layout: element (tx, ty, tz)
 * of (n/2+1, n, n) holds the frequency (tx - n/2, ty - n/2, tz - n/2). Output
 * bin (idx, idy, idz), in the same layout when outputzerocentered, else in FFT
 * order read at the negated centred frequency as the reference does, is the
 * input interpolated at transpose(Euler(angle)) applied to its centred
 * frequency, 0 where that lies at radius >= n/2, taken from the Hermitian mate
 * (conjugate) where its x is positive. mode GTOMX_INTERP_LINEAR: trilinear;
 * GTOMX_INTERP_CUBIC: the cubic B-spline of the prefiltered input
 * (d_CubicBSplinePrefilter3D on a copy, cubicTex3D). The reference's texture
 * unit quantises the weights of its fetches to 1/256; here they are exact float,
 * and texel indices are clamped. Output: nangles * (n/2+1)*n*n values. Synchronous. */
int gtomx_rotate3d_ft(const gtomx_complex* d_volume, gtomx_complex* d_output, gtomx_int3 dims,
                      const gtomx_float3* h_angles, unsigned nangles, int mode, int outputzerocentered,
                      void* stream);
/* The same on real-valued half spectra (src/Transformation/Rotation.cu:325-374,
 * kernel :606-654): no conjugation. */
int gtomx_rotate3d_ft_real(const float* d_volume, float* d_output, gtomx_int3 dims, const gtomx_float3* h_angles,
                           unsigned nangles, int mode, int outputzerocentered, void* stream);

/* d_Rotate2DFT, include/Transformation.cuh:52 / src/Transformation/Rotation.cu:200-256
 * (kernel :485-539): `batch` zero-centred half spectra of dims.x x dims.y images
 * (dims.z == 1), item b rotated by h_angles[b] (radians): output bin (idx, idy)
 * takes the input at R(-angle) (idx - dims.x/2, idy - dims.y/2), 0 beyond
 * maxfreq, read at the Hermitian mate (imaginary part negated) when its x >
 * 1e-5; written in place or, when outputzerocentered is 0, at the FFT-order bin
 * (dims.x/2 - idx, dims.y-1 - ((idy + dims.y/2 - 1) % dims.y)). mode as for
 * gtomx_rotate3d_ft (bilinear, or the cubic B-spline of the prefiltered input);
 * exact weights, clamped indices. */
int gtomx_rotate2d_ft(const gtomx_complex* d_input, gtomx_complex* d_output, gtomx_int3 dims, const float* h_angles,
                      float maxfreq, int mode, int outputzerocentered, unsigned batch, void* stream);

/* ------------------------------------------------------------------------ */
/* Peak extraction (the consumer of the search output)                       */
/* ------------------------------------------------------------------------ */

/* d_LocalPeaks, include/Correlation.cuh:81 / src/Correlation/LocalPeaks.cu:24-62
 * (kernel :98-143): for each of `batch` volumes, the voxels whose value is not
 * below threshold and not smaller than any voxel within the Euclidean ball of
 * radius localextent (clamped at the volume faces, no wrap). Peaks of volume
 * b are written as (x, y, z) in raster order (z, then y, then x) to
 * h_peaks + b*capacity, at most `capacity` of them; h_counts[b] receives the
 * full count, so a caller whose capacity was too small can call again.
 * Synchronous: returns after the results are on the host. */
int gtomx_local_peaks(const float* d_input, gtomx_int3 dims, int localextent, float threshold,
                      unsigned batch, gtomx_int3* h_peaks, size_t capacity, size_t* h_counts,
                      void* stream);

/* T_PEAK_MODE, include/Correlation.cuh:44-52. */
enum { GTOMX_PEAK_INTEGER = 1, GTOMX_PEAK_SUBCOARSE = 2, GTOMX_PEAK_SUBFINE = 3 };

/* d_Peak, include/Correlation.cuh:65 / src/Correlation/Peak.cu:26-216: the
 * maximum of each of `batch` maps of dims (the first in raster order on ties)
 * as (x, y, z) and value, optionally refined: SUBCOARSE moves each coordinate
 * to the sinc-interpolated maximum (1/512 steps) on the line through the
 * maximum and evaluates the sinc interpolation there; SUBFINE up-scales the
 * 5^3 (3D) or 9^2 (2D) region around it in Fourier space by 63 (3D) or 105
 * (2D) and moves to that maximum. The cuFFT plan arguments of the reference
 * have no counterpart. dims with y = 1 and z > 1 are rejected. Synchronous. */
int gtomx_peak(const float* d_input, gtomx_float3* d_positions, float* d_values, gtomx_int3 dims, int mode,
               unsigned batch, void* stream);

/* d_PeakOne2D, include/Correlation.cuh:67 / src/Correlation/Peak.cu:218-225
 * (kernel :380-454): per image (dims.z = 1), the maximum of the centred
 * dimsregion plus the value-weighted centroid of its 3x3 neighbourhood, minus
 * (dims.x/2, dims.y/2) when subtractcenter is set; z of each position is 0.
 * Asynchronous on `stream`. */
int gtomx_peak_one_2d(const float* d_input, gtomx_float3* d_positions, float* d_values, gtomx_int3 dims,
                      gtomx_int3 dimsregion, int subtractcenter, unsigned batch, void* stream);

/* d_SubpixelMax, include/Correlation.cuh:83 / src/Correlation/LocalPeaks.cu:64-91
 * (kernel :145-178): for every shift s = (i + 1/2)/subpixsteps - 1/2 per
 * dimension, d_output = max(d_output, sum over the 13^ndims window of input *
 * sinc(s - offset)), at the voxels the reference's launch covers (x from 6 to
 * min(nx - 9, 128 ceil((nx - 15)/128) + 5), y from 6 to ny - 10, z from 6 to
 * nz - 10 in 3D); nothing else is written, so d_output must hold the starting
 * values. 2D and 3D maps, subpixsteps in [1, 64]. Synchronous. */
int gtomx_subpixel_max(const float* d_input, float* d_output, gtomx_int3 dims, int subpixsteps, void* stream);

/* ------------------------------------------------------------------------ */
/* MRC files (the on-disk format of tomograms, templates and output maps)    */
/* ------------------------------------------------------------------------ */

/* HeaderMRC, include/IO.cuh:26-96: the 1024-byte MRC header, field for field
 * (little-endian). cell* are the reference's `pixelsize`, mx..mz its
 * `griddimensions`, nsymbt its `extendedbytes`. */
typedef struct {
  int32_t nx, ny, nz;
  int32_t mode;                 /* 0 byte, 1 short, 2 float, 3 short complex, 4 float complex, 6 ushort, 16 RGB */
  int32_t startx, starty, startz;
  int32_t mx, my, mz;
  float cellx, celly, cellz;
  float alpha, beta, gamma;
  int32_t mapc, mapr, maps;
  float dmin, dmax, dmean;
  int32_t ispg;
  int32_t nsymbt;
  int16_t creatid;
  uint8_t extra1[30];
  int16_t nint, nreal;
  uint8_t extra2[28];
  int16_t idtype, lens, nd1, nd2, vd1, vd2;
  float tiltoriginal[3], tiltcurrent[3], origin[3];
  char cmap[4];
  uint8_t stamp[4];
  float rms;
  int32_t nlabl;
  char labels[10][80];
} gtomx_mrc_header;

/* HeaderMRC() defaults (include/IO.cuh:72-95), plus the "MAP " identifier and a
 * little-endian machine stamp, which the reference leaves unset. */
void gtomx_mrc_header_default(gtomx_mrc_header* header);
/* ReadMRCHeader, include/IO.cuh:166 / src/IO/mrc.cpp:38-66. */
int gtomx_mrc_read_header(const char* path, gtomx_mrc_header* header);
/* ReadMRC, include/IO.cuh:165 / src/IO/mrc.cpp:7-36: the raw data of z section nframe (nframe >= 0)
 * or of the whole volume (nframe < 0), in the file's mode, after the header
 * and the extended header. h_data must hold `capacity` >= the data bytes;
 * header (may be NULL) receives the header. */
int gtomx_mrc_read(const char* path, int nframe, void* h_data, size_t capacity, gtomx_mrc_header* header);
/* WriteMRC(void*, HeaderMRC, path), include/IO.cuh:168 / src/IO/mrc.cpp:68-83: the header and
 * nx*ny*nz voxels of its mode. No extended header is written, so nsymbt is
 * written as 0 (the reference writes the field as given). */
int gtomx_mrc_write(const void* h_data, const gtomx_mrc_header* header, const char* path);
/* WriteMRC(tfloat*, int3, path), include/IO.cuh:169 / src/IO/mrc.cpp:85-106: float volume with the
 * default header and its min/max. */
int gtomx_mrc_write_f32(const float* h_data, gtomx_int3 dims, const char* path);
/* d_WriteMRC(tfloat*, int3, path), include/IO.cuh:171 / src/IO/mrc.cpp:108-115. Synchronous. */
int gtomx_mrc_write_device_f32(const float* d_data, gtomx_int3 dims, const char* path, void* stream);
/* MixedToDeviceTfloat(void*, tfloat*, MRC_DATATYPE, size_t),
 * src/Helpers/Memory.cu:424-436: modes 0 (as unsigned bytes), 1, 2 and 6 to
 * float on the device; the raw bytes cross PCIe and are widened on the GPU.
 * Synchronous. */
int gtomx_mrc_to_device_f32(const void* h_data, int mode, size_t elements, float* d_output, void* stream);

/* ------------------------------------------------------------------------ */
/* The template-matching path                                                */
/* ------------------------------------------------------------------------ */

typedef struct gtomx_pick_plan gtomx_pick_plan;

/* Scratch and configuration for d_PickSubTomograms on cubes of side n =
 * dimsvolume.x, or on n x n images (dimsvolume.z = 1: the reference's 2D
 * branch, central sections of the projector), against a projector of side
 * p = dimsprojector.x. The plan owns every device buffer the search needs
 * (the reference allocates them per call, src/Correlation/SubTomograms.cu:44-56).
 * batch is the number of orientations processed per pass (0 selects the
 * default: 256 for volumes below 128, 128 up to 256, 32 or fewer above
 * (spectra of about 8 GiB), 240 for images as in the reference, fewer above
 * 256 (about 1 GiB)). Volume sides: 2^a and 3*2^a from 8 to 256 (fused
 * passes), 512, 768 and 1024 (long-row pass), and every other side from 8 to
 * 1024, odd ones included (the same form on chirp-z transforms). Image sides: 2^a and 3*2^a
 * from 8 to 256, and 512, 768, 1024, 1536, 2048, 3072 and 4096. */
int gtomx_pick_plan_create(gtomx_pick_plan** plan, gtomx_int3 dimsprojector,
                           gtomx_int3 dimsvolume, unsigned batch);
int gtomx_pick_plan_destroy(gtomx_pick_plan* plan);
/* Orientations per pass and device bytes owned by the plan. */
int gtomx_pick_plan_info(const gtomx_pick_plan* plan, unsigned* batch, size_t* device_bytes);

/* Row-pass kernel of the 2D search on images above 256 (no reference
 * counterpart; the same operations in another schedule, so the results agree
 * up to the rounding of contracted multiply-adds):
 * GTOMX_ROWS_PLAIN loads each orientation's spectrum rows when it starts,
 * GTOMX_ROWS_PREFETCH loads the next orientation's rows into registers while
 * the current one is transformed (more registers per thread, fewer resident
 * waves), GTOMX_ROWS_ONE (images of side 2048, 3072 or 4096) runs one row per
 * workgroup with every twiddle from LDS tables staged once per workgroup.
 * The micrograph autotuner picks one per image side. */
#define GTOMX_ROWS_PLAIN 0
#define GTOMX_ROWS_PREFETCH 1
#define GTOMX_ROWS_ONE 2
int gtomx_pick_plan_set_rows_variant(gtomx_pick_plan* plan, int variant);

/* Kernel timing: when enabled, every launch of the search kernels is bracketed
 * by HIP events on the plan's stream; gtomx_pick_plan_timings synchronises and
 * returns, per stage k in {0: project+z, 1: plane correction + y, 2: x+max},
 * the summed milliseconds and launch count since the last reset. */
int gtomx_pick_plan_enable_timing(gtomx_pick_plan* plan, int enable);
int gtomx_pick_plan_timings(gtomx_pick_plan* plan, double* ms3, unsigned long long* launches3,
                            int reset);

/* d_PickSubTomograms, include/Correlation.cuh:97-109 /
 * src/Correlation/SubTomograms.cu:19-125, with the device-array projector
 * (the texture pair is replaced by d_projectorft, a (p/2+1, p, p) half
 * spectrum in FFT order). For every volume v and orientation o:
 *   R = project(projector, transpose(Euler(h_angles[o])) * oversample) * ctf_v,
 *   R' = R normalised as C2R -> mean0/std1 -> R2C -> unit RMS,
 *   C = C2R(E_v * conj(R')), best = max over o (strict '>', first wins),
 * then best /= local std of C2R(E_v) under a soft sphere of maskradius.
 * d_ctf holds (n/2+1)*n*n floats; ctf_volume_stride is the element offset
 * between the CTFs of consecutive volumes (0: one CTF for all volumes, as in
 * the reference). d_experimentalft is not modified. d_bestindex is int32;
 * d_bestangle (optional, may be NULL) receives the same index as float, the
 * reference's output type. Asynchronous on `stream`. */
int gtomx_pick_subtomograms(gtomx_pick_plan* plan, const gtomx_complex* d_projectorft,
                            float projectoroversample, gtomx_int3 dimsprojector,
                            const gtomx_complex* d_experimentalft, const float* d_ctf,
                            size_t ctf_volume_stride, gtomx_int3 dimsvolume, unsigned nvolumes,
                            const gtomx_float3* h_angles, unsigned nangles, float maskradius,
                            float* d_bestcorrelation, int32_t* d_bestindex, float* d_bestangle,
                            void* stream);

/* Resumable pieces of the same search (no reference counterpart; they split
 * d_PickSubTomograms at its natural seams so a caller can checkpoint or
 * spread an orientation list over several calls):
 *  - begin: fills best with -1e30 and the index with 0;
 *  - update: searches h_angles[0..nangles) and records indices as
 *    angle_offset + o, keeping strictly better maxima only;
 *  - finish: divides best by the local standard deviation of C2R(E_v).
 * begin; update(list[0:k], 0); update(list[k:], k); finish equals one call of
 * gtomx_pick_subtomograms on the whole list. */
int gtomx_pick_begin(gtomx_pick_plan* plan, gtomx_int3 dimsvolume, unsigned nvolumes,
                     float* d_bestcorrelation, int32_t* d_bestindex, void* stream);
int gtomx_pick_update(gtomx_pick_plan* plan, const gtomx_complex* d_projectorft,
                      float projectoroversample, gtomx_int3 dimsprojector,
                      const gtomx_complex* d_experimentalft, const float* d_ctf,
                      size_t ctf_volume_stride, gtomx_int3 dimsvolume, unsigned nvolumes,
                      const gtomx_float3* h_angles, unsigned nangles, unsigned angle_offset,
                      float* d_bestcorrelation, int32_t* d_bestindex, void* stream);
int gtomx_pick_finish(gtomx_pick_plan* plan, const gtomx_complex* d_experimentalft,
                      gtomx_int3 dimsvolume, unsigned nvolumes, float maskradius,
                      float* d_bestcorrelation, int32_t* d_bestindex, float* d_bestangle,
                      void* stream);

/* Grouped update over 32^3 tiles (no reference counterpart; the learned search
 * below is built on it). h_angles holds ngroups * group orientations; group g
 * (orientations g*group .. g*group + group - 1, whose indices are their
 * positions in h_angles) is searched as one batch, and only on the tiles set in
 * its mask: h_tilemasks holds, per volume and group, (n/32)^2 bytes, byte
 * tz*(n/32) + ty with bit tx for the tile of voxels [32tx, 32tx+32) x
 * [32ty, ...) x [32tz, ...). Voxels of unset tiles keep their values; groups
 * with empty masks are skipped. Needs n^3 volumes with n a multiple of 32 (at
 * most 256) and group <= the plan's batch. Between begin and finish like
 * gtomx_pick_update; the values it writes are those of gtomx_pick_update on
 * the same orientations, restricted to the tiles. */
int gtomx_pick_update_tiles(gtomx_pick_plan* plan, const gtomx_complex* d_projectorft,
                            float projectoroversample, gtomx_int3 dimsprojector,
                            const gtomx_complex* d_experimentalft, const float* d_ctf,
                            size_t ctf_volume_stride, gtomx_int3 dimsvolume, unsigned nvolumes,
                            const gtomx_float3* h_angles, unsigned ngroups, unsigned group,
                            const uint8_t* h_tilemasks, float* d_bestcorrelation,
                            int32_t* d_bestindex, void* stream);

/* ------------------------------------------------------------------------ */
/* Micrographs of uneven sizes in one call                                   */
/* ------------------------------------------------------------------------ */

/* One micrograph of a batch: its R2C spectrum (n x (n/2+1), as
 * gtomx_fft_r2c writes it), its own CTF on the same half grid (e.g. from
 * gtomx_ctf_2d, astigmatism included), the local-std radius, and the outputs
 * (n x n best and int32 index, as gtomx_pick_subtomograms writes them). */
typedef struct {
  int n;
  const gtomx_complex* d_spectrum;
  const float* d_ctf;
  float maskradius;
  float* d_best;
  int32_t* d_index;
} gtomx_micrograph;

/* The 2D branch of d_PickSubTomograms (src/Correlation/SubTomograms.cu:34-36,
 * include/Correlation.cuh:97-109) over images of several sizes: one search plan
 * per side in h_sides (h_batches: orientations per pass for each side, 0 for the
 * plan default; h_variants: GTOMX_ROWS_* per side, NULL for GTOMX_ROWS_PLAIN;
 * normally both from the autotuner's table), each with device slots for one
 * image. The object owns a stream of its own (default priority). */
typedef struct gtomx_mpick gtomx_mpick;
int gtomx_mpick_create(gtomx_mpick** out, int projector_side, const int* h_sides, const unsigned* h_batches,
                       const int* h_variants, unsigned nsides);
int gtomx_mpick_destroy(gtomx_mpick* mp);
/* Batches and row variants per side (in h_sides order; either may be NULL),
 * graphs captured and graph replays so far. */
int gtomx_mpick_info(const gtomx_mpick* mp, unsigned* h_batches, int* h_variants, unsigned long long* captures,
                     unsigned long long* replays);
/* Searches every image over h_angles (oversample projector_side / n for an
 * image of side n; indices are positions in h_angles) with the projector
 * d_projectorft (projector_side^3 half spectrum, FFT order). Each image's result
 * equals gtomx_pick_subtomograms on that image alone. With use_graphs the
 * orientation loop of each side is captured once as a hipGraph (again when the
 * projector or the orientation list changes) and replayed per image. The work
 * runs on the object's stream, ordered after `stream`'s prior work, and
 * `stream` waits for it; the call returns without synchronising. */
int gtomx_mpick_run(gtomx_mpick* mp, const gtomx_complex* d_projectorft, const gtomx_micrograph* h_images,
                    unsigned nimages, const gtomx_float3* h_angles, unsigned nangles, int use_graphs,
                    void* stream);
/* Autotuner measurement for one image side: for each candidate (batch
 * h_batches[c], row variant h_variants[c] or GTOMX_ROWS_PLAIN when NULL), the
 * orientation loop over h_angles (one warm-up, then `repeats` timed runs with
 * events on `stream`), reported as milliseconds per orientation (infinity for a
 * batch whose plan does not fit). Synchronous. */
int gtomx_mpick_tune(int projector_side, int n, const gtomx_complex* d_projectorft, const unsigned* h_batches,
                     const int* h_variants, unsigned ncand, const gtomx_float3* h_angles, unsigned nangles,
                     int repeats, float* h_ms_per_orientation, void* stream);

/* ------------------------------------------------------------------------ */
/* Split-precision transforms on the matrix cores (no reference counterpart) */
/* ------------------------------------------------------------------------ */

/* Operand types of the radix-16 matrix-core stages: hi/lo split fp16 or bf16
 * (hi*hi + hi*lo + lo*hi, fp32 accumulation), or hi alone (no compensation). */
#define GTOMX_SPLIT_FP16 0
#define GTOMX_SPLIT_BF16 1
#define GTOMX_SINGLE_FP16 2
#define GTOMX_SINGLE_BF16 3

/* gtomx_pick_update (the search over h_angles between begin and finish) on 64^3
 * volumes with every per-orientation inverse transform (the C2R of
 * src/Correlation/SubTomograms.cu:81, which the reference runs through cuFFT,
 * src/FFT/IFFT.cu:48-56) done as 64 = 16 x 4 with the radix-16 stage on the
 * matrix cores in `precision`; projection, CTF, conj-multiply, unit-RMS scale
 * and max/argmax as in gtomx_pick_update, so only the transform's arithmetic
 * differs. Asynchronous on `stream`. */
int gtomx_pick_update_split(gtomx_pick_plan* plan, int precision, const gtomx_complex* d_projectorft,
                            float projectoroversample, gtomx_int3 dimsprojector,
                            const gtomx_complex* d_experimentalft, const float* d_ctf,
                            size_t ctf_volume_stride, gtomx_int3 dimsvolume, unsigned nvolumes,
                            const gtomx_float3* h_angles, unsigned nangles, unsigned angle_offset,
                            float* d_bestcorrelation, int32_t* d_bestindex, void* stream);

/* The transform of gtomx_pick_update_split alone: nlines contiguous lines of 64
 * complex values, out[k] = sum_n in[n] exp(+2 pi i k n / 64) (unnormalised
 * inverse); in may equal out. Asynchronous on `stream`. */
int gtomx_split_dft64(const gtomx_complex* d_in, gtomx_complex* d_out, unsigned nlines, int precision,
                      void* stream);

/* ------------------------------------------------------------------------ */
/* Learned orientation proposals (no reference counterpart)                  */
/* ------------------------------------------------------------------------ */

/* A 3D CNN over the 32^3 tiles of a volume (4 conv 3^3 layers of 16 channels
 * with ReLU and average pooling, a linear head) scoring 500 coarse directions
 * and one "no peak" class; bf16 weights and activations, fp32 accumulation,
 * forward and backward on the matrix cores. Parameters are
 * gtomx_pnet_param_count() floats (fp32 masters). Trained on the search's own
 * argmax labels (direction = orientation index / in-plane steps). */
typedef struct gtomx_pnet gtomx_pnet;
int gtomx_pnet_param_count(void);
int gtomx_pnet_create(gtomx_pnet** net, unsigned max_tiles);
int gtomx_pnet_destroy(gtomx_pnet* net);
/* Synchronous copies of all parameters; set also resets the optimiser state. */
int gtomx_pnet_set_params(gtomx_pnet* net, const float* h_params, size_t count, void* stream);
int gtomx_pnet_get_params(gtomx_pnet* net, float* h_params, size_t count, void* stream);
/* Network input: n standardised 32^3 float tiles, or every tile of an n^3
 * volume (n a multiple of 32) standardised by its mean and std. */
int gtomx_pnet_load_tiles(gtomx_pnet* net, const float* d_tiles, unsigned n, void* stream);
int gtomx_pnet_load_volume(gtomx_pnet* net, const float* d_volume, gtomx_int3 dims, void* stream);
/* Logits of the first n loaded tiles (d_logits optional: n x 512 floats, 501 used). */
int gtomx_pnet_forward(gtomx_pnet* net, unsigned n, float* d_logits, void* stream);
/* Softmax cross-entropy against d_labels (class per tile, 500 = no peak): the
 * gradients of every parameter (optional) and per-tile losses (optional). */
int gtomx_pnet_gradients(gtomx_pnet* net, const int32_t* d_labels, unsigned n, float* d_grads,
                         float* d_loss, void* stream);
/* One Adam step (0.9, 0.999, 1e-8) with L2 weight decay wd on the loaded tiles. */
int gtomx_pnet_train_step(gtomx_pnet* net, const int32_t* d_labels, unsigned n, float lr, float wd,
                          float* d_loss, void* stream);
/* From the last forward pass: per tile the top-k directions, the top-1
 * softmax probability and the top class, and (each optional) the summed
 * probability of the top-k directions and the probability of "no peak". */
int gtomx_pnet_propose(gtomx_pnet* net, unsigned topk, int32_t* d_dirs, float* d_top1,
                       int32_t* d_class, float* d_mass, float* d_nopeak, void* stream);
/* The head's GEMM alone (tests of the matrix-core operand layout). Synchronous. */
int gtomx_pnet_debug_gemm(const float* d_a, const float* d_b, float* d_c, int M, int N, int K,
                          void* stream);

typedef struct {
  unsigned tiles;           /* tiles over all volumes */
  unsigned exhaustive;      /* top-1 probability below the confidence: every direction */
  unsigned shortlisted;     /* a confident direction: its top-k directions */
  unsigned skipped;         /* confidently no peak: not searched */
  unsigned groups_searched; /* (volume, direction) groups run */
} gtomx_learned_stats;

/* The learned search (cfg. d_PickSubTomograms, include/Correlation.cuh:97-109,
 * whose argmax it is trained on): h_angles holds ndirections (500) x npsi
 * orientations, direction-major. The network scores every 32^3 tile of
 * C2R(E_v). A tile's proposal is "no peak" or its top-k directions, whichever
 * has the larger probability under the network (P(no peak) against the summed
 * probability of the k directions); a proposal below `confidence` falls back to
 * every orientation, a confident "no peak" tile is not searched, and any other
 * tile is searched over the npsi orientations of its top-k directions.
 * Searched voxels hold the result of gtomx_pick_subtomograms restricted to
 * their tile's orientations (same kernels, same values); voxels of tiles that
 * were not searched hold 0 and index -1. Synchronises `stream` once per
 * volume (the proposals choose the launches) and before returning. */
int gtomx_pick_subtomograms_learned(gtomx_pick_plan* plan, gtomx_pnet* net,
                                    const gtomx_complex* d_projectorft, float projectoroversample,
                                    gtomx_int3 dimsprojector, const gtomx_complex* d_experimentalft,
                                    const float* d_ctf, size_t ctf_volume_stride,
                                    gtomx_int3 dimsvolume, unsigned nvolumes,
                                    const gtomx_float3* h_angles, unsigned ndirections,
                                    unsigned npsi, unsigned topk, float confidence,
                                    float maskradius, float* d_bestcorrelation,
                                    int32_t* d_bestindex, float* d_bestangle,
                                    gtomx_learned_stats* h_stats, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GTOMX_H */
