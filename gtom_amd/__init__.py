"""Python host binding of the MI355X template-matching library (libgtomx.so).

The product is the C ABI declared in include/gtomx.h; this module binds it with
ctypes and exposes functions named after the reference's C++ API
(d_PickSubTomograms, d_rlnProject, d_FFTR2C, ...) that take torch tensors as
device memory. Every call goes to the HIP library; there is no fallback: if the
library is missing or the call fails, a RuntimeError is raised.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgtomx.so")
_lib = None


class GtomxError(RuntimeError):
    def __init__(self, status, where, message):
        super().__init__(f"{where}: status {status} ({message})")
        self.status = status


class MrcHeader(ctypes.Structure):
    """gtomx_mrc_header: HeaderMRC (include/IO.cuh:26-96), the 1024-byte MRC header."""
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("mode", ctypes.c_int32),
                ("startx", ctypes.c_int32), ("starty", ctypes.c_int32), ("startz", ctypes.c_int32),
                ("mx", ctypes.c_int32), ("my", ctypes.c_int32), ("mz", ctypes.c_int32),
                ("cellx", ctypes.c_float), ("celly", ctypes.c_float), ("cellz", ctypes.c_float),
                ("alpha", ctypes.c_float), ("beta", ctypes.c_float), ("gamma", ctypes.c_float),
                ("mapc", ctypes.c_int32), ("mapr", ctypes.c_int32), ("maps", ctypes.c_int32),
                ("dmin", ctypes.c_float), ("dmax", ctypes.c_float), ("dmean", ctypes.c_float),
                ("ispg", ctypes.c_int32), ("nsymbt", ctypes.c_int32),
                ("creatid", ctypes.c_int16), ("extra1", ctypes.c_uint8 * 30),
                ("nint", ctypes.c_int16), ("nreal", ctypes.c_int16), ("extra2", ctypes.c_uint8 * 28),
                ("idtype", ctypes.c_int16), ("lens", ctypes.c_int16), ("nd1", ctypes.c_int16),
                ("nd2", ctypes.c_int16), ("vd1", ctypes.c_int16), ("vd2", ctypes.c_int16),
                ("tiltoriginal", ctypes.c_float * 3), ("tiltcurrent", ctypes.c_float * 3),
                ("origin", ctypes.c_float * 3), ("cmap", ctypes.c_char * 4), ("stamp", ctypes.c_uint8 * 4),
                ("rms", ctypes.c_float), ("nlabl", ctypes.c_int32), ("labels", (ctypes.c_char * 80) * 10)]


class int3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int)]


class float3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class CtfParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("pixelsize", "Cs", "voltage", "defocus", "amplitude", "phaseshift", "Bfactor", "scale")]

    @classmethod
    def make(cls, pixelsize=1e-10, Cs=2e-3, voltage=300e3, defocus=-3e-6, amplitude=0.07,
             phaseshift=0.0, Bfactor=0.0, scale=1.0):
        return cls(pixelsize, Cs, voltage, defocus, amplitude, phaseshift, Bfactor, scale)


OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_UNSUPPORTED = 2
ERR_OUT_OF_MEMORY = 3
ERR_DEVICE = 4

_P = ctypes.c_void_p
_I = ctypes.c_int
_U = ctypes.c_uint
_F = ctypes.c_float
_SZ = ctypes.c_size_t

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGS = {
    "gtomx_version": [],
    "gtomx_status_string": [_I],
    "gtomx_last_error": [],
    "gtomx_debug_fail_alloc_after": [_I],
    "gtomx_equal_angular_spacing": [_F, _F, _F, _F, _F, _F, _F, _P, _I, ctypes.POINTER(_I)],
    "gtomx_fibonacci_angles": [_I, _I, _P],
    "gtomx_projector_matrices": [_P, _U, _F, _P],
    "gtomx_fft_r2c": [_P, _P, _I, int3, _U, _P],
    "gtomx_ifft_c2r": [_P, _P, _I, int3, _U, _I, _P],
    "gtomx_pad": [_P, _P, int3, int3, _F, _U, _P],
    "gtomx_remap_full2fullfft": [_P, _P, int3, _U, _P],
    "gtomx_project": [_P, int3, _P, int3, _U, _P, _U, _P],
    "gtomx_project_angles": [_P, int3, _P, int3, _P, _F, _U, _P],
    "gtomx_ctf_3d": [_P, _U, int3, _P, _P],
    "gtomx_complex_multiply_by_vector": [_P, _P, _P, _SZ, _U, _P],
    "gtomx_complex_multiply_by_conj_vector": [_P, _P, _P, _SZ, _U, _P],
    "gtomx_update_max_argmax": [_P, _SZ, _U, _I, _P, _P, _P],
    "gtomx_divide_safe_by_vector": [_P, _P, _P, _SZ, _U, _P],
    "gtomx_sphere_mask": [_P, _P, int3, _F, _F, _U, _P],
    "gtomx_local_std": [_P, int3, _P, _F, _P, _P, _P],
    "gtomx_local_peaks": [_P, int3, _I, _F, _U, _P, _SZ, _P, _P],
    "gtomx_fft_full_pad": [_P, _P, int3, int3, _U, _P],
    "gtomx_fft_resize": [_I, _I, _P, _P, int3, int3, _U, _P],
    "gtomx_sphere_mask_ex": [_P, _P, int3, _F, _F, _P, _I, _I, _U, _P],
    "gtomx_peak": [_P, _P, _P, int3, _I, _U, _P],
    "gtomx_peak_one_2d": [_P, _P, _P, int3, int3, _I, _U, _P],
    "gtomx_subpixel_max": [_P, _P, int3, _I, _P],
    "gtomx_ctf_simulate": [_P, _P, _P, _P, _U, _I, _I, _U, _P],
    "gtomx_ctf_2d": [_P, _U, int3, _P, _P],
    "gtomx_ctf_tilt_params_grid": [_P, int3, int3, _P, _U, _P],
    "gtomx_rotate3d_ft": [_P, _P, int3, _P, _U, _I, _I, _P],
    "gtomx_rotate3d_ft_real": [_P, _P, int3, _P, _U, _I, _I, _P],
    "gtomx_pick_subtomograms_diff2": [_P, _F, int3, _P, _P, int3, _U, int3, _P, _U, _P, _P, _P],
    "gtomx_mrc_header_default": [_P],
    "gtomx_mrc_read_header": [ctypes.c_char_p, _P],
    "gtomx_mrc_read": [ctypes.c_char_p, _I, _P, _SZ, _P],
    "gtomx_mrc_write": [_P, _P, ctypes.c_char_p],
    "gtomx_mrc_write_f32": [_P, int3, ctypes.c_char_p],
    "gtomx_mrc_write_device_f32": [_P, int3, ctypes.c_char_p, _P],
    "gtomx_mrc_to_device_f32": [_P, _I, _SZ, _P, _P],
    "gtomx_norm_mean01std": [_P, _P, _SZ, _U, _P],
    "gtomx_norm_monolithic": [_P, _P, _SZ, _U, _P],
    "gtomx_norm_monolithic_noalloc": [_P, _P, _SZ, _U, _P],
    "gtomx_sum_f32": [_P, _P, _SZ, _U, _P],
    "gtomx_sum_f64": [_P, _P, _SZ, _U, _P],
    "gtomx_sum_i32": [_P, _P, _SZ, _U, _P],
    "gtomx_norm_ft_monolithic": [_P, _P, _SZ, _U, _P],
    "gtomx_picker_create": [ctypes.POINTER(_P), _P, int3, _P, int3, _U, _P],
    "gtomx_picker_destroy": [_P],
    "gtomx_picker_set_image": [_P, _P, _P, _P],
    "gtomx_picker_correlate": [_P, _F, _P, _P, _P],
    "gtomx_picker_correlate_angles": [_P, _P, _U, _P, _P, _P],
    "gtomx_prob_picker_create": [ctypes.POINTER(_P), _P, int3, _U, _P, _I, _I, int3, _U, _P],
    "gtomx_prob_picker_destroy": [_P],
    "gtomx_prob_picker_set_image": [_P, _P, _P, _P],
    "gtomx_prob_picker_correlate": [_P, _U, _F, _P, _P, _P, _P],
    "gtomx_prob_picker_count_angles": [_I, _F, ctypes.POINTER(_I)],
    "gtomx_debug_rand_outputs": [_U, ctypes.c_longlong, _I, _P],
    "gtomx_rotate2d_ft": [_P, _P, int3, _P, _F, _I, _I, _U, _P],
    "gtomx_pick_plan_create": [ctypes.POINTER(_P), int3, int3, _U],
    "gtomx_pick_plan_destroy": [_P],
    "gtomx_pick_plan_set_rows_variant": [_P, ctypes.c_int],
    "gtomx_pick_plan_info": [_P, ctypes.POINTER(_U), ctypes.POINTER(_SZ)],
    "gtomx_pick_plan_enable_timing": [_P, _I],
    "gtomx_pick_plan_timings": [_P, _P, _P, _I],
    "gtomx_pick_subtomograms": [_P, _P, _F, int3, _P, _P, _SZ, int3, _U, _P, _U, _F, _P, _P, _P, _P],
    "gtomx_pick_begin": [_P, int3, _U, _P, _P, _P],
    "gtomx_pick_update": [_P, _P, _F, int3, _P, _P, _SZ, int3, _U, _P, _U, _U, _P, _P, _P],
    "gtomx_pick_finish": [_P, _P, int3, _U, _F, _P, _P, _P, _P],
    "gtomx_pick_update_tiles": [_P, _P, _F, int3, _P, _P, _SZ, int3, _U, _P, _U, _U, _P, _P, _P, _P],
    "gtomx_pick_update_split": [_P, ctypes.c_int, _P, _F, int3, _P, _P, _SZ, int3, _U, _P, _U, _U, _P, _P, _P],
    "gtomx_split_dft64": [_P, _P, _U, ctypes.c_int, _P],
    "gtomx_mpick_create": [_P, ctypes.c_int, _P, _P, _P, _U],
    "gtomx_exact2d_weighting": [_P, int3, _P, _P, _P, _U, _F, ctypes.c_int, _U, _P],
    "gtomx_tomopicker_create": [_P, _P, int3, _P, ctypes.c_int, int3, _U, _P],
    "gtomx_tomopicker_destroy": [_P],
    "gtomx_tomopicker_set_image": [_P, _P, _P, _P, _P],
    "gtomx_tomopicker_correlate": [_P, _P, _P, int3, _F, _P],
    "gtomx_tomopicker_correlate_angles": [_P, _P, _U, _P, _P, int3, _P, _P, _P, _P],
    "gtomx_mpick_destroy": [_P],
    "gtomx_mpick_info": [_P, _P, _P, _P, _P],
    "gtomx_mpick_run": [_P, _P, _P, _U, _P, _U, ctypes.c_int, _P],
    "gtomx_mpick_tune": [ctypes.c_int, ctypes.c_int, _P, _P, _P, _U, _P, _U, ctypes.c_int, _P, _P],
    "gtomx_pick_subtomograms_learned": [_P, _P, _P, _F, int3, _P, _P, _SZ, int3, _U, _P, _U, _U, _U, _F, _F, _P,
                                        _P, _P, _P, _P],
    "gtomx_pnet_param_count": [],
    "gtomx_pnet_create": [ctypes.POINTER(_P), _U],
    "gtomx_pnet_destroy": [_P],
    "gtomx_pnet_set_params": [_P, _P, _SZ, _P],
    "gtomx_pnet_get_params": [_P, _P, _SZ, _P],
    "gtomx_pnet_load_tiles": [_P, _P, _U, _P],
    "gtomx_pnet_load_volume": [_P, _P, int3, _P],
    "gtomx_pnet_forward": [_P, _U, _P, _P],
    "gtomx_pnet_gradients": [_P, _P, _U, _P, _P, _P],
    "gtomx_pnet_train_step": [_P, _P, _U, _F, _F, _P, _P],
    "gtomx_pnet_propose": [_P, _U, _P, _P, _P, _P, _P, _P],
    "gtomx_pnet_debug_gemm": [_P, _P, _P, _I, _I, _I, _P],
}
_RESTYPES = {"gtomx_version": ctypes.c_char_p, "gtomx_status_string": ctypes.c_char_p,
             "gtomx_last_error": ctypes.c_char_p, "gtomx_debug_fail_alloc_after": None}


def _load_torch_runtime_first():
    """PyTorch wheels carry their own HIP runtime under the same soname. Importing
    torch before libgtomx.so is loaded makes the library bind to that runtime, so
    the process holds one HIP/HSA runtime and torch's device pointers are valid in
    the library's calls. Loading libgtomx.so first would map a second runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Loads libgtomx.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gtom_amd.build`")
        _load_torch_runtime_first()
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGS)


def last_error():
    return lib().gtomx_last_error().decode()


def check(status, where):
    if status != OK:
        raise GtomxError(status, where, last_error())
    return status


def i3(d):
    if isinstance(d, int3):
        return d
    d = tuple(int(v) for v in d)
    return int3(*(d + (1,) * (3 - len(d))))


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    import torch
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch tensor")
    if not t.is_cuda:
        raise ValueError("tensor must be on the GPU")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _complex_view(t):
    import torch
    if t.dtype == torch.complex64:
        return t
    raise TypeError("complex tensors must be complex64 (interleaved float pairs)")


# ---------------------------------------------------------------------------
# host utilities
# ---------------------------------------------------------------------------
def equal_angular_spacing(phirange, thetarange, psirange, increment):
    """GetEqualAngularSpacing: (n, 3) float32 array of (rot, tilt, psi) in radians."""
    L = lib()
    cnt = _I(0)
    a = [np.float32(phirange[0]), np.float32(phirange[1]), np.float32(thetarange[0]),
         np.float32(thetarange[1]), np.float32(psirange[0]), np.float32(psirange[1]),
         np.float32(increment)]
    check(L.gtomx_equal_angular_spacing(*a, None, 0, ctypes.byref(cnt)), "equal_angular_spacing")
    out = np.empty((cnt.value, 3), np.float32)
    check(L.gtomx_equal_angular_spacing(*a, out.ctypes.data, cnt.value, ctypes.byref(cnt)),
          "equal_angular_spacing")
    return out


def fibonacci_angles(ndirs, npsi):
    out = np.empty((ndirs * npsi, 3), np.float32)
    check(lib().gtomx_fibonacci_angles(ndirs, npsi, out.ctypes.data), "fibonacci_angles")
    return out


def projector_matrices(angles, oversample):
    a = np.ascontiguousarray(angles, np.float32).reshape(-1, 3)
    out = np.empty((a.shape[0], 3, 3), np.float32)
    check(lib().gtomx_projector_matrices(a.ctypes.data, a.shape[0], oversample, out.ctypes.data),
          "projector_matrices")
    return out


# ---------------------------------------------------------------------------
# device operations (torch tensors as device memory)
# ---------------------------------------------------------------------------
def d_FFTR2C(x, ndims=3, stream=None):
    """Unnormalised R2C of real (..., nz, ny, nx) float32 -> complex64 (..., nz, ny, nx/2+1)."""
    import torch
    nz, ny, nx = x.shape[-3:] if ndims == 3 else ((1,) + tuple(x.shape[-2:]) if ndims == 2 else (1, 1, x.shape[-1]))
    batch = x.numel() // (nx * ny * nz)
    out = torch.empty(x.shape[:-1] + (nx // 2 + 1,), dtype=torch.complex64, device=x.device)
    check(lib().gtomx_fft_r2c(_ptr(x), _ptr(out), ndims, int3(nx, ny, nz), batch, _stream(stream)),
          "gtomx_fft_r2c")
    return out


def d_IFFTC2R(X, nx, ndims=3, renormalize=True, stream=None):
    """C2R of complex64 half spectra to width nx (renormalize: multiply by 1/N)."""
    import torch
    X = _complex_view(X)
    shp = X.shape[:-1] + (nx,)
    nz, ny = (shp[-3], shp[-2]) if ndims == 3 else ((1, shp[-2]) if ndims == 2 else (1, 1))
    batch = int(np.prod(shp)) // (nx * ny * nz)
    out = torch.empty(shp, dtype=torch.float32, device=X.device)
    check(lib().gtomx_ifft_c2r(_ptr(X), _ptr(out), ndims, int3(nx, ny, nz), batch, int(bool(renormalize)),
                               _stream(stream)), "gtomx_ifft_c2r")
    return out


def d_Pad(x, outdims, value=0.0, stream=None):
    import torch
    ind = tuple(x.shape[-3:])
    batch = x.numel() // int(np.prod(ind))
    out = torch.empty(x.shape[:-3] + tuple(outdims), dtype=torch.float32, device=x.device)
    check(lib().gtomx_pad(_ptr(x), _ptr(out), int3(ind[2], ind[1], ind[0]),
                          int3(outdims[2], outdims[1], outdims[0]), value, batch, _stream(stream)), "gtomx_pad")
    return out


def d_RemapFull2FullFFT(x, stream=None):
    import torch
    d = tuple(x.shape[-3:])
    batch = x.numel() // int(np.prod(d))
    out = torch.empty_like(x)
    check(lib().gtomx_remap_full2fullfft(_ptr(x), _ptr(out), int3(d[2], d[1], d[0]), batch, _stream(stream)),
          "gtomx_remap_full2fullfft")
    return out


def prepare_projector(template, p, stream=None):
    """Projector half spectrum R2C(ifftshift(pad_centre(template, p))) (complex64, (p, p, p/2+1))."""
    padded = d_Pad(template, (p, p, p), 0.0, stream)
    return d_FFTR2C(d_RemapFull2FullFFT(padded, stream), 3, stream)


def d_rlnProject(volume_ft, p, n, matrices=None, angles=None, oversample=1.0, rmax=None, stream=None, ndims=3):
    """Project a (p, p, p/2+1) half spectrum to `batch` (n, n, n/2+1) half spectra, or with
    ndims=2 to (n, n/2+1) central sections."""
    import torch
    if matrices is None:
        matrices = torch.from_numpy(projector_matrices(angles, oversample)).to(volume_ft.device)
    matrices = matrices.reshape(-1, 9).contiguous().float()
    batch = matrices.shape[0]
    nz = n if ndims == 3 else 1
    shape = (batch, n, n, n // 2 + 1) if ndims == 3 else (batch, n, n // 2 + 1)
    out = torch.empty(shape, dtype=torch.complex64, device=volume_ft.device)
    check(lib().gtomx_project(_ptr(volume_ft), int3(p, p, p), _ptr(out), int3(n, n, nz),
                              n // 2 if rmax is None else rmax, _ptr(matrices), batch, _stream(stream)),
          "gtomx_project")
    return out


INTERP_LINEAR, INTERP_CUBIC = 1, 2


def d_Rotate3DFT(volume, angles, outputzerocentered=True, stream=None, mode=INTERP_LINEAR):
    """d_Rotate3DFT (src/Transformation/Rotation.cu:263-374): a (n, n, n/2+1)
    zero-centred half spectrum (complex64 or float32 CUDA tensor) rotated by each
    (rot, tilt, psi) of `angles`, linear or cubic; returns (nangles, n, n, n/2+1)
    of the same dtype."""
    import numpy as np
    import torch
    a = np.ascontiguousarray(np.asarray(angles, np.float32).reshape(-1, 3))
    n = volume.shape[0]
    out = torch.empty((len(a),) + tuple(volume.shape), dtype=volume.dtype, device=volume.device)
    fn = lib().gtomx_rotate3d_ft if volume.is_complex() else lib().gtomx_rotate3d_ft_real
    check(fn(_ptr(volume), _ptr(out), int3(n, n, n), ctypes.c_void_p(a.ctypes.data), len(a), int(mode),
             int(bool(outputzerocentered)), _stream(stream)), "gtomx_rotate3d_ft")
    return out


def d_Rotate2DFT(image, angles, maxfreq, outputzerocentered=True, stream=None, mode=INTERP_LINEAR):
    """d_Rotate2DFT, linear (src/Transformation/Rotation.cu:200-256): (batch, n, n/2+1)
    zero-centred complex64 half spectra, item b rotated by angles[b] (radians);
    returns a new tensor of the same shape."""
    import torch
    a = np.ascontiguousarray(np.asarray(angles, np.float32).reshape(-1))
    x = image if image.dim() == 3 else image.unsqueeze(0)
    if x.shape[0] != len(a):
        raise ValueError("one angle per image")
    ny, nh = x.shape[-2:]
    out = torch.empty_like(x)
    check(lib().gtomx_rotate2d_ft(_ptr(x), _ptr(out), int3(2 * (nh - 1), ny, 1), ctypes.c_void_p(a.ctypes.data),
                                  float(maxfreq), int(mode), int(bool(outputzerocentered)), len(a), _stream(stream)),
          "gtomx_rotate2d_ft")
    return out


def ctf_3d(params, n, device="cuda", stream=None, ndims=3):
    """Radially symmetric CTFs on (n, n, n/2+1) half spectra, or (n, n/2+1) with ndims=2."""
    import torch
    params = list(params) if isinstance(params, (list, tuple)) else [params]
    arr = (CtfParams * len(params))(*params)
    nz = n if ndims == 3 else 1
    shape = (len(params), n, n, n // 2 + 1) if ndims == 3 else (len(params), n, n // 2 + 1)
    out = torch.empty(shape, dtype=torch.float32, device=device)
    check(lib().gtomx_ctf_3d(arr, len(params), int3(n, n, nz), _ptr(out), _stream(stream)), "gtomx_ctf_3d")
    return out


class CtfParamsFull(ctypes.Structure):
    """gtomx_ctf_params_full: all of CTFParams (include/CTF.cuh:11-43), SI units."""
    _fields_ = [(f, ctypes.c_float) for f in ("pixelsize", "pixeldelta", "pixelangle", "Cs", "voltage", "defocus",
                                              "astigmatismangle", "defocusdelta", "amplitude", "Bfactor",
                                              "Bfactordelta", "Bfactorangle", "scale", "phaseshift")]

    @classmethod
    def make(cls, **kw):
        d = dict(pixelsize=1e-10, pixeldelta=0.0, pixelangle=0.0, Cs=2e-3, voltage=300e3, defocus=-3e-6,
                 astigmatismangle=0.0, defocusdelta=0.0, amplitude=0.07, Bfactor=0.0, Bfactordelta=0.0,
                 Bfactorangle=0.0, scale=1.0, phaseshift=0.0)
        d.update(kw)
        return cls(**d)


class CtfTiltParams(ctypes.Structure):
    """gtomx_ctf_tilt_params: CTFTiltParams (include/CTF.cuh:136-158), radians."""
    _fields_ = [("imageangle", ctypes.c_float), ("stageangle", ctypes.c_float * 2),
                ("specimenangle", ctypes.c_float * 2), ("centerparams", CtfParamsFull)]

    @classmethod
    def make(cls, centerparams, imageangle=0.0, stageangle=(0.0, 0.0), specimenangle=(0.0, 0.0)):
        return cls(float(imageangle), (ctypes.c_float * 2)(*stageangle), (ctypes.c_float * 2)(*specimenangle),
                   centerparams)


def ctf_tilt_params_grid(tilt, dimsimage, dimsregion, positions):
    """CTFTiltParams::GetParamsGrid2D: a CtfParamsFull per region origin (x, y)."""
    n = len(positions)
    pos = (int3 * max(1, n))(*[int3(int(p[0]), int(p[1]), 0) for p in positions])
    out = (CtfParamsFull * max(1, n))()
    check(lib().gtomx_ctf_tilt_params_grid(ctypes.byref(tilt), int3(int(dimsimage[0]), int(dimsimage[1]), 1),
                                           int3(int(dimsregion[0]), int(dimsregion[1]), 1), pos, n, out),
          "gtomx_ctf_tilt_params_grid")
    return [out[i] for i in range(n)]


def d_CTFSimulate(params, addresses, gammacorrection=None, amplitudesquared=False, ignorefirstpeak=False,
                  stream=None):
    """d_CTFSimulate (src/CTF/Simulate.cu:22-47): addresses is an (n, 2) float32 CUDA tensor of
    (k [cycles/pixel], angle); returns (batch, n)."""
    import torch
    params = list(params) if isinstance(params, (list, tuple)) else [params]
    arr = (CtfParamsFull * len(params))(*params)
    n = addresses.shape[0]
    out = torch.empty((len(params), n), dtype=torch.float32, device=addresses.device)
    check(lib().gtomx_ctf_simulate(arr, _ptr(addresses), _ptr(gammacorrection), _ptr(out), n,
                                   int(amplitudesquared), int(ignorefirstpeak), len(params), _stream(stream)),
          "gtomx_ctf_simulate")
    return out


def ctf_2d(params, n, device="cuda", stream=None):
    """Astigmatic CTFs on (n, n/2+1) half spectra, one per parameter set."""
    import torch
    params = list(params) if isinstance(params, (list, tuple)) else [params]
    arr = (CtfParamsFull * len(params))(*params)
    out = torch.empty((len(params), n, n // 2 + 1), dtype=torch.float32, device=device)
    check(lib().gtomx_ctf_2d(arr, len(params), int3(n, n, 1), _ptr(out), _stream(stream)), "gtomx_ctf_2d")
    return out


def d_UpdateMaxArgmax(corr, best, bestidx, batchoffset, stream=None):
    batch = corr.shape[0]
    L = best.numel()
    check(lib().gtomx_update_max_argmax(_ptr(corr), L, batch, batchoffset, _ptr(best), _ptr(bestidx),
                                        _stream(stream)), "gtomx_update_max_argmax")


def d_LocalPeaks(vol, localextent, threshold, stream=None):
    """d_LocalPeaks (src/Correlation/LocalPeaks.cu:24-62) on a float32 CUDA
    tensor of shape (..., z, y, x): one (k, 3) int64 numpy array of (x, y, z)
    per volume, in raster order."""
    import numpy as np
    nz, ny, nx = vol.shape[-3:]
    batch = max(1, vol.numel() // (nx * ny * nz))
    counts = np.zeros(batch, np.uint64)
    cap = 4096
    while True:
        peaks = np.zeros((batch, cap, 3), np.int32)
        check(lib().gtomx_local_peaks(_ptr(vol), int3(nx, ny, nz), int(localextent), float(threshold), batch,
                                      ctypes.c_void_p(peaks.ctypes.data), cap,
                                      ctypes.c_void_p(counts.ctypes.data), _stream(stream)),
              "gtomx_local_peaks")
        need = int(counts.max())
        if need <= cap:
            break
        cap = need
    return [peaks[b, :int(counts[b])].astype(np.int64) for b in range(batch)]


PEAK_INTEGER, PEAK_SUBCOARSE, PEAK_SUBFINE = 1, 2, 3
_PEAK_MODES = {"integer": 1, "subcoarse": 2, "subfine": 3}


def d_Peak(maps, mode, ndims=3, stream=None):
    """d_Peak (src/Correlation/Peak.cu:26-216) on a float32 CUDA tensor holding
    maps of shape (..., z, y, x) (ndims=3) or (..., y, x) (ndims=2). mode is
    "integer", "subcoarse", "subfine" or 1/2/3. Returns (positions (batch, 3)
    as (x, y, z), values (batch,)) as float32 CUDA tensors."""
    import torch
    shape = maps.shape[-ndims:]
    nx, ny = shape[-1], (shape[-2] if ndims >= 2 else 1)
    nz = shape[-3] if ndims == 3 else 1
    batch = max(1, maps.numel() // (nx * ny * nz))
    pos = torch.empty((batch, 3), dtype=torch.float32, device=maps.device)
    val = torch.empty(batch, dtype=torch.float32, device=maps.device)
    check(lib().gtomx_peak(_ptr(maps), _ptr(pos), _ptr(val), int3(nx, ny, nz), int(_PEAK_MODES.get(mode, mode)),
                           batch, _stream(stream)), "gtomx_peak")
    return pos, val


def d_PeakOne2D(images, region, subtractcenter, stream=None):
    """d_PeakOne2D (src/Correlation/Peak.cu:218-225) on a float32 CUDA tensor of
    shape (..., y, x); region = (rx, ry). Returns (positions (batch, 3), values)."""
    import torch
    ny, nx = images.shape[-2:]
    batch = max(1, images.numel() // (nx * ny))
    pos = torch.empty((batch, 3), dtype=torch.float32, device=images.device)
    val = torch.empty(batch, dtype=torch.float32, device=images.device)
    check(lib().gtomx_peak_one_2d(_ptr(images), _ptr(pos), _ptr(val), int3(nx, ny, 1),
                                  int3(int(region[0]), int(region[1]), 1), int(bool(subtractcenter)), batch,
                                  _stream(stream)), "gtomx_peak_one_2d")
    return pos, val


def d_SubpixelMax(vol, out, subpixsteps, stream=None):
    """d_SubpixelMax (src/Correlation/LocalPeaks.cu:64-91): `out` (float32 CUDA
    tensor of the shape of `vol`, (z, y, x) or (y, x)) is raised in place to the
    running maximum over the sub-sample shifts at the voxels the reference
    covers."""
    nx, ny = vol.shape[-1], vol.shape[-2]
    nz = vol.shape[-3] if vol.dim() == 3 else 1
    check(lib().gtomx_subpixel_max(_ptr(vol), _ptr(out), int3(nx, ny, nz), int(subpixsteps), _stream(stream)),
          "gtomx_subpixel_max")
    return out


def d_Sum(x, batch=1, stream=None):
    """d_Sum (src/Generics/Sum.cu:99-139) of `batch` equal parts of a float32, float64
    or int32 CUDA tensor; returns a tensor of `batch` sums of the same dtype."""
    import torch
    fn = {torch.float32: lib().gtomx_sum_f32, torch.float64: lib().gtomx_sum_f64,
          torch.int32: lib().gtomx_sum_i32}.get(x.dtype)
    if fn is None:
        raise TypeError("d_Sum takes float32, float64 or int32")
    out = torch.zeros(batch, dtype=x.dtype, device=x.device)
    check(fn(_ptr(x), _ptr(out), x.numel() // batch, batch, _stream(stream)), "gtomx_sum")
    return out


def d_NormMonolithic(x, batch=1, stream=None):
    """d_NormMonolithic (src/ImageManipulation/Norm.cu:76-91) of `batch` equal parts of a
    float32 CUDA tensor; returns a new tensor."""
    import torch
    out = torch.empty_like(x)
    check(lib().gtomx_norm_monolithic(_ptr(x), _ptr(out), x.numel() // batch, batch, _stream(stream)),
          "gtomx_norm_monolithic")
    return out


NOALLOC_MAX_BATCH = 256  # GTOMX_NOALLOC_MAX_BATCH


def d_NormMonolithic_noalloc(x, out, batch=1, stream=None):
    """d_NormMonolithic through the allocation-free entry point: writes `out` (a float32 CUDA
    tensor shaped like x) without allocating device memory or synchronising the stream."""
    check(lib().gtomx_norm_monolithic_noalloc(_ptr(x), _ptr(out), x.numel() // batch, batch, _stream(stream)),
          "gtomx_norm_monolithic_noalloc")
    return out


def d_NormFTMonolithic(x, batch=1, stream=None):
    """d_NormFTMonolithic (src/ImageManipulation/Norm.cu:147-154) of `batch` equal parts of a
    complex64 CUDA tensor; returns a new tensor."""
    import torch
    out = torch.empty_like(x)
    check(lib().gtomx_norm_ft_monolithic(_ptr(x), _ptr(out), x.numel() // batch, batch, _stream(stream)),
          "gtomx_norm_ft_monolithic")
    return out


def d_LocalStd(vol, radius, with_mean=False, fouriermask=None, stream=None):
    """d_LocalStd (src/Generics/LocalStd.cu:10-118) of a float32 CUDA volume (nz, ny, nx):
    sigma, and the mean if asked, under a soft sphere of `radius`; `fouriermask`
    is None or a float32 CUDA half spectrum (nz, ny, nx//2+1) in FFT order."""
    import torch
    nz, ny, nx = vol.shape[-3:]
    if fouriermask is not None and tuple(fouriermask.shape[-3:]) != (nz, ny, nx // 2 + 1):
        raise ValueError("fouriermask must be the half spectrum (nz, ny, nx//2+1)")
    sd = torch.empty_like(vol)
    mean = torch.empty_like(vol) if with_mean else None
    check(lib().gtomx_local_std(_ptr(vol), int3(nx, ny, nz), _ptr(fouriermask), radius, _ptr(sd), _ptr(mean),
                                _stream(stream)),
          "gtomx_local_std")
    return (sd, mean) if with_mean else sd


ROWS_PLAIN = 0  # GTOMX_ROWS_PLAIN
ROWS_PREFETCH = 1  # GTOMX_ROWS_PREFETCH
ROWS_ONE = 2  # GTOMX_ROWS_ONE (images of side 2048, 3072 or 4096)


class PickPlan:
    """Owns a gtomx_pick_plan (scratch for d_PickSubTomograms on n^3 volumes)."""

    def __init__(self, p, n, batch=0, ndims=3):
        """ndims 3: n^3 volumes; ndims 2: n x n images (the reference's 2D branch)."""
        self.p, self.n, self.ndims = p, n, ndims
        self.nz = n if ndims == 3 else 1
        h = _P()
        check(lib().gtomx_pick_plan_create(ctypes.byref(h), int3(p, p, p), int3(n, n, self.nz), batch),
              "gtomx_pick_plan_create")
        self.handle = h
        b, nbytes = _U(0), _SZ(0)
        check(lib().gtomx_pick_plan_info(h, ctypes.byref(b), ctypes.byref(nbytes)), "gtomx_pick_plan_info")
        self.batch, self.device_bytes = b.value, nbytes.value

    def close(self):
        if self.handle:
            lib().gtomx_pick_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def enable_timing(self, on=True):
        check(lib().gtomx_pick_plan_enable_timing(self.handle, int(bool(on))), "enable_timing")

    def set_rows_variant(self, variant):
        """ROWS_PLAIN, ROWS_PREFETCH or ROWS_ONE (sides 2048, 3072, 4096): the row-pass kernel of the 2D search on images
        above 256."""
        check(lib().gtomx_pick_plan_set_rows_variant(self.handle, int(variant)), "gtomx_pick_plan_set_rows_variant")

    def timings(self, reset=True):
        ms = (ctypes.c_double * 3)()
        cnt = (ctypes.c_ulonglong * 3)()
        check(lib().gtomx_pick_plan_timings(self.handle, ms, cnt, int(bool(reset))), "timings")
        return list(ms), list(cnt)


_MRC_DTYPES = {0: np.uint8, 1: np.int16, 2: np.float32, 3: np.dtype([("re", "<i2"), ("im", "<i2")]),
               4: np.complex64, 6: np.uint16}


def read_mrc_header(path):
    """ReadMRCHeader (src/IO/mrc.cpp:38-66): an MrcHeader."""
    h = MrcHeader()
    check(lib().gtomx_mrc_read_header(os.fsencode(path), ctypes.byref(h)), "gtomx_mrc_read_header")
    return h


def read_mrc(path, nframe=-1):
    """ReadMRC (src/IO/mrc.cpp:7-36): (header, numpy array (z, y, x) in the file's mode);
    one z section when nframe >= 0."""
    h = read_mrc_header(path)
    if h.mode not in _MRC_DTYPES:
        raise GtomxError(2, "read_mrc", f"MRC mode {h.mode} is not read into numpy here")
    dt = np.dtype(_MRC_DTYPES[h.mode])
    shape = (1 if nframe >= 0 else h.nz, h.ny, h.nx)
    out = np.empty(shape, dt)
    check(lib().gtomx_mrc_read(os.fsencode(path), int(nframe), ctypes.c_void_p(out.ctypes.data), out.nbytes,
                               ctypes.byref(h)), "gtomx_mrc_read")
    return h, out


def write_mrc(path, data, header=None):
    """WriteMRC: float32 data with the default header and its min/max
    (src/IO/mrc.cpp:85-106), or any supported mode with the given header
    (src/IO/mrc.cpp:68-83)."""
    a = np.ascontiguousarray(data)
    if header is None:
        a = np.ascontiguousarray(a, np.float32)
        nz, ny, nx = a.shape[-3:] if a.ndim >= 3 else (1,) * (3 - a.ndim) + a.shape
        check(lib().gtomx_mrc_write_f32(ctypes.c_void_p(a.ctypes.data), int3(nx, ny, nz), os.fsencode(path)),
              "gtomx_mrc_write_f32")
    else:
        check(lib().gtomx_mrc_write(ctypes.c_void_p(a.ctypes.data), ctypes.byref(header), os.fsencode(path)),
              "gtomx_mrc_write")


def mrc_header_default():
    h = MrcHeader()
    lib().gtomx_mrc_header_default(ctypes.byref(h))
    return h


def d_WriteMRC(x, path, stream=None):
    """d_WriteMRC (src/IO/mrc.cpp:108-115) of a float32 CUDA tensor (z, y, x)."""
    nz, ny, nx = x.shape[-3:]
    check(lib().gtomx_mrc_write_device_f32(_ptr(x), int3(nx, ny, nz), os.fsencode(path), _stream(stream)),
          "gtomx_mrc_write_device_f32")


def mrc_to_device(data, mode, device="cuda", stream=None):
    """MixedToDeviceTfloat (src/Helpers/Memory.cu:424-436): MRC mode 0/1/2/6 data to a
    float32 CUDA tensor, widened on the GPU."""
    import torch
    a = np.ascontiguousarray(data)
    out = torch.empty(a.shape, dtype=torch.float32, device=device)
    check(lib().gtomx_mrc_to_device_f32(ctypes.c_void_p(a.ctypes.data), int(mode), a.size, _ptr(out),
                                        _stream(stream)), "gtomx_mrc_to_device_f32")
    return out


def d_FFTFullPad(x, newshape, stream=None):
    """d_FFTFullPad (src/FFT/FFTResize.cu:92-110) of real volumes (..., z, y, x) in FFT
    order to newshape (z, y, x)."""
    import torch
    oz, oy, ox = x.shape[-3:]
    nz, ny, nx = newshape
    batch = max(1, x.numel() // (ox * oy * oz))
    out = torch.empty(tuple(x.shape[:-3]) + (nz, ny, nx), dtype=torch.float32, device=x.device)
    check(lib().gtomx_fft_full_pad(_ptr(x), _ptr(out), int3(ox, oy, oz), int3(nx, ny, nz), batch, _stream(stream)),
          "gtomx_fft_full_pad")
    return out


_FFT_RESIZE = {"crop": 1, "pad": 2, "full_crop": 3, "full_pad": 4}


def d_FFTResize(op, x, olddims, newdims, stream=None):
    """d_FFTCrop ("crop"), d_FFTPad ("pad") on half spectra (..., z, y, x/2+1), or
    d_FFTFullCrop ("full_crop"), d_FFTFullPad ("full_pad") on full arrays (..., z, y, x),
    float32 or complex64 CUDA tensors; dims are (x, y, z) (src/FFT/FFTResize.cu)."""
    import torch
    k = _FFT_RESIZE[op]
    half = k <= 2
    ox, oy, oz = olddims
    nx, ny, nz = newdims
    per = (ox // 2 + 1 if half else ox) * oy * oz
    batch = max(1, x.numel() // per)
    shape = (nz, ny, nx // 2 + 1 if half else nx)
    out = torch.empty((batch,) + shape, dtype=x.dtype, device=x.device)
    check(lib().gtomx_fft_resize(k, int(x.is_complex()), _ptr(x), _ptr(out), int3(ox, oy, oz), int3(nx, ny, nz),
                                 batch, _stream(stream)), "gtomx_fft_resize")
    return out


def d_SphereMask(x, radius, sigma, center=None, decentered=False, ndims=3, stream=None):
    """d_SphereMask (src/Masking/SphereMask.cu:18-106) of float32 or complex64 CUDA
    tensors (..., z, y, x) (ndims=3) or (..., y, x); center (x, y, z) or None; returns a
    new tensor."""
    import torch
    shape = x.shape[-ndims:]
    nx, ny = shape[-1], shape[-2]
    nz = shape[-3] if ndims == 3 else 1
    batch = max(1, x.numel() // (nx * ny * nz))
    out = torch.empty_like(x)
    c = float3(*[float(v) for v in center]) if center is not None else None
    check(lib().gtomx_sphere_mask_ex(_ptr(x), _ptr(out), int3(nx, ny, nz), float(radius), float(sigma),
                                     ctypes.byref(c) if c is not None else None, int(bool(decentered)),
                                     int(x.is_complex()), batch, _stream(stream)), "gtomx_sphere_mask_ex")
    return out


def d_NormMean01Std(x, batch=1, out=None, stream=None):
    """d_Norm with T_NORM_MEAN01STD (src/ImageManipulation/Norm.cu:36-70) over `batch`
    equal parts of x."""
    import torch
    out = torch.empty_like(x) if out is None else out
    check(lib().gtomx_norm_mean01std(_ptr(x), _ptr(out), x.numel() // batch, batch, _stream(stream)),
          "gtomx_norm_mean01std")
    return out


class Picker:
    """The Picker class (include/Correlation.cuh:136-164): initialize with a cubic
    reference (z, y, x) and a volume shape, or a square reference (y, x) and an
    image shape (the 2D branch: in-plane orientations); set_image; perform_correlation."""

    def __init__(self, ref, imageshape, refmask=None, batch=0, stream=None):
        nr = ref.shape[-1]
        nrz = ref.shape[0] if ref.dim() == 3 else 1
        nz, ny, nx = tuple(imageshape) if len(imageshape) == 3 else (1,) + tuple(imageshape)
        h = _P()
        check(lib().gtomx_picker_create(ctypes.byref(h), _ptr(ref), int3(nr, nr, nrz), _ptr(refmask),
                                        int3(nx, ny, nz), batch, _stream(stream)), "gtomx_picker_create")
        self.handle = h
        self.imageshape = tuple(imageshape)

    def close(self):
        if self.handle:
            lib().gtomx_picker_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_image(self, image, ctf, stream=None):
        """Normalises `image` in place and keeps `ctf`; afterwards `image` holds its
        square, as in the reference. Keep `ctf` alive while correlating."""
        self._ctf = ctf
        check(lib().gtomx_picker_set_image(self.handle, _ptr(image), _ptr(ctf), _stream(stream)),
              "gtomx_picker_set_image")

    def perform_correlation(self, best, bestangle, anglestep=None, angles=None, stream=None):
        """Updates best (z, y, x) and bestangle (z, y, x, 3) in place over the
        orientations of GetEqualAngularSpacing(anglestep) or an explicit list."""
        if angles is not None:
            a = _angles_arr(angles)
            check(lib().gtomx_picker_correlate_angles(self.handle, ctypes.c_void_p(a.ctypes.data), len(a),
                                                      _ptr(best), _ptr(bestangle), _stream(stream)),
                  "gtomx_picker_correlate_angles")
        else:
            check(lib().gtomx_picker_correlate(self.handle, float(anglestep), _ptr(best), _ptr(bestangle),
                                               _stream(stream)), "gtomx_picker_correlate")


class ProbabilisticPicker:
    """The ProbabilisticPicker class (include/Correlation.cuh:166-207) with circular
    masks: refs and masks are (nrefs, z, y, x) or (nrefs, y, x) float32 CUDA
    tensors; imageshape (z, y, x) or (y, x)."""

    def __init__(self, refs, refmasks, imageshape, lowpassfreq, doctf=False, ismaskcircular=True, stream=None):
        nd = len(imageshape)
        rshape = tuple(refs.shape[-nd:])
        nrefs = refs.numel() // int(np.prod(rshape))
        dr = int3(rshape[-1], rshape[-2], rshape[0] if nd == 3 else 1)
        di = int3(imageshape[-1], imageshape[-2], imageshape[0] if nd == 3 else 1)
        h = _P()
        check(lib().gtomx_prob_picker_create(ctypes.byref(h), _ptr(refs), dr, nrefs, _ptr(refmasks),
                                             int(bool(ismaskcircular)), int(bool(doctf)), di, int(lowpassfreq),
                                             _stream(stream)), "gtomx_prob_picker_create")
        self.handle = h
        self.ndims = nd
        self.imageshape = tuple(imageshape)

    def close(self):
        if self.handle:
            lib().gtomx_prob_picker_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_image(self, image, ctf=None, stream=None):
        check(lib().gtomx_prob_picker_set_image(self.handle, _ptr(image), _ptr(ctf), _stream(stream)),
              "gtomx_prob_picker_set_image")

    def count_angles(self, anglestep):
        c = _I()
        check(lib().gtomx_prob_picker_count_angles(self.ndims, float(anglestep), ctypes.byref(c)),
              "gtomx_prob_picker_count_angles")
        return c.value

    def perform_correlation(self, n, anglestep, bestccf, bestangle, bestref, stream=None):
        """Updates bestccf (image shape, float32), bestangle (image shape + (3,),
        float32) and bestref (image shape, int32) in place for reference n."""
        check(lib().gtomx_prob_picker_correlate(self.handle, int(n), float(anglestep), _ptr(bestccf),
                                                _ptr(bestangle), _ptr(bestref), _stream(stream)),
              "gtomx_prob_picker_correlate")


def _angles_arr(angles):
    a = np.ascontiguousarray(angles, np.float32).reshape(-1, 3)
    return a


def d_PickSubTomograms(plan, projector_ft, oversample, exp_ft, ctf, angles, maskradius,
                       ctf_volume_stride=0, best=None, bestidx=None, bestangle=None, stream=None):
    """Exhaustive template match; exp_ft is (nvol, n, n, n/2+1) complex64 for volumes
    or (nvol, n, n/2+1) for images (a 2D plan).

    Returns (best, bestidx[, bestangle]) with best normalised by the local std.
    """
    import torch
    n, p, nz = plan.n, plan.p, plan.nz
    nvol = exp_ft.numel() // (nz * n * (n // 2 + 1))
    dev = exp_ft.device
    shape = (nvol, n, n, n) if nz > 1 else (nvol, n, n)
    if best is None:
        best = torch.empty(shape, dtype=torch.float32, device=dev)
    if bestidx is None:
        bestidx = torch.empty(shape, dtype=torch.int32, device=dev)
    a = _angles_arr(angles)
    check(lib().gtomx_pick_subtomograms(plan.handle, _ptr(projector_ft), oversample, int3(p, p, p), _ptr(exp_ft),
                                        _ptr(ctf), ctf_volume_stride, int3(n, n, nz), nvol, a.ctypes.data,
                                        a.shape[0], maskradius, _ptr(best), _ptr(bestidx), _ptr(bestangle),
                                        _stream(stream)), "gtomx_pick_subtomograms")
    return best, bestidx


def d_PickSubTomogramsDiff2(projector_ft, p, oversample, exp_ft, ctf, relevant, angles, stream=None):
    """d_PickSubTomogramsDiff2 (src/Correlation/SubTomograms.cu:165-313): exp_ft and ctf are
    (nvol, n, n, n/2+1); relevant = (rz, ry, rx). Returns (best score, orientation index
    as float), each (nvol, rz, ry, rx)."""
    import torch
    n = exp_ft.shape[-2]
    nvol = exp_ft.numel() // (n * n * (n // 2 + 1))
    rz, ry, rx = relevant
    best = torch.empty((nvol, rz, ry, rx), dtype=torch.float32, device=exp_ft.device)
    ang = torch.empty_like(best)
    a = _angles_arr(angles)
    check(lib().gtomx_pick_subtomograms_diff2(_ptr(projector_ft), oversample, int3(p, p, p), _ptr(exp_ft), _ptr(ctf),
                                              int3(n, n, n), nvol, int3(rx, ry, rz), a.ctypes.data, a.shape[0],
                                              _ptr(best), _ptr(ang), _stream(stream)),
          "gtomx_pick_subtomograms_diff2")
    return best, ang


def pick_begin(plan, nvol, best, bestidx, stream=None):
    n = plan.n
    check(lib().gtomx_pick_begin(plan.handle, int3(n, n, plan.nz), nvol, _ptr(best), _ptr(bestidx), _stream(stream)),
          "gtomx_pick_begin")


def pick_update(plan, projector_ft, oversample, exp_ft, ctf, angles, angle_offset, best, bestidx,
                ctf_volume_stride=0, stream=None):
    n, p = plan.n, plan.p
    nvol = exp_ft.numel() // (plan.nz * n * (n // 2 + 1))
    a = _angles_arr(angles)
    check(lib().gtomx_pick_update(plan.handle, _ptr(projector_ft), oversample, int3(p, p, p), _ptr(exp_ft),
                                  _ptr(ctf), ctf_volume_stride, int3(n, n, plan.nz), nvol, a.ctypes.data, a.shape[0],
                                  angle_offset, _ptr(best), _ptr(bestidx), _stream(stream)), "gtomx_pick_update")


def pick_finish(plan, exp_ft, maskradius, best, bestidx, bestangle=None, stream=None):
    n = plan.n
    nvol = exp_ft.numel() // (plan.nz * n * (n // 2 + 1))
    check(lib().gtomx_pick_finish(plan.handle, _ptr(exp_ft), int3(n, n, plan.nz), nvol, maskradius, _ptr(best),
                                  _ptr(bestidx), _ptr(bestangle), _stream(stream)), "gtomx_pick_finish")


# operand types of the split-precision matrix-core stages (GTOMX_SPLIT_* / GTOMX_SINGLE_*)
SPLIT_FP16, SPLIT_BF16, SINGLE_FP16, SINGLE_BF16 = 0, 1, 2, 3
PRECISIONS = {"fp16": SPLIT_FP16, "bf16": SPLIT_BF16, "fp16-hi": SINGLE_FP16, "bf16-hi": SINGLE_BF16}


def pick_update_split(plan, precision, projector_ft, oversample, exp_ft, ctf, angles, angle_offset, best, bestidx,
                      ctf_volume_stride=0, stream=None):
    """gtomx_pick_update_split: gtomx_pick_update on 64^3 volumes with the radix-16 stages of every
    per-orientation inverse transform on the matrix cores in `precision` (a PRECISIONS value)."""
    n, p = plan.n, plan.p
    nvol = exp_ft.numel() // (plan.nz * n * (n // 2 + 1))
    a = _angles_arr(angles)
    check(lib().gtomx_pick_update_split(plan.handle, int(precision), _ptr(projector_ft), oversample, int3(p, p, p),
                                        _ptr(exp_ft), _ptr(ctf), ctf_volume_stride, int3(n, n, plan.nz), nvol,
                                        a.ctypes.data, a.shape[0], angle_offset, _ptr(best), _ptr(bestidx),
                                        _stream(stream)), "gtomx_pick_update_split")


def d_PickSubTomogramsSplit(plan, precision, projector_ft, oversample, exp_ft, ctf, angles, maskradius,
                            ctf_volume_stride=0, stream=None):
    """The exhaustive search (begin, update, finish) with the split-precision update; returns (best, bestidx)."""
    import torch
    n = plan.n
    nvol = exp_ft.numel() // (plan.nz * n * (n // 2 + 1))
    best = torch.empty((nvol, n, n, n), dtype=torch.float32, device=exp_ft.device)
    idx = torch.empty((nvol, n, n, n), dtype=torch.int32, device=exp_ft.device)
    pick_begin(plan, nvol, best, idx, stream=stream)
    pick_update_split(plan, precision, projector_ft, oversample, exp_ft, ctf, angles, 0, best, idx,
                      ctf_volume_stride=ctf_volume_stride, stream=stream)
    pick_finish(plan, exp_ft, maskradius, best, idx, stream=stream)
    return best, idx


def split_dft64(x, precision, stream=None):
    """Unnormalised inverse 64-point DFTs of the rows of a complex64 CUDA tensor (..., 64) through the
    split-precision matrix-core transform; returns a new tensor."""
    import torch
    assert x.dtype == torch.complex64 and x.shape[-1] == 64 and x.is_contiguous()
    out = torch.empty_like(x)
    check(lib().gtomx_split_dft64(_ptr(x), _ptr(out), x.numel() // 64, int(precision), _stream(stream)),
          "gtomx_split_dft64")
    return out


def d_Exact2DWeighting(weights, angles, imageweights=None, indices=None, maxfreq=None, stream=None):
    """d_Exact2DWeighting (src/Projection/Weighting.cu:21-55) in place on float32 CUDA filters
    (batch, ny, nx/2+1) of the images `indices` (default all) of a tilt series."""
    batch, ny, nh = weights.shape
    nx = 2 * (nh - 1)
    a = _angles_arr(angles)
    ind = np.ascontiguousarray(np.arange(batch) if indices is None else indices, np.int32)
    check(lib().gtomx_exact2d_weighting(_ptr(weights), int3(nx, ny, 1), ind.ctypes.data, a.ctypes.data,
                                        _ptr(imageweights), a.shape[0], float(nx if maxfreq is None else maxfreq), 0,
                                        batch, _stream(stream)), "gtomx_exact2d_weighting")
    return weights


class TomoPicker:
    """The TomoPicker class (include/Correlation.cuh:209-257): initialize with a cubic reference and
    mask (z, y, x) and the tilt-image shape (ny, nx) and count; set_image; perform_correlation."""

    def __init__(self, ref, mask, imageshape, nimages, circular=True, stream=None):
        nr = ref.shape[-1]
        ny, nx = imageshape
        h = _P()
        check(lib().gtomx_tomopicker_create(ctypes.byref(h), _ptr(ref), int3(nr, nr, nr), _ptr(mask),
                                            int(bool(circular)), int3(nx, ny, 1), int(nimages), _stream(stream)),
              "gtomx_tomopicker_create")
        self.handle = h

    def close(self):
        if self.handle:
            lib().gtomx_tomopicker_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_image(self, images, tilt_angles, weights, stream=None):
        """images (nimages, ny, nx) float32 CUDA, weighted in place as in the reference."""
        a = _angles_arr(tilt_angles)
        w = np.ascontiguousarray(weights, np.float32)
        check(lib().gtomx_tomopicker_set_image(self.handle, _ptr(images), a.ctypes.data, w.ctypes.data,
                                               _stream(stream)), "gtomx_tomopicker_set_image")

    def perform_correlation(self, corr, corrangle=None, anglestep=None, angles=None, index=None, sum1=None,
                            sum2=None, stream=None):
        """Updates corr (z, y, x) (and corrangle (z, y, x, 3), index, sum1, sum2 when given) over
        GetEqualAngularSpacing(anglestep) or an explicit orientation list."""
        cz, cy, cx = corr.shape
        if angles is not None:
            a = _angles_arr(angles)
            check(lib().gtomx_tomopicker_correlate_angles(self.handle, a.ctypes.data, a.shape[0], _ptr(corr),
                                                          _ptr(corrangle), int3(cx, cy, cz), _ptr(index), _ptr(sum1),
                                                          _ptr(sum2), _stream(stream)),
                  "gtomx_tomopicker_correlate_angles")
        else:
            check(lib().gtomx_tomopicker_correlate(self.handle, _ptr(corr), _ptr(corrangle), int3(cx, cy, cz),
                                                   float(anglestep), _stream(stream)), "gtomx_tomopicker_correlate")


class Micrograph(ctypes.Structure):
    """gtomx_micrograph."""
    _fields_ = [("n", ctypes.c_int), ("d_spectrum", ctypes.c_void_p), ("d_ctf", ctypes.c_void_p),
                ("maskradius", ctypes.c_float), ("d_best", ctypes.c_void_p), ("d_index", ctypes.c_void_p)]


class MicrographBatch:
    """gtomx_mpick: the 2D search over micrographs of several sides, one plan per side (batches from
    `batches` and row variants from `variants`, e.g. the autotuner's table), per-side orientation loops
    captured as hipGraphs."""

    def __init__(self, p, sides, batches=None, variants=None):
        sides = [int(n) for n in sides]
        h = _P()
        s = (ctypes.c_int * len(sides))(*sides)
        b = (ctypes.c_uint * len(sides))(*(batches or [0] * len(sides)))
        v = (ctypes.c_int * len(sides))(*(variants or [ROWS_PLAIN] * len(sides)))
        check(lib().gtomx_mpick_create(ctypes.byref(h), int(p), s, b, v, len(sides)), "gtomx_mpick_create")
        self.handle, self.p, self.sides = h, int(p), sides

    def close(self):
        if self.handle:
            lib().gtomx_mpick_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        b = (ctypes.c_uint * len(self.sides))()
        v = (ctypes.c_int * len(self.sides))()
        cap = ctypes.c_ulonglong()
        rep = ctypes.c_ulonglong()
        check(lib().gtomx_mpick_info(self.handle, b, v, ctypes.byref(cap), ctypes.byref(rep)), "gtomx_mpick_info")
        return {"batches": dict(zip(self.sides, list(b))), "variants": dict(zip(self.sides, list(v))),
                "captures": cap.value, "replays": rep.value}

    def run(self, projector_ft, images, angles, use_graphs=True, stream=None):
        """images: list of dicts with 'spectrum' (complex64 (n, n/2+1)), 'ctf' (float32 (n, n/2+1)),
        'maskradius', 'best' (float32 (n, n)) and 'index' (int32 (n, n)) CUDA tensors."""
        arr = (Micrograph * len(images))()
        for k, im in enumerate(images):
            n = im["spectrum"].shape[-2]
            arr[k] = Micrograph(n, _ptr(im["spectrum"]), _ptr(im["ctf"]), float(im["maskradius"]), _ptr(im["best"]),
                                _ptr(im["index"]))
        a = _angles_arr(angles)
        check(lib().gtomx_mpick_run(self.handle, _ptr(projector_ft), arr, len(images), a.ctypes.data, a.shape[0],
                                    1 if use_graphs else 0, _stream(stream)), "gtomx_mpick_run")


def mpick_tune(p, n, projector_ft, batches, angles, repeats=1, stream=None, variants=None):
    """Milliseconds per orientation of the 2D search's orientation loop on an n x n image for each
    candidate (batches[c], variants[c] or ROWS_PLAIN) (gtomx_mpick_tune; infinity where the plan does
    not fit)."""
    a = _angles_arr(angles)
    b = (ctypes.c_uint * len(batches))(*[int(x) for x in batches])
    v = (ctypes.c_int * len(batches))(*(variants or [ROWS_PLAIN] * len(batches)))
    out = (ctypes.c_float * len(batches))()
    check(lib().gtomx_mpick_tune(int(p), int(n), _ptr(projector_ft), b, v, len(batches), a.ctypes.data, a.shape[0],
                                 int(repeats), out, _stream(stream)), "gtomx_mpick_tune")
    return [float(x) for x in out]


class LearnedStats(ctypes.Structure):
    """gtomx_learned_stats."""
    _fields_ = [("tiles", ctypes.c_uint), ("exhaustive", ctypes.c_uint), ("shortlisted", ctypes.c_uint),
                ("skipped", ctypes.c_uint), ("groups_searched", ctypes.c_uint)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


def pick_update_tiles(plan, projector_ft, oversample, exp_ft, ctf, angles, group, tilemasks, best, bestidx,
                      ctf_volume_stride=0, stream=None):
    """gtomx_pick_update_tiles: angles (ngroups*group, 3); tilemasks uint8 (nvol, ngroups, nt, nt) with
    bit tx of [.., tz, ty] selecting the 32^3 tile (tx, ty, tz)."""
    n, p = plan.n, plan.p
    nvol = exp_ft.numel() // (n * n * (n // 2 + 1))
    a = _angles_arr(angles)
    ngroups = a.shape[0] // group
    m = np.ascontiguousarray(tilemasks, np.uint8)
    nt = n // 32
    assert m.size == nvol * ngroups * nt * nt, (m.shape, nvol, ngroups, nt)
    check(lib().gtomx_pick_update_tiles(plan.handle, _ptr(projector_ft), oversample, int3(p, p, p), _ptr(exp_ft),
                                        _ptr(ctf), ctf_volume_stride, int3(n, n, n), nvol, a.ctypes.data, ngroups,
                                        group, m.ctypes.data, _ptr(best), _ptr(bestidx), _stream(stream)),
          "gtomx_pick_update_tiles")


def d_PickSubTomogramsLearned(plan, net, projector_ft, oversample, exp_ft, ctf, angles, maskradius, npsi=18,
                              topk=16, confidence=0.5, ctf_volume_stride=0, best=None, bestidx=None,
                              bestangle=None, stream=None):
    """The learned search (gtomx_pick_subtomograms_learned): returns (best, bestidx, stats dict)."""
    import torch
    n, p = plan.n, plan.p
    nvol = exp_ft.numel() // (n * n * (n // 2 + 1))
    dev = exp_ft.device
    if best is None:
        best = torch.empty((nvol, n, n, n), dtype=torch.float32, device=dev)
    if bestidx is None:
        bestidx = torch.empty((nvol, n, n, n), dtype=torch.int32, device=dev)
    a = _angles_arr(angles)
    st = LearnedStats()
    check(lib().gtomx_pick_subtomograms_learned(plan.handle, net.handle, _ptr(projector_ft), oversample,
                                                int3(p, p, p), _ptr(exp_ft), _ptr(ctf), ctf_volume_stride,
                                                int3(n, n, n), nvol, a.ctypes.data, a.shape[0] // npsi, npsi, topk,
                                                confidence, maskradius, _ptr(best), _ptr(bestidx), _ptr(bestangle),
                                                ctypes.byref(st), _stream(stream)),
          "gtomx_pick_subtomograms_learned")
    return best, bestidx, st.as_dict()
