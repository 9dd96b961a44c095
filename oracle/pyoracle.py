"""ctypes wrapper around the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, always as the checker or the timed host
baseline, never as part of the product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_d = ctypes.c_double
_f = ctypes.c_float
_i = ctypes.c_int
_ll = ctypes.c_longlong
_p = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_num_threads.restype = _i
        L.orc_set_num_threads.argtypes = [_i]
        L.orc_fft_c2c_1d_f64.argtypes = [_p, _i, _i]
        for nm in ("orc_r2c_3d_f64", "orc_c2r_3d_f64", "orc_r2c_3d_f32", "orc_c2r_3d_f32"):
            getattr(L, nm).argtypes = [_p, _p, _i, _i, _i, _i]
        L.orc_euler_matrix.argtypes = [_d, _d, _d, _p]
        L.orc_projector_matrix.argtypes = [_d, _d, _d, _d, _p]
        L.orc_project_f64.argtypes = [_p, _i, _p, _i, _p, _i]
        L.orc_equal_angular_spacing.argtypes = [_f, _f, _f, _f, _f, _f, _f, _p, _i]
        L.orc_equal_angular_spacing.restype = _i
        L.orc_ctf3_f64.argtypes = [_p, _i, _d, _d, _d, _d, _d, _d, _d, _d]
        L.orc_sphere_mask_f64.argtypes = [_p, _i, _i, _i, _d, _d]
        L.orc_remap_full2fullfft_f64.argtypes = [_p, _p, _i, _i, _i]
        L.orc_local_std_f64.argtypes = [_p, _i, _i, _i, _d, _p, _p, _p]
        L.orc_update_max_argmax_f64.argtypes = [_p, _ll, _i, _i, _p, _p]
        L.orc_local_peaks_f32.argtypes = [_p, _i, _i, _i, _i, _f, _p, _ll]
        L.orc_local_peaks_f32.restype = _ll
        L.orc_diff2_f64.argtypes = [_p, _i, _d, _p, _p, _i, _i, _i, _i, _p, _i, _p, _p, _p]
        L.orc_ctf_simulate_f64.argtypes = [_p, _p, _p, _p, _i, _i, _i, _i]
        L.orc_full_pad_f64.argtypes = [_p, _p, _i, _i, _i, _i, _i, _i]
        L.orc_peak_f64.argtypes = [_p, _i, _i, _i, _i, _i, _p, _p]
        L.orc_ctf_tilt_zgrid_f64.argtypes = [_p, _d, _d, _i, _i, _i, _i, _p, _i, _p]
        L.orc_rotate3d_ft_f64.argtypes = [_p, _p, _i, _p, _i, _i, _i]
        L.orc_norm_monolithic_f64.argtypes = [_p, _p, _ll, _i, _i]
        L.orc_fft_resize_f64.argtypes = [_i, _p, _p, _p, _p, _i]
        L.orc_sphere_mask_ex_f64.argtypes = [_p, _i, _i, _i, _d, _d, _p, _i]
        L.orc_peak_one_2d_f64.argtypes = [_p, _i, _i, _i, _i, _i, _i, _p, _p]
        L.orc_subpixel_max_f64.argtypes = [_p, _p, _i, _i, _i, _i]
        L.orc_norm_mean01std_f64.argtypes = [_p, _p, _ll]
        L.orc_picker_f64.argtypes = [_p, _i, _i, _p, _i, _i, _i, _p, _p, _i, _p, _p, _p]
        L.orc_exact2d_weighting_f64.argtypes = [_p, _i, _i, _p, _p, _i, _p, _i, _d]
        L.orc_tomo_picker_f64.argtypes = [_p, _p, _i, _i, _p, _i, _i, _i, _p, _p, _p, _i, _i, _i, _i, _p, _p, _p,
                                          _p, _p]
        pick_args = [_p, _i, _d, _p, _p, _i, _i, _i, _p, _i, _d, _p, _p, _p, _p, _p]
        L.orc_pick_f64.argtypes = pick_args
        L.orc_pick_f64.restype = _d
        L.orc_pick_f32.argtypes = pick_args
        L.orc_pick_f32.restype = _d
        L.orc_pick2d_f64.argtypes = pick_args
        L.orc_pick2d_f64.restype = _d
        L.orc_project2d_f64.argtypes = [_p, _i, _p, _i, _p, _i]
        L.orc_rotate2d_ft_f64.argtypes = [_p, _p, _i, _i, _d, _d, _i, _i]
        L.orc_prob_picker_f64.argtypes = [_p, _p, _p, _p, _i, _i, _p, _p, _f, _i, _p, _p, _p, _p, _i]
        L.orc_bspline_prefilter_f64.argtypes = [_p, _i, _i, _i]
        L.orc_rotate_cubic_f64.argtypes = [_p, _p, _i, _i, _i, _p]
        L.orc_gaussrand_f32.argtypes = [ctypes.c_uint, _i, _p, _i]
        L.orc_pnet_param_count.restype = _i
        L.orc_pnet_forward_f64.argtypes = [_p, _p, _i, _i, _p]
        L.orc_pnet_gradients_f64.argtypes = [_p, _p, _p, _i, _i, _p, _p]
        L.orc_pnet_gradients_f64.restype = _d
        _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(t):
    lib().orc_set_num_threads(int(t))


def fft_1d(x, sign=-1):
    """Unnormalised complex DFT of a 1D array (sign -1 forward, +1 inverse)."""
    a = _c(x, np.complex128).copy()
    lib().orc_fft_c2c_1d_f64(_ptr(a), a.size, sign)
    return a


def r2c(vol, dtype=np.float64):
    """Unnormalised forward R2C of a (nz, ny, nx) or (batch, nz, ny, nx) volume."""
    v = _c(vol, dtype)
    shp = v.shape if v.ndim == 4 else (1,) + v.shape
    b, nz, ny, nx = shp
    cdt = np.complex128 if dtype == np.float64 else np.complex64
    out = np.empty((b, nz, ny, nx // 2 + 1), cdt)
    fn = lib().orc_r2c_3d_f64 if dtype == np.float64 else lib().orc_r2c_3d_f32
    fn(_ptr(v), _ptr(out), nx, ny, nz, b)
    return out if vol.ndim == 4 else out[0]


def c2r(spec, nx, dtype=np.float64):
    """Unnormalised inverse C2R (real-part semantics) to width nx."""
    cdt = np.complex128 if dtype == np.float64 else np.complex64
    s = _c(spec, cdt)
    shp = s.shape if s.ndim == 4 else (1,) + s.shape
    b, nz, ny, nh = shp
    assert nh == nx // 2 + 1
    out = np.empty((b, nz, ny, nx), dtype)
    fn = lib().orc_c2r_3d_f64 if dtype == np.float64 else lib().orc_c2r_3d_f32
    fn(_ptr(s), _ptr(out), nx, ny, nz, b)
    return out if spec.ndim == 4 else out[0]


def euler_matrix(rot, tilt, psi):
    A = np.empty(9)
    lib().orc_euler_matrix(rot, tilt, psi, _ptr(A))
    return A.reshape(3, 3)


def projector_matrix(rot, tilt, psi, s):
    M = np.empty(9)
    lib().orc_projector_matrix(rot, tilt, psi, s, _ptr(M))
    return M.reshape(3, 3)


def project(vol_ft, p, n, M, rmax=None):
    """Project the (p/2+1, p, p) half spectrum with row-major matrix M to (n/2+1, n, n)."""
    v = _c(vol_ft, np.complex128)
    assert v.shape == (p, p, p // 2 + 1)
    out = np.empty((n, n, n // 2 + 1), np.complex128)
    Mm = _c(M, np.float64).reshape(9)
    lib().orc_project_f64(_ptr(v), p, _ptr(out), n, _ptr(Mm), n // 2 if rmax is None else int(rmax))
    return out


def project2d(vol_ft, p, n, M, rmax=None):
    """2D central section: the (p/2+1, p, p) half spectrum sampled at M (x, y, 0) onto (n, n/2+1)."""
    v = _c(vol_ft, np.complex128)
    assert v.shape == (p, p, p // 2 + 1)
    out = np.empty((n, n // 2 + 1), np.complex128)
    Mm = _c(M, np.float64).reshape(9)
    lib().orc_project2d_f64(_ptr(v), p, _ptr(out), n, _ptr(Mm), n // 2 if rmax is None else int(rmax))
    return out


def equal_angular_spacing(phirange, thetarange, psirange, increment):
    L = lib()
    args = [np.float32(phirange[0]), np.float32(phirange[1]), np.float32(thetarange[0]),
            np.float32(thetarange[1]), np.float32(psirange[0]), np.float32(psirange[1]),
            np.float32(increment)]
    cnt = L.orc_equal_angular_spacing(*args, None, 0)
    out = np.empty((cnt, 3), np.float32)
    L.orc_equal_angular_spacing(*args, _ptr(out), cnt)
    return out


def ctf3(n, pixelsize_m, voltage_V, Cs_m, defocus_m, amplitude, phaseshift=0.0, bfactor_m2=0.0,
         scale=1.0):
    out = np.empty((n, n, n // 2 + 1))
    lib().orc_ctf3_f64(_ptr(out), n, pixelsize_m, voltage_V, Cs_m, defocus_m, amplitude,
                       phaseshift, bfactor_m2, scale)
    return out


def sphere_mask(shape, radius, sigma=1.0):
    nz, ny, nx = shape
    out = np.empty(shape)
    lib().orc_sphere_mask_f64(_ptr(out), nx, ny, nz, radius, sigma)
    return out


def remap_full2fullfft(vol):
    v = _c(vol, np.float64)
    out = np.empty_like(v)
    nz, ny, nx = v.shape
    lib().orc_remap_full2fullfft_f64(_ptr(v), _ptr(out), nx, ny, nz)
    return out


def local_std(vol, radius, with_mean=False, fouriermask=None):
    v = _c(vol, np.float64)
    nz, ny, nx = v.shape
    fm = None if fouriermask is None else _c(fouriermask, np.float64)
    if fm is not None:
        assert fm.shape == (nz, ny, nx // 2 + 1)
    sd = np.empty_like(v)
    mean = np.empty_like(v) if with_mean else None
    lib().orc_local_std_f64(_ptr(v), nx, ny, nz, radius, _ptr(sd), _ptr(mean) if with_mean else None,
                            None if fm is None else _ptr(fm))
    return (sd, mean) if with_mean else sd


def update_max_argmax(corr, best, bestidx, batchoffset):
    c = _c(corr, np.float64)
    b = best  # updated in place
    assert b.dtype == np.float64 and bestidx.dtype == np.int32
    batch = c.shape[0]
    L = c[0].size
    lib().orc_update_max_argmax_f64(_ptr(c), L, batch, batchoffset, _ptr(b), _ptr(bestidx))


def local_peaks(vol, localextent, threshold):
    """d_LocalPeaks on one float32 volume (nz, ny, nx): (k, 3) int array of (x, y, z) in raster order."""
    v = _c(vol, np.float32)
    nz, ny, nx = v.shape
    cap = 4096
    while True:
        out = np.empty(cap, np.int64)
        k = lib().orc_local_peaks_f32(_ptr(v), nx, ny, nz, int(localextent), float(threshold), _ptr(out), cap)
        if k <= cap:
            break
        cap = k
    lin = out[:k]
    return np.stack([lin % nx, (lin // nx) % ny, lin // (nx * ny)], axis=1).astype(np.int64)


def diff2(projft, p, oversample, expft, ctf, n, relevant, angles, with_second=False):
    """d_PickSubTomogramsDiff2 on one n^3 volume; relevant = (rz, ry, rx).
    Returns (best score, orientation index[, second-best score]), each of shape relevant."""
    P = _c(projft, np.complex128)
    E = _c(expft, np.complex128)
    C = _c(ctf, np.float64)
    A = _c(angles, np.float64).reshape(-1, 3)
    rz, ry, rx = relevant
    best = np.empty(relevant, np.float64)
    idx = np.empty(relevant, np.int32)
    second = np.empty(relevant, np.float64)
    lib().orc_diff2_f64(_ptr(P), p, oversample, _ptr(E), _ptr(C), n, rx, ry, rz, _ptr(A), len(A), _ptr(best),
                        _ptr(idx), _ptr(second))
    return (best, idx, second) if with_second else (best, idx)


CTF_FIELDS = ("pixelsize", "pixeldelta", "pixelangle", "Cs", "voltage", "defocus", "astigmatismangle",
              "defocusdelta", "amplitude", "Bfactor", "Bfactordelta", "Bfactorangle", "scale", "phaseshift")
CTF_DEFAULTS = dict(pixelsize=1e-10, pixeldelta=0.0, pixelangle=0.0, Cs=2e-3, voltage=300e3, defocus=-3e-6,
                    astigmatismangle=0.0, defocusdelta=0.0, amplitude=0.07, Bfactor=0.0, Bfactordelta=0.0,
                    Bfactorangle=0.0, scale=1.0, phaseshift=0.0)


def ctf_simulate(params, addresses, gammacorr=None, ampsq=False, ignorefirst=False):
    """d_CTFSimulate: params is a list of dicts (CTFParams fields, SI units), addresses
    (n, 2) of (k [cycles/pixel], angle). Returns (batch, n) float64."""
    P = np.array([[dict(CTF_DEFAULTS, **p)[f] for f in CTF_FIELDS] for p in params], np.float64)
    A = _c(addresses, np.float64).reshape(-1, 2)
    out = np.empty((len(P), len(A)), np.float64)
    g = None if gammacorr is None else _c(gammacorr, np.float64)
    lib().orc_ctf_simulate_f64(_ptr(P), _ptr(A), None if g is None else _ptr(g), _ptr(out), len(A), int(ampsq),
                               int(ignorefirst), len(P))
    return out


def ctf_tilt_zgrid(imageangle, stageangle, specimenangle, defocus, pixelsize, dimsimage, dimsregion, positions):
    """CTFTiltParams::GetZGrid2D: defocus under each region centre; positions (n, 2 or 3) of (x, y[, z])."""
    a = np.array([imageangle, stageangle[0], stageangle[1], specimenangle[0], specimenangle[1]], np.float64)
    p = np.zeros((len(positions), 3), np.int32)
    p[:, :np.asarray(positions).shape[1]] = positions
    z = np.empty(len(p), np.float64)
    lib().orc_ctf_tilt_zgrid_f64(_ptr(a), float(defocus), float(pixelsize), int(dimsimage[0]), int(dimsimage[1]),
                                 int(dimsregion[0]), int(dimsregion[1]), _ptr(p), len(p), _ptr(z))
    return z


def rotate3d_ft(vol, M, zerocentred=True, cubic=False):
    """d_Rotate3DFT (linear) of one centred half spectrum (n, n, n/2+1), complex or
    real, by the row-major matrix M = transpose(Euler)."""
    v = np.asarray(vol)
    cplx = np.iscomplexobj(v)
    src = np.ascontiguousarray(v, np.complex128).view(np.float64) if cplx else _c(v, np.float64)
    n = v.shape[0]
    out = np.empty_like(src)
    lib().orc_rotate3d_ft_f64(_ptr(src), _ptr(out), n, _ptr(_c(M, np.float64)), int(zerocentred), 2 if cplx else 1,
                              int(bool(cubic)))
    return out.view(np.complex128) if cplx else out


def ctf_addresses_2d(n):
    """(k, angle) of the (n, n/2+1) half-spectrum bins: k = |(x, y)|/n, angle = atan2(y, x)."""
    y = np.fft.fftfreq(n, 1 / n)[:, None]
    x = np.arange(n // 2 + 1)[None, :]
    yy, xx = np.broadcast_arrays(y, x)
    return np.stack([np.hypot(xx, yy) / n, np.arctan2(yy, xx)], axis=-1).reshape(-1, 2)


PEAK_MODES = {"integer": 1, "subcoarse": 2, "subfine": 3}


def peak(maps, mode, dims=None):
    """d_Peak on float32 maps: maps is (z, y, x) or (batch, z, y, x) (or (y, x) /
    (batch, y, x) with dims=2). Returns (positions (batch, 3) as (x, y, z), values (batch,))."""
    m = _c(maps, np.float32)
    nd = dims or (m.ndim if m.ndim <= 3 else 3)
    if m.ndim == nd:
        m = m[None]
    shape = m.shape[1:]
    nx = shape[-1]
    ny = shape[-2] if nd >= 2 else 1
    nz = shape[-3] if nd >= 3 else 1
    b = m.shape[0]
    pos = np.empty((b, 3), np.float64)
    val = np.empty(b, np.float64)
    lib().orc_peak_f64(_ptr(m), nx, ny, nz, PEAK_MODES.get(mode, mode), b, _ptr(pos), _ptr(val))
    return pos, val


def peak_one_2d(images, region, subtractcenter):
    """d_PeakOne2D on (batch, ny, nx) float32 images, region (rx, ry)."""
    m = _c(images, np.float32)
    if m.ndim == 2:
        m = m[None]
    b, ny, nx = m.shape
    pos = np.empty((b, 3), np.float64)
    val = np.empty(b, np.float64)
    lib().orc_peak_one_2d_f64(_ptr(m), nx, ny, int(region[0]), int(region[1]), int(subtractcenter), b, _ptr(pos),
                              _ptr(val))
    return pos, val


def subpixel_max(vol, out, steps):
    """d_SubpixelMax: vol (z, y, x) or (y, x) float32; out (same shape, float64) is updated in place."""
    v = _c(vol, np.float32)
    assert out.dtype == np.float64 and out.flags.c_contiguous and out.shape == v.shape
    nx = v.shape[-1]
    ny = v.shape[-2]
    nz = v.shape[-3] if v.ndim == 3 else 1
    lib().orc_subpixel_max_f64(_ptr(v), _ptr(out), nx, ny, nz, int(steps))
    return out


def full_pad(vol, newshape):
    """d_FFTFullPad of a real (z, y, x) volume in FFT order to newshape (z, y, x)."""
    v = _c(vol, np.float64)
    oz, oy, ox = v.shape
    nz, ny, nx = newshape
    out = np.empty(newshape, np.float64)
    lib().orc_full_pad_f64(_ptr(v), _ptr(out), ox, oy, oz, nx, ny, nz)
    return out


def sphere_mask_ex(shape, radius, sigma, center=None, decentered=False):
    """d_SphereMask values for a (z, y, x) or (y, x) shape; center (x, y, z) or None = dims/2."""
    nz, ny, nx = (1,) + tuple(shape) if len(shape) == 2 else tuple(shape)
    c = np.array(center if center is not None else (nx // 2, ny // 2, nz // 2), np.float32)
    out = np.empty((nz, ny, nx), np.float64)
    lib().orc_sphere_mask_ex_f64(_ptr(out), nx, ny, nz, float(radius), float(sigma), _ptr(c), int(decentered))
    return out.reshape(shape)


FFT_RESIZE = {"crop": 1, "pad": 2, "full_crop": 3, "full_pad": 4}


def fft_resize(op, a, olddims, newdims):
    """d_FFTCrop/d_FFTPad (half spectra, a of shape (z, y, x/2+1)) or d_FFTFullCrop/
    d_FFTFullPad (full arrays (z, y, x)); dims are (x, y, z). Real or complex."""
    cplx = np.iscomplexobj(a)
    src = np.ascontiguousarray(a, np.complex128).view(np.float64) if cplx else _c(a, np.float64)
    k = FFT_RESIZE[op]
    nxs = newdims[0] // 2 + 1 if k <= 2 else newdims[0]
    shape = (newdims[2], newdims[1], nxs)
    out = np.empty(shape + ((2,) if cplx else ()), np.float64)
    od = np.array(olddims, np.int32)
    nd = np.array(newdims, np.int32)
    lib().orc_fft_resize_f64(k, _ptr(src), _ptr(out), _ptr(od), _ptr(nd), 2 if cplx else 1)
    return out.view(np.complex128).reshape(shape) if cplx else out


def norm_monolithic(a):
    """d_NormMonolithic over the last axis group: a is (batch, ...) float32; each item normalised."""
    v = _c(a, np.float32)
    b = v.shape[0]
    n = v.size // b
    out = np.empty(v.shape, np.float64)
    lib().orc_norm_monolithic_f64(_ptr(v), _ptr(out), n, b, 0)
    return out


def norm_ft_monolithic(a):
    """d_NormFTMonolithic: a is (batch, ...) complex; each item divided by its RMS magnitude."""
    v = np.ascontiguousarray(a, np.complex64).view(np.float32)
    b = v.shape[0]
    n = v.size // (2 * b)
    out = np.empty(v.shape, np.float64)
    lib().orc_norm_monolithic_f64(_ptr(v), _ptr(out), n, b, 1)
    return out.view(np.complex128)


def norm_mean01std(a):
    """d_Norm with T_NORM_MEAN01STD (population std, double accumulation)."""
    v = _c(a, np.float64)
    out = np.empty_like(v)
    lib().orc_norm_mean01std_f64(_ptr(v), _ptr(out), v.size)
    return out


def exact2d_weighting(nx, ny, angles, weights=None, indices=None, maxfreq=None, start=None):
    """d_Exact2DWeighting (src/Projection/Weighting.cu:21-55) for FFT-order images: the
    (batch, ny, nx/2+1) filters of the images `indices` (default all) of a tilt series with
    (rot, tilt, psi) `angles`, multiplied into `start` (default ones)."""
    A = _c(angles, np.float64).reshape(-1, 3)
    n = len(A)
    w = _c(np.ones(n) if weights is None else weights, np.float64)
    ind = np.ascontiguousarray(np.arange(n) if indices is None else indices, np.int32)
    out = _c(np.ones((len(ind), ny, nx // 2 + 1)) if start is None else start, np.float64).copy()
    lib().orc_exact2d_weighting_f64(_ptr(out), nx, ny, _ptr(A), _ptr(w), n, _ptr(ind), len(ind),
                                    float(nx if maxfreq is None else maxfreq))
    return out


def tomo_picker(ref, mask, circular, images, tilt, weights, angles, dimscorr, corr=None):
    """The TomoPicker pipeline (Initialize, SetImage, PerformCorrelation over `angles`).
    images (nimages, ny, nx) are returned filtered as SetImage leaves them. Returns
    dict(corr, idx, sum1, sum2, images); corr starts at -1e30 unless given."""
    r = _c(ref, np.float64)
    m = _c(mask, np.float64)
    nr = r.shape[-1]
    im = _c(images, np.float64).copy()
    nimg, ny, nx = im.shape
    T = _c(tilt, np.float64).reshape(-1, 3)
    w = _c(weights, np.float64)
    A = _c(angles, np.float64).reshape(-1, 3)
    cx, cy, cz = dimscorr
    c = np.full((cz, cy, cx), -1e30) if corr is None else _c(corr, np.float64).copy()
    idx = np.full((cz, cy, cx), -1, np.int32)
    s1 = np.zeros((cz, cy, cx))
    s2 = np.zeros((cz, cy, cx))
    c2 = np.full((cz, cy, cx), -1e30)
    lib().orc_tomo_picker_f64(_ptr(r), _ptr(m), nr, int(bool(circular)), _ptr(im), nx, ny, nimg, _ptr(T), _ptr(w),
                              _ptr(A), len(A), cx, cy, cz, _ptr(c), _ptr(idx), _ptr(s1), _ptr(s2), _ptr(c2))
    return dict(corr=c, idx=idx, sum1=s1, sum2=s2, images=im, corr2=c2)


def picker(ref, image, ctf, angles):
    """The Picker pipeline (Initialize, SetImage, PerformCorrelation) over an explicit
    angle list: a cubic reference with a (z, y, x) image, or a square 2D reference
    with a (y, x) image (the 2D branch). Returns dict(best, idx, top2) with best/top2
    started at -1e30."""
    r = _c(ref, np.float64)
    nr = r.shape[-1]
    nrz = r.shape[0] if r.ndim == 3 else 1
    im = _c(image, np.float64)
    nz, ny, nx = im.shape if im.ndim == 3 else (1,) + im.shape
    c = _c(ctf, np.float64)
    A = _c(angles, np.float64).reshape(-1, 3)
    best = np.full(im.shape, -1e30)
    top2 = np.full(im.shape, -1e30)
    idx = np.zeros(im.shape, np.int32)
    lib().orc_picker_f64(_ptr(r), nr, nrz, _ptr(im), nx, ny, nz, _ptr(c), _ptr(A), len(A), _ptr(best), _ptr(idx),
                         _ptr(top2))
    return dict(best=best, idx=idx, top2=top2)


def pick(projft, p, oversample, expft, ctf, n, angles, maskradius, ctf_per_volume=False,
         dtype=np.float64, ndims=3):
    """Run the whole path on n^3 volumes (ndims 3) or n^2 images (ndims 2, float64 only).
    Returns dict(best, idx, top1, top2, sigma, loop_seconds)."""
    cdt = np.complex128 if dtype == np.float64 else np.complex64
    P = _c(projft, cdt)
    E = _c(expft, cdt)
    nvol = E.shape[0] if E.ndim == ndims + 1 else 1
    C = _c(ctf, dtype)
    A = _c(angles, np.float64).reshape(-1, 3)
    shape = (nvol,) + (n,) * ndims
    best = np.empty(shape, dtype)
    idx = np.empty(shape, np.int32)
    top1 = np.empty(shape, dtype)
    top2 = np.empty(shape, dtype)
    sigma = np.empty(shape, dtype)
    if ndims == 2:
        assert dtype == np.float64
        fn = lib().orc_pick2d_f64
    else:
        fn = lib().orc_pick_f64 if dtype == np.float64 else lib().orc_pick_f32
    secs = fn(_ptr(P), p, oversample, _ptr(E), _ptr(C), int(bool(ctf_per_volume)), n, nvol,
              _ptr(A), A.shape[0], maskradius, _ptr(best), _ptr(idx), _ptr(top1), _ptr(top2),
              _ptr(sigma))
    return dict(best=best, idx=idx, top1=top1, top2=top2, sigma=sigma, loop_seconds=secs)


def rotate2d_ft(img, angle, maxfreq, zerocentred=True, cubic=False):
    """d_Rotate2DFT (linear) of one zero-centred complex half spectrum (ny, nx/2+1)."""
    v = np.ascontiguousarray(img, np.complex128)
    ny, nh = v.shape
    out = np.empty_like(v)
    lib().orc_rotate2d_ft_f64(_ptr(v.view(np.float64)), _ptr(out.view(np.float64)), 2 * (nh - 1), ny, float(angle),
                              float(maxfreq), int(zerocentred), int(bool(cubic)))
    return out


def gaussrand(n, seed=123, reset=True):
    """n draws of the oracle's gaussrand after srand(seed); reset restarts its pair phase."""
    out = np.empty(n, np.float32)
    lib().orc_gaussrand_f32(seed, int(reset), _ptr(out), n)
    return out


def bspline_prefilter(vol):
    """d_CubicBSplinePrefilter3D / 2D of a copy of vol ((nz, ny, nx) or (ny, nx))."""
    v = np.array(vol, np.float64, copy=True, order="C")
    shp = v.shape if v.ndim == 3 else (1,) + v.shape
    lib().orc_bspline_prefilter_f64(_ptr(v), shp[2], shp[1], shp[0])
    return v


def rotate_cubic(coef, M):
    """d_Rotate3D / d_Rotate2D (T_INTERP_CUBIC, FFT-order output) of B-spline
    coefficients; M is the 3x3 transpose(Euler) or the 2x2 R(-angle)."""
    c = _c(coef, np.float64)
    shp = c.shape if c.ndim == 3 else (1,) + c.shape
    m = np.zeros(9)
    Mi = np.asarray(M, np.float64)
    if Mi.shape == (2, 2):
        m[[0, 1, 3, 4]] = Mi.ravel()
    else:
        m[:] = Mi.ravel()
    out = np.empty_like(c)
    lib().orc_rotate_cubic_f64(_ptr(c), _ptr(out), shp[2], shp[1], shp[0], _ptr(m))
    return out


def prob_picker(ref, mask, image, lowpass, anglestep, ctf=None, refid=0, best=None, circular=True):
    """ProbabilisticPicker, one reference: returns (best, bestangle, bestref,
    top2), arrays of the image shape (bestangle + (3,)). Starts from best
    (default -1e30 everywhere). circular=False rotates the mask per orientation."""
    r = _c(ref, np.float64)
    m = _c(mask, np.float64)
    im = _c(image, np.float64)
    nd = im.ndim
    dr = np.array(list(r.shape[::-1]) + ([1] if nd == 2 else []), np.int32)
    di = np.array(list(im.shape[::-1]) + ([1] if nd == 2 else []), np.int32)
    b = np.full(im.shape, -1e30) if best is None else _c(best, np.float64).copy()
    ang = np.zeros(im.shape + (3,))
    bref = np.zeros(im.shape, np.int32)
    top2 = np.full(im.shape, -np.inf)
    c = None if ctf is None else _c(ctf, np.float64)
    lib().orc_prob_picker_f64(_ptr(r), _ptr(m), _ptr(dr), _ptr(di), int(lowpass), int(ctf is not None), _ptr(im),
                              None if c is None else _ptr(c), float(anglestep), int(refid), _ptr(b), _ptr(ang),
                              _ptr(bref), _ptr(top2), int(bool(circular)))
    return b, ang, bref, top2


def pnet_param_count():
    return int(lib().orc_pnet_param_count())


def pnet_forward(params, tiles, emulate=False):
    """Logits (n, 501) of the proposal network (oracle/pnet.cpp) for tiles (n, 32, 32, 32)."""
    P = _c(params, np.float64)
    T = _c(tiles, np.float64)
    n = T.shape[0]
    out = np.empty((n, 501), np.float64)
    lib().orc_pnet_forward_f64(_ptr(P), _ptr(T), n, int(bool(emulate)), _ptr(out))
    return out


def pnet_gradients(params, tiles, labels, emulate=False, plant_unflipped=False):
    """(mean loss, gradients (nparams,), per-tile losses (n,)) of the softmax cross-entropy.
    plant_unflipped: a deliberate mistake in the data gradient (for the tests that must catch it)."""
    P = _c(params, np.float64)
    T = _c(tiles, np.float64)
    L = _c(labels, np.int32)
    n = T.shape[0]
    g = np.empty(P.shape[0], np.float64)
    losses = np.empty(n, np.float64)
    flags = int(bool(emulate)) | (2 if plant_unflipped else 0)
    tot = lib().orc_pnet_gradients_f64(_ptr(P), _ptr(T), _ptr(L), n, flags, _ptr(g), _ptr(losses))
    return tot, g, losses
