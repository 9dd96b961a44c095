"""R2C / C2R transforms with extents that have no native kernel (odd lengths,
other prime factors, up to 2048) by the chirp-z form, alone and mixed with
native axes, against numpy's FP64 pocketfft (R2C, and C2R of Hermitian
spectra) and the oracle's real-part C2R (non-Hermitian half spectra). cuFFT
behind d_FFTR2C / d_IFFTC2R (src/FFT/FFT.cu:17-42, src/FFT/IFFT.cu:8-56) plans
any extent. Tolerance 4e-6 normwise, as for the native transforms."""
import numpy as np
import pytest

from tests._util import rel_linf, to_np

pytestmark = pytest.mark.gpu

TOL = 4e-6

SHAPES = [  # (nz, ny, nx); x, y, z each native (N) or chirp-z (C); the convolution length M is the
    # shortest 2^a or 3 * 2^a >= 2L - 1
    (1, 1, 5),  # C
    (1, 1, 100),  # C
    (1, 1, 2047),  # C, the longest chirp-z length (4096-point convolution)
    (1, 75, 100),  # C, C
    (1, 64, 100),  # C x, N y
    (1, 100, 64),  # N x, C y
    (1, 7, 4096),  # N x (two-pass rows), C y
    (20, 30, 50),  # C, C, C
    (27, 64, 36),  # C x, N y, C z
    (40, 8, 200),  # C x, N y, C z
    (3, 2, 1000),  # tiny y and z
    (1, 96, 384),  # C x (384 has native columns but no native row kernel), N y
    (1, 150, 320),  # C x and C y on 3 * 2^a convolution lengths: 768 and 384 (two-pass columns)
    (1, 2, 1500),  # C x on a 3072-point convolution
    (6, 384, 384),  # C x, N y (two-pass), C z
]


def _nd(shape):
    nz, ny, nx = shape
    return 3 if nz > 1 else (2 if ny > 1 else 1)


def _dev(torch, a, shape, cuda):
    """a is (batch, nz, ny, n); the transform's own rank drops the unit axes."""
    nz, ny = shape[:2]
    keep = {3: (nz, ny), 2: (ny,), 1: ()}[_nd(shape)]
    return torch.from_numpy(a).to(cuda).reshape(a.shape[:1] + keep + a.shape[-1:])


@pytest.mark.parametrize("shape", SHAPES)
def test_fft_any_r2c_matches_numpy(gx, cuda, shape):
    import torch
    nz, ny, nx = shape
    rng = np.random.default_rng(nx + 3 * ny + 7 * nz)
    x = rng.standard_normal((2, nz, ny, nx)).astype(np.float32)
    X = gx.d_FFTR2C(_dev(torch, x, shape, cuda), ndims=_nd(shape))
    ref = np.fft.rfftn(x.astype(np.float64), axes=(1, 2, 3))
    err = rel_linf(to_np(X).reshape(ref.shape), ref)
    print(f"R2C {shape}: {err:.2e}")
    assert err < TOL


@pytest.mark.parametrize("shape", SHAPES)
def test_fft_any_roundtrip(gx, cuda, shape):
    """C2R of the R2C (renormalised) gives the input back: numpy's irfftn of the same spectrum."""
    import torch
    nz, ny, nx = shape
    rng = np.random.default_rng(11 * nx + ny + nz)
    x = rng.standard_normal((2, nz, ny, nx)).astype(np.float32)
    X = gx.d_FFTR2C(_dev(torch, x, shape, cuda), ndims=_nd(shape))
    y = gx.d_IFFTC2R(X, nx, ndims=_nd(shape), renormalize=True)
    err = rel_linf(to_np(y).reshape(x.shape), x.astype(np.float64))
    print(f"round trip {shape}: {err:.2e}")
    assert err < TOL


@pytest.mark.parametrize("shape", [(1, 30, 51), (5, 12, 70), (1, 1, 9)])
def test_fft_any_c2r_nonhermitian(gx, orc, cuda, shape):
    """Arbitrary half spectra: the real part of the Hermitian-completed inverse (the imaginary
    parts of the x = 0 and, for even widths, x = nx/2 bins are not read), input unmodified."""
    import torch
    nz, ny, nx = shape
    rng = np.random.default_rng(nx * 5 + ny)
    X = (rng.standard_normal((1, nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((1, nz, ny, nx // 2 + 1)))
    X = X.astype(np.complex64)
    Xd = _dev(torch, X, shape, cuda)
    x = gx.d_IFFTC2R(Xd, nx, ndims=_nd(shape), renormalize=False)
    ref = orc.c2r(X[0].astype(np.complex128), nx)
    err = rel_linf(to_np(x).reshape(ref.shape), ref)
    print(f"C2R {shape}: {err:.2e}")
    assert err < TOL
    assert np.array_equal(to_np(Xd).reshape(X.shape), X)


@pytest.mark.parametrize("shape,r", [((30, 40, 50), 6.0), ((1, 75, 100), 9.0), ((21, 33, 45), 5.0)])
def test_local_std_any_extent_matches_oracle(gx, orc, cuda, shape, r):
    """d_LocalStd (src/Generics/LocalStd.cu:10-118) on extents without a native transform, odd ones
    included, against the FP64 oracle (its transforms are exact DFTs at these sizes)."""
    import torch
    rng = np.random.default_rng(sum(shape))
    m = (3.0 + rng.standard_normal(shape)).astype(np.float32)
    x = torch.from_numpy(m).to(cuda)
    sd, mean = gx.d_LocalStd(x, r, with_mean=True)
    torch.cuda.synchronize()
    rsd, rmean = orc.local_std(m.astype(np.float64), r, with_mean=True)
    err = rel_linf(to_np(sd).reshape(rsd.shape), rsd)
    print(f"local std {shape}: {err:.2e}")
    assert err < 1e-5
    assert rel_linf(to_np(mean).reshape(rmean.shape), rmean) < 1e-5
