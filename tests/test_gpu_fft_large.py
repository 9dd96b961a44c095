"""Generic R2C / C2R transforms with axes above 256 (two-pass, four-step
decomposition): 2^a up to 4096 and 3 * 2^a up to 3072, on every axis, against
numpy's FP64 pocketfft (R2C) and the oracle's real-part C2R (non-Hermitian
half spectra). Tolerance 4e-6 normwise: the FP32 error grows with log2 of the
transform size, and these reach 2^24 points."""
import numpy as np
import pytest

from tests._util import rel_linf, to_np

pytestmark = pytest.mark.gpu

TOL = 4e-6

SHAPES = [  # (nz, ny, nx)
    (1, 512, 512),
    (1, 384, 768),
    (1, 24, 3072),
    (1, 1536, 48),
    (1, 1024, 2048),
    (768, 8, 12),
    (16, 512, 1024),
    (1, 1, 4096),
]


def _nd(shape):
    nz, ny, nx = shape
    return 3 if nz > 1 else (2 if ny > 1 else 1)


def _as_tensor(torch, x, shape, cuda):
    """x is (batch, nz, ny, nx); the transform's own rank drops the unit axes."""
    nz, ny, nx = shape
    keep = {3: (nz, ny, nx), 2: (ny, nx), 1: (nx,)}[_nd(shape)]
    return torch.from_numpy(x).to(cuda).reshape(x.shape[:1] + keep)


@pytest.mark.parametrize("shape", SHAPES)
def test_fft_r2c_large_axes(gx, cuda, shape):
    import torch
    nz, ny, nx = shape
    rng = np.random.default_rng(nx + 3 * ny + 7 * nz)
    x = rng.standard_normal((2, nz, ny, nx)).astype(np.float32)
    nd = _nd(shape)
    X = gx.d_FFTR2C(_as_tensor(torch, x, shape, cuda), ndims=nd)
    ref = np.fft.rfftn(x.astype(np.float64), axes=(1, 2, 3))
    assert rel_linf(to_np(X).reshape(ref.shape), ref) < TOL


@pytest.mark.parametrize("shape", [(1, 512, 1024), (768, 8, 16), (1, 8, 3072), (1, 2048, 512)])
def test_fft_c2r_large_axes_nonhermitian(gx, orc, cuda, shape):
    """Arbitrary half spectra: the real part of the Hermitian-completed inverse,
    which reads only the real parts of the x = 0 and x = nx/2 bins."""
    import torch
    nz, ny, nx = shape
    rng = np.random.default_rng(nx * 5 + ny)
    X = (rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1)))
    X = X.astype(np.complex64)
    Xd = torch.from_numpy(X).to(cuda)
    if nz == 1:
        Xd = Xd.reshape(ny, nx // 2 + 1)
    x = gx.d_IFFTC2R(Xd, nx, ndims=3 if nz > 1 else 2, renormalize=False)
    ref = orc.c2r(X.astype(np.complex128), nx)
    assert rel_linf(to_np(x).reshape(ref.shape), ref) < TOL
    assert np.array_equal(to_np(Xd).reshape(X.shape), X)  # the input is left unmodified


def test_fft_roundtrip_and_parseval_4096(gx, cuda):
    """A 4096^2 micrograph: C2R_n(R2C(x)) = x and Parseval over the half spectrum."""
    import torch
    g = torch.Generator(device=cuda).manual_seed(5)
    x = torch.randn((4096, 4096), device=cuda, generator=g)
    X = gx.d_FFTR2C(x, ndims=2)
    y = gx.d_IFFTC2R(X, 4096, ndims=2, renormalize=True)
    assert (y - x).abs().max().item() / x.abs().max().item() < 4e-6
    w = torch.full((4096 // 2 + 1,), 2.0, device=cuda, dtype=torch.float64)
    w[0] = 1.0
    w[-1] = 1.0
    e_freq = ((X.abs().double() ** 2) * w).sum().item() / x.numel()
    e_real = (x.double() ** 2).sum().item()
    assert abs(e_freq - e_real) / e_real < 1e-5


def test_fft_batched_3072_planes_match_numpy(gx, cuda):
    """A batch of 3072 x 3072 images (3 * 2^10 on both axes)."""
    import torch
    rng = np.random.default_rng(11)
    x = rng.standard_normal((2, 3072, 3072)).astype(np.float32)
    X = gx.d_FFTR2C(torch.from_numpy(x).to(cuda), ndims=2)
    ref = np.fft.rfft2(x.astype(np.float64))
    assert rel_linf(to_np(X), ref) < TOL


@pytest.mark.parametrize("shape,r", [((1, 1024, 1024), 24.0), ((1, 1024, 2048), 40.0)])
def test_local_std_micrograph_matches_oracle(gx, orc, cuda, shape, r):
    """d_LocalStd on micrograph-sized images (its transforms take the two-pass
    and LDS-row paths), against the FP64 oracle at 1e-5 normwise."""
    import torch
    rng = np.random.default_rng(shape[2] + int(r))
    m = (rng.standard_normal(shape) * 3 + 1).astype(np.float32)
    sd, mean = gx.d_LocalStd(torch.from_numpy(m).to(cuda), r, with_mean=True)
    rsd, rmean = orc.local_std(m.astype(np.float64), r, with_mean=True)
    assert rel_linf(to_np(sd), rsd) < 1e-5
    assert rel_linf(to_np(mean), rmean) < 1e-5
