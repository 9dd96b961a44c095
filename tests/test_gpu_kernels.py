"""Parity of the HIP building blocks against the FP64 oracle (needs a GPU)."""
import numpy as np
import pytest

from tests._util import rel_linf, to_np

pytestmark = pytest.mark.gpu

FFT_TOL = 2e-6  # normwise relative, FP32 transform vs FP64


@pytest.mark.parametrize("shape", [(8, 8, 8), (16, 16, 16), (8, 16, 32), (32, 32, 32), (64, 64, 64),
                                   (128, 32, 16), (16, 256, 8), (96, 96, 96), (12, 24, 48), (192, 8, 24)])
def test_fft_r2c_matches_oracle(gx, orc, cuda, shape):
    import torch
    rng = np.random.default_rng(hash(shape) % 2**32)
    x = rng.standard_normal((2,) + shape).astype(np.float32)
    X = gx.d_FFTR2C(torch.from_numpy(x).to(cuda))
    ref = orc.r2c(x.astype(np.float64))
    assert rel_linf(to_np(X), ref) < FFT_TOL


@pytest.mark.parametrize("shape", [(1, 32, 32), (1, 256, 256), (1, 1, 64)])
def test_fft_r2c_2d_1d(gx, orc, cuda, shape):
    import torch
    rng = np.random.default_rng(7)
    x = rng.standard_normal(shape).astype(np.float32)
    nd = 2 if shape[1] > 1 else 1
    X = gx.d_FFTR2C(torch.from_numpy(x).to(cuda).reshape(shape[1:]) if nd == 2 else torch.from_numpy(x).to(cuda).reshape(-1), ndims=nd)
    ref = np.fft.rfftn(x.astype(np.float64).reshape(shape[1:] if nd == 2 else (-1,)))
    assert rel_linf(to_np(X), ref) < FFT_TOL


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 16, 8), (64, 64, 64), (8, 128, 256), (96, 48, 96),
                                   (24, 192, 12)])
def test_fft_c2r_nonhermitian_matches_oracle(gx, orc, cuda, shape):
    """Arbitrary (non-Hermitian) half spectra: the result is the real part of the completed inverse."""
    import torch
    nz, ny, nx = shape
    rng = np.random.default_rng(nx * 7 + ny)
    X = (rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1)))
    Xd = torch.from_numpy(X.astype(np.complex64)).to(cuda)
    x = gx.d_IFFTC2R(Xd, nx, renormalize=False)
    ref = orc.c2r(X.astype(np.complex64).astype(np.complex128), nx)
    assert rel_linf(to_np(x), ref) < FFT_TOL
    # the input is left unmodified
    assert np.array_equal(to_np(Xd), X.astype(np.complex64))
    xn = gx.d_IFFTC2R(Xd, nx, renormalize=True)
    assert rel_linf(to_np(xn), ref / (nx * ny * nz)) < FFT_TOL


def test_fft_roundtrip_and_parseval_256(gx, cuda):
    """At the full 256^3 size: C2R_n(R2C(x)) = x and Parseval over the half spectrum."""
    import torch
    g = torch.Generator(device=cuda).manual_seed(3)
    x = torch.randn((256, 256, 256), device=cuda, generator=g)
    X = gx.d_FFTR2C(x)
    y = gx.d_IFFTC2R(X, 256, renormalize=True)
    err = (y - x).abs().max().item() / x.abs().max().item()
    assert err < 2e-6
    w = torch.full((256 // 2 + 1,), 2.0, device=cuda, dtype=torch.float64)
    w[0] = 1.0
    w[-1] = 1.0
    e_freq = ((X.abs().double() ** 2) * w).sum().item() / x.numel()
    e_real = (x.double() ** 2).sum().item()
    assert abs(e_freq - e_real) / e_real < 1e-5


@pytest.mark.parametrize("p,n,s", [(32, 32, 1.0), (64, 32, 2.0), (48, 64, 0.75), (96, 256, 0.375)])
def test_project_matches_oracle(gx, orc, cuda, p, n, s):
    import torch
    rng = np.random.default_rng(p * 1000 + n)
    vol = orc.r2c(rng.standard_normal((p, p, p)))
    angles = rng.uniform(0, 2 * np.pi, (3, 3))
    angles[:, 1] = rng.uniform(0, np.pi, 3)
    mats = gx.projector_matrices(angles.astype(np.float32), s)
    out = gx.d_rlnProject(torch.from_numpy(vol.astype(np.complex64)).to(cuda), p, n,
                          matrices=torch.from_numpy(mats).to(cuda))
    for b in range(3):
        ref = orc.project(vol.astype(np.complex64).astype(np.complex128), p, n, mats[b].astype(np.float64))
        assert rel_linf(to_np(out[b]), ref) < 2e-5, b


def test_ctf_matches_oracle(gx, orc, cuda):
    n = 64
    params = [gx.CtfParams.make(pixelsize=10e-10, Cs=2.7e-3, voltage=300e3, defocus=-3e-6, amplitude=0.07),
              gx.CtfParams.make(pixelsize=10e-10, Cs=2.7e-3, voltage=300e3, defocus=-4.2e-6, amplitude=0.1)]
    c = to_np(gx.ctf_3d(params, n, device=cuda))
    for k, (df, a) in enumerate([(-3e-6, 0.07), (-4.2e-6, 0.1)]):
        ref = orc.ctf3(n, 10e-10, 300e3, 2.7e-3, df, a)
        assert np.abs(c[k] - ref).max() < 2e-6


def test_update_max_argmax_exact(gx, orc, cuda):
    """Strict '>' with the first maximum winning, including planted ties."""
    import torch
    rng = np.random.default_rng(11)
    L, B = 4099, 7
    corr = rng.integers(-50, 50, (B, L)).astype(np.float32)  # many ties
    best = np.full(L, -1e30, np.float32)
    idx = np.zeros(L, np.int32)
    dbest = torch.from_numpy(best.copy()).to(cuda)
    didx = torch.from_numpy(idx.copy()).to(cuda)
    gx.d_UpdateMaxArgmax(torch.from_numpy(corr).to(cuda), dbest, didx, 5)
    rb = best.astype(np.float64)
    orc.update_max_argmax(corr.astype(np.float64), rb, idx, 5)
    assert np.array_equal(to_np(didx), idx)
    assert np.array_equal(to_np(dbest).astype(np.float64), rb)


@pytest.mark.parametrize("n,r", [(32, 6.0), (64, 14.0)])
def test_local_std_matches_oracle(gx, orc, cuda, n, r):
    import torch
    rng = np.random.default_rng(n)
    m = (rng.standard_normal((n, n, n)) * 3 + 1).astype(np.float32)
    sd, mean = gx.d_LocalStd(torch.from_numpy(m).to(cuda), r, with_mean=True)
    rsd, rmean = orc.local_std(m.astype(np.float64), r, with_mean=True)
    assert rel_linf(to_np(sd), rsd) < 1e-5
    assert rel_linf(to_np(mean), rmean) < 1e-5


@pytest.mark.parametrize("n,r,scale", [(32, 6.0, 1.0), (48, 9.0, 0.6)])
def test_local_std_fourier_mask_matches_oracle(gx, orc, cuda, n, r, scale):
    """d_LocalStd with a Fourier mask: an anisotropic low-pass (different width
    per axis) whose DC value scales the mask sum."""
    import torch
    rng = np.random.default_rng(100 + n)
    m = (rng.standard_normal((n, n, n)) * 3 + 1).astype(np.float32)
    f = np.fft.fftfreq(n)
    fz, fy, fx = np.meshgrid(f, f, np.fft.rfftfreq(n), indexing="ij")
    fm = (scale * np.exp(-(fx ** 2 / 0.02 + fy ** 2 / 0.05 + fz ** 2 / 0.11))).astype(np.float32)
    sd, mean = gx.d_LocalStd(torch.from_numpy(m).to(cuda), r, with_mean=True,
                             fouriermask=torch.from_numpy(fm).to(cuda))
    rsd, rmean = orc.local_std(m.astype(np.float64), r, with_mean=True, fouriermask=fm.astype(np.float64))
    assert rel_linf(to_np(sd), rsd) < 1e-5
    assert rel_linf(to_np(mean), rmean) < 1e-5
    # and the mask is applied: without it sigma differs
    assert rel_linf(to_np(gx.d_LocalStd(torch.from_numpy(m).to(cuda), r)), rsd) > 1e-2


def test_pad_remap_match_oracle(gx, orc, cuda):
    import torch
    rng = np.random.default_rng(5)
    t = rng.standard_normal((16, 16, 16)).astype(np.float32)
    pd = to_np(gx.d_Pad(torch.from_numpy(t).to(cuda), (32, 32, 32), 0.0))
    ref = np.zeros((32, 32, 32), np.float32)
    ref[8:24, 8:24, 8:24] = t
    assert np.array_equal(pd, ref)
    rm = to_np(gx.d_RemapFull2FullFFT(torch.from_numpy(ref).to(cuda)))
    assert np.array_equal(rm, orc.remap_full2fullfft(ref.astype(np.float64)).astype(np.float32))
    assert np.array_equal(rm, np.fft.ifftshift(ref))


@pytest.mark.parametrize("shape,radius,sigma,center,dec,cplx", [
    ((16, 18, 20), 6.0, 2.0, None, False, False), ((16, 18, 20), 5.5, 0.0, (7.5, 9.25, 6.75), False, True),
    ((16, 18, 20), 7.0, 3.0, None, True, True), ((1, 24, 30), 9.0, 1.5, (10.5, 12.0, 0.0), False, False)])
def test_sphere_mask_ex_matches_oracle(gx, orc, cuda, shape, radius, sigma, center, dec, cplx):
    import torch
    rng = np.random.default_rng(shape[2])
    x = rng.standard_normal((2,) + shape).astype(np.float32)
    if cplx:
        x = (x + 1j * rng.standard_normal(x.shape)).astype(np.complex64)
    got = to_np(gx.d_SphereMask(torch.from_numpy(x).to(cuda), radius, sigma, center, dec))
    m = orc.sphere_mask_ex(shape, radius, sigma, center, dec)
    assert np.abs(got - x * m[None]).max() <= 1e-6 * np.abs(x).max()
