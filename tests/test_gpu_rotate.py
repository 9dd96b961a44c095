"""d_Rotate3DFT (linear and cubic) and d_Rotate2DFT (cubic) on the GPU against the FP64 oracle: complex and real
half spectra, zero-centred and FFT-order output, within 1e-5 of the input's
largest magnitude (float positions and weights). Bins whose frequency lies
exactly on the radius n/2 are left out: the rotated radius equals n/2 up to
rounding there, so the '>=' cut keeps or zeroes them by the last bit (in the
reference's float arithmetic too)."""
import numpy as np
import pytest

from tests._util import to_np

pytestmark = pytest.mark.gpu


def _off_radius(n, zerocentred):
    s = lambda i: np.where(i <= n // 2, i, i - n)
    g, gx = np.arange(n), np.arange(n // 2 + 1)
    z, y, x = np.meshgrid(g, g, gx, indexing="ij")
    c = (x - n // 2, y - n // 2, z - n // 2) if zerocentred else (-x, -s(y), -s(z))
    return c[0] ** 2 + c[1] ** 2 + c[2] ** 2 != n * n // 4


@pytest.mark.parametrize("n", [16, 32, 64])
@pytest.mark.parametrize("zerocentred", [True, False])
def test_rotate3d_ft_matches_oracle(gx, orc, cuda, n, zerocentred):
    import torch
    rng = np.random.default_rng(n)
    vol = rng.standard_normal((n, n, n))
    half = np.fft.fftshift(np.fft.fftn(vol))[:, :, :n // 2 + 1].astype(np.complex64)
    angles = np.concatenate([[[0, 0, 0], [np.pi / 2, 0, 0]], rng.uniform(-np.pi, np.pi, (4, 3))]).astype(np.float32)
    got = to_np(gx.d_Rotate3DFT(torch.from_numpy(half).to(cuda), angles, zerocentred))
    scale = np.abs(half).max()
    keep = _off_radius(n, zerocentred)
    for k, a in enumerate(angles):
        ref = orc.rotate3d_ft(half, orc.projector_matrix(*a.astype(np.float64), 1.0), zerocentred)
        assert np.abs(got[k] - ref)[keep].max() <= 1e-5 * scale, k
    # real-valued spectra (e.g. CTFs or masks in Fourier space)
    amp = np.ascontiguousarray(np.abs(half))
    gr = to_np(gx.d_Rotate3DFT(torch.from_numpy(amp).to(cuda), angles[2:4], zerocentred))
    for k, a in enumerate(angles[2:4]):
        ref = orc.rotate3d_ft(amp, orc.projector_matrix(*a.astype(np.float64), 1.0), zerocentred)
        assert np.abs(gr[k] - ref)[keep].max() <= 1e-5 * scale


@pytest.mark.parametrize("n", [16, 32])
@pytest.mark.parametrize("zerocentred", [True, False])
def test_rotate3d_ft_cubic_matches_oracle(gx, orc, cuda, n, zerocentred):
    """T_INTERP_CUBIC: the cubic B-spline of the prefiltered spectrum (64 taps), within
    2e-5 of the input's largest magnitude (the coefficients exceed the samples)."""
    import torch
    rng = np.random.default_rng(100 + n)
    vol = rng.standard_normal((n, n, n))
    half = np.fft.fftshift(np.fft.fftn(vol))[:, :, :n // 2 + 1].astype(np.complex64)
    angles = np.concatenate([[[0, 0, 0]], rng.uniform(-np.pi, np.pi, (3, 3))]).astype(np.float32)
    got = to_np(gx.d_Rotate3DFT(torch.from_numpy(half).to(cuda), angles, zerocentred, mode=gx.INTERP_CUBIC))
    scale = np.abs(half).max()
    for k, a in enumerate(angles):
        ref = orc.rotate3d_ft(half.astype(np.complex128), orc.projector_matrix(*a.astype(np.float64), 1.0),
                              zerocentred, cubic=True)
        assert np.abs(got[k] - ref).max() <= 2e-5 * scale, k
    amp = np.ascontiguousarray(np.abs(half))
    gr = to_np(gx.d_Rotate3DFT(torch.from_numpy(amp).to(cuda), angles[1:3], zerocentred, mode=gx.INTERP_CUBIC))
    for k, a in enumerate(angles[1:3]):
        ref = orc.rotate3d_ft(amp.astype(np.float64), orc.projector_matrix(*a.astype(np.float64), 1.0), zerocentred,
                              cubic=True)
        assert np.abs(gr[k] - ref).max() <= 2e-5 * scale


@pytest.mark.parametrize("zerocentred", [True, False])
def test_rotate2d_ft_cubic_matches_oracle(gx, orc, cuda, zerocentred):
    import torch
    rng = np.random.default_rng(7 + zerocentred)
    n, nb = 32, 3
    half = (rng.standard_normal((nb, n, n // 2 + 1)) + 1j * rng.standard_normal((nb, n, n // 2 + 1))).astype(np.complex64)
    angles = rng.uniform(-np.pi, np.pi, nb).astype(np.float32)
    got = to_np(gx.d_Rotate2DFT(torch.from_numpy(half).to(cuda), angles, n / 2 - 0.5, zerocentred,
                                mode=gx.INTERP_CUBIC))
    for b in range(nb):
        ref = orc.rotate2d_ft(half[b].astype(np.complex128), float(angles[b]), n / 2 - 0.5, zerocentred, cubic=True)
        assert np.abs(got[b] - ref).max() <= 2e-5 * np.abs(half[b]).max()
