"""Pins the oracle's R2C / C2R against numpy's pocketfft and closed forms."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "fft_golden.npz")


def test_oracle_matches_committed_golden_vectors(orc):
    g = np.load(GOLDEN)
    for k, (nz, ny, nx) in enumerate(g["shapes"]):
        X = orc.r2c(g[f"r2c_in_{k}"])
        ref = g[f"r2c_out_{k}"]
        assert np.abs(X - ref).max() <= 1e-12 * np.abs(ref).max()
        x = orc.c2r(g[f"c2r_in_{k}"], int(nx))
        ref = g[f"c2r_out_{k}"]
        assert np.abs(x - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("shape", [(16, 16, 16), (32, 8, 4), (12, 24, 48), (3, 5, 6)])
def test_oracle_matches_numpy_live(orc, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    assert np.allclose(orc.r2c(x), np.fft.rfftn(x), rtol=0, atol=1e-11)
    X = rng.standard_normal(shape[:-1] + (shape[-1] // 2 + 1,)) * (1 + 1j)
    assert np.allclose(orc.c2r(X, shape[-1]), np.fft.irfftn(X, s=shape, axes=(0, 1, 2)) * np.prod(shape),
                       rtol=0, atol=1e-10)


def test_closed_forms(orc):
    n = 16
    # delta at the origin -> constant spectrum
    d = np.zeros((n, n, n))
    d[0, 0, 0] = 1.0
    assert np.allclose(orc.r2c(d), 1.0)
    # a cosine along x with frequency f -> two bins of n^3/2 (one in the half spectrum)
    f = 3
    x = np.cos(2 * np.pi * f * np.arange(n) / n)[None, None, :] * np.ones((n, n, n))
    X = orc.r2c(x)
    expect = np.zeros_like(X)
    expect[0, 0, f] = n ** 3 / 2
    assert np.allclose(X, expect, atol=1e-9)
    # a shifted delta -> linear phase exp(-2 pi i k.r / n)
    r = (2, 5, 7)  # (x, y, z)
    d = np.zeros((n, n, n))
    d[r[2], r[1], r[0]] = 1.0
    kz, ky, kx = np.meshgrid(np.arange(n), np.arange(n), np.arange(n // 2 + 1), indexing="ij")
    phase = np.exp(-2j * np.pi * (kx * r[0] + ky * r[1] + kz * r[2]) / n)
    assert np.allclose(orc.r2c(d), phase, atol=1e-12)


def test_parseval_and_roundtrip(orc):
    rng = np.random.default_rng(1)
    n = 16
    x = rng.standard_normal((n, n, n))
    X = orc.r2c(x)
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    assert np.isclose((np.abs(X) ** 2 * w).sum() / n ** 3, (x ** 2).sum(), rtol=1e-12)
    assert np.allclose(orc.c2r(X, n) / n ** 3, x, atol=1e-13)


def test_c2r_is_real_part_of_completed_inverse(orc):
    """A non-Hermitian x=0 plane: C2R_u equals Re(IFFT3(completed spectrum)) * N."""
    rng = np.random.default_rng(2)
    n = 8
    X = rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))
    full = np.zeros((n, n, n), complex)
    full[:, :, : n // 2 + 1] = X
    for kx in range(n // 2 + 1, n):
        full[:, :, kx] = np.conj(X[(-np.arange(n)) % n][:, (-np.arange(n)) % n][:, :, n - kx])
    ref = np.real(np.fft.ifftn(full)) * n ** 3
    assert np.allclose(orc.c2r(X, n), ref, atol=1e-11)
