"""Pins the oracle's d_Rotate3DFT restatement (oracle/oracle.cpp rotate3d_ft).

The reference ships no test for it. Checks: the identity is the sphere-masked
input (zero-centred output) or its read at the negated frequency (FFT-order
output); for a real volume, the rotated zero-centred half spectrum equals
numpy's full FFT evaluated at M k, exactly for quarter turns (grid points) and
to interpolation accuracy through scipy's trilinear map_coordinates on the full
grid for general rotations, which pins the layout, the rotation direction, and
the flip to the Hermitian mate with conjugation."""
import numpy as np
import pytest

from oracle import pyoracle as po

N = 16


def _centred_half(vol):
    """Zero-centred half layout: element (tz, ty, tx) = F(tx - n/2, ty - n/2, tz - n/2)."""
    n = vol.shape[0]
    F = np.fft.fftshift(np.fft.fftn(vol))  # index i <-> frequency i - n/2 on every axis
    return np.ascontiguousarray(F[:, :, :n // 2 + 1]), F


def _centred_coords(n):
    g = np.arange(n) - n // 2
    gx = np.arange(n // 2 + 1) - n // 2
    z, y, x = np.meshgrid(g, g, gx, indexing="ij")
    return np.stack([x, y, z], axis=-1).astype(np.float64)


def test_identity_is_the_sphere_masked_input():
    rng = np.random.default_rng(0)
    vol = rng.standard_normal((N, N, N))
    half, _ = _centred_half(vol)
    out = po.rotate3d_ft(half, np.eye(3))
    c = _centred_coords(N)
    inside = (c ** 2).sum(-1) < N * N / 4
    assert np.array_equal(out[inside], half[inside]) and np.all(out[~inside] == 0)


def test_identity_fft_order_output_reads_the_negated_frequency():
    rng = np.random.default_rng(1)
    half = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    out = po.rotate3d_ft(half, np.eye(3), zerocentred=False)
    s = lambda i: i if i <= N // 2 else i - N
    for kz in range(N):
        for ky in range(N):
            for kx in range(N // 2 + 1):
                fx, fy, fz = -kx, -s(ky), -s(kz)
                want = half[fz + N // 2, fy + N // 2, fx + N // 2] if fx * fx + fy * fy + fz * fz < N * N / 4 else 0
                assert out[kz, ky, kx] == want


@pytest.mark.parametrize("angles", [(np.pi / 2, 0, 0), (0, np.pi / 2, 0), (0, 0, -np.pi / 2),
                                    (np.pi / 2, np.pi / 2, 0), (np.pi, np.pi / 2, np.pi / 2)])
def test_quarter_turns_equal_the_full_fft_at_rotated_points(angles):
    rng = np.random.default_rng(2)
    vol = rng.standard_normal((N, N, N))
    half, F = _centred_half(vol)
    M = po.projector_matrix(*angles, 1.0)
    out = po.rotate3d_ft(half, M)
    c = _centred_coords(N)
    p = np.rint(c @ M.T).astype(int)  # integer for quarter turns
    inside = (c ** 2).sum(-1) < N * N / 4
    idx = p[inside] + N // 2
    want = F[idx[:, 2], idx[:, 1], idx[:, 0]]
    assert np.allclose(out[inside], want, atol=1e-9)


@pytest.mark.parametrize("seed", range(3))
def test_general_rotation_matches_trilinear_on_the_full_grid(seed):
    """Away from the sphere's edge all 8 corners exist on the full grid, where
    interpolating at p equals interpolating the Hermitian mate at -p and conjugating."""
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(10 + seed)
    vol = rng.standard_normal((N, N, N))
    half, F = _centred_half(vol)
    angles = rng.uniform(-np.pi, np.pi, 3)
    M = po.projector_matrix(*angles, 1.0)
    out = po.rotate3d_ft(half, M)
    c = _centred_coords(N)
    p = c @ M.T
    sel = ((c ** 2).sum(-1) < N * N / 4) & (np.abs(p) < N / 2 - 1.5).all(-1)
    q = (p[sel] + N // 2).T[::-1]  # (z, y, x) grid coordinates
    want = map_coordinates(F.real, q, order=1) + 1j * map_coordinates(F.imag, q, order=1)
    assert sel.sum() > 500
    assert np.allclose(out[sel], want, atol=1e-9 * np.abs(F).max())


def test_real_valued_spectrum_has_no_conjugation():
    rng = np.random.default_rng(5)
    vol = rng.standard_normal((N, N, N))
    half, _ = _centred_half(vol)
    M = po.projector_matrix(0.3, 1.1, -0.4, 1.0)
    c = po.rotate3d_ft(half, M)
    r = po.rotate3d_ft(np.ascontiguousarray(np.abs(half)), M)
    i_re = po.rotate3d_ft(np.ascontiguousarray(half.real), M)
    assert np.allclose(i_re, c.real, atol=1e-12)  # linear in the real part
    assert np.all(r >= -1e-12)  # interpolating a non-negative field, no sign flip


# ---------------------------------------------------------------- cubic interpolation
def test_cubic_identity_reproduces_the_input_inside():
    """Cubic B-spline interpolation of the prefiltered spectrum at the grid points
    gives the samples back, away from the array faces (where the prefilter's
    truncated start sum and the clamped taps sit)."""
    rng = np.random.default_rng(30)
    half = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    out = po.rotate3d_ft(half, np.eye(3), cubic=True)
    c = _centred_coords(N)
    t = c + N // 2
    inside = ((c ** 2).sum(-1) < N * N / 4) & np.all(t >= 2, -1) & (t[..., 1] <= N - 3) & (t[..., 2] <= N - 3) & \
        (t[..., 0] <= N // 2 - 2)
    assert inside.sum() > 300
    assert np.allclose(out[inside], half[inside], atol=1e-6)


@pytest.mark.parametrize("seed", range(2))
def test_cubic_rotation_equals_scipy_spline_of_the_coefficients(seed):
    """Bins whose rotated position needs no flip hold scipy's order-3 spline of the
    prefiltered Re and Im at that texel (the prefilter is pinned separately)."""
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(40 + seed)
    half = rng.standard_normal((N, N, N // 2 + 1)) + 1j * rng.standard_normal((N, N, N // 2 + 1))
    M = po.projector_matrix(*rng.uniform(-np.pi, np.pi, 3), 1.0)
    out = po.rotate3d_ft(half, M, cubic=True)
    cre, cim = po.bspline_prefilter(half.real), po.bspline_prefilter(half.imag)
    c = _centred_coords(N)
    p = c @ M.T
    t = p + N // 2
    sel = ((c ** 2).sum(-1) < N * N / 4) & (p[..., 0] < -1e-3) & np.all(t > 2.5, -1) & (t[..., 0] < N // 2 - 2.5) & \
        (t[..., 1] < N - 3.5) & (t[..., 2] < N - 3.5)
    q = [t[sel][:, 2], t[sel][:, 1], t[sel][:, 0]]
    want = map_coordinates(cre, q, order=3, prefilter=False) + 1j * map_coordinates(cim, q, order=3, prefilter=False)
    assert sel.sum() > 100
    assert np.allclose(out[sel], want, atol=1e-12)


def test_cubic_2d_identity_reproduces_the_input_inside():
    rng = np.random.default_rng(31)
    n = 24
    half = rng.standard_normal((n, n // 2 + 1)) + 1j * rng.standard_normal((n, n // 2 + 1))
    out = po.rotate2d_ft(half, 0.0, n / 2, cubic=True)
    y, x = np.meshgrid(np.arange(n), np.arange(n // 2 + 1), indexing="ij")
    inside = (np.hypot(x - n // 2, y - n // 2) <= n / 2) & (x >= 2) & (x <= n // 2 - 2) & (y >= 2) & (y <= n - 3)
    assert np.allclose(out[inside], half[inside], atol=1e-6)
