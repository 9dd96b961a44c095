"""Pins the oracle's ProbabilisticPicker restatement (oracle/oracle.cpp prob_picker,
rotate2d_ft, OrcGaussRand).

- gaussrand against the C runtime's rand() read through ctypes and Box-Muller in
  float32 (src/Helpers/Misc.cpp:14-31), the pair identity z_sin^2 + z_cos^2 =
  -2 ln U, and the phase that carries across srand calls;
- d_Rotate2DFT against numpy's FFT of real images: identity, quarter turns, and
  bilinear interpolation on the full grid for general angles;
- the whole picker at one orientation (identity), 2D and 3D, against a numpy
  program built from numpy's FFTs, frequency-indexed pad/crop and the
  reference's formulas (src/Correlation/ProbabilisticPicker.cu), including the
  zero-frequency reads of RemapHalfFFT2HalfKernel (src/FFT/FFTRemap.cu).
"""
import ctypes
import ctypes.util
import math

import numpy as np
import pytest

from oracle import pyoracle as po

RAND_MAX = 2147483647  # glibc


def _libc():
    return ctypes.CDLL(ctypes.util.find_library("c"))


def _np_gauss(count, seed=123):
    """gaussrand pairs from libc rand() with float32 arithmetic."""
    libc = _libc()
    libc.srand(seed)
    out = []
    twopi = np.float32(2) * np.float32(math.pi)
    while len(out) < count:
        U = np.float32((libc.rand() + 1.0) / (RAND_MAX + 2.0))
        V = np.float32(libc.rand() / (RAND_MAX + 1.0))
        m = np.sqrt(np.float32(-2) * np.log(U))
        out.append(m * np.sin(twopi * V))
        out.append(m * np.cos(twopi * V))
    return np.array(out[:count], np.float32)


# ---------------------------------------------------------------- gaussrand
def test_gaussrand_follows_libc_rand_and_box_muller():
    got = po.gaussrand(2000, seed=123, reset=True)
    want = _np_gauss(2000, 123)
    assert np.allclose(got, want, rtol=2e-6, atol=2e-6)
    # pairs share U: sin^2 + cos^2 = 1 times -2 ln U > 0
    r2 = got[0::2].astype(np.float64) ** 2 + got[1::2].astype(np.float64) ** 2
    assert np.all(r2 > 0)


def test_gaussrand_statistics():
    z = po.gaussrand(200000, seed=7, reset=True).astype(np.float64)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.02


def test_gaussrand_phase_carries_across_srand():
    """An odd number of draws leaves the cos half of a pair pending; after
    srand the next draw is that cos with the old U, V (static state)."""
    a = po.gaussrand(3, seed=123, reset=True)  # sin, cos, sin
    b = po.gaussrand(1, seed=123, reset=False)  # the pending cos
    fresh = _np_gauss(4, 123)
    assert np.isclose(b[0], fresh[3], rtol=2e-6) and np.isclose(a[2], fresh[2], rtol=2e-6)
    po.gaussrand(1, seed=1, reset=True)  # leave the phase at 0


# ---------------------------------------------------------------- d_Rotate2DFT
N = 16


def _centred_half(img):
    n = img.shape[0]
    F = np.fft.fftshift(np.fft.fft2(img))  # index i <-> frequency i - n/2 on both axes
    return np.ascontiguousarray(F[:, :n // 2 + 1]), F


def _coords(n):
    y, x = np.meshgrid(np.arange(n) - n // 2, np.arange(n // 2 + 1) - n // 2, indexing="ij")
    return x.astype(float), y.astype(float)


def test_rotate2d_identity_is_the_disc_masked_input():
    rng = np.random.default_rng(0)
    half, _ = _centred_half(rng.standard_normal((N, N)))
    out = po.rotate2d_ft(half, 0.0, N / 2 - 1.5)
    x, y = _coords(N)
    inside = np.sqrt(x * x + y * y) <= N / 2 - 1.5
    assert np.array_equal(out[inside], half[inside]) and np.all(out[~inside] == 0)


def test_rotate2d_fft_order_output_reads_the_negated_frequency():
    rng = np.random.default_rng(1)
    half = rng.standard_normal((N, N // 2 + 1)) + 1j * rng.standard_normal((N, N // 2 + 1))
    out = po.rotate2d_ft(half, 0.0, N / 2, zerocentred=False)
    s = lambda i: i if i <= N // 2 else i - N
    for ky in range(N):
        for kx in range(N // 2 + 1):
            fx, fy = -kx, -s(ky)
            want = half[fy + N // 2, fx + N // 2] if math.hypot(fx, fy) <= N / 2 else 0
            assert out[ky, kx] == want


@pytest.mark.parametrize("angle", [math.pi / 2, -math.pi / 2, math.pi])
def test_rotate2d_quarter_turns_equal_the_full_fft(angle):
    """Output bin k holds F(R(-angle) k) (positions are integers up to the float
    rounding of cos and sin, which moves the bilinear weights by ~1e-7)."""
    rng = np.random.default_rng(2)
    half, F = _centred_half(rng.standard_normal((N, N)))
    out = po.rotate2d_ft(half, angle, N / 2 - 1)
    x, y = _coords(N)
    c, s = math.cos(angle), math.sin(angle)
    px, py = np.rint(c * x + s * y).astype(int), np.rint(-s * x + c * y).astype(int)
    inside = np.sqrt(x * x + y * y) <= N / 2 - 1
    want = F[py[inside] + N // 2, px[inside] + N // 2]
    assert np.allclose(out[inside], want, atol=1e-5 * np.abs(F).max())


@pytest.mark.parametrize("seed", range(3))
def test_rotate2d_general_angle_is_bilinear_on_the_full_grid(seed):
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(10 + seed)
    half, F = _centred_half(rng.standard_normal((N, N)))
    angle = rng.uniform(-math.pi, math.pi)
    out = po.rotate2d_ft(half, angle, N / 2)
    x, y = _coords(N)
    c, s = np.float32(math.cos(-angle)), np.float32(math.sin(-angle))
    px, py = c * x - s * y, s * x + c * y
    sel = (np.sqrt(x * x + y * y) <= N / 2) & (np.abs(px) < N / 2 - 1.5) & (np.abs(py) < N / 2 - 1.5)
    q = np.stack([py[sel] + N // 2, px[sel] + N // 2])
    want = map_coordinates(F.real, q, order=1) + 1j * map_coordinates(F.imag, q, order=1)
    assert sel.sum() > 80
    assert np.allclose(out[sel], want, atol=1e-9 * np.abs(F).max())


# ---------------------------------------------------------------- the picker, one orientation
def _sf(n):
    """Signed FFT-order frequencies of an axis of length n."""
    return np.where(np.arange(n) < n // 2 + 1, np.arange(n), np.arange(n) - n)


def _band(shape_full, lp):
    """|k| < lp on the numpy half spectrum of a real array of shape_full."""
    axes = [_sf(n) for n in shape_full[:-1]] + [np.arange(shape_full[-1] // 2 + 1)]
    g = np.meshgrid(*axes, indexing="ij")
    return sum(a.astype(np.int64) ** 2 for a in g) < lp * lp


def _np_prob_identity(ref, mask, image, lp, G, ctf=None, plain_shift=False):
    """ProbabilisticPicker at the identity orientation, from numpy primitives;
    G holds the gaussrand draws (five per element of the reference, in order).
    plain_shift replaces the reference's half remap by an fftshift."""
    di, dr = image.shape, ref.shape
    nd = image.ndim
    n = 2 * lp
    nh = n // 2 + 1
    Vi, Er = image.size, ref.size

    def pad(a):
        out = np.zeros(di)
        out[tuple(slice(D // 2 - d // 2, D // 2 - d // 2 + d) for d, D in zip(dr, di))] = a
        return out

    R = np.fft.ifftshift(pad(ref))
    F = np.fft.rfftn(R)
    # FFTCrop to n: keep the frequencies |f| <= n/2 of the full axes, x < n/2+1
    full_idx = [np.array([i if i < n // 2 + 1 else i - n + D for i in range(n)]) for D in di[:-1]]
    C = F[np.ix_(*full_idx, np.arange(nh))]
    C = C * _band((n,) * nd, lp)
    # RemapHalfFFT2HalfKernel: output row i reads row signed(i) + n/2 of the
    # buffer [C, then R's floats as pairs]; for i = n/2 that is row n.
    buf = np.concatenate([C.ravel(), R.ravel().view(np.complex128)[C.size:]])
    rows = (_sf(n) + n // 2)
    if nd == 2:
        j = rows[:, None] * nh + np.arange(nh)[None, :]
    else:
        j = (rows[:, None, None] * n + rows[None, :, None]) * nh + np.arange(nh)[None, None, :]
    Hc = buf[j]
    if plain_shift:
        Hc = np.fft.fftshift(C, axes=tuple(range(nd - 1)))
    # rotation by the identity into FFT order: bin k reads the centred bin -k
    # within the sphere (3D: |k|^2 < n^2/4; 2D: |k| <= n/2)
    ks = [_sf(n)] * (nd - 1) + [np.arange(nh)]
    g = np.meshgrid(*ks, indexing="ij")
    src = tuple(-a + n // 2 for a in g)
    r2 = sum(a.astype(float) ** 2 for a in g)
    keep = r2 < n * n / 4 if nd == 3 else np.sqrt(r2) <= n / 2
    rot = np.where(keep, Hc[src], 0)
    if ctf is not None:
        rot = rot * ctf[np.ix_(*full_idx, np.arange(nh))]
    # FFTPad: frequencies f in [-n/2, n/2) on y/z and [0, n/2) on x keep their place
    P = np.zeros(di[:-1] + (di[-1] // 2 + 1,), complex)
    f = np.arange(-n // 2, n // 2)
    P[np.ix_(*[f % D for D in di[:-1]], np.arange(n // 2))] = rot[np.ix_(*[f % n] * (nd - 1), np.arange(n // 2))]
    M = np.fft.ifftshift(pad(mask))
    r = np.fft.irfftn(P, s=di, axes=tuple(range(nd))) * M
    # image side
    imn = (image - image.mean()) / image.std()
    I = np.fft.rfftn(imn) * _band(di, lp)
    I2 = np.fft.rfftn(imn * imn) * _band(di, lp)
    corr = np.fft.irfftn(I * np.conj(np.fft.rfftn(r)), s=di, axes=tuple(range(nd))) * Vi
    IM = np.fft.rfftn(np.fft.ifftshift(pad(1.0 - mask)))
    # FFTFullCrop of r and M to dims_ref, in the crop's order
    crop = [np.concatenate([np.arange(d // 2), np.arange(D - (d - d // 2), D)]) for d, D in zip(dr, di)]
    refc, maskc = r[np.ix_(*crop)].ravel(), M[np.ix_(*crop)].ravel()
    G = np.asarray(G, np.float64).reshape(Er, 5)
    samples = maskc.sum()
    epr = (refc[:, None] ** 2 + 2.0 * refc[:, None] * G).sum(0)
    smr, smr2 = refc.sum() / samples, (refc ** 2).sum() / samples
    ep = math.exp(epr.sum() / 5.0 / (2.0 * samples))
    fs = 1.0 / (Vi * (Er - samples))
    smean = np.fft.irfftn(I * np.conj(IM), s=di, axes=tuple(range(nd))) * Vi * fs
    sstd = np.sqrt(np.maximum(0, np.fft.irfftn(I2 * np.conj(IM), s=di, axes=tuple(range(nd))) * Vi * fs - smean ** 2))
    nf = 1.0 / (Vi * samples)
    d = (-2.0 * nf * corr + 2.0 * smean * smr) / np.maximum(1e-7, sstd) + smr2
    return (np.exp(-d / 2.0) - 1.0) / (ep - 1.0)


def _disc(shape, radius):
    g = np.meshgrid(*[np.arange(s) - s // 2 for s in shape], indexing="ij")
    return (np.sqrt(sum(a.astype(float) ** 2 for a in g)) < radius).astype(float)


@pytest.mark.parametrize("nd,doctf", [(2, False), (2, True), (3, False)])
def test_picker_at_identity_matches_numpy(nd, doctf):
    rng = np.random.default_rng(40 + nd)
    if nd == 2:
        di, dr, lp, step = (32, 32), (12, 12), 6, 7.0  # ceil(2 pi / 7) = 1 orientation
    else:
        di, dr, lp, step = (16, 16, 16), (8, 8, 8), 4, 4.0  # tilt, rot and psi all 0 only
    ref = rng.standard_normal(dr)
    mask = _disc(dr, dr[0] / 2 - 1)
    image = rng.standard_normal(di) + 0.5
    ctf = None
    if doctf:
        ctf = 0.2 + rng.uniform(0, 1, di[:-1] + (di[-1] // 2 + 1,))
    po.gaussrand(0, reset=True)  # the pair phase starts at 0
    best, ang, bref, _ = po.prob_picker(ref, mask, image, lp, step, ctf=ctf, refid=3)
    # the same draws (pinned against libc rand() above; numpy's float32 log/sin
    # differ from the C library's by an ulp, which moves the score by ~1e-7)
    G = po.gaussrand(5 * ref.size, seed=123, reset=True)
    want = _np_prob_identity(ref, mask, image, lp, G, ctf)
    assert np.allclose(best, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert np.all(bref == 3) and np.all(ang == 0)
    # with numpy's own draws the score agrees to the noise's float rounding
    want_np = _np_prob_identity(ref, mask, image, lp, _np_gauss(5 * ref.size), ctf)
    assert np.abs(best - want_np).max() <= 1e-6 * np.abs(want).max()


def test_picker_zero_frequency_reads_matter():
    """The remap's reads past the cropped spectrum change the score: with a plain
    fftshift in its place the scores differ."""
    rng = np.random.default_rng(47)
    di, dr, lp = (32, 32), (12, 12), 6
    ref = rng.standard_normal(dr)
    mask = _disc(dr, 5.0)
    image = rng.standard_normal(di)
    po.gaussrand(0, reset=True)
    best, _, _, _ = po.prob_picker(ref, mask, image, lp, 7.0)
    G = po.gaussrand(5 * ref.size, seed=123, reset=True)
    quirk = _np_prob_identity(ref, mask, image, lp, G)
    plain = _np_prob_identity(ref, mask, image, lp, G, plain_shift=True)
    assert np.allclose(best, quirk, rtol=1e-12, atol=1e-12 * np.abs(quirk).max())
    assert np.abs(best - plain).max() > 1e-2 * np.abs(plain).max()


# ---------------------------------------------------------------- non-circular masks
@pytest.mark.parametrize("shape", [(40, 36), (24, 30), (20, 18, 22), (32, 30, 34)])
def test_bspline_prefilter_interior_equals_scipy(shape):
    """Away from the ends (where the reference's truncated clamped start sum sits)
    the coefficients equal scipy's exact cubic spline prefilter, and evaluating
    them at the grid points gives the samples back."""
    from scipy.ndimage import spline_filter
    rng = np.random.default_rng(len(shape))
    v = rng.standard_normal(shape)
    c = po.bspline_prefilter(v)
    want = spline_filter(v, order=3, mode="mirror")
    inner = tuple(slice(14, s - 14) for s in shape)
    if all(s > 28 for s in shape):
        assert np.allclose(c[inner], want[inner], atol=1e-7)
    # interpolation at the grid points: (c[i-1] + 4 c[i] + c[i+1]) / 6 along every axis
    back = c.copy()
    for ax in range(len(shape)):
        back = (np.roll(back, 1, ax) + 4 * back + np.roll(back, -1, ax)) / 6
    inner2 = tuple(slice(2, s - 2) for s in shape)
    assert np.allclose(back[inner2], v[inner2], atol=1e-6)
    # a constant stays that constant up to the truncated start sum
    k = po.bspline_prefilter(np.full(shape, 2.5))
    assert np.allclose(k, 2.5, atol=2.5e-6)


def test_rotate_cubic_identity_is_the_fft_shifted_samples():
    rng = np.random.default_rng(3)
    v = np.zeros((16, 16, 16))
    v[4:12, 4:12, 4:12] = rng.standard_normal((8, 8, 8))
    out = po.rotate_cubic(po.bspline_prefilter(v), np.eye(3))
    assert np.allclose(out, np.fft.ifftshift(v), atol=1e-10)  # content within the first 12 samples: exact
    v2 = np.zeros((20, 20))
    v2[5:15, 5:15] = rng.standard_normal((10, 10))
    out2 = po.rotate_cubic(po.bspline_prefilter(v2), np.eye(2))
    # content past the 12-sample start sum: its truncation leaves ~z^13 ~ 4e-8 per sample
    assert np.allclose(out2, np.fft.ifftshift(v2), atol=1e-6)


@pytest.mark.parametrize("seed", range(2))
def test_rotate_cubic_matches_scipy_spline_evaluation(seed):
    """Output voxel i holds the cubic spline of the coefficients at
    M (x - n/2) + n/2, x = (i + n/2) % n (scipy's order-3 map_coordinates of
    the same coefficients without prefiltering), away from the clamped faces."""
    from scipy.ndimage import map_coordinates
    rng = np.random.default_rng(20 + seed)
    n = 24
    coef = rng.standard_normal((n, n, n))
    M = po.projector_matrix(*rng.uniform(-np.pi, np.pi, 3), 1.0)
    out = po.rotate_cubic(coef, M)
    x = (np.arange(n) + n // 2) % n - n // 2
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    c = np.stack([X, Y, Z], -1).astype(float)
    p = c @ M.T + n // 2  # (x, y, z) grid positions
    sel = np.all((p > 2.5) & (p < n - 3.5), axis=-1)
    want = map_coordinates(coef, [p[sel][:, 2], p[sel][:, 1], p[sel][:, 0]], order=3, prefilter=False, mode="nearest")
    assert sel.sum() > 1000
    assert np.allclose(out[sel], want, atol=1e-12)


@pytest.mark.parametrize("nd", [2, 3])
def test_noncircular_mask_at_identity_equals_the_circular_path(nd):
    """At the identity orientation the rotated mask is the FFT-shifted padded
    mask, so the per-orientation inverse mask and solvent statistics equal the
    circular path's, for any mask (an ellipsoid here)."""
    rng = np.random.default_rng(50 + nd)
    if nd == 2:
        di, dr, lp, step = (32, 32), (12, 12), 6, 7.0
    else:
        di, dr, lp, step = (16, 16, 16), (8, 8, 8), 4, 4.0
    g = np.meshgrid(*[np.arange(s) - s // 2 for s in dr], indexing="ij")
    radii = [2.5 + i for i in range(nd)]
    mask = (sum((a / r) ** 2 for a, r in zip(g, radii)) < 1).astype(float)
    ref = rng.standard_normal(dr)
    image = rng.standard_normal(di)
    po.gaussrand(0, reset=True)
    bc, _, _, _ = po.prob_picker(ref, mask, image, lp, step, circular=True)
    po.gaussrand(0, reset=True)
    bn, _, _, _ = po.prob_picker(ref, mask, image, lp, step, circular=False)
    assert np.allclose(bn, bc, rtol=1e-9, atol=1e-9 * np.abs(bc).max())
