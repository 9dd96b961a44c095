"""Pins the oracle's Picker restatement (oracle/oracle.cpp: full_pad,
norm_mean01std, picker) without the reference, which has no runnable Picker:
closed forms for the padding and the normalisation, an independent numpy
composition of one orientation, and planted rotated copies found at their
shift and orientation."""
import numpy as np
import pytest

from oracle import pyoracle as po


def _numpy_full_pad(a, newshape):
    """Even sizes: centre (fftshift), zero-pad symmetrically, back to FFT order."""
    c = np.fft.fftshift(a)
    pad = [((n - o) // 2, (n - o) - (n - o) // 2) for o, n in zip(a.shape, newshape)]
    return np.fft.ifftshift(np.pad(c, pad))


@pytest.mark.parametrize("old,new", [((8, 8, 8), (16, 16, 16)), ((4, 6, 8), (8, 12, 20)), ((8, 8, 8), (8, 8, 8))])
def test_full_pad_matches_centred_padding(old, new):
    a = np.random.default_rng(1).standard_normal(old)
    assert np.array_equal(po.full_pad(a, new), _numpy_full_pad(a, new))


def test_full_pad_keeps_origin_and_negative_offsets():
    a = np.zeros((8, 8, 8))
    a[0, 0, 0] = 1.0
    a[7, 0, 1] = 2.0  # offset (x, y, z) = (1, 0, -1)
    out = po.full_pad(a, (16, 16, 16))
    assert out[0, 0, 0] == 1.0 and out[15, 0, 1] == 2.0 and out.sum() == 3.0


def test_norm_mean01std_is_population_standardisation():
    x = np.random.default_rng(2).standard_normal(1000) * 3 + 5
    y = po.norm_mean01std(x)
    assert np.allclose(y, (x - x.mean()) / x.std(), rtol=0, atol=1e-12)
    assert abs(y.mean()) < 1e-12 and abs(y.std() - 1) < 1e-12
    c = np.full(10, 4.0)
    assert np.array_equal(po.norm_mean01std(c), np.zeros(10))  # std 0: mean removed only


def _blob(n, seed):
    rng = np.random.default_rng(seed)
    g = np.arange(n) - n // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    v = np.zeros((n, n, n))
    for _ in range(4):
        c = rng.uniform(-n / 6, n / 6, 3)
        s = rng.uniform(1.0, 2.0, 3)
        v += np.exp(-(((x - c[0]) / s[0]) ** 2 + ((y - c[1]) / s[1]) ** 2 + ((z - c[2]) / s[2]) ** 2) / 2)
    return v


def test_identity_orientation_matches_numpy_composition():
    nr, n = 16, 32
    ref = _blob(nr, 3)
    image = np.random.default_rng(4).standard_normal((n, n, n))
    ctf = np.random.default_rng(5).uniform(0.5, 1.0, (nr, nr, nr // 2 + 1))
    out = po.picker(ref, image, ctf, np.zeros((1, 3)))
    # numpy: project with the identity = the stored spectrum inside the sphere
    F = np.fft.rfftn(np.fft.ifftshift(ref))
    k = np.fft.fftfreq(nr, 1 / nr)
    kz, ky, kx = np.meshgrid(k, k, np.arange(nr // 2 + 1), indexing="ij")
    R = F * (kx ** 2 + ky ** 2 + kz ** 2 <= (nr // 2) ** 2) * ctf
    r = np.fft.irfftn(R, s=(nr, nr, nr), axes=(0, 1, 2))
    p = _numpy_full_pad(r, (n, n, n))
    p = (p - p.mean()) / p.std()
    im = (image - image.mean()) / image.std()
    C = np.fft.irfftn(np.fft.rfftn(im) * np.conj(np.fft.rfftn(p)), s=(n, n, n), axes=(0, 1, 2)) * n ** 3
    assert np.max(np.abs(out["best"] - C)) <= 1e-9 * np.max(np.abs(C))


@pytest.mark.parametrize("k", [0, 5])
def test_planted_rotated_copy_found_at_shift_and_orientation(k):
    """The image holds the reference rotated by angle k (made by the same
    projector at the reference size) and shifted by s: the best CC is at s with index k."""
    nr, n = 16, 32
    ref = _blob(nr, 7)
    tau = 2 * np.pi
    angles = po.equal_angular_spacing((0, tau), (0, np.pi), (0, tau), np.deg2rad(45.0))[:12].astype(np.float64)
    F = np.fft.rfftn(np.fft.ifftshift(ref))
    M = po.projector_matrix(*angles[k], 1.0)
    R = po.project(F, nr, nr, M)
    rot = np.fft.irfftn(R, s=(nr, nr, nr), axes=(0, 1, 2))
    image = np.roll(_numpy_full_pad(rot, (n, n, n)), (3, -2, 5), axis=(0, 1, 2))
    image += 0.01 * np.random.default_rng(8).standard_normal(image.shape)
    out = po.picker(ref, image, np.ones((nr, nr, nr // 2 + 1)), angles)
    pos = np.unravel_index(np.argmax(out["best"]), out["best"].shape)
    assert pos == (3, n - 2, 5)
    assert out["idx"][pos] == k


def _blob2d(n, seed):
    rng = np.random.default_rng(seed)
    g = np.arange(n) - n // 2
    y, x = np.meshgrid(g, g, indexing="ij")
    v = np.zeros((n, n))
    for _ in range(4):
        c = rng.uniform(-n / 6, n / 6, 2)
        s = rng.uniform(1.0, 2.0, 2)
        v += np.exp(-(((x - c[0]) / s[0]) ** 2 + ((y - c[1]) / s[1]) ** 2) / 2)
    return v


def test_2d_identity_orientation_matches_numpy_composition():
    """2D branch (square reference, image): the identity in-plane orientation is the
    circle-masked reference spectrum through the same composition in numpy."""
    nr, n = 16, 48
    ref = _blob2d(nr, 3)
    image = np.random.default_rng(4).standard_normal((n, n))
    ctf = np.random.default_rng(5).uniform(0.5, 1.0, (nr, nr // 2 + 1))
    out = po.picker(ref, image, ctf, np.zeros((1, 3)))
    F = np.fft.rfft2(np.fft.ifftshift(ref))
    k = np.fft.fftfreq(nr, 1 / nr)
    ky, kx = np.meshgrid(k, np.arange(nr // 2 + 1), indexing="ij")
    r = np.fft.irfft2(F * (kx ** 2 + ky ** 2 <= (nr // 2) ** 2) * ctf, s=(nr, nr))
    p = np.fft.ifftshift(np.pad(np.fft.fftshift(r), (n - nr) // 2))
    p = (p - p.mean()) / p.std()
    im = (image - image.mean()) / image.std()
    C = np.fft.irfft2(np.fft.rfft2(im) * np.conj(np.fft.rfft2(p)), s=(n, n)) * n * n
    assert out["best"].shape == (n, n)
    assert np.max(np.abs(out["best"] - C)) <= 1e-9 * np.max(np.abs(C))


def test_2d_quarter_turn_picks_the_rotated_copy():
    """An exact real-space quarter turn of the reference about its centre, built by
    index arithmetic (no projector), planted at a shift: of the in-plane steps
    0, 90, 180 and 270 degrees the oracle picks 90 (psi rotates the reference by
    +psi: r(x, y) = f(-y, x) about the origin), at the shift."""
    nr, n = 16, 48
    ref = _blob2d(nr, 9)
    f = np.fft.ifftshift(ref)                     # origin at index 0
    yy, xx = np.meshgrid(np.arange(nr), np.arange(nr), indexing="ij")
    rot = f[xx, (-yy) % nr]                       # rot[y, x] = f[y' = x, x' = -y]
    rot = np.fft.fftshift(rot)
    img = np.zeros((n, n))
    img[(n - nr) // 2:(n + nr) // 2, (n - nr) // 2:(n + nr) // 2] = rot
    img = np.fft.ifftshift(img)
    img = np.roll(img, (5, -7), axis=(0, 1)) + 0.01 * np.random.default_rng(10).standard_normal((n, n))
    angles = np.array([[0, 0, a] for a in (0.0, np.pi / 2, np.pi, 1.5 * np.pi)])
    out = po.picker(ref, img, np.ones((nr, nr // 2 + 1)), angles)
    b = out["best"]
    pos = np.unravel_index(np.argmax(b), b.shape)
    assert pos == (5, n - 7)
    assert out["idx"][pos] == 1
