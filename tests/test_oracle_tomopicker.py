"""Pins the oracle's TomoPicker restatement (oracle/oracle.cpp: exact2d_weighting,
tomo_picker) without the reference, whose TomoPicker throws on construction
(src/Correlation/TomoPicker.cu:23) and whose projections and correlation call
are commented out (:223-229, :264-269, :297):

- d_Exact2DWeighting (src/Projection/Weighting.cu:92-133) in closed form: one
  untilted image gives w inside the frequency circle and 0 outside; with images
  whose normals are x and y, the untilted image's filter is w0^2 / (w0 + w1 [x =
  0] + w2 [y = 0]) (sinc vanishes at every other integer frequency), which fixes
  which FFT row and column each term lands on; a 30-degree tilt gives the
  fractional sinc(1/2) = 2/pi at x = 1;
- the whole pipeline on one untilted image: a planted projection of the
  reference (numpy's sum along z, not the projector) is found at its image
  position with a masked, locally normalised correlation near 1 (Pearson), and
  every z of the correlation volume above that pixel holds it;
- the search orientation: a quarter turn of that projection about its centre
  (index arithmetic) is picked by psi = 90 degrees and not by 270.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def test_weighting_single_untilted_image_is_w_inside_circle():
    nx, ny = 32, 24
    out = po.exact2d_weighting(nx, ny, np.zeros((1, 3)), weights=np.array([0.7]), maxfreq=9.0)[0]
    fy = np.fft.fftfreq(ny, 1 / ny)[:, None]
    fx = np.arange(nx // 2 + 1)[None, :]
    inside = fx ** 2 + fy ** 2 <= 81
    assert np.allclose(out[inside], 0.7, rtol=0, atol=1e-15)
    assert np.all(out[~inside] == 0)


def test_weighting_row_and_column_terms():
    nx, ny = 32, 32
    w = np.array([1.0, 0.5, 0.25])
    # normals: (0,0,1) untilted; Euler(0, pi/2, 0) row 2 = (1, 0, 0); Euler(pi/2, pi/2, 0) row 2 = (0, 1, 0)
    angles = np.array([[0, 0, 0], [0, np.pi / 2, 0], [np.pi / 2, np.pi / 2, 0]])
    out = po.exact2d_weighting(nx, ny, angles, weights=w, indices=[0], maxfreq=float(nx))[0]
    want = np.full((ny, nx // 2 + 1), w[0])
    want[:, 0] = w[0] ** 2 / (w[0] + w[1])
    want[0, :] = w[0] ** 2 / (w[0] + w[2])
    want[0, 0] = w[0] ** 2 / (w[0] + w[1] + w[2])
    assert np.max(np.abs(out - want)) < 1e-12


def test_weighting_fractional_sinc():
    nx, ny = 16, 16
    angles = np.array([[0, 0, 0], [0, np.pi / 6, 0]])  # second normal (sin 30, 0, cos 30)
    out = po.exact2d_weighting(nx, ny, angles, indices=[0], maxfreq=float(nx))[0]
    s = abs(np.sin(np.pi * 0.5) / (np.pi * 0.5))  # x = 1: dot = -sin(30) * 1
    assert abs(out[3, 1] - 1.0 / (1.0 + s)) < 1e-12  # any row: the second normal has no y part


def _ref(nr, seed):
    rng = np.random.default_rng(seed)
    g = np.arange(nr) - nr // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    v = np.zeros((nr, nr, nr))
    for _ in range(5):
        c = rng.uniform(-nr / 5, nr / 5, 3)
        s = rng.uniform(1.2, 2.2, 3)
        v += np.exp(-(((x - c[0]) / s[0]) ** 2 + ((y - c[1]) / s[1]) ** 2 + ((z - c[2]) / s[2]) ** 2) / 2)
    return v


def _plant(proj, n, pos, noise, seed):
    """Image (n, n) with the centred nr x nr projection's centre at pixel pos (x, y)."""
    nr = proj.shape[0]
    img = np.zeros((n, n))
    img[:nr, :nr] = proj
    img = np.roll(img, (pos[1] - nr // 2, pos[0] - nr // 2), axis=(0, 1))
    return img + noise * np.random.default_rng(seed).standard_normal((n, n))


@pytest.mark.parametrize("pos", [(20, 13), (30, 18)])
def test_untilted_planted_projection_found(pos):
    nr, n = 16, 48
    ref = _ref(nr, 1)
    mask = np.ones((nr, nr, nr))
    proj = ref.sum(axis=0)
    img = _plant(proj, n, pos, 0.05, 2)[None]
    cx, cy, cz = 24, 20, 3
    out = po.tomo_picker(ref, mask, True, img, np.zeros((1, 3)), np.ones(1), np.zeros((1, 3)), (cx, cy, cz))
    c = out["corr"]
    # voxel (x, y, z) samples pixel (x - cx/2 + n/2, y - cy/2 + n/2)
    vx, vy = pos[0] + cx // 2 - n // 2, pos[1] + cy // 2 - n // 2
    if 0 <= vx < cx and 0 <= vy < cy:
        for z in range(cz):
            assert np.unravel_index(np.argmax(c[z]), c[z].shape) == (vy, vx)
            assert c[z, vy, vx] > 0.97
            assert out["idx"][z, vy, vx] == 0
    assert np.allclose(out["images"], img)  # one untilted image: the filter is 1
    assert np.allclose(out["sum1"], np.where(c > -1e29, c, 0))  # one orientation: the sum is the value


def test_quarter_turn_picks_psi_90():
    nr, n = 16, 48
    ref = _ref(nr, 3)
    proj = np.fft.ifftshift(ref.sum(axis=0))
    yy, xx = np.meshgrid(np.arange(nr), np.arange(nr), indexing="ij")
    rot = np.fft.fftshift(proj[xx, (-yy) % nr])  # rot[y, x] = proj[y' = x, x' = -y] about the centre
    img = _plant(rot, n, (24, 24), 0.05, 4)[None]
    angles = np.array([[0, 0, a] for a in (0.0, np.pi / 2, np.pi, 1.5 * np.pi)])
    out = po.tomo_picker(ref, np.ones((nr, nr, nr)), True, img, np.zeros((1, 3)), np.ones(1), angles, (8, 8, 1))
    c = out["corr"][0]
    assert np.unravel_index(np.argmax(c), c.shape) == (4, 4)
    assert out["idx"][0, 4, 4] == 1 and c[4, 4] > 0.97
