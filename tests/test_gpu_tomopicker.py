"""TomoPicker (gtom_amd/csrc/tomopicker.hip) and d_Exact2DWeighting against the
FP64 oracle (oracle/oracle.cpp: exact2d_weighting, tomo_picker).

Tolerances: the weighting filters to 2e-5 relative (float sinc); the weighted
tilt images to 1e-5 normwise; the correlation volume (the maximum over
orientations of the back-projected, locally normalised tilt correlations) and
the sums of that mean and its square to 1e-4 normwise; the orientation index
identical wherever the oracle's best-to-second margin exceeds 1e-4 * max|corr|.
The reference's TomoPicker does not run (src/Correlation/TomoPicker.cu:23)."""
import numpy as np
import pytest

from tests._util import rel_linf, to_np

pytestmark = pytest.mark.gpu
TAU = 2 * np.pi


def _tilts(n, span=50.0):
    return np.array([[0.0, np.deg2rad(t), 0.0] for t in np.linspace(-span, span, n)], np.float32)


def test_exact2d_weighting_matches_oracle(gx, orc, cuda):
    import torch
    nx, ny = 64, 48
    tilts = _tilts(9)
    tilts[:, 0] = np.linspace(0, 0.3, 9)  # some rot as well
    w = np.random.default_rng(3).uniform(0.5, 1.5, 9).astype(np.float32)
    start = np.random.default_rng(4).uniform(0.5, 2.0, (4, ny, nx // 2 + 1)).astype(np.float32)
    ind = [6, 4, 7, 8]  # as in the reference, filter b uses image b's frame and weight whatever the index
    got = torch.from_numpy(start).to(cuda)
    gx.d_Exact2DWeighting(got, tilts, imageweights=torch.from_numpy(w).to(cuda), indices=ind, maxfreq=30.0)
    want = orc.exact2d_weighting(nx, ny, tilts.astype(np.float64), w.astype(np.float64), ind, 30.0,
                                 start=start.astype(np.float64))
    g = to_np(got)
    assert np.all((g == 0) == (want == 0))
    nz = want != 0
    assert np.max(np.abs(g[nz] - want[nz]) / np.abs(want[nz])) <= 2e-5


def _sphere(nr, r):
    g = np.arange(nr) - nr // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return (x * x + y * y + z * z <= r * r).astype(np.float32)


def _ellipsoid(nr):
    g = np.arange(nr) - nr // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    return ((x / 7.0) ** 2 + (y / 5.0) ** 2 + (z / 4.0) ** 2 <= 1).astype(np.float32)


def _blob(nr, seed):
    rng = np.random.default_rng(seed)
    g = np.arange(nr) - nr // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    v = np.zeros((nr, nr, nr))
    for _ in range(5):
        c = rng.uniform(-nr / 5, nr / 5, 3)
        s = rng.uniform(1.2, 2.2, 3)
        v += np.exp(-(((x - c[0]) / s[0]) ** 2 + ((y - c[1]) / s[1]) ** 2 + ((z - c[2]) / s[2]) ** 2) / 2)
    return v.astype(np.float32)


@pytest.mark.parametrize("circular,axis", [(True, "y"), (False, "y"), (True, "rotated"), (False, "rotated"),
                                           (True, "common"), (False, "common-steep")])
def test_tomopicker_matches_oracle(gx, orc, cuda, circular, axis):
    """axis "y": every tilt about y alone (the update reads one pixel row per image); "rotated":
    three of the five tilt axes turned in-plane by 0.1 ... 0.3 rad (the general bilinear form) and
    two about y; "common" / "common-steep": every axis turned by 0.25 / 1.2 rad (one second row for
    all images: lanes along x / along y over sheared rows)."""
    import torch
    nr, nx, ny, nimg = 16, 64, 64, 5
    ref = _blob(nr, 5)
    mask = _sphere(nr, 7) if circular else _ellipsoid(nr)
    rng = np.random.default_rng(6)
    images = rng.standard_normal((nimg, ny, nx)).astype(np.float32)
    tilts = _tilts(nimg, 40.0)
    if axis == "rotated":
        tilts[[0, 2, 3], 0] = [0.1, 0.2, 0.3]
    elif axis == "common":
        tilts[:, 0] = 0.25
    elif axis == "common-steep":
        tilts[:, 0] = 1.2
    w = rng.uniform(0.5, 1.0, nimg).astype(np.float32)
    angles = orc.equal_angular_spacing((0, TAU), (0, np.pi), (0, TAU), np.deg2rad(60.0))[:10].astype(np.float32)
    dc = (32, 24, 16)
    tp = gx.TomoPicker(torch.from_numpy(ref).to(cuda), torch.from_numpy(mask).to(cuda), (ny, nx), nimg,
                       circular=circular)
    img = torch.from_numpy(images).to(cuda)
    tp.set_image(img, tilts, w)
    corr = torch.full((dc[2], dc[1], dc[0]), -1e30, device=cuda)
    ang = torch.zeros(corr.shape + (3,), device=cuda)
    idx = torch.full(corr.shape, -1, dtype=torch.int32, device=cuda)
    s1 = torch.empty_like(corr)
    s2 = torch.empty_like(corr)
    tp.perform_correlation(corr, ang, angles=angles, index=idx, sum1=s1, sum2=s2)
    torch.cuda.synchronize()
    tp.close()
    ref64 = orc.tomo_picker(ref.astype(np.float64), mask.astype(np.float64), circular, images.astype(np.float64),
                            tilts.astype(np.float64), w.astype(np.float64), angles.astype(np.float64), dc)
    assert rel_linf(to_np(img), ref64["images"]) <= 1e-5
    c = to_np(corr)
    err = rel_linf(c, ref64["corr"])
    assert err <= 1e-4, err
    assert rel_linf(to_np(s1), ref64["sum1"]) <= 1e-4
    assert rel_linf(to_np(s2), ref64["sum2"]) <= 1e-4
    sure = (ref64["corr"] - ref64["corr2"]) > 1e-4 * np.abs(ref64["corr"]).max()
    assert sure.sum() > 0.9 * sure.size
    assert np.array_equal(to_np(idx)[sure], ref64["idx"][sure])
    assert np.array_equal(to_np(ang)[sure], angles[ref64["idx"][sure]])


def test_tomopicker_non_native_sizes_match_oracle(gx, orc, cuda):
    """A 10^3 reference (padded projector 20^3) and 50 x 36 tilt images: extents without a native
    transform, which take the chirp-z form; the same tolerances as above."""
    import torch
    nr, nx, ny, nimg = 10, 50, 36, 4
    ref = _blob(nr, 15)
    mask = _sphere(nr, 4)
    rng = np.random.default_rng(16)
    images = rng.standard_normal((nimg, ny, nx)).astype(np.float32)
    tilts = _tilts(nimg, 45.0)
    tilts[:, 0] = 0.2
    w = rng.uniform(0.5, 1.0, nimg).astype(np.float32)
    angles = orc.equal_angular_spacing((0, TAU), (0, np.pi), (0, TAU), np.deg2rad(60.0))[:6].astype(np.float32)
    dc = (30, 20, 12)
    tp = gx.TomoPicker(torch.from_numpy(ref).to(cuda), torch.from_numpy(mask).to(cuda), (ny, nx), nimg)
    img = torch.from_numpy(images).to(cuda)
    tp.set_image(img, tilts, w)
    corr = torch.full((dc[2], dc[1], dc[0]), -1e30, device=cuda)
    idx = torch.full(corr.shape, -1, dtype=torch.int32, device=cuda)
    tp.perform_correlation(corr, angles=angles, index=idx)
    torch.cuda.synchronize()
    tp.close()
    ref64 = orc.tomo_picker(ref.astype(np.float64), mask.astype(np.float64), True, images.astype(np.float64),
                            tilts.astype(np.float64), w.astype(np.float64), angles.astype(np.float64), dc)
    assert rel_linf(to_np(img), ref64["images"]) <= 1e-5
    err = rel_linf(to_np(corr), ref64["corr"])
    print(f"TomoPicker 10^3 / 50 x 36: max error {err:.2e}")
    assert err <= 1e-4, err
    sure = (ref64["corr"] - ref64["corr2"]) > 1e-4 * np.abs(ref64["corr"]).max()
    assert sure.sum() > 0.5 * sure.size
    assert np.array_equal(to_np(idx)[sure], ref64["idx"][sure])


def test_tomopicker_anglestep_finds_planted_projection(gx, orc, cuda):
    """PerformCorrelation(anglestep) on one untilted image holding the reference's
    projection along z (numpy's sum) at a pixel: the voxels above it hold the
    best correlation (near 1) with the identity orientation, GetEqualAngularSpacing's first."""
    import torch
    nr, n = 16, 64
    ref = _blob(nr, 9)
    proj = ref.sum(axis=0)
    img = np.zeros((n, n), np.float32)
    img[:nr, :nr] = proj
    pos = (37, 22)
    img = np.roll(img, (pos[1] - nr // 2, pos[0] - nr // 2), axis=(0, 1))
    img = (img + 0.05 * np.random.default_rng(10).standard_normal((n, n))).astype(np.float32)
    tp = gx.TomoPicker(torch.from_numpy(ref).to(cuda), torch.ones((nr, nr, nr), device=cuda), (n, n), 1)
    tp.set_image(torch.from_numpy(img[None]).to(cuda), np.zeros((1, 3), np.float32), np.ones(1, np.float32))
    corr = torch.full((4, 32, 32), -1e30, device=cuda)
    ang = torch.full(corr.shape + (3,), 9.0, device=cuda)
    tp.perform_correlation(corr, ang, anglestep=float(np.deg2rad(90.0)))
    torch.cuda.synchronize()
    tp.close()
    c, a = to_np(corr), to_np(ang)
    vx, vy = pos[0] + 16 - n // 2, pos[1] + 16 - n // 2
    for z in range(4):
        assert np.unravel_index(np.argmax(c[z]), c[z].shape) == (vy, vx)
        assert c[z, vy, vx] > 0.97
        assert np.array_equal(a[z, vy, vx], np.zeros(3, np.float32))
