"""The Picker class and its two standalone pieces on the GPU against the oracle
(oracle/oracle.cpp: full_pad, norm_mean01std, picker).

Tolerances: FullPad is a copy (bit-exact); mean0/std1 within 1e-6 of the FP64
result; the Picker's best-CC map within 1e-4 normwise of the FP64 oracle
(max|gpu - fp64| / max|fp64|), and the orientation triple identical wherever
the oracle's top-2 margin exceeds 1e-4 * max|best|.
"""
import numpy as np
import pytest

from tests._util import rel_linf, to_np

pytestmark = pytest.mark.gpu

TAU = 2 * np.pi


@pytest.mark.parametrize("old,new", [((8, 8, 8), (16, 16, 16)), ((4, 6, 8), (8, 12, 20)),
                                     ((5, 7, 9), (9, 11, 16)), ((16, 16, 16), (16, 16, 16))])
def test_full_pad_matches_oracle(gx, orc, cuda, old, new):
    import torch
    a = np.random.default_rng(sum(old)).standard_normal((2,) + old).astype(np.float32)
    got = to_np(gx.d_FFTFullPad(torch.from_numpy(a).to(cuda), new))
    for b in range(2):
        assert np.array_equal(got[b], orc.full_pad(a[b], new).astype(np.float32))


def test_full_pad_in_place(gx, orc, cuda):
    import torch
    a = np.random.default_rng(3).standard_normal((8, 8, 8)).astype(np.float32)
    buf = torch.zeros(16 ** 3, dtype=torch.float32, device=cuda)
    buf[:512] = torch.from_numpy(a.ravel()).to(cuda)
    gx.check(gx.lib().gtomx_fft_full_pad(gx._ptr(buf), gx._ptr(buf), gx.int3(8, 8, 8), gx.int3(16, 16, 16), 1,
                                         gx._stream()), "full_pad")
    torch.cuda.synchronize()
    assert np.array_equal(to_np(buf).reshape(16, 16, 16), orc.full_pad(a, (16, 16, 16)).astype(np.float32))


def test_norm_mean01std_matches_oracle(gx, orc, cuda):
    import torch
    x = (np.random.default_rng(5).standard_normal((3, 40000)) * 7 + 3).astype(np.float32)
    x[2] = 1.5  # constant: std 0, only the mean is removed
    got = to_np(gx.d_NormMean01Std(torch.from_numpy(x).to(cuda), batch=3))
    for b in range(3):
        ref = orc.norm_mean01std(x[b].astype(np.float64))
        assert np.max(np.abs(got[b] - ref)) <= 1e-6 * max(1.0, np.max(np.abs(ref)))


def _blob(n, seed):
    rng = np.random.default_rng(seed)
    g = np.arange(n) - n // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    v = np.zeros((n, n, n))
    for _ in range(5):
        c = rng.uniform(-n / 6, n / 6, 3)
        s = rng.uniform(1.0, 2.2, 3)
        v += np.exp(-(((x - c[0]) / s[0]) ** 2 + ((y - c[1]) / s[1]) ** 2 + ((z - c[2]) / s[2]) ** 2) / 2)
    return v.astype(np.float32)


def _run_picker(gx, cuda, ref, image, ctf, angles=None, anglestep=None, batch=0):
    import torch
    pk = gx.Picker(torch.from_numpy(ref).to(cuda), image.shape, batch=batch)
    img = torch.from_numpy(image).to(cuda)
    c = torch.from_numpy(ctf).to(cuda)
    pk.set_image(img, c)
    best = torch.full(image.shape, -1e30, dtype=torch.float32, device=cuda)
    ang = torch.zeros(image.shape + (3,), dtype=torch.float32, device=cuda)
    pk.perform_correlation(best, ang, anglestep=anglestep, angles=angles)
    torch.cuda.synchronize()
    pk.close()
    return to_np(best), to_np(ang), to_np(img)


@pytest.mark.parametrize("shape,batch", [((32, 32, 32), 0), ((24, 48, 32), 5), ((128, 128, 128), 0)])
def test_picker_matches_oracle(gx, orc, cuda, shape, batch):
    nr = 16
    ref = _blob(nr, 11)
    rng = np.random.default_rng(12)
    image = rng.standard_normal(shape).astype(np.float32)
    ctf = rng.uniform(-1, 1, (nr, nr, nr // 2 + 1)).astype(np.float32)
    angles = orc.equal_angular_spacing((0, TAU), (0, np.pi), (0, TAU), np.deg2rad(60.0))[:23]
    best, ang, img_after = _run_picker(gx, cuda, ref, image, ctf, angles=angles, batch=batch)
    ref64 = orc.picker(ref.astype(np.float64), image.astype(np.float64), ctf.astype(np.float64),
                       angles.astype(np.float64))
    err = rel_linf(best, ref64["best"])
    assert err <= 1e-4, err
    margin = ref64["best"] - ref64["top2"]
    sure = margin > 1e-4 * np.max(np.abs(ref64["best"]))
    assert sure.sum() > 0.9 * sure.size
    assert np.array_equal(ang[sure], angles[ref64["idx"][sure]])
    # SetImage leaves the square of the normalised image in the caller's buffer
    norm = orc.norm_mean01std(image.astype(np.float64))
    assert np.max(np.abs(img_after - norm ** 2)) <= 1e-5 * np.max(norm ** 2)


def test_picker_anglestep_finds_planted_copy(gx, orc, cuda):
    """PerformCorrelation(anglestep) over GetEqualAngularSpacing: a copy of the
    reference rotated by one of those orientations and shifted is found at its
    shift with that orientation's triple."""
    nr, n = 16, 32
    ref = _blob(nr, 21)
    angles = orc.equal_angular_spacing((0, TAU), (0, np.pi), (0, TAU), np.deg2rad(60.0))
    k = 17
    F = np.fft.rfftn(np.fft.ifftshift(ref.astype(np.float64)))
    R = orc.project(F, nr, nr, orc.projector_matrix(*angles[k].astype(np.float64), 1.0))
    rot = np.fft.irfftn(R, s=(nr, nr, nr), axes=(0, 1, 2))
    image = np.roll(orc.full_pad(rot, (n, n, n)), (4, -3, 6), axis=(0, 1, 2))
    image = (image + 0.01 * np.random.default_rng(22).standard_normal(image.shape)).astype(np.float32)
    best, ang, _ = _run_picker(gx, cuda, ref, image, np.ones((nr, nr, nr // 2 + 1), np.float32),
                               anglestep=float(np.deg2rad(60.0)))
    pos = np.unravel_index(np.argmax(best), best.shape)
    assert pos == (4, n - 3, 6)
    assert np.array_equal(ang[pos], angles[k])


def _blob2d(n, seed):
    rng = np.random.default_rng(seed)
    g = np.arange(n) - n // 2
    y, x = np.meshgrid(g, g, indexing="ij")
    v = np.zeros((n, n))
    for _ in range(4):
        c = rng.uniform(-n / 6, n / 6, 2)
        s = rng.uniform(1.0, 2.0, 2)
        v += np.exp(-(((x - c[0]) / s[0]) ** 2 + ((y - c[1]) / s[1]) ** 2) / 2)
    return v.astype(np.float32)


@pytest.mark.parametrize("shape,batch", [((64, 48), 0), ((256, 256), 5)])
def test_picker_2d_matches_oracle(gx, orc, cuda, shape, batch):
    """The 2D branch (square reference, image, in-plane orientations;
    src/Correlation/Picker.cu:163-171) against the oracle, same tolerance as 3D."""
    nr = 16
    ref = _blob2d(nr, 13)
    rng = np.random.default_rng(14)
    image = rng.standard_normal(shape).astype(np.float32)
    ctf = rng.uniform(-1, 1, (nr, nr // 2 + 1)).astype(np.float32)
    angles = np.array([[0.0, 0.0, a] for a in np.arange(12) * (TAU / 12)], np.float32)
    best, ang, img_after = _run_picker(gx, cuda, ref, image, ctf, angles=angles, batch=batch)
    ref64 = orc.picker(ref.astype(np.float64), image.astype(np.float64), ctf.astype(np.float64),
                       angles.astype(np.float64))
    err = rel_linf(best, ref64["best"])
    assert err <= 1e-4, err
    sure = (ref64["best"] - ref64["top2"]) > 1e-4 * np.max(np.abs(ref64["best"]))
    assert sure.sum() > 0.9 * sure.size
    assert np.array_equal(ang[sure], angles[ref64["idx"][sure]])
    norm = orc.norm_mean01std(image.astype(np.float64))
    assert np.max(np.abs(img_after - norm ** 2)) <= 1e-5 * np.max(norm ** 2)


def test_picker_2d_anglestep_finds_quarter_turn(gx, orc, cuda):
    """PerformCorrelation(anglestep = 90 degrees) on 2D data runs the in-plane steps
    0, 90, 180, 270 degrees: an exact quarter turn of the reference (index
    arithmetic), planted at a shift, is found there with psi = 90 degrees."""
    nr, n = 16, 64
    ref = _blob2d(nr, 15)
    f = np.fft.ifftshift(ref)
    yy, xx = np.meshgrid(np.arange(nr), np.arange(nr), indexing="ij")
    rot = np.fft.fftshift(f[xx, (-yy) % nr])
    img = np.zeros((n, n), np.float32)
    img[(n - nr) // 2:(n + nr) // 2, (n - nr) // 2:(n + nr) // 2] = rot
    img = np.roll(np.fft.ifftshift(img), (6, -9), axis=(0, 1))
    img = (img + 0.01 * np.random.default_rng(16).standard_normal(img.shape)).astype(np.float32)
    best, ang, _ = _run_picker(gx, cuda, ref, img, np.ones((nr, nr // 2 + 1), np.float32),
                               anglestep=float(np.pi / 2))
    pos = np.unravel_index(np.argmax(best), best.shape)
    assert pos == (6, n - 9)
    assert np.array_equal(ang[pos], np.array([0.0, 0.0, np.float32(2 * np.pi) / 4], np.float32))


def test_picker_non_native_sizes_match_oracle(gx, orc, cuda):
    """A 10^3 reference and a 20 x 30 x 36 image: extents without a native transform take the
    chirp-z form with full (not banded) transforms; the same bounds as above."""
    nr = 10
    ref = _blob(nr, 13)
    rng = np.random.default_rng(14)
    image = rng.standard_normal((20, 30, 36)).astype(np.float32)
    ctf = rng.uniform(-1, 1, (nr, nr, nr // 2 + 1)).astype(np.float32)
    angles = orc.equal_angular_spacing((0, TAU), (0, np.pi), (0, TAU), np.deg2rad(60.0))[:11]
    best, ang, _ = _run_picker(gx, cuda, ref, image, ctf, angles=angles, batch=4)
    ref64 = orc.picker(ref.astype(np.float64), image.astype(np.float64), ctf.astype(np.float64),
                       angles.astype(np.float64))
    err = rel_linf(best, ref64["best"])
    print(f"Picker 10^3 / 20 x 30 x 36: max error {err:.2e}")
    assert err <= 1e-4, err
    sure = ref64["best"] - ref64["top2"] > 1e-4 * np.max(np.abs(ref64["best"]))
    assert sure.sum() > 0.9 * sure.size
    assert np.array_equal(ang[sure], angles[ref64["idx"][sure]])
