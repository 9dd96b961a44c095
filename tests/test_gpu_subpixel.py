"""Sub-voxel peak location on the GPU (d_Peak, d_PeakOne2D, d_SubpixelMax)
against the FP64 oracle (oracle/oracle.cpp). Integer positions are index data
and must be identical. Sub-coarse positions lie on a 1/512 grid and sub-fine
positions on a 1/63 (3D) or 1/105 (2D) grid: they must agree to one grid step
(a step differs only where two grid points hold values equal to float
rounding). Values must agree to 1e-5 relative to the map's largest magnitude."""
import numpy as np
import pytest

from tests._util import to_np

pytestmark = pytest.mark.gpu


def _peaked_maps(shape, batch, seed, width=1.5):
    rng = np.random.default_rng(seed)
    maps = 0.3 * rng.standard_normal((batch,) + shape)
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    for b in range(batch):
        c = [rng.uniform(4, s - 4) for s in shape]
        r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
        maps[b] += 5.0 * np.exp(-r2 / (2 * width ** 2))
    return maps.astype(np.float32)


def _check(gx, orc, cuda, maps, ndims, mode):
    import torch
    pos, val = gx.d_Peak(torch.from_numpy(maps).to(cuda), mode, ndims=ndims)
    pos, val = to_np(pos).astype(np.float64), to_np(val).astype(np.float64)
    rpos, rval = orc.peak(maps, mode, dims=ndims)
    step = {"integer": 0.0, "subcoarse": 1 / 512, "subfine": 1 / (63 if ndims == 3 else 105)}[mode]
    assert np.abs(pos - rpos).max() <= step + 1e-5, (mode, pos, rpos)
    scale = np.abs(maps).max()
    assert np.abs(val - rval).max() <= 1e-5 * scale, (mode, val, rval)
    return pos, rpos


@pytest.mark.parametrize("mode", ["integer", "subcoarse", "subfine"])
@pytest.mark.parametrize("shape", [(40, 36, 30), (64, 64, 64)])
def test_peak_3d_matches_oracle(gx, orc, cuda, mode, shape):
    pos, rpos = _check(gx, orc, cuda, _peaked_maps(shape, 3, seed=sum(shape)), 3, mode)
    # off-grid disagreements would show as many mismatches; require most to be exact
    assert np.mean(np.all(np.abs(pos - rpos) <= 1e-5, axis=1)) >= 2 / 3


@pytest.mark.parametrize("mode", ["integer", "subcoarse", "subfine"])
@pytest.mark.parametrize("shape", [(20, 20), (96, 80), (256, 256)])
def test_peak_2d_matches_oracle(gx, orc, cuda, mode, shape):
    _check(gx, orc, cuda, _peaked_maps(shape, 4, seed=shape[0] + 7), 2, mode)


def test_peak_reference_case_2_3(gx, orc, cuda):
    """The reference's own test data recipe (GTOMTest/Peak.cpp cases 2-3): the
    GPU agrees with the oracle, and sub-fine stays within the reference's 0.02."""
    from tests.test_oracle_subpixel import _shifted_correlation
    maps = np.stack([_shifted_correlation(20, (1.5, 2.5), s) for s in range(6)])
    _check(gx, orc, cuda, maps, 2, "subcoarse")
    pos, _ = _check(gx, orc, cuda, maps, 2, "subfine")
    assert np.abs(pos[:, :2] - [11.5, 12.5]).mean(axis=1).max() <= 0.02


def test_peak_1d_and_ties(gx, orc, cuda):
    lines = _peaked_maps((50,), 3, seed=3)
    for mode in ("integer", "subcoarse", "subfine"):
        _check(gx, orc, cuda, lines, 1, mode)
    # plateaus: the first maximum in raster order
    q = np.zeros((2, 9, 10, 11), np.float32)
    q[0, 3, 4, 5] = q[0, 7, 1, 2] = q[0, 3, 4, 9] = 2.0
    q[1] = -1.0
    _check(gx, orc, cuda, q, 3, "integer")


def test_peak_full_size_map(gx, orc, cuda):
    """A 256^3 map (the size of the search's output) in all three modes."""
    import torch
    rng = np.random.default_rng(11)
    m = (0.2 * rng.standard_normal((256, 256, 256))).astype(np.float32)
    g = np.arange(-6, 7)
    zz, yy, xx = np.meshgrid(g, g, g, indexing="ij")
    m[140 - 6:140 + 7, 70 - 6:70 + 7, 200 - 6:200 + 7] += 4.0 * np.exp(
        -((xx - 0.3) ** 2 + (yy + 0.2) ** 2 + (zz - 0.4) ** 2) / 4.0)
    true = np.array([200.3, 69.8, 140.4])
    # accuracy of each mode on a noisy Gaussian: the lines of sub-coarse sample it
    # well; sub-fine's 5^3 window is not periodic, so its Fourier up-scaling rings
    for mode, tol in (("integer", 0.5), ("subcoarse", 0.1), ("subfine", 0.25)):
        pos, _ = _check(gx, orc, cuda, m[None], 3, mode)
        assert np.abs(pos[0] - true).max() < tol, mode


@pytest.mark.parametrize("shape,steps", [((40, 36), 3), ((24, 143), 2), ((24, 22, 20), 2), ((30, 19, 17), 1)])
def test_subpixel_max_matches_oracle(gx, orc, cuda, shape, steps):
    import torch
    rng = np.random.default_rng(steps + shape[0])
    v = rng.standard_normal(shape).astype(np.float32)
    init = np.full(shape, -2.5, np.float32)
    out = torch.from_numpy(init.copy()).to(cuda)
    gx.d_SubpixelMax(torch.from_numpy(v).to(cuda), out, steps)
    got = to_np(out).astype(np.float64)
    ref = orc.subpixel_max(v, init.astype(np.float64), steps)
    assert np.abs(got - ref).max() <= 2e-5 * np.abs(ref).max()
    assert np.array_equal(got == -2.5, ref == -2.5)  # the same voxels written


def test_subpixel_max_small_maps_write_nothing(gx, cuda):
    import torch
    v = torch.ones((15, 40), device=cuda)
    out = torch.full((15, 40), -1.0, device=cuda)
    gx.d_SubpixelMax(v, out, 2)
    assert bool((out == -1.0).all())


@pytest.mark.parametrize("dims,region,sub", [((24, 20), (10, 8), True), ((64, 48), (64, 48), False),
                                             ((7, 5), (3, 3), True)])
def test_peak_one_2d_matches_oracle(gx, orc, cuda, dims, region, sub):
    import torch
    rng = np.random.default_rng(dims[0])
    imgs = rng.random((5, dims[1], dims[0])).astype(np.float32)
    pos, val = gx.d_PeakOne2D(torch.from_numpy(imgs).to(cuda), region, sub)
    rpos, rval = orc.peak_one_2d(imgs, region, sub)
    assert np.abs(to_np(pos) - rpos).max() <= 1e-5 * max(dims)
    assert np.array_equal(to_np(val).astype(np.float64), rval)


@pytest.mark.parametrize("i", [10, 14, 18, 22, 24])
def test_integer_peak_reference_max_case(gx, cuda, i):
    """GTOMTest/Max.cpp: (rand() % 100000) * 1e-5 - 0.5 with one element set to
    0.500001; the maximum's index is found exactly (a seeded generator stands in
    for rand())."""
    import torch
    rng = np.random.default_rng(i + 1)
    size = (1 << i) - int(rng.integers(0, i))
    h = (rng.integers(0, 100000, size) * np.float32(0.00001) - np.float32(0.5)).astype(np.float32)
    k = int(rng.integers(0, size))
    h[k] = np.float32(0.500001)
    pos, val = gx.d_Peak(torch.from_numpy(h).to(cuda), "integer", ndims=1)
    assert to_np(pos)[0].tolist() == [float(k), 0.0, 0.0] and float(to_np(val)[0]) == float(h[k])
