"""Pins the oracle's d_LocalPeaks restatement (oracle/oracle.cpp, orc_local_peaks_f32).

The reference ships no test for it, so the checks are: an independent numpy
formulation (shifted whole-array comparisons against a -inf margin instead of
a clamped per-voxel scan), and known answers that single out the rule's
details: ties do not reject (strict '<'), the threshold itself qualifies,
the faces clamp (no wrap), and the neighbourhood is the Euclidean ball.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def _numpy_peaks(vol, r, thr):
    nz, ny, nx = vol.shape
    pad = np.full((nz + 2 * r, ny + 2 * r, nx + 2 * r), -np.inf, np.float32)
    pad[r:r + nz, r:r + ny, r:r + nx] = vol
    ok = ~(vol < thr)
    for dz in range(-r, r + 1):
        for dy in range(-r, r + 1):
            for dx in range(-r, r + 1):
                d2 = dx * dx + dy * dy + dz * dz
                if d2 == 0 or d2 > r * r:
                    continue
                nb = pad[r + dz:r + dz + nz, r + dy:r + dy + ny, r + dx:r + dx + nx]
                ok &= ~(vol < nb)
    zyx = np.argwhere(ok)  # raster order
    return zyx[:, ::-1].astype(np.int64)


@pytest.mark.parametrize("r", [0, 1, 2, 3, 5])
def test_matches_numpy_formulation(r):
    rng = np.random.default_rng(r + 3)
    # quantised values make ties common
    vol = np.round(rng.standard_normal((11, 13, 9)) * 3).astype(np.float32)
    for thr in (-1e30, 0.0, 2.0):
        got = po.local_peaks(vol, r, thr)
        ref = _numpy_peaks(vol, r, thr)
        assert np.array_equal(got, ref), (r, thr)


def test_single_delta_and_threshold():
    vol = np.zeros((8, 9, 10), np.float32)
    vol[3, 4, 5] = 2.0
    assert po.local_peaks(vol, 2, 1.0).tolist() == [[5, 4, 3]]
    # the threshold itself qualifies (the kernel rejects value < threshold)
    assert po.local_peaks(vol, 2, 2.0).tolist() == [[5, 4, 3]]
    assert po.local_peaks(vol, 2, 2.0001).shape == (0, 3)


def test_plateau_ties_are_all_peaks():
    vol = np.zeros((6, 6, 6), np.float32)
    vol[2, 2, 2:4] = 1.0  # two equal neighbours
    got = po.local_peaks(vol, 3, 0.5).tolist()
    assert got == [[2, 2, 2], [3, 2, 2]]


def test_faces_clamp_without_wrap():
    vol = np.zeros((5, 5, 16), np.float32)
    vol[2, 2, 0] = 1.0
    vol[2, 2, 15] = 3.0  # across the periodic boundary, 1 voxel away if it wrapped
    got = po.local_peaks(vol, 2, 0.5).tolist()
    assert got == [[0, 2, 2], [15, 2, 2]]


def test_ball_not_box():
    # a larger value at (2, 2, 0) offset: squared distance 8 > 2^2, outside the
    # radius-2 ball although inside its bounding box
    vol = np.zeros((7, 7, 7), np.float32)
    vol[3, 3, 3] = 1.0
    vol[3, 5, 5] = 2.0
    assert [3, 3, 3] in po.local_peaks(vol, 2, 0.5).tolist()
    assert [3, 3, 3] not in po.local_peaks(vol, 3, 0.5).tolist()


def test_gaussian_blobs_peak_at_centres():
    z, y, x = np.meshgrid(np.arange(20), np.arange(22), np.arange(24), indexing="ij")
    centres = [(4, 5, 6), (15, 16, 6), (10, 10, 17)]
    vol = np.zeros(x.shape, np.float64)
    for i, (cx, cy, cz) in enumerate(centres):
        vol += (1 + 0.1 * i) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / 8.0)
    got = po.local_peaks(vol.astype(np.float32), 4, 0.5).tolist()
    assert got == sorted([list(c) for c in centres], key=lambda c: (c[2], c[1], c[0]))
