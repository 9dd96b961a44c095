"""d_PickSubTomogramsDiff2 on the GPU against the FP64 oracle (a literal sum of
|R ctf e^{i k.s} - E|^2 over the half spectrum). Tolerance: the best-score map
within 1e-5 normwise (max|gpu - fp64| / max|fp64|); the orientation index
identical wherever the oracle's margin to the second-best score exceeds
1e-5 * max|best|."""
import numpy as np
import pytest

from tests._util import rel_linf, to_np

pytestmark = pytest.mark.gpu
TAU = 2 * np.pi


def _setup(gx, orc, cuda, n, box, nvol, seed):
    import torch
    rng = np.random.default_rng(seed)
    g = np.arange(box) - box // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    t = sum(np.exp(-(((x - c[0]) / 1.5) ** 2 + ((y - c[1]) / 2.2) ** 2 + ((z - c[2]) / 1.3) ** 2) / 2)
            for c in rng.uniform(-box / 5, box / 5, (5, 3))).astype(np.float32)
    p = 2 * box
    projft = gx.prepare_projector(torch.from_numpy(t).to(cuda), p)
    vols = rng.standard_normal((nvol, n, n, n)).astype(np.float32)
    E = gx.d_FFTR2C(torch.from_numpy(vols).to(cuda))
    ctf = torch.from_numpy(rng.uniform(-1, 1, (nvol, n, n, n // 2 + 1)).astype(np.float32)).to(cuda)
    return projft, p, E, ctf


@pytest.mark.parametrize("rel", [(4, 5, 6), (1, 8, 3)])
def test_diff2_matches_oracle(gx, orc, cuda, rel):
    n, box = 32, 16
    projft, p, E, ctf = _setup(gx, orc, cuda, n, box, nvol=2, seed=sum(rel))
    angles = orc.equal_angular_spacing((0, TAU), (0, np.pi), (0, TAU), np.deg2rad(60.0))[:40]
    best, ang = gx.d_PickSubTomogramsDiff2(projft, p, p / n, E, ctf, rel, angles)
    best, ang = to_np(best), to_np(ang)
    P = to_np(projft).astype(np.complex128)
    for v in range(2):
        rb, ri, r2 = orc.diff2(P, p, p / n, to_np(E[v]).astype(np.complex128), to_np(ctf[v]).astype(np.float64), n,
                               rel, angles.astype(np.float64), with_second=True)
        assert rel_linf(best[v], rb) <= 1e-5
        sure = (r2 - rb) > 1e-5 * np.abs(rb).max()
        assert sure.sum() > 0.5 * sure.size  # noise volumes: many near-ties
        assert np.array_equal(ang[v][sure], ri[sure].astype(np.float32))


def test_diff2_planted_shift_found(gx, orc, cuda):
    """E = R_k ctf e^{i k.s} for one orientation and one relevant shift scores ~0 there."""
    import torch
    n, box, k = 32, 16, 11
    projft, p, _, ctf = _setup(gx, orc, cuda, n, box, nvol=1, seed=9)
    angles = orc.equal_angular_spacing((0, TAU), (0, np.pi), (0, TAU), np.deg2rad(60.0))[:20]
    R = to_np(gx.d_rlnProject(projft, p, n, angles=angles[k:k + 1], oversample=p / n))[0].astype(np.complex128)
    R = R * to_np(ctf[0])
    rel = (5, 5, 5)
    vz, vy, vx = 4, 0, 2
    s = -TAU * (np.array([vx, vy, vz]) - 2) / n
    kz = np.fft.fftfreq(n, 1 / n)[:, None, None]
    ky = np.fft.fftfreq(n, 1 / n)[None, :, None]
    kx = np.arange(n // 2 + 1)[None, None, :]
    Ev = (R * np.exp(1j * (kx * s[0] + ky * s[1] + kz * s[2]))).astype(np.complex64)
    E = torch.from_numpy(Ev[None]).to(cuda)
    best, ang = gx.d_PickSubTomogramsDiff2(projft, p, p / n, E, ctf, rel, angles)
    best, ang = to_np(best)[0], to_np(ang)[0]
    assert np.unravel_index(np.argmin(best), rel) == (vz, vy, vx)
    assert ang[vz, vy, vx] == k
    assert best[vz, vy, vx] < 1e-4 * np.sum(np.abs(Ev) ** 2)
