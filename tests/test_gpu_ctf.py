"""d_CTFSimulate and the astigmatic 2D CTF on the GPU against the FP64 oracle
(oracle/oracle.cpp ctf_at): within 2e-6 absolute (float output of a double
evaluation)."""
import numpy as np
import pytest

from tests._util import to_np

pytestmark = pytest.mark.gpu

CASES = [dict(pixelsize=1.2e-10, defocus=-2.1e-6, defocusdelta=0.3e-6, astigmatismangle=0.7),
         dict(pixelsize=2e-10, defocus=-1.0e-6, phaseshift=0.5, Bfactor=-80e-20, Bfactordelta=20e-20,
              Bfactorangle=-0.4, pixeldelta=0.05e-10, pixelangle=1.1, scale=0.8),
         dict(pixelsize=0.9e-10, defocus=-3.5e-6, amplitude=0.1, Cs=2.7e-3, voltage=200e3)]


def _f32(cases):
    """The parameters as the float32 CTFParams fields hold them (the oracle gets the same inputs)."""
    return [{k: float(np.float32(v)) for k, v in c.items()} for c in cases]


@pytest.mark.parametrize("ampsq,first", [(False, False), (True, False), (False, True)])
def test_ctf_simulate_matches_oracle(gx, orc, cuda, ampsq, first):
    import torch
    rng = np.random.default_rng(5)
    addr = np.stack([rng.uniform(0, 0.5, 5000), rng.uniform(-np.pi, np.pi, 5000)], 1).astype(np.float32)
    gcorr = rng.uniform(-0.2, 0.2, 5000).astype(np.float32)
    got = to_np(gx.d_CTFSimulate([gx.CtfParamsFull.make(**c) for c in CASES], torch.from_numpy(addr).to(cuda),
                                 torch.from_numpy(gcorr).to(cuda), ampsq, first))
    ref = orc.ctf_simulate(_f32(CASES), addr.astype(np.float64), gcorr.astype(np.float64), ampsq, first)
    assert np.max(np.abs(got - ref)) <= 2e-6


def test_ctf_2d_matches_oracle(gx, orc, cuda):
    n = 96
    got = to_np(gx.ctf_2d([gx.CtfParamsFull.make(**c) for c in CASES], n, device=cuda))
    ref = orc.ctf_simulate(_f32(CASES), orc.ctf_addresses_2d(n)).reshape(len(CASES), n, n // 2 + 1)
    assert np.max(np.abs(got - ref)) <= 2e-6


def test_tilted_region_ctfs_match_oracle(gx, orc, cuda):
    """Per-region CTFs of a tilted specimen: GetParamsGrid2D on the host, then the
    2D half-spectrum CTF of each region on the GPU, against the oracle's
    d_CTFSimulate at the same addresses with the oracle's own tilted defoci."""
    centre = gx.CtfParamsFull.make(pixelsize=2e-10, defocus=-3e-6, defocusdelta=0.2e-6, astigmatismangle=0.7)
    tilt = gx.CtfTiltParams.make(centre, imageangle=0.3, stageangle=(0.2, 0.6))
    pos = [(0, 0), (896, 0), (448, 448), (0, 896), (896, 896)]
    params = gx.ctf_tilt_params_grid(tilt, (1024, 1024), (128, 128), pos)
    got = to_np(gx.ctf_2d(params, 128, device=cuda))
    f32 = lambda v: float(np.float32(v))
    z = orc.ctf_tilt_zgrid(f32(0.3), (f32(0.2), f32(0.6)), (0.0, 0.0), f32(centre.defocus), f32(centre.pixelsize),
                           (1024, 1024), (128, 128), pos)
    assert np.ptp(z) > 1e-8  # the plane's tilt separates the regions' defoci
    fields = {f: f32(getattr(centre, f)) for f, _ in gx.CtfParamsFull._fields_}
    ref = orc.ctf_simulate([dict(fields, defocus=f32(zi)) for zi in z], orc.ctf_addresses_2d(128))
    assert np.abs(got.reshape(len(pos), -1) - ref).max() <= 2e-5
