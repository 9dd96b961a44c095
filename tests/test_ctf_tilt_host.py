"""CTFTiltParams::GetParamsGrid2D (include/CTF.cuh:160-194) in the library
(host code, no device work) against the oracle's restatement: defocus per
region centre to float rounding, every other field copied from centerparams."""
import numpy as np
import pytest


@pytest.mark.parametrize("image,stage,spec", [(0.0, (0.0, 0.5), (0.0, 0.0)), (0.8, (1.2, -0.3), (0.4, 0.15)),
                                              (-0.5, (0.0, 0.0), (2.0, 0.6))])
def test_params_grid_matches_oracle(gx, orc, image, stage, spec):
    centre = gx.CtfParamsFull.make(pixelsize=1.7e-10, defocus=-2.2e-6, defocusdelta=0.3e-6, astigmatismangle=0.4,
                                   Bfactor=-20e-20)
    tilt = gx.CtfTiltParams.make(centre, imageangle=image, stageangle=stage, specimenangle=spec)
    rng = np.random.default_rng(5)
    pos = rng.integers(0, 4096 - 256, (40, 2))
    got = gx.ctf_tilt_params_grid(tilt, (4096, 3072), (256, 256), pos)
    f32 = lambda v: float(np.float32(v))
    want = orc.ctf_tilt_zgrid(f32(image), tuple(map(f32, stage)), tuple(map(f32, spec)), f32(centre.defocus),
                              f32(centre.pixelsize), (4096, 3072), (256, 256), pos)
    assert np.allclose([p.defocus for p in got], want, rtol=2e-7, atol=0)
    for p in got:
        for f, _ in gx.CtfParamsFull._fields_:
            if f != "defocus":
                assert getattr(p, f) == getattr(centre, f)
