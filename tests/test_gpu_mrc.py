"""MRC data to and from the device: MixedToDeviceTfloat widening on the GPU
(exact for every mode) and d_WriteMRC read back with numpy."""
import os

import numpy as np
import pytest

from tests._util import to_np

pytestmark = pytest.mark.gpu
POINT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "point.rec")


@pytest.mark.parametrize("mode,dtype", [(0, np.uint8), (1, np.int16), (2, np.float32), (6, np.uint16)])
def test_mode_to_float_on_device(gx, cuda, mode, dtype):
    rng = np.random.default_rng(mode)
    info = np.iinfo(dtype) if mode != 2 else None
    a = (rng.integers(info.min, info.max, 100003, endpoint=True).astype(dtype) if info
         else rng.standard_normal(100003).astype(dtype))
    got = to_np(gx.mrc_to_device(a, mode, device=cuda))
    assert np.array_equal(got, a.astype(np.float32))


def test_reference_map_to_device_and_back(gx, cuda, tmp_path):
    h, vol = gx.read_mrc(POINT)
    d = gx.mrc_to_device(vol, h.mode, device=cuda)
    p = str(tmp_path / "back.mrc")
    gx.d_WriteMRC(d, p)
    raw = open(p, "rb").read()
    back = np.frombuffer(raw[1024:], "<f4").reshape(64, 64, 64)
    assert np.array_equal(back, vol)
    f = np.frombuffer(raw[:1024], "<f4")
    assert (f[19], f[20]) == (vol.min(), vol.max())
