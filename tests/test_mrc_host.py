"""MRC I/O (host side of libgtomx.so; no GPU needed). The header layout is
checked against the MRC format's byte offsets, reading against a data file
from the reference's own test data (tests/golden/point.rec, a 64^3 float
map), independently parsed with numpy, and writing by round trips read back
with numpy."""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
POINT = os.path.join(HERE, "golden", "point.rec")


@pytest.fixture(scope="module")
def gxh():
    import gtom_amd as gx
    gx.lib()
    return gx


def test_header_layout_matches_the_format(gxh):
    H = gxh.MrcHeader
    assert ctypes.sizeof(H) == 1024
    off = {f: getattr(H, f).offset for f, _ in H._fields_}
    # MRC2014 word offsets: NX 0, MODE 12, NXSTART 16, MX 28, CELLA 40, CELLB 52,
    # MAPC 64, DMIN 76, ISPG 88, NSYMBT 92, ORIGIN 196, MAP 208, MACHST 212, RMS 216, NLABL 220, LABEL 224
    for f, o in dict(nx=0, mode=12, startx=16, mx=28, cellx=40, alpha=52, mapc=64, dmin=76, ispg=88, nsymbt=92,
                     origin=196, cmap=208, stamp=212, rms=216, nlabl=220, labels=224).items():
        assert off[f] == o, (f, off[f], o)


def test_reads_reference_test_map(gxh):
    raw = open(POINT, "rb").read()
    ints = np.frombuffer(raw[:1024], "<i4")
    floats = np.frombuffer(raw[:1024], "<f4")
    h, vol = gxh.read_mrc(POINT)
    assert (h.nx, h.ny, h.nz, h.mode) == tuple(ints[:4]) == (64, 64, 64, 2)
    assert (h.cellx, h.dmin, h.dmax) == (floats[10], floats[19], floats[20])
    assert h.cmap == b"MAP "
    ref = np.frombuffer(raw[1024 + ints[23]:], "<f4").reshape(64, 64, 64)
    assert np.array_equal(vol, ref)
    assert vol.min() == h.dmin and vol.max() == h.dmax
    # one z section
    _, sec = gxh.read_mrc(POINT, nframe=37)
    assert np.array_equal(sec[0], ref[37])


@pytest.mark.parametrize("mode,dtype", [(0, np.uint8), (1, np.int16), (2, np.float32), (6, np.uint16)])
def test_write_read_round_trip(gxh, tmp_path, mode, dtype):
    rng = np.random.default_rng(mode)
    data = (rng.standard_normal((5, 6, 7)) * 50 + (100 if mode in (0, 6) else 0)).astype(dtype)
    h = gxh.mrc_header_default()
    h.nx, h.ny, h.nz, h.mode = 7, 6, 5, mode
    h.nsymbt = 80  # written as 0: no extended header follows
    p = str(tmp_path / f"m{mode}.mrc")
    gxh.write_mrc(p, data, header=h)
    raw = open(p, "rb").read()
    assert len(raw) == 1024 + data.nbytes
    assert np.frombuffer(raw[:1024], "<i4")[23] == 0
    h2, back = gxh.read_mrc(p)
    assert (h2.nx, h2.ny, h2.nz, h2.mode) == (7, 6, 5, mode)
    assert np.array_equal(back, data)
    assert np.array_equal(np.frombuffer(raw[1024:], data.dtype).reshape(data.shape), data)


def test_write_float_sets_min_max_and_defaults(gxh, tmp_path):
    data = np.linspace(-3, 5, 4 * 5 * 6, dtype=np.float32).reshape(4, 5, 6)
    p = str(tmp_path / "f.mrc")
    gxh.write_mrc(p, data)
    h, back = gxh.read_mrc(p)
    assert (h.dmin, h.dmax) == (-3.0, 5.0)
    assert (h.mapc, h.mapr, h.maps, h.mx, h.cellx, h.cmap) == (1, 2, 3, 1, 1.0, b"MAP ")
    assert np.array_equal(back, data)


def test_errors(gxh, tmp_path):
    with pytest.raises(gxh.GtomxError):
        gxh.read_mrc_header(str(tmp_path / "missing.mrc"))
    p = str(tmp_path / "short.mrc")
    open(p, "wb").write(b"\0" * 100)
    with pytest.raises(gxh.GtomxError):
        gxh.read_mrc_header(p)
    with pytest.raises(gxh.GtomxError):
        gxh.read_mrc(POINT, nframe=64)
