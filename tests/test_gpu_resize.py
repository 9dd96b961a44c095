"""Fourier-space resizes on the GPU against the oracle: pure index maps, so
bit-exact, for real and complex arrays, odd and even sizes, batches, and in
place (input buffer sized for the larger dims)."""
import numpy as np
import pytest

from tests._util import to_np

pytestmark = pytest.mark.gpu

CASES = [("crop", (16, 12, 10), (8, 6, 5)), ("crop", (9, 7, 11), (5, 7, 3)),
         ("pad", (8, 6, 5), (16, 12, 10)), ("pad", (5, 7, 3), (15, 21, 9)),
         ("full_crop", (12, 10, 8), (8, 6, 5)), ("full_crop", (9, 7, 11), (4, 7, 3)),
         ("full_pad", (6, 5, 4), (10, 9, 8)), ("full_pad", (7, 5, 3), (12, 10, 8))]


def _shape(op, d):
    return (d[2], d[1], d[0] // 2 + 1 if op in ("crop", "pad") else d[0])


@pytest.mark.parametrize("op,old,new", CASES)
@pytest.mark.parametrize("cplx", [False, True])
def test_resize_matches_oracle(gx, orc, cuda, op, old, new, cplx):
    import torch
    rng = np.random.default_rng(len(op) + old[0])
    a = rng.standard_normal((2,) + _shape(op, old)).astype(np.float32)
    if cplx:
        a = (a + 1j * rng.standard_normal(a.shape)).astype(np.complex64)
    got = to_np(gx.d_FFTResize(op, torch.from_numpy(a).to(cuda), old, new))
    for b in range(2):
        ref = orc.fft_resize(op, a[b], old, new)
        assert np.array_equal(got[b].astype(ref.dtype), ref), (op, b)


def test_resize_in_place(gx, orc, cuda):
    import torch
    op, old, new = "pad", (8, 6, 5), (16, 12, 10)
    rng = np.random.default_rng(9)
    a = (rng.standard_normal(_shape(op, old)) + 1j * rng.standard_normal(_shape(op, old))).astype(np.complex64)
    buf = torch.zeros(int(np.prod(_shape(op, new))), dtype=torch.complex64, device=cuda)
    buf[:a.size] = torch.from_numpy(a.ravel()).to(cuda)
    gx.check(gx.lib().gtomx_fft_resize(2, 1, gx._ptr(buf), gx._ptr(buf), gx.int3(*old), gx.int3(*new), 1, None),
             "gtomx_fft_resize")
    assert np.array_equal(to_np(buf).reshape(_shape(op, new)).astype(np.complex128), orc.fft_resize(op, a, old, new))
