"""d_NormMonolithic and d_NormFTMonolithic on the GPU against the FP64 oracle:
within 2e-6 of the output scale (float data, float mean and scale), with the
zero output for a constant item bit-exact; d_Sum against numpy's float64 sum
(float: 1e-6 relative, double: 1e-12, int: exact with 32-bit wrap-around)."""
import numpy as np
import pytest

from tests._util import to_np

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(3, 100), (2, 64, 64, 64), (1, 256, 256, 256)])
def test_norm_monolithic_matches_oracle(gx, orc, cuda, shape):
    import torch
    rng = np.random.default_rng(len(shape))
    x = (rng.standard_normal(shape) * 3.0 + 1.0).astype(np.float32)
    got = to_np(gx.d_NormMonolithic(torch.from_numpy(x).to(cuda), batch=shape[0]))
    ref = orc.norm_monolithic(x)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


def test_norm_monolithic_constant_item_is_zero(gx, cuda):
    import torch
    x = torch.full((2, 5000), 1.5, device=cuda)
    x[1] = torch.arange(5000, device=cuda, dtype=torch.float32)
    y = gx.d_NormMonolithic(x, batch=2)
    assert bool((y[0] == 0).all()) and float(y[1].std(unbiased=False)) == pytest.approx(1.0, rel=1e-5)


@pytest.mark.parametrize("shape", [(4, 33), (2, 256, 256, 129)])
def test_norm_ft_monolithic_matches_oracle(gx, orc, cuda, shape):
    import torch
    rng = np.random.default_rng(shape[-1])
    x = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64) * 7.0
    got = to_np(gx.d_NormFTMonolithic(torch.from_numpy(x).to(cuda), batch=shape[0]))
    ref = orc.norm_ft_monolithic(x)
    assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


def test_sum_matches_numpy(gx, cuda):
    import torch
    rng = np.random.default_rng(11)
    x = rng.standard_normal((3, 1 << 20)).astype(np.float32)
    got = to_np(gx.d_Sum(torch.from_numpy(x).to(cuda), batch=3))
    assert np.allclose(got, x.astype(np.float64).sum(axis=1), rtol=1e-6, atol=1e-3)
    d = rng.standard_normal((2, 100003))
    got = to_np(gx.d_Sum(torch.from_numpy(d).to(cuda), batch=2))
    assert np.allclose(got, d.sum(axis=1), rtol=1e-12, atol=1e-9)
    i = rng.integers(-2**31, 2**31 - 1, (2, 5000)).astype(np.int32)  # the sums overflow 32 bits
    got = to_np(gx.d_Sum(torch.from_numpy(i).to(cuda), batch=2))
    want = ((i.astype(np.int64).sum(axis=1) + 2**31) % 2**32 - 2**31).astype(np.int32)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("i", [2, 7, 12, 17, 20, 24])
def test_sum_reference_case(gx, cuda, i):
    """GTOMTest/Sum.cpp: 2^i - 1 values (rand() % 100000) * 1e-5 summed within 1e-6
    (relative) of their float32 Kahan sum (a seeded generator stands in for rand();
    beyond 5000 values the exact sum, which the Kahan sum equals to float rounding)."""
    import torch
    size = (1 << i) - 1
    h = (np.random.default_rng(i).integers(0, 100000, size) * np.float32(0.00001)).astype(np.float32)
    s, c = np.float32(h[0]), np.float32(0)
    for v in h[1:]:  # Kahan summation in float32, as the reference test
        y = np.float32(v - c)
        t = np.float32(s + y)
        c = np.float32(np.float32(t - s) - y)
        s = t
    if size > 5000:  # the loop above is slow in Python beyond a few thousand; use the exact sum there
        s = np.float32(h.astype(np.float64).sum())
    got = float(to_np(gx.d_Sum(torch.from_numpy(h).to(cuda)))[0])
    assert abs(got - float(s)) / float(s) <= 1e-6
