"""The element-wise exports on their own (the search fuses them), called through
the C ABI with batch > 1, in place and out of place:

- gtomx_complex_multiply_by_vector: out[b][i] = in[b][i] * v[i]
  (src/Generics/Arithmetics.cu:192-197, 275-286), bit-exact against the two float32
  products per element that numpy forms;
- gtomx_complex_multiply_by_conj_vector: out[b][i] = in[b][i] * conj(v[i])
  (src/Generics/Arithmetics.cu:220-225, 334-344), within 2 float ulps of the
  float64 product (a fused multiply-add may round once where numpy rounds twice);
  a swapped conjugation fails by the full magnitude;
- gtomx_divide_safe_by_vector: out = |d| > 1e-15 ? in / d : 0
  (src/Generics/Arithmetics.cu:478-493), bit-exact against numpy's float32
  division, including divisors at and just above 1e-15 and negative ones.

d_LocalStd on a map with a large constant offset (mean 100, sigma 1) against the
FP64 oracle: the variance M*conv2 - conv1^2 is formed after the map's mean is
subtracted, so sigma keeps the zero-mean test's accuracy (1e-5 of its scale)."""
import numpy as np
import pytest

from tests._util import rel_linf, to_np

pytestmark = pytest.mark.gpu


def _call(gx, name, *args):
    gx.check(getattr(gx.lib(), name)(*args), name)


@pytest.mark.parametrize("inplace", [False, True])
def test_complex_multiply_by_vector(gx, cuda, inplace):
    import torch
    rng = np.random.default_rng(1)
    n, batch = 100003, 3
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64) * 50
    v = rng.standard_normal(n).astype(np.float32)
    xd = torch.from_numpy(x).to(cuda)
    vd = torch.from_numpy(v).to(cuda)
    out = xd if inplace else torch.empty_like(xd)
    _call(gx, "gtomx_complex_multiply_by_vector", gx._ptr(xd), gx._ptr(vd), gx._ptr(out), n, batch, gx._stream())
    torch.cuda.synchronize()
    got = to_np(out)
    assert np.array_equal(got.real, x.real * v[None, :]) and np.array_equal(got.imag, x.imag * v[None, :])


@pytest.mark.parametrize("inplace", [False, True])
def test_complex_multiply_by_conj_vector(gx, cuda, inplace):
    import torch
    rng = np.random.default_rng(2)
    n, batch = 65537, 4
    x = (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n))).astype(np.complex64)
    v = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    xd = torch.from_numpy(x).to(cuda)
    vd = torch.from_numpy(v).to(cuda)
    out = xd if inplace else torch.empty_like(xd)
    _call(gx, "gtomx_complex_multiply_by_conj_vector", gx._ptr(xd), gx._ptr(vd), gx._ptr(out), n, batch,
          gx._stream())
    torch.cuda.synchronize()
    want = x.astype(np.complex128) * np.conj(v.astype(np.complex128))[None, :]
    got = to_np(out).astype(np.complex128)
    ulp = np.spacing(np.maximum(np.abs(want.real), np.abs(want.imag)).astype(np.float32)).astype(np.float64)
    assert (np.abs(got.real - want.real) <= 2 * ulp + 1e-30).all()
    assert (np.abs(got.imag - want.imag) <= 2 * ulp + 1e-30).all()
    wrong = x.astype(np.complex128) * v.astype(np.complex128)[None, :]
    assert np.abs(wrong - got).max() > 0.5 * np.abs(want).max()


@pytest.mark.parametrize("inplace", [False, True])
def test_divide_safe_by_vector(gx, cuda, inplace):
    import torch
    rng = np.random.default_rng(3)
    n, batch = 4099, 3
    x = rng.standard_normal((batch, n)).astype(np.float32)
    d = rng.standard_normal(n).astype(np.float32)
    d[:8] = np.array([0.0, 1e-15, -1e-15, 1e-16, np.nextafter(np.float32(1e-15), np.float32(1)), -2e-15, 1e-30,
                      -0.0], np.float32)
    xd = torch.from_numpy(x).to(cuda)
    dd = torch.from_numpy(d).to(cuda)
    out = xd if inplace else torch.empty_like(xd)
    _call(gx, "gtomx_divide_safe_by_vector", gx._ptr(xd), gx._ptr(dd), gx._ptr(out), n, batch, gx._stream())
    torch.cuda.synchronize()
    ok = np.abs(d) > np.float32(1e-15)
    with np.errstate(divide="ignore", invalid="ignore"):
        want = np.where(ok[None, :], x / d[None, :], np.float32(0)).astype(np.float32)
    got = to_np(out)
    assert np.array_equal(got, want)
    assert not ok[:4].any() and ok[4] and ok[5]


def test_local_std_with_large_offset(gx, orc, cuda):
    import torch
    rng = np.random.default_rng(4)
    v = (100.0 + rng.standard_normal((48, 48, 48))).astype(np.float32)
    sd, mean = gx.d_LocalStd(torch.from_numpy(v).to(cuda), 6.0, with_mean=True)
    torch.cuda.synchronize()
    rsd, rmean = orc.local_std(v, 6.0, with_mean=True)
    assert rel_linf(to_np(sd), rsd) <= 1e-5
    assert rel_linf(to_np(mean), rmean) <= 1e-6
    # the spread is of order 1 here, not 1e-2 (what float cancellation at mean 100 would leave)
    assert 0.5 < float(rsd.mean()) < 1.5
