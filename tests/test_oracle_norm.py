"""Pins the oracle's d_NormMonolithic / d_NormFTMonolithic restatements
(oracle/oracle.cpp norm_monolithic, norm_ft_monolithic) against numpy: the
population standardisation, the zero output where the spread is zero, and
division by the RMS magnitude (zero input stays zero through the 1e-20 floor)."""
import numpy as np

from oracle import pyoracle as po


def test_norm_monolithic_is_numpy_standardisation():
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((3, 7, 11)) * 2.5 + 4.0).astype(np.float32)
    got = po.norm_monolithic(x)
    x64 = x.astype(np.float64)
    want = (x64 - x64.mean(axis=(1, 2), keepdims=True)) / x64.std(axis=(1, 2), keepdims=True)
    assert np.allclose(got, want, atol=1e-10)


def test_norm_monolithic_zero_spread_gives_zero():
    x = np.full((2, 1000), 1.5, np.float32)
    assert np.all(po.norm_monolithic(x) == 0.0)


def test_norm_ft_monolithic_divides_by_rms():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal((2, 5, 9)) + 1j * rng.standard_normal((2, 5, 9))).astype(np.complex64)
    got = po.norm_ft_monolithic(x)
    x128 = x.astype(np.complex128)
    want = x128 / np.sqrt((np.abs(x128) ** 2).mean(axis=(1, 2), keepdims=True))
    assert np.allclose(got, want, atol=1e-12)
    assert np.all(po.norm_ft_monolithic(np.zeros((1, 4), np.complex64)) == 0)
