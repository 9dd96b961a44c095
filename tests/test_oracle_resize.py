"""Pins the oracle's Fourier-space resizes (oracle/oracle.cpp fft_resize):
index formulations written independently with numpy (the kept index ranges
per axis), and the physics they serve: cropping the half spectrum of a
band-limited real signal and inverting at the new size gives the signal on
the coarser grid; padding the half spectrum of an odd-sized signal and
inverting at the larger size interpolates through the original samples; the
complex full pad is the real one applied to both parts."""
import numpy as np
import pytest

from oracle import pyoracle as po


def _keep_full(o, n):
    return np.concatenate([np.arange(n // 2), np.arange(o - (n - n // 2), o)])


def _keep_half_rows(o, n):
    return np.concatenate([np.arange(n // 2 + 1), np.arange(o - (n - n // 2 - 1), o)])


@pytest.mark.parametrize("old,new", [((12, 10, 8), (8, 6, 5)), ((9, 7, 11), (4, 7, 3))])
def test_full_crop_keeps_the_lowest_frequencies(old, new):
    rng = np.random.default_rng(0)
    a = rng.standard_normal(old[::-1])
    got = po.fft_resize("full_crop", a, old, new)
    want = a[np.ix_(_keep_full(old[2], new[2]), _keep_full(old[1], new[1]), _keep_full(old[0], new[0]))]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("old,new", [((12, 10, 8), (8, 6, 4)), ((9, 7, 11), (5, 7, 3))])
def test_half_crop_index_ranges(old, new):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((old[2], old[1], old[0] // 2 + 1)) + 1j * rng.standard_normal((old[2], old[1], old[0] // 2 + 1))
    got = po.fft_resize("crop", a, old, new)
    want = a[np.ix_(_keep_half_rows(old[2], new[2]), _keep_half_rows(old[1], new[1]), np.arange(new[0] // 2 + 1))]
    assert np.array_equal(got, want)


def _bandlimited(shape, kmax, rng):
    """Real periodic signal with frequencies |k| < kmax on every axis."""
    F = np.zeros(shape, complex)
    for _ in range(20):
        k = [int(rng.integers(-kmax + 1, kmax)) for _ in shape]
        c = rng.standard_normal() + 1j * rng.standard_normal()
        F[tuple(k)] += c
        F[tuple(-np.array(k))] += np.conj(c)
    return np.fft.ifftn(F).real


def test_half_crop_downsamples_a_bandlimited_signal():
    rng = np.random.default_rng(2)
    old, new = (16, 12, 8), (8, 6, 4)  # (x, y, z); factor 2
    sig = _bandlimited(old[::-1], 2, rng)  # |k| < 2 < new/2 on every axis
    half = np.fft.rfftn(sig)
    cropped = po.fft_resize("crop", half, old, new)
    small = np.fft.irfftn(cropped, s=new[::-1], axes=(0, 1, 2)) * (np.prod(new) / np.prod(old))
    assert np.allclose(small, sig[::2, ::2, ::2], atol=1e-12)


@pytest.mark.parametrize("old,new", [((5, 7, 3), (15, 21, 9)), ((8, 6, 4), (16, 12, 8))])
def test_half_pad_interpolates_through_the_samples(old, new):
    rng = np.random.default_rng(3)
    f = [n // o for o, n in zip(old, new)]
    # odd sizes: any real signal; even sizes: keep the old Nyquist empty (band-limited)
    sig = rng.standard_normal(old[::-1]) if all(o % 2 for o in old) else _bandlimited(old[::-1], min(old) // 2, rng)
    half = np.fft.rfftn(sig)
    padded = po.fft_resize("pad", half, old, new)
    big = np.fft.irfftn(padded, s=new[::-1], axes=(0, 1, 2)) * (np.prod(new) / np.prod(old))
    assert np.allclose(big[::f[2], ::f[1], ::f[0]], sig, atol=1e-12)


def test_complex_full_pad_is_the_real_one_on_both_parts():
    rng = np.random.default_rng(4)
    old, new = (6, 5, 4), (10, 9, 8)
    a = rng.standard_normal(old[::-1]) + 1j * rng.standard_normal(old[::-1])
    got = po.fft_resize("full_pad", a, old, new)
    assert np.array_equal(got.real, po.full_pad(a.real, new[::-1]))
    assert np.array_equal(got.imag, po.full_pad(a.imag, new[::-1]))
