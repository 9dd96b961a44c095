"""d_LocalPeaks on the GPU against the oracle (oracle/oracle.cpp,
orc_local_peaks_f32). Peak lists are index data: they must be identical,
element for element and in order."""
import numpy as np
import pytest

from tests._util import to_np

pytestmark = pytest.mark.gpu


def _smooth_field(shape, sigma, seed):
    import scipy.ndimage as ndi
    rng = np.random.default_rng(seed)
    f = ndi.gaussian_filter(rng.standard_normal(shape), sigma, mode="wrap")
    return (f / f.std()).astype(np.float32)


@pytest.mark.parametrize("r", [0, 1, 2, 3, 6])
def test_quantised_random_matches_oracle(gx, orc, cuda, r):
    import torch
    rng = np.random.default_rng(40 + r)
    vol = np.round(rng.standard_normal((2, 19, 23, 29)) * 2).astype(np.float32)  # many ties
    got = gx.d_LocalPeaks(torch.from_numpy(vol).to(cuda), r, 1.0)
    assert len(got) == 2
    for b in range(2):
        ref = orc.local_peaks(vol[b], r, 1.0)
        assert np.array_equal(got[b], ref), (r, b, len(got[b]), len(ref))


@pytest.mark.parametrize("shape,r,thr", [((64, 64, 64), 4, 1.5), ((96, 80, 72), 8, 2.0),
                                         ((256, 256, 256), 10, 3.0)])
def test_smooth_field_matches_oracle(gx, orc, cuda, shape, r, thr):
    import torch
    vol = _smooth_field(shape, 2.0, seed=sum(shape))
    got = gx.d_LocalPeaks(torch.from_numpy(vol).to(cuda), r, thr)[0]
    ref = orc.local_peaks(vol, r, thr)
    assert len(ref) > 0
    assert np.array_equal(got, ref)


def test_more_peaks_than_first_capacity(gx, orc, cuda):
    """Every voxel is a peak at radius 0 and threshold -inf: 8000 > the binding's first capacity."""
    import torch
    vol = np.random.default_rng(3).standard_normal((20, 20, 20)).astype(np.float32)
    got = gx.d_LocalPeaks(torch.from_numpy(vol).to(cuda), 0, -np.inf)[0]
    assert got.shape == (8000, 3)
    assert np.array_equal(got, orc.local_peaks(vol, 0, -np.inf))


def test_search_output_peaks_at_planted_copies(gx, orc, cuda):
    """The consumer step: peaks of a small search's normalised max-CC map match
    the oracle's peaks of the GPU map (same input), and include the planted
    positions."""
    import torch
    from gtom_amd import synth
    w = synth.cfg1(seed=5, n=32, box=16, ncopies=2, snr=2.0, maskradius=8)
    n, p = w["n"], w["p"]
    projft = gx.prepare_projector(torch.from_numpy(w["template"]).to(cuda), p)
    E = gx.d_FFTR2C(torch.from_numpy(w["tomogram"][None]).to(cuda))
    ctf = torch.ones((n // 2 + 1) * n * n, dtype=torch.float32, device=cuda)
    plan = gx.PickPlan(p, n, 0)
    best, idx = gx.d_PickSubTomograms(plan, projft, w["oversample"], E, ctf, w["angles"][:64], w["maskradius"])
    plan.close()
    torch.cuda.synchronize()
    got = gx.d_LocalPeaks(best, 4, float(best.max().item()) * 0.5)[0]
    ref = orc.local_peaks(best.cpu().numpy()[0], 4, float(best.max().item()) * 0.5)
    assert np.array_equal(got, ref)


def test_reference_scenario_two_upscaled_deltas(gx, orc, cuda):
    """GTOMTest/LocalPeaks.cpp (which asserts nothing): deltas at (1, 1) and (3, 6) of
    an 8 x 8 image, Fourier-upscaled to 1769 x 1769 (d_Scale: R2C, d_FFTPad, C2R / 64;
    the 1769-point inverse here with numpy), mean0/std1, local peaks with extent 100
    and threshold 4.6: exactly the two deltas, within 2 voxels of their upscaled
    positions, the same list as the oracle."""
    import torch
    img = np.zeros((8, 8))
    img[1, 1] = img[6, 3] = 1.0
    half = np.fft.rfftn(img).astype(np.complex64)
    padded = to_np(gx.d_FFTResize("pad", torch.from_numpy(half[None]).to(cuda), (8, 8, 1), (1769, 1769, 1)))[0, 0]
    up = (np.fft.irfftn(padded.astype(np.complex128), s=(1769, 1769), axes=(0, 1)) * (1769 * 1769) / 64.0)
    up = up.astype(np.float32)
    normed = gx.d_NormMean01Std(torch.from_numpy(up).to(cuda))
    got = gx.d_LocalPeaks(normed.reshape(1, 1769, 1769), 100, 4.6)[0]
    ref = orc.local_peaks(orc.norm_mean01std(up).astype(np.float32).reshape(1, 1769, 1769), 100, 4.6)
    assert np.array_equal(got, ref)
    want = np.array([[1 * 1769 / 8, 1 * 1769 / 8], [3 * 1769 / 8, 6 * 1769 / 8]])
    assert len(got) == 2 and np.abs(got[:, :2] - want).max() <= 2
