"""ProbabilisticPicker and d_Rotate2DFT on the GPU against the FP64 oracle.

Tolerance: scores within 1e-4 normwise; the orientation identical wherever the
oracle's best beats its second-best orientation by more than 1e-4 * max|best|;
the reference index exact. d_Rotate2DFT within 1e-5 of the spectrum's maximum,
away from bins whose radius or flip test sits within 1e-3 of its threshold
(there the value is interpolated in float on the device and in double in the
oracle, which take the same float decision). Larger images (256^2, 64^3) are
checked against the oracle as well."""
import math

import numpy as np
import pytest

from tests._util import rel_linf, to_np

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("zc", [True, False])
def test_rotate2d_matches_oracle(gx, orc, cuda, zc):
    import torch
    rng = np.random.default_rng(3 + zc)
    n, nb = 32, 4
    half = (rng.standard_normal((nb, n, n // 2 + 1)) + 1j * rng.standard_normal((nb, n, n // 2 + 1))).astype(np.complex64)
    angles = rng.uniform(-math.pi, math.pi, nb).astype(np.float32)
    maxfreq = n / 2 - 0.5
    got = to_np(gx.d_Rotate2DFT(torch.from_numpy(half).to(cuda), angles, maxfreq, zc))
    y, x = np.meshgrid(np.arange(n) - n // 2, np.arange(n // 2 + 1) - n // 2, indexing="ij")
    for b in range(nb):
        want = orc.rotate2d_ft(half[b].astype(np.complex128), float(angles[b]), maxfreq, zc)
        c, s = math.cos(-angles[b]), math.sin(-angles[b])
        px, py = c * x - s * y, s * x + c * y
        ok = (np.abs(np.hypot(px, py) - maxfreq) > 1e-3) & (np.abs(px - 1e-5) > 1e-3)
        if not zc:  # the output bin of thread (idx, idy) moves to the FFT-order place
            ok = ok[(n - 1 - (np.arange(n) + n // 2 - 1) % n)[:, None], (n // 2 - np.arange(n // 2 + 1))[None, :]]
        assert np.abs(got[b] - want)[ok].max() <= 1e-5 * np.abs(half[b]).max()


def _disc(shape, radius):
    g = np.meshgrid(*[np.arange(s) - s // 2 for s in shape], indexing="ij")
    return (np.sqrt(sum(a.astype(float) ** 2 for a in g)) < radius).astype(np.float32)


def _run_gpu(gx, cuda, refs, masks, image, lp, step, ctf, n, best=None, ang=None, bref=None):
    import torch
    pp = gx.ProbabilisticPicker(torch.from_numpy(refs).to(cuda), torch.from_numpy(masks).to(cuda), image.shape, lp,
                                doctf=ctf is not None)
    pp.set_image(torch.from_numpy(image).to(cuda), None if ctf is None else torch.from_numpy(ctf).to(cuda))
    shape = image.shape
    b = torch.full(shape, -1e30, device=cuda) if best is None else best
    a = torch.zeros(shape + (3,), device=cuda) if ang is None else ang
    r = torch.full(shape, -1, dtype=torch.int32, device=cuda) if bref is None else bref
    pp.perform_correlation(n, step, b, a, r)
    torch.cuda.synchronize()
    nang = pp.count_angles(step)
    pp.close()
    return b, a, r, nang


@pytest.mark.parametrize("nd,doctf", [(2, False), (2, True), (3, False)])
def test_probabilistic_picker_matches_oracle(gx, orc, cuda, nd, doctf):
    rng = np.random.default_rng(60 + nd + doctf)
    if nd == 2:
        di, dr, lp, step = (64, 64), (16, 16), 8, 0.5  # 13 in-plane orientations
    else:
        di, dr, lp, step = (32, 32, 32), (12, 12, 12), 6, 1.2  # 50 orientations
    refs = rng.standard_normal((2,) + dr).astype(np.float32)
    masks = np.stack([_disc(dr, dr[0] / 2 - 1)] * 2)
    image = (rng.standard_normal(di) + 0.3).astype(np.float32)
    ctf = (0.1 + rng.uniform(0, 1, di[:-1] + (di[-1] // 2 + 1,))).astype(np.float32) if doctf else None
    b, a, r, nang = _run_gpu(gx, cuda, refs, masks, image, lp, step, ctf, 1)
    assert nang == (13 if nd == 2 else 50)
    orc.gaussrand(0, reset=True)
    ob, oa, oref, top2 = orc.prob_picker(refs[1].astype(np.float64), masks[1].astype(np.float64),
                                         image.astype(np.float64), lp, step,
                                         ctf=None if ctf is None else ctf.astype(np.float64), refid=1)
    gb = to_np(b).astype(np.float64)
    assert rel_linf(gb, ob) <= 1e-4
    assert np.all(to_np(r) == 1)
    sel = (ob - top2) > 1e-4 * np.abs(ob).max()
    assert sel.mean() > 0.9
    assert np.array_equal(to_np(a)[sel], oa[sel].astype(np.float32))


def test_probabilistic_picker_second_reference_updates_where_it_wins(gx, orc, cuda):
    """Two references in turn into the same outputs: the second replaces the first
    exactly where its score is higher (the C runtime's rand() is reseeded per call)."""
    rng = np.random.default_rng(71)
    di, dr, lp, step = (48, 48), (12, 12), 6, 0.7
    refs = rng.standard_normal((2,) + dr).astype(np.float32)
    masks = np.stack([_disc(dr, 5.0)] * 2)
    image = rng.standard_normal(di).astype(np.float32)
    b, a, r, _ = _run_gpu(gx, cuda, refs, masks, image, lp, step, None, 0)
    b, a, r, _ = _run_gpu(gx, cuda, refs, masks, image, lp, step, None, 1, b, a, r)
    orc.gaussrand(0, reset=True)
    b0, _, _, _ = orc.prob_picker(refs[0].astype(np.float64), masks[0].astype(np.float64), image.astype(np.float64),
                                  lp, step, refid=0)
    b1, _, _, _ = orc.prob_picker(refs[1].astype(np.float64), masks[1].astype(np.float64), image.astype(np.float64),
                                  lp, step, refid=1)
    want = np.maximum(b0, b1)
    assert rel_linf(to_np(b).astype(np.float64), want) <= 1e-4
    clear = np.abs(b1 - b0) > 1e-4 * np.abs(want).max()
    assert np.array_equal(to_np(r)[clear], (b1 > b0)[clear].astype(np.int32))


@pytest.mark.parametrize("nd", [2, 3])
def test_larger_images_match_oracle(gx, orc, cuda, nd):
    """256 x 256 images and 64^3 volumes. The expected probability ratio of an
    orientation comes from the cropped template, whose energy falls with the
    image size faster than its gaussrand term, so ratio - 1 is small and of
    either sign; the score divides by it, and the max over orientations mixes
    both signs. Parity is the check here, not a planted peak."""
    rng = np.random.default_rng(90 + nd)
    if nd == 2:
        di, dr, lp, step = (256, 256), (40, 40), 24, 0.25  # 26 orientations
    else:
        di, dr, lp, step = (64, 64, 64), (24, 24, 24), 16, 1.0  # 84 orientations
    refs = rng.standard_normal((1,) + dr).astype(np.float32)
    masks = _disc(dr, dr[0] / 2 - 2)[None]
    image = rng.standard_normal(di).astype(np.float32)
    b, a, r, nang = _run_gpu(gx, cuda, refs, masks, image, lp, step, None, 0)
    orc.gaussrand(0, reset=True)
    ob, oa, _, top2 = orc.prob_picker(refs[0].astype(np.float64), masks[0].astype(np.float64),
                                      image.astype(np.float64), lp, step, refid=0)
    assert rel_linf(to_np(b).astype(np.float64), ob) <= 1e-4
    sel = (ob - top2) > 1e-4 * np.abs(ob).max()
    assert sel.mean() > 0.9
    assert np.array_equal(to_np(a)[sel], oa[sel].astype(np.float32))
    assert np.all(to_np(r) == 0)


def test_gaussrand_phase_carries_between_calls(gx, orc, cuda):
    """5 * 225 * 13 draws per call is odd: the second call starts with the cos
    half of the previous call's last pair, as gaussrand's statics do. The two
    calls leave the phase even again."""
    rng = np.random.default_rng(95)
    di, dr, lp, step = (48, 48), (15, 15), 6, 0.5
    refs = rng.standard_normal((2,) + dr).astype(np.float32)
    masks = np.stack([_disc(dr, 6.0)] * 2)
    image = rng.standard_normal(di).astype(np.float32)
    b, a, r, nang = _run_gpu(gx, cuda, refs, masks, image, lp, step, None, 0)
    assert (5 * refs[0].size * nang) % 2 == 1
    b, a, r, _ = _run_gpu(gx, cuda, refs, masks, image, lp, step, None, 1, b, a, r)
    orc.gaussrand(0, reset=True)
    b0, _, _, _ = orc.prob_picker(refs[0].astype(np.float64), masks[0].astype(np.float64), image.astype(np.float64),
                                  lp, step, refid=0)
    b1, _, _, _ = orc.prob_picker(refs[1].astype(np.float64), masks[1].astype(np.float64), image.astype(np.float64),
                                  lp, step, refid=1)
    want = np.maximum(b0, b1)
    assert rel_linf(to_np(b).astype(np.float64), want) <= 1e-4
    # the second call with a fresh phase gives another answer
    orc.gaussrand(0, reset=True)
    b1f, _, _, _ = orc.prob_picker(refs[1].astype(np.float64), masks[1].astype(np.float64), image.astype(np.float64),
                                   lp, step, refid=1)
    assert rel_linf(b1f, b1) > 1e-3


@pytest.mark.parametrize("nd", [2, 3])
def test_noncircular_mask_matches_oracle(gx, orc, cuda, nd):
    """A non-circular (ellipsoidal) mask, rotated per orientation by cubic
    B-splines into FFT order, with per-orientation solvent statistics."""
    rng = np.random.default_rng(110 + nd)
    if nd == 2:
        di, dr, lp, step = (64, 64), (16, 16), 8, 0.5  # 13 orientations
    else:
        di, dr, lp, step = (32, 32, 32), (12, 12, 12), 6, 1.2  # 50 orientations
    g = np.meshgrid(*[np.arange(s) - s // 2 for s in dr], indexing="ij")
    radii = [3.0 + 1.5 * i for i in range(nd)]
    mask = (sum((a / r) ** 2 for a, r in zip(g, radii)) < 1).astype(np.float32)
    refs = rng.standard_normal((1,) + dr).astype(np.float32)
    image = rng.standard_normal(di).astype(np.float32)
    import torch
    pp = gx.ProbabilisticPicker(torch.from_numpy(refs).to(cuda), torch.from_numpy(mask[None]).to(cuda), di, lp,
                                ismaskcircular=False)
    pp.set_image(torch.from_numpy(image).to(cuda))
    b = torch.full(di, -1e30, device=cuda)
    a = torch.zeros(di + (3,), device=cuda)
    r = torch.full(di, -1, dtype=torch.int32, device=cuda)
    pp.perform_correlation(0, step, b, a, r)
    torch.cuda.synchronize()
    pp.close()
    orc.gaussrand(0, reset=True)
    ob, oa, _, top2 = orc.prob_picker(refs[0].astype(np.float64), mask.astype(np.float64), image.astype(np.float64),
                                      lp, step, refid=0, circular=False)
    assert rel_linf(to_np(b).astype(np.float64), ob) <= 1e-4
    sel = (ob - top2) > 1e-4 * np.abs(ob).max()
    assert sel.mean() > 0.9
    assert np.array_equal(to_np(a)[sel], oa[sel].astype(np.float32))
    assert np.all(to_np(r) == 0)


@pytest.mark.parametrize("di", [(1024, 1024), (1536, 1536)])
def test_micrograph_sizes_match_oracle(gx, orc, cuda, di):
    """Micrograph-sized images, whose transforms run in two passes per axis
    (1024 = 32 x 32 and 1536 = 48 x 32 along y; rows of 512 and 768 complex)."""
    rng = np.random.default_rng(di[0])
    dr, lp, step = (32, 32), 16, 0.5  # 13 orientations
    refs = rng.standard_normal((1,) + dr).astype(np.float32)
    masks = _disc(dr, 14.0)[None]
    image = rng.standard_normal(di).astype(np.float32)
    b, a, r, nang = _run_gpu(gx, cuda, refs, masks, image, lp, step, None, 0)
    assert nang == 13
    orc.gaussrand(0, reset=True)
    ob, oa, _, top2 = orc.prob_picker(refs[0].astype(np.float64), masks[0].astype(np.float64),
                                      image.astype(np.float64), lp, step, refid=0)
    assert rel_linf(to_np(b).astype(np.float64), ob) <= 1e-4
    sel = (ob - top2) > 1e-4 * np.abs(ob).max()
    assert sel.mean() > 0.9
    assert np.array_equal(to_np(a)[sel], oa[sel].astype(np.float32))
    assert np.all(to_np(r) == 0)


def test_bench_size_3d_matches_oracle(gx, orc, cuda):
    """The measured 3D configuration: a 128^3 volume, a masked 32^3 reference,
    low-pass 24 (rotation at 48^3, the first C2R over its 25 nonzero columns),
    here at anglestep 1.2 (50 orientations)."""
    rng = np.random.default_rng(128)
    di, dr, lp, step = (128, 128, 128), (32, 32, 32), 24, 1.2
    mask = _disc(dr, 14.0)
    refs = (rng.standard_normal((1,) + dr) * mask).astype(np.float32)
    image = rng.standard_normal(di).astype(np.float32)
    b, a, r, nang = _run_gpu(gx, cuda, refs, mask[None], image, lp, step, None, 0)
    assert nang == 50
    orc.gaussrand(0, reset=True)
    ob, oa, _, top2 = orc.prob_picker(refs[0].astype(np.float64), mask.astype(np.float64), image.astype(np.float64),
                                      lp, step, refid=0)
    assert rel_linf(to_np(b).astype(np.float64), ob) <= 1e-4
    sel = (ob - top2) > 1e-4 * np.abs(ob).max()
    assert sel.mean() > 0.9
    assert np.array_equal(to_np(a)[sel], oa[sel].astype(np.float32))


@pytest.mark.parametrize("nd", [2, 3])
def test_non_native_sizes_match_oracle(gx, orc, cuda, nd):
    """Images of 100 x 90 and volumes of 30 x 30 x 40 with references of 20^2 / 10^3: extents without
    a native transform take the chirp-z form with full (not banded) transforms."""
    rng = np.random.default_rng(70 + nd)
    if nd == 2:
        di, dr, lp, step = (90, 100), (20, 20), 10, 0.5
    else:
        di, dr, lp, step = (40, 30, 30), (10, 10, 10), 5, 1.2
    refs = rng.standard_normal((1,) + dr).astype(np.float32)
    masks = _disc(dr, dr[0] / 2 - 1)[None]
    image = rng.standard_normal(di).astype(np.float32)
    b, a, r, nang = _run_gpu(gx, cuda, refs, masks, image, lp, step, None, 0)
    orc.gaussrand(0, reset=True)
    ob, oa, _, top2 = orc.prob_picker(refs[0].astype(np.float64), masks[0].astype(np.float64),
                                      image.astype(np.float64), lp, step, refid=0)
    err = rel_linf(to_np(b).astype(np.float64), ob)
    print(f"ProbabilisticPicker {di} / {dr}: max error {err:.2e} over {nang} orientations")
    assert err <= 1e-4
    sel = (ob - top2) > 1e-4 * np.abs(ob).max()
    assert sel.mean() > 0.9
    assert np.array_equal(to_np(a)[sel], oa[sel].astype(np.float32))
