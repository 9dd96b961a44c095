"""Micrographs of uneven sizes in one call (gtomx_mpick): one plan per side,
the per-side orientation loop captured once as a hipGraph and replayed per
image, each image with its own astigmatic CTF.

- Every image's result equals gtomx_pick_subtomograms on that image alone, bit
  for bit (same kernels), with graphs and without; the graph of a side is
  replayed once per image of that side with the object's eager copies and the
  local-std division between replays, and a second call replays them again.
- A reduced uneven set (1024^2 and 2048^2 micrographs, 40 orientations) against
  the FP64 oracle: max-CC within 1e-4 normwise, argmax identical wherever the
  oracle's top-2 margin exceeds 1e-4 * max|top1| (the search's bar).
- The autotuner reports a finite time per orientation for every candidate batch
  that fits.
The reference batches only images of one size (src/Correlation/SubTomograms.cu:34-36)."""
import numpy as np
import pytest

from tests._util import argmax_agreement, rel_linf, to_np

pytestmark = pytest.mark.gpu

CC_TOL = 1e-4


def _blob(box, seed):
    rng = np.random.default_rng(seed)
    g = np.arange(box) - box // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    t = np.zeros((box, box, box))
    for _ in range(8):
        c = rng.uniform(-box / 4, box / 4, 3)
        s = rng.uniform(1.0, 2.5, 3)
        t += np.exp(-(((x - c[0]) / s[0]) ** 2 + ((y - c[1]) / s[1]) ** 2 + ((z - c[2]) / s[2]) ** 2) / 2)
    return t.astype(np.float32)


def _images(gx, cuda, sides, seed, pixelsize=1e-10):
    import torch
    rng = np.random.default_rng(seed)
    out = []
    for k, n in enumerate(sides):
        img = torch.from_numpy(rng.standard_normal((1, n, n)).astype(np.float32)).to(cuda)
        E = gx.d_FFTR2C(img, ndims=2)[0].contiguous()
        par = gx.CtfParamsFull.make(pixelsize=pixelsize, Cs=2.7e-3, voltage=300e3,
                                    defocus=-float(rng.uniform(0.8e-6, 3.5e-6)),
                                    defocusdelta=float(rng.uniform(0, 100e-9)),
                                    astigmatismangle=float(rng.uniform(0, np.pi)), amplitude=0.07)
        ctf = gx.ctf_2d([par], n, device=cuda)[0].contiguous()
        out.append({"spectrum": E, "ctf": ctf, "maskradius": n / 8,
                    "best": torch.empty((n, n), device=cuda), "index": torch.empty((n, n), dtype=torch.int32,
                                                                                   device=cuda)})
    return out


def _alone(gx, p, projft, im, angles, batch):
    import torch
    n = im["spectrum"].shape[0]
    plan = gx.PickPlan(p, n, batch, ndims=2)
    best, idx = gx.d_PickSubTomograms(plan, projft, p / n, im["spectrum"][None], im["ctf"][None], angles,
                                      im["maskradius"])
    torch.cuda.synchronize()
    plan.close()
    return to_np(best)[0], to_np(idx)[0]


@pytest.mark.parametrize("use_graphs", [True, False])
def test_mpick_equals_each_image_alone(gx, cuda, use_graphs):
    import torch
    box = 16
    p = 2 * box
    projft = gx.prepare_projector(torch.from_numpy(_blob(box, 3)).to(cuda), p)
    sides = [512, 768, 512, 256, 512, 768]
    images = _images(gx, cuda, sides, seed=5)
    angles = gx.fibonacci_angles(10, 6)
    batches = {256: 24, 512: 16, 768: 8}
    # the long-row sides run the prefetching row kernel here, the plain one in _alone: same results
    mb = gx.MicrographBatch(p, [256, 512, 768], [batches[n] for n in (256, 512, 768)],
                            [gx.ROWS_PLAIN, gx.ROWS_PREFETCH, gx.ROWS_PREFETCH])
    for call in range(2):
        for im in images:
            im["best"].fill_(float("nan"))
        mb.run(projft, images, angles, use_graphs=use_graphs)
        torch.cuda.synchronize()
        for k, im in enumerate(images):
            wb, wi = _alone(gx, p, projft, im, angles, batches[sides[k]])
            assert np.array_equal(to_np(im["best"]), wb), (call, k, sides[k])
            assert np.array_equal(to_np(im["index"]), wi), (call, k, sides[k])
    info = mb.info()
    mb.close()
    assert info["variants"] == {256: gx.ROWS_PLAIN, 512: gx.ROWS_PREFETCH, 768: gx.ROWS_PREFETCH}, info
    if use_graphs:
        assert info["captures"] == 3 and info["replays"] == 2 * len(sides), info
    else:
        assert info["captures"] == 0 and info["replays"] == 0, info


@pytest.mark.parametrize("variants", [(0, 0), (1, 1), (0, 2)])
def test_mpick_uneven_set_matches_oracle(gx, orc, cuda, variants):
    import torch
    box = 32
    p = 2 * box
    projft = gx.prepare_projector(torch.from_numpy(_blob(box, 7)).to(cuda), p)
    sides = [2048, 1024]
    images = _images(gx, cuda, sides, seed=8)
    angles = gx.fibonacci_angles(8, 5)
    mb = gx.MicrographBatch(p, [1024, 2048], variants=list(variants))
    mb.run(projft, images, angles)
    torch.cuda.synchronize()
    mb.close()
    P = to_np(projft).astype(np.complex128)
    for k, im in enumerate(images):
        n = sides[k]
        ref = orc.pick(P, p, p / n, to_np(im["spectrum"])[None].astype(np.complex128),
                       to_np(im["ctf"]).astype(np.float64), n, angles.astype(np.float64), im["maskradius"], ndims=2)
        best, idx = to_np(im["best"]), to_np(im["index"])
        err = rel_linf(best, ref["best"][0])
        agree, cnt = argmax_agreement(idx, ref["idx"][0], ref["top1"][0], ref["top2"][0], CC_TOL)
        print(f"mpick {n}^2 rows variants {variants}: max-CC error {err:.2e}, argmax agreement {agree} on {cnt} pixels")
        assert err <= CC_TOL, (n, err)
        assert cnt > 0 and agree == 1.0, (n, agree, cnt)


def test_mpick_tune_times_every_candidate(gx, cuda):
    import torch
    box = 16
    p = 2 * box
    projft = gx.prepare_projector(torch.from_numpy(_blob(box, 4)).to(cuda), p)
    ms = gx.mpick_tune(p, 1024, projft, [4, 8, 16, 8], gx.fibonacci_angles(4, 4), repeats=1,
                       variants=[gx.ROWS_PLAIN] * 3 + [gx.ROWS_PREFETCH])
    print("mpick tune 1024^2 (ms per orientation, batches 4/8/16 plain, 8 prefetch):", ms)
    assert all(np.isfinite(m) and m > 0 for m in ms), ms


def test_rows_variant_rejects_unknown_value(gx, cuda):
    plan = gx.PickPlan(32, 512, 8, ndims=2)
    plan.set_rows_variant(gx.ROWS_PREFETCH)
    with pytest.raises(RuntimeError, match="rows variant"):
        plan.set_rows_variant(3)
    with pytest.raises(RuntimeError, match="4096"):
        plan.set_rows_variant(gx.ROWS_ONE)  # the one-row pass takes 2048^2, 3072^2 and 4096^2 images only
    plan.close()
