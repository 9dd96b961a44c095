"""d_PickSubTomograms on volume sides without a fused search (every other side
up to 1024: 40, 45, 63, 100, 200, 384, ...; odd sides have no x = n/2 plane): pass 1 over the whole half spectrum,
the plane correction and scale as in the fused form, the z and y inverses and
the row C2R by the chirp-z transforms (native where the length has one), and
the in-order max/argmax over the real maps. The reference takes any size cuFFT
plans (src/Correlation/SubTomograms.cu:19-125).

- 40^3, 100^3 and the odd 45^3 and 63^3 against the FP64 oracle over two batches, with a CTF: max-CC
  within 1e-4 normwise, argmax identical wherever the oracle's top-2 margin
  exceeds 1e-4 * max|top1|.
- 200^3 and 384^3, a size-independent property: a copy planted at a search
  orientation peaks at its centre with that orientation's index."""
import numpy as np
import pytest

from tests._util import argmax_agreement, rel_linf, to_np

pytestmark = pytest.mark.gpu


def _volume(rng, n, template, pos, angle, amplitude, noise):
    from gtom_amd import synth
    sig = synth.plant((n, n, n), template, [pos], [tuple(float(a) for a in angle)], amplitude)
    return (sig + noise * rng.standard_normal((n, n, n), dtype=np.float32)).astype(np.float32)


@pytest.mark.parametrize("n,box,nang,batch", [(40, 10, 5, 3), (100, 16, 4, 3), (45, 10, 4, 3), (63, 12, 3, 2)])
def test_generic_side_matches_oracle(gx, orc, cuda, n, box, nang, batch):
    import torch
    from gtom_amd import synth
    rng = np.random.default_rng(n)
    p = 2 * box
    template = synth.blob_template(box, rng, radius=box / 3)
    angles = gx.fibonacci_angles(100, 6)[rng.choice(600, nang, replace=False)]
    tomo = _volume(rng, n, template, (n // 2 + 3, n // 3, n // 2 - 5), angles[1], 1.0, 1.0)
    ctf_np = synth.ctf3_numpy(n, pixelsize=10e-10, defocus=-3e-6).astype(np.float32)
    projft = gx.prepare_projector(torch.from_numpy(template).to(cuda), p)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(tomo[None]).to(cuda))
    ctf = torch.from_numpy(ctf_np).to(cuda)
    plan = gx.PickPlan(p, n, batch=batch)
    best, idx = gx.d_PickSubTomograms(plan, projft, p / n, exp_ft, ctf, angles, n / 6)
    torch.cuda.synchronize()
    plan.close()
    ref = orc.pick(to_np(projft).astype(np.complex128), p, p / n, to_np(exp_ft).astype(np.complex128),
                   ctf_np.astype(np.float64), n, angles.astype(np.float64), n / 6)
    err = rel_linf(to_np(best), ref["best"])
    agree, cnt = argmax_agreement(to_np(idx), ref["idx"], ref["top1"], ref["top2"], 1e-4)
    print(f"{n}^3 x {nang} orientations: max-CC error {err:.2e}, argmax agreement {agree} on {cnt} voxels")
    assert err <= 1e-4
    assert cnt > 0 and agree == 1.0, (agree, cnt)


@pytest.mark.parametrize("n,box", [(200, 24), (384, 32)])
def test_generic_side_planted_copy_found(gx, cuda, n, box):
    import torch
    from gtom_amd import synth
    rng = np.random.default_rng(n + 1)
    p = 2 * box
    angles = gx.fibonacci_angles(500, 18)[::300][:24]
    target = 17
    template = synth.blob_template(box, rng, radius=box / 3)
    pos = (n // 2 - 11, n // 3, 2 * n // 3)
    tomo = _volume(rng, n, template, pos, angles[target], 1.0, 0.05)
    projft = gx.prepare_projector(torch.from_numpy(template).to(cuda), p)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(tomo[None]).to(cuda))
    del tomo
    ctf = torch.ones((n, n, n // 2 + 1), device=cuda)
    plan = gx.PickPlan(p, n, batch=10)  # batches of 10, 10 and 4
    best, idx = gx.d_PickSubTomograms(plan, projft, p / n, exp_ft, ctf, angles, box / 2)
    torch.cuda.synchronize()
    plan.close()
    b = to_np(best)[0]
    z, y, x = np.unravel_index(np.argmax(b), b.shape)
    assert (x, y, z) == pos
    assert to_np(idx)[0, z, y, x] == target
