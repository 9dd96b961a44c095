"""d_PickSubTomograms at the benchmark's block size: 256^3 blocks searched with
a 48^3 template through its 96^3 projector at oversample 0.375 (the centred
sampler and the occupancy bound that only n = 256 uses), against the FP64 oracle on 40 orientations in two batches,
and by a size-independent property over more orientations: a copy planted at
a search orientation peaks at its centre with that orientation's index.

Tolerance as for every search test: max-CC within 1e-4 normwise, argmax
identical wherever the oracle's top-2 margin exceeds 1e-4 * max|top1|."""
import numpy as np
import pytest

from tests._util import argmax_agreement, rel_linf, to_np

pytestmark = pytest.mark.gpu


def test_full_size_block_matches_oracle(gx, orc, cuda):
    import torch
    from gtom_amd import synth
    w = synth.cfg2(seed=5, shape=(256, 256, 256), ncopies=6)
    n, p, s = w["n"], w["p"], w["oversample"]
    angles = w["angles"][::225][:40]  # spread over the 9000 directions x psi; a full batch of 32 and 8
    ctf_np = synth.ctf3_numpy(n, pixelsize=w["pixelsize"], defocus=-w["defoci"][0]).astype(np.float32)
    projft = gx.prepare_projector(torch.from_numpy(w["template"]).to(cuda), p)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(w["blocks"]).to(cuda))
    ctf = torch.from_numpy(ctf_np).to(cuda)
    plan = gx.PickPlan(p, n, 32)  # a full batch of 32 and a partial one of 8
    best, idx = gx.d_PickSubTomograms(plan, projft, s, exp_ft, ctf, angles, w["maskradius"])
    torch.cuda.synchronize()
    plan.close()
    ref = orc.pick(to_np(projft).astype(np.complex128), p, s, to_np(exp_ft).astype(np.complex128),
                   ctf_np.astype(np.float64), n, angles.astype(np.float64), w["maskradius"])
    assert rel_linf(to_np(best), ref["best"]) <= 1e-4
    agree, cnt = argmax_agreement(to_np(idx), ref["idx"], ref["top1"], ref["top2"], 1e-4)
    assert cnt > 0 and agree == 1.0, (agree, cnt)


def test_full_size_planted_copy_found(gx, cuda):
    import torch
    from gtom_amd import synth
    rng = np.random.default_rng(21)
    n, box = 256, 48
    p = 2 * box
    angles = gx.fibonacci_angles(500, 18)[::141][:64]  # 64 orientations, two batches of 32
    target = 45
    template = synth.blob_template(box, rng, radius=14.0)
    pos = (131, 77, 190)
    sig = synth.plant((n, n, n), template, [pos], [tuple(float(a) for a in angles[target])], 1.0)
    tomo = (sig + 0.05 * rng.standard_normal((n, n, n))).astype(np.float32)
    projft = gx.prepare_projector(torch.from_numpy(template).to(cuda), p)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(tomo[None]).to(cuda))
    ctf = torch.ones((n, n, n // 2 + 1), device=cuda)
    plan = gx.PickPlan(p, n, 32)
    best, idx = gx.d_PickSubTomograms(plan, projft, p / n, exp_ft, ctf, angles, 24.0)
    torch.cuda.synchronize()
    plan.close()
    b = to_np(best)[0]
    z, y, x = np.unravel_index(np.argmax(b), b.shape)
    assert (x, y, z) == pos
    assert to_np(idx)[0, z, y, x] == target


def test_tail_batches_with_per_volume_ctfs_match_oracle(gx, orc, cuda):
    import torch
    from gtom_amd import synth
    w = synth.cfg2(seed=7)
    n, p, s = w["n"], w["p"], w["oversample"]
    H = (n // 2 + 1) * n * n
    assert w["blocks"].shape[0] == 2 and len(w["angles"]) == 9000
    off = 8960
    angles = w["angles"][off:]
    ctf_np = np.stack([synth.ctf3_numpy(n, pixelsize=w["pixelsize"], defocus=-d) for d in w["defoci"]]).astype(
        np.float32)
    projft = gx.prepare_projector(torch.from_numpy(w["template"]).to(cuda), p)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(w["blocks"]).to(cuda))
    ctf = torch.from_numpy(ctf_np).to(cuda)
    plan = gx.PickPlan(p, n, 32)  # a full batch 8960..8991 and a partial one 8992..8999
    assert plan.batch == 32 and (len(angles) % plan.batch) == 8
    best = torch.empty((2, n, n, n), dtype=torch.float32, device=cuda)
    idx = torch.empty((2, n, n, n), dtype=torch.int32, device=cuda)
    gx.pick_begin(plan, 2, best, idx)
    gx.pick_update(plan, projft, s, exp_ft, ctf, angles, off, best, idx, ctf_volume_stride=H)
    gx.pick_finish(plan, exp_ft, w["maskradius"], best, idx)
    torch.cuda.synchronize()
    plan.close()
    ref = orc.pick(to_np(projft).astype(np.complex128), p, s, to_np(exp_ft).astype(np.complex128),
                   ctf_np.astype(np.float64), n, angles.astype(np.float64), w["maskradius"], ctf_per_volume=True)
    for v in range(2):
        assert rel_linf(to_np(best)[v], ref["best"][v]) <= 1e-4, v
        agree, cnt = argmax_agreement(to_np(idx)[v], ref["idx"][v] + off, ref["top1"][v], ref["top2"][v], 1e-4)
        assert cnt > 0 and agree == 1.0, (v, agree, cnt)
    got_idx = to_np(idx)
    assert got_idx.min() >= off and got_idx.max() <= off + 39 and (got_idx >= 8992).any()


def test_benchmark_call_shape_matches_oracle(gx, orc, cuda):
    """The benchmark's own call shape: the two 256^3 blocks of the workload with one CTF each
    (gtomx_ctf_3d, ctf_volume_stride = H), the default batch of 128, and the last 40 of the 9000
    orientations at their own indices: the final partial batch 8960..8999, as the whole-list call
    splits them (8960 = 70 x 128). The oracle gets the same CTF array per volume."""
    import torch
    from gtom_amd import synth
    w = synth.cfg2(seed=9)
    n, p, s = w["n"], w["p"], w["oversample"]
    H = (n // 2 + 1) * n * n
    first = 8960
    angles = w["angles"][first:]
    assert len(w["angles"]) == 9000 and len(angles) == 40
    projft = gx.prepare_projector(torch.from_numpy(w["template"]).to(cuda), p)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(w["blocks"]).to(cuda))
    ctf = gx.ctf_3d([gx.CtfParams.make(pixelsize=w["pixelsize"], Cs=2.7e-3, voltage=300e3, defocus=-float(d),
                                       amplitude=0.07) for d in w["defoci"]], n, device=cuda)
    plan = gx.PickPlan(p, n)
    assert plan.batch == 128 and 9000 % plan.batch == len(angles)
    nvol = w["blocks"].shape[0]
    best = torch.empty((nvol, n, n, n), dtype=torch.float32, device=cuda)
    idx = torch.empty((nvol, n, n, n), dtype=torch.int32, device=cuda)
    gx.pick_begin(plan, nvol, best, idx)
    gx.pick_update(plan, projft, s, exp_ft, ctf, angles, first, best, idx, ctf_volume_stride=H)
    gx.pick_finish(plan, exp_ft, w["maskradius"], best, idx)
    torch.cuda.synchronize()
    plan.close()
    ref = orc.pick(to_np(projft).astype(np.complex128), p, s, to_np(exp_ft).astype(np.complex128),
                   to_np(ctf).reshape(nvol, -1).astype(np.float64), n, angles.astype(np.float64), w["maskradius"],
                   ctf_per_volume=True)
    b, i = to_np(best), to_np(idx)
    assert i.min() >= first and i.max() < 9000
    for v in range(nvol):
        err = rel_linf(b[v], ref["best"][v])
        agree, cnt = argmax_agreement(i[v] - first, ref["idx"][v], ref["top1"][v], ref["top2"][v], 1e-4)
        print(f"block {v}: max-CC error {err:.2e}, argmax agreement {agree} on {cnt} voxels")
        assert err <= 1e-4, (v, err)
        assert cnt > 0 and agree == 1.0, (v, agree, cnt)
