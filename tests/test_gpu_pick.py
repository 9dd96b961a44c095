"""End-to-end parity of d_PickSubTomograms on the GPU against the FP64 oracle.

Tolerance (stated once for every test here): the normalised max-CC map must
agree with the FP64 oracle to 1e-4 normwise (max|gpu - fp64| / max|fp64|), and
the argmax orientation must be identical wherever the oracle's top-2 margin
exceeds 1e-4 * max|top1|.
"""
import numpy as np
import pytest

from tests._util import argmax_agreement, rel_linf, to_np

pytestmark = pytest.mark.gpu

CC_TOL = 1e-4


def _run_both(gx, orc, cuda, w, ctf_list, nang=None, batch=0):
    import torch
    n, p, s = w["n"], w["p"], w["oversample"]
    tomos = w["tomos"]
    angles = w["angles"] if nang is None else w["angles"][:nang]
    template = torch.from_numpy(w["template"]).to(cuda)
    projft = gx.prepare_projector(template, p)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(tomos).to(cuda))
    ctf = torch.from_numpy(np.stack(ctf_list).astype(np.float32)).to(cuda)
    stride = ctf[0].numel() if len(ctf_list) > 1 else 0
    plan = gx.PickPlan(p, n, batch)
    best, idx = gx.d_PickSubTomograms(plan, projft, s, exp_ft, ctf, angles, w["maskradius"],
                                      ctf_volume_stride=stride)
    torch.cuda.synchronize()
    # the oracle gets the same float32 inputs, promoted to double
    P = to_np(projft).astype(np.complex128)
    E = to_np(exp_ft).astype(np.complex128)
    C = np.stack(ctf_list).astype(np.float32).astype(np.float64)
    ref = orc.pick(P, p, s, E, C if len(ctf_list) > 1 else C[0], n, angles.astype(np.float64),
                   w["maskradius"], ctf_per_volume=len(ctf_list) > 1)
    plan.close()
    return to_np(best), to_np(idx), ref


def _check(best, idx, ref):
    err = rel_linf(best, ref["best"])
    assert err <= CC_TOL, err
    agree, cnt = argmax_agreement(idx, ref["idx"], ref["top1"], ref["top2"], CC_TOL)
    assert cnt > 0
    assert agree == 1.0, (agree, cnt)


def _small_workload(gx, n=32, box=16, seed=4, nvol=1):
    from gtom_amd import synth
    out = []
    for v in range(nvol):
        w = synth.cfg1(seed=seed + v, n=n, box=box, ncopies=2, snr=0.3, maskradius=n / 4)
        out.append(w)
    w = dict(out[0])
    w["tomos"] = np.stack([o["tomogram"] for o in out])
    return w


def test_pick_small_matches_oracle(gx, orc, cuda):
    w = _small_workload(gx)
    ctf = [synth_ctf(orc, w["n"], -3e-6)]
    best, idx, ref = _run_both(gx, orc, cuda, w, ctf, nang=100)
    _check(best, idx, ref)


def test_pick_two_volumes_per_volume_ctf(gx, orc, cuda):
    w = _small_workload(gx, nvol=2, seed=9)
    ctf = [synth_ctf(orc, w["n"], -2.5e-6), synth_ctf(orc, w["n"], -4e-6)]
    best, idx, ref = _run_both(gx, orc, cuda, w, ctf, nang=60, batch=7)
    _check(best, idx, ref)


@pytest.mark.parametrize("factor", [0.8, 1.5])
def test_pick_other_oversampling(gx, orc, cuda, factor):
    """Oversampling other than p/n. At 0.8 * p/n every sample stays inside the
    projector (the centred projector layout); at 1.5 * p/n samples run past
    its Nyquist and read the reference's wrapped and linear addresses (the
    reference's own layout)."""
    w = _small_workload(gx, seed=13)
    w["oversample"] = w["oversample"] * factor
    ctf = [synth_ctf(orc, w["n"], -3e-6)]
    best, idx, ref = _run_both(gx, orc, cuda, w, ctf, nang=40)
    _check(best, idx, ref)


@pytest.mark.parametrize("batch", [1, 5, 16])
def test_pick_batch_sizes_agree(gx, orc, cuda, batch):
    """Results do not depend on the number of orientations per pass (bit-exact)."""
    import torch
    w = _small_workload(gx, seed=21)
    ctf = [synth_ctf(orc, w["n"], -3e-6)]
    b1, i1, ref = _run_both(gx, orc, cuda, w, ctf, nang=37, batch=batch)
    _check(b1, i1, ref)
    b2, i2, _ = _run_both(gx, orc, cuda, w, ctf, nang=37, batch=3)
    assert np.array_equal(i1, i2)
    assert np.array_equal(b1, b2)


def test_pick_cfg1_matches_oracle(gx, orc, cuda):
    """The 64^3 configuration: 3 planted copies, 30-degree search (588 orientations)."""
    from gtom_amd import synth
    w = synth.cfg1(seed=1)
    w["tomos"] = w["tomogram"][None]
    ctf = [synth_ctf(orc, 64, -3e-6)]
    best, idx, ref = _run_both(gx, orc, cuda, w, ctf)
    assert ref["idx"].shape == (1, 64, 64, 64)
    _check(best, idx, ref)


def test_pick_resume_equals_single_call(gx, cuda):
    """begin; update(first k); update(rest, offset k); finish == one call, bit for bit."""
    import torch
    from gtom_amd import synth
    w = synth.cfg1(seed=3, n=32, box=16, ncopies=2, snr=0.3, maskradius=8.0)
    n, p, s = w["n"], w["p"], w["oversample"]
    ang = w["angles"][:50]
    projft = gx.prepare_projector(torch.from_numpy(w["template"]).to(cuda), p)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(w["tomogram"][None]).to(cuda))
    ctf = torch.ones((n, n, n // 2 + 1), device=cuda)
    plan = gx.PickPlan(p, n, 4)
    b1, i1 = gx.d_PickSubTomograms(plan, projft, s, exp_ft, ctf, ang, 8.0)
    b2 = torch.empty_like(b1)
    i2 = torch.empty_like(i1)
    gx.pick_begin(plan, 1, b2, i2)
    gx.pick_update(plan, projft, s, exp_ft, ctf, ang[:23], 0, b2, i2)
    gx.pick_update(plan, projft, s, exp_ft, ctf, ang[23:], 23, b2, i2)
    ba = torch.empty_like(b1)
    gx.pick_finish(plan, exp_ft, 8.0, b2, i2, ba)
    torch.cuda.synchronize()
    assert torch.equal(i1, i2)
    assert torch.equal(b1, b2)
    assert torch.equal(ba, i2.float())
    plan.close()


def test_planted_copy_found_at_position_and_index(gx, cuda):
    """A copy planted at a search orientation peaks at its centre with that orientation's index."""
    import torch
    from gtom_amd import synth
    rng = np.random.default_rng(8)
    n, box = 64, 32
    angles = gx.equal_angular_spacing((0, 2 * np.pi), (0, np.pi), (0, 2 * np.pi), np.deg2rad(30))
    target = 137
    template = synth.blob_template(box, rng)
    pos = (30, 25, 33)
    sig = synth.plant((n, n, n), template, [pos], [tuple(float(a) for a in angles[target])], 1.0)
    tomo = (sig + 0.05 * rng.standard_normal((n, n, n))).astype(np.float32)
    projft = gx.prepare_projector(torch.from_numpy(template).to(cuda), 2 * box)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(tomo[None]).to(cuda))
    ctf = torch.ones((n, n, n // 2 + 1), device=cuda)
    plan = gx.PickPlan(2 * box, n)
    best, idx = gx.d_PickSubTomograms(plan, projft, 1.0, exp_ft, ctf, angles, 14.0)
    b = to_np(best)[0]
    z, y, x = np.unravel_index(np.argmax(b), b.shape)
    assert (x, y, z) == pos
    assert to_np(idx)[0, z, y, x] == target
    plan.close()


def synth_ctf(orc, n, defocus):
    return orc.ctf3(n, 10e-10, 300e3, 2.7e-3, defocus, 0.07)
