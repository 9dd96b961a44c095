"""The learned search (gtomx_pick_subtomograms_learned) end to end.

- With confidence 1 every tile falls back to the exhaustive search (no softmax
  maximum of an untrained network reaches 1), and the result equals
  d_PickSubTomograms bit for bit: the same kernels run over every direction
  group with all tile bits set.
- With confidence 0 and a network that puts every tile in the "no peak" class,
  nothing is searched: every voxel holds 0 and index -1, and no group runs.
- With the committed trained network, on held-out blocks of the learned
  configuration (SNR 0.3), at the recall operating point: at least 99% of the
  exhaustive search's peaks keep their exact value and orientation while the
  network skips tiles, and its detection beats a scorer without information.
The labels come from UpdateTomoCorrKernel's argmax semantics
(src/Correlation/SubTomograms.cu:138-158); the network has no reference counterpart."""
import os

import numpy as np
import pytest

from tests._util import to_np

pytestmark = pytest.mark.gpu


def _small(gx, cuda, npsi=1):
    import torch
    from gtom_amd import synth
    w = synth.cfg1(seed=11, n=64, box=16, ncopies=2, snr=0.5, maskradius=8.0)
    n, p, s = w["n"], w["p"], w["oversample"]
    projft = gx.prepare_projector(torch.from_numpy(w["template"]).to(cuda), p)
    E = gx.d_FFTR2C(torch.from_numpy(w["tomogram"][None]).to(cuda))
    ctf = torch.from_numpy(synth.ctf3_numpy(n, pixelsize=10e-10, defocus=-3e-6).astype(np.float32)).to(cuda)
    angles = gx.fibonacci_angles(500, npsi)
    return w, projft, E, ctf, angles


def test_learned_all_fallback_equals_exhaustive(gx, cuda):
    import torch
    from gtom_amd import proposal as pr
    w, projft, E, ctf, angles = _small(gx, cuda)
    plan = gx.PickPlan(w["p"], w["n"], 8)
    b0, i0 = gx.d_PickSubTomograms(plan, projft, w["oversample"], E, ctf, angles, w["maskradius"])
    net = pr.ProposalNet(max_tiles=8, seed=1)
    b1, i1, st = gx.d_PickSubTomogramsLearned(plan, net, projft, w["oversample"], E, ctf, angles, w["maskradius"],
                                              npsi=1, topk=16, confidence=1.0)
    torch.cuda.synchronize()
    net.close()
    plan.close()
    assert st["exhaustive"] == 8 and st["groups_searched"] == 500, st
    assert np.array_equal(to_np(b0), to_np(b1))
    assert np.array_equal(to_np(i0), to_np(i1))


def test_learned_confident_no_peak_searches_nothing(gx, cuda):
    import torch
    from gtom_amd import proposal as pr
    w, projft, E, ctf, angles = _small(gx, cuda)
    params = pr.init_params(2)
    off, ln = pr.LAYERS["bf"]
    params[off + pr.NO_PEAK] = 100.0  # the "no peak" logit dominates every tile
    net = pr.ProposalNet(max_tiles=8, params=params)
    plan = gx.PickPlan(w["p"], w["n"], 8)
    b, i, st = gx.d_PickSubTomogramsLearned(plan, net, projft, w["oversample"], E, ctf, angles, w["maskradius"],
                                            npsi=1, confidence=0.0)
    torch.cuda.synchronize()
    net.close()
    plan.close()
    assert st["skipped"] == 8 and st["groups_searched"] == 0, st
    assert np.all(to_np(b) == 0) and np.all(to_np(i) == -1)


_WEIGHTS = os.path.join(os.path.dirname(__file__), "..", "gtom_amd", "proposal_weights.npz")


@pytest.mark.skipif(not os.path.exists(_WEIGHTS), reason="no trained network committed")
def test_trained_network_recall_on_heldout_block(gx, cuda):
    """At the recall operating point (a tile is skipped only when the network gives it "no peak"
    with at least the calibrated confidence; every other tile is searched over every direction),
    a held-out block keeps at least 99% of the exhaustive search's peaks with their exact value and
    orientation, while the network removes tiles from the search."""
    from gtom_amd import proposal as pr
    from gtom_amd import train_proposal as tp
    w = tp.Workload()
    net = pr.ProposalNet(max_tiles=512, params=np.load(tp.WEIGHTS)["params"])
    res = tp.recall(w, net, [tp.HELDOUT_SEEDS[0] + 49], tp.RECALL_TOPK, tp.RECALL_CONFIDENCE)[0]
    net.close()
    print("held-out block:", res)
    assert res["peaks"] > 0
    assert res["found"] >= 0.99 * res["peaks"], res
    assert res["stats"]["skipped"] > 0, res


@pytest.mark.skipif(not os.path.exists(_WEIGHTS), reason="no trained network committed")
def test_trained_network_gain_over_no_network(gx, cuda):
    """What the network adds over searching without it: its P(peak) ranks the tiles that hold an
    exhaustive peak above the tiles with no peak nearby (area under the ROC, 0.5 without
    information; 0.95 over the 50 held-out blocks), and at the recall point the tiles it skips
    hold far fewer peaks than the same number of tiles skipped at random would."""
    from gtom_amd import proposal as pr
    from gtom_amd import train_proposal as tp
    w = tp.Workload()
    net = pr.ProposalNet(max_tiles=512, params=np.load(tp.WEIGHTS)["params"])
    blocks = [tp.block_scores(w, net, tp.HELDOUT_SEEDS[0] + k) for k in range(3)]
    net.close()
    auc = tp.detection_auc(blocks)
    lost = peaks = skipped = tiles = 0
    for pk, pnp, _ in blocks:
        skip = pnp >= max(0.5, tp.RECALL_CONFIDENCE)
        skipped += int(skip.sum())
        tiles += len(pnp)
        peaks += len(pk)
        lost += sum(int(skip[tp.tile_of(x, y, z)]) for x, y, z, _, _ in pk)
    random_lost = peaks * skipped / tiles  # expected loss of skipping as many tiles at random
    print(f"detection AUC {auc:.4f}; recall point: {skipped}/{tiles} tiles skipped, {lost} of {peaks} peaks in "
          f"them (random skipping: {random_lost:.2f} expected)")
    assert auc >= 0.9
    assert skipped > 0
    assert lost <= 0.25 * random_lost
