"""The proposal network's kernels (gtom_amd/csrc/learn.hip) against its FP64
oracle (oracle/pnet.cpp) with the oracle rounding to bf16 where the GPU stores
bf16, so only fp32 accumulation order (and the rare bf16 rounding flip it
causes) separates them:

- the head's matrix-core GEMM (v_mfma_f32_16x16x32_bf16 operand layout) equals
  numpy's product of the bf16-rounded operands to 1e-5 normwise, on shapes that
  are not multiples of the 16 x 16 x 32 block;
- logits to 2e-3 normwise, per-tile losses to 1e-3;
- every parameter array's gradient to 2e-2 of its largest entry;
- the top-k directions equal numpy's ranking of the logits (ties to the lower
  index, the "no peak" class excluded), the top-1 probability the softmax
  maximum, the top-k mass the summed softmax of those directions and P(no peak)
  the softmax of class 500;
- Adam steps on one batch drive its loss down (the network can fit labels).
Tiles are 32^3; the network has no counterpart in the reference."""
import numpy as np
import pytest
from scipy import ndimage

from tests._util import rel_linf, to_np

pytestmark = pytest.mark.gpu


def _bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _tiles(rng, n):
    t = rng.standard_normal((n, 32, 32, 32))
    return (ndimage.gaussian_filter(t, (0, 1.2, 1.2, 1.2)) * 3.0).astype(np.float32)


def _params(seed):
    from gtom_amd import proposal as pr
    rng = np.random.default_rng(seed)
    p = pr.init_params(seed)
    for b in ("b1", "b2", "b3", "b4", "bf"):
        off, ln = pr.LAYERS[b]
        p[off:off + ln] = rng.standard_normal(ln) * 0.05
    return p


@pytest.mark.parametrize("shape", [(16, 16, 32), (37, 501, 1024), (250, 70, 45)])
def test_head_gemm_matches_numpy(gx, cuda, shape):
    import torch
    from gtom_amd import proposal as pr
    M, N, K = shape
    rng = np.random.default_rng(M)
    a = rng.standard_normal((M, K)).astype(np.float32)
    b = rng.standard_normal((K, N)).astype(np.float32)
    c = to_np(pr.debug_gemm(torch.from_numpy(a).to(cuda), torch.from_numpy(b).to(cuda)))
    want = _bf16(a).astype(np.float64) @ _bf16(b).astype(np.float64)
    assert rel_linf(c, want) <= 1e-5
    # and within bf16 rounding of the float product
    assert rel_linf(c, a.astype(np.float64) @ b.astype(np.float64)) <= 2e-2


def test_forward_and_gradients_match_oracle(gx, orc, cuda):
    import torch
    from gtom_amd import proposal as pr
    rng = np.random.default_rng(5)
    params = _params(11)
    tiles = _tiles(rng, 5)
    labels = np.array([0, 499, pr.NO_PEAK, 123, 250], np.int32)
    net = pr.ProposalNet(max_tiles=8, params=params)
    net.load_tiles(torch.from_numpy(tiles).to(cuda))
    logits = to_np(net.forward(5))
    want = orc.pnet_forward(params.astype(np.float64), tiles, emulate=True)
    assert rel_linf(logits, want) <= 2e-3
    g, loss = net.gradients(torch.from_numpy(labels).to(cuda))
    torch.cuda.synchronize()
    _, gw, lw = orc.pnet_gradients(params.astype(np.float64), tiles, labels, emulate=True)
    assert np.abs(to_np(loss) - lw).max() <= 1e-3 * max(1.0, np.abs(lw).max())
    g = to_np(g)
    for name, (off, ln) in pr.LAYERS.items():
        assert rel_linf(g[off:off + ln], gw[off:off + ln]) <= 2e-2, name
    net.close()


def test_proposals_rank_the_logits(gx, cuda):
    import torch
    from gtom_amd import proposal as pr
    rng = np.random.default_rng(6)
    net = pr.ProposalNet(max_tiles=40, params=_params(12))
    net.load_tiles(torch.from_numpy(_tiles(rng, 40)).to(cuda))
    logits = to_np(net.forward(40)).astype(np.float64)
    dirs, top1, cls, mass, pnp = net.propose(40, topk=16, full=True)
    dirs, top1, cls, mass, pnp = to_np(dirs), to_np(top1), to_np(cls), to_np(mass), to_np(pnp)
    for r in range(40):
        order = sorted(range(pr.NDIR), key=lambda k: (-logits[r, k], k))[:16]
        assert list(dirs[r]) == order
        p = np.exp(logits[r] - logits[r].max())
        p /= p.sum()
        assert cls[r] == int(np.argmax(p)) and abs(top1[r] - p.max()) <= 1e-5
        assert abs(mass[r] - p[order].sum()) <= 1e-5 and abs(pnp[r] - p[pr.NO_PEAK]) <= 1e-6
    net.close()


def test_training_fits_a_batch(gx, cuda):
    import torch
    from gtom_amd import proposal as pr
    rng = np.random.default_rng(7)
    net = pr.ProposalNet(max_tiles=32, seed=3)
    tiles = torch.from_numpy(_tiles(rng, 32)).to(cuda)
    labels = torch.from_numpy(rng.integers(0, pr.NCLS, 32).astype(np.int32)).to(cuda)
    net.load_tiles(tiles)
    first = float(net.train_step(labels, 1e-3).mean())
    for _ in range(60):
        last = float(net.train_step(labels, 1e-3).mean())
    assert np.isfinite(last) and last < 0.5 * first, (first, last)
    net.load_tiles(tiles)
    net.forward(32)
    _, _, cls = net.propose(32)
    assert (to_np(cls) == to_np(labels)).mean() >= 0.9
    net.close()
