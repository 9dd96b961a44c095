"""The proposal network's FP64 oracle (oracle/pnet.cpp), pinned on the CPU:

- forward: an independent formulation with scipy.ndimage.correlate (zero
  padding; tap (dz, dy, dx) weights input voxel (z+dz, y+dy, x+dx)), numpy
  block means for the pooling and a dense product for the head, to 1e-12;
  the kernel applied as a convolution (flipped taps) fails;
- gradients: central finite differences of the double-precision loss at
  random coordinates of every weight and bias array, to 1e-4 of the array's
  largest gradient (a step can cross a ReLU kink of one of the 65,536 voxels a
  bias reaches, which moves the difference quotient by ~1e-5); the data
  gradient taken with unflipped taps (planted in the oracle) fails by far more;
- bf16 emulation: rounding changes the logits by less than 3% and the
  gradients by less than 10% (it only moves values to neighbouring bf16
  numbers).
No counterpart exists in the reference (the network is the build's own)."""
import numpy as np
import pytest
from scipy import ndimage

from gtom_amd import proposal as pr


def _tiles(rng, n):
    # blobs on noise, so the activations are not all one sign
    t = rng.standard_normal((n, 32, 32, 32))
    return ndimage.gaussian_filter(t, (0, 1.2, 1.2, 1.2)) * 3.0


def _layer(params, name):
    off, ln = pr.LAYERS[name]
    return params[off:off + ln]


def _numpy_forward(params, tile, flip=False):
    P = params.astype(np.float64)

    def conv(x, w, b, cin):
        # x: (cin, S, S, S); w taps t = (dz+1)*9 + (dy+1)*3 + (dx+1)
        out = []
        for co in range(16):
            acc = np.full(x.shape[1:], b[co])
            for ci in range(cin):
                if cin == 1:
                    k = w.reshape(16, 27)[co]
                else:
                    k = w.reshape(27, 16, 16)[:, co, ci]
                k = k.reshape(3, 3, 3)
                if flip:
                    k = k[::-1, ::-1, ::-1]
                acc = acc + ndimage.correlate(x[ci], k, mode="constant", cval=0.0)
            out.append(np.maximum(acc, 0))
        return np.stack(out)

    def pool(x):
        c, s = x.shape[0], x.shape[1] // 2
        return x.reshape(c, s, 2, s, 2, s, 2).mean(axis=(2, 4, 6))

    a1 = conv(tile[None], _layer(P, "w1"), _layer(P, "b1"), 1)
    a2 = conv(pool(a1), _layer(P, "w2"), _layer(P, "b2"), 16)
    a3 = conv(pool(a2), _layer(P, "w3"), _layer(P, "b3"), 16)
    a4 = conv(a3, _layer(P, "w4"), _layer(P, "b4"), 16)
    feat = np.transpose(pool(a4), (1, 2, 3, 0)).reshape(-1)  # (z, y, x, c)
    W = _layer(P, "wf").reshape(pr.NCLS, pr.NFEAT)
    return W @ feat + _layer(P, "bf")


def test_forward_matches_independent_formulation(orc):
    rng = np.random.default_rng(1)
    params = pr.init_params(3).astype(np.float64)
    params[pr.OFF_BF:] = rng.standard_normal(pr.NCLS) * 0.1
    for b in ("b1", "b2", "b3", "b4"):
        off, ln = pr.LAYERS[b]
        params[off:off + ln] = rng.standard_normal(ln) * 0.05
    tiles = _tiles(rng, 2)
    got = orc.pnet_forward(params, tiles)
    for k in range(2):
        want = _numpy_forward(params, tiles[k])
        assert np.abs(got[k] - want).max() <= 1e-12 * max(1.0, np.abs(want).max())
    wrong = _numpy_forward(params, tiles[0], flip=True)
    assert np.abs(got[0] - wrong).max() > 1e-3 * np.abs(got[0]).max()


def test_gradients_match_finite_differences(orc):
    rng = np.random.default_rng(2)
    params = pr.init_params(5).astype(np.float64)
    for b in ("b1", "b2", "b3", "b4"):
        off, ln = pr.LAYERS[b]
        params[off:off + ln] = rng.standard_normal(ln) * 0.05
    tiles = _tiles(rng, 2)
    labels = np.array([17, pr.NO_PEAK], np.int32)
    loss, g, _ = orc.pnet_gradients(params, tiles, labels)
    def loss_of(p):
        lg = orc.pnet_forward(p, tiles)
        m = lg.max(axis=1, keepdims=True)
        lse = (m[:, 0] + np.log(np.exp(lg - m).sum(axis=1)))
        return float(np.mean(lse - lg[np.arange(len(labels)), labels]))

    assert abs(loss_of(params) - loss) <= 1e-12 * abs(loss)

    def fd(j, eps):
        p1 = params.copy()
        p1[j] += eps
        p2 = params.copy()
        p2[j] -= eps
        return (loss_of(p1) - loss_of(p2)) / (2 * eps)

    for name, (off, ln) in pr.LAYERS.items():
        scale = np.abs(g[off:off + ln]).max()
        assert scale > 0, name
        picks = rng.choice(ln, size=4, replace=False)
        # include the largest gradient of the array
        picks = np.append(picks, np.argmax(np.abs(g[off:off + ln])))
        for i in picks:
            # two step sizes: a kink crossed by one step is rarely crossed by the other
            errs = [abs(fd(off + i, eps) - g[off + i]) for eps in (1e-6, 1e-7)]
            assert min(errs) <= 1e-4 * scale + 1e-10, (name, i, errs, g[off + i])
    _, gw, _ = orc.pnet_gradients(params, tiles, labels, plant_unflipped=True)
    off, ln = pr.LAYERS["w2"]
    assert np.abs(gw[off:off + ln] - g[off:off + ln]).max() > 0.05 * np.abs(g[off:off + ln]).max()
    off, ln = pr.LAYERS["w4"]
    assert np.array_equal(gw[off:off + ln], g[off:off + ln])  # the last layer's gradient needs no data gradient


def test_emulated_rounding_stays_close(orc):
    rng = np.random.default_rng(3)
    params = pr.init_params(7).astype(np.float64)
    tiles = _tiles(rng, 2)
    labels = np.array([3, 250], np.int32)
    l64 = orc.pnet_forward(params, tiles)
    lbf = orc.pnet_forward(params, tiles, emulate=True)
    assert 0 < np.abs(l64 - lbf).max() <= 0.03 * np.abs(l64).max()
    _, g64, _ = orc.pnet_gradients(params, tiles, labels)
    _, gbf, _ = orc.pnet_gradients(params, tiles, labels, emulate=True)
    for name, (off, ln) in pr.LAYERS.items():
        a, b = g64[off:off + ln], gbf[off:off + ln]
        assert np.abs(a - b).max() <= 0.1 * np.abs(a).max(), name
