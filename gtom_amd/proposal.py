"""The orientation-proposal network (gtom_amd/csrc/learn.hip) from Python.

A 3D CNN over the 32^3 tiles of a tomogram block scores the 500 coarse
directions of the search's orientation set (Fibonacci-sphere directions, each
searched with 18 in-plane steps) and a "no peak" class. It is trained on the
exhaustive search's own labels: for each tile, the direction (argmax index // 18,
the argmax of UpdateCorrelationKernel, src/Correlation/SubTomograms.cu:138-158)
of the strongest peak of the search's normalised map within the tile (or a few
voxels around it), or "no peak". Training, inference and the shortlisting all
run in the library's kernels; this module holds device memory and the loops.
"""
import ctypes

import numpy as np

from . import _P, _ptr, _stream, check, lib

TILE = 32
NDIR = 500
NCLS = 501
NO_PEAK = 500
CPAD = 512
NFEAT = 1024
_NTAP = 27
_PC = 16

# parameter layout (fp32), as in learn.hip
OFF_W1 = 0
OFF_B1 = OFF_W1 + _PC * _NTAP
OFF_W2 = OFF_B1 + _PC
OFF_B2 = OFF_W2 + _NTAP * _PC * _PC
OFF_W3 = OFF_B2 + _PC
OFF_B3 = OFF_W3 + _NTAP * _PC * _PC
OFF_W4 = OFF_B3 + _PC
OFF_B4 = OFF_W4 + _NTAP * _PC * _PC
OFF_WF = OFF_B4 + _PC
OFF_BF = OFF_WF + NCLS * NFEAT
NPARAM = OFF_BF + NCLS

LAYERS = {"w1": (OFF_W1, _PC * _NTAP), "b1": (OFF_B1, _PC), "w2": (OFF_W2, _NTAP * _PC * _PC), "b2": (OFF_B2, _PC),
          "w3": (OFF_W3, _NTAP * _PC * _PC), "b3": (OFF_B3, _PC), "w4": (OFF_W4, _NTAP * _PC * _PC),
          "b4": (OFF_B4, _PC), "wf": (OFF_WF, NCLS * NFEAT), "bf": (OFF_BF, NCLS)}


def init_params(seed=0):
    """He-normal convolution weights, a small normal head, zero biases (float32)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    p[OFF_W1:OFF_B1] = rng.standard_normal(OFF_B1 - OFF_W1) * np.sqrt(2.0 / _NTAP)
    for w, b in ((OFF_W2, OFF_B2), (OFF_W3, OFF_B3), (OFF_W4, OFF_B4)):
        p[w:b] = rng.standard_normal(b - w) * np.sqrt(2.0 / (_NTAP * _PC))
    p[OFF_WF:OFF_BF] = rng.standard_normal(OFF_BF - OFF_WF) * np.sqrt(1.0 / NFEAT)
    return p


class ProposalNet:
    """Owns a gtomx_pnet: parameters, optimiser state and the activations of up
    to `max_tiles` tiles per pass."""

    def __init__(self, max_tiles=512, params=None, seed=0):
        h = _P()
        check(lib().gtomx_pnet_create(ctypes.byref(h), max_tiles), "gtomx_pnet_create")
        self.handle = h
        self.max_tiles = max_tiles
        assert lib().gtomx_pnet_param_count() == NPARAM
        self.set_params(init_params(seed) if params is None else params)

    def close(self):
        if self.handle:
            lib().gtomx_pnet_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, stream=None):
        a = np.ascontiguousarray(params, np.float32)
        assert a.shape == (NPARAM,)
        check(lib().gtomx_pnet_set_params(self.handle, a.ctypes.data, NPARAM, _stream(stream)),
              "gtomx_pnet_set_params")

    def get_params(self, stream=None):
        a = np.empty(NPARAM, np.float32)
        check(lib().gtomx_pnet_get_params(self.handle, a.ctypes.data, NPARAM, _stream(stream)),
              "gtomx_pnet_get_params")
        return a

    def load_tiles(self, tiles, stream=None):
        """tiles: float32 CUDA tensor (n, 32, 32, 32), already standardised."""
        n = tiles.shape[0]
        check(lib().gtomx_pnet_load_tiles(self.handle, _ptr(tiles), n, _stream(stream)), "gtomx_pnet_load_tiles")
        return n

    def load_volume(self, volume, stream=None):
        """Every 32^3 tile of an n^3 float32 CUDA volume, standardised by its mean and std."""
        from . import int3
        n = volume.shape[-1]
        check(lib().gtomx_pnet_load_volume(self.handle, _ptr(volume), int3(n, n, n), _stream(stream)),
              "gtomx_pnet_load_volume")
        return (n // TILE) ** 3

    def forward(self, n, stream=None):
        """Logits (n, 501) of the loaded tiles (a CUDA tensor)."""
        import torch
        out = torch.empty((n, CPAD), dtype=torch.float32, device="cuda")
        check(lib().gtomx_pnet_forward(self.handle, n, _ptr(out), _stream(stream)), "gtomx_pnet_forward")
        return out[:, :NCLS]

    def gradients(self, labels, stream=None):
        """(gradients (NPARAM,), per-tile losses (n,)) of the loaded tiles, CUDA tensors."""
        import torch
        n = labels.shape[0]
        g = torch.empty(NPARAM, dtype=torch.float32, device="cuda")
        loss = torch.empty(n, dtype=torch.float32, device="cuda")
        check(lib().gtomx_pnet_gradients(self.handle, _ptr(labels), n, _ptr(g), _ptr(loss), _stream(stream)),
              "gtomx_pnet_gradients")
        return g, loss

    def train_step(self, labels, lr, wd=0.0, stream=None):
        """One Adam step on the loaded tiles; returns the per-tile losses (CUDA tensor)."""
        import torch
        n = labels.shape[0]
        loss = torch.empty(n, dtype=torch.float32, device="cuda")
        check(lib().gtomx_pnet_train_step(self.handle, _ptr(labels), n, lr, wd, _ptr(loss), _stream(stream)),
              "gtomx_pnet_train_step")
        return loss

    def propose(self, n, topk=16, stream=None, full=False):
        """From the last forward pass: (directions (n, topk), top-1 probability (n,), top class (n,)),
        with full=True also (summed probability of the top-k directions (n,), P(no peak) (n,))."""
        import torch
        dirs = torch.empty((n, topk), dtype=torch.int32, device="cuda")
        top1 = torch.empty(n, dtype=torch.float32, device="cuda")
        cls = torch.empty(n, dtype=torch.int32, device="cuda")
        mass = torch.empty(n, dtype=torch.float32, device="cuda")
        pnp = torch.empty(n, dtype=torch.float32, device="cuda")
        check(lib().gtomx_pnet_propose(self.handle, topk, _ptr(dirs), _ptr(top1), _ptr(cls), _ptr(mass), _ptr(pnp),
                                       _stream(stream)), "gtomx_pnet_propose")
        return (dirs, top1, cls, mass, pnp) if full else (dirs, top1, cls)


def debug_gemm(a, b, stream=None):
    """C = A . B through the head's matrix-core GEMM (bf16 operands, fp32 result)."""
    import torch
    M, K = a.shape
    K2, N = b.shape
    assert K == K2
    c = torch.empty((M, N), dtype=torch.float32, device=a.device)
    check(lib().gtomx_pnet_debug_gemm(_ptr(a), _ptr(b), _ptr(c), M, N, K, _stream(stream)), "gtomx_pnet_debug_gemm")
    return c
