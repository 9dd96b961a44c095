"""Shared checks for the parity tests."""
import numpy as np


def rel_linf(a, b):
    """Normwise relative L-infinity error max|a - b| / max|b|."""
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def to_np(t):
    return t.detach().cpu().numpy()


def argmax_agreement(gpu_idx, orc_idx, top1, top2, tol):
    """Fraction of voxels with a top-2 margin above tol*max|top1| whose index agrees, and their count."""
    scale = np.abs(top1).max()
    sel = (top1 - top2) > tol * scale
    n = int(sel.sum())
    if n == 0:
        return 1.0, 0
    agree = (gpu_idx[sel] == orc_idx[sel]).mean()
    return float(agree), n
