"""The allocation-free normalisation (gtomx_norm_monolithic_noalloc) against
gtomx_norm_monolithic, which runs the same kernels on per-call scratch: equal
bit for bit on two calls in a row, the second smaller (fewer elements and a
smaller batch, so stale partials of the first call would show), and a batch
above GTOMX_NOALLOC_MAX_BATCH is refused before any device work. The entry point
has no counterpart in the reference (d_NormMonolithic is
src/ImageManipulation/Norm.cu:76-91)."""
import numpy as np
import pytest

from tests._util import to_np

pytestmark = pytest.mark.gpu


def test_noalloc_norm_equals_allocating_norm(gx, cuda):
    import torch
    rng = np.random.default_rng(17)
    for batch, n in ((6, 300_001), (3, 1_000)):
        x = torch.from_numpy((rng.standard_normal((batch, n)) * 4.0 + 2.5).astype(np.float32)).to(cuda)
        want = to_np(gx.d_NormMonolithic(x, batch=batch))
        out = torch.full_like(x, float("nan"))
        gx.d_NormMonolithic_noalloc(x, out, batch=batch)
        torch.cuda.synchronize()
        got = to_np(out)
        assert np.array_equal(got, want), (batch, n, np.abs(got - want).max())


def test_noalloc_norm_refuses_batch_above_maximum(gx, cuda):
    import torch
    x = torch.zeros((gx.NOALLOC_MAX_BATCH + 1, 8), device=cuda)
    out = torch.empty_like(x)
    with pytest.raises(gx.GtomxError, match="GTOMX_NOALLOC_MAX_BATCH"):
        gx.d_NormMonolithic_noalloc(x, out, batch=gx.NOALLOC_MAX_BATCH + 1)
