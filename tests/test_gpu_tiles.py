"""The grouped search over 32^3 tiles (gtomx_pick_update_tiles), the pruned
form the learned search runs: group g of `group` orientations is searched only
on the tiles set in its mask. Checked bit for bit against the plain update
(gtomx_pick_update) run group by group: every voxel of a tile holds the first
maximum over the groups whose mask sets that tile, in increasing orientation
order, with those groups' indices; tiles set in no mask keep the values begin
wrote. The masks reach pass 2 (planes and rows skipped) and pass 3 (rows
skipped, voxels of unset tiles untouched), so random masks at the benchmark's
256^3 shape and at 64^3 cover both the skipping and the per-voxel masking."""
import numpy as np
import pytest

from tests._util import to_np

pytestmark = pytest.mark.gpu


def _setup(gx, cuda, n, box, seed):
    import torch
    from gtom_amd import synth
    rng = np.random.default_rng(seed)
    template = synth.blob_template(box, rng, radius=min(14.0, box / 2 - 6))
    vol = rng.standard_normal((1, n, n, n)).astype(np.float32)
    p = 2 * box
    projft = gx.prepare_projector(torch.from_numpy(template).to(cuda), p)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(vol).to(cuda))
    ctf = torch.from_numpy(synth.ctf3_numpy(n, pixelsize=10e-10, defocus=-3e-6).astype(np.float32)).to(cuda)
    return rng, p, projft, exp_ft, ctf


@pytest.mark.parametrize("n,box,ngroups,group", [(64, 16, 5, 6), (256, 48, 4, 18)])
def test_grouped_update_equals_plain_update_per_tile(gx, cuda, n, box, ngroups, group):
    import torch
    rng, p, projft, exp_ft, ctf = _setup(gx, cuda, n, box, n)
    s = p / n
    angles = gx.fibonacci_angles(500, 18)[::37][:ngroups * group]
    nt = n // 32
    masks = (rng.random((1, ngroups, nt, nt, nt)) < 0.4)
    masks[0, 1] = False  # an empty group is skipped
    masks[0, :, nt - 1, 0, nt - 1] = False  # and one tile is set in no group
    packed = np.zeros((1, ngroups, nt, nt), np.uint8)
    for tx in range(nt):
        packed |= (masks[..., tx].astype(np.uint8) << tx)
    plan = gx.PickPlan(p, n, batch=max(group, 8))
    best = torch.empty((1, n, n, n), dtype=torch.float32, device=cuda)
    idx = torch.empty_like(best, dtype=torch.int32)
    gx.pick_begin(plan, 1, best, idx)
    gx.pick_update_tiles(plan, projft, s, exp_ft, ctf, angles, group, packed, best, idx)
    torch.cuda.synchronize()
    got_b, got_i = to_np(best)[0], to_np(idx)[0]
    # plain update per group, combined on the host in group order (strict '>')
    want_b = np.full((n, n, n), -1e30, np.float32)
    want_i = np.zeros((n, n, n), np.int32)
    tz, ty, tx = np.meshgrid(np.arange(n) // 32, np.arange(n) // 32, np.arange(n) // 32, indexing="ij")
    for g in range(ngroups):
        bg = torch.empty_like(best)
        ig = torch.empty_like(idx)
        gx.pick_begin(plan, 1, bg, ig)
        gx.pick_update(plan, projft, s, exp_ft, ctf, angles[g * group:(g + 1) * group], g * group, bg, ig)
        torch.cuda.synchronize()
        bg, ig = to_np(bg)[0], to_np(ig)[0]
        on = masks[0, g][tz, ty, tx]
        upd = on & (bg > want_b)
        want_b[upd] = bg[upd]
        want_i[upd] = ig[upd]
    plan.close()
    assert np.array_equal(got_b, want_b)
    assert np.array_equal(got_i, want_i)
    untouched = ~masks[0].any(axis=0)[tz, ty, tx]
    assert untouched.any() and (got_b[untouched] == -1e30).all()
