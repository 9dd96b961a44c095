"""The 2D branch of d_PickSubTomograms (n x n images, central sections of the
3D template; src/Correlation/SubTomograms.cu:33-36) on the GPU against the FP64
oracle, with the tolerance of the 3D search: max-CC within 1e-4 normwise, the
argmax orientation identical wherever the oracle's top-2 margin exceeds
1e-4 * max|top1|. Also the 2D forms of the standalone projector and CTF."""
import numpy as np
import pytest

from tests._util import argmax_agreement, rel_linf, to_np

pytestmark = pytest.mark.gpu
TAU = 2 * np.pi


def _template(box, seed):
    rng = np.random.default_rng(seed)
    g = np.arange(box) - box // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    t = np.zeros((box, box, box))
    for _ in range(8):
        c = rng.uniform(-box / 4, box / 4, 3)
        s = rng.uniform(1.0, 2.5, 3)
        t += np.exp(-(((x - c[0]) / s[0]) ** 2 + ((y - c[1]) / s[1]) ** 2 + ((z - c[2]) / s[2]) ** 2) / 2)
    return t.astype(np.float32)


def _images(orc, gx, cuda, n, box, nimg, angles, seed):
    """Images holding a few shifted central sections of the template plus noise."""
    import torch
    rng = np.random.default_rng(seed)
    p = 2 * box
    projft = gx.prepare_projector(torch.from_numpy(_template(box, seed)).to(cuda), p)
    P = to_np(projft).astype(np.complex128)
    imgs = []
    for i in range(nimg):
        img = np.zeros((n, n))
        for _ in range(3):
            k = int(rng.integers(len(angles)))
            R = orc.project2d(P, p, n, orc.projector_matrix(*angles[k].astype(np.float64), p / n))
            img += np.roll(np.fft.irfft2(R, s=(n, n)), tuple(rng.integers(0, n, 2)), axis=(0, 1))
        img = img / img.std() + 1.0 * rng.standard_normal((n, n))
        imgs.append(img)
    return projft, P, np.stack(imgs).astype(np.float32)


def _run(gx, orc, cuda, n, box, nimg, nang, batch, seed, ctf_per_image=False, oversample=None, rows_variant=0,
         oracle=True):
    import torch
    angles = orc.equal_angular_spacing((0, TAU), (0, np.pi), (0, TAU), np.deg2rad(30.0))[:nang]
    projft, P, imgs = _images(orc, gx, cuda, n, box, nimg, angles, seed)
    p = 2 * box
    E = gx.d_FFTR2C(torch.from_numpy(imgs).to(cuda), ndims=2)
    params = [gx.CtfParams.make(pixelsize=5e-10, Cs=2.7e-3, voltage=300e3, defocus=-(1.5 + 0.5 * i) * 1e-6,
                                amplitude=0.07) for i in range(nimg if ctf_per_image else 1)]
    ctf = gx.ctf_3d(params, n, device=cuda, ndims=2)
    stride = ctf[0].numel() if ctf_per_image else 0
    plan = gx.PickPlan(p, n, batch, ndims=2)
    if rows_variant:
        plan.set_rows_variant(rows_variant)
    ov = p / n if oversample is None else oversample
    best, idx = gx.d_PickSubTomograms(plan, projft, ov, E, ctf, angles, n / 8, ctf_volume_stride=stride)
    torch.cuda.synchronize()
    plan.close()
    if not oracle:
        return to_np(best), to_np(idx), None
    C = to_np(ctf).astype(np.float64)
    ref = orc.pick(P, p, ov, to_np(E).astype(np.complex128), C if ctf_per_image else C[0], n,
                   angles.astype(np.float64), n / 8, ctf_per_volume=ctf_per_image, ndims=2)
    return to_np(best), to_np(idx), ref


def _check(best, idx, ref):
    assert rel_linf(best, ref["best"]) <= 1e-4
    agree, cnt = argmax_agreement(idx, ref["idx"], ref["top1"], ref["top2"], 1e-4)
    assert cnt > 0 and agree == 1.0, (agree, cnt)


@pytest.mark.parametrize("n,box", [(32, 16), (64, 16)])
def test_pick2d_matches_oracle(gx, orc, cuda, n, box):
    best, idx, ref = _run(gx, orc, cuda, n, box, nimg=2, nang=100, batch=0, seed=n)
    _check(best, idx, ref)


def test_pick2d_256_several_batches_per_image_ctf(gx, orc, cuda):
    """256^2 images with their own CTFs, 300 orientations: one full pass of 240 and a partial one."""
    best, idx, ref = _run(gx, orc, cuda, 256, 32, nimg=2, nang=300, batch=0, seed=7, ctf_per_image=True)
    _check(best, idx, ref)


def test_pick2d_batch_size_does_not_change_results(gx, orc, cuda):
    b1, i1, _ = _run(gx, orc, cuda, 64, 16, nimg=1, nang=77, batch=240, seed=3)
    b2, i2, _ = _run(gx, orc, cuda, 64, 16, nimg=1, nang=77, batch=7, seed=3)
    assert np.array_equal(i1, i2) and np.array_equal(b1, b2)


@pytest.mark.parametrize("n,nang,ctf_per_image", [(512, 60, False), (768, 40, True), (1024, 40, False)])
def test_pick2d_micrograph_sizes_match_oracle(gx, orc, cuda, n, nang, ctf_per_image):
    """Images above 256 (the generic-transform form of the 2D branch: every
    half-spectrum bin projected, the batch through the two-pass / LDS-row C2R,
    maps folded into best / index in orientation order)."""
    best, idx, ref = _run(gx, orc, cuda, n, 32, nimg=2 if ctf_per_image else 1, nang=nang, batch=0, seed=n,
                          ctf_per_image=ctf_per_image)
    _check(best, idx, ref)


@pytest.mark.parametrize("n", [2048, 3072, 4096])
def test_pick2d_one_row_pass_matches_oracle(gx, orc, cuda, n):
    """The one-row-per-workgroup row pass with LDS twiddles (GTOMX_ROWS_ONE) against the oracle,
    and against the plain row pass (the same operations, compiled into a different kernel: equal
    up to the rounding of contracted multiply-adds), over two batches (4 + 2)."""
    best, idx, ref = _run(gx, orc, cuda, n, 32, nimg=1, nang=6, batch=4, seed=11, rows_variant=gx.ROWS_ONE)
    _check(best, idx, ref)
    b0, i0, _ = _run(gx, orc, cuda, n, 32, nimg=1, nang=6, batch=4, seed=11, oracle=False)
    err = rel_linf(best, b0)
    agree, cnt = argmax_agreement(idx, i0, ref["top1"], ref["top2"], 1e-5)
    print(f"{n}^2 one-row vs plain row pass: max difference {err:.2e}, argmax agreement {agree} on {cnt} pixels")
    assert err <= 1e-6
    assert cnt > 0 and agree == 1.0


def test_pick2d_micrograph_uncentred_projector_matches_oracle(gx, orc, cuda):
    """An oversample above p/n samples the projector beyond its centred copy
    (the plain projector layout) in the fused projection pass."""
    best, idx, ref = _run(gx, orc, cuda, 512, 32, nimg=1, nang=40, batch=0, seed=9, oversample=0.25)
    _check(best, idx, ref)


def test_pick2d_micrograph_batch_size_does_not_change_results(gx, orc, cuda):
    b1, i1, _ = _run(gx, orc, cuda, 512, 32, nimg=1, nang=50, batch=0, seed=5)
    b2, i2, _ = _run(gx, orc, cuda, 512, 32, nimg=1, nang=50, batch=7, seed=5)
    assert np.array_equal(i1, i2) and np.array_equal(b1, b2)


def test_project2d_matches_oracle(gx, orc, cuda):
    import torch
    box, n = 16, 48
    p = 2 * box
    projft = gx.prepare_projector(torch.from_numpy(_template(box, 5)).to(cuda), p)
    angles = orc.equal_angular_spacing((0, TAU), (0, np.pi), (0, TAU), np.deg2rad(45.0))[:9]
    got = to_np(gx.d_rlnProject(projft, p, n, angles=angles, oversample=p / n, ndims=2))
    P = to_np(projft).astype(np.complex128)
    for k in range(len(angles)):
        ref = orc.project2d(P, p, n, orc.projector_matrix(*angles[k].astype(np.float64), p / n))
        assert rel_linf(got[k], ref) <= 2e-5


def test_ctf2d_is_the_central_plane_of_the_3d_ctf(gx, cuda):
    params = [gx.CtfParams.make(pixelsize=1e-9, Cs=2.7e-3, voltage=300e3, defocus=-3e-6, amplitude=0.07)]
    c2 = to_np(gx.ctf_3d(params, 64, device=cuda, ndims=2))[0]
    c3 = to_np(gx.ctf_3d(params, 64, device=cuda))[0]
    assert np.array_equal(c2, c3[0])
