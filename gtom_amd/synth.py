"""Seeded synthetic template-matching workloads (host side, numpy).

A workload is a tomogram of Gaussian noise with rotated, CTF-filtered copies
of a random-blob template planted at integer positions, plus the orientation
list to search. The template is rotated by t_rot(x) = t(A^T (x - c)), the
real-space counterpart of the projector's Fourier-space sampling at A^T k, so a
copy planted with orientation A peaks at its centre c and at A's index.
"""
import math

import numpy as np
from scipy import ndimage


def euler_matrix(rot, tilt, psi):
    """ZYZ matrix in RELION's Euler_angles2matrix layout (src/Helpers/Euler.cu:39-68), double."""
    ca, sa = math.cos(rot), math.sin(rot)
    cb, sb = math.cos(tilt), math.sin(tilt)
    cg, sg = math.cos(psi), math.sin(psi)
    cc, cs, sc, ss = cb * ca, cb * sa, sb * ca, sb * sa
    return np.array([[cg * cc - sg * sa, cg * cs + sg * ca, -cg * sb],
                     [-sg * cc - cg * sa, -sg * cs + cg * ca, sg * sb],
                     [sc, ss, cb]])


def blob_template(box, rng, nblobs=12, sigma=(2.0, 4.0), radius=14.0):
    """Sum of isotropic Gaussians with centres inside `radius` of the box centre (z, y, x order)."""
    c = box // 2
    g = np.arange(box, dtype=np.float64) - c
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    t = np.zeros((box, box, box))
    for _ in range(nblobs):
        while True:
            p = rng.uniform(-radius, radius, 3)
            if np.dot(p, p) <= radius * radius:
                break
        s = rng.uniform(*sigma)
        # p is (x, y, z)
        t += np.exp(-((X - p[0]) ** 2 + (Y - p[1]) ** 2 + (Z - p[2]) ** 2) / (2 * s * s))
    return t.astype(np.float32)


def random_orientation(rng):
    """Uniform random rotation as ZYZ (rot, tilt, psi)."""
    return (rng.uniform(0, 2 * np.pi), math.acos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi))


def rotated_copy(template, A, order=3):
    """t_rot(x) = t(A^T x) about the box centre, same box."""
    box = template.shape[0]
    c = box // 2
    g = np.arange(box, dtype=np.float64) - c
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()])  # (x, y, z)
    src = A.T @ pts + c
    coords = np.stack([src[2], src[1], src[0]])  # map_coordinates wants (z, y, x)
    out = ndimage.map_coordinates(template.astype(np.float64), coords, order=order, mode="constant", cval=0.0)
    return out.reshape(template.shape).astype(np.float32)


def ctf3_numpy(n, pixelsize=1e-10, voltage=300e3, Cs=2.7e-3, defocus=-3e-6, amplitude=0.07, phaseshift=0.0):
    """Radially symmetric 3D CTF on the rfftn half spectrum (numpy axis order z, y, x), double.

    Same model as gtomx_ctf_3d (include/CTF.cuh:100-134, 195-223), used here only to
    filter the synthetic signal.
    """
    ny_ = 1.0 / (n * pixelsize * 1e10)
    lam = 12.2643247 / math.sqrt(voltage * (1.0 + voltage * 0.978466e-6))
    K1 = math.pi * lam
    K2 = math.pi / 2 * (Cs * 1e10) * lam ** 3
    K3 = math.atan(amplitude / math.sqrt(1 - amplitude ** 2))
    f = np.fft.fftfreq(n) * n
    fx = np.arange(n // 2 + 1, dtype=np.float64)
    Z, Y, X = np.meshgrid(f, f, fx, indexing="ij")
    k2 = (X ** 2 + Y ** 2 + Z ** 2) * ny_ ** 2
    gamma = K1 * (defocus * 1e10) * k2 + K2 * k2 * k2 - phaseshift - K3
    return -np.sin(gamma)


def plant(shape, template, positions, orientations, amplitude):
    """Signal volume (z, y, x) with rotated copies centred at positions (x, y, z), periodic."""
    sig = np.zeros(shape, np.float64)
    box = template.shape[0]
    c = box // 2
    for pos, ang in zip(positions, orientations):
        cp = rotated_copy(template, euler_matrix(*ang)) * amplitude
        zs = (np.arange(box) - c + pos[2]) % shape[0]
        ys = (np.arange(box) - c + pos[1]) % shape[1]
        xs = (np.arange(box) - c + pos[0]) % shape[2]
        sig[np.ix_(zs, ys, xs)] += cp
    return sig


def snr_amplitude(template, snr, noise_sigma=1.0, support_frac=0.05):
    t = template.astype(np.float64)
    sup = t > support_frac * t.max()
    v = t[sup].var()
    return math.sqrt(snr * noise_sigma ** 2 / v) if v > 0 else 0.0


def ctf_filter(vol, ctf_half):
    return np.fft.irfftn(np.fft.rfftn(vol) * ctf_half, s=vol.shape)


def cfg1(seed=1, n=64, box=32, ncopies=3, snr=0.1, angle_step_deg=30.0, maskradius=14.0,
         pixelsize=10e-10, defocus=-3e-6, orientations=None):
    """64^3 tomogram, 32^3 blob template, CTF 300 kV / Cs 2.7 mm / 10 A/px / 3 um, 30 deg search."""
    from . import equal_angular_spacing
    rng = np.random.default_rng(seed)
    template = blob_template(box, rng)
    angles = equal_angular_spacing((0, 2 * np.pi), (0, np.pi), (0, 2 * np.pi), np.deg2rad(angle_step_deg))
    if orientations is None:
        orientations = [random_orientation(rng) for _ in range(ncopies)]
    positions = [tuple(int(v) for v in rng.integers(box // 2, n - box // 2, 3)) for _ in range(ncopies)]
    amp = snr_amplitude(template, snr)
    sig = plant((n, n, n), template, positions, orientations, amp)
    ctf = ctf3_numpy(n, pixelsize=pixelsize, defocus=defocus)
    tomo = ctf_filter(sig, ctf) + rng.standard_normal((n, n, n))
    p = 2 * box
    return dict(n=n, p=p, oversample=p / n, template=template, tomogram=tomo.astype(np.float32),
                angles=angles, positions=positions, orientations=orientations, maskradius=maskradius,
                ctf_params=dict(pixelsize=pixelsize, voltage=300e3, Cs=2.7e-3, defocus=defocus,
                                amplitude=0.07), amplitude=amp)


def cfg2(seed=2, shape=(256, 256, 512), block=256, box=48, ncopies=40, snr=0.05, ndirs=500, npsi=18,
         maskradius=24.0, pixelsize=10e-10, margin=24):
    """Two 256^3 blocks side by side in x (a 512 x 256 x 256 tomogram), 48^3 template,
    500x18 Fibonacci orientations; copies are kept `margin` voxels from every face of
    their block, because the correlation is circular.

    Returns the blocks (nblocks, 256, 256, 256) float32 and per-block defoci.
    """
    from . import fibonacci_angles
    rng = np.random.default_rng(seed)
    template = blob_template(box, rng, radius=min(14.0, box / 2 - 6))
    nz, ny, nx = shape
    nblocks = nx // block
    positions, orientations = [], []
    for i in range(ncopies):
        b = i % nblocks
        x = int(rng.integers(b * block + margin, (b + 1) * block - margin))
        y = int(rng.integers(margin, block - margin))
        z = int(rng.integers(margin, block - margin))
        positions.append((x, y, z))
        orientations.append(random_orientation(rng))
    amp = snr_amplitude(template, snr)
    defoci = rng.uniform(2e-6, 5e-6, nblocks)
    blocks = np.empty((nblocks, block, block, block), np.float32)
    for b in range(nblocks):
        sel = [i for i in range(ncopies) if positions[i][0] // block == b]
        pos = [(positions[i][0] - b * block, positions[i][1], positions[i][2]) for i in sel]
        sig = plant((block, block, block), template, pos, [orientations[i] for i in sel], amp)
        ctf = ctf3_numpy(block, pixelsize=pixelsize, defocus=-defoci[b])
        blocks[b] = (ctf_filter(sig, ctf) + rng.standard_normal((block, block, block))).astype(np.float32)
    p = 2 * box
    return dict(n=block, p=p, oversample=p / block, template=template, blocks=blocks,
                angles=fibonacci_angles(ndirs, npsi), positions=positions, orientations=orientations,
                maskradius=maskradius, defoci=defoci, pixelsize=pixelsize, amplitude=amp)


# ---------------------------------------------------------------------------
# 256^3 blocks generated on the GPU (training and evaluation of the proposal
# network need thousands of them)
# ---------------------------------------------------------------------------
def proposal_template(box=48):
    """The fixed 48^3 random-blob template of the learned-proposal workloads (the
    template cfg2(seed=2) draws first)."""
    return blob_template(box, np.random.default_rng(2), radius=min(14.0, box / 2 - 6))


def rotate_trilinear(template, mats):
    """Rotated copies t(A^T (x - c) + c) of a (box, box, box) CUDA template for each
    3x3 matrix in `mats` (k, 3, 3), by trilinear interpolation, zero outside."""
    import torch
    import torch.nn.functional as F
    box = template.shape[0]
    c = box // 2
    dev = template.device
    g = torch.arange(box, device=dev, dtype=torch.float32) - c
    Z, Y, X = torch.meshgrid(g, g, g, indexing="ij")
    pts = torch.stack([X, Y, Z], -1).reshape(-1, 3)  # (x, y, z)
    A = torch.as_tensor(np.asarray(mats), dtype=torch.float32, device=dev)
    src = pts[None] @ A + c  # row (A^T p)^T = p^T A
    grid = (src / (box - 1) * 2 - 1).reshape(len(A), box, box, box, 3)
    inp = template[None, None].expand(len(A), 1, box, box, box)
    return F.grid_sample(inp, grid, mode="bilinear", padding_mode="zeros", align_corners=True)[:, 0]


def tomogram_block_gpu(seed, template, amplitude, n=256, ncopies=20, margin=24, pixelsize=10e-10,
                       defocus_range=(2e-6, 5e-6), device="cuda"):
    """One n^3 block: Gaussian noise (sigma 1) plus `ncopies` copies of the template
    at integer positions `margin` from the faces and uniform random orientations,
    the signal scaled by `amplitude` and filtered by the block's CTF (defocus
    uniform in defocus_range, 10 A/px, 300 kV, Cs 2.7 mm, amplitude contrast 0.07).
    Positions, orientations and defocus come from numpy's generator at `seed`, the
    noise from torch's CUDA generator at `seed`. Returns (block, positions (x, y, z),
    orientations (rot, tilt, psi), defocus in metres)."""
    import torch
    from . import CtfParams, ctf_3d, d_FFTR2C, d_IFFTC2R
    rng = np.random.default_rng(seed)
    positions = rng.integers(margin, n - margin, (ncopies, 3))
    orientations = [random_orientation(rng) for _ in range(ncopies)]
    defocus = float(rng.uniform(*defocus_range))
    box = template.shape[0]
    h = box // 2
    copies = rotate_trilinear(template, [euler_matrix(*a) for a in orientations]) * amplitude
    sig = torch.zeros((n, n, n), dtype=torch.float32, device=device)
    for k, (x, y, z) in enumerate(positions):
        sig[z - h:z - h + box, y - h:y - h + box, x - h:x - h + box] += copies[k]
    ctf = ctf_3d([CtfParams.make(pixelsize=pixelsize, Cs=2.7e-3, voltage=300e3, defocus=-defocus, amplitude=0.07)],
                 n, device=device)
    spec = d_FFTR2C(sig[None])
    spec *= ctf
    sig = d_IFFTC2R(spec, n)[0]
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    sig += torch.randn((n, n, n), generator=gen, device=device)
    return sig, positions, orientations, defocus
