"""Pins the oracle's CTF, normalisation, local sigma, max/argmax and the whole path."""
import math

import numpy as np
import pytest


# ---------------------------------------------------------------- CTF
def test_ctf_at_zero_frequency_is_amplitude_contrast(orc):
    for amp in (0.07, 0.1, 0.2):
        c = orc.ctf3(32, 10e-10, 300e3, 2.7e-3, -3e-6, amp)
        assert math.isclose(c[0, 0, 0], amp, rel_tol=1e-12)


def test_ctf_zeros_follow_the_defocus_quadratic(orc):
    """With lambda(300 kV) = 0.019687 A, gamma(k) = pi*lambda*df*k^2 + pi/2*Cs*lambda^3*k^4 - atan(A/sqrt(1-A^2))
    crosses -m*pi at the roots of a quadratic in k^2; the sampled CTF changes sign between the
    neighbouring integer radii along x."""
    n, px, df, cs, amp = 128, 2e-10, -3e-6, 2.7e-3, 0.07
    c = orc.ctf3(n, px, 300e3, cs, df, amp)[0, 0, :]
    lam = 0.019687
    K1, K2 = math.pi * lam, math.pi / 2 * (cs * 1e10) * lam ** 3
    K3 = math.atan(amp / math.sqrt(1 - amp ** 2))
    ny = 1 / (n * px * 1e10)
    found = 0
    for m in range(0, 6):
        # K2 u^2 + K1 df u - K3 + m pi = 0, u = k^2 (take the small positive root)
        a, b, cc = K2, K1 * df * 1e10, -K3 + m * math.pi
        disc = b * b - 4 * a * cc
        if disc < 0:
            continue
        u = (-b - math.sqrt(disc)) / (2 * a)
        if u <= 0:
            continue
        r = math.sqrt(u) / ny
        lo = int(math.floor(r))
        if 1 <= lo < n // 2 - 1 and abs(r - lo) > 0.05 and abs(r - lo - 1) > 0.05:
            assert c[lo] * c[lo + 1] < 0, (m, r)
            found += 1
    assert found >= 3


def test_ctf_is_radial(orc):
    n = 32
    c = orc.ctf3(n, 10e-10, 300e3, 2.7e-3, -3e-6, 0.07)
    assert math.isclose(c[0, 0, 5], c[5, 0, 0], rel_tol=1e-14)  # x = 5 vs z = 5
    assert math.isclose(c[0, 3, 4], c[(-4) % n, 0, 3], rel_tol=1e-14)


# ---------------------------------------------------------------- spectral normalisation
def _herm0(R, n):
    """Hermitian projection of the x = 0 and x = n/2 planes, DC set to zero."""
    out = R.copy()
    for ix in (0, n // 2):
        plane = R[:, :, ix]
        mate = np.conj(plane[(-np.arange(n)) % n][:, (-np.arange(n)) % n])
        out[:, :, ix] = 0.5 * (plane + mate)
    out[0, 0, 0] = 0
    return out


def test_literal_normalisation_equals_fourier_form(orc):
    """C2R -> (x - mean)/std -> R2C -> /RMS (the reference sequence, src/Correlation/SubTomograms.cu:66-70)
    equals Herm(R) with DC zeroed, scaled to unit RMS over the half spectrum (the GPU's form)."""
    rng = np.random.default_rng(7)
    n = 16
    R = rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))
    r = orc.c2r(R, n)
    r = (r - r.mean()) / r.std()
    Y = orc.r2c(r)
    Y /= max(math.sqrt((np.abs(Y) ** 2).sum() / Y.size), 1e-20)
    F = _herm0(R, n)
    F /= math.sqrt((np.abs(F) ** 2).sum() / F.size)
    assert np.abs(Y - F).max() < 1e-12
    # and the Hermitian part matters: without it the two differ
    G = R.copy()
    G[0, 0, 0] = 0
    G /= math.sqrt((np.abs(G) ** 2).sum() / G.size)
    assert np.abs(Y - G).max() > 1e-3


# ---------------------------------------------------------------- local sigma
def test_local_std_matches_brute_force_window(orc):
    rng = np.random.default_rng(8)
    n, rad = 16, 3.5
    m = rng.standard_normal((n, n, n)) * 2 + 0.5
    sd, mean = orc.local_std(m, rad, with_mean=True)
    # soft-sphere weight at integer offset r (centre n/2, raised-cosine edge of width 1)
    c = n // 2
    W = orc.sphere_mask((n, n, n), rad, 1.0)
    offs = [(dz - c, dy - c, dx - c, W[dz, dy, dx]) for dz in range(n) for dy in range(n) for dx in range(n)
            if W[dz, dy, dx] > 0]
    M = sum(w for *_, w in offs)
    s1 = np.zeros_like(m)
    s2 = np.zeros_like(m)
    for dz, dy, dx, w in offs:
        sh = np.roll(m, (-dz, -dy, -dx), axis=(0, 1, 2))  # sh[x] = m[x + d]
        s1 += w * sh
        s2 += w * sh * sh
    ref_mean = s1 / M
    ref_sd = np.sqrt(np.maximum(s2 / M - ref_mean ** 2, 0))
    assert np.abs(mean - ref_mean).max() < 1e-10
    assert np.abs(sd - ref_sd).max() < 1e-10


def _aniso_lowpass(n, scale=1.0, widths=(0.02, 0.05, 0.11)):
    """A real, even half-spectrum filter (FFT order, x fastest) with a different
    width on each axis (x, y, z), so that a transposed layout changes the result."""
    f = np.fft.fftfreq(n)
    fz, fy, fx = np.meshgrid(f, f, np.fft.rfftfreq(n), indexing="ij")
    return scale * np.exp(-(fx ** 2 / widths[0] + fy ** 2 / widths[1] + fz ** 2 / widths[2]))


def test_local_std_fourier_mask_matches_brute_force_window(orc):
    """With a Fourier mask the window is the sphere filtered by it, k =
    irfftn(rfftn(ifftshift(sphere)) * fm), weighting offsets in FFT order; mean
    and sigma are the k-weighted moments with M = sum k."""
    rng = np.random.default_rng(18)
    n, rad = 12, 3.0
    m = rng.standard_normal((n, n, n)) * 2 + 0.5
    fm = _aniso_lowpass(n, 0.8)
    sd, mean = orc.local_std(m, rad, with_mean=True, fouriermask=fm)
    W = orc.sphere_mask((n, n, n), rad, 1.0)
    k = np.fft.irfftn(np.fft.rfftn(np.fft.ifftshift(W)) * fm, s=(n, n, n), axes=(0, 1, 2))
    M = k.sum()
    assert math.isclose(M, 0.8 * W.sum(), rel_tol=1e-12)  # the DC bin scales the sum
    s1 = np.zeros_like(m)
    s2 = np.zeros_like(m)
    for dz in range(n):
        for dy in range(n):
            for dx in range(n):
                sh = np.roll(m, (-dz, -dy, -dx), axis=(0, 1, 2))  # sh[x] = m[x + d]
                s1 += k[dz, dy, dx] * sh
                s2 += k[dz, dy, dx] * sh * sh
    ref_mean = s1 / M
    ref_sd = np.sqrt(np.maximum(s2 / M - ref_mean ** 2, 0))
    assert np.abs(mean - ref_mean).max() < 1e-10
    assert np.abs(sd - ref_sd).max() < 1e-10
    # the axis widths matter: x and z swapped give another window
    sd_t = orc.local_std(m, rad, fouriermask=_aniso_lowpass(n, 0.8, (0.11, 0.05, 0.02)))
    assert np.abs(sd_t - sd).max() > 1e-3


def test_local_std_fourier_mask_neutral_and_scale_invariant(orc):
    rng = np.random.default_rng(19)
    n, rad = 16, 4.0
    m = rng.standard_normal((n, n, n)) + 1.0
    sd0, mean0 = orc.local_std(m, rad, with_mean=True)
    sd1, mean1 = orc.local_std(m, rad, with_mean=True, fouriermask=np.ones((n, n, n // 2 + 1)))
    assert np.abs(sd1 - sd0).max() < 1e-12 and np.abs(mean1 - mean0).max() < 1e-12
    # a constant filter scales both moments and the mask sum alike
    sdc, meanc = orc.local_std(m, rad, with_mean=True, fouriermask=np.full((n, n, n // 2 + 1), 3.5))
    assert np.abs(sdc - sd0).max() < 1e-12 and np.abs(meanc - mean0).max() < 1e-12


def test_sphere_mask_edge(orc):
    n, rad = 32, 6.0
    W = orc.sphere_mask((n, n, n), rad, 1.0)
    c = n // 2
    assert W[c, c, c] == 1.0 and W[c, c, c + 5] == 1.0
    assert W[c, c, c + 7] == 0.0
    # distance 6.5: (cos(0.5 pi) + 1) / 2 = 0.5 at a non-grid point; check a grid point at distance sqrt(40)
    d = math.sqrt(40)
    assert math.isclose(W[c, c + 2, c + 6], (math.cos((d - rad) * math.pi) + 1) / 2, rel_tol=1e-12)


# ---------------------------------------------------------------- max / argmax
def test_update_max_argmax_first_maximum_wins(orc):
    rng = np.random.default_rng(9)
    B, L = 6, 500
    corr = rng.integers(0, 4, (B, L)).astype(float)
    best = np.full(L, -1e30)
    idx = np.zeros(L, np.int32)
    orc.update_max_argmax(corr, best, idx, 10)
    for i in range(L):
        col = corr[:, i]
        j = int(np.flatnonzero(col == col.max())[0])
        assert best[i] == col.max() and idx[i] == 10 + j


# ---------------------------------------------------------------- whole path
def _small_inputs(orc, n, rng):
    p = n
    r2 = ((np.indices((p, p, p)) - p // 2) ** 2).sum(0)
    template = rng.standard_normal((p, p, p)) * (r2 < (p // 4) ** 2)
    P = orc.r2c(np.fft.ifftshift(template))
    tomo = rng.standard_normal((n, n, n))
    E = orc.r2c(tomo)
    ctf = orc.ctf3(n, 10e-10, 300e3, 2.7e-3, -3e-6, 0.07)
    return P, p, tomo, E, ctf


def test_single_orientation_equals_direct_correlation(orc):
    """top1 for one orientation equals the circular cross-correlation sum_r e(r + x) t'(r), with
    t' the normalised, CTF-filtered projection in real space (pins the conj order / shift sign)."""
    rng = np.random.default_rng(10)
    n = 12
    P, p, tomo, E, ctf = _small_inputs(orc, n, rng)
    ang = np.array([[0.3, 0.9, 1.7]])
    res = orc.pick(P, p, 1.0, E[None], ctf, n, ang, 3.0)
    R = orc.project(P, p, n, orc.projector_matrix(*ang[0], 1.0)) * ctf
    t = orc.c2r(R, n)
    t = (t - t.mean()) / t.std()
    Y = orc.r2c(t)
    Y /= math.sqrt((np.abs(Y) ** 2).sum() / Y.size)
    tp = orc.c2r(Y, n) / n ** 3  # real-space template, centred at the origin
    e = tomo  # C2R_u(E * conj(Y))(x) = N * sum_r tomo(x + r) tp(r)
    cc = np.zeros((n, n, n))
    for z in range(n):
        for y in range(n):
            for x in range(n):
                cc[z, y, x] = (np.roll(e, (-z, -y, -x), axis=(0, 1, 2)) * tp).sum()
    cc *= n ** 3
    assert np.abs(res["top1"][0] - cc).max() < 1e-9 * np.abs(cc).max()


def test_planted_copy_is_found(orc):
    """A template planted at a known position with orientation index o peaks there with index o."""
    rng = np.random.default_rng(11)
    n, box = 32, 16
    from gtom_amd import synth
    t = synth.blob_template(box, rng, radius=5.0, sigma=(1.5, 2.5))
    angles = orc.equal_angular_spacing((0, 2 * np.pi), (0, np.pi), (0, 2 * np.pi), np.deg2rad(30))[:120]
    target, pos = 77, (13, 18, 11)
    sig = synth.plant((n, n, n), t, [pos], [tuple(float(a) for a in angles[target])], 1.0)
    tomo = sig + 0.02 * rng.standard_normal((n, n, n))
    p = 2 * box
    padded = np.zeros((p, p, p))
    padded[p // 2 - box // 2: p // 2 + box // 2, p // 2 - box // 2: p // 2 + box // 2,
           p // 2 - box // 2: p // 2 + box // 2] = t
    P = orc.r2c(np.fft.ifftshift(padded))
    E = orc.r2c(tomo)[None]
    ctf = np.ones((n, n, n // 2 + 1))
    res = orc.pick(P, p, p / n, E, ctf, n, angles.astype(float), 8.0)
    b = res["best"][0]
    z, y, x = np.unravel_index(np.argmax(b), b.shape)
    assert (x, y, z) == pos
    assert res["idx"][0, z, y, x] == target


def test_float32_twin_agrees_with_float64(orc):
    rng = np.random.default_rng(12)
    n = 16
    P, p, tomo, E, ctf = _small_inputs(orc, n, rng)
    ang = rng.uniform(0, 3, (5, 3))
    r64 = orc.pick(P, p, 1.0, E[None], ctf, n, ang, 4.0)
    r32 = orc.pick(P.astype(np.complex64), p, 1.0, E[None].astype(np.complex64), ctf.astype(np.float32), n, ang,
                   4.0, dtype=np.float32)
    assert np.abs(r32["best"] - r64["best"]).max() < 1e-4 * np.abs(r64["best"]).max()


def test_pick2d_planted_projection_found():
    """The 2D branch: an image holding the central section of the template at orientation k,
    shifted, is matched at that shift with index k."""
    from oracle import pyoracle as po
    rng = np.random.default_rng(41)
    n, box = 32, 16
    p = 2 * box
    g = np.arange(box) - box // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    t = np.zeros((box, box, box))
    for _ in range(6):
        c = rng.uniform(-4, 4, 3)
        t += np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / 4.0)
    padded = np.zeros((p, p, p))
    padded[p // 2 - box // 2:p // 2 + box // 2, p // 2 - box // 2:p // 2 + box // 2,
           p // 2 - box // 2:p // 2 + box // 2] = t
    F = np.fft.rfftn(np.fft.ifftshift(padded))
    tau = 2 * np.pi
    angles = po.equal_angular_spacing((0, tau), (0, np.pi), (0, tau), np.deg2rad(60.0))[:20].astype(np.float64)
    k = 13
    R = po.project2d(F, p, n, po.projector_matrix(*angles[k], p / n))
    img = np.roll(np.fft.irfft2(R, s=(n, n)), (5, -7), axis=(0, 1))
    img = img / img.std() + 0.05 * rng.standard_normal(img.shape)
    E = np.fft.rfft2(img)
    out = po.pick(F, p, p / n, E[None], np.ones((n, n // 2 + 1)), n, angles, 6.0, ndims=2)
    pos = np.unravel_index(np.argmax(out["top1"][0]), (n, n))
    assert pos == (5, n - 7)
    assert out["idx"][0][pos] == k


def _diff2_setup(seed=51, n=16, p=16, nang=6):
    from oracle import pyoracle as po
    rng = np.random.default_rng(seed)
    g = np.arange(p) - p // 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    t = sum(np.exp(-(((x - c[0]) / 1.5) ** 2 + ((y - c[1]) / 2.0) ** 2 + ((z - c[2]) / 1.2) ** 2) / 2)
            for c in rng.uniform(-3, 3, (4, 3)))
    P = np.fft.rfftn(np.fft.ifftshift(t))
    tau = 2 * np.pi
    angles = po.equal_angular_spacing((0, tau), (0, np.pi), (0, tau), np.deg2rad(90.0))[:nang].astype(np.float64)
    ctf = rng.uniform(0.3, 1.0, (n, n, n // 2 + 1))
    return po, rng, P, angles, ctf


def test_diff2_matches_fft_formulation():
    """sum|R c e^{ik.s} - E|^2 = sum|Rc|^2 + sum|E|^2 - 2 Re X(d), X = forward DFT of
    Rc conj(E) on the half grid (zero for kx > n/2) evaluated at d = index - r/2."""
    po, rng, P, angles, ctf = _diff2_setup()
    n, p = 16, 16
    E = np.fft.rfftn(rng.standard_normal((n, n, n)))
    rel = (3, 4, 5)  # (rz, ry, rx)
    best, idx = po.diff2(P, p, p / n, E, ctf, n, rel, angles)
    scores = []
    for a in range(len(angles)):
        R = po.project(P, p, n, po.projector_matrix(*angles[a], p / n)) * ctf
        full = np.zeros((n, n, n), np.complex128)
        full[:, :, :n // 2 + 1] = R * np.conj(E)
        X = np.fft.fftn(full)
        dz, dy, dx = [np.arange(r) - r // 2 for r in rel]
        Xd = X[np.ix_(dz % n, dy % n, dx % n)]
        scores.append(np.sum(np.abs(R) ** 2) + np.sum(np.abs(E) ** 2) - 2 * Xd.real)
    scores = np.stack(scores)
    assert np.allclose(best, scores.min(axis=0), rtol=1e-9, atol=0)
    assert np.array_equal(idx, scores.argmin(axis=0))


def test_diff2_planted_shift_scores_zero():
    """E built as R_k ctf exp(i k.s0) for orientation k and the shift of one relevant voxel:
    that voxel scores 0 with index k."""
    po, rng, P, angles, ctf = _diff2_setup(seed=52)
    n, p, k = 16, 16, 4
    rel = (4, 4, 4)
    vz, vy, vx = 1, 3, 0  # voxel index in the relevant box
    s = -2 * np.pi * (np.array([vx, vy, vz]) - np.array([2, 2, 2])) / n
    R = po.project(P, p, n, po.projector_matrix(*angles[k], p / n)) * ctf
    kz = np.fft.fftfreq(n, 1 / n)[:, None, None]
    ky = np.fft.fftfreq(n, 1 / n)[None, :, None]
    kx = np.arange(n // 2 + 1)[None, None, :]
    E = R * np.exp(1j * (kx * s[0] + ky * s[1] + kz * s[2]))
    best, idx = po.diff2(P, p, p / n, E, ctf, n, rel, angles)
    assert best[vz, vy, vx] < 1e-18 * np.sum(np.abs(E) ** 2)
    assert idx[vz, vy, vx] == k
    assert np.argmin(best) == np.ravel_multi_index((vz, vy, vx), rel)
