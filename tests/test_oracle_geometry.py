"""Pins the oracle's Euler matrices, orientation lists and projector."""
import math

import numpy as np
import pytest
from scipy import ndimage


def rz(t):  # passive rotation about z, RELION convention
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1.0]])


def ry(t):
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, 0, -s], [0, 1, 0], [s, 0, c]])


@pytest.mark.parametrize("ang", [(0.3, 1.1, -2.0), (np.deg2rad(-13), np.deg2rad(-20), np.deg2rad(-161)),
                                 (0, 0, 0), (2.5, np.pi, 0.7)])
def test_euler_is_zyz_composition(orc, ang):
    A = orc.euler_matrix(*ang)
    assert np.allclose(A, rz(ang[2]) @ ry(ang[1]) @ rz(ang[0]), atol=1e-15)
    assert np.allclose(A @ A.T, np.eye(3), atol=1e-14)
    assert math.isclose(np.linalg.det(A), 1.0, rel_tol=1e-14)


def test_euler_known_vector():
    """Matrix3Euler(0, 20 deg, 0) applied to x (GTOMTest/Euler.cpp:17-19) gives (cos 20, 0, sin 20)."""
    from oracle import pyoracle as orc
    A = orc.euler_matrix(0.0, np.deg2rad(20), 0.0)
    assert np.allclose(A @ [1, 0, 0], [math.cos(np.deg2rad(20)), 0, math.sin(np.deg2rad(20))], atol=1e-15)


@pytest.mark.parametrize("deg,count", [(30, 588), (15, 4536), (12.5, 7616)])
def test_equal_angular_spacing_counts(orc, deg, count):
    a = orc.equal_angular_spacing((0, 2 * np.pi), (0, np.pi), (0, 2 * np.pi), np.deg2rad(deg))
    assert a.shape == (count, 3)
    # first ring is the pole with a single rot step; psi runs fastest
    npsi = int(math.floor((360 - deg) / deg + 1e-6)) + 1  # psi in [0, 2 pi - increment]
    assert np.all(a[:npsi, 1] == 0) and np.all(a[:npsi, 0] == 0)
    assert a[npsi, 1] > 0
    assert np.allclose(np.diff(a[:npsi, 2]), np.float32(np.deg2rad(deg)), atol=1e-6)


def test_projector_matrix_is_scaled_transpose(orc):
    ang = (0.4, 2.2, -1.3)
    M = orc.projector_matrix(*ang, 0.375)
    assert np.allclose(M, 0.375 * orc.euler_matrix(*ang).T, atol=1e-16)


def _projector(orc, rng, p):
    return orc.r2c(rng.standard_normal((p, p, p)))


def _signed(i, n):
    return i if i <= n // 2 else i - n


def test_project_identity_returns_stored_bins(orc):
    rng = np.random.default_rng(3)
    n = p = 16
    P = _projector(orc, rng, p)
    out = orc.project(P, p, n, np.eye(3))
    for iz in range(n):
        for iy in range(n):
            for ix in range(n // 2 + 1):
                z, y = _signed(iz, n), _signed(iy, n)
                if ix * ix + y * y + z * z > (n // 2) ** 2:
                    assert out[iz, iy, ix] == 0
                else:
                    assert out[iz, iy, ix] == P[iz, iy, ix]


def test_project_quarter_turn_permutes_bins(orc):
    """rot = 90 deg about z maps q = (k_y, -k_x, k_z); negative q_x is taken from the Hermitian mate."""
    rng = np.random.default_rng(4)
    n = p = 16
    P = _projector(orc, rng, p)
    M = orc.projector_matrix(np.pi / 2, 0.0, 0.0, 1.0)
    out = orc.project(P, p, n, M)
    for iz in range(n):
        for iy in range(n):
            for ix in range(n // 2 + 1):
                x, y, z = ix, _signed(iy, n), _signed(iz, n)
                if x * x + y * y + z * z > (n // 2) ** 2:
                    continue
                q = np.rint(M @ [x, y, z]).astype(int)
                if q[0] < 0:
                    ref = np.conj(P[(-q[2]) % p, (-q[1]) % p, -q[0]])
                else:
                    ref = P[q[2] % p, q[1] % p, q[0]]
                assert abs(out[iz, iy, ix] - ref) < 1e-12 * abs(P).max()


def test_project_matches_scipy_trilinear(orc):
    """Independent trilinear interpolation (scipy, periodic in y/z) of the projector at q = M k."""
    rng = np.random.default_rng(5)
    n, p = 16, 32
    P = _projector(orc, rng, p)
    M = orc.projector_matrix(0.7, 1.2, -0.4, 1.7)  # |q| <= 1.7 * 8 < p/2 - 1: no x wrap
    out = orc.project(P, p, n, M)
    iz, iy, ix = np.meshgrid(np.arange(n), np.arange(n), np.arange(n // 2 + 1), indexing="ij")
    z = np.where(iz <= n // 2, iz, iz - n)
    y = np.where(iy <= n // 2, iy, iy - n)
    x = ix
    inside = x * x + y * y + z * z <= (n // 2) ** 2
    q = np.einsum("ij,jzyx->izyx", M, np.stack([x, y, z]).astype(float))
    flip = q[0] < -1e-5
    q = np.where(flip[None], -q, q)
    coords = np.stack([np.mod(q[2], p), np.mod(q[1], p), q[0]]).reshape(3, -1)
    # extend x by one column so that x1 = p/2 + 1 never leaves the array (weights are ~0 there)
    Pe = np.concatenate([P, P[:, :, :1]], axis=2)
    re = ndimage.map_coordinates(Pe.real, coords, order=1, mode="grid-wrap").reshape(q.shape[1:])
    im = ndimage.map_coordinates(Pe.imag, coords, order=1, mode="grid-wrap").reshape(q.shape[1:])
    ref = np.where(flip, re - 1j * im, re + 1j * im)
    ref = np.where(inside, ref, 0)
    assert np.abs(out - ref).max() < 1e-10 * np.abs(P).max()


def _gauss_ft_template(orc, n, A, sig):
    """R2C(ifftshift(t_rot)) for t_rot(r) = g(A^T r), g an anisotropic Gaussian at the box centre."""
    c = n // 2
    g = np.arange(n) - c
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    pts = np.stack([X, Y, Z]).reshape(3, -1)
    src = A.T @ pts
    v = np.exp(-0.5 * ((src[0] / sig[0]) ** 2 + (src[1] / sig[1]) ** 2 + (src[2] / sig[2]) ** 2))
    return orc.r2c(np.fft.ifftshift(v.reshape(n, n, n)))


def test_project_rotated_gaussian_matches_closed_form(orc):
    """Projecting the oversampled FT of an anisotropic Gaussian at rotation A equals the FT of the
    Gaussian rotated in real space by A (t(A^T r)), within interpolation error; the transposed
    rotation does not. Pins the rotation direction and the oversampling factor."""
    n, p = 32, 64
    sig = (1.6, 2.4, 3.6)
    ang = (0.6, 1.0, 2.1)
    A = orc.euler_matrix(*ang)
    P = _gauss_ft_template(orc, p, np.eye(3), sig)  # unrotated template in the p box
    out = orc.project(P, p, n, orc.projector_matrix(*ang, p / n))
    ref = _gauss_ft_template(orc, n, A, sig)
    iz, iy, ix = np.meshgrid(np.arange(n), np.arange(n), np.arange(n // 2 + 1), indexing="ij")
    z = np.where(iz <= n // 2, iz, iz - n)
    y = np.where(iy <= n // 2, iy, iy - n)
    inside = ix ** 2 + y ** 2 + z ** 2 <= (n // 2) ** 2
    err = np.abs(out - ref)[inside].max() / np.abs(ref).max()
    assert err < 0.02, err
    wrong = _gauss_ft_template(orc, n, A.T, sig)
    err_t = np.abs(out - wrong)[inside].max() / np.abs(ref).max()
    assert err_t > 0.2, err_t


def test_project_reads_next_row_with_zero_weight(orc):
    """q_x = p/2 exactly: x1 = p/2 + 1 is read at the next linear address with weight 0."""
    rng = np.random.default_rng(6)
    n = p = 16
    P = _projector(orc, rng, p)
    out = orc.project(P, p, n, np.eye(3))
    assert out[0, 0, n // 2] == P[0, 0, p // 2]


def test_project2d_is_the_central_plane_of_the_3d_projection(orc):
    """Project3DArraytoNDKernel<2> samples M (x, y, 0) with the test x^2 + y^2 <= rmax^2,
    i.e. exactly the z = 0 plane of the 3D projection."""
    rng = np.random.default_rng(31)
    p, n = 24, 16
    F = np.fft.rfftn(rng.standard_normal((p, p, p)))
    for k in range(4):
        M = orc.projector_matrix(*rng.uniform(0, 2 * np.pi, 3), p / n)
        assert np.array_equal(orc.project2d(F, p, n, M), orc.project(F, p, n, M)[0])


def test_project2d_identity_is_the_transform_of_the_z_sum(orc):
    """Central-section theorem on the grid (p = n, identity): the 2D projection is the 2D
    transform of the volume summed along z, inside the disc."""
    rng = np.random.default_rng(32)
    n = 16
    vol = rng.standard_normal((n, n, n))
    got = orc.project2d(np.fft.rfftn(vol), n, n, np.eye(3))
    ref = np.fft.rfft2(vol.sum(axis=0))
    ky = np.fft.fftfreq(n, 1 / n)[:, None]
    kx = np.arange(n // 2 + 1)[None, :]
    inside = kx ** 2 + ky ** 2 <= (n // 2) ** 2
    assert np.allclose(got[inside], ref[inside], atol=1e-9 * np.abs(ref).max())
    assert np.all(got[~inside] == 0)


def test_ctf_simulate_symmetric_case_is_the_3d_ctf_plane(orc):
    """Without astigmatism, on the 2D half-spectrum addresses, d_CTFSimulate equals the
    z = 0 plane of the radially symmetric 3D CTF (already pinned by its closed form)."""
    n, px = 32, 2e-10
    got = orc.ctf_simulate([dict(pixelsize=px, defocus=-2e-6, Cs=2.7e-3)], orc.ctf_addresses_2d(n))[0]
    ref = orc.ctf3(n, px, 300e3, 2.7e-3, -2e-6, 0.07)[0]
    assert np.allclose(got.reshape(n, n // 2 + 1), ref, atol=1e-12)


def test_ctf_simulate_astigmatism_axes(orc):
    """Along the astigmatism angle the defocus is defocus + delta/2, at 90 degrees to it
    defocus - delta/2 (CTFParamsLean halves defocusdelta)."""
    k = np.linspace(0, 0.5, 50)
    ang, d, dd = 0.3, -2.5e-6, 0.4e-6
    astig = dict(pixelsize=1.5e-10, defocus=d, defocusdelta=dd, astigmatismangle=ang)
    along = orc.ctf_simulate([astig], np.stack([k, np.full_like(k, ang)], 1))[0]
    across = orc.ctf_simulate([astig], np.stack([k, np.full_like(k, ang + np.pi / 2)], 1))[0]
    sym_hi = orc.ctf_simulate([dict(pixelsize=1.5e-10, defocus=d + dd / 2)], np.stack([k, np.zeros_like(k)], 1))[0]
    sym_lo = orc.ctf_simulate([dict(pixelsize=1.5e-10, defocus=d - dd / 2)], np.stack([k, np.zeros_like(k)], 1))[0]
    assert np.allclose(along, sym_hi, atol=1e-12) and np.allclose(across, sym_lo, atol=1e-12)


def test_ctf_simulate_flags_and_bfactor(orc):
    k = np.linspace(0, 0.5, 64)
    a = np.stack([k, np.zeros_like(k)], 1)
    base = orc.ctf_simulate([dict(pixelsize=2e-10)], a)[0]
    assert np.allclose(orc.ctf_simulate([dict(pixelsize=2e-10)], a, ampsq=True)[0], np.abs(base))
    b = orc.ctf_simulate([dict(pixelsize=2e-10, Bfactor=-50e-20)], a)[0]
    assert np.allclose(b, base * np.exp(-50 * 0.25 * (k / 2.0) ** 2))
    first = orc.ctf_simulate([dict(pixelsize=2e-10)], a, ignorefirst=True)[0]
    assert first[0] == 1.0 and np.allclose(first[-1], base[-1])


def _tilt_positions(imx, imy, rx, ry):
    return [(x, y) for y in range(0, imy - ry + 1, (imy - ry) // 3) for x in range(0, imx - rx + 1, (imx - rx) // 4)]


def test_ctf_tilt_zgrid_untilted_is_constant(orc):
    pos = _tilt_positions(1024, 768, 128, 128)
    z = orc.ctf_tilt_zgrid(0.7, (0.3, 0.0), (1.1, 0.0), -2.5e-6, 1.5e-10, (1024, 768), (128, 128), pos)
    assert np.allclose(z, -2.5e-6, rtol=0, atol=1e-18)


@pytest.mark.parametrize("axis,theta,imageangle", [(0.0, 0.4, 0.0), (np.pi / 2, -0.3, 0.0), (0.6, 0.25, 0.9)])
def test_ctf_tilt_zgrid_single_tilt_is_the_rotated_plane(orc, axis, theta, imageangle):
    """One tilt: the height of the plane whose normal is scipy's rotation of e_z
    about the in-plane axis, at the image-rotated region centre."""
    from scipy.spatial.transform import Rotation
    imx, imy, rx, ry, px, d0 = 1024, 768, 128, 96, 2e-10, -3e-6
    pos = _tilt_positions(imx, imy, rx, ry)
    n = Rotation.from_rotvec(theta * np.array([np.cos(axis), np.sin(axis), 0.0])).apply([0, 0, 1.0])
    c, s = np.cos(imageangle), np.sin(imageangle)
    want = []
    for x, y in pos:
        fx, fy = (x + rx / 2 - imx / 2) * px, (y + ry / 2 - imy / 2) * px
        gx, gy = c * fx - s * fy, s * fx + c * fy
        want.append(d0 - (gx * n[0] + gy * n[1]) / n[2])
    z = orc.ctf_tilt_zgrid(imageangle, (axis, theta), (0.0, 0.0), d0, px, (imx, imy), (rx, ry), pos)
    assert np.allclose(z, want, rtol=0, atol=1e-20)


@pytest.mark.parametrize("sa,st,pa,pt", [(0.4, 0.35, 1.3, -0.2), (0.0, 0.0, np.pi / 2, -0.3), (0.9, 0.0, 0.2, 0.4)])
def test_ctf_tilt_zgrid_specimen_tilt_keeps_the_reference_matrix(orc, sa, st, pa, pt):
    """A specimen tilt: the stage matrix's [0][0] element is c + a_x + a_x(1 - c)
    (src/Helpers/Euler.cu:319) instead of Rodrigues' c + a_x^2 (1 - c) (even for
    an untilted stage, where it is 1 + a_x), so the normal's x component exceeds
    scipy's composition by (a_x + a_x(1-c)(1 - a_x)) v_x, v = the specimen-rotated e_z."""
    from scipy.spatial.transform import Rotation
    ax = np.array([np.cos(sa), np.sin(sa), 0.0])
    v = Rotation.from_rotvec(pt * np.array([np.cos(pa), np.sin(pa), 0.0])).apply([0, 0, 1.0])
    n = Rotation.from_rotvec(st * ax).apply(v)
    c = np.cos(st)
    n[0] += (ax[0] + ax[0] * (1 - c) * (1 - ax[0])) * v[0]
    imx, imy, rx, ry, px, d0 = 512, 512, 64, 64, 3e-10, -2e-6
    pos = _tilt_positions(imx, imy, rx, ry)
    want = [d0 - ((x + rx / 2 - imx / 2) * px * n[0] + (y + ry / 2 - imy / 2) * px * n[1]) / n[2] for x, y in pos]
    z = orc.ctf_tilt_zgrid(0.0, (sa, st), (pa, pt), d0, px, (imx, imy), (rx, ry), pos)
    assert np.allclose(z, want, rtol=0, atol=1e-20)


def _numpy_sphere(shape, radius, sigma, center=None, decentered=False):
    nz, ny, nx = shape
    c = center if center is not None else (nx // 2, ny // 2, nz // 2)
    def d(n, cv):
        i = np.arange(n)
        return np.where(i < n // 2 + 1, i, i - n) if decentered else np.trunc(i - np.float32(cv)).astype(np.int64)
    zz = d(nz, c[2]) if nz > 1 else np.zeros(nz, np.int64)
    r = np.sqrt((d(nx, c[0])[None, None, :] ** 2 + d(ny, c[1])[None, :, None] ** 2
                 + zz[:, None, None] ** 2).astype(np.float64))
    soft = np.maximum(0, (np.cos(np.minimum(1, (r - radius) / sigma) * np.pi) + 1) / 2) if sigma > 0 else 0 * r
    return np.where(r < radius, 1.0, soft)


@pytest.mark.parametrize("shape,radius,sigma,center,dec", [
    ((16, 18, 20), 6.0, 2.0, None, False), ((16, 18, 20), 5.5, 0.0, (7.5, 9.25, 6.75), False),
    ((16, 18, 20), 7.0, 3.0, None, True), ((1, 24, 30), 9.0, 1.5, (10.5, 12.0, 0.0), False)])
def test_sphere_mask_ex_matches_numpy(orc, shape, radius, sigma, center, dec):
    got = orc.sphere_mask_ex(shape, radius, sigma, center, dec)
    assert np.allclose(got, _numpy_sphere(shape, radius, sigma, center, dec), atol=1e-12)


def test_sphere_mask_ex_known_answers(orc):
    shape = (12, 14, 16)
    # an integer centre moved by (2, -3, 1) rolls a mask that stays clear of the faces
    small = orc.sphere_mask_ex(shape, 2.5, 1.0)
    moved = orc.sphere_mask_ex(shape, 2.5, 1.0, center=(16 // 2 + 2, 14 // 2 - 3, 12 // 2 + 1))
    assert np.array_equal(moved, np.roll(small, (1, -3, 2), axis=(0, 1, 2)))
    centred = orc.sphere_mask_ex(shape, 4.5, 2.0)
    # decentered = the centred mask moved to the origin (ifftshift) for even sizes
    dec = orc.sphere_mask_ex(shape, 4.5, 2.0, decentered=True)
    assert np.array_equal(dec, np.fft.ifftshift(centred))
