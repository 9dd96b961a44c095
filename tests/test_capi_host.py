"""The C ABI without a GPU: the library loads, exports what include/gtomx.h declares,
its host-only functions agree with the oracle, and argument validation returns
error codes before any device work."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    hdr = open(os.path.join(ROOT, "include", "gtomx.h")).read()
    return set(re.findall(r"\b(gtomx_[a-z0-9_]+)\(", hdr))


def test_library_exports_every_declared_symbol(gx):
    out = subprocess.run(["nm", "-D", "--defined-only", gx.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = declared() - syms
    assert not missing, missing
    assert declared() == set(gx.exported_symbols())


def test_library_targets_gfx950(gx):
    data = open(gx.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_one_hip_runtime_in_process(gx):
    maps = open("/proc/self/maps").read()
    hip = {line.split()[-1] for line in maps.splitlines() if "libamdhip64" in line}
    assert len(hip) == 1, hip


def test_status_strings(gx):
    L = gx.lib()
    assert L.gtomx_status_string(0) == b"ok"
    assert L.gtomx_status_string(2) == b"unsupported"
    assert b"gfx950" in L.gtomx_version()


@pytest.mark.parametrize("deg", [30, 15, 12.5])
def test_equal_angular_spacing_matches_oracle(gx, orc, deg):
    a = gx.equal_angular_spacing((0, 2 * np.pi), (0, np.pi), (0, 2 * np.pi), np.deg2rad(deg))
    b = orc.equal_angular_spacing((0, 2 * np.pi), (0, np.pi), (0, 2 * np.pi), np.deg2rad(deg))
    assert np.array_equal(a, b)


def test_fibonacci_angles(gx):
    a = gx.fibonacci_angles(500, 18)
    assert a.shape == (9000, 3)
    i = np.arange(500)
    z = 1 - (2 * i + 1) / 500
    assert np.allclose(np.cos(a[::18, 1].astype(np.float64)), z, atol=1e-6)
    assert np.allclose(a[:18, 2], 2 * np.pi * np.arange(18) / 18, atol=1e-6)
    # directions are the rot/tilt pairs; each repeated for every psi
    assert np.all(a[:18, 0] == a[0, 0]) and np.all(a[:18, 1] == a[0, 1])


def test_projector_matrices_match_oracle(gx, orc):
    rng = np.random.default_rng(0)
    ang = rng.uniform(-3, 3, (20, 3)).astype(np.float32)
    M = gx.projector_matrices(ang, 0.375)
    for k in range(20):
        ref = orc.projector_matrix(*ang[k].astype(np.float64), 0.375)
        assert np.abs(M[k] - ref).max() < 1e-7


def test_validation_errors_without_device_work(gx):
    L = gx.lib()
    i3 = gx.int3
    # unsupported transform length, bad ndims, null buffers, zero batch
    assert L.gtomx_fft_r2c(None, None, 3, i3(100, 100, 2049), 1, None) == gx.ERR_UNSUPPORTED
    # the standalone transforms take 2^a up to 4096, 3 * 2^a up to 3072 and any other extent up
    # to 2048 (chirp-z); 100 and 320 now pass the size check and stop at the NULL buffer
    for n, ok in [(4096, True), (3072, True), (384, True), (8192, False), (6144, False), (320, True), (100, True),
                  (7, True), (2048, True), (2049, False), (2500, False)]:
        st = L.gtomx_fft_r2c(None, None, 2, i3(n, 16, 1), 1, None)
        assert st == (gx.ERR_INVALID_ARGUMENT if ok else gx.ERR_UNSUPPORTED), n
        assert ("NULL" in gx.last_error()) == ok
    assert L.gtomx_fft_r2c(None, None, 2, i3(64, 64, 64), 1, None) == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_fft_r2c(None, None, 3, i3(64, 64, 64), 1, None) == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_ifft_c2r(ctypes.c_void_p(16), ctypes.c_void_p(16), 3, i3(64, 64, 64), 0, 1, None) \
        == gx.ERR_INVALID_ARGUMENT
    assert "batch" in gx.last_error()
    h = ctypes.c_void_p()
    assert L.gtomx_pick_plan_create(ctypes.byref(h), i3(96, 96, 96), i3(64, 64, 32), 0) == gx.ERR_UNSUPPORTED
    assert L.gtomx_pick_plan_create(ctypes.byref(h), i3(95, 95, 95), i3(64, 64, 64), 0) == gx.ERR_UNSUPPORTED
    # the fused 3D search stops at 256 and the generic form takes every other side up to 1024 (sides
    # above 1024 are refused); images (dims.z == 1) go to 4096 or 3 * 2^10
    assert L.gtomx_pick_plan_create(ctypes.byref(h), i3(96, 96, 96), i3(1025, 1025, 1025), 0) == gx.ERR_UNSUPPORTED
    assert L.gtomx_pick_plan_create(ctypes.byref(h), i3(96, 96, 96), i3(1026, 1026, 1026), 0) == gx.ERR_UNSUPPORTED
    assert L.gtomx_pick_plan_create(ctypes.byref(h), i3(96, 96, 96), i3(2048, 2048, 2048), 0) == gx.ERR_UNSUPPORTED
    assert L.gtomx_pick_plan_create(ctypes.byref(h), i3(96, 96, 96), i3(8192, 8192, 1), 0) == gx.ERR_UNSUPPORTED
    assert L.gtomx_pick_plan_create(ctypes.byref(h), i3(96, 96, 96), i3(640, 640, 1), 0) == gx.ERR_UNSUPPORTED
    assert L.gtomx_pick_plan_create(ctypes.byref(h), i3(96, 96, 96), i3(384, 384, 1), 0) == gx.ERR_UNSUPPORTED
    assert L.gtomx_pick_plan_create(ctypes.byref(h), i3(96, 96, 96), i3(64, 64, 64), 1000) == \
        gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_pick_plan_create(None, i3(96, 96, 96), i3(64, 64, 64), 0) == gx.ERR_INVALID_ARGUMENT
    cnt = ctypes.c_int()
    assert L.gtomx_equal_angular_spacing(0, 1, 0, 1, 0, 1, 0, None, 0, ctypes.byref(cnt)) == \
        gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_fibonacci_angles(0, 18, None) == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_pick_plan_info(None, None, None) == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_pick_update(None, None, 1.0, i3(1, 1, 1), None, None, 0, i3(64, 64, 64), 1, None, 1, 0, None,
                               None, None) == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_pick_plan_destroy(None) == gx.OK
    assert L.gtomx_project(None, i3(8, 8, 8), None, i3(8, 8, 8), 4, None, 1, None) == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_update_max_argmax(None, 10, 1, 0, None, None, None) == gx.ERR_INVALID_ARGUMENT
    ctf = gx.CtfParams.make(pixelsize=0.0)
    assert L.gtomx_ctf_3d(ctypes.byref(ctf), 1, i3(16, 16, 16), ctypes.c_void_p(16), None) == \
        gx.ERR_INVALID_ARGUMENT


def test_prob_picker_host_paths(gx, orc):
    """The orientation count is the reference's list; creation rejects what it
    cannot run before touching the device."""
    L = gx.lib()
    i3 = gx.int3
    c = ctypes.c_int()
    for step in (0.5, 0.3, 1.0):
        assert L.gtomx_prob_picker_count_angles(2, step, ctypes.byref(c)) == gx.OK
        assert c.value == int(np.ceil(np.float32(2 * np.pi) / np.float32(step)))
    for step in (1.2, 0.5):
        assert L.gtomx_prob_picker_count_angles(3, step, ctypes.byref(c)) == gx.OK
        assert c.value == len(orc.equal_angular_spacing((0, 2 * np.pi), (0, np.pi), (0, 2 * np.pi), step))
    assert L.gtomx_prob_picker_count_angles(3, 1.2, ctypes.byref(c)) == gx.OK and c.value == 50
    assert L.gtomx_prob_picker_count_angles(4, 1.0, ctypes.byref(c)) == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_prob_picker_count_angles(2, 0.0, ctypes.byref(c)) == gx.ERR_INVALID_ARGUMENT
    h = ctypes.c_void_p()
    fake = ctypes.c_void_p(256)
    # an image too small next to the low-pass box: the remap would read unwritten memory
    assert L.gtomx_prob_picker_create(ctypes.byref(h), fake, i3(8, 8, 1), 1, fake, 1, 0, i3(16, 16, 1), 8, None) \
        == gx.ERR_UNSUPPORTED
    # low-pass box larger than the image, reference larger than the image, mixed dimensionality
    assert L.gtomx_prob_picker_create(ctypes.byref(h), fake, i3(8, 8, 1), 1, fake, 1, 0, i3(32, 32, 1), 20, None) \
        == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_prob_picker_create(ctypes.byref(h), fake, i3(40, 40, 1), 1, fake, 1, 0, i3(32, 32, 1), 4, None) \
        == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_prob_picker_create(ctypes.byref(h), fake, i3(8, 8, 8), 1, fake, 1, 0, i3(32, 32, 1), 4, None) \
        == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_prob_picker_create(ctypes.byref(h), None, i3(8, 8, 1), 1, fake, 1, 0, i3(32, 32, 1), 4, None) \
        == gx.ERR_INVALID_ARGUMENT
    assert L.gtomx_prob_picker_destroy(None) == gx.OK


@pytest.mark.parametrize("seed", [123, 1, 987654])
def test_rand_jump_ahead_equals_libc_rand(gx, seed):
    """The library's jump-ahead restatement of glibc's rand(), read at several
    offsets, equals the C runtime's own sequence after srand(seed)."""
    import ctypes.util
    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.srand(seed)
    total = 140000
    want = np.array([libc.rand() for _ in range(total)], np.int32)
    L = gx.lib()
    for first, n in ((0, 1000), (1, 31), (344, 500), (131071, 2000), (total - 700, 700)):
        out = np.empty(n, np.int32)
        L.gtomx_debug_rand_outputs(seed, first, n, out.ctypes.data_as(ctypes.c_void_p))
        assert np.array_equal(out, want[first:first + n]), (seed, first)
