"""The split-precision matrix-core transforms (gtom_amd/csrc/split.hip) and the
search that runs on them, against FP64.

- gtomx_split_dft64 against numpy's inverse DFT in complex128 on lines whose
  magnitudes span 1e-6 .. 1e6 (the per-column power-of-two scaling): the
  per-line error max|out - ref| / max|ref| stays below 5e-6 (fp16 hi/lo) and
  1e-4 (bf16 hi/lo), and the compensation gains at least 10x over hi alone.
- The 64^3 configuration (3 planted copies, 588 orientations, CTF) searched
  with the split transforms in fp16 and in bf16: the normalised max-CC map
  within 1e-4 normwise of the FP64 oracle and the argmax identical wherever the
  oracle's top-2 margin exceeds 1e-4 * max|top1| (the plain path's bar), and
  the uncompensated fp16 search measurably worse than the split one.
Each test prints its errors, which are the figures quoted for the precision
configuration. The reference runs these transforms through cuFFT
(src/FFT/IFFT.cu:48-56); it has no reduced-precision form.
"""
import numpy as np
import pytest

from tests._util import argmax_agreement, rel_linf, to_np

pytestmark = pytest.mark.gpu

CC_TOL = 1e-4


def _line_errors(out, ref):
    return np.abs(out - ref).max(axis=1) / np.abs(ref).max(axis=1)


def test_split_dft64_error_per_precision(gx, cuda):
    import torch
    rng = np.random.default_rng(21)
    n = 4096
    x = (rng.standard_normal((n, 64)) + 1j * rng.standard_normal((n, 64)))
    x *= 10.0 ** rng.uniform(-6, 6, size=(n, 1))
    x[:16, 1::2] = 0  # sparse lines: some subsequences are all zero
    x = x.astype(np.complex64)
    ref = np.fft.ifft(x.astype(np.complex128), axis=1) * 64
    xt = torch.from_numpy(x).to(cuda)
    errs = {}
    for name, prec in gx.PRECISIONS.items():
        out = to_np(gx.split_dft64(xt, prec)).astype(np.complex128)
        errs[name] = float(_line_errors(out, ref).max())
    print("split dft64 per-line error vs FP64:", {k: f"{v:.2e}" for k, v in errs.items()})
    assert errs["fp16"] <= 5e-6, errs
    assert errs["bf16"] <= 1e-4, errs
    assert errs["fp16"] * 10 <= errs["fp16-hi"], errs
    assert errs["bf16"] * 10 <= errs["bf16-hi"], errs


def test_split_dft64_in_place(gx, cuda):
    import torch
    rng = np.random.default_rng(2)
    x = (rng.standard_normal((37, 64)) + 1j * rng.standard_normal((37, 64))).astype(np.complex64)
    xt = torch.from_numpy(x).to(cuda)
    want = to_np(gx.split_dft64(xt, gx.SPLIT_FP16))
    gx.check(gx.lib().gtomx_split_dft64(gx._ptr(xt), gx._ptr(xt), 37, gx.SPLIT_FP16, None), "in place")
    torch.cuda.synchronize()
    assert np.array_equal(to_np(xt), want)


@pytest.fixture(scope="module")
def cfg1_case(gx, orc, cuda):
    """The 64^3 configuration and its FP64 oracle result (computed once)."""
    import torch
    from gtom_amd import synth
    w = synth.cfg1(seed=1)
    n, p, s = w["n"], w["p"], w["oversample"]
    projft = gx.prepare_projector(torch.from_numpy(w["template"]).to(cuda), p)
    exp_ft = gx.d_FFTR2C(torch.from_numpy(w["tomogram"][None]).to(cuda))
    ctf_np = orc.ctf3(n, 10e-10, 300e3, 2.7e-3, -3e-6, 0.07).astype(np.float32)
    ctf = torch.from_numpy(ctf_np).to(cuda)
    ref = orc.pick(to_np(projft).astype(np.complex128), p, s, to_np(exp_ft).astype(np.complex128),
                   ctf_np.astype(np.float64), n, w["angles"].astype(np.float64), w["maskradius"])
    return w, projft, exp_ft, ctf, ref


def _search(gx, w, projft, exp_ft, ctf, prec):
    import torch
    plan = gx.PickPlan(w["p"], w["n"], 64)
    best, idx = gx.d_PickSubTomogramsSplit(plan, prec, projft, w["oversample"], exp_ft, ctf, w["angles"],
                                           w["maskradius"])
    torch.cuda.synchronize()
    plan.close()
    return to_np(best), to_np(idx)


@pytest.mark.parametrize("name", ["fp16", "bf16"])
def test_split_search_cfg1_matches_oracle(gx, cuda, cfg1_case, name):
    w, projft, exp_ft, ctf, ref = cfg1_case
    best, idx = _search(gx, w, projft, exp_ft, ctf, gx.PRECISIONS[name])
    err = rel_linf(best, ref["best"])
    agree, cnt = argmax_agreement(idx, ref["idx"], ref["top1"], ref["top2"], CC_TOL)
    print(f"split search ({name}): max-CC error vs FP64 {err:.2e}, argmax agreement {agree:.6f} on {cnt} voxels")
    assert err <= CC_TOL, err
    assert cnt > 0 and agree == 1.0, (agree, cnt)


def test_split_search_compensation_matters(gx, cuda, cfg1_case):
    w, projft, exp_ft, ctf, ref = cfg1_case
    e_split = rel_linf(_search(gx, w, projft, exp_ft, ctf, gx.SPLIT_FP16)[0], ref["best"])
    e_hi = rel_linf(_search(gx, w, projft, exp_ft, ctf, gx.SINGLE_FP16)[0], ref["best"])
    e_bhi = rel_linf(_search(gx, w, projft, exp_ft, ctf, gx.SINGLE_BF16)[0], ref["best"])
    print(f"split search max-CC error vs FP64: fp16 hi/lo {e_split:.2e}, fp16 hi {e_hi:.2e}, bf16 hi {e_bhi:.2e}")
    assert e_split * 10 <= e_hi, (e_split, e_hi)
    assert e_hi < e_bhi
