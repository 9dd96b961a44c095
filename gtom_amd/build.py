"""Builds gtom_amd/libgtomx.so (hand-written HIP kernels + C ABI) for gfx950.

Each source is compiled with hipcc to an object in gtom_amd/_build/ and the
objects are linked into one shared library next to this file. Run as
`python -m gtom_amd.build` or through __graft_entry__.build().
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgtomx.so")
SOURCES = ["runtime.cpp", "fft.hip", "ops.hip", "pick.hip", "peaks.hip", "picker.hip", "mrc.hip", "diff2.hip", "peak_subvoxel.hip", "rotate.hip", "norm.hip", "resize.hip", "probpicker.hip", "learn.hip", "split.hip", "mpick.hip", "tomopicker.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GTOMX_ARCH", "gfx950")
# Packed FP32 instructions stay disabled in all device code.
# Code object v5 loads on every ROCm >= 5 runtime, including the one bundled with PyTorch.
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-mcode-object-version=5",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-I", os.path.join(HERE, "..", "include")]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for o, err in results:
            warn = "\n".join(l for l in err.splitlines() if "packed-fp32-ops" not in l and l.strip())
            if warn:
                print(f"[{os.path.basename(o)}]\n{warn}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
