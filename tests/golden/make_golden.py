"""Writes tests/golden/fft_golden.npz: inputs and numpy (pocketfft, float64) outputs
for the R2C / C2R conventions of the path.

R2C is numpy.fft.rfftn (unnormalised forward). C2R_u is numpy.fft.irfftn times N:
numpy transforms the y and z axes as complex inverses and then takes the last
axis as a real inverse that ignores the imaginary parts of its DC and Nyquist
bins, which is the real part of the Hermitian-completed inverse. The C2R inputs
are deliberately non-Hermitian on the x = 0 and x = n/2 planes.
Run: python tests/golden/make_golden.py
"""
import os

import numpy as np

SHAPES = [(8, 8, 8), (4, 8, 16), (6, 12, 24), (1, 16, 16), (2, 3, 8)]


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for k, (nz, ny, nx) in enumerate(SHAPES):
        x = rng.standard_normal((nz, ny, nx))
        out[f"r2c_in_{k}"] = x
        out[f"r2c_out_{k}"] = np.fft.rfftn(x)
        X = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
        out[f"c2r_in_{k}"] = X
        out[f"c2r_out_{k}"] = np.fft.irfftn(X, s=(nz, ny, nx), axes=(0, 1, 2)) * (nz * ny * nx)
    out["shapes"] = np.array(SHAPES)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
