"""Writes tests/golden/fft_golden.npz: numpy FP64 FFT fixtures for the FFF transforms.

Forward: rfftn over (x, y, z) of a seeded real field, divided by Nx*Ny*Nz
(the forward normalisation of the FFF basis).  Inverse: irfftn of a seeded
random half spectrum, multiplied by Nx*Ny*Nz (unscaled inverse); numpy's
c2r ignores the imaginary parts of the kz = 0 and kz = Nz/2 coefficients.
numpy (pocketfft) is an independent library, not the reference.
"""
import os

import numpy as np

SHAPES = [(8, 8, 8), (16, 8, 32), (4, 8, 16), (16, 16, 16), (6, 10, 12), (15, 9, 20)]


def main():
    out = {}
    for n, shp in enumerate(SHAPES):
        rng = np.random.default_rng(100 + n)
        nx, ny, nz = shp
        r = rng.standard_normal(shp)
        out[f"real_{n}"] = r
        out[f"spec_{n}"] = np.fft.rfftn(r, axes=(0, 1, 2)) / (nx * ny * nz)
        S = rng.standard_normal((nx, ny, nz // 2 + 1)) + 1j * rng.standard_normal((nx, ny, nz // 2 + 1))
        out[f"S_{n}"] = S
        out[f"inv_{n}"] = np.fft.irfftn(S, s=shp, axes=(0, 1, 2)) * (nx * ny * nz)
        out[f"shape_{n}"] = np.array(shp)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
