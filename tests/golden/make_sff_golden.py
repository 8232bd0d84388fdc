"""Writes tests/golden/sff_golden.npz: scipy.fft DCT-II / DST-II reference
vectors for the SFF transform convention (sample grid x_j = (j + 1/2) L / N,
A_m = DCT-II_m / (2N) for 'C', A_m = DST-II_{m-1} / (2N) for 'S' with
A_0 = 0), applied along x after the FFF r2c along z and complex FFT along y
(scaled 1/(Ny Nz)).  Inputs are seeded random real fields."""
import os

import numpy as np
from scipy import fft as sf

HERE = os.path.dirname(os.path.abspath(__file__))


def sff_reference(r, parity):
    nx, ny, nz = r.shape
    B = np.fft.fft(np.fft.rfft(r, axis=2), axis=1) / (ny * nz)
    if parity == 1:
        return sf.dct(B, type=2, axis=0) / (2 * nx)
    Y = sf.dst(B, type=2, axis=0) / (2 * nx)
    A = np.zeros_like(Y)
    A[1:] = Y[:-1]
    return A


def main():
    rng = np.random.default_rng(20240611)
    out = {}
    for shape in [(8, 4, 8), (16, 8, 4)]:
        r = rng.standard_normal(shape)
        key = "x".join(map(str, shape))
        out[f"real_{key}"] = r
        out[f"sin_{key}"] = sff_reference(r, 0)
        out[f"cos_{key}"] = sff_reference(r, 1)
    np.savez_compressed(os.path.join(HERE, "sff_golden.npz"), **out)


if __name__ == "__main__":
    main()
