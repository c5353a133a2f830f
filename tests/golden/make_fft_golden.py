"""Writes tests/golden/fft_vectors.npz: seeded complex inputs and their FP64
DFTs computed by numpy's pocketfft (numpy 2.2) for 1-D lengths with factors
2, 3, 5 and one 3-D array.  Run once: python tests/golden/make_fft_golden.py
"""
import os

import numpy as np

SIZES = [1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 20, 24, 25, 27, 30, 32, 40, 45, 48, 50, 60, 64, 75, 80, 96,
         100, 125, 128, 160, 200, 256, 400]

if __name__ == "__main__":
    rng = np.random.default_rng(2024)
    out = {}
    for n in SIZES:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        out["x%d" % n] = x
        out["X%d" % n] = np.fft.fft(x)
    x3 = rng.standard_normal((6, 10, 12)) + 1j * rng.standard_normal((6, 10, 12))
    out["x3d"] = x3
    out["X3d"] = np.fft.fftn(x3)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_vectors.npz"), **out)
    print("wrote", len(SIZES), "vectors")
