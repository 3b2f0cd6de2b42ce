"""Writes fft_golden.npz: seeded complex inputs and their unnormalised forward
and backward 3D DFTs computed by numpy.fft (pocketfft, FP64).  numpy's sign
convention (forward exp(-2 pi i k.x/N)) is the one of G.c:209-273.

Run from the repo root:  python tests/golden/make_fft_golden.py
"""
import os

import numpy as np

SHAPES = [(8, 6, 4), (6, 6, 6), (16, 2, 16), (12, 9, 8), (32, 4, 3), (5, 10, 3)]


def main():
    out = {}
    rng = np.random.default_rng(20261202)
    for i, s in enumerate(SHAPES):
        x = rng.standard_normal(s) + 1j * rng.standard_normal(s)
        out[f"x{i}"] = x
        out[f"fwd{i}"] = np.fft.fftn(x)
        out[f"bwd{i}"] = np.fft.ifftn(x) * np.prod(s)
    out["shapes"] = np.array(SHAPES)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
