"""Writes tests/golden/r2r_golden.npz: known-answer vectors for the four FFTW
r2r kinds the reference uses (quasi/quasi.cpp:449-495).

SciPy's unnormalised scipy.fft.dct/dst of types 2 and 3 are the same
definitions as FFTW's REDFT10/REDFT01 and RODFT10/RODFT01 (FFTW 3.3 manual,
"What FFTW Really Computes"); FFTW itself is not installed here.

Run:  python tests/golden/make_r2r_golden.py   (numpy + scipy, seed 20261203)
"""
import os

import numpy as np
import scipy
from scipy.fft import dct, dst

LENGTHS = (1, 2, 3, 8, 12, 32, 48, 64, 96, 128, 256, 512)
KINDS = {4: (dct, 3), 5: (dct, 2), 8: (dst, 3), 9: (dst, 2)}  # FFTW kind values
SEED = 20261203


def main():
    rng = np.random.default_rng(SEED)
    out = {"lengths": np.array(LENGTHS), "seed": np.array(SEED),
           "scipy_version": np.array(scipy.__version__)}
    for n in LENGTHS:
        x = rng.standard_normal(n)
        out[f"x_{n}"] = x
        for kind, (fn, t) in KINDS.items():
            out[f"y_{n}_{kind}"] = fn(x, type=t, norm=None)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "r2r_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
