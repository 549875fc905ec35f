"""Writes the golden FFT vectors in this folder (numpy.fft = pocketfft, FP64).

Run: python tests/golden/gen_golden.py
Each fixture <name>.npz holds
  x      : complex64 input (rows, cols)
  fwd    : fft2(x)                         (unscaled forward, FP64)
  inv    : ifft2(x) * rows * cols          (GLS_FFT_INVERSE, unnormalised)
  fwd_shift : fftshift(fft2(x))            (GLS_FFT_SHIFT)
  inv_scale : ifft2(x)                     (GLS_FFT_INVERSE | GLS_FFT_SCALE)
Inputs: re, im i.i.d. U[-1,1) from numpy.random.default_rng(seed) (PCG64),
seed 20261203 plus the fixture index.  "cfg1_256x256" is the 256x256 plain
small workload; fixtures above 8192 elements keep only x and fwd.
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = [("16x16", 16, 16), ("32x32", 32, 32), ("cfg1_256x256", 256, 256), ("1x32", 1, 32),
         ("64x1", 64, 1), ("64x128", 64, 128), ("8x4", 8, 4)]
SEED = 20261203


def make(idx, rows, cols):
    rng = np.random.default_rng(SEED + idx)
    x = (rng.uniform(-1, 1, (rows, cols)) + 1j * rng.uniform(-1, 1, (rows, cols))).astype(np.complex64)
    xd = x.astype(np.complex128)
    axes = (0, 1)
    return dict(x=x, fwd=np.fft.fft2(xd, axes=axes), inv=np.fft.ifft2(xd, axes=axes) * rows * cols,
                fwd_shift=np.fft.fftshift(np.fft.fft2(xd, axes=axes), axes=axes),
                inv_scale=np.fft.ifft2(xd, axes=axes))


def main():
    for i, (name, r, c) in enumerate(CASES):
        d = make(i, r, c)
        if r * c > 8192:   # keep large fixtures small: input and forward output only
            d = {"x": d["x"], "fwd": d["fwd"]}
        np.savez_compressed(os.path.join(HERE, f"fft_{name}.npz"), **d)
    print("wrote", len(CASES), "fixtures")


if __name__ == "__main__":
    main()
