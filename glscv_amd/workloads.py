"""Synthetic inputs of the measured workloads (numpy, seeded)."""
from __future__ import annotations

import numpy as np


def tones_2d(rows: int, cols: int, seed: int = 0, n_tones: int = 8, noise_sigma: float = 1e-2,
             return_tones: bool = False):
    """Sum of `n_tones` complex sinusoids with random (non-integer) 2D
    frequencies, random phases and amplitudes log-uniform over 1e-3..1e3, plus
    complex Gaussian noise of standard deviation `noise_sigma` per component.
    The spectrum spans about twelve orders of magnitude in power: a test of an
    FFT's dynamic range.  Built in FP64 (each tone is an outer product of two
    1D exponentials) and returned as complex64; with `return_tones`, also the
    list of (fy, fx, amplitude) of the tones (a forward FFT puts tone f's
    energy around bin f)."""
    rng = np.random.default_rng(seed)
    y = np.arange(rows, dtype=np.float64)[:, None]
    x = np.arange(cols, dtype=np.float64)[None, :]
    out = np.zeros((rows, cols), dtype=np.complex128)
    tones = []
    for _ in range(n_tones):
        fy, fx = rng.uniform(0, rows), rng.uniform(0, cols)
        amp = 10.0 ** rng.uniform(-3, 3)
        ph = rng.uniform(0, 2 * np.pi)
        out += amp * np.exp(1j * (2 * np.pi * fy * y / rows + ph)) * np.exp(1j * 2 * np.pi * fx * x / cols)
        tones.append((fy, fx, amp))
    out += noise_sigma * (rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols)))
    out = out.astype(np.complex64)
    return (out, tones) if return_tones else out
