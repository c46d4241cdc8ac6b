"""FP64 direct NDFT -- the closed-form definition of the transform.

y_m = N^{-d/2} sum_r x(r) exp(-2 pi i k_m . r), r = gen_grd(im_size), i.e.
r_i = (n_i - N_i//2) / N_i (reference utils.py:166-200; matrix_nufft
fourier.py:637-640; tests/test_nufft_scaling.py:61-62).
"""

from __future__ import annotations

import math

import numpy as np


def pixel_grid(im_size) -> np.ndarray:
    lins = [(np.arange(n) - n // 2) / n for n in im_size]
    g = np.meshgrid(*lins, indexing="ij")
    return np.stack(g, axis=-1).reshape(-1, len(im_size))


def ndft_forward(x: np.ndarray, k: np.ndarray, im_size, chunk: int = 4096) -> np.ndarray:
    """x (..., *im_size) complex, k (M, d) -> (..., M) complex128."""
    d = len(im_size)
    r = pixel_grid(im_size)
    R = r.shape[0]
    xb = np.asarray(x, dtype=np.complex128).reshape(-1, R)
    k = np.asarray(k, dtype=np.float64).reshape(-1, d)
    out = np.empty((xb.shape[0], k.shape[0]), dtype=np.complex128)
    for m0 in range(0, k.shape[0], chunk):
        ph = k[m0:m0 + chunk] @ r.T
        out[:, m0:m0 + chunk] = xb @ np.exp(-2j * np.pi * ph).T
    out /= math.sqrt(R)
    return out.reshape(*np.shape(x)[:-d], k.shape[0])


def ndft_adjoint(y: np.ndarray, k: np.ndarray, im_size, chunk: int = 4096) -> np.ndarray:
    """y (..., M) complex, k (M, d) -> (..., *im_size) complex128."""
    d = len(im_size)
    r = pixel_grid(im_size)
    R = r.shape[0]
    k = np.asarray(k, dtype=np.float64).reshape(-1, d)
    yb = np.asarray(y, dtype=np.complex128).reshape(-1, k.shape[0])
    out = np.zeros((yb.shape[0], R), dtype=np.complex128)
    for m0 in range(0, k.shape[0], chunk):
        ph = k[m0:m0 + chunk] @ r.T
        out += yb[:, m0:m0 + chunk] @ np.exp(2j * np.pi * ph)
    out /= math.sqrt(R)
    return out.reshape(*np.shape(y)[:-1], *im_size)
