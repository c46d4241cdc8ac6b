"""Pipe-Menon density compensation -- TEST INFRASTRUCTURE ONLY (part of the oracle).

``density_compensation`` (algs.py:20-106) hands the trajectory to
``sigpy.mri.pipe_menon_dcf`` (sigpy: a third-party dependency, listed without a
version in environment.yml:24 and not installed here). Its published algorithm
(Pipe & Menon, MRM 41:179-186, 1999, in the form sigpy implements it) is restated:

    w = 1;  max_iter times:  w <- w / |G^H G w|
    G: Kaiser-Bessel gridding onto img_shape, width 4, beta 8, kernel per dimension
       I0(beta sqrt(1 - u^2)) with u = (x - k) / (width / 2), |u| <= 1 (I0 from
       Abramowitz & Stegun 9.8.1/9.8.2); coordinates in grid units, grid indices
       wrapping mod N, no oversampling and no shift
    dcf = w / max(w)                                              (algs.py:105)

G and G^H come from ``interp_ref`` (the reference's pure-torch gridding and
interpolation, pinned there against a brute-force sum); ``gram_dense`` builds
G^H G as a dense matrix from the kernel definition alone, so the iteration can
be checked without either.
"""

from __future__ import annotations

import itertools
import math

import numpy as np
import torch

from . import interp_ref
from .kb import i0_as_np

WIDTH = 4.0
BETA = 8.0


def pipe_menon(coord: torch.Tensor, img_shape, max_iter: int = 30, width: float = WIDTH, beta: float = BETA,
               normalize: bool = True, dtype=torch.float64) -> torch.Tensor:
    """dcf of shape coord.shape[:-1] (float64 arithmetic by default; float32 = the
    reference's arithmetic)."""
    D = coord.shape[-1]
    flat = coord.reshape(-1, D).to(torch.float32)
    cdt = torch.complex128 if dtype == torch.float64 else torch.complex64
    w = torch.ones(flat.shape[0], dtype=dtype)
    for _ in range(max_iter):
        g = interp_ref.grid(w[None].to(cdt), flat, tuple(img_shape), width, beta, dtype=dtype)
        ghg = interp_ref.ungrid(g, flat, width, beta, dtype=dtype)[0]
        w = w / ghg.abs().to(dtype)
    if normalize:
        w = w / w.max()
    return w.reshape(coord.shape[:-1])


def gridding_matrix(coord: torch.Tensor, img_shape, width: float = WIDTH, beta: float = BETA) -> np.ndarray:
    """Dense G^T (M, prod N) in float64 straight from the kernel definition: entry
    (m, g) sums, over every periodic image of grid point g within width/2 of sample
    m in each dimension, prod_d I0(beta sqrt(1 - u_d^2)). Tiny inputs only."""
    D = coord.shape[-1]
    t = coord.reshape(-1, D).to(torch.float32).numpy().astype(np.float64)
    G = tuple(int(n) for n in img_shape)
    r = width / 2.0
    A = np.zeros((t.shape[0], math.prod(G)))
    for m in range(t.shape[0]):
        per_dim = []
        for d in range(D):
            pts = []
            for g in range(int(math.floor(t[m, d] - r)) - 1, int(math.ceil(t[m, d] + r)) + 2):
                u = (g - t[m, d]) / r
                if abs(u) <= 1.0:
                    pts.append((g % G[d], float(i0_as_np(beta * math.sqrt(max(1.0 - u * u, 0.0))))))
            per_dim.append(pts)
        for combo in itertools.product(*per_dim):
            A[m, np.ravel_multi_index(tuple(c[0] for c in combo), G)] += math.prod(c[1] for c in combo)
    return A


def pipe_menon_dense(coord: torch.Tensor, img_shape, max_iter: int = 30, width: float = WIDTH,
                     beta: float = BETA, normalize: bool = True) -> np.ndarray:
    """The same iteration with G^H G = A A^T formed densely (float64)."""
    A = gridding_matrix(coord, img_shape, width, beta)
    gram = A @ A.T
    w = np.ones(A.shape[0])
    for _ in range(max_iter):
        w = w / np.abs(gram @ w)
    if normalize:
        w = w / w.max()
    return w.reshape(tuple(coord.shape[:-1]))
