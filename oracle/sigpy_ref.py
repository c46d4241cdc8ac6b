"""sigpy's Kaiser-Bessel interpolation and gridding -- TEST INFRASTRUCTURE ONLY.

``sigpy_nufft`` (fourier.py:901-963) calls ``sigpy.interp.interpolate`` and
``sigpy.interp.gridding`` with ``kernel='kaiser_bessel'`` (sigpy: third-party, listed
without a version in environment.yml:24, not installed here). Their published
definition, restated: for a sample at k (grid units, already scaled and shifted),
every grid point x with ceil(k_d - w/2) <= x_d <= floor(k_d + w/2) in each dimension
contributes prod_d KB((x_d - k_d) / (w/2)) times grid[x mod G] (interpolate), and
gridding is its adjoint. KB(u) = I0(beta sqrt(1 - u^2)) for |u| <= 1 with the
Abramowitz & Stegun I0, 0 outside. Indices wrap mod G; the coordinate is not wrapped.

This differs from the reference's pure-torch interpolation (``interp_ref``) only for
kernel widths below 2, where that path also masks |u| <= w/2 (kernels.py:119, 129).
"""

from __future__ import annotations

import itertools
import math

import torch

from .kb import kaiser_bessel_torch


def _patch(coords, width, beta, dtype):
    """Per sample and patch offset: flat grid index and weight (0 outside the support)."""
    D = coords.shape[-1]
    k = coords.to(torch.float32).to(dtype)
    r = torch.tensor(float(width) / 2.0, dtype=dtype)
    lower = torch.ceil(k - r).to(torch.int64)
    P = int(math.floor(float(width))) + 1
    offs = torch.tensor(list(itertools.product(range(P), repeat=D)), dtype=torch.int64)
    x = lower[:, None, :] + offs[None]                      # (M, P^D, D)
    u = (x.to(dtype) - k[:, None, :]) / r
    inside = u.abs() <= 1
    # beta: a number, or one per dimension (sigpy's `param` tuple)
    b = torch.as_tensor(beta, dtype=dtype) if isinstance(beta, (tuple, list)) else beta
    kb = kaiser_bessel_torch(torch.where(inside, u, torch.zeros_like(u)), b)
    w = torch.where(inside, kb, torch.zeros_like(kb)).prod(-1)
    return x, w


def _flat(x, grid_size):
    G = torch.tensor(grid_size, dtype=torch.int64)
    xi = torch.remainder(x, G)
    idx = torch.zeros(xi.shape[:-1], dtype=torch.int64)
    for d in range(len(grid_size)):
        idx = idx * grid_size[d] + xi[..., d]
    return idx


def interpolate(grid, coords, width, beta, dtype=torch.float32):
    """grid (B, *G) -> (B, M)."""
    D = coords.shape[-1]
    G = tuple(grid.shape[-D:])
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    x, w = _patch(coords.reshape(-1, D), width, beta, dtype)
    flat = grid.reshape(grid.shape[0], -1).to(cdt)
    vals = flat[:, _flat(x, G)]                              # (B, M, P^D)
    return (vals * w[None]).sum(-1)


def gridding(vals, coords, grid_size, width, beta, dtype=torch.float32):
    """vals (B, M) -> (B, *grid_size), the adjoint of ``interpolate``."""
    D = coords.shape[-1]
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    x, w = _patch(coords.reshape(-1, D), width, beta, dtype)
    idx = _flat(x, tuple(grid_size)).reshape(-1)
    B = vals.shape[0]
    out = torch.zeros((B, math.prod(grid_size)), dtype=cdt)
    contrib = (vals.to(cdt)[:, :, None] * w[None]).reshape(B, -1)
    out.index_add_(1, idx, contrib)
    return out.reshape(B, *grid_size)
