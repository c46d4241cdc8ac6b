"""Reference CPU interpolation / gridding (torch on CPU).

Restates the reference's pure-torch path, which is what ``triton_interp``
runs whenever tensors are not on a GPU (ungrid.py:82-123, grid.py:86-127):
  ``ungrid_torch``  ungrid.py:897-957   patch gather over (w+1)^d points
  ``grid_torch``    grid.py:556-624     index_put_(accumulate=True) scatter
  ``weights_torch`` kernels.py:101-140  norm '1', NaN -> 0, circular wrap
  shape handling    prep_ungrid_shapes ungrid.py:851-894, prep_grid_shapes grid.py:508-553
``dtype`` selects the arithmetic (float32 = the reference's; float64 = a
higher-precision twin used to measure rounding).
"""

from __future__ import annotations

import math

import torch

from .kb import kaiser_bessel_torch, kaiser_bessel_torch_safe


def _weights(locs, grid_locs, radius, beta, grid_size, safe=False):
    """kernels.py:101-140 for norm '1' / circular: separable product, non-finite -> 0."""
    dlocs = (locs[:, None] - grid_locs) / radius
    kb = kaiser_bessel_torch_safe if safe else kaiser_bessel_torch
    w = kb(dlocs, beta).prod(dim=-1)
    if not safe:
        w = torch.where(torch.isfinite(w), w, torch.zeros_like(w))
    mask = (torch.abs(dlocs) <= radius).all(dim=-1)
    idx = torch.remainder(grid_locs, grid_size).to(torch.int64)
    return w, idx, mask


def ungrid_torch(vals, locs, grid_size, width, beta, batch_size=2 ** 20, dtype=torch.float32, safe=False):
    """vals (B, *G) complex, locs (npts, d) already wrapped -> (B, npts) (ungrid.py:897-957)."""
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    vals = vals.to(cdt)
    locs = locs.to(dtype) if locs.dtype != torch.float64 else locs
    nbatch = vals.shape[0]
    ndim = locs.shape[-1]
    npts = locs.shape[0]
    gs = torch.tensor(grid_size)
    diff = torch.meshgrid(*(torch.arange(int(w) + 1) for w in width), indexing="ij")
    diff = torch.stack(diff, dim=-1).reshape(-1, ndim)
    radius = torch.tensor(width, dtype=dtype) / 2
    lower = torch.ceil(locs - radius).to(torch.int64)
    out = torch.zeros(nbatch, npts, dtype=cdt)
    for p0 in range(0, npts, batch_size):
        p1 = min(p0 + batch_size, npts)
        gl = lower[p0:p1, None] + diff
        w, idx, mask = _weights(locs[p0:p1], gl, radius, beta, gs, safe)
        sl = (slice(None),) + tuple(idx[..., i] for i in range(ndim))
        out[:, p0:p1] = torch.sum(w * vals[sl] * mask, dim=-1)
    return out


def grid_torch(vals, locs, grid_size, width, beta, batch_size=2 ** 20, dtype=torch.float32, safe=False):
    """vals (B, npts) complex, locs (npts, d) wrapped -> (B, *G) (grid.py:556-624)."""
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    vals = vals.to(cdt)
    locs = locs.to(dtype) if locs.dtype != torch.float64 else locs
    ndim = locs.shape[-1]
    nbatch = vals.shape[0]
    gs = torch.tensor(grid_size)
    diff = torch.meshgrid(*(torch.arange(int(w) + 1) for w in width), indexing="ij")
    diff = torch.stack(diff, dim=-1).reshape(-1, ndim)
    radius = torch.tensor(width, dtype=dtype) / 2
    lower = torch.ceil(locs - radius).to(torch.int64)
    out = torch.zeros(nbatch, *grid_size, dtype=cdt)
    npts = vals.shape[1]
    bidx = torch.arange(nbatch).view(-1, 1, 1)
    for p0 in range(0, npts, batch_size):
        p1 = min(p0 + batch_size, npts)
        gl = lower[p0:p1, None] + diff
        w, idx, mask = _weights(locs[p0:p1], gl, radius, beta, gs, safe)
        patches = w * mask * vals[:, p0:p1, None]
        b = bidx.expand(-1, p1 - p0, patches.shape[-1])
        out.index_put_((b, *(idx[..., i] for i in range(ndim))), patches, accumulate=True)
    return out


def _width(width, ndim):
    return (float(width),) * ndim if isinstance(width, (int, float)) else tuple(float(w) for w in width)


def _coord_dtype(locs, coords64):
    return torch.float64 if coords64 else torch.float32


def ungrid(vals, locs, width, beta=1.0, dtype=torch.float32, batch_size=2 ** 20, coords64=False, safe=False):
    """Full ``ungrid`` surface (ungrid.py:34-63): vals (..., *G), locs (..., d) in grid units.

    ``coords64`` keeps coordinates in float64 (the reference rounds them to fp32),
    for finite-difference checks; ``safe`` masks before the square root so that
    autograd through the weights is finite."""
    ndim = locs.shape[-1]
    grid_size = tuple(vals.shape[-ndim:])
    batch_shape = vals.shape[:-ndim]
    vflat = vals.reshape(-1, *grid_size)
    wl = torch.remainder(locs.to(_coord_dtype(locs, coords64)), torch.tensor(grid_size))
    out = ungrid_torch(vflat, wl.reshape(-1, ndim), grid_size, _width(width, ndim), beta,
                       batch_size=batch_size, dtype=dtype, safe=safe)
    return out.reshape(*batch_shape, *locs.shape[:-1])


def grid(vals, locs, grid_size, width, beta=1.0, dtype=torch.float32, batch_size=2 ** 20, coords64=False,
         safe=False):
    """Full ``grid`` surface (grid.py:32-67)."""
    ndim = len(grid_size)
    locs_batch = locs.shape[:-1]
    batch_shape = vals.shape[: vals.ndim - len(locs_batch)]
    vflat = vals.reshape(int(math.prod(batch_shape)), -1)
    wl = torch.remainder(locs.to(_coord_dtype(locs, coords64)), torch.tensor(grid_size))
    out = grid_torch(vflat, wl.reshape(-1, ndim), tuple(grid_size), _width(width, ndim), beta,
                     batch_size=batch_size, dtype=dtype, safe=safe)
    return out.reshape(*batch_shape, *grid_size)


def ungrid_bruteforce(vals, locs, width, beta):
    """Pure-Python definition on tiny inputs: every grid point g with |g_d - t_d| <= w/2
    (g taken in all periodic images) contributes prod_d I0(beta sqrt(1 - u_d^2)) vals[g mod G]."""
    import numpy as np

    from .kb import i0_as_np

    ndim = locs.shape[-1]
    G = vals.shape[-ndim:]
    v = vals.reshape(-1, *G).numpy().astype(np.complex128)
    t = torch.remainder(locs.to(torch.float32), torch.tensor(G)).reshape(-1, ndim).numpy().astype(np.float64)
    r = width / 2.0
    out = np.zeros((v.shape[0], t.shape[0]), dtype=np.complex128)
    for m in range(t.shape[0]):
        pts = []
        for d in range(ndim):
            cand = [g for g in range(int(math.floor(t[m, d] - r)) - 1, int(math.ceil(t[m, d] + r)) + 2)
                    if abs(g - t[m, d]) <= r]
            pts.append([(g % G[d], float(i0_as_np(beta * math.sqrt(max(1 - ((g - t[m, d]) / r) ** 2, 0.0)))))
                        for g in cand])
        for combo in __import__("itertools").product(*pts):
            idx = tuple(c[0] for c in combo)
            w = math.prod(c[1] for c in combo)
            out[:, m] += w * v[(slice(None),) + idx]
    return out
