"""Coordinate scaling, wrapping, lower corners and tile binning (numpy, bit-level).

Scaling follows ``triton_nufft._scale_trj`` (fourier.py:420-430):
    t = fl32(fl32(k * fl32(ceil(s N)/N)) + ceil(s N)//2)
wrapping ``prep_ungrid_shapes`` (ungrid.py:864-865) / ``prep_grid_shapes``
(grid.py:516-517): t = torch.remainder(t, G), which for fp32 is fmod(t, G)
plus G when the result is non-zero and negative; and the patch corner
``get_neighborhood`` (kernels.py:531-536) / ``ungrid_torch`` (ungrid.py:930):
lower = ceil(fl32(t - w/2)). The tile binning (lower corner, shifted by
``wrap`` and wrapped mod G, divided by the tile edge; within a tile ordered by
the corner's position in the tile, then by sample index) has no reference
counterpart: it is the build's own plan and is defined here.
"""

from __future__ import annotations

import math

import numpy as np

f32 = np.float32


def scale_coords(k: np.ndarray, scale, shift) -> np.ndarray:
    k = np.asarray(k, dtype=f32)
    out = np.empty_like(k)
    for d in range(k.shape[-1]):
        v = (k[..., d] * f32(scale[d])).astype(f32)
        out[..., d] = (v + f32(shift[d])).astype(f32)
    return out


def wrap_coords(t: np.ndarray, grid_size) -> np.ndarray:
    t = np.asarray(t, dtype=f32)
    out = np.empty_like(t)
    for d, G in enumerate(grid_size):
        with np.errstate(invalid="ignore"):
            m = np.fmod(t[..., d], f32(G)).astype(f32)
            out[..., d] = np.where((m != 0) & (m < 0), (m + f32(G)).astype(f32), m)
    return out


def lower_corner(t: np.ndarray, width: float) -> np.ndarray:
    with np.errstate(invalid="ignore"):
        return np.ceil((t - f32(width / 2.0)).astype(f32))


def tile_keys(t: np.ndarray, grid_size, width, tile, wrap, with_local: bool = False):
    """Linear tile id per sample (non-finite samples get ntiles) and, with
    ``with_local``, the position of the wrapped lower corner inside its tile (0 for
    non-finite samples): row-major in 1D/2D; in 3D ordered by 1x2x4 cell of corners
    first (corner plane, row // 2, column // 4), then by position in the cell."""
    ntile = [(g + tile - 1) // tile for g in grid_size]
    ntiles = math.prod(ntile)
    ok = np.all(np.isfinite(t), axis=-1)
    lo = lower_corner(t, width)
    lo = np.where(np.isfinite(lo), lo, 0).astype(np.int64)
    key = np.zeros(t.shape[0], dtype=np.int64)
    loc = np.zeros(t.shape[0], dtype=np.int64)
    a = []
    for d, G in enumerate(grid_size):
        lw = np.mod(lo[:, d] + wrap[d], G)
        key = key * ntile[d] + lw // tile
        a.append(lw % tile)
        loc = loc * tile + lw % tile
    if len(grid_size) == 3:
        # cell-major: cell (a0, a1 // 2, a2 // 4), then (a1 % 2, a2 % 4)
        loc = (((a[0] * (tile // 2) + a[1] // 2) * (tile // 4) + a[2] // 4) * 2 + a[1] % 2) * 4 + a[2] % 4
    key = np.where(ok, key, ntiles)
    if with_local:
        return key, np.where(ok, loc, 0), ntiles
    return key, ntiles


def bin_plan(coords, grid_size, width, tile, wrap=None, scale=None, shift=None):
    """Returns (tsort (M*D,), perm (M,), tile_start (ntiles+2,), keys)."""
    D = len(grid_size)
    k = np.asarray(coords, dtype=f32).reshape(-1, D)
    t = scale_coords(k, scale, shift) if scale is not None else k
    t = wrap_coords(t, grid_size)
    wrap = wrap if wrap is not None else (0,) * D
    key, loc, ntiles = tile_keys(t, grid_size, width, tile, wrap, with_local=True)
    perm = np.lexsort((np.arange(key.size), loc, key))  # by tile, corner position, sample index
    counts = np.bincount(key, minlength=ntiles + 1)
    start = np.concatenate([[0], np.cumsum(counts)])
    return t[perm].reshape(-1), perm.astype(np.int32), start.astype(np.int32), key
