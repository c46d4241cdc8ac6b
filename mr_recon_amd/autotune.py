"""Autotuner for the work decomposition of the KB interp / gridding kernels.

Two knobs vary with the trajectory. ``samples_per_item``: small items spread hot
tiles (radial centres) over more blocks, large items cut the per-item subgrid
loads and the gridding scratch. The tile edge (2D: 8, 16 or 32 grid points; the
kernels are compiled for each): small tiles keep the LDS subgrids small and the
blocks many, large ones cut the subgrid halo (a 16 x 16 tile's subgrid at P = 5 is
1.56x its area, a 32 x 32 one's 1.27x) and with it the gridding scratch.
``tune_work`` times interp + gridding on the caller's own trajectory for every
candidate pair and keeps the fastest; results are cached in a JSON table keyed by
(ndim, kernel points, grid size, samples/tile bucket, batch bucket, tile choice)
so later plans of the same class skip the search.
"""

from __future__ import annotations

import json
import math
import os
import threading
from pathlib import Path
from typing import Iterable, Optional

import torch

from . import _lib, ops
from .plan import SamplePlan, kernel_points

CANDIDATES = (256, 512, 1024, 2048, 4096, 8192)
TILES = {1: (256,), 2: (8, 16, 32), 3: (8,)}  # tile edges the kernels are compiled for
_CACHE_PATH = Path(os.environ.get("MR_RECON_AMD_TUNE_CACHE",
                                  Path.home() / ".cache" / "mr_recon_amd" / "tune.json"))
_lock = threading.Lock()
_table: Optional[dict] = None


def _load() -> dict:
    global _table
    if _table is None:
        try:
            _table = json.loads(_CACHE_PATH.read_text())
        except (OSError, ValueError):
            _table = {}
    return _table


def _save() -> None:
    try:
        _CACHE_PATH.parent.mkdir(parents=True, exist_ok=True)
        _CACHE_PATH.write_text(json.dumps(_table, indent=1, sort_keys=True))
    except OSError:
        pass  # read-only home: keep the in-memory table


def tune_key(ndim: int, width: float, grid_size, npts: int, nbatch: int) -> str:
    ntiles = math.prod(-(-g // {1: 256, 2: 16, 3: 8}[ndim]) for g in grid_size)
    dens = max(1, npts // max(1, ntiles))
    return (f"d{ndim}-p{kernel_points(width)}-g{'x'.join(map(str, grid_size))}-"
            f"s{1 << int(math.log2(dens))}-b{1 << int(math.log2(max(1, nbatch)))}")


def _time(fn, reps: int = 3) -> float:
    fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune_work(coords: torch.Tensor, grid_size, width: float, beta: float, nbatch: int, *, scale=None, shift=None,
              wrap=None, tiles: Optional[Iterable[int]] = None, samples: Optional[Iterable[int]] = None,
              tune_tile: bool = True, use_cache: bool = True, reps: int = 3) -> tuple[int, int]:
    """Fastest (tile edge, samples_per_item) for interp + gridding of ``nbatch`` grids on this
    trajectory. ``tiles`` / ``samples`` restrict the candidates; without ``tune_tile`` the
    tile is the default (or the one given)."""
    D = len(grid_size)
    M = int(math.prod(coords.shape[:-1]))
    L = _lib.lib()
    tiles = tuple(tiles) if tiles is not None else (TILES[D] if tune_tile else (int(L.mra_default_tile(D)),))
    samples = tuple(samples) if samples is not None else CANDIDATES
    key = tune_key(D, width, grid_size, M, nbatch) + "-t" + "_".join(map(str, tiles)) + \
        "-s" + "_".join(map(str, samples))
    with _lock:
        table = _load()
        if use_cache and key in table:
            return int(table[key]["tile"]), int(table[key]["samples_per_item"])
    dev = coords.device
    grid = torch.randn((nbatch, *grid_size), dtype=torch.complex64, device=dev)
    ksp = torch.randn((nbatch, M), dtype=torch.complex64, device=dev)
    out = torch.empty_like(grid)
    times = {}
    for T in tiles:
        for S in samples:
            plan = SamplePlan(coords, grid_size, width, beta, scale=scale, shift=shift, wrap=wrap, tile=T,
                              samples_per_item=S)
            scratch = torch.empty(ops.gridding_scratch_elems(plan, nbatch), dtype=torch.complex64, device=dev)
            t_i = _time(lambda: ops.interp(grid, plan, out=ksp), reps)
            t_g = _time(lambda: ops.gridding(ksp, plan, out, scratch=scratch), reps)
            times[(T, S)] = t_i + t_g
            del scratch, plan
    best = min(times, key=times.get)
    with _lock:
        table = _load()
        table[key] = {"tile": int(best[0]), "samples_per_item": int(best[1]),
                      "ms": {f"{t}/{s}": round(v, 4) for (t, s), v in times.items()}}
        _save()
    return int(best[0]), int(best[1])


def tune_samples_per_item(coords: torch.Tensor, grid_size, width: float, beta: float, nbatch: int, *,
                          scale=None, shift=None, wrap=None, candidates: Iterable[int] = CANDIDATES,
                          use_cache: bool = True, reps: int = 3) -> int:
    """Fastest ``samples_per_item`` for interp + gridding of ``nbatch`` grids on this
    trajectory, at the default tile edge."""
    return tune_work(coords, grid_size, width, beta, nbatch, scale=scale, shift=shift, wrap=wrap,
                     samples=candidates, tune_tile=False, use_cache=use_cache, reps=reps)[1]


def tuned_table() -> dict:
    with _lock:
        return dict(_load())
