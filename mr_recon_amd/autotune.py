"""Autotuner for the work decomposition of the KB interp / gridding kernels.

The tile edge T is fixed per dimensionality at compile time (3D 8, 2D 16, 1D
256); what varies with the trajectory is how many sorted samples a work item
takes (``samples_per_item``): small items spread hot tiles (radial centres)
over more blocks, large items cut the per-item subgrid loads and the gridding
scratch. ``tune_samples_per_item`` times interp + gridding on the caller's own
trajectory for each candidate and keeps the fastest; results are cached in a
JSON table keyed by (ndim, kernel points, grid size, samples/tile bucket,
batch bucket) so later plans of the same class skip the search.
"""

from __future__ import annotations

import json
import math
import os
import threading
from pathlib import Path
from typing import Iterable, Optional

import torch

from . import ops
from .plan import SamplePlan, kernel_points

CANDIDATES = (256, 512, 1024, 2048, 4096, 8192)
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


def tune_samples_per_item(coords: torch.Tensor, grid_size, width: float, beta: float, nbatch: int, *,
                          scale=None, shift=None, wrap=None, candidates: Iterable[int] = CANDIDATES,
                          use_cache: bool = True, reps: int = 3) -> int:
    """Fastest ``samples_per_item`` for interp + gridding of ``nbatch`` grids on this trajectory."""
    D = len(grid_size)
    M = int(math.prod(coords.shape[:-1]))
    key = tune_key(D, width, grid_size, M, nbatch)
    with _lock:
        table = _load()
        if use_cache and key in table:
            return int(table[key]["samples_per_item"])
    dev = coords.device
    grid = torch.randn((nbatch, *grid_size), dtype=torch.complex64, device=dev)
    ksp = torch.randn((nbatch, M), dtype=torch.complex64, device=dev)
    out = torch.empty_like(grid)
    times = {}
    for S in candidates:
        plan = SamplePlan(coords, grid_size, width, beta, scale=scale, shift=shift, wrap=wrap, samples_per_item=S)
        scratch = torch.empty(ops.gridding_scratch_elems(plan, nbatch), dtype=torch.complex64, device=dev)
        t_i = _time(lambda: ops.interp(grid, plan, out=ksp), reps)
        t_g = _time(lambda: ops.gridding(ksp, plan, out, scratch=scratch), reps)
        times[S] = t_i + t_g
        del scratch, plan
    best = min(times, key=times.get)
    with _lock:
        table = _load()
        table[key] = {"samples_per_item": int(best), "ms": {str(k): round(v, 4) for k, v in times.items()}}
        _save()
    return int(best)


def tuned_table() -> dict:
    with _lock:
        return dict(_load())
