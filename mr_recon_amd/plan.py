"""Sample plans: tile binning of one trajectory against one oversampled grid.

A plan is built once per (trajectory, grid geometry, kernel width) and reused by
every interp / gridding / gradient call on that trajectory, so the CG loop pays
for binning once. Building it runs a stable LSD radix sort of the samples by
(tile, corner position in the tile) -- two or three counting-sort passes of a
per-chunk histogram, a scan and a stable scatter -- and two small passes for the
work items; the only host synchronisation is the work-item count.

Coordinate handling is bit-exact with the reference:
  * ``triton_nufft._scale_trj`` (fourier.py:420-430): t = fl32(fl32(k * fl32(G/N)) + G//2)
  * ``prep_ungrid_shapes`` / ``prep_grid_shapes`` (ungrid.py:864-865, grid.py:516-517):
    t = torch.remainder(t, G)
  * ``get_neighborhood`` (kernels.py:531-536): lower = ceil(t - w/2)
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr

MAX_DIGIT = 8192  # largest radix of one sort pass (mra_bin_build)


def binning_chunk(npts: int, ntiles: int, tile: int, ndim: int) -> int:
    """Samples per radix-sort chunk. One wave walks a chunk in order, so the chunks are the
    parallelism of each scatter pass (4096-sample chunks left a 131 k-sample trajectory 32
    waves); each pass's histogram holds digits x chunks counts, written and scanned whole.
    Aim for about 1024 chunks of 128..4096 samples, with at most 2^24 counts per histogram
    (measured: a 5.12 M-sample 3D plan takes 0.84 ms at 4096, 1.4 ms at 1024, 7 ms at 128)."""
    digits = max(tile ** ndim, min(ntiles + 1, MAX_DIGIT))
    c = -(-npts // 1024)
    c = min(max(-(-c // 64) * 64, 128), 4096)
    while c < 4096 and digits * -(-npts // c) > 1 << 24:
        c *= 2
    return min(c, 4096)


def default_samples_per_item(npts: int) -> int:
    """Work-item size: about a thousand items per trajectory, 256..8192 samples each."""
    s = -(-npts // 1024)
    s = -(-s // 256) * 256
    return int(min(max(s, 256), 8192))


def kernel_points(width: float) -> int:
    """Grid points per dimension a KB patch of width ``width`` can touch."""
    return int(math.floor(float(width))) + 1


@dataclass(frozen=True)
class GeomKey:
    nbatch: int
    ksp_stride: int
    out_scale: float
    grid_stride: int
    dcf_plan_order: bool = False
    ksp_plan_order: bool = False


class SamplePlan:
    """Tile-sorted samples of one trajectory on one grid.

    Parameters
    ----------
    coords : (..., D) float32 device tensor, reference units (see ``scale``)
    grid_size : tuple of D ints, the (oversampled) grid the samples live on
    width : kernel width in grid points (same in every dim)
    beta : Kaiser-Bessel shape parameter (or one per dimension)
    scale, shift : per-dim fp32 scale and shift applied before wrapping (the
        NUFFT's ``_scale_trj``); ``None`` means coordinates are already in grid
        units (the plain ``ungrid``/``grid`` surface)
    wrap : per-dim integer added to a grid index before it wraps mod G. The
        NUFFT uses ``-(G//2)`` so kernels address the unshifted FFT output.
    tile : tile edge in grid points (default: the edge selected for this thread, see
        ``_lib.tile_scope``; kernels are compiled for 2D edges 8, 16, 32, 1D 256, 3D 8)
    samples_per_item : max samples per work item (default: about 1k items)
    """

    def __init__(self, coords: torch.Tensor, grid_size, width: float, beta: float, *, scale=None, shift=None,
                 wrap=None, tile: int | None = None, samples_per_item: int | None = None):
        _lib.require_device(coords)
        L = lib()
        D = len(grid_size)
        if coords.shape[-1] != D:
            raise ValueError(f"coords last dim {coords.shape[-1]} != grid ndim {D}")
        if D < 1 or D > 3:
            raise ValueError("the HIP NUFFT supports 1D, 2D and 3D grids")
        if coords.dtype != torch.float32:
            raise TypeError(f"coordinates must be float32, got {coords.dtype}")
        self.device = coords.device
        self.ndim = D
        self.grid_size = tuple(int(g) for g in grid_size)
        self.width = float(width)
        self.P = kernel_points(width)
        if not 2 <= self.P <= 9:
            raise ValueError(f"kernel width {width} unsupported: floor(width)+1 must be in 2..9")
        if isinstance(beta, (tuple, list)):  # one KB shape per dimension (sigpy's `param` tuple)
            if len(beta) != D:
                raise ValueError(f"{len(beta)} betas for a {D}-dimensional grid")
            self.betas = tuple(float(b) for b in beta)
        else:
            self.betas = (float(beta),) * D
        self.beta = self.betas[0]
        self.coords_shape = tuple(coords.shape[:-1])
        self.npts = int(math.prod(self.coords_shape))
        if self.npts <= 0:
            raise ValueError("trajectory has no samples")
        self.tile = int(tile or L.mra_tile(D))
        with _lib.tile_scope(D, self.tile):  # raises for an edge without kernels
            pass
        self.ntile = [(g + self.tile - 1) // self.tile for g in self.grid_size]
        self.ntiles = int(math.prod(self.ntile))
        self.S = int(samples_per_item or default_samples_per_item(self.npts))
        if self.S < 1:
            raise ValueError("samples_per_item must be positive")
        self.wrap = tuple(int(w) for w in (wrap if wrap is not None else (0,) * D))

        flat = coords.reshape(-1, D).contiguous()
        dev = self.device
        chunk = binning_chunk(self.npts, self.ntiles, self.tile, D)
        wwords = int(L.mra_bin_workspace(D, self.npts, self.ntiles, self.tile, chunk))
        if wwords < 0:
            check(1)
        work = torch.empty(wwords, dtype=torch.int32, device=dev)
        self.tsort = torch.empty(self.npts * D, dtype=torch.float32, device=dev)
        self.perm = torch.empty(self.npts, dtype=torch.int32, device=dev)
        nkeys = self.ntiles + 1
        self.tile_start = torch.empty(nkeys + 1, dtype=torch.int32, device=dev)
        apply = scale is not None
        sc = _lib.floats(scale if apply else [1.0] * D)
        sh = _lib.floats(shift if shift is not None else [0.0] * D)
        st = stream_ptr(dev)
        check(L.mra_bin_build(D, ptr(flat), self.npts, _lib.ints(self.grid_size), sc, sh, int(apply),
                              self.width, self.tile, _lib.ints(self.wrap), chunk, ptr(work), ptr(self.tsort),
                              ptr(self.perm), ptr(self.tile_start), st))
        self.max_items = nkeys + (self.npts + self.S - 1) // self.S
        self.items = torch.empty((self.max_items, 4), dtype=torch.int32, device=dev)
        nitems_ws = torch.empty(nkeys + 1, dtype=torch.int32, device=dev)
        item_off = torch.empty(nkeys + 1, dtype=torch.int32, device=dev)
        scan_ws2 = torch.empty(L.mra_scan_workspace(nkeys + 1), dtype=torch.int32, device=dev)
        check(L.mra_bin_items(ptr(self.tile_start), self.ntiles, self.S, 0, ptr(nitems_ws), ptr(item_off),
                              ptr(scan_ws2), ptr(self.items), self.max_items, st))
        # item_off[t] is the first work item of tile t (t = ntiles: non-finite samples)
        self.tile_items = item_off
        # KB weights (D x P per sorted sample) and subgrid offsets, once per trajectory
        self.wsort = torch.empty(self.npts * D * self.P, dtype=torch.float32, device=dev)
        self.lbase = torch.empty(self.npts, dtype=torch.int32, device=dev)
        with _lib.tile_scope(D, self.tile):
            check(L.mra_plan_weights_dims(D, self.P, ptr(self.tsort), self.npts, _lib.ints(self.grid_size),
                                          _lib.ints(self.wrap), self.width, _lib.floats(self.betas),
                                          ptr(self.wsort), ptr(self.lbase), st))
            self.subgrid = int(L.mra_subgrid_points(D, self.P))
            self.max_coil_chunk = int(L.mra_max_coil_chunk(D, self.P))
        # The only host synchronisation of a plan: the number of work items, which
        # sizes the launches and the gridding scratch.
        self.nitems = int(item_off[nkeys].item())
        # The workspaces go back to the caching allocator here; PyTorch only
        # hands them out again to work ordered after these kernels on this stream.
        del work, nitems_ws, scan_ws2, flat
        self._geoms: dict[GeomKey, torch.Tensor] = {}

    def coil_chunk(self, nbatch: int) -> int:
        """Coils per block: as many subgrids as the per-block LDS budget holds."""
        return max(1, min(nbatch, self.max_coil_chunk))

    def geom(self, nbatch: int, ksp_stride: int | None = None, out_scale: float = 1.0,
             grid_stride: int | None = None, dcf_plan_order: bool = False,
             ksp_plan_order: bool = False) -> torch.Tensor:
        """Device Geom table (one entry) for a call on ``nbatch`` grids; ``dcf_plan_order``:
        the gridding dcf comes in sorted (plan) order (``dcf[perm]``); ``ksp_plan_order``:
        k-space (interp output, gridding input) and dcf are all in sorted order."""
        gstride = int(grid_stride if grid_stride is not None else math.prod(self.grid_size))
        kstride = int(ksp_stride if ksp_stride is not None else self.npts)
        key = GeomKey(int(nbatch), kstride, float(out_scale), gstride, bool(dcf_plan_order), bool(ksp_plan_order))
        g = self._geoms.get(key)
        if g is None:
            g = make_geoms([dict(grid_off=0, grid_stride=gstride, ksp_off=0, ksp_stride=kstride, samp_off=0,
                                 nbatch=int(nbatch), npts=self.npts, grid_size=self.grid_size, tile=self.tile,
                                 wrap=self.wrap, beta=self.beta, width=self.width, out_scale=out_scale,
                                 item_base=0, scratch_off=0, tile_items_off=0,
                                 dcf_plan_order=bool(dcf_plan_order), ksp_plan_order=bool(ksp_plan_order))],
                           self.device)
            self._geoms[key] = g
        return g


    def geom_host(self, nbatch: int, out_scale: float = 1.0):
        """Host Geom record of a one-problem call on ``nbatch`` grids (what ``geom`` puts in
        device memory), for the no-allocation gridding entry point."""
        return pack_geoms([dict(grid_off=0, grid_stride=math.prod(self.grid_size), ksp_off=0, ksp_stride=self.npts,
                                samp_off=0, nbatch=int(nbatch), npts=self.npts, grid_size=self.grid_size,
                                tile=self.tile, wrap=self.wrap, beta=self.beta, width=self.width,
                                out_scale=out_scale)])


def pack_geoms(entries: list[dict]):
    """Pack Geom records (see common.h) into a host ctypes byte buffer."""
    L = lib()
    gsz = L.mra_geom_size()
    buf = (C.c_ubyte * (gsz * len(entries)))()
    for i, e in enumerate(entries):
        D = len(e["grid_size"])
        with _lib.tile_scope(D, e["tile"]):
            check(L.mra_geom_fill(C.addressof(buf) + i * gsz, e["grid_off"], e["grid_stride"], e["ksp_off"],
                                  e["ksp_stride"], e["samp_off"], e["nbatch"], e["npts"], D,
                                  _lib.ints(e["grid_size"]), e["tile"], _lib.ints(e["wrap"]), e["beta"], e["width"],
                                  e["out_scale"], e.get("item_base", 0), e.get("scratch_off", 0),
                                  e.get("tile_items_off", 0)))
        if e.get("dcf_plan_order"):
            check(L.mra_geom_set_dcf_plan_order(C.addressof(buf) + i * gsz, 1))
        if e.get("ksp_plan_order"):
            check(L.mra_geom_set_ksp_plan_order(C.addressof(buf) + i * gsz, 1))
    return buf


def make_geoms(entries: list[dict], device) -> torch.Tensor:
    """Pack Geom records (see common.h) into a device byte tensor."""
    host = torch.frombuffer(bytearray(bytes(pack_geoms(entries))), dtype=torch.uint8)
    return host.to(device)
