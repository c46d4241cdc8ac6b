"""Many uneven CG-SENSE problems in one batched, graph-captured pipeline.

The reference reconstructs problems one at a time: a Python loop of
``sense_linop`` + ``CG_SENSE_recon`` (linops.py:1041-1308, recons.py:77-156), each
CG iteration paying Python, launch and host-sync costs (algs.py:845, 859) that
dominate for 2D problems. ``BatchedSense`` runs every stage of SENSE for all
problems at once:

  pad/apodize/expand + FFT   one fused kernel chain per image-size *group* (the LDS
                             FFT length is a compile-time parameter), writing into
                             that group's part of one shared grid buffer
  KB interp                  ONE launch over the concatenated work items of every
                             problem of every group; each item names its problem,
                             whose Geom record gives its grid size, grid and sample
                             offsets and coil count
  KB gridding, phase 1       ONE launch over the same items into one scratch buffer
  KB gridding, reduce        one launch per group (its tile count sizes the grid)
  iFFT + crop/reduce         one fused kernel chain per group, coil ranges per image
  CG                         segmented FP64 reductions, one step size per problem

Problems may differ in image size, coil count, sample count and trajectory.
``capture`` records the whole fixed-iteration CG of every problem into one HIP
graph, so a batch of reconstructions is one graph launch.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Sequence

import torch

from . import _lib, ops
from ._lib import check, lib, ptr, stream_ptr
from .dtypes import complex_dtype
from .fourier import _fftn, _ifftn, hip_nufft
from .plan import make_geoms


@dataclass
class Problem:
    trj: torch.Tensor            # (*trj, 2) device, reference units
    mps: torch.Tensor            # (C, N, N) device
    dcf: Optional[torch.Tensor] = None  # (*trj,) device


class _Group:
    """Problems of one image size: their plans, maps and transform constants. Offsets into
    the batch-wide buffers (grid, k-space, samples, items, scratch) are set by BatchedSense."""

    def __init__(self, idx: list[int], problems: Sequence[Problem], N: int, oversamp: float, width: float,
                 samples_per_item: Optional[int], tile: Optional[int] = None):
        self.idx = idx
        self.N = N
        self.im = (N, N)
        dev = problems[idx[0]].trj.device
        self.device = dev
        cmax = max(int(problems[i].mps.shape[0]) for i in idx)
        self.nufft = hip_nufft(self.im, oversamp=oversamp, width=width, samples_per_item=samples_per_item,
                               tune_batch=cmax, tile=tile)
        self.G = self.nufft.os_shape()
        self.npts_grid = math.prod(self.G)
        self.apod = self.nufft.apod_table(dev)
        self.s_fwd = 1.0 / math.sqrt(N * N)
        self.s_interp = 1.0 / width ** 2
        self.plans = [self.nufft.plan(problems[i].trj.reshape(-1, 2)) for i in idx]
        self.P = self.plans[0].P
        self.tile = self.plans[0].tile
        self.ntiles = self.plans[0].ntiles
        self.subgrid = self.plans[0].subgrid
        self.C = [int(problems[i].mps.shape[0]) for i in idx]
        self.M = [p.npts for p in self.plans]
        self.nimg = len(idx)
        self.ncoil = sum(self.C)
        self.maxC = max(self.C)
        self.mps = torch.cat([problems[i].mps.to(complex_dtype).reshape(-1, N, N) for i in idx]).contiguous()
        self.coil_off = torch.tensor([0] + list(_cumsum(self.C)), dtype=torch.int32, device=dev)
        self.img_index = torch.repeat_interleave(torch.arange(self.nimg, dtype=torch.int32),
                                                 torch.tensor(self.C)).to(dev)
        # k-space offsets of this group's problems relative to the group's block
        self.ksp_off = [0] + list(_cumsum([c * m for c, m in zip(self.C, self.M)]))
        self.ksp_len = self.ksp_off[-1]
        self.fused = ops.fused_supported(self.G)  # pad+FFT / iFFT+crop in the transform passes
        # set by BatchedSense
        self.prob0 = 0       # global index of this group's first problem
        self.grid_base = 0   # complex64 offset of this group's grids in the shared buffer
        self.ksp_base = 0    # complex64 offset of this group's k-space block

    def grid_view(self, buf: torch.Tensor) -> torch.Tensor:
        return buf[self.grid_base:self.grid_base + self.ncoil * self.npts_grid].view(self.ncoil, *self.G)


def _cumsum(v):
    s = 0
    for x in v:
        s += x
        yield s


class BatchedSense:
    """Stacked SENSE operators of many 2D problems (grouped by image size for the transforms,
    one launch across all groups for interpolation and gridding).

    k-space of the whole batch is one vector (``ksp_len`` complex64): the groups' blocks in
    group order, within a block problem after problem, each (C_i, M_i). Images are one
    tensor (nimg, N, N) per group."""

    def __init__(self, problems: Sequence[Problem], oversamp: float = 1.25, width: float = 4,
                 samples_per_item: Optional[int] = None, tile=None):
        """``tile``: the 2D tile edge of every problem's plan (one launch serves them all):
        None for the default, an int, or "auto" -- each compiled edge is tried on the
        batch's own normal operator and the fastest kept."""
        self.tile_times = None  # ms per normal-operator application for each tried edge
        if tile == "auto":
            tile, self.tile_times = self._tune_tile(problems, oversamp, width, samples_per_item)
        self.problems = list(problems)
        sizes = {}
        for i, pr in enumerate(self.problems):
            _lib.require_device(pr.trj, pr.mps)
            if pr.mps.ndim != 3 or pr.mps.shape[1] != pr.mps.shape[2]:
                raise ValueError("batched problems are 2D with square images: mps (C, N, N)")
            sizes.setdefault(int(pr.mps.shape[1]), []).append(i)
        self.groups = [_Group(idx, self.problems, N, oversamp, width, samples_per_item, tile)
                       for N, idx in sorted(sizes.items())]
        g0 = self.groups[0]
        if any(g.P != g0.P or g.tile != g0.tile for g in self.groups):
            raise ValueError("batched problems must share the kernel width and tile")
        self.P = g0.P
        self.tile = g0.tile
        self.device = g0.device
        self._assemble()
        self._graph = None

    @staticmethod
    def _tune_tile(problems, oversamp, width, samples_per_item, reps: int = 3):
        from .autotune import TILES

        times = {}
        for T in TILES[2]:
            B = BatchedSense(problems, oversamp, width, samples_per_item, tile=T)
            xs = [torch.randn(pr.mps.shape[1:], dtype=complex_dtype, device=pr.mps.device) for pr in problems]
            Xs = B.stack_images(xs)
            B.normal(Xs)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(reps):
                B.normal(Xs)
            b.record()
            torch.cuda.synchronize()
            times[T] = a.elapsed_time(b) / reps
            del B, Xs
        return min(times, key=times.get), times

    def _assemble(self):
        """One Geom table, item list, sample arrays, scratch and grid buffer for the batch."""
        L = lib()
        dev = self.device
        entries, items, tsort, perm, wsort, lbase, tiles, dcfs = [], [], [], [], [], [], [], []
        prob = grid_base = ksp_base = samp = item_base = scratch = tile_off = 0
        maxC = 0
        for g in self.groups:
            g.prob0, g.grid_base, g.ksp_base = prob, grid_base, ksp_base
            coil0 = 0
            for j, (i, p) in enumerate(zip(g.idx, g.plans)):
                C, M = g.C[j], g.M[j]
                entries.append(dict(grid_off=grid_base + coil0 * g.npts_grid, grid_stride=g.npts_grid,
                                    ksp_off=ksp_base + g.ksp_off[j], ksp_stride=M, samp_off=samp, nbatch=C, npts=M,
                                    grid_size=g.G, tile=p.tile, wrap=tuple(-(x // 2) for x in g.G),
                                    beta=g.nufft.beta, width=g.nufft.width, out_scale=g.s_interp,
                                    item_base=item_base, scratch_off=scratch, tile_items_off=tile_off))
                it = p.items[: p.nitems].clone()
                it[:, 0] = prob
                items.append(it)
                tiles.append(p.tile_items[: p.ntiles + 2] + item_base)
                tsort.append(p.tsort)
                perm.append(p.perm)
                wsort.append(p.wsort)
                lbase.append(p.lbase)
                d = self.problems[i].dcf
                dcfs.append(d)
                coil0 += C
                prob += 1
                samp += M
                item_base += p.nitems
                scratch += p.nitems * C * g.subgrid
                tile_off += p.ntiles + 2
                maxC = max(maxC, C)
            grid_base += g.ncoil * g.npts_grid
            ksp_base += g.ksp_len
        self.nprob = prob
        self.nitems = item_base
        self.ksp_len = ksp_base
        self.maxC = maxC
        self.geoms = make_geoms(entries, dev)
        # the normal operator keeps its intermediate k-space in plan (sorted) order, so the
        # gather stores it and gridding loads it contiguously
        self.geoms_po = make_geoms([dict(e, ksp_plan_order=True) for e in entries], dev)
        self.items = torch.cat(items).contiguous()
        self.tile_items = torch.cat(tiles).to(torch.int32).contiguous()
        self.tsort = torch.cat(tsort)
        self.perm = torch.cat(perm)
        self.wsort = torch.cat(wsort)
        self.lbase = torch.cat(lbase)
        self.dcf = self.dcf_po = None
        if any(d is not None for d in dcfs):
            ms = [m for g in self.groups for m in g.M]
            full = [(d if d is not None else torch.ones(m, device=dev)).reshape(-1).float() for d, m in zip(dcfs, ms)]
            self.dcf = torch.cat(full).contiguous()
            self.dcf_po = torch.cat([d[pm.long()] for d, pm in zip(full, perm)]).contiguous()
        with _lib.tile_scope(2, self.tile):
            self.cc = max(1, min(maxC, L.mra_max_coil_chunk(2, self.P)))
        self.nchunks = -(-maxC // self.cc)
        self.scratch = torch.empty(max(1, scratch), dtype=complex_dtype, device=dev)
        self.gridbuf = torch.empty(max(1, grid_base), dtype=complex_dtype, device=dev)

    # -- operators: images Xs (one (nimg, N, N) tensor per group) <-> k-space Y (ksp_len) --
    def forward(self, Xs: Sequence[torch.Tensor], plan_order: bool = False) -> torch.Tensor:
        """``plan_order``: k-space in each problem's sorted (plan) order instead of sample order."""
        st = stream_ptr(self.device)
        for g, X in zip(self.groups, Xs):
            view = g.grid_view(self.gridbuf)
            if g.fused:
                ops.pad_fft(X, g.im, g.G, smaps=g.mps, apod=g.apod, scale=g.s_fwd, img_index=g.img_index,
                            nq=g.ncoil, out=view)
            else:
                check(lib().mra_pad_expand_batched(2, _lib.ints(g.im), _lib.ints(g.G), None, ptr(X), g.nimg,
                                                   ptr(g.mps), g.ncoil, ptr(g.img_index), ptr(g.apod), g.s_fwd,
                                                   ptr(view), st))
                view.copy_(_fftn(view, 2))
        Y = torch.empty(self.ksp_len, dtype=complex_dtype, device=self.device)
        geoms = self.geoms_po if plan_order else self.geoms
        with _lib.tile_scope(2, self.tile):
            check(lib().mra_interp(2, self.P, ptr(self.gridbuf), ptr(Y), ptr(self.wsort), ptr(self.lbase),
                                   ptr(self.perm), ptr(geoms), ptr(self.items), self.nitems, self.cc, self.nchunks,
                                   st))
        return Y

    def adjoint(self, Y: torch.Tensor, scale: float = 1.0, plan_order: bool = False) -> list[torch.Tensor]:
        st = stream_ptr(self.device)
        if Y.numel() != self.ksp_len or Y.dtype != complex_dtype or not Y.is_contiguous():
            raise ValueError(f"adjoint: k-space must be contiguous complex64 with {self.ksp_len} elements")
        geoms, dcf = (self.geoms_po, self.dcf_po) if plan_order else (self.geoms, self.dcf)
        with _lib.tile_scope(2, self.tile):
            check(lib().mra_grid_partial(2, self.P, ptr(Y), ptr(dcf), ptr(self.wsort), ptr(self.lbase),
                                         ptr(self.perm), ptr(geoms), ptr(self.items), self.nitems, self.maxC,
                                         ptr(self.scratch), st))
            for g in self.groups:
                check(lib().mra_grid_reduce(2, self.P, ptr(self.gridbuf), ptr(self.scratch), ptr(self.geoms),
                                            g.prob0, g.nimg, ptr(self.tile_items), g.ntiles, g.maxC, 0, st))
        Xs = []
        for g in self.groups:
            view = g.grid_view(self.gridbuf)
            X = torch.empty((g.nimg, *g.im), dtype=complex_dtype, device=self.device)
            if g.fused:
                ops.ifft_crop(view, g.im, g.G, g.nimg, smaps=g.mps, apod=g.apod, scale=g.s_fwd * scale, out=X,
                              coil_off=g.coil_off)
            else:
                gi = _ifftn(view, 2)
                check(lib().mra_crop_reduce_batched(2, _lib.ints(g.im), _lib.ints(g.G), None, ptr(gi), g.nimg,
                                                    ptr(g.mps), ptr(g.coil_off), ptr(g.apod), g.s_fwd * scale,
                                                    ptr(X), st))
            Xs.append(X)
        return Xs

    def normal(self, Xs: Sequence[torch.Tensor], scale: float = 1.0) -> list[torch.Tensor]:
        return self.adjoint(self.forward(Xs, plan_order=True), scale, plan_order=True)

    # stacking helpers -------------------------------------------------------------------
    def ksp_slice(self, i: int) -> slice:
        """Where problem i's k-space (C_i x M_i) lies in the batch's k-space vector."""
        for g in self.groups:
            if i in g.idx:
                j = g.idx.index(i)
                return slice(g.ksp_base + g.ksp_off[j], g.ksp_base + g.ksp_off[j + 1])
        raise IndexError(i)

    def stack_ksp(self, ksp: Sequence[torch.Tensor]) -> torch.Tensor:
        return torch.cat([ksp[i].to(complex_dtype).reshape(-1) for g in self.groups for i in g.idx])

    def stack_images(self, xs: Sequence[torch.Tensor]) -> list[torch.Tensor]:
        return [torch.stack([xs[i].to(complex_dtype) for i in g.idx]) for g in self.groups]

    def unstack_images(self, Xs: Sequence[torch.Tensor]) -> list[torch.Tensor]:
        out = [None] * len(self.problems)
        for g, X in zip(self.groups, Xs):
            for j, i in enumerate(g.idx):
                out[i] = X[j]
        return out

    # reconstruction ---------------------------------------------------------------------
    def _cg(self, Y, num_iters, lamda_l2, max_eigen, states):
        """CG-SENSE (recons.py:77-156, fixed iteration count) on every problem at once; writes
        into states. Each operator application covers all groups."""
        inv = 1.0 / max_eigen
        st = stream_ptr(self.device)
        ahb = self.adjoint(Y, max_eigen ** -0.5)
        for (x, *_), a in zip(states, ahb):
            x.copy_(a)
        ax = self.normal([s[0] for s in states], inv)
        for g, (x, r, p, part, sc), a, axg in zip(self.groups, states, ahb, ax):
            check(lib().mra_cg_init_seg(ptr(a), ptr(axg), ptr(x), ptr(r), ptr(p), g.N * g.N, g.nimg,
                                        float(lamda_l2), ptr(part), ptr(sc), st))
        for _ in range(num_iters):
            ap = self.normal([s[2] for s in states], inv)
            for g, (x, r, p, part, sc), apg in zip(self.groups, states, ap):
                check(lib().mra_cg_step_seg(ptr(x), ptr(r), ptr(p), ptr(apg), g.N * g.N, g.nimg, float(lamda_l2),
                                            ptr(part), ptr(sc), st))

    def _states(self):
        nb = lib().mra_cg_seg_blocks()
        st = []
        for g in self.groups:
            shp = (g.nimg, *g.im)
            mk = lambda: torch.empty(shp, dtype=complex_dtype, device=g.device)  # noqa: E731
            st.append((mk(), mk(), mk(), torch.empty(g.nimg * nb, dtype=torch.float64, device=g.device),
                       torch.empty(g.nimg * 8, dtype=torch.float64, device=g.device)))
        return st

    def cg_sense(self, ksp: Sequence[torch.Tensor], num_iters: int = 10, lamda_l2: float = 0.0,
                 max_eigen: float = 1.0) -> list[torch.Tensor]:
        """Eager batched CG-SENSE; returns one image per problem."""
        Y = self.stack_ksp(ksp)
        states = self._states()
        self._cg(Y, num_iters, lamda_l2, max_eigen, states)
        return self.unstack_images([s[0] for s in states])

    def capture(self, num_iters: int = 10, lamda_l2: float = 0.0, max_eigen: float = 1.0):
        """Record the whole batched CG-SENSE in one HIP graph. Fill ``self.inputs`` (the
        batch's stacked k-space) and call ``replay()``; results in ``self.outputs``."""
        self.inputs = torch.zeros(self.ksp_len, dtype=complex_dtype, device=self.device)
        self._cap_states = self._states()
        self._cap_args = (num_iters, lamda_l2, max_eigen)
        # warm up outside capture: FFT plans, allocator pools
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self._cg(self.inputs, num_iters, lamda_l2, max_eigen, self._cap_states)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self._cg(self.inputs, num_iters, lamda_l2, max_eigen, self._cap_states)
        self._graph = graph
        self.outputs = [st[0] for st in self._cap_states]
        return graph

    def replay(self, ksp: Optional[Sequence[torch.Tensor]] = None) -> list[torch.Tensor]:
        if self._graph is None:
            raise RuntimeError("call capture() first")
        if ksp is not None:
            self.inputs.copy_(self.stack_ksp(ksp))
        self._graph.replay()
        return self.unstack_images(self.outputs)


__all__ = ["Problem", "BatchedSense"]
