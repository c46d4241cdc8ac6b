"""Many uneven CG-SENSE problems in one batched, graph-captured pipeline.

The reference reconstructs problems one at a time: a Python loop of
``sense_linop`` + ``CG_SENSE_recon`` (linops.py:1041-1308, recons.py:77-156), each
CG iteration paying Python, launch and host-sync costs (algs.py:845, 859) that
dominate for 2D problems. ``BatchedSense`` stacks problems that share an image
size into one *group* and runs every stage of SENSE for all of them in one launch:

  pad/apodize/expand  one kernel over all coils of the group (coil -> image table)
  FFT                 one batched rocFFT over all coils of the group
  KB interp           one launch over the concatenated work items; each item
                      names its problem, whose Geom record gives its grid,
                      samples and coil count
  KB gridding         one partial launch + one reduce launch for the group
  crop/reduce         one kernel, coil ranges per image
  CG                  segmented FP64 reductions, one step size per problem

Problems may differ in coil count, sample count and trajectory; groups differ in
image size. ``capture`` records the whole fixed-iteration CG of every group
into one HIP graph, so a batch of reconstructions is one graph launch.
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
    """Problems of one image size, stacked."""

    def __init__(self, idx: list[int], problems: Sequence[Problem], N: int, oversamp: float, width: float,
                 samples_per_item: Optional[int], plans: Optional[list] = None):
        L = lib()
        self.idx = idx
        self.N = N
        self.im = (N, N)
        self.d = 2
        dev = problems[idx[0]].trj.device
        self.device = dev
        cmax = max(int(problems[i].mps.shape[0]) for i in idx)
        self.nufft = hip_nufft(self.im, oversamp=oversamp, width=width, samples_per_item=samples_per_item,
                               tune_batch=cmax)
        self.G = self.nufft.os_shape()
        self.apod = self.nufft.apod_table(dev)
        self.s_fwd = 1.0 / math.sqrt(N * N)
        self.s_interp = 1.0 / width ** 2
        self.s_adj = self.s_fwd * self.s_interp
        self.plans = plans or [self.nufft.plan(problems[i].trj.reshape(-1, 2)) for i in idx]
        self.P = self.plans[0].P
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
        self.ksp_off = [0] + list(_cumsum([c * m for c, m in zip(self.C, self.M)]))
        self.samp_off = [0] + list(_cumsum(self.M))
        nit = [p.nitems for p in self.plans]
        self.item_base = [0] + list(_cumsum(nit))
        self.nitems = self.item_base[-1]
        scr = [n * c * self.subgrid for n, c in zip(nit, self.C)]
        self.scratch_off = [0] + list(_cumsum(scr))
        self.tsort = torch.cat([p.tsort for p in self.plans])
        self.perm = torch.cat([p.perm for p in self.plans])
        self.wsort = torch.cat([p.wsort for p in self.plans])
        self.lbase = torch.cat([p.lbase for p in self.plans])
        items = []
        for j, p in enumerate(self.plans):
            it = p.items[: p.nitems].clone()
            it[:, 0] = j
            items.append(it)
        self.items = torch.cat(items).contiguous()
        self.tile_items = torch.cat([p.tile_items[: p.ntiles + 2] + self.item_base[j]
                                     for j, p in enumerate(self.plans)]).to(torch.int32).contiguous()
        dcfs = [problems[i].dcf for i in idx]
        self.dcf = None
        if any(x is not None for x in dcfs):
            self.dcf = torch.cat([(x if x is not None else torch.ones(m, device=dev)).reshape(-1).float()
                                  for x, m in zip(dcfs, self.M)]).contiguous()
        self.cc = max(1, min(self.maxC, L.mra_max_coil_chunk(2, self.P)))
        self.nchunks = -(-self.maxC // self.cc)
        entries = []
        for j in range(self.nimg):
            entries.append(dict(grid_off=int(self.coil_off[j]) * math.prod(self.G), grid_stride=math.prod(self.G),
                                ksp_off=self.ksp_off[j], ksp_stride=self.M[j], samp_off=self.samp_off[j],
                                nbatch=self.C[j], npts=self.M[j], grid_size=self.G, tile=self.plans[j].tile,
                                wrap=tuple(-(g // 2) for g in self.G), beta=self.nufft.beta,
                                width=self.nufft.width, out_scale=self.s_interp, item_base=self.item_base[j],
                                scratch_off=self.scratch_off[j], tile_items_off=j * (self.ntiles + 2)))
        self.geoms = make_geoms(entries, dev)
        self.scratch = torch.empty(max(1, self.scratch_off[-1]), dtype=complex_dtype, device=dev)
        self.ksp_len = self.ksp_off[-1]
        self.fused = ops.fused_supported(self.G)  # pad+FFT / iFFT+crop in the transform passes

    # -- operators on the stacked images X (nimg, N, N) and k-space Y (sum C_i M_i) --------
    def forward(self, X: torch.Tensor) -> torch.Tensor:
        if self.fused:
            g = ops.pad_fft(X, self.im, self.G, smaps=self.mps, apod=self.apod, scale=self.s_fwd,
                            img_index=self.img_index, nq=self.ncoil)
        else:
            g = torch.empty((self.ncoil, *self.G), dtype=complex_dtype, device=self.device)
            check(lib().mra_pad_expand_batched(2, _lib.ints(self.im), _lib.ints(self.G), None, ptr(X), self.nimg,
                                               ptr(self.mps), self.ncoil, ptr(self.img_index), ptr(self.apod),
                                               self.s_fwd, ptr(g), stream_ptr(self.device)))
            g = _fftn(g, 2)
        Y = torch.empty(self.ksp_len, dtype=complex_dtype, device=self.device)
        check(lib().mra_interp(2, self.P, ptr(g), ptr(Y), ptr(self.wsort), ptr(self.lbase), ptr(self.perm),
                               ptr(self.geoms), ptr(self.items), self.nitems, self.cc, self.nchunks,
                               stream_ptr(self.device)))
        return Y

    def adjoint(self, Y: torch.Tensor, scale: float = 1.0) -> torch.Tensor:
        g = torch.empty((self.ncoil, *self.G), dtype=complex_dtype, device=self.device)
        st = stream_ptr(self.device)
        check(lib().mra_grid_partial(2, self.P, ptr(Y), ptr(self.dcf), ptr(self.wsort), ptr(self.lbase),
                                     ptr(self.perm), ptr(self.geoms), ptr(self.items), self.nitems, self.maxC,
                                     ptr(self.scratch), st))
        check(lib().mra_grid_reduce(2, self.P, ptr(g), ptr(self.scratch), ptr(self.geoms), 0, self.nimg,
                                    ptr(self.tile_items), self.ntiles, self.maxC, 0, st))
        X = torch.empty((self.nimg, *self.im), dtype=complex_dtype, device=self.device)
        if self.fused:
            ops.ifft_crop(g, self.im, self.G, self.nimg, smaps=self.mps, apod=self.apod, scale=self.s_fwd * scale,
                          out=X, coil_off=self.coil_off)
            return X
        g = _ifftn(g, 2)
        check(lib().mra_crop_reduce_batched(2, _lib.ints(self.im), _lib.ints(self.G), None, ptr(g), self.nimg,
                                            ptr(self.mps), ptr(self.coil_off), ptr(self.apod), self.s_fwd * scale,
                                            ptr(X), st))
        return X

    def normal(self, X: torch.Tensor, scale: float = 1.0) -> torch.Tensor:
        return self.adjoint(self.forward(X), scale)


def _cumsum(v):
    s = 0
    for x in v:
        s += x
        yield s


class BatchedSense:
    """Stacked SENSE operators of many 2D problems (grouped by image size)."""

    def __init__(self, problems: Sequence[Problem], oversamp: float = 1.25, width: float = 4,
                 samples_per_item: Optional[int] = None):
        self.problems = list(problems)
        sizes = {}
        for i, pr in enumerate(self.problems):
            _lib.require_device(pr.trj, pr.mps)
            if pr.mps.ndim != 3 or pr.mps.shape[1] != pr.mps.shape[2]:
                raise ValueError("batched problems are 2D with square images: mps (C, N, N)")
            sizes.setdefault(int(pr.mps.shape[1]), []).append(i)
        self.groups = [_Group(idx, self.problems, N, oversamp, width, samples_per_item)
                       for N, idx in sorted(sizes.items())]
        self._graph = None

    # stacking helpers -------------------------------------------------------------------
    def stack_ksp(self, ksp: Sequence[torch.Tensor]) -> list[torch.Tensor]:
        return [torch.cat([ksp[i].to(complex_dtype).reshape(-1) for i in g.idx]) for g in self.groups]

    def unstack_images(self, Xs: Sequence[torch.Tensor]) -> list[torch.Tensor]:
        out = [None] * len(self.problems)
        for g, X in zip(self.groups, Xs):
            for j, i in enumerate(g.idx):
                out[i] = X[j]
        return out

    # reconstruction ---------------------------------------------------------------------
    def _cg(self, Ys, num_iters, lamda_l2, max_eigen, states):
        """CG-SENSE (recons.py:77-156, fixed iteration count) on every group; writes into states."""
        inv = 1.0 / max_eigen
        for g, Y, stt in zip(self.groups, Ys, states):
            x, r, p, part, sc = stt
            ahb = g.adjoint(Y, max_eigen ** -0.5)
            x.copy_(ahb)
            seg = g.N * g.N
            ax = g.normal(x, inv)
            check(lib().mra_cg_init_seg(ptr(ahb), ptr(ax), ptr(x), ptr(r), ptr(p), seg, g.nimg, float(lamda_l2),
                                        ptr(part), ptr(sc), stream_ptr(g.device)))
            for _ in range(num_iters):
                ap = g.normal(p, inv)
                check(lib().mra_cg_step_seg(ptr(x), ptr(r), ptr(p), ptr(ap), seg, g.nimg, float(lamda_l2),
                                            ptr(part), ptr(sc), stream_ptr(g.device)))

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
        Ys = self.stack_ksp(ksp)
        states = self._states()
        self._cg(Ys, num_iters, lamda_l2, max_eigen, states)
        return self.unstack_images([s[0] for s in states])

    def capture(self, num_iters: int = 10, lamda_l2: float = 0.0, max_eigen: float = 1.0):
        """Record the whole batched CG-SENSE in one HIP graph. Fill ``self.inputs`` (one
        stacked k-space tensor per group) and call ``replay()``; results in ``self.outputs``."""
        self.inputs = [torch.zeros(g.ksp_len, dtype=complex_dtype, device=g.device) for g in self.groups]
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
            for buf, Y in zip(self.inputs, self.stack_ksp(ksp)):
                buf.copy_(Y)
        self._graph.replay()
        return self.unstack_images(self.outputs)


__all__ = ["Problem", "BatchedSense"]
