"""Linear operators on the NUFFT path: ``linop`` base, ``batching_params``, ``sense_linop``,
the low-rank coil / imperfection model (``imperf_coil_lowrank``), the subspace model
(``subspace_linop``) and the type-3 NUFFT (``type3_nufft``, ``type3_nufft_naive``).

Mirrors the reference (linops.py:22-47, 1041-1308) call for call. With the HIP
NUFFT and a single imperfection segment (the default), every SENSE
application is one fused pipeline over all coils:

  forward : pad+apodize+expand (x * S_c)  ->  FFT  ->  KB interp            (C, M)
  adjoint : KB gridding (* dcf)  ->  iFFT  ->  crop+deapodize+reduce (sum conj(S_c) .)
  normal  : pad+expand  ->  FFT  ->  * PSF  ->  iFFT  ->  crop+reduce    (use_toeplitz)

Coils are processed together rather than one at a time (the reference default
``coil_batch_size=1``): a 32-coil 160^3 grid is 1 GiB, small against HBM; coil
chunks are only formed when the grids and their gridding scratch would exceed
``MAX_GRID_BYTES``. Other
NUFFTs, or spatial/temporal imperfection functions with more than one segment,
take the reference's generic loop through ``nufft.forward``/``adjoint``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import torch
import torch.nn as nn

from . import _lib, ops
from .dtypes import complex_dtype, real_dtype
from .fourier import NUFFT, fft, hip_nufft, ifft, sigpy_nufft, toeplitz_unshifted
from .pad import PadLast
from .plan import SamplePlan
from .utils import batch_iterator

MAX_GRID_BYTES = 32 << 30  # coil chunking threshold for the fused path (grids + gridding scratch)


@dataclass
class batching_params:
    coil_batch_size: Optional[int] = 1
    sub_batch_size: Optional[int] = 1
    field_batch_size: Optional[int] = 1
    toeplitz_batch_size: Optional[int] = 1
    img_batch_size: Optional[int] = 1


class linop(nn.Module):
    """Generic linear operator: ishape, oshape, forward, adjoint, normal (linops.py:30-47)."""

    def __init__(self, ishape, oshape):
        super().__init__()
        self.ishape = ishape
        self.oshape = oshape

    def forward(self, *args):
        raise NotImplementedError

    def adjoint(self, *args):
        raise NotImplementedError

    def normal(self, *args):
        raise NotImplementedError


class type3_nufft_naive(linop):
    """y(t) = sum_r x(r) exp(-2 pi i phi(r) . alpha(t)) by an explicit encoding matrix
    (linops.py:49-121); phis (B, *im_size), alphas (B, *trj_size)."""

    def __init__(self, phis: torch.Tensor, alphas: torch.Tensor):
        super().__init__(phis.shape[1:], alphas.shape[1:])
        B = phis.shape[0]
        assert phis.shape[0] == alphas.shape[0], "phis and alphas must have same number of bases (B)"
        self.phis = phis.reshape((B, -1))
        self.alphas = alphas.reshape((B, -1))

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        N = x.shape[0]
        enc = torch.exp(-2j * torch.pi * (self.phis.T @ self.alphas))
        return (x.reshape((N, -1)) @ enc).reshape((N, *self.oshape))

    def adjoint(self, y: torch.Tensor) -> torch.Tensor:
        N = y.shape[0]
        enc = torch.exp(2j * torch.pi * (self.alphas.T @ self.phis))
        return (y.reshape((N, -1)) @ enc).reshape((N, *self.ishape))

    def normal(self, x: torch.Tensor) -> torch.Tensor:
        return self.adjoint(self.forward(x))


class type3_nufft(linop):
    """Type-3 NUFFT y_k = sum_n x_n exp(-2 pi i phi_n . alpha_k) (linops.py:123-425), after
    Barnett et al. (SIAM J. Sci. Comput. 2019): KB gridding of the image samples at
    rescaled phi onto an oversampled grid, a type-2 NUFFT of that grid at rescaled
    alpha, deapodization and midpoint phases.

    The construction (centring, rescaling to [-1/2, 1/2], grid sizes, one KB beta per
    dimension, the inner ``sigpy_nufft`` with its defaults) is the reference's
    (:157-207). Every application runs on the HIP kernels: the image-domain gridding
    and interpolation through one sample plan with per-dimension betas, the inner
    NUFFT fused as in ``hip_nufft``, and the per-sample phase / apodization factors --
    precomputed once here -- as one complex multiply kernel on each side.
    ``use_toep`` convolves the gridded image with the inner NUFFT's PSF instead of
    running it forward and back (:332-425).
    """

    def __init__(self, phis: torch.Tensor, alphas: torch.Tensor, oversamp: float = 2.0, width: float = 4.0,
                 use_toep: Optional[bool] = False):
        B, *im_size = phis.shape
        B_, *trj_size = alphas.shape
        assert B == B_, "phis and alphas must have same number of bases (B)"
        assert phis.device == alphas.device, "phis and alphas must be on same device"
        _lib.require_device(phis, alphas)
        super().__init__(tuple(im_size), tuple(trj_size))
        R, T = math.prod(im_size), math.prod(trj_size)
        phis_flt = phis.reshape((B, R)).to(real_dtype)
        alphas_flt = alphas.reshape((B, T)).to(real_dtype)
        # centre both sets and scale phi into [-1/2, 1/2] (alpha scaled inversely)
        phis_mp = (phis_flt.min(dim=1).values + phis_flt.max(dim=1).values) / 2
        alphas_mp = (alphas_flt.min(dim=1).values + alphas_flt.max(dim=1).values) / 2
        phis_c = phis_flt - phis_mp[:, None]
        alphas_c = alphas_flt - alphas_mp[:, None]
        scales = phis_c.abs().max(dim=1).values * 2
        phis_c /= scales[:, None]
        phis_mp /= scales
        alphas_c *= scales[:, None]
        alphas_mp *= scales
        self.phis, self.alphas, self.phis_mp, self.alphas_mp = phis_c, alphas_c, phis_mp, alphas_mp
        self.torch_dev = phis.device
        self.B = B
        # grid of the image-domain gridding (:190-203)
        self.grd_S = self.alphas.abs().max(dim=1).values
        self.grd_W = width
        n_os = torch.ceil(2 * self.grd_S * oversamp + self.grd_W).long()
        self.grd_os = n_os / (n_os / oversamp).round()
        self.grd_N = n_os / self.grd_os
        self.grd_gamma = n_os / (2 * self.grd_os * self.grd_S)
        self.grd_beta = math.pi * (((self.grd_W / self.grd_os) * (self.grd_os - 0.5)) ** 2 - 0.8) ** 0.5
        self.grd_N_os = tuple((self.grd_N * self.grd_os).ceil().long().tolist())
        self.nft = sigpy_nufft(self.grd_N_os)
        # image samples on that grid: one plan, KB width W, one beta per dimension
        sc = (self.grd_os * self.grd_N).ceil() / self.grd_N
        sh = (self.grd_os * self.grd_N).ceil() // 2
        img_trj = (sc * (self.phis.T * self.grd_N / self.grd_gamma) + sh).contiguous()
        betas = [float(b) for b in self.grd_beta.tolist()]
        if B <= 3:
            self.img_plan = SamplePlan(img_trj, self.grd_N_os, float(width), betas)
        else:  # 4 or 5 bases: the direct kernels, coordinates wrapped once here
            self.img_plan = None
            G = torch.tensor(self.grd_N_os, dtype=torch.float32, device=img_trj.device)
            self._img_t = torch.remainder(img_trj.to(torch.float32), G).contiguous()
            self._img_betas = betas
            if use_toep:
                raise NotImplementedError("type3_nufft: Toeplitz kernels with more than 3 bases")
        self._trj = (self.alphas.T * self.grd_gamma).contiguous()          # inner NUFFT trajectory (T, B)
        apod = self.apod(self.alphas.T * self.grd_gamma / self.grd_N / self.grd_os, self.grd_beta,
                         self.grd_W).prod(dim=-1)
        root = float(math.prod(self.grd_N_os)) ** 0.5
        two_pi_i = 2j * math.pi
        # forward: x * f_in -> grid / W^B -> NUFFT -> * f_out
        self._f_in = torch.exp(-two_pi_i * (self.phis.T @ self.alphas_mp)).to(complex_dtype)
        self._f_out = (apod * root * torch.exp(-two_pi_i * ((self.alphas.T + self.alphas_mp) @ self.phis_mp))
                       ).to(complex_dtype)
        # adjoint: y * g_in -> adjoint NUFFT -> interp / W^B -> * g_out
        self._g_in = (apod * root * torch.exp(two_pi_i * (self.alphas.T @ self.phis_mp))).to(complex_dtype)
        self._g_out = torch.exp(two_pi_i * ((self.phis.T + self.phis_mp) @ self.alphas_mp)).to(complex_dtype)
        if use_toep:
            self.kerns = self.nft.calc_teoplitz_kernels(trj=self._trj[None], weights=apod[None] ** 2)
            self._n_in = torch.exp(-two_pi_i * ((self.phis.T + self.phis_mp) @ self.alphas_mp)).to(complex_dtype)
        else:
            self.kerns = None

    @staticmethod
    def apod(x, beta, W):
        """KB apodization (:209-212)."""
        a = (beta ** 2 - (math.pi * W * x) ** 2) ** 0.5
        return a / torch.sinh(a)

    def _direct(self, adjoint: bool, v: torch.Tensor, n: int) -> torch.Tensor:
        B, G = self.B, self.grd_N_os
        P = int(math.floor(self.grd_W)) + 1
        M = self._img_t.shape[0]
        args = (B, P, _lib.ptr(v), _lib.ptr(self._img_t), M, n, _lib.ints(G), _lib.floats(self._img_betas),
                float(self.grd_W), 1.0 / self.grd_W ** B)
        if adjoint:
            out = torch.zeros((n, *G), dtype=complex_dtype, device=v.device)
            _lib.check(_lib.lib().mra_grid_direct(*args, _lib.ptr(out), _lib.stream_ptr(v.device)))
        else:
            out = torch.empty((n, M), dtype=complex_dtype, device=v.device)
            _lib.check(_lib.lib().mra_interp_direct(*args, _lib.ptr(out), _lib.stream_ptr(v.device)))
        return out

    def _grid(self, x: torch.Tensor) -> torch.Tensor:
        n = x.shape[0]
        if self.img_plan is None:
            return self._direct(True, x.contiguous(), n)
        out = torch.empty((n, *self.grd_N_os), dtype=complex_dtype, device=x.device)
        return ops.gridding(x, self.img_plan, out, out_scale=1.0 / self.grd_W ** self.B)

    def _interp(self, g: torch.Tensor) -> torch.Tensor:
        if self.img_plan is None:
            return self._direct(False, g.contiguous(), g.shape[0])
        return ops.interp(g, self.img_plan, out_scale=1.0 / self.grd_W ** self.B)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """x (N, *im_size) -> (N, *trj_size) (:214-268)."""
        N = x.shape[0]
        xm = ops.cmul_bcast(x.reshape(N, -1).to(complex_dtype).clone(), self._f_in)
        g = self._grid(xm)
        y = self.nft.forward(g[None], self._trj[None])[0].contiguous()
        return ops.cmul_bcast(y, self._f_out).reshape((N, *self.oshape))

    def adjoint(self, y: torch.Tensor) -> torch.Tensor:
        """y (N, *trj_size) -> (N, *im_size) (:270-330)."""
        N = y.shape[0]
        ym = ops.cmul_bcast(y.reshape(N, -1).to(complex_dtype).clone(), self._g_in)
        g = self.nft.adjoint(ym[None], self._trj[None])[0].contiguous()
        x = self._interp(g)
        return ops.cmul_bcast(x, self._g_out).reshape((N, *self.ishape))

    def normal(self, x: torch.Tensor) -> torch.Tensor:
        """A^H A x; with Toeplitz kernels the inner NUFFT pair is one PSF convolution (:332-425)."""
        if self.kerns is None:
            return self.adjoint(self.forward(x))
        N = x.shape[0]
        xm = ops.cmul_bcast(x.reshape(N, -1).to(complex_dtype).clone(), self._n_in)
        g = self._grid(xm)
        g = self.nft.normal_toeplitz(g[None], self.kerns)[0] * float(math.prod(self.grd_N_os))
        out = self._interp(g.contiguous())
        return ops.cmul_bcast(out, self._g_out).reshape((N, *self.ishape))


class sense_linop(linop):
    """SENSE model A x = [F(S_c x)]_c with optional DCF in the adjoint and a Toeplitz normal."""

    def __init__(self, trj: torch.Tensor, mps: torch.Tensor, dcf: Optional[torch.Tensor] = None,
                 nufft: Optional[NUFFT] = None, use_toeplitz: Optional[bool] = False,
                 spatial_funcs: Optional[torch.Tensor] = None, temporal_funcs: Optional[torch.Tensor] = None,
                 bparams: Optional[batching_params] = None):
        im_size = tuple(mps.shape[1:])
        trj_size = tuple(trj.shape[:-1])
        ncoils = mps.shape[0]
        super().__init__(im_size, (ncoils, *trj_size))
        dev = trj.device
        assert mps.device == dev
        if nufft is None:
            nufft = sigpy_nufft(im_size)  # the reference's default (linops.py:1086-1087)
        if dcf is None:
            dcf = torch.ones(trj_size, dtype=real_dtype, device=dev)
        else:
            assert dcf.device == dev
        generic_funcs = spatial_funcs is not None or temporal_funcs is not None
        if spatial_funcs is None:
            spatial_funcs = torch.ones((1,) * (len(im_size) + 1), dtype=complex_dtype, device=dev)
        if temporal_funcs is None:
            temporal_funcs = torch.ones((1,) * (len(trj_size) + 1), dtype=complex_dtype, device=dev)
        bparams = batching_params() if bparams is None else bparams

        trj = nufft.rescale_trajectory(trj).type(real_dtype)
        dcf = dcf.type(real_dtype).contiguous()
        mps = mps.type(complex_dtype).contiguous()

        self.im_size = im_size
        self.trj_size = trj_size
        self.use_toeplitz = use_toeplitz
        self.trj = trj
        self.mps = mps
        self.dcf = dcf
        self.nufft = nufft
        self.bparams = bparams
        self.spatial_funcs = spatial_funcs
        self.temporal_funcs = temporal_funcs
        self.torch_dev = dev
        L = max(spatial_funcs.shape[0], temporal_funcs.shape[0])
        self.fused = isinstance(nufft, hip_nufft) and L == 1 and not generic_funcs

        if use_toeplitz:
            if L == 1:
                self.toep_kerns = nufft.calc_teoplitz_kernels(trj[None], dcf[None])[0]
            else:
                w = torch.einsum("a...,b...->ab...", temporal_funcs.conj(), temporal_funcs)
                w = w.reshape(L * L, *trj_size) * dcf[None]
                ks = [nufft.calc_teoplitz_kernels(trj[None], w[b1:b2])
                      for b1, b2 in batch_iterator(w.shape[0], bparams.toeplitz_batch_size)]
                self.toep_kerns = torch.cat(ks, 0).reshape(L, L, *ks[0].shape[1:])
        else:
            self.toep_kerns = None

        if self.fused:
            _lib.require_device(trj, mps)
            if nufft.tune_batch is None:
                nufft.tune_batch = ncoils
            self.plan = nufft.plan(trj)
            self.os_shape = nufft.os_shape()
            self.apod = nufft.apod_table(dev)
            d = len(im_size)
            self._fwd_scale = 1.0 / math.sqrt(math.prod(im_size))
            self._interp_scale = 1.0 / nufft.width ** d
            self._adj_scale = 1.0 / (nufft.width ** d * math.sqrt(math.prod(im_size)))
            if use_toeplitz:
                self.padder = PadLast(self.toep_kerns.shape[-d:], im_size)
                self.toep_unshifted = toeplitz_unshifted(self.toep_kerns)
            # dcf in plan order: gridding reads it contiguously instead of through perm
            self._dcf_plan = self.dcf.reshape(-1)[self.plan.perm.long()].contiguous()
            # per coil: its oversampled (or PSF-sized) grid and its gridding scratch (one
            # subgrid per work item, 2.2x the grid for the 3D kooshball at 128^3)
            per_coil = 8 * (max(math.prod(self.os_shape), math.prod(self.toep_kerns.shape[-d:]) if use_toeplitz else 0)
                            + self.plan.nitems * self.plan.subgrid)
            self.coil_chunk = max(1, min(ncoils, MAX_GRID_BYTES // per_coil))

        self.timers = None  # {stage: [(start_event, end_event), ...]} when profiling

    # ------------------------------------------------------------------
    def _mark(self, stage: str, fn, *args, **kwargs):
        """Run fn; when ``timers`` is a dict, bracket it with HIP events on the current stream."""
        if self.timers is None:
            return fn(*args, **kwargs)
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        out = fn(*args, **kwargs)
        e.record()
        self.timers.setdefault(stage, []).append((s, e))
        return out

    def _coil_chunks(self):
        return batch_iterator(self.mps.shape[0], self.coil_chunk)

    def forward(self, img: torch.Tensor) -> torch.Tensor:
        """img (*im_size) -> ksp (C, *trj_size)."""
        if not self.fused:
            return self._forward_generic(img)
        C = self.mps.shape[0]
        M = self.plan.npts
        ksp = torch.empty((C, M), dtype=complex_dtype, device=self.torch_dev)
        x = img.to(complex_dtype).reshape(1, *self.im_size)
        for c0, c1 in self._coil_chunks():
            g = self._mark("pad_fft", ops.pad_fft, x, self.im_size, self.os_shape, smaps=self.mps[c0:c1],
                           apod=self.apod, scale=self._fwd_scale)
            g = g.view(c1 - c0, *self.os_shape)
            self._mark("interp", ops.interp, g, self.plan, out_scale=self._interp_scale, out=ksp[c0:c1])
        return ksp.reshape(C, *self.trj_size)

    def adjoint(self, ksp: torch.Tensor) -> torch.Tensor:
        """ksp (C, *trj_size) -> img (*im_size), with the DCF applied to ksp first."""
        if not self.fused:
            return self._adjoint_generic(ksp)
        C = self.mps.shape[0]
        y = ksp.to(complex_dtype).reshape(C, -1)
        img = torch.empty((1, *self.im_size), dtype=complex_dtype, device=self.torch_dev)
        for i, (c0, c1) in enumerate(self._coil_chunks()):
            g = torch.empty((c1 - c0, *self.os_shape), dtype=complex_dtype, device=self.torch_dev)
            sc = self._mark("grid_partial", ops.gridding_partial, y[c0:c1], self.plan, c1 - c0,
                            dcf=self._dcf_plan, dcf_plan_order=True)
            self._mark("grid_reduce", ops.gridding_reduce, sc, self.plan, g)
            del sc
            self._mark("ifft_crop", ops.ifft_crop, g, self.im_size, self.os_shape, 1, smaps=self.mps[c0:c1],
                       apod=self.apod, scale=self._adj_scale, out=img, accumulate=i > 0)
        return img[0]

    def normal(self, img: torch.Tensor) -> torch.Tensor:
        """A^H A img (through the PSF kernels when use_toeplitz)."""
        if not self.use_toeplitz:
            if not self.fused:
                return self.adjoint(self.forward(img))
            return self._normal_nufft(img)
        if not self.fused:
            return self._normal_generic(img)
        d = len(self.im_size)
        P = tuple(self.toep_kerns.shape[-d:])
        x = img.to(complex_dtype).reshape(1, *self.im_size)
        out = torch.empty((1, *self.im_size), dtype=complex_dtype, device=self.torch_dev)
        cen = self.padder.center
        for i, (c0, c1) in enumerate(self._coil_chunks()):
            ops.toeplitz_normal(x, self.im_size, P, self.mps[c0:c1], self.toep_unshifted, center=cen, out=out,
                                accumulate=i > 0)
        return out[0]

    def _normal_nufft(self, img: torch.Tensor) -> torch.Tensor:
        """adjoint(forward(img)) with the intermediate k-space in plan (sorted) order: the
        gather stores and gridding loads it contiguously instead of through the sample
        permutation (which scatters 8-byte accesses over the whole k-space vector). The
        same sums in the same order as adjoint(forward(img))."""
        C = self.mps.shape[0]
        M = self.plan.npts
        ksp = torch.empty((C, M), dtype=complex_dtype, device=self.torch_dev)
        x = img.to(complex_dtype).reshape(1, *self.im_size)
        out = torch.empty((1, *self.im_size), dtype=complex_dtype, device=self.torch_dev)
        for c0, c1 in self._coil_chunks():
            g = self._mark("pad_fft", ops.pad_fft, x, self.im_size, self.os_shape, smaps=self.mps[c0:c1],
                           apod=self.apod, scale=self._fwd_scale)
            g = g.view(c1 - c0, *self.os_shape)
            geo = self.plan.geom(c1 - c0, M, self._interp_scale, ksp_plan_order=True)
            self._mark("interp", ops.interp, g, self.plan, out_scale=self._interp_scale, out=ksp[c0:c1], geom=geo)
        for i, (c0, c1) in enumerate(self._coil_chunks()):
            g = torch.empty((c1 - c0, *self.os_shape), dtype=complex_dtype, device=self.torch_dev)
            sc = self._mark("grid_partial", ops.gridding_partial, ksp[c0:c1], self.plan, c1 - c0,
                            dcf=self._dcf_plan, dcf_plan_order=True, ksp_plan_order=True)
            self._mark("grid_reduce", ops.gridding_reduce, sc, self.plan, g)
            del sc
            self._mark("ifft_crop", ops.ifft_crop, g, self.im_size, self.os_shape, 1, smaps=self.mps[c0:c1],
                       apod=self.apod, scale=self._adj_scale, out=out, accumulate=i > 0)
        return out[0]

    # -- reference loops (any NUFFT, L >= 1 segments): linops.py:1134-1308 ------------
    def _forward_generic(self, img):
        L = self.spatial_funcs.shape[0]
        nc = self.mps.shape[0]
        ksp = torch.zeros((nc, *self.trj_size), dtype=complex_dtype, device=self.torch_dev)
        for c, d in batch_iterator(nc, self.bparams.coil_batch_size):
            sx = self.mps[c:d] * img
            for l1, l2 in batch_iterator(L, self.bparams.field_batch_size):
                sbx = sx[:, None, ...] * self.spatial_funcs[l1:l2]
                fsbx = self.nufft.forward(sbx[None], self.trj[None])[0]
                ksp[c:d] += (fsbx * self.temporal_funcs[l1:l2]).sum(dim=-len(self.trj_size) - 1)
        return ksp

    def _adjoint_generic(self, ksp):
        L = self.temporal_funcs.shape[0]
        nc = self.mps.shape[0]
        img = torch.zeros(self.im_size, dtype=complex_dtype, device=self.torch_dev)
        for c, d in batch_iterator(nc, self.bparams.coil_batch_size):
            mps = self.mps[c:d]
            kw = ksp[c:d] * self.dcf[None]
            for l1, l2 in batch_iterator(L, self.bparams.field_batch_size):
                hwy = kw[:, None, ...] * self.temporal_funcs[l1:l2].conj()
                fhwy = self.nufft.adjoint(hwy[None], self.trj[None])[0]
                sfhwy = torch.einsum("cl...,c...->l...", fhwy, mps.conj())
                img += (sfhwy * self.spatial_funcs[l1:l2].conj()).sum(dim=-len(self.im_size) - 1)
        return img

    def _normal_generic(self, img):
        from .fourier import fft, ifft

        L = self.temporal_funcs.shape[0]
        nc = self.mps.shape[0]
        d = len(self.im_size)
        inner = 1 if d == 3 else L
        padder = PadLast(self.toep_kerns.shape[-d:], self.im_size)
        kern = self.toep_kerns if self.toep_kerns.ndim == 2 + d else self.toep_kerns[None, None]
        out = torch.zeros_like(img, dtype=complex_dtype)
        dims = tuple(range(-d, 0))
        for c, dd in batch_iterator(nc, self.bparams.coil_batch_size):
            mps = self.mps[c:dd]
            sx = mps * img
            for l1, l2 in batch_iterator(L, self.bparams.field_batch_size):
                fsbx = fft(padder.forward(sx[:, None, ...] * self.spatial_funcs[l1:l2]), dim=dims)
                for m1, m2 in batch_iterator(L, inner):
                    mf = torch.einsum("ml...,cl...->cm...", kern[m1:m2, l1:l2], fsbx)
                    f = padder.adjoint(ifft(mf, dim=dims))
                    b = (f * self.spatial_funcs[m1:m2].conj()).sum(dim=-d - 1)
                    out += (b * mps.conj()).sum(dim=0)
        return out


class imperf_coil_lowrank(linop):
    """Coils and imperfections folded into one low-rank model (linops.py:525-759): the
    k-space of coil c is sum_l h[l, c] * F(B_l x), spatial functions B (L, *im_size),
    temporal functions h (L, C, *trj_size).

    One NUFFT call transforms all L images B_l x (the reference loops over segment
    batches); the contraction with h is one HIP pass (``mra_lowrank_expand``) and the
    adjoint's conj(h)-weighted, dcf-weighted coil sum another (``mra_lowrank_reduce``).
    The Toeplitz normal uses the L x L PSF kernels of the weights
    sum_c conj(h[l1, c]) h[l2, c] dcf (as the reference builds them, :575-590) and returns
    sum_lo conj(B_lo) crop(iFFT(sum_li K[lo, li] FFT(pad(B_li x)))), i.e. A^H A up to the
    Toeplitz approximation. (The reference's Toeplitz branch, :738-757, adds the (L, *im)
    array of un-weighted segment images into a (*im) image, which fails to broadcast.)
    ``nufft`` defaults to ``sigpy_nufft`` as in the reference.
    """

    def __init__(self, trj: torch.Tensor, spatial_funcs: torch.Tensor, temporal_funcs: torch.Tensor,
                 dcf: Optional[torch.Tensor] = None, nufft: Optional[NUFFT] = None,
                 use_toeplitz: Optional[bool] = False, bparams: Optional[batching_params] = None):
        im_size = tuple(spatial_funcs.shape[1:])
        oshape = tuple(temporal_funcs.shape[1:])
        super().__init__(im_size, oshape)
        L = spatial_funcs.shape[0]
        C = temporal_funcs.shape[1]
        torch_dev = trj.device
        assert L == temporal_funcs.shape[0]
        assert temporal_funcs.device == torch_dev
        assert spatial_funcs.device == torch_dev
        _lib.require_device(trj, spatial_funcs, temporal_funcs)
        if nufft is None:
            nufft = sigpy_nufft(im_size)
        if dcf is None:
            dcf = torch.ones(trj.shape[:-1], dtype=real_dtype, device=torch_dev)
        else:
            assert dcf.device == torch_dev
        trj = nufft.rescale_trajectory(trj).type(real_dtype)
        self.L, self.C = L, C
        self.im_size = im_size
        self.trj_size = tuple(trj.shape[:-1])
        self.use_toeplitz = use_toeplitz
        self.spatial_funcs = spatial_funcs.type(complex_dtype).contiguous()
        self.temporal_funcs = temporal_funcs.type(complex_dtype).contiguous()
        if tuple(self.temporal_funcs.shape[2:]) != self.trj_size:
            raise ValueError("temporal_funcs must have shape (L, C, *trj_size)")
        self.trj = trj
        self.dcf = dcf.type(real_dtype).contiguous()
        self.nufft = nufft
        self.bparams = bparams if bparams is not None else batching_params()
        self.torch_dev = torch_dev
        self.M = int(math.prod(self.trj_size))
        if use_toeplitz:
            h = self.temporal_funcs.reshape(L, C, -1)
            w = torch.einsum("acm,bcm->abm", h.conj(), h).reshape(L * L, *self.trj_size) * self.dcf
            kerns = nufft.calc_teoplitz_kernels(trj[None], w)                        # (L L) *im_os
            self.toep_kerns = kerns.reshape(L, L, *kerns.shape[1:])
        else:
            self.toep_kerns = None

    def forward(self, img: torch.Tensor) -> torch.Tensor:
        """img (*im_size) -> ksp (C, *trj_size)."""
        bx = img.type(complex_dtype) * self.spatial_funcs                            # L *im
        fbx = self.nufft.forward(bx[None], self.trj[None])[0].reshape(self.L, self.M).contiguous()
        ksp = torch.empty((self.C, *self.trj_size), dtype=complex_dtype, device=self.torch_dev)
        _lib.check(_lib.lib().mra_lowrank_expand(_lib.ptr(fbx), _lib.ptr(self.temporal_funcs), self.L, self.C,
                                                 self.M, _lib.ptr(ksp), _lib.stream_ptr(self.torch_dev)))
        return ksp

    def adjoint(self, ksp: torch.Tensor) -> torch.Tensor:
        """ksp (C, *trj_size) -> img (*im_size), with the dcf applied to ksp first."""
        y = ops._c64(ksp.type(complex_dtype), "ksp")
        if y.numel() != self.C * self.M:
            raise ValueError(f"ksp must have shape ({self.C}, *{self.trj_size})")
        hdy = torch.empty((self.L, *self.trj_size), dtype=complex_dtype, device=self.torch_dev)
        _lib.check(_lib.lib().mra_lowrank_reduce(_lib.ptr(y), _lib.ptr(self.temporal_funcs), _lib.ptr(self.dcf),
                                                 self.L, self.C, self.M, _lib.ptr(hdy),
                                                 _lib.stream_ptr(self.torch_dev)))
        f = self.nufft.adjoint(hdy[None], self.trj[None])[0]                          # L *im
        return (f * self.spatial_funcs.conj()).sum(dim=0)

    def normal(self, img: torch.Tensor) -> torch.Tensor:
        if not self.use_toeplitz:
            return self.adjoint(self.forward(img))
        d = len(self.im_size)
        im_os = tuple(self.toep_kerns.shape[-d:])
        padder = PadLast(im_os, self.im_size)
        dims = tuple(range(-d, 0))
        F = fft(padder.forward(img.type(complex_dtype) * self.spatial_funcs), dim=dims)   # L *os
        out = torch.zeros(self.im_size, dtype=complex_dtype, device=img.device)
        for lo in range(self.L):
            m = (self.toep_kerns[lo] * F).sum(dim=0)
            out += padder.adjoint(ifft(m, dim=dims)) * self.spatial_funcs[lo].conj()
        return out


class subspace_linop(linop):
    """Subspace SENSE model (linops.py:1496-1772): the k-space of coil c at time point t is
    sum_l T_l(t) sum_k phi_k(t) F(S_c B_l alpha_k), with phi (nsub, ntr) indexed by the
    trajectory's last batch axis, alpha (nsub, *im_size) the coefficient images and
    optional imperfection segments B_l (spatial) / T_l (temporal).

    All coils, coefficients and segments go through one HIP NUFFT call per application
    (the reference loops over coil / coefficient / segment batches); the contraction
    with phi and the segment sums are elementwise torch work on the NUFFT's output. The
    Toeplitz normal builds the nsub x nsub PSF kernels of conj(phi_k1) phi_k2 dcf as the
    reference does (:1563-1580) and mixes the coefficient images through them.
    ``nufft`` defaults to ``hip_nufft`` (the reference defaults to torchkbnufft).
    """

    def __init__(self, im_size: tuple, trj: torch.Tensor, mps: torch.Tensor, phi: torch.Tensor,
                 dcf: Optional[torch.Tensor] = None, nufft: Optional[NUFFT] = None,
                 use_toeplitz: Optional[bool] = False, spatial_funcs: Optional[torch.Tensor] = None,
                 temporal_funcs: Optional[torch.Tensor] = None, bparams: Optional[batching_params] = None):
        ishape = (phi.shape[0], *im_size)
        oshape = (mps.shape[0], *trj.shape[:-1])
        super().__init__(ishape, oshape)
        torch_dev = trj.device
        assert phi.device == torch_dev
        assert mps.device == torch_dev
        _lib.require_device(trj, mps, phi)
        if nufft is None:
            nufft = hip_nufft(im_size)
        if dcf is None:
            dcf = torch.ones(trj.shape[:-1], dtype=real_dtype, device=torch_dev)
        else:
            assert dcf.device == torch_dev
        if spatial_funcs is None:
            spatial_funcs = torch.ones((1,) * len(ishape), dtype=complex_dtype, device=torch_dev)
        if temporal_funcs is None:
            temporal_funcs = torch.ones((1,) * len(oshape), dtype=complex_dtype, device=torch_dev)
        trj = nufft.rescale_trajectory(trj).type(real_dtype)
        self.im_size = tuple(im_size)
        self.trj_size = tuple(trj.shape[:-1])
        self.use_toeplitz = use_toeplitz
        self.trj = trj
        self.phi = phi.type(complex_dtype)
        self.mps = mps.type(complex_dtype)
        self.dcf = dcf.type(real_dtype)
        self.nufft = nufft
        self.bparams = bparams if bparams is not None else batching_params()
        self.torch_dev = torch_dev
        self.spatial_funcs = spatial_funcs.type(complex_dtype)
        self.temporal_funcs = temporal_funcs.type(complex_dtype)
        if use_toeplitz:
            if math.prod(spatial_funcs.shape) > 1:
                raise NotImplementedError
            K = self.phi.shape[0]
            pp = self.phi.conj()[:, None, :] * self.phi[None]                   # nsub nsub ntr
            w = pp.reshape(K * K, *([1] * (len(self.trj_size) - 1)), pp.shape[-1]) * self.dcf
            kerns = nufft.calc_teoplitz_kernels(trj[None], w[None])[0]         # (K K) *im_os
            self.toep_kerns = kerns.reshape(K, K, *kerns.shape[1:])
        else:
            self.toep_kerns = None

    def _phi_b(self) -> torch.Tensor:
        """phi (nsub, ntr) broadcast against (..., nsub, nseg, *trj_size)."""
        K = self.phi.shape[0]
        return self.phi.reshape(K, 1, *([1] * (len(self.trj_size) - 1)), -1)

    def forward(self, alphas: torch.Tensor) -> torch.Tensor:
        """alphas (nsub, *im_size) -> ksp (nc, *trj_size)."""
        d = len(self.im_size)
        mw = self.mps[:, None] * self.spatial_funcs                                 # nc nseg *im
        sx = mw[:, None] * alphas.type(complex_dtype)[None, :, None]                 # nc nsub nseg *im
        nc, K, L = sx.shape[:3]
        fsx = self.nufft.forward(sx.reshape(1, nc, K, L, *self.im_size), self.trj[None])[0]   # nc nsub nseg *trj
        pfsx = (fsx * self._phi_b()).sum(dim=1)                                      # nc nseg *trj
        return (pfsx * self.temporal_funcs).sum(dim=1)

    def adjoint(self, ksp: torch.Tensor) -> torch.Tensor:
        """ksp (nc, *trj_size) -> alphas (nsub, *im_size)."""
        wy = (ksp * self.dcf)[:, None] * self.temporal_funcs.conj()                 # nc nseg *trj
        pwy = wy[:, None] * self._phi_b().conj()                                     # nc nsub nseg *trj
        f = self.nufft.adjoint(pwy[None], self.trj[None])[0]                         # nc nsub nseg *im
        sf = (f * self.mps.conj()[:, None, None]).sum(dim=0)                         # nsub nseg *im
        return (sf * self.spatial_funcs.conj()).sum(dim=1)

    def normal(self, alphas: torch.Tensor) -> torch.Tensor:
        if not self.use_toeplitz:
            return self.adjoint(self.forward(alphas))
        d = len(self.im_size)
        K = self.phi.shape[0]
        im_os = tuple(self.toep_kerns.shape[-d:])
        padder = PadLast(im_os, self.im_size)
        dims = tuple(range(-d, 0))
        sx = self.mps[:, None] * alphas.type(complex_dtype)[None]                   # nc nsub *im
        F = fft(padder.forward(sx), dim=dims)                                        # nc nsub *os
        out = torch.zeros((K, *self.im_size), dtype=complex_dtype, device=alphas.device)
        for i in range(K):
            m = (self.toep_kerns[i][None] * F).sum(dim=1)                            # nc *os
            out[i] = (padder.adjoint(ifft(m, dim=dims)) * self.mps.conj()).sum(dim=0)
        return out
