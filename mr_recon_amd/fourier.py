"""NUFFT call surface on the HIP kernels.

Mirrors the reference ``mr_recon.fourier`` for the hot path:
  ``fft`` / ``ifft``            fourier.py:37-61   (centred FFTs, plain torch)
  ``calc_toep_kernel_helper``   fourier.py:63-105
  ``NUFFT``                     fourier.py:107-352 (base class, ``normal_toeplitz``)
  ``hip_nufft``                 fourier.py:354-582 (``triton_nufft``; alias kept)
  ``sigpy_nufft``               fourier.py:728-1033 (same kernels, sigpy's apodization)
  ``gridded_nufft``             fourier.py:1213-1326 (nearest-grid gather / scatter)

Transform definition (the reference's): forward
    ksp = interp_KB( FFT( pad_G( apod * img / sqrt(prod N) ) ) ) / w^d
with the pad centred (utils.resize) and the FFT centred (ifftshift / fftshift),
and the adjoint its exact conjugate transpose. ``forward`` and ``adjoint`` run
fused: one pad/apodize kernel writes the oversampled grid straight into the
unshifted FFT layout, rocFFT runs in place, and the KB kernels address that
layout through an index offset, so no shift, pad or crop copy is made. The
split methods (``forward_FT_only`` etc.) keep the reference's centred layout of
the intermediate grid so they compose with code written against it.
"""

from __future__ import annotations

import math
from collections import OrderedDict
from math import ceil
from typing import Optional

import numpy as np
import torch
import torch.fft as fft_torch
import torch.nn as nn

from . import _lib, ops
from .dtypes import complex_dtype, real_dtype
from .gridfft import _fftn, _ifftn, _lds_fft  # noqa: F401  (in-place grid FFTs)
from .pad import PadLast
from .plan import SamplePlan
from .utils import batch_iterator, resize


def fft(x, dim=None, oshape=None, norm="ortho"):
    """Centred FFT: fftshift(fftn(ifftshift(x))) (reference fourier.py:37-47)."""
    if oshape is not None:
        x = resize(x, oshape)
    x = fft_torch.ifftshift(x, dim=dim)
    x = fft_torch.fftn(x, dim=dim, norm=norm)
    return fft_torch.fftshift(x, dim=dim)


def ifft(x, dim=None, oshape=None, norm="ortho"):
    """Centred inverse FFT (reference fourier.py:50-61)."""
    if oshape is not None:
        x = resize(x, oshape)
    x = fft_torch.ifftshift(x, dim=dim)
    x = fft_torch.ifftn(x, dim=dim, norm=norm)
    return fft_torch.fftshift(x, dim=dim)


def calc_toep_kernel_helper(nufft_adj_os: callable, trj: torch.Tensor, weights: Optional[torch.Tensor] = None):
    """PSF kernels: centred FFT of the adjoint NUFFT of ``weights`` (reference fourier.py:63-105)."""
    dev = trj.device
    if weights is None:
        weights = torch.ones(trj.shape[:-1], dtype=real_dtype, device=dev)
    else:
        assert weights.device == dev
    trj_batch = trj.shape[1:-1]
    if trj.shape[0] == 1 and weights.shape[0] != 1:
        trj = trj.expand((weights.shape[0], *trj.shape[1:]))
    N = trj.shape[0]
    d = trj.shape[-1]
    ksp = torch.ones((N, 1, *trj_batch), device=dev, dtype=complex_dtype)
    ksp_weighted = ksp * weights[:, None, ...]
    img = nufft_adj_os(ksp_weighted, trj)[:, 0, ...]
    return fft(img, dim=tuple(range(-d, 0)))


class NUFFT(nn.Module):
    """Forward NUFFT e^{-j 2 pi k . r}, k_i in [-N_i/2, N_i/2], r_i in [-1/2, 1/2] (fourier.py:107-352)."""

    def __init__(self, im_size: tuple):
        super().__init__()
        self.im_size = tuple(im_size)

    def rescale_trajectory(self, trj: torch.Tensor) -> torch.Tensor:
        return trj

    def forward_FT_only(self, img):
        raise NotImplementedError

    def adjoint_iFT_only(self, ksp_os):
        raise NotImplementedError

    def adjoint_grid_only(self, ksp, trj):
        raise NotImplementedError

    def forward_interp_only(self, ksp_os, trj):
        raise NotImplementedError

    def forward(self, img: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        return self.forward_interp_only(self.forward_FT_only(img), trj)

    def adjoint(self, ksp: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        return self.adjoint_iFT_only(self.adjoint_grid_only(ksp, trj))

    def calc_teoplitz_kernels(self, trj, weights=None, os_factor=2.0):
        raise NotImplementedError

    def normal_toeplitz(self, img: torch.Tensor, toeplitz_kernels: torch.Tensor) -> torch.Tensor:
        """A^H A img through the PSF kernels (reference fourier.py:302-352).

        Same composition as the reference (PadLast pad, centred ortho FFT,
        multiply, centred ortho iFFT, crop), run as fused pad+FFT -> multiply
        kernel -> fused iFFT+crop in the unshifted layout.
        """
        _lib.require_device(img, toeplitz_kernels)
        N = img.shape[0]
        d = len(self.im_size)
        im_size_os = tuple(toeplitz_kernels.shape[1:])
        img_batch = img.shape[1:-d]
        B = int(math.prod(img_batch))
        padder = PadLast(im_size_os, self.im_size)
        cen = padder.center
        out = torch.empty((N, B, *self.im_size), dtype=complex_dtype, device=img.device)
        for n in range(N):
            kern = toeplitz_unshifted(toeplitz_kernels[n])
            x = img[n].reshape(B, *self.im_size).to(complex_dtype)
            g = ops.pad_fft(x, self.im_size, im_size_os, center=cen)
            ops.cmul_bcast(g, kern)
            ops.ifft_crop(g, self.im_size, im_size_os, B, center=cen, out=out[n])
        return out.reshape(N, *img_batch, *self.im_size)


def toeplitz_unshifted(kern: torch.Tensor) -> torch.Tensor:
    """Centred ortho-normalised PSF kernel -> multiplier for unnormalised FFTs in the unshifted layout."""
    d = kern.ndim
    P = int(math.prod(kern.shape))
    return (fft_torch.ifftshift(kern, dim=tuple(range(-d, 0))) / P).to(complex_dtype).contiguous()


def kb_beta(width: float, oversamp: float) -> float:
    """Beatty et al. shape parameter, as the reference (fourier.py:372)."""
    return float(np.pi * (((width / oversamp) * (oversamp - 0.5)) ** 2 - 0.8) ** 0.5)


class _PlanCache:
    """LRU of sample plans keyed by the trajectory tensor's memory and version.

    Each entry holds a reference to the trajectory tensor, so its memory cannot
    be recycled for another trajectory while the plan is cached; in-place edits
    bump the version counter and miss the cache.
    """

    def __init__(self, size: int = 4):
        self.size = size
        self.d: OrderedDict = OrderedDict()

    def get(self, trj: torch.Tensor, extra, build):
        key = (trj.data_ptr(), trj._version, tuple(trj.shape), tuple(trj.stride()), trj.dtype, str(trj.device),
               extra)
        hit = self.d.get(key)
        if hit is not None:
            self.d.move_to_end(key)
            return hit[0]
        plan = build()
        self.d[key] = (plan, trj)
        while len(self.d) > self.size:
            self.d.popitem(last=False)
        return plan


class hip_nufft(NUFFT):
    """Kaiser-Bessel NUFFT on hand-written HIP kernels (drop-in for ``triton_nufft``, fourier.py:354-582)."""

    def __init__(self, im_size: tuple, oversamp: Optional[float] = 1.25, width: Optional[int] = 6,
                 beta: Optional[float] = None, apodize: Optional[bool] = True,
                 samples_per_item=None, tune_batch: Optional[int] = None, tile=None):
        """``samples_per_item``: work-item size of the KB kernels (None: a size heuristic;
        "auto": timed by the autotuner on the first plan of each trajectory class, for
        ``tune_batch`` grids per call). ``tile``: tile edge of the plans (None: the default
        for the dimensionality; an int; "auto": chosen by the autotuner together with the
        work-item size, 2D only -- 1D and 3D have one compiled edge)."""
        super().__init__(im_size)
        self.oversamp = oversamp
        self.width = width
        self.apodize = apodize
        self.beta = kb_beta(width, oversamp) if beta is None else beta
        self.samples_per_item = samples_per_item
        self.tune_batch = tune_batch
        self.tile = tile
        self._plans = _PlanCache()
        self._apod_cache: dict = {}
        self._scale_cache: dict = {}

    # -- geometry ------------------------------------------------------
    def os_shape(self, im_size=None):
        im_size = self.im_size if im_size is None else im_size
        return tuple(ceil(self.oversamp * n) for n in im_size)

    def _scale_shift(self, im_size):
        """Per-dim (fl32(ceil(sN)/N), ceil(sN)//2) of ``_scale_trj`` (fourier.py:420-430)."""
        sc = [float(np.float32(ceil(self.oversamp * n) / n)) for n in im_size]
        sh = [float(ceil(self.oversamp * n) // 2) for n in im_size]
        return sc, sh

    def _apodization_func(self, x: torch.Tensor) -> torch.Tensor:
        """sqrt(beta^2 - (pi w x)^2) / sinh(sqrt(...))   (fourier.py:376-396)."""
        apod = (self.beta ** 2 - (np.pi * self.width * x) ** 2) ** 0.5
        return apod / torch.sinh(apod)

    def apod_table(self, device, im_size=None) -> Optional[torch.Tensor]:
        """Concatenated per-dim apodization (pixel n: x = (n - N//2) / ceil(s N))."""
        if not self.apodize:
            return None
        im_size = tuple(self.im_size if im_size is None else im_size)
        key = (str(device), im_size)
        t = self._apod_cache.get(key)
        if t is None:
            parts = []
            for n in im_size:
                crds = torch.arange(n, dtype=torch.float64) - n // 2
                parts.append(self._apodization_func(crds / ceil(self.oversamp * n)))
            t = torch.cat(parts).to(torch.float32).to(device)
            self._apod_cache[key] = t
        return t

    def scale_table(self, device, im_size=None) -> torch.Tensor:
        """Per-dim trajectory scale fl32(ceil(sN)/N) (dt/dk of the plan coordinates) as a
        float32 device tensor, cached per (device, image size): building it is a host-to-device
        copy, which waits for the stream."""
        im_size = tuple(self.im_size if im_size is None else im_size)
        key = (str(device), im_size)
        t = self._scale_cache.get(key)
        if t is None:
            t = torch.tensor(self._scale_shift(im_size)[0], dtype=torch.float32).to(device)
            self._scale_cache[key] = t
        return t

    def apodize_img(self, img: torch.Tensor) -> torch.Tensor:
        """In-place apodization of the trailing image dims (fourier.py:398-418)."""
        ndim = len(self.im_size)
        for i in range(-ndim, 0):
            crds = torch.arange(img.shape[i], device=img.device) - img.shape[i] // 2
            tup = (slice(None),) + (None,) * (-i - 1)
            img *= self._apodization_func(crds / ceil(self.oversamp * img.shape[i]))[tup]
        return img

    def plan(self, trj_n: torch.Tensor, grid_size=None, centred: bool = False, im_size=None) -> SamplePlan:
        """Sample plan of one trajectory (trj_n: (*trj_batch, d), reference units).

        ``centred`` plans address a centred grid (the public split methods);
        otherwise the unshifted FFT output (the fused paths).
        """
        im_size = tuple(self.im_size if im_size is None else im_size)
        G = tuple(self.os_shape(im_size) if grid_size is None else grid_size)
        wrap = (0,) * len(G) if centred else tuple(-(g // 2) for g in G)
        sc, sh = self._scale_shift(im_size)
        t = trj_n.detach()
        if t.dtype != torch.float32:
            t = t.to(torch.float32)

        def build():
            S, T = self.samples_per_item, self.tile
            if S == "auto" or T == "auto":
                from .autotune import tune_work

                T, S = tune_work(t.reshape(-1, len(G)), G, self.width, self.beta, self.tune_batch or 1,
                                 scale=sc, shift=sh, wrap=wrap, tiles=None if T in (None, "auto") else (T,),
                                 samples=None if S in (None, "auto") else (S,), tune_tile=T == "auto")
            return SamplePlan(t, G, self.width, self.beta, scale=sc, shift=sh, wrap=wrap, tile=T,
                              samples_per_item=S)

        return self._plans.get(t, (G, wrap, im_size, self.width, self.beta, self.samples_per_item, self.tile),
                               build)

    # -- split methods (reference layout) -------------------------------
    # -- 4D / 5D: the reference's composition with torch FFTs and the direct KB kernels -------
    def _scaled_coords(self, trj_n: torch.Tensor, im_size) -> torch.Tensor:
        """_scale_trj (fourier.py:420-430): t = fl32(fl32(k * fl32(G/N)) + G//2), per dim."""
        sc, sh = self._scale_shift(im_size)
        t = trj_n.to(torch.float32)
        return (t * torch.tensor(sc, dtype=torch.float32, device=t.device)) + torch.tensor(
            sh, dtype=torch.float32, device=t.device)

    def _direct(self, adjoint: bool, v: torch.Tensor, t: torch.Tensor, G, nbatch: int) -> torch.Tensor:
        d = len(G)
        M = int(math.prod(t.shape[:-1]))
        tw = torch.remainder(t.reshape(-1, d), torch.tensor(G, dtype=torch.float32, device=t.device)).contiguous()
        P = int(math.floor(self.width)) + 1
        args = (d, P, _lib.ptr(v), _lib.ptr(tw), M, nbatch, _lib.ints(G), _lib.floats([self.beta] * d),
                float(self.width), 1.0 / self.width ** d)
        if adjoint:
            out = torch.zeros((nbatch, *G), dtype=complex_dtype, device=v.device)
            _lib.check(_lib.lib().mra_grid_direct(*args, _lib.ptr(out), _lib.stream_ptr(v.device)))
        else:
            out = torch.empty((nbatch, M), dtype=complex_dtype, device=v.device)
            _lib.check(_lib.lib().mra_interp_direct(*args, _lib.ptr(out), _lib.stream_ptr(v.device)))
        return out

    def _forward_FT_generic(self, img: torch.Tensor) -> torch.Tensor:
        d = len(self.im_size)
        x = img.to(complex_dtype).clone()
        if self.apodize:
            x = self.apodize_img(x)
        x = x / math.prod(x.shape[-d:]) ** 0.5
        x = resize(x, tuple(x.shape[:-d]) + self.os_shape(tuple(x.shape[-d:])))
        return fft(x, dim=tuple(range(-d, 0)), norm=None)

    def _adjoint_iFT_generic(self, ksp_os: torch.Tensor) -> torch.Tensor:
        d = len(self.im_size)
        G = tuple(ksp_os.shape[-d:])
        img = ifft(ksp_os.to(complex_dtype), dim=tuple(range(-d, 0)), norm=None)
        img = resize(img, tuple(img.shape[:-d]) + self.im_size) * (math.prod(G) / math.prod(self.im_size) ** 0.5)
        return self.apodize_img(img) if self.apodize else img

    def forward_FT_only(self, img: torch.Tensor) -> torch.Tensor:
        """apodize -> / sqrt(prod N) -> centred zero-pad -> centred FFT (norm=None)."""
        _lib.require_device(img)
        d = len(self.im_size)
        if d > 3:
            return self._forward_FT_generic(img)
        im = tuple(img.shape[-d:])
        G = self.os_shape(im)
        lead = img.shape[:-d]
        x = img.reshape(-1, *im).to(complex_dtype)
        g = ops.pad_fft(x, im, G, apod=self.apod_table(img.device, im), scale=1.0 / math.sqrt(math.prod(im)))
        g = fft_torch.fftshift(g, dim=tuple(range(-d, 0)))
        return g.reshape(*lead, *G)

    def forward_interp_only(self, ksp_os: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        """KB interpolation of a centred grid at ``trj``, / w^d."""
        _lib.require_device(ksp_os, trj)
        d = len(self.im_size)
        G = tuple(ksp_os.shape[-d:])
        N = trj.shape[0]
        lead = ksp_os.shape[1:-d]
        B = int(math.prod(lead))
        trj_batch = trj.shape[1:-1]
        M = int(math.prod(trj_batch))
        out = torch.empty((N, B, M), dtype=complex_dtype, device=ksp_os.device)
        src = ksp_os.to(complex_dtype).reshape(N, B, *G)
        if d > 3:
            im = (N, *lead, *self.im_size)
            for i in range(N):
                out[i] = self._direct(False, src[i].contiguous(), self._scaled_coords(trj[i], im[-d:]), G, B)
            return out.reshape(N, *lead, *trj_batch)
        for i in range(N):
            plan = self.plan(trj[i], grid_size=G, centred=True)
            ops.interp(src[i], plan, out_scale=1.0 / self.width ** d, out=out[i])
        return out.reshape(N, *lead, *trj_batch)

    def adjoint_grid_only(self, ksp: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        """KB gridding onto a centred grid, / w^d."""
        _lib.require_device(ksp, trj)
        d = trj.shape[-1]
        N = trj.shape[0]
        trj_batch = trj.shape[1:-1]
        M = int(math.prod(trj_batch))
        lead = ksp.shape[1: ksp.ndim - len(trj_batch)]
        B = int(math.prod(lead))
        G = self.os_shape()
        out = torch.empty((N, B, *G), dtype=complex_dtype, device=ksp.device)
        src = ksp.to(complex_dtype).reshape(N, B, M)
        if d > 3:
            for i in range(N):
                out[i] = self._direct(True, src[i].contiguous(), self._scaled_coords(trj[i], self.im_size), G, B)
            return out.reshape(N, *lead, *G)
        for i in range(N):
            plan = self.plan(trj[i], grid_size=G, centred=True)
            ops.gridding(src[i], plan, out[i], out_scale=1.0 / self.width ** d)
        return out.reshape(N, *lead, *G)

    def adjoint_iFT_only(self, ksp_os: torch.Tensor) -> torch.Tensor:
        """centred iFFT (norm=None) -> crop -> * prod(G)/sqrt(prod N) -> apodize."""
        _lib.require_device(ksp_os)
        d = len(self.im_size)
        if d > 3:
            return self._adjoint_iFT_generic(ksp_os)
        G = tuple(ksp_os.shape[-d:])
        lead = ksp_os.shape[:-d]
        g = fft_torch.ifftshift(ksp_os.to(complex_dtype), dim=tuple(range(-d, 0))).contiguous()
        g = g.reshape(-1, *G)
        img = ops.ifft_crop(g, self.im_size, G, g.shape[0], apod=self.apod_table(ksp_os.device),
                            scale=1.0 / math.sqrt(math.prod(self.im_size)))
        return img.reshape(*lead, *self.im_size)

    # -- fused transforms ------------------------------------------------
    def forward(self, img: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        """img (N, *img_batch, *im_size), trj (N, *trj_batch, d) -> ksp (N, *img_batch, *trj_batch)."""
        _lib.require_device(img, trj)
        d = len(self.im_size)
        if d > 3:
            return self.forward_interp_only(self.forward_FT_only(img), trj)
        N = trj.shape[0]
        img_batch = img.shape[1:-d]
        B = int(math.prod(img_batch))
        trj_batch = trj.shape[1:-1]
        M = int(math.prod(trj_batch))
        G = self.os_shape()
        apod = self.apod_table(img.device)
        out = torch.empty((N, B, M), dtype=complex_dtype, device=img.device)
        x = img.to(complex_dtype).reshape(img.shape[0], B, *self.im_size)
        for n in range(N):
            plan = self.plan(trj[n])
            g = ops.pad_fft(x[n if x.shape[0] > 1 else 0], self.im_size, G, apod=apod,
                            scale=1.0 / math.sqrt(math.prod(self.im_size)))
            ops.interp(g, plan, out_scale=1.0 / self.width ** d, out=out[n])
        return out.reshape(N, *img_batch, *trj_batch)

    def adjoint(self, ksp: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        """ksp (N, *ksp_batch, *trj_batch), trj (N, *trj_batch, d) -> img (N, *ksp_batch, *im_size)."""
        _lib.require_device(ksp, trj)
        d = len(self.im_size)
        if d > 3:
            return self.adjoint_iFT_only(self.adjoint_grid_only(ksp, trj))
        N = trj.shape[0]
        trj_batch = trj.shape[1:-1]
        M = int(math.prod(trj_batch))
        ksp_batch = ksp.shape[1: ksp.ndim - len(trj_batch)]
        B = int(math.prod(ksp_batch))
        G = self.os_shape()
        apod = self.apod_table(ksp.device)
        out = torch.empty((N, B, *self.im_size), dtype=complex_dtype, device=ksp.device)
        y = ksp.to(complex_dtype).reshape(ksp.shape[0], B, M)
        scale = 1.0 / (self.width ** d * math.sqrt(math.prod(self.im_size)))
        for n in range(N):
            plan = self.plan(trj[n])
            g = torch.empty((B, *G), dtype=complex_dtype, device=ksp.device)
            ops.gridding(y[n if y.shape[0] > 1 else 0], plan, g)
            ops.ifft_crop(g, self.im_size, G, B, apod=apod, scale=scale, out=out[n])
        return out.reshape(N, *ksp_batch, *self.im_size)

    def calc_teoplitz_kernels(self, trj: torch.Tensor, weights: Optional[torch.Tensor] = None,
                              os_factor: Optional[float] = 2.0):
        """Centred PSF kernels (N, *round(os_factor * im_size)) (fourier.py:552-582)."""
        im_size_os = tuple(round(i * os_factor) for i in self.im_size)
        nufft_os = hip_nufft(im_size=im_size_os, oversamp=self.oversamp, width=self.width,
                             samples_per_item=self.samples_per_item, tune_batch=1)
        return calc_toep_kernel_helper(nufft_os.adjoint, trj * os_factor, weights) * (os_factor ** len(self.im_size))


# The reference class name, so code written against mr_recon.fourier.triton_nufft
# switches by changing the import only.
triton_nufft = hip_nufft


def sigpy_beta(width: float, oversamp: float) -> float:
    """sigpy_nufft's default beta (fourier.py:740-747): the usual KB formula, or 1.0
    when its radicand is negative (with the reference's warning unless width == 1)."""
    if ((width / oversamp) * (oversamp - 0.5)) ** 2 - 0.8 < 0:
        if width != 1:
            print("WARNING: beta is set to 1.0")
        return 1.0
    return kb_beta(width, oversamp)


class sigpy_nufft(hip_nufft):
    """Drop-in for the reference's ``sigpy_nufft`` (fourier.py:728-1033) on the HIP kernels.

    The reference runs sigpy's Kaiser-Bessel ``interpolate`` / ``gridding`` on the
    scaled trajectory (fourier.py:901-963): the same transform as ``triton_nufft`` --
    apodize, / sqrt(prod N), centred pad to ceil(s N), centred FFT (norm=None), KB
    interpolation / width^d -- so ``forward``, ``adjoint`` and the split methods are
    ``hip_nufft``'s fused kernels. What differs is kept here: the beta default falls
    back to 1.0 (``sigpy_beta``), the apodization divides by sinh + 1e-12 and continues
    to sqrt(-a) / sin(sqrt(-a)) where beta^2 < (pi w x)^2 (fourier.py:817-841), and the
    Toeplitz kernels come from a ``sigpy_nufft`` of the oversampled size.
    """

    def __init__(self, im_size: tuple, oversamp: Optional[float] = 1.25, width: Optional[int] = 6,
                 beta: Optional[float] = None, apodize: Optional[bool] = True,
                 samples_per_item=None, tune_batch: Optional[int] = None):
        super().__init__(im_size, oversamp=oversamp, width=width,
                         beta=sigpy_beta(width, oversamp) if beta is None else beta, apodize=apodize,
                         samples_per_item=samples_per_item, tune_batch=tune_batch)

    def _apodization_func_kb(self, x: torch.Tensor, beta: float, width: float) -> torch.Tensor:
        """KB apodization with its continuation past beta = pi w |x| (fourier.py:817-841)."""
        eps = 1e-12
        arg = beta ** 2 - (np.pi * width * x) ** 2
        pos = arg.clamp(min=0).sqrt()
        pos = pos / (torch.sinh(pos) + eps)
        neg = (-arg.clamp(max=0)).sqrt()
        neg = neg / (torch.sin(neg) + eps)
        return pos + neg

    def _apodization_func_rect(self, x: torch.Tensor) -> torch.Tensor:
        """Apodization of a rectangular kernel (fourier.py:800-815)."""
        return 1 / torch.sinc(self.width * x / self.oversamp)

    def _apodization_func(self, x: torch.Tensor) -> torch.Tensor:
        return self._apodization_func_kb(x, self.beta, self.width)

    def calc_teoplitz_kernels(self, trj: torch.Tensor, weights: Optional[torch.Tensor] = None,
                              os_factor: Optional[float] = 2.0):
        """Centred PSF kernels (N, *round(os_factor * im_size)) (fourier.py:1007-1033); the
        oversampled operator is a new ``sigpy_nufft`` with the default beta."""
        im_size_os = tuple(round(i * os_factor) for i in self.im_size)
        nufft_os = sigpy_nufft(im_size=im_size_os, oversamp=self.oversamp, width=self.width,
                               samples_per_item=self.samples_per_item, tune_batch=1)
        return calc_toep_kernel_helper(nufft_os.adjoint, trj * os_factor, weights) * (os_factor ** len(self.im_size))

    def optimal_beta(self, nbetas: Optional[int] = 100, num_sigs: Optional[int] = 1000,
                     torch_dev: Optional[torch.device] = None) -> float:
        """The beta on torch.linspace(0.5, 20, nbetas) whose 1D NUFFT, least-squares
        scaled, best matches the exact DFT of ``num_sigs`` random signals at N offsets
        within +-0.5/oversamp of the grid (fourier.py:749-798). The signals are drawn
        from torch's CPU generator as in the reference (its default device is the CPU);
        the transforms run on ``torch_dev`` (default: the current GPU)."""
        from .utils import gen_grd

        dev = torch.device("cuda", torch.cuda.current_device()) if torch_dev is None else torch.device(torch_dev)
        N = max(self.im_size)
        kdevs = torch.linspace(-0.5, 0.5, N) / self.oversamp
        betas = torch.linspace(0.5, 20, nbetas)
        rnd_sigs = torch.randn((num_sigs, N), dtype=complex_dtype)
        rs = gen_grd((N,))[:, 0]
        gt = torch.exp(-2j * np.pi * (kdevs[:, None] @ rs[None, :]))
        vals_gt = (rnd_sigs @ gt.T).to(dev)
        sigs = rnd_sigs.to(dev)
        trj = kdevs[None, :, None].to(dev)
        errs = []
        for beta in betas:
            nft = sigpy_nufft((self.im_size[0],), self.oversamp, self.width, beta=float(beta),
                              samples_per_item=self.samples_per_item)
            vals = nft.forward(sigs[None], trj)[0]
            scale = (vals.conj() * vals_gt).sum() / (vals.conj() * vals).sum()
            errs.append((vals * scale - vals_gt).norm())
        return betas[int(torch.stack(errs).argmin())].item()

class gridded_nufft(NUFFT):
    """Nearest-grid NUFFT (fourier.py:1213-1326): samples rounded to the nearest point of
    a ``round(oversamp * N)`` grid, a centred ortho FFT, and a gather (forward) or
    scatter-add (adjoint) at those points, times oversamp^(d/2).

    ``rescale_trajectory`` turns grid-unit coordinates into the integer grid indices the
    other methods take (as the reference's callers do). The gather and scatter are the
    HIP ``mra_nn_gather`` / ``mra_nn_scatter`` kernels at one flat index per sample.
    ``forward`` / ``adjoint`` run the fused pad+FFT and iFFT+crop kernels in the unshifted
    layout (the index carries the centring), and the adjoint sums each grid point's samples
    in chunks of at most 64 over a stable sort by index, one atomic per chunk sum
    (``mra_nn_scatter_chunks``);
    the split methods keep the reference's centred intermediate grid, torch's FFT and
    the atomic scatter, as does any grid the fused transforms do not cover.
    """

    def __init__(self, im_size: tuple, oversamp: Optional[float] = 1.0):
        super().__init__(im_size)
        self.im_size_os = tuple(round(i * oversamp) for i in self.im_size)
        self.oversamp = oversamp
        self._tables = _PlanCache()

    def rescale_trajectory(self, trj: torch.Tensor) -> torch.Tensor:
        """k -> round(k * oversamp + G/2) mod G as int32 (fourier.py:1222-1233)."""
        trj_rs = trj * self.oversamp
        for i in range(trj_rs.shape[-1]):
            n_over_2 = self.im_size_os[i] / 2
            trj_rs[..., i] = (trj_rs[..., i] + n_over_2).round() % self.im_size_os[i]
        return trj_rs.type(torch.int32)

    def _flat_index(self, trj_n: torch.Tensor, G, unshifted: bool = False) -> torch.Tensor:
        """Flat index of each sample's grid point (cached per trajectory tensor)."""
        return self._tables.get(trj_n, ("idx", tuple(G), unshifted),
                                lambda: self._build_flat_index(trj_n, G, unshifted))

    def _scatter_tables(self, trj_n: torch.Tensor, G):
        """Sorted-chunk tables of the fused adjoint's scatter (cached per trajectory tensor):
        built once, so repeated adjoints on one trajectory run no sort and no host sync."""
        return self._tables.get(trj_n, ("chunks", tuple(G)),
                                lambda: self._chunks(self._flat_index(trj_n, G, unshifted=True)))

    @staticmethod
    def _build_flat_index(trj_n: torch.Tensor, G, unshifted: bool = False) -> torch.Tensor:
        """Flat index of each sample's grid point; indices follow Python's rule (-G..G-1).
        ``unshifted``: position in the unshifted FFT layout, where centred index i sits
        at (i - G//2) mod G."""
        t = trj_n.reshape(-1, len(G)).to(torch.int64)
        Gt = torch.tensor(G, dtype=torch.int64, device=t.device)
        if bool(((t < -Gt) | (t >= Gt)).any()):
            raise IndexError("gridded_nufft: trajectory index out of range of the grid")
        t = torch.remainder(t - Gt // 2 if unshifted else t, Gt)
        idx = torch.zeros(t.shape[0], dtype=torch.int64, device=t.device)
        for d in range(len(G)):
            idx = idx * G[d] + t[:, d]
        return idx.to(torch.int32).contiguous()

    CHUNK = 64  # samples per scatter thread at most

    @staticmethod
    def _chunks(idx: torch.Tensor, chunk: int = CHUNK):
        """Samples sorted by flat index (stable), each run of one index cut into chunks of at
        most ``chunk``: the sort permutation and per chunk its first / past-last sorted
        position and grid index."""
        sidx, perm = torch.sort(idx, stable=True)
        head = torch.ones_like(sidx, dtype=torch.bool)
        head[1:] = sidx[1:] != sidx[:-1]
        starts = torch.nonzero(head).reshape(-1)
        ends = torch.cat([starts[1:], starts.new_tensor([idx.numel()])])
        nch = (ends - starts + chunk - 1) // chunk
        run = torch.repeat_interleave(torch.arange(starts.numel(), device=idx.device), nch)
        first = torch.cumsum(nch, 0) - nch
        k = torch.arange(run.numel(), device=idx.device) - first[run]
        cstart = starts[run] + k * chunk
        cend = torch.minimum(cstart + chunk, ends[run])
        return (perm.to(torch.int32).contiguous(), cstart.to(torch.int32).contiguous(),
                cend.to(torch.int32).contiguous(), sidx[starts][run].contiguous())

    def forward_FT_only(self, img: torch.Tensor) -> torch.Tensor:
        """centred resize to the oversampled grid -> centred ortho FFT (fourier.py:1235-1248)."""
        d = len(self.im_size)
        img_os = resize(img, tuple(img.shape[:-d]) + self.im_size_os)
        return fft(img_os, dim=tuple(range(-d, 0)))

    def forward_interp_only(self, ksp_os: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        """ksp_os (N, *batch, *G) at integer trj (N, *trj_batch, d) * oversamp^(d/2)."""
        _lib.require_device(ksp_os, trj)
        d = trj.shape[-1]
        G = tuple(ksp_os.shape[-d:])
        N = trj.shape[0]
        batch = tuple(ksp_os.shape[1:-d])
        B = int(math.prod(batch))
        M = int(math.prod(trj.shape[1:-1]))
        out = torch.empty((N, B, M), dtype=complex_dtype, device=ksp_os.device)
        scale = float(self.oversamp ** (d / 2))
        for i in range(N):
            g = ksp_os[i].to(complex_dtype).contiguous()
            idx = self._flat_index(trj[i], G)
            _lib.check(_lib.lib().mra_nn_gather(_lib.ptr(g), int(math.prod(G)), _lib.ptr(idx), M, B, scale,
                                                _lib.ptr(out[i]), _lib.stream_ptr(g.device)))
        return out.reshape(N, *batch, *trj.shape[1:-1])

    def adjoint_iFT_only(self, ksp_os: torch.Tensor) -> torch.Tensor:
        """centred ortho iFFT -> centred crop (fourier.py:1267-1276)."""
        d = len(self.im_size)
        img_os = ifft(ksp_os, dim=tuple(range(-d, 0)))
        return resize(img_os, tuple(img_os.shape[:-d]) + self.im_size)

    def adjoint_grid_only(self, ksp: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        """ksp (N, *batch, *trj_batch) added into the oversampled grid at integer trj."""
        _lib.require_device(ksp, trj)
        d = trj.shape[-1]
        G = self.im_size_os
        N = trj.shape[0]
        nt = trj.ndim - 2
        batch = tuple(ksp.shape[1:ksp.ndim - nt])
        B = int(math.prod(batch))
        M = int(math.prod(trj.shape[1:-1]))
        out = torch.zeros((N, B, *G), dtype=complex_dtype, device=ksp.device)
        scale = float(self.oversamp ** (d / 2))
        for i in range(N):
            y = ksp[i].reshape(B, M).to(complex_dtype).contiguous()
            idx = self._flat_index(trj[i], G)
            _lib.check(_lib.lib().mra_nn_scatter(_lib.ptr(y), _lib.ptr(idx), M, B, scale, _lib.ptr(out[i]),
                                                 int(math.prod(G)), _lib.stream_ptr(y.device)))
        return out.reshape(N, *batch, *G)

    def _fused(self) -> bool:
        return all(g >= n for g, n in zip(self.im_size_os, self.im_size)) and ops.fused_supported(self.im_size_os)

    def forward(self, img: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        """forward_interp_only(forward_FT_only(img)); fused: one pad+FFT kernel chain writes the
        unshifted spectrum and the gather reads it at the shifted index."""
        if not self._fused():
            return self.forward_interp_only(self.forward_FT_only(img), trj)
        _lib.require_device(img, trj)
        d = len(self.im_size)
        G = self.im_size_os
        N = trj.shape[0]
        batch = tuple(img.shape[1:-d])
        B = int(math.prod(batch))
        M = int(math.prod(trj.shape[1:-1]))
        out = torch.empty((N, B, M), dtype=complex_dtype, device=img.device)
        for n in range(N):
            x = img[n if img.shape[0] > 1 else 0].reshape(B, *self.im_size).to(complex_dtype)
            g = ops.pad_fft(x, self.im_size, G, scale=1.0 / math.sqrt(math.prod(G)))
            idx = self._flat_index(trj[n], G, unshifted=True)
            _lib.check(_lib.lib().mra_nn_gather(_lib.ptr(g), int(math.prod(G)), _lib.ptr(idx), M, B,
                                                float(self.oversamp ** (d / 2)), _lib.ptr(out[n]),
                                                _lib.stream_ptr(g.device)))
        return out.reshape(N, *batch, *trj.shape[1:-1])

    def adjoint(self, ksp: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        """adjoint_iFT_only(adjoint_grid_only(ksp)); fused: scatter into the unshifted layout,
        then one iFFT+crop kernel chain."""
        if not self._fused():
            return self.adjoint_iFT_only(self.adjoint_grid_only(ksp, trj))
        _lib.require_device(ksp, trj)
        d = len(self.im_size)
        G = self.im_size_os
        N = trj.shape[0]
        nt = trj.ndim - 2
        batch = tuple(ksp.shape[1:ksp.ndim - nt])
        B = int(math.prod(batch))
        M = int(math.prod(trj.shape[1:-1]))
        out = torch.empty((N, B, *self.im_size), dtype=complex_dtype, device=ksp.device)
        for n in range(N):
            y = ksp[n].reshape(B, M).to(complex_dtype).contiguous()
            perm, cstart, cend, cidx = self._scatter_tables(trj[n], G)
            g = torch.zeros((B, *G), dtype=complex_dtype, device=ksp.device)
            _lib.check(_lib.lib().mra_nn_scatter_chunks(_lib.ptr(y), _lib.ptr(perm), _lib.ptr(cstart),
                                                        _lib.ptr(cend), _lib.ptr(cidx), cidx.numel(), M, B,
                                                        float(self.oversamp ** (d / 2)), _lib.ptr(g),
                                                        int(math.prod(G)), _lib.stream_ptr(y.device)))
            ops.ifft_crop(g, self.im_size, G, B, scale=1.0 / math.sqrt(math.prod(G)), out=out[n])
        return out.reshape(N, *batch, *self.im_size)

    def calc_teoplitz_kernels(self, trj: torch.Tensor, weights: Optional[torch.Tensor] = None,
                              os_factor: Optional[float] = None):
        """PSF kernels on the oversampled grid (fourier.py:1296-1326; os_factor unused)."""
        os_factor = self.oversamp
        im_size_os = tuple(round(i * os_factor) for i in self.im_size)
        nufft_os = gridded_nufft(im_size_os, oversamp=1.0)
        return calc_toep_kernel_helper(nufft_os.adjoint, trj.type(torch.int32), weights) * (
            os_factor ** len(self.im_size))


class _LowRankNNUFFT(NUFFT):
    """Nearest-grid transforms with a low-rank correction of the grid-deviation phase:
    forward = s * sum_l h_l * multi_index(FFT(pad(b_l x))), adjoint = s * sum_l conj(b_l)
    crop(iFFT(multi_grid(conj(h_l) y))), with temporal functions ``h`` (L, *trj_batch),
    spatial functions ``b`` (L, *im_size) and s = ``grid_oversamp`` (the common form of
    chebyshev_nufft and svd_nufft, fourier.py:1487-1572, 1703-1774). The L spectra of one image
    come from one fused pad+FFT launch over L maps (b); the weighted gather and the chunked
    weighted scatter are ``mra_nn_gather_sum`` / ``mra_nn_scatter_sum_chunks``; the adjoint's L
    inverse transforms and the sum with conj(b) are one fused iFFT+crop launch. All L terms
    are applied in one call (``n_batch_size`` is kept for interface parity)."""

    def _padder(self) -> PadLast:
        raise NotImplementedError

    def _index(self, trj_n: torch.Tensor, G):
        return self._tables.get(trj_n, ("idx", tuple(G)),
                                lambda: gridded_nufft._build_flat_index(trj_n, G, unshifted=True))

    def _scatter_tables(self, trj_n: torch.Tensor, G):
        return self._tables.get(trj_n, ("chunks", tuple(G)),
                                lambda: gridded_nufft._chunks(self._index(trj_n, G)))

    def _check(self, trj):
        if getattr(self, "h", None) is None:
            raise RuntimeError(f"{type(self).__name__}: call rescale_trajectory(trj) first (it builds the "
                               "basis functions)")
        M = int(math.prod(trj.shape[1:-1]))
        if self.h.shape[0] != self.b.shape[0] or int(math.prod(self.h.shape[1:])) != M:
            raise ValueError(f"{type(self).__name__}: trajectory does not match the one rescale_trajectory was given")

    def forward(self, img: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        """img (N, *batch, *im_size), trj (N, *trj_batch, d) int32 from rescale_trajectory."""
        _lib.require_device(img, trj)
        self._check(trj)
        d = len(self.im_size)
        pad = self._padder()
        G = pad.pad_im_size
        N = trj.shape[0]
        batch = tuple(img.shape[1:-d])
        B = int(math.prod(batch))
        L = self.h.shape[0]
        M = int(math.prod(trj.shape[1:-1]))
        out = torch.empty((N, B, M), dtype=complex_dtype, device=img.device)
        for n in range(N):
            x = img[n if img.shape[0] > 1 else 0].reshape(B, *self.im_size).to(complex_dtype)
            g = ops.pad_fft(x, self.im_size, G, smaps=self.b, scale=1.0 / math.sqrt(math.prod(G)), center=pad.center)
            idx = self._index(trj[n], G)
            _lib.check(_lib.lib().mra_nn_gather_sum(_lib.ptr(g), int(math.prod(G)), _lib.ptr(idx), _lib.ptr(self.h),
                                                    L, M, B, float(self.grid_oversamp), _lib.ptr(out[n]),
                                                    _lib.stream_ptr(g.device)))
        return out.reshape(N, *batch, *trj.shape[1:-1])

    def adjoint(self, ksp: torch.Tensor, trj: torch.Tensor) -> torch.Tensor:
        """ksp (N, *batch, *trj_batch), trj (N, *trj_batch, d) int32 -> (N, *batch, *im_size)."""
        _lib.require_device(ksp, trj)
        self._check(trj)
        d = len(self.im_size)
        pad = self._padder()
        G = pad.pad_im_size
        N = trj.shape[0]
        nt = trj.ndim - 2
        batch = tuple(ksp.shape[1:ksp.ndim - nt])
        B = int(math.prod(batch))
        L = self.h.shape[0]
        M = int(math.prod(trj.shape[1:-1]))
        out = torch.empty((N, B, *self.im_size), dtype=complex_dtype, device=ksp.device)
        for n in range(N):
            y = ksp[n].reshape(B, M).to(complex_dtype).contiguous()
            perm, cstart, cend, cidx = self._scatter_tables(trj[n], G)
            g = torch.zeros((B * L, *G), dtype=complex_dtype, device=ksp.device)
            _lib.check(_lib.lib().mra_nn_scatter_sum_chunks(_lib.ptr(y), _lib.ptr(self.h), L, _lib.ptr(perm),
                                                            _lib.ptr(cstart), _lib.ptr(cend), _lib.ptr(cidx),
                                                            cidx.numel(), M, B, 1.0, _lib.ptr(g),
                                                            int(math.prod(G)), _lib.stream_ptr(y.device)))
            ops.ifft_crop(g, self.im_size, G, B, smaps=self.b, scale=self.grid_oversamp / math.sqrt(math.prod(G)),
                          out=out[n], center=pad.center)
        return out.reshape(N, *batch, *self.im_size)


class chebyshev_nufft(_LowRankNNUFFT):
    """NUFFT from a Chebyshev expansion of the grid deviations (fourier.py:1574-1774; after
    arXiv:1701.04492, Lemma A.3 of Townsend's thesis): each sample is rounded to its nearest
    point of a ``round(grid_oversamp * N)`` grid and the residual phase
    exp(-2 pi i (k - k_grid) r) is expanded as sum_l h_l(k) b_l(r), l over
    ``n_cheby_per_dim``^d products of one-dimensional terms. Forward:
    sum_l h_l * multi_index(FFT(pad(b_l x))); adjoint: sum_l conj(b_l) crop(iFFT(multi_grid(
    conj(h_l) y))); both times ``grid_oversamp``.

    ``rescale_trajectory`` computes h for the trajectory it is given and b for the image
    (as the reference does) and returns the clamped grid indices the other methods take.
    """

    def __init__(self, im_size: tuple, n_cheby_per_dim: Optional[int] = 5, grid_oversamp: Optional[float] = 1.0,
                 n_batch_size: Optional[int] = None):
        super().__init__(im_size)
        grd_os_size = tuple(round(i * grid_oversamp) for i in self.im_size)
        self.grid_oversamp = grid_oversamp
        self.grog_padder = PadLast(grd_os_size, im_size)
        self.gamma = 1 / (2 * grid_oversamp)
        self.n_cheby_per_dim = n_cheby_per_dim
        self.n_batch_size = n_cheby_per_dim ** len(im_size) if n_batch_size is None else n_batch_size
        self._tables = _PlanCache()

    @staticmethod
    def cheby_weights(p, r, gamma):
        """Coefficient of T_r in the expansion of the p-th term (fourier.py:1601-1608):
        4 i^r J_{(r+p)/2}(-gamma pi / 2) J_{(r-p)/2}(-gamma pi / 2) when r - p is even, else 0."""
        from scipy.special import jv

        if (abs(p - r) % 2) == 0:
            return 4 * (1j ** r) * jv((r + p) / 2, -gamma * np.pi / 2) * jv((r - p) / 2, -gamma * np.pi / 2)
        return 0

    def compute_spatial_funcs(self, im_size: tuple) -> torch.Tensor:
        """b[i, p](r) = sum_q s_p s_q T_q(2 r_i) c(p, q), s_0 = 1/2, else 1 (fourier.py:1610-1639);
        (d, n, *im_size) complex64 on the CPU."""
        n = self.n_cheby_per_dim
        d = len(im_size)
        grd = _gen_grd(im_size)
        b = torch.zeros((d, n, *im_size), dtype=complex_dtype)
        for i in range(d):
            r = grd[..., i]
            for lp in range(n):
                sp_ = 0.5 if lp == 0 else 1.0
                for lq in range(n):
                    sq = 0.5 if lq == 0 else 1.0
                    b[i, lp] += torch.cos(lq * torch.arccos(2 * r)) * self.cheby_weights(lp, lq, self.gamma) * (sq * sp_)
        return b

    def compute_temporal_funcs(self, trj: torch.Tensor) -> torch.Tensor:
        """h[i, p](k) = T_p((k_i - round(k_i s) / s) / gamma) (fourier.py:1641-1664); (d, n, *trj)
        complex64 on the trajectory's device. The normalised deviation is clamped to [-1, 1]
        before arccos: mathematically it lies there, but for grid_oversamp != 1 fp32 rounding can
        put it an ulp outside, where the reference's arccos returns NaN."""
        n = self.n_cheby_per_dim
        d = trj.shape[-1]
        dev = trj.device
        trj_dev = trj - torch.round(trj * self.grid_oversamp) / self.grid_oversamp
        h = torch.zeros((d, n, *trj.shape[:-1]), dtype=complex_dtype, device=dev)
        for i in range(d):
            k = trj_dev[..., i]
            for lp in range(n):
                h[i, lp] = torch.cos(lp * torch.arccos(torch.clamp(k / self.gamma, -1.0, 1.0)))
        return h

    def rescale_trajectory(self, trj: torch.Tensor) -> torch.Tensor:
        """Clamped nearest grid indices (fourier.py:1666-1693); also sets the L temporal
        functions ``h`` (L, *trj_batch) and spatial functions ``b`` (L, *im_size)."""
        trj_rs = trj * self.grid_oversamp
        G = self.grog_padder.pad_im_size
        for i in range(trj_rs.shape[-1]):
            trj_rs[..., i] = torch.clamp(trj_rs[..., i] + G[i] / 2, 0, G[i] - 1)
        trj_rs = torch.round(trj_rs).type(torch.int32)
        h = self.compute_temporal_funcs(trj)
        b = self.compute_spatial_funcs(self.im_size).to(trj.device)
        d = trj.shape[-1]
        h = _outer_terms([h[i] for i in range(d)]).reshape(-1, *trj.shape[:-1])
        b = _outer_terms([b[i] for i in range(d)]).reshape(-1, *self.im_size)
        self.h = h.contiguous()
        self.b = b.contiguous()
        return trj_rs

    def _padder(self) -> PadLast:
        return self.grog_padder


class svd_nufft(_LowRankNNUFFT):
    """NUFFT from a low-rank SVD model of the grid deviations (fourier.py:1328-1572): samples
    are rounded to the nearest point of a ``round(grid_oversamp * N)`` grid (wrapping) and the
    residual phase exp(-2 pi i dk r) is approximated by the ``n_svd`` leading singular pairs of
    the matrix exp(-2 pi i k_m r_n) on the ``svd_mx_size`` pixel grid, interpolated to the
    samples and pixels with cubic splines.

    The reference finds the eigenvectors of mx^H mx by LOBPCG on the full prod(svd_mx_size)^2
    matrix. That matrix is the Kronecker product of one-dimensional matrices, so here each
    dimension's eigenpairs come from a small FP64 eigendecomposition and the n_svd largest
    products are taken (same eigenspaces; a 3D 50^3 model needs no 125 000^2 matrix). When the
    n_svd-th eigenvalue is degenerate with the next one the reference's operator depends on
    LOBPCG's random start; here ties are taken in dimension order.
    """

    def __init__(self, im_size: tuple, grid_oversamp: Optional[float] = 1.0, n_svd: Optional[int] = 5,
                 n_batch_size: Optional[int] = None, svd_mx_size: Optional[tuple] = None):
        super().__init__(im_size)
        os_size = tuple(round(i * grid_oversamp) for i in self.im_size)
        self.grid_oversamp = grid_oversamp
        self.padder = PadLast(os_size, im_size)
        self.n_svd = n_svd
        self.n_batch_size = n_svd if n_batch_size is None else n_batch_size
        self.svd_mx_size = (50,) * len(im_size) if svd_mx_size is None else tuple(svd_mx_size)
        assert len(self.svd_mx_size) == len(im_size), "svd_mx_size must match im_size in length"
        self._tables = _PlanCache()

    def _padder(self) -> PadLast:
        return self.padder

    @staticmethod
    def spatial_interp(spatial_input: torch.Tensor, coords: torch.Tensor, order: Optional[int] = 3,
                       mode: Optional[str] = "nearest") -> torch.Tensor:
        """Cubic-spline interpolation of each of the N grids of ``spatial_input`` (N, d0..dK-1)
        at ``coords`` (*crds, K) in index units -> (N, *crds) (fourier.py:1351-1411, which calls
        cupyx map_coordinates(order=3, mode='nearest')). The spline coefficients are
        scipy.ndimage.spline_filter of the input padded by 12 edge values (what map_coordinates
        does for mode 'nearest'), computed in FP64 on the host; the 4^K-tap evaluation at the
        coordinates runs on the coordinates' device."""
        import numpy as _np
        from scipy.ndimage import spline_filter

        if order != 3 or mode != "nearest":
            raise NotImplementedError("spatial_interp implements order=3, mode='nearest' (the reference's call)")
        K = coords.shape[-1]
        if spatial_input.ndim != K + 1:
            raise ValueError("Input tensor must have K spatial dimensions.")
        npad = 12
        a = spatial_input.detach().cpu().numpy().astype(_np.complex128)
        coef = _np.empty((a.shape[0], *[n + 2 * npad for n in a.shape[1:]]), dtype=_np.complex128)
        for i in range(a.shape[0]):
            re = spline_filter(_np.pad(a[i].real, npad, mode="edge"), 3, output=_np.float64, mode="nearest")
            im = spline_filter(_np.pad(a[i].imag, npad, mode="edge"), 3, output=_np.float64, mode="nearest")
            coef[i] = re + 1j * im
        dev = coords.device
        P = coef.shape[1:]
        cf = torch.from_numpy(coef.reshape(coef.shape[0], -1)).to(dev, torch.complex128)
        x = coords.reshape(-1, K).to(torch.float64) + npad
        start = torch.floor(x).to(torch.int64) - 1
        y = x - torch.floor(x)
        w = torch.stack([(1 - y) ** 3 / 6, (3 * y ** 3 - 6 * y ** 2 + 4) / 6,
                         (-3 * y ** 3 + 3 * y ** 2 + 3 * y + 1) / 6, y ** 3 / 6], dim=-1)  # (npts, K, 4)
        strides = [math.prod(P[i + 1:]) for i in range(K)]
        out = torch.zeros((cf.shape[0], x.shape[0]), dtype=torch.complex128, device=dev)
        for taps in np.ndindex(*(4,) * K):
            flat = torch.zeros(x.shape[0], dtype=torch.int64, device=dev)
            wt = torch.ones(x.shape[0], dtype=torch.float64, device=dev)
            for i, j in enumerate(taps):
                flat += torch.clamp(start[:, i] + j, 0, P[i] - 1) * strides[i]
                wt = wt * w[:, i, j]
            out += cf[:, flat] * wt
        return out.reshape(cf.shape[0], *coords.shape[:-1]).to(complex_dtype)

    def compute_svd_funcs(self, trj: torch.Tensor, batch_size: Optional[int] = None):
        """(temporal functions (n_svd, *trj_batch), conj of the spatial functions (n_svd, *im_size))
        (fourier.py:1413-1470)."""
        d = len(self.im_size)
        ss = self.svd_mx_size
        lams, vecs, mxs = [], [], []
        for n in ss:
            r = (np.arange(n) - n // 2) / n
            m = np.exp(-2j * np.pi * np.outer(r, r))
            lam, vec = np.linalg.eigh(m.conj().T @ m)
            lams.append(lam[::-1])
            vecs.append(vec[:, ::-1])
            mxs.append(m)
        keep = [min(n, self.n_svd) for n in ss]
        cand = sorted(np.ndindex(*keep), key=lambda t: -math.prod(lams[i][t[i]] for i in range(d)))[: self.n_svd]
        ev = np.stack([_kron_vectors([vecs[i][:, t[i]] for i in range(d)]) for t in cand]).reshape(-1, *ss)
        tv = np.stack([_kron_vectors([mxs[i] @ vecs[i][:, t[i]] for i in range(d)]) for t in cand]).reshape(-1, *ss)
        dev = trj.device
        sz = torch.tensor(ss, dtype=torch.float64, device=dev)
        spatial = self.spatial_interp(torch.from_numpy(ev), (_gen_grd(self.im_size).to(dev, torch.float64) + 0.5) * sz)
        trj_dev = trj - (trj * self.grid_oversamp).round() / self.grid_oversamp
        temporal = self.spatial_interp(torch.from_numpy(tv), (0.5 + trj_dev.to(torch.float64) * self.grid_oversamp) * sz)
        return temporal, spatial.conj()

    def rescale_trajectory(self, trj: torch.Tensor) -> torch.Tensor:
        """Basis functions for this trajectory, then round(k s + G/2) mod G (fourier.py:1472-1485)."""
        temporal, spatial = self.compute_svd_funcs(trj)
        self.temporal_funcs, self.spatial_funcs = temporal, spatial
        self.h = temporal.contiguous()
        self.b = spatial.contiguous()
        trj_rs = trj * self.grid_oversamp
        G = self.padder.pad_im_size
        for i in range(trj_rs.shape[-1]):
            trj_rs[..., i] = (trj_rs[..., i] + G[i] / 2).round() % G[i]
        return torch.round(trj_rs).type(torch.int32)


def _kron_vectors(vs):
    out = vs[0]
    for v in vs[1:]:
        out = np.multiply.outer(out, v)
    return out.reshape(-1)


def _gen_grd(im_size) -> torch.Tensor:
    """Pixel grid r_i = (n_i - N_i//2) / N_i, shape (*im_size, d) (utils.py:166-200)."""
    lins = [torch.arange(-(n // 2), n // 2 + (n % 2)) / n for n in im_size]
    return torch.stack(torch.meshgrid(*lins, indexing="ij"), dim=-1).to(real_dtype)


def _outer_terms(fs):
    """Products over dims of per-dim terms: fs[i] (n_i, *S) -> (n_0, ..., n_{d-1}, *S)."""
    out = fs[0]
    for f in fs[1:]:
        out = out[:, None] * f[None]
        out = out.reshape(-1, *f.shape[1:])
    return out.reshape(*[f.shape[0] for f in fs], *fs[0].shape[1:])


__all__ = ["fft", "ifft", "calc_toep_kernel_helper", "NUFFT", "hip_nufft", "triton_nufft", "kb_beta",
           "sigpy_nufft", "sigpy_beta", "gridded_nufft", "chebyshev_nufft", "svd_nufft",
           "toeplitz_unshifted", "batch_iterator"]
