"""Reference NUFFT (``triton_nufft``) restated on CPU torch.

Follows fourier.py:354-582 step by step (apodize in three in-place passes,
divide by sqrt(prod N), centred ``resize`` pad, centred FFT with norm=None,
``_scale_trj``, per-N-batch interpolation, divide by w^d; and the mirror for the
adjoint), ``calc_toep_kernel_helper`` (fourier.py:63-105), ``normal_toeplitz``
(fourier.py:302-352) with ``PadLast`` (pad.py:5-37), ``fft``/``ifft``
(fourier.py:37-61) and ``resize`` (utils.py:41-84).

``dtype=torch.float64`` runs the same algorithm in double precision.
"""

from __future__ import annotations

import math
from math import ceil

import numpy as np
import torch
import torch.fft as tfft
import torch.nn.functional as F

from . import interp_ref, sigpy_ref


def fft_c(x, dim, norm="ortho"):
    return tfft.fftshift(tfft.fftn(tfft.ifftshift(x, dim=dim), dim=dim, norm=norm), dim=dim)


def ifft_c(x, dim, norm="ortho"):
    return tfft.fftshift(tfft.ifftn(tfft.ifftshift(x, dim=dim), dim=dim, norm=norm), dim=dim)


def resize(x, oshape):
    """utils.py:41-84: centred pad/crop, ishift = max(i//2 - o//2, 0), oshift = max(o//2 - i//2, 0)."""
    ish, osh = list(x.shape), list(oshape)
    if ish == osh:
        return x.reshape(oshape)
    isf = [max(i // 2 - o // 2, 0) for i, o in zip(ish, osh)]
    osf = [max(o // 2 - i // 2, 0) for i, o in zip(ish, osh)]
    cp = [min(i - si, o - so) for i, si, o, so in zip(ish, isf, osh, osf)]
    out = torch.zeros(osh, dtype=x.dtype)
    out[tuple(slice(s, s + c) for s, c in zip(osf, cp))] = x[tuple(slice(s, s + c) for s, c in zip(isf, cp))]
    return out


class PadLastRef:
    """pad.py:5-37 (symmetric (P-N)//2 zero pad of the trailing dims, and the crop)."""

    def __init__(self, pad_im_size, im_size):
        self.im_size, self.pad_im_size = tuple(im_size), tuple(pad_im_size)
        self.q = [(p - n) // 2 for p, n in zip(pad_im_size, im_size)]

    def forward(self, x):
        pad = []
        for q in reversed(self.q):
            pad += [q, q]
        return F.pad(x, pad)

    def adjoint(self, y):
        d = len(self.im_size)
        sl = tuple(slice(q, -q) if q > 0 else slice(None) for q in self.q)
        return y[(slice(None),) * (y.ndim - d) + sl]


class TritonNufftRef:
    """fourier.py:354-582 on CPU."""

    def __init__(self, im_size, oversamp=1.25, width=6, beta=None, apodize=True, dtype=torch.float32,
                 coords64=False, safe=False):
        self.im_size = tuple(im_size)
        self.oversamp = oversamp
        self.width = width
        self.apodize = apodize
        self.beta = np.pi * (((width / oversamp) * (oversamp - 0.5)) ** 2 - 0.8) ** 0.5 if beta is None else beta
        self.dtype = dtype
        self.cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128
        # coords64: scale and wrap trajectories in float64 instead of the reference's
        # fp32 (finite differences need it); safe: autograd-safe KB masking.
        self.coords64 = coords64
        self.safe = safe

    def _apodization_func(self, x):
        apod = (self.beta ** 2 - (np.pi * self.width * x) ** 2) ** 0.5
        return apod / torch.sinh(apod)

    def apodize_img(self, img):
        nd = len(self.im_size)
        for i in range(-nd, 0):
            crds = (torch.arange(img.shape[i]) - img.shape[i] // 2).to(self.dtype)
            tup = (slice(None),) + (None,) * (-i - 1)
            img *= self._apodization_func(crds / ceil(self.oversamp * img.shape[i]))[tup]
        return img

    def _scale_trj(self, trj, shape):
        out = trj.clone().to(torch.float64 if self.coords64 else torch.float32)
        nd = trj.shape[-1]
        for i in range(-nd, 0):
            out[..., i] *= ceil(self.oversamp * shape[i]) / shape[i]
            out[..., i] += ceil(self.oversamp * shape[i]) // 2
        return out

    def forward_FT_only(self, img):
        nd = len(self.im_size)
        os_shape = list(img.shape)[:-nd] + [ceil(self.oversamp * i) for i in img.shape[-nd:]]
        x = img.clone().to(self.cdtype)
        if self.apodize:
            x = self.apodize_img(x)
        x /= np.prod(x.shape[-nd:]) ** 0.5
        x = resize(x, os_shape)
        return fft_c(x, dim=tuple(range(-nd, 0)), norm=None)

    def forward_interp_only(self, ksp_os, trj):
        nd = len(self.im_size)
        img_shape = (*ksp_os.shape[:-nd], *self.im_size)
        N = trj.shape[0]
        t = self._scale_trj(trj, img_shape)
        outs = []
        for i in range(N):
            outs.append(interp_ref.ungrid(ksp_os[i], t[i], (self.width,) * nd, beta=self.beta, dtype=self.dtype,
                                       coords64=self.coords64, safe=self.safe))
        return torch.stack(outs, 0) / self.width ** nd

    def adjoint_iFT_only(self, ksp_os):
        nd = len(self.im_size)
        oshape = (*ksp_os.shape[:-nd], *self.im_size)
        os_shape = list(oshape)[:-nd] + [ceil(self.oversamp * i) for i in oshape[-nd:]]
        img_os = ifft_c(ksp_os, dim=tuple(range(-nd, 0)), norm=None)
        out = resize(img_os, oshape)
        out = out * (np.prod(os_shape[-nd:]) / np.prod(oshape[-nd:]) ** 0.5)
        if self.apodize:
            out = self.apodize_img(out)
        return out

    def adjoint_grid_only(self, ksp, trj):
        nd = trj.shape[-1]
        N = trj.shape[0]
        oshape = (N, *ksp.shape[1:-(trj.ndim - 2)], *self.im_size)
        os_shape = list(oshape)[:-nd] + [ceil(self.oversamp * i) for i in oshape[-nd:]]
        t = self._scale_trj(trj, oshape)
        outs = []
        for i in range(N):
            outs.append(interp_ref.grid(ksp[i], t[i], tuple(os_shape[-nd:]), (self.width,) * nd, beta=self.beta,
                                     dtype=self.dtype, coords64=self.coords64, safe=self.safe))
        return torch.stack(outs, 0).reshape(os_shape) / self.width ** nd

    def forward(self, img, trj):
        return self.forward_interp_only(self.forward_FT_only(img), trj)

    def adjoint(self, ksp, trj):
        return self.adjoint_iFT_only(self.adjoint_grid_only(ksp, trj))

    def calc_teoplitz_kernels(self, trj, weights=None, os_factor=2.0):
        im_size_os = tuple(round(i * os_factor) for i in self.im_size)
        nos = TritonNufftRef(im_size_os, self.oversamp, self.width, dtype=self.dtype)
        return calc_toep_kernel_helper(nos.adjoint, trj * os_factor, weights, self.cdtype) * (
            os_factor ** len(self.im_size))

    def normal_toeplitz(self, img, kerns):
        """fourier.py:302-352."""
        N = img.shape[0]
        nd = len(self.im_size)
        im_os = kerns.shape[1:]
        flat = img.reshape((N, -1, *self.im_size)).to(self.cdtype)
        pad = PadLastRef(im_os, self.im_size)
        out = torch.zeros_like(flat)
        dims = tuple(range(-nd, 0))
        for n in range(N):
            for i in range(flat.shape[1]):
                f = pad.forward(flat[n:n + 1, i:i + 1])
                f = fft_c(f, dims) * kerns[n:n + 1, None]
                out[n:n + 1, i:i + 1] = pad.adjoint(ifft_c(f, dims))
        return out.reshape(img.shape)


class SigpyNufftRef(TritonNufftRef):
    """fourier.py:728-1033 on CPU: the transform of TritonNufftRef with sigpy's
    interpolate / gridding (``sigpy_ref``; fourier.py:901-963), sigpy_nufft's beta default (1.0 when the formula's radicand is negative,
    :740-747), its apodization (sinh + 1e-12, continued by sqrt(-a)/sin(sqrt(-a)),
    :817-841) and Toeplitz kernels from a default-beta instance (:1007-1033)."""

    def __init__(self, im_size, oversamp=1.25, width=6, beta=None, apodize=True, dtype=torch.float32):
        if beta is None and ((width / oversamp) * (oversamp - 0.5)) ** 2 - 0.8 < 0:
            beta = 1.0
        super().__init__(im_size, oversamp, width, beta=beta, apodize=apodize, dtype=dtype)

    def forward_interp_only(self, ksp_os, trj):
        """fourier.py:901-933: sigpy's interpolate per trajectory, / width^d."""
        nd = len(self.im_size)
        t = self._scale_trj(trj, (*ksp_os.shape[:-nd], *self.im_size))
        outs = []
        for i in range(trj.shape[0]):
            g = ksp_os[i]
            bshape = g.shape[:-nd]
            y = sigpy_ref.interpolate(g.reshape(-1, *g.shape[-nd:]), t[i].reshape(-1, nd), self.width, self.beta,
                                      dtype=self.dtype)
            outs.append(y.reshape(*bshape, *t[i].shape[:-1]))
        return torch.stack(outs, 0) / self.width ** nd

    def adjoint_grid_only(self, ksp, trj):
        """fourier.py:935-963: sigpy's gridding per trajectory, / width^d."""
        nd = trj.shape[-1]
        N = trj.shape[0]
        oshape = (N, *ksp.shape[1:-(trj.ndim - 2)], *self.im_size)
        os_shape = list(oshape)[:-nd] + [ceil(self.oversamp * i) for i in oshape[-nd:]]
        t = self._scale_trj(trj, oshape)
        M = math.prod(trj.shape[1:-1])
        outs = []
        for i in range(N):
            y = ksp[i].reshape(-1, M)
            outs.append(sigpy_ref.gridding(y, t[i].reshape(-1, nd), tuple(os_shape[-nd:]), self.width, self.beta,
                                           dtype=self.dtype))
        return torch.stack(outs, 0).reshape(os_shape) / self.width ** nd

    def _apodization_func(self, x):
        eps = 1e-12
        arg = self.beta ** 2 - (np.pi * self.width * x) ** 2
        pos = arg.clamp(min=0).sqrt()
        pos = pos / (torch.sinh(pos) + eps)
        neg = (-arg.clamp(max=0)).sqrt()
        neg = neg / (torch.sin(neg) + eps)
        return pos + neg

    def calc_teoplitz_kernels(self, trj, weights=None, os_factor=2.0):
        im_size_os = tuple(round(i * os_factor) for i in self.im_size)
        nos = SigpyNufftRef(im_size_os, self.oversamp, self.width, dtype=self.dtype)
        return calc_toep_kernel_helper(nos.adjoint, trj * os_factor, weights, self.cdtype) * (
            os_factor ** len(self.im_size))


class GriddedNufftRef:
    """fourier.py:1213-1326 on CPU: rescale to round(k s + G/2) mod G, centred resize and
    ortho FFT, then torch advanced indexing (multi_index, indexing.py:29-51) or
    index_add_ (multi_grid, indexing.py:53-79), times s^(d/2)."""

    def __init__(self, im_size, oversamp=1.0, dtype=torch.float32):
        self.im_size = tuple(im_size)
        self.im_size_os = tuple(round(i * oversamp) for i in self.im_size)
        self.oversamp = oversamp
        self.cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128

    def rescale_trajectory(self, trj):
        t = trj * self.oversamp
        for i in range(t.shape[-1]):
            t[..., i] = (t[..., i] + self.im_size_os[i] / 2).round() % self.im_size_os[i]
        return t.type(torch.int32)

    def forward(self, img, trj):
        d = len(self.im_size)
        g = fft_c(resize(img.to(self.cdtype), tuple(img.shape[:-d]) + self.im_size_os), tuple(range(-d, 0)))
        outs = []
        for i in range(trj.shape[0]):
            t = trj[i].long()
            outs.append(g[i][(Ellipsis,) + tuple(t[..., j] for j in range(d))])
        return torch.stack(outs, 0) * self.oversamp ** (d / 2)

    def adjoint(self, ksp, trj):
        d = trj.shape[-1]
        G = self.im_size_os
        nt = trj.ndim - 2
        out = torch.zeros((trj.shape[0], *ksp.shape[1:ksp.ndim - nt], *G), dtype=self.cdtype)
        for i in range(trj.shape[0]):
            t = trj[i].long().reshape(-1, d)
            flat = torch.zeros(t.shape[0], dtype=torch.int64)
            for j in range(d):
                flat = flat * G[j] + t[:, j]
            y = ksp[i].to(self.cdtype).reshape(-1, t.shape[0])
            o = torch.zeros((y.shape[0], math.prod(G)), dtype=self.cdtype)
            o.index_add_(1, flat, y)
            out[i] = o.reshape(out.shape[1:])
        g = out * self.oversamp ** (d / 2)
        img = ifft_c(g, tuple(range(-d, 0)))
        return resize(img, tuple(img.shape[:-d]) + self.im_size)


def calc_toep_kernel_helper(nufft_adj_os, trj, weights=None, cdtype=torch.complex64):
    """fourier.py:63-105."""
    if weights is None:
        weights = torch.ones(trj.shape[:-1], dtype=torch.float32)
    trj_batch = trj.shape[1:-1]
    if trj.shape[0] == 1 and weights.shape[0] != 1:
        trj = trj.expand((weights.shape[0], *trj.shape[1:]))
    N = trj.shape[0]
    d = trj.shape[-1]
    ksp = torch.ones((N, 1, *trj_batch), dtype=cdtype) * weights[:, None, ...]
    img = nufft_adj_os(ksp, trj)[:, 0, ...]
    return fft_c(img, dim=tuple(range(-d, 0)))
