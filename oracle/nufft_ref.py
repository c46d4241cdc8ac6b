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


class ChebyshevNufftRef:
    """fourier.py:1574-1774 on CPU: Chebyshev expansion of the grid-deviation phase. Basis
    functions as compute_spatial_funcs / compute_temporal_funcs (:1610-1664) with the Bessel
    coefficients of cheby_weights (:1601-1608, scipy.special.jv); rescale_trajectory clamps
    and rounds (:1666-1693); forward / adjoint use PadLast, centred ortho FFTs and
    multi_index / multi_grid (indexing.py:29-79), times grid_oversamp (:1703-1774)."""

    def __init__(self, im_size, n_cheby_per_dim=5, grid_oversamp=1.0, dtype=torch.float32):
        self.im_size = tuple(im_size)
        self.G = tuple(round(i * grid_oversamp) for i in self.im_size)
        self.grid_oversamp = grid_oversamp
        self.gamma = 1 / (2 * grid_oversamp)
        self.n = n_cheby_per_dim
        self.dtype = dtype
        self.cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128
        self.pad = PadLastRef(self.G, self.im_size)

    @staticmethod
    def cheby_weights(p, r, gamma):
        from scipy.special import jv

        if (abs(p - r) % 2) == 0:
            return 4 * (1j ** r) * jv((r + p) / 2, -gamma * np.pi / 2) * jv((r - p) / 2, -gamma * np.pi / 2)
        return 0

    def spatial_funcs(self):
        d = len(self.im_size)
        lins = [(torch.arange(-(n // 2), n // 2 + (n % 2)) / n).to(self.dtype) for n in self.im_size]
        grd = torch.stack(torch.meshgrid(*lins, indexing="ij"), dim=-1)
        b = torch.zeros((d, self.n, *self.im_size), dtype=self.cdtype)
        for i in range(d):
            r = grd[..., i]
            for lp in range(self.n):
                for lq in range(self.n):
                    s = (0.5 if lp == 0 else 1.0) * (0.5 if lq == 0 else 1.0)
                    b[i, lp] += torch.cos(lq * torch.arccos(2 * r)) * self.cheby_weights(lp, lq, self.gamma) * s
        return b

    def temporal_funcs(self, trj):
        d = trj.shape[-1]
        dev = trj - torch.round(trj * self.grid_oversamp) / self.grid_oversamp
        h = torch.zeros((d, self.n, *trj.shape[:-1]), dtype=self.cdtype)
        for i in range(d):
            for lp in range(self.n):
                # clamped to [-1, 1] as the product does (fp32 rounding can leave it an ulp outside)
                h[i, lp] = torch.cos(lp * torch.arccos(torch.clamp(dev[..., i] / self.gamma, -1.0, 1.0)))
        return h

    @staticmethod
    def _outer(fs):
        out = fs[0]
        for f in fs[1:]:
            out = torch.einsum("a...,b...->ab...", out, f).reshape(-1, *f.shape[1:])
        return out

    def rescale_trajectory(self, trj):
        t = trj * self.grid_oversamp
        for i in range(t.shape[-1]):
            t[..., i] = torch.clamp(t[..., i] + self.G[i] / 2, 0, self.G[i] - 1)
        t = torch.round(t).type(torch.int32)
        d = trj.shape[-1]
        h, b = self.temporal_funcs(trj), self.spatial_funcs()
        self.h = self._outer([h[i] for i in range(d)]).reshape(-1, *trj.shape[:-1])
        self.b = self._outer([b[i] for i in range(d)]).reshape(-1, *self.im_size)
        return t

    def forward(self, img, trj):
        d = len(self.im_size)
        xb = img.to(self.cdtype).unsqueeze(-d - 1) * self.b
        ksp_os = fft_c(self.pad.forward(xb), dim=tuple(range(-d, 0)))
        outs = []
        for i in range(trj.shape[0]):
            t = trj[i].long()
            ki = ksp_os[i][(Ellipsis,) + tuple(t[..., j] for j in range(d))]  # (*batch, L, *trj)
            outs.append((ki * self.h).sum(-self.h.ndim))
        return torch.stack(outs, 0) * self.grid_oversamp

    def adjoint(self, ksp, trj):
        d = trj.shape[-1]
        nt = trj.ndim - 2
        outs = []
        for i in range(trj.shape[0]):
            t = trj[i].long().reshape(-1, d)
            flat = torch.zeros(t.shape[0], dtype=torch.int64)
            for j in range(d):
                flat = flat * self.G[j] + t[:, j]
            y = ksp[i].to(self.cdtype).unsqueeze(-nt - 1) * self.h.conj()  # (*batch, L, *trj)
            yf = y.reshape(-1, t.shape[0])
            o = torch.zeros((yf.shape[0], math.prod(self.G)), dtype=self.cdtype)
            o.index_add_(1, flat, yf)
            g = o.reshape(*y.shape[:y.ndim - nt], *self.G)
            img = self.pad.adjoint(ifft_c(g, dim=tuple(range(-d, 0))))
            outs.append((img * self.b.conj()).sum(-self.b.ndim))
        return torch.stack(outs, 0) * self.grid_oversamp


class SvdNufftRef:
    """fourier.py:1328-1572 on CPU: low-rank SVD model of the grid-deviation phase. The matrix
    mx[m, n] = exp(-2 pi i k_m . r_n) on the svd_mx_size pixel grid (k = r, gen_grd), the n_svd
    largest eigenvectors of mx^H mx (here numpy's eigh in FP64; the reference runs LOBPCG),
    temporal functions mx @ evecs; both interpolated with cubic splines, mode 'nearest'
    (scipy.ndimage.map_coordinates, the CPU twin of the reference's cupyx call, :1351-1411) at
    (gen_grd(im_size) + 1/2) * svd_size and (1/2 + (k - round(k s)/s) s) * svd_size
    (:1413-1470); rescale_trajectory wraps round(k s + G/2) mod G (:1472-1485);
    forward / adjoint as :1487-1572."""

    def __init__(self, im_size, grid_oversamp=1.0, n_svd=5, svd_mx_size=None, dtype=torch.float32):
        self.im_size = tuple(im_size)
        self.G = tuple(round(i * grid_oversamp) for i in self.im_size)
        self.grid_oversamp = grid_oversamp
        self.n_svd = n_svd
        self.svd_size = (50,) * len(im_size) if svd_mx_size is None else tuple(svd_mx_size)
        self.cdtype = torch.complex64 if dtype == torch.float32 else torch.complex128
        self.pad = PadLastRef(self.G, self.im_size)

    @staticmethod
    def _grd(size):
        lins = [(np.arange(n) - n // 2) / n for n in size]
        return np.stack(np.meshgrid(*lins, indexing="ij"), axis=-1)

    @staticmethod
    def spatial_interp(inp, coords):
        """inp (N, *dims) complex numpy, coords (..., K) -> (N, ...) (map_coordinates, order 3,
        mode 'nearest', on the real and imaginary parts)."""
        from scipy.ndimage import map_coordinates

        flat = coords.reshape(-1, coords.shape[-1]).T
        out = np.empty((inp.shape[0], flat.shape[1]), dtype=np.complex128)
        for i in range(inp.shape[0]):
            re = map_coordinates(inp[i].real, flat, order=3, mode="nearest")
            im = map_coordinates(inp[i].imag, flat, order=3, mode="nearest")
            out[i] = re + 1j * im
        return out.reshape(inp.shape[0], *coords.shape[:-1])

    def svd_funcs(self, trj):
        ss = self.svd_size
        rs = self._grd(ss).reshape(-1, len(ss))
        mx = np.exp(-2j * np.pi * (rs @ rs.T))
        lam, vec = np.linalg.eigh(mx.conj().T @ mx)
        order = np.argsort(lam)[::-1][: self.n_svd]
        evecs = vec[:, order].T.reshape(self.n_svd, *ss)
        temp = (mx @ vec[:, order]).T.reshape(self.n_svd, *ss)
        self.eigvals = lam[order]
        sz = np.array(ss, dtype=np.float64)
        sp = self.spatial_interp(evecs, (self._grd(self.im_size) + 0.5) * sz)
        t = trj.detach().cpu().numpy().astype(np.float64)
        dev = t - np.round(t * self.grid_oversamp) / self.grid_oversamp
        tp = self.spatial_interp(temp, (0.5 + dev * self.grid_oversamp) * sz)
        return torch.from_numpy(tp).to(self.cdtype), torch.from_numpy(sp).conj().to(self.cdtype)

    def rescale_trajectory(self, trj):
        self.h, self.b = self.svd_funcs(trj)
        t = trj * self.grid_oversamp
        for i in range(t.shape[-1]):
            t[..., i] = (t[..., i] + self.G[i] / 2).round() % self.G[i]
        return torch.round(t).type(torch.int32)

    def forward(self, img, trj):
        d = len(self.im_size)
        xb = img.to(self.cdtype).unsqueeze(-d - 1) * self.b
        ksp_os = fft_c(self.pad.forward(xb), dim=tuple(range(-d, 0)))
        outs = []
        for i in range(trj.shape[0]):
            t = trj[i].long()
            ki = ksp_os[i][(Ellipsis,) + tuple(t[..., j] for j in range(d))]
            outs.append((ki * self.h).sum(-self.h.ndim))
        return torch.stack(outs, 0) * self.grid_oversamp

    adjoint = ChebyshevNufftRef.adjoint
