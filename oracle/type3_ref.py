"""Type-3 NUFFT -- TEST INFRASTRUCTURE ONLY (part of the oracle).

Restates ``type3_nufft`` (linops.py:123-425): y_k = sum_n x_n exp(-2 pi i phi_n . alpha_k)
by centring and rescaling phi and alpha (:165-188), KB gridding of the image samples
at scaled phi with one beta per dimension (sigpy gridding, ``sigpy_ref``; :226-253),
a ``sigpy_nufft`` (default parameters) of that grid at scaled alpha (:255-257), KB
apodization and the midpoint phases (:259-268); the adjoint mirrors it (:270-330) and
the Toeplitz normal convolves the gridded image with the PSF of the inner NUFFT
(:332-425). ``naive`` is type3_nufft_naive (linops.py:49-121), the exact sum.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import sigpy_ref
from .nufft_ref import SigpyNufftRef, fft_c, ifft_c, resize


def naive_forward(phis, alphas, x):
    """linops.py:72-89 in complex128: x (N, *im) -> (N, *trj)."""
    B = phis.shape[0]
    P = phis.reshape(B, -1).to(torch.float64)
    A = alphas.reshape(B, -1).to(torch.float64)
    enc = torch.exp(-2j * torch.pi * (P.T @ A))
    return (x.reshape(x.shape[0], -1).to(torch.complex128) @ enc).reshape(x.shape[0], *alphas.shape[1:])


def naive_adjoint(phis, alphas, y):
    B = phis.shape[0]
    P = phis.reshape(B, -1).to(torch.float64)
    A = alphas.reshape(B, -1).to(torch.float64)
    enc = torch.exp(2j * torch.pi * (A.T @ P))
    return (y.reshape(y.shape[0], -1).to(torch.complex128) @ enc).reshape(y.shape[0], *phis.shape[1:])


class Type3Ref:
    def __init__(self, phis, alphas, oversamp=2.0, width=4.0, use_toep=False):
        B, *im_size = phis.shape
        _, *trj_size = alphas.shape
        self.ishape, self.oshape = tuple(im_size), tuple(trj_size)
        R, T = math.prod(im_size), math.prod(trj_size)
        pf = phis.reshape(B, R).to(torch.float32)
        af = alphas.reshape(B, T).to(torch.float32)
        p_mp = (pf.min(dim=1).values + pf.max(dim=1).values) / 2
        a_mp = (af.min(dim=1).values + af.max(dim=1).values) / 2
        pc = pf - p_mp[:, None]
        ac = af - a_mp[:, None]
        sc = pc.abs().max(dim=1).values * 2
        pc /= sc[:, None]
        p_mp /= sc
        ac *= sc[:, None]
        a_mp *= sc
        self.phis, self.alphas, self.phis_mp, self.alphas_mp, self.B = pc, ac, p_mp, a_mp, B
        S = ac.abs().max(dim=1).values
        self.W = width
        N_os = torch.ceil(2 * S * oversamp + width).long()
        self.os = N_os / (N_os / oversamp).round()
        self.N = N_os / self.os
        self.gamma = N_os / (2 * self.os * S)
        self.beta = np.pi * (((width / self.os) * (self.os - 0.5)) ** 2 - 0.8) ** 0.5
        self.N_os = tuple((self.N * self.os).ceil().long().tolist())
        self.nft = SigpyNufftRef(self.N_os)
        self.kerns = None
        if use_toep:
            aw = self.apod(ac.T * self.gamma / self.N / self.os).prod(dim=-1)
            self.kerns = self.nft.calc_teoplitz_kernels(ac.T[None] * self.gamma, aw[None] ** 2)

    def apod(self, x):
        a = (self.beta ** 2 - (np.pi * self.W * x) ** 2) ** 0.5
        return a / torch.sinh(a)

    def _img_trj(self):
        scales = (self.os * self.N).ceil() / self.N
        shifts = (self.os * self.N).ceil() // 2
        return scales * (self.phis.T * self.N / self.gamma) + shifts

    def _grid(self, x):
        return sigpy_ref.gridding(x, self._img_trj(), self.N_os, self.W, self.beta.tolist()) / self.W ** self.B

    def _interp(self, g):
        return sigpy_ref.interpolate(g, self._img_trj(), self.W, self.beta.tolist()) / self.W ** self.B

    def forward(self, x):
        n = x.shape[0]
        xm = x.reshape(n, -1).to(torch.complex64) * torch.exp(-2j * torch.pi * (self.phis.T @ self.alphas_mp))
        g = self._grid(xm)
        trj = self.alphas.T[None] * self.gamma
        y = self.nft.forward(g[None], trj)[0] * np.prod(self.N_os) ** 0.5
        y = y * self.apod(self.alphas.T * self.gamma / self.N / self.os).prod(dim=-1)
        y = y * torch.exp(-2j * torch.pi * ((self.alphas.T + self.alphas_mp) @ self.phis_mp))
        return y.reshape(n, *self.oshape)

    def adjoint(self, y):
        n = y.shape[0]
        ym = y.reshape(n, -1).to(torch.complex64) * torch.exp(2j * torch.pi * (self.alphas.T @ self.phis_mp))
        ym = ym * self.apod(self.alphas.T * self.gamma / self.N / self.os).prod(dim=-1)
        trj = self.alphas.T[None] * self.gamma
        g = self.nft.adjoint(ym[None], trj)[0] * np.prod(self.N_os) ** 0.5
        x = self._interp(g)
        x = x * torch.exp(2j * torch.pi * ((self.phis.T + self.phis_mp) @ self.alphas_mp))
        return x.reshape(n, *self.ishape)

    def normal(self, x):
        if self.kerns is None:
            return self.adjoint(self.forward(x))
        n, B = x.shape[0], self.B
        phz = (self.phis.T + self.phis_mp) @ self.alphas_mp
        g = self._grid(x.reshape(n, -1).to(torch.complex64) * torch.exp(-2j * torch.pi * phz))
        zp = resize(g, [n] + [2 * s for s in self.N_os])
        dims = tuple(range(-B, 0))
        g = resize(ifft_c(fft_c(zp, dims) * self.kerns, dims), (n, *self.N_os)) * np.prod(self.N_os)
        out = self._interp(g) * torch.exp(2j * torch.pi * phz)
        return out.reshape(n, *self.ishape)
