"""SENSE linop, CG, power method and CG-SENSE driver restated on CPU torch.

  ``SenseRef``              linops.py:1041-1308 (coil batch 1, L = 1 segment)
  ``conjugate_gradient``    algs.py:779-872
  ``power_method_operator`` algs.py:364-399
  ``cg_sense_recon``        recons.py:77-156
"""

from __future__ import annotations

import torch

from .nufft_ref import PadLastRef, fft_c, ifft_c


class SenseRef:
    def __init__(self, trj, mps, dcf=None, nufft=None, use_toeplitz=False):
        self.im_size = tuple(mps.shape[1:])
        self.trj_size = tuple(trj.shape[:-1])
        self.nufft = nufft
        self.cdt = nufft.cdtype
        self.mps = mps.to(self.cdt)
        self.trj = trj.to(torch.float32)
        self.dcf = torch.ones(self.trj_size, dtype=torch.float32) if dcf is None else dcf.to(torch.float32)
        self.use_toeplitz = use_toeplitz
        if use_toeplitz:
            self.toep_kerns = nufft.calc_teoplitz_kernels(self.trj[None], self.dcf[None])[0]
        self.ishape = self.im_size

    def forward(self, img):
        nc = self.mps.shape[0]
        ksp = torch.zeros((nc, *self.trj_size), dtype=self.cdt)
        for c in range(nc):
            sx = self.mps[c:c + 1] * img
            ksp[c:c + 1] += self.nufft.forward(sx[None, :, None], self.trj[None])[0][:, 0]
        return ksp

    def adjoint(self, ksp):
        nc = self.mps.shape[0]
        img = torch.zeros(self.im_size, dtype=self.cdt)
        for c in range(nc):
            y = ksp[c:c + 1] * self.dcf[None]
            f = self.nufft.adjoint(y[None, :, None], self.trj[None])[0][:, 0]
            img += (f * self.mps[c:c + 1].conj()).sum(0)
        return img

    def normal(self, img):
        if not self.use_toeplitz:
            return self.adjoint(self.forward(img))
        d = len(self.im_size)
        pad = PadLastRef(self.toep_kerns.shape[-d:], self.im_size)
        dims = tuple(range(-d, 0))
        out = torch.zeros_like(img, dtype=self.cdt)
        for c in range(self.mps.shape[0]):
            f = fft_c(pad.forward((self.mps[c] * img)[None]), dims)
            f = ifft_c(f * self.toep_kerns, dims)
            out += (pad.adjoint(f)[0] * self.mps[c].conj())
        return out


def conjugate_gradient(AHA, AHb, num_iters=10, lamda_l2=0.0, tolerance=1e-8, return_resids=False):
    """algs.py:779-872 with P = identity and weights = None."""
    def aha_w(x):
        return AHA(x) + lamda_l2 * x

    x0 = AHb.clone()
    if num_iters == 0:
        return x0
    r = AHb - aha_w(x0)
    z = r
    p = z.clone()
    resids = []
    for _ in range(num_iters):
        Ap = aha_w(p)
        pAp = torch.real(torch.sum(p.conj() * Ap)).item()
        rz = torch.real(torch.sum(r.conj() * z))
        alpha = rz / pAp
        x0 = x0 + alpha * p
        r = r - alpha * Ap
        rnrm = torch.norm(r)
        resids.append(rnrm.item())
        if rnrm < tolerance:
            break
        z = r
        beta = torch.real(torch.sum(r.conj() * z)) / rz
        p = z + beta * p
    return (resids, x0) if return_resids else x0


def power_method_operator(A, x0, num_iter=15):
    """algs.py:364-399."""
    for _ in range(num_iter):
        z = A(x0)
        ll = torch.norm(z)
        x0 = z / ll
    return x0, ll.item()


def cg_sense_recon(A, ksp, max_iter=15, lamda_l2=0.0, max_eigen=None, tolerance=1e-8):
    """recons.py:77-156."""
    if max_eigen is None:
        x0 = torch.randn(A.ishape, dtype=A.cdt)
        _, max_eigen = power_method_operator(A.normal, x0)
        max_eigen *= 1.01
    AHb = A.adjoint(ksp.to(A.cdt)) / (max_eigen ** 0.5)
    if max_iter == 0:
        return AHb
    return conjugate_gradient(lambda x: A.normal(x) / max_eigen, AHb, num_iters=max_iter, lamda_l2=lamda_l2,
                              tolerance=tolerance)


class SubspaceRef:
    """``subspace_linop`` (linops.py:1496-1772) with every batch size 1: k-space of coil c
    is sum_l T_l * sum_k phi_k(tr) F(S_c B_l alpha_k); the adjoint mirrors it with the
    dcf; the Toeplitz normal mixes the coefficient images through the K x K PSF
    kernels of the weights conj(phi_k1) phi_k2 dcf (:1563-1580, 1711-1772)."""

    def __init__(self, trj, mps, phi, dcf=None, nufft=None, use_toeplitz=False, spatial_funcs=None,
                 temporal_funcs=None):
        self.nufft = nufft
        self.cdt = nufft.cdtype
        self.im_size = tuple(mps.shape[1:])
        self.trj_size = tuple(trj.shape[:-1])
        self.trj = trj.to(torch.float32)
        self.mps = mps.to(self.cdt)
        self.phi = phi.to(self.cdt)
        self.dcf = torch.ones(self.trj_size, dtype=torch.float32) if dcf is None else dcf.to(torch.float32)
        d, nt = len(self.im_size), len(self.trj_size)
        self.spatial = (torch.ones((1,) * (1 + d), dtype=self.cdt) if spatial_funcs is None
                        else spatial_funcs.to(self.cdt))
        self.temporal = (torch.ones((1,) * (1 + nt), dtype=self.cdt) if temporal_funcs is None
                         else temporal_funcs.to(self.cdt))
        self.use_toeplitz = use_toeplitz
        if use_toeplitz:
            K = phi.shape[0]
            pp = self.phi.conj()[:, None, :] * self.phi[None]                     # k1 k2 ntr
            w = pp.reshape(K * K, *([1] * (nt - 1)), pp.shape[-1]) * self.dcf
            kerns = [nufft.calc_teoplitz_kernels(self.trj[None], w[None, j:j + 1])[0, 0] for j in range(K * K)]
            self.toep_kerns = torch.stack(kerns, 0).reshape(K, K, *kerns[0].shape)

    def _phi(self, k):
        return self.phi[k].reshape(*([1] * (len(self.trj_size) - 1)), -1)

    def forward(self, alphas):
        nc, K, L = self.mps.shape[0], self.phi.shape[0], self.spatial.shape[0]
        ksp = torch.zeros((nc, *self.trj_size), dtype=self.cdt)
        for c in range(nc):
            for l in range(max(L, self.temporal.shape[0])):
                sl = self.spatial[min(l, L - 1)]
                tl = self.temporal[min(l, self.temporal.shape[0] - 1)]
                acc = torch.zeros(self.trj_size, dtype=self.cdt)
                for k in range(K):
                    sx = (self.mps[c] * sl * alphas[k])[None, None]
                    acc += self.nufft.forward(sx, self.trj[None])[0, 0] * self._phi(k)
                ksp[c] += acc * tl
        return ksp

    def adjoint(self, ksp):
        nc, K, L = self.mps.shape[0], self.phi.shape[0], self.spatial.shape[0]
        out = torch.zeros((K, *self.im_size), dtype=self.cdt)
        for c in range(nc):
            y = ksp[c] * self.dcf
            for l in range(max(L, self.temporal.shape[0])):
                sl = self.spatial[min(l, L - 1)]
                wy = y * self.temporal[min(l, self.temporal.shape[0] - 1)].conj()
                for k in range(K):
                    f = self.nufft.adjoint((wy * self._phi(k).conj())[None, None], self.trj[None])[0, 0]
                    out[k] += f * self.mps[c].conj() * sl.conj()
        return out

    def normal(self, alphas):
        if not self.use_toeplitz:
            return self.adjoint(self.forward(alphas))
        d = len(self.im_size)
        K = self.phi.shape[0]
        pad = PadLastRef(self.toep_kerns.shape[-d:], self.im_size)
        dims = tuple(range(-d, 0))
        out = torch.zeros((K, *self.im_size), dtype=self.cdt)
        for c in range(self.mps.shape[0]):
            F = fft_c(pad.forward((self.mps[c] * alphas.to(self.cdt))[None]), dims)[0]   # K *os
            for i in range(K):
                m = (self.toep_kerns[i] * F).sum(0)
                out[i] += pad.adjoint(ifft_c(m[None], dims))[0] * self.mps[c].conj()
        return out


class SegmentedSenseRef:
    """``sense_linop`` with L > 1 imperfection segments (linops.py:1134-1308), every batch size
    1: k-space of coil c is sum_l T_l * F(S_c B_l x); the adjoint weights by the dcf and conj(T_l)
    and sums conj(S_c) conj(B_l) F^H; the Toeplitz normal mixes the segments through the L x L PSF
    kernels of conj(T_l1) T_l2 dcf (:1103-1117, 1265-1306)."""

    def __init__(self, trj, mps, spatial_funcs, temporal_funcs, dcf=None, nufft=None, use_toeplitz=False):
        self.nufft = nufft
        self.cdt = nufft.cdtype
        self.im_size = tuple(mps.shape[1:])
        self.trj_size = tuple(trj.shape[:-1])
        self.trj = trj.to(torch.float32)
        self.mps = mps.to(self.cdt)
        self.B = spatial_funcs.to(self.cdt)
        self.T = temporal_funcs.to(self.cdt)
        self.dcf = torch.ones(self.trj_size, dtype=torch.float32) if dcf is None else dcf.to(torch.float32)
        self.ishape = self.im_size
        if use_toeplitz:
            L = self.T.shape[0]
            kerns = [[nufft.calc_teoplitz_kernels(self.trj[None], (self.T[a].conj() * self.T[b] * self.dcf)[None])[0]
                      for b in range(L)] for a in range(L)]
            self.toep_kerns = torch.stack([torch.stack(r, 0) for r in kerns], 0)
        else:
            self.toep_kerns = None

    def forward(self, img):
        ksp = torch.zeros((self.mps.shape[0], *self.trj_size), dtype=self.cdt)
        for c in range(self.mps.shape[0]):
            for l in range(self.B.shape[0]):
                sbx = (self.mps[c] * self.B[l] * img)[None, None]
                ksp[c] += self.nufft.forward(sbx, self.trj[None])[0, 0] * self.T[l]
        return ksp

    def adjoint(self, ksp):
        img = torch.zeros(self.im_size, dtype=self.cdt)
        for c in range(self.mps.shape[0]):
            y = ksp[c] * self.dcf
            for l in range(self.T.shape[0]):
                f = self.nufft.adjoint((y * self.T[l].conj())[None, None], self.trj[None])[0, 0]
                img += f * self.mps[c].conj() * self.B[l].conj()
        return img

    def normal(self, img):
        if self.toep_kerns is None:
            return self.adjoint(self.forward(img))
        d = len(self.im_size)
        pad = PadLastRef(self.toep_kerns.shape[-d:], self.im_size)
        dims = tuple(range(-d, 0))
        L = self.B.shape[0]
        out = torch.zeros(self.im_size, dtype=self.cdt)
        for c in range(self.mps.shape[0]):
            F = [fft_c(pad.forward((self.mps[c] * self.B[l] * img)[None]), dims)[0] for l in range(L)]
            for m in range(L):
                acc = sum(self.toep_kerns[m, l] * F[l] for l in range(L))
                out += pad.adjoint(ifft_c(acc[None], dims))[0] * self.B[m].conj() * self.mps[c].conj()
        return out


class ImperfCoilLowrankRef:
    """``imperf_coil_lowrank`` (linops.py:525-759), batch sizes 1: k-space of coil c is
    sum_l h[l, c] * F(B_l x) (:614-652); the adjoint sums conj(B_l) F^H(sum_c conj(h[l, c]) dcf
    y_c) (:654-714). The normal is adjoint(forward) (the reference's Toeplitz branch does not
    run, see the product's docstring)."""

    def __init__(self, trj, spatial_funcs, temporal_funcs, dcf=None, nufft=None):
        self.nufft = nufft
        self.cdt = nufft.cdtype
        self.B = spatial_funcs.to(self.cdt)
        self.h = temporal_funcs.to(self.cdt)
        self.trj = trj.to(torch.float32)
        self.trj_size = tuple(trj.shape[:-1])
        self.im_size = tuple(spatial_funcs.shape[1:])
        self.dcf = torch.ones(self.trj_size, dtype=torch.float32) if dcf is None else dcf.to(torch.float32)

    def forward(self, img):
        L, C = self.h.shape[:2]
        ksp = torch.zeros((C, *self.trj_size), dtype=self.cdt)
        for l in range(L):
            f = self.nufft.forward((img * self.B[l])[None, None], self.trj[None])[0, 0]
            for c in range(C):
                ksp[c] += f * self.h[l, c]
        return ksp

    def adjoint(self, ksp):
        L, C = self.h.shape[:2]
        img = torch.zeros(self.im_size, dtype=self.cdt)
        for l in range(L):
            hdy = sum(ksp[c] * self.dcf * self.h[l, c].conj() for c in range(C))
            img += self.nufft.adjoint(hdy[None, None], self.trj[None])[0, 0] * self.B[l].conj()
        return img

    def normal(self, img):
        return self.adjoint(self.forward(img))
