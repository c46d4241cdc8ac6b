"""Differentiable NUFFT / SENSE with a hand-written trajectory gradient.

The reference differentiates only with respect to values: ``InterpolateFn`` and
``InterpolateAdjointFn`` pair ungrid with grid and return ``None`` for the
coordinates (triton_interp/interp.py:49-75, 130-155). Here each operator is a
``torch.autograd.Function`` whose backward is the exact adjoint for the values
and, for the trajectory, the derivative of the Kaiser-Bessel weights in the
gather (``mra_interp_grad``):

  forward  y_c = A(S_c x),   A = interp_t . FFT . pad . apod / sqrt(prod N),  interp / w^d
     dL/dt_m = sum_c Re( conj(g_cm) * sum_p dW_p/dt * F_c[p] ) / w^d,   F_c = FFT(pad(apod S_c x))/sqrt(prod N)
  adjoint  x = sum_c conj(S_c) A^H(dcf * y_c)
     dL/dt_m = sum_c Re( conj(dcf_m y_cm) * sum_p dW_p/dt * H_c[p] ) / w^d,  H_c = FFT(pad(apod S_c gx))/sqrt(prod N)

with t = k * (ceil(s N)/N) + ceil(s N)//2, so dL/dk = dL/dt * ceil(s N)/N. The
forward's oversampled grids F_c are kept for backward up to KEEP_GRIDS_BYTES and
recomputed (one pad + FFT) beyond; H_c is formed in backward. ``unrolled_cg_sense`` runs CG-SENSE iterations as differentiable
torch ops around these operators, each iteration optionally under
``torch.utils.checkpoint``.

Complex gradients follow PyTorch's convention: for a real loss L the gradient
of a complex tensor z is dL/dRe z + i dL/dIm z.
"""

from __future__ import annotations

import math

import torch
import torch.utils.checkpoint as ckpt
from torch.autograd import Function

from . import _lib, ops
from .dtypes import complex_dtype
from .fourier import hip_nufft


# SenseNormalFn keeps the forward's oversampled coil grids for the trajectory gradient up to
# this size (one copy per application kept by autograd, e.g. per unrolled CG iteration)
KEEP_GRIDS_BYTES = 256 << 20


class _Ctx:
    """Everything one differentiable SENSE application needs for one trajectory."""

    def __init__(self, nufft: hip_nufft, trj: torch.Tensor, mps: torch.Tensor | None, dcf: torch.Tensor | None):
        _lib.require_device(trj)
        self.nufft = nufft
        self.im_size = tuple(nufft.im_size)
        self.d = len(self.im_size)
        self.G = nufft.os_shape()
        self.trj_shape = tuple(trj.shape[:-1])
        self.plan = nufft.plan(trj.detach().reshape(-1, self.d))
        self.mps = None if mps is None else mps.to(complex_dtype).contiguous()
        self.C = 1 if mps is None else mps.shape[0]
        self.dcf = None if dcf is None else dcf.to(torch.float32).reshape(-1).contiguous()
        # plan (sorted) order copy: the normal operator keeps its k-space in that order
        self.dcf_po = None if self.dcf is None else self.dcf[self.plan.perm.long()].contiguous()
        self.apod = nufft.apod_table(trj.device)
        self.s_fwd = 1.0 / math.sqrt(math.prod(self.im_size))
        self.s_interp = 1.0 / nufft.width ** self.d
        self.s_adj = self.s_interp * self.s_fwd
        self.dt_dk = nufft.scale_table(trj.device, self.im_size)

    def grids(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """FFT(pad(apod * S_c x)) / sqrt(prod N): (C, *G) in the unshifted layout (new, or
        written into ``out``)."""
        x = x.to(complex_dtype).reshape(1, *self.im_size)
        g = ops.pad_fft(x, self.im_size, self.G, smaps=self.mps, apod=self.apod, scale=self.s_fwd, out=out)
        return g.reshape(self.C, *self.G)

    def forward(self, x):
        return self.gather(self.grids(x))

    def gather(self, grids, plan_order=False, out=None):
        """Interpolation of the oversampled grids: (C, M) in sample order, or in plan order
        (new, or written into ``out``)."""
        if out is None:
            out = torch.empty((self.C, self.plan.npts), dtype=complex_dtype, device=grids.device)
        geo = self.plan.geom(self.C, self.plan.npts, self.s_interp, ksp_plan_order=True) if plan_order else None
        return ops.interp(grids, self.plan, out_scale=self.s_interp, out=out, geom=geo)

    def adjoint(self, y, dcf, plan_order=False):
        g = torch.empty((self.C, *self.G), dtype=complex_dtype, device=y.device)
        ops.gridding(y.to(complex_dtype).reshape(self.C, -1), self.plan, g, dcf=dcf, plan_order=plan_order)
        return ops.ifft_crop(g, self.im_size, self.G, 1, smaps=self.mps, apod=self.apod, scale=self.s_adj)[0]

    def dk(self, grids, a, plan_order=False):
        """dL/dk (M, d) from the KB-derivative gather of ``grids`` against coefficients ``a``
        (sample order, or plan order), summed over their batch (C grids, or several terms'
        grids stacked)."""
        B = grids.shape[0]
        gt = ops.interp_grad(grids, a.to(complex_dtype).reshape(B, -1).contiguous(), self.plan,
                             out_scale=self.s_interp, a_plan_order=plan_order)
        return (gt * self.dt_dk).reshape(*self.trj_shape, self.d)


class SenseForwardFn(Function):
    """(x (*im), trj (*trj, d)) -> y (C, *trj): y_c = NUFFT(S_c x)."""

    @staticmethod
    def forward(ctx, x, trj, nufft, mps):
        c = _Ctx(nufft, trj, mps, None)
        grids = c.grids(x.detach())
        y = c.gather(grids)
        ctx.c = c
        keep = ctx.needs_input_grad[1] and grids.numel() * grids.element_size() <= KEEP_GRIDS_BYTES
        ctx.save_for_backward(x.detach(), grids if keep else None)
        return y.reshape(c.C, *c.trj_shape)

    @staticmethod
    def backward(ctx, gy):
        c = ctx.c
        x, grids_x = ctx.saved_tensors
        gx = gt = None
        gy = gy.to(complex_dtype).contiguous()
        if ctx.needs_input_grad[0]:
            gx = c.adjoint(gy, None).reshape(x.shape)
            if not torch.is_complex(x):
                gx = gx.real
        if ctx.needs_input_grad[1]:
            # the forward's grids, or their recompute when they were too large to keep
            gt = c.dk(grids_x if grids_x is not None else c.grids(x), gy)
        return gx, gt, None, None


class SenseAdjointFn(Function):
    """(y (C, *trj), trj) -> x (*im): x = sum_c conj(S_c) NUFFT^H(dcf * y_c)."""

    @staticmethod
    def forward(ctx, y, trj, nufft, mps, dcf):
        c = _Ctx(nufft, trj, mps, dcf)
        x = c.adjoint(y.detach(), c.dcf)
        ctx.c = c
        ctx.save_for_backward(y.detach())
        return x

    @staticmethod
    def backward(ctx, gx):
        c = ctx.c
        (y,) = ctx.saved_tensors
        gy = gt = None
        gx = gx.to(complex_dtype).contiguous()
        if ctx.needs_input_grad[0]:
            gy = c.forward(gx).reshape(y.shape)
            if c.dcf is not None:
                gy = gy * c.dcf.reshape(c.trj_shape)
        if ctx.needs_input_grad[1]:
            a = y.to(complex_dtype).reshape(c.C, -1)
            if c.dcf is not None:
                a = a * c.dcf
            gt = c.dk(c.grids(gx), a)
        return gy, gt, None, None, None


class SenseNormalFn(Function):
    """(x (*im), trj) -> A^H D A x in one function, the k-space in between kept in plan
    (sorted) order so the gather stores it and gridding loads it contiguously; G_x (the
    oversampled grids of x) is kept for backward up to KEEP_GRIDS_BYTES. Backward
    (D real, A^H D A self-adjoint), with g the output cotangent and G_g = FFT(pad(S g)):
        gx = A^H D A g
        gt = dk(G_g, D A x) + dk(G_x, D A g)      (the adjoint's and the forward's k-terms)
    G_g is computed once and serves both the gather D A g and the first k-term; A x is
    saved from the forward. When G_x is kept, G_g is written beside it and A x beside A g,
    so the two k-terms are one KB-derivative gather over 2C grids."""

    @staticmethod
    def forward(ctx, x, trj, nufft, mps, dcf):
        c = _Ctx(nufft, trj, mps, dcf)
        need_t = ctx.needs_input_grad[1]
        gbytes = c.C * math.prod(c.G) * 8
        keep = need_t and 2 * gbytes <= KEEP_GRIDS_BYTES
        dev = x.device
        # [G_g | G_x]: G_g is formed in backward
        g2 = torch.empty((2, c.C, *c.G), dtype=complex_dtype, device=dev) if keep else None
        grids = c.grids(x.detach(), out=g2[1] if keep else None)
        # [A x | A g] in plan order: A g is formed in backward
        y2 = torch.empty((2, c.C, c.plan.npts), dtype=complex_dtype, device=dev) if need_t else None
        y = c.gather(grids, plan_order=True, out=y2[0] if need_t else None)
        out = c.adjoint(y, c.dcf_po, plan_order=True)
        ctx.c = c
        ctx.save_for_backward(x.detach(), y if y2 is None else y2, g2)
        return out

    @staticmethod
    def backward(ctx, gout):
        c = ctx.c
        x, ys, g2 = ctx.saved_tensors
        gout = gout.to(complex_dtype).contiguous()
        gx = gt = None
        need_x, need_t = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if need_x or need_t:
            Gg = c.grids(gout, out=g2[0] if g2 is not None else None)
            # A g, plan order (beside A x when the trajectory gradient is needed)
            ag = c.gather(Gg, plan_order=True, out=ys[1] if need_t else None)
            if need_x:
                gx = c.adjoint(ag, c.dcf_po, plan_order=True).reshape(x.shape)
                if not torch.is_complex(x):
                    gx = gx.real
            if need_t:
                # [D A x | D A g]: the coefficients of the adjoint's and the forward's k-terms
                dys = ys * c.dcf_po if c.dcf_po is not None else ys
                if g2 is not None:
                    gt = c.dk(g2.reshape(2 * c.C, *c.G), dys.reshape(2 * c.C, -1), plan_order=True)
                else:
                    gt = c.dk(Gg, dys[0], plan_order=True) + c.dk(c.grids(x), dys[1], plan_order=True)
        return gx, gt, None, None, None


def nufft_forward(img: torch.Tensor, trj: torch.Tensor, nufft: hip_nufft) -> torch.Tensor:
    """Differentiable single-image NUFFT: img (*im_size), trj (*trj, d) -> ksp (*trj)."""
    return SenseForwardFn.apply(img, trj, nufft, None)[0]


def nufft_adjoint(ksp: torch.Tensor, trj: torch.Tensor, nufft: hip_nufft) -> torch.Tensor:
    """Differentiable adjoint NUFFT: ksp (*trj), trj (*trj, d) -> img (*im_size)."""
    return SenseAdjointFn.apply(ksp[None], trj, nufft, None, None)


def sense_forward(x, trj, mps, nufft: hip_nufft):
    return SenseForwardFn.apply(x, trj, nufft, mps)


def sense_adjoint(y, trj, mps, nufft: hip_nufft, dcf=None):
    return SenseAdjointFn.apply(y, trj, nufft, mps, dcf)


def sense_normal(x, trj, mps, nufft: hip_nufft, dcf=None):
    """A^H D A x, differentiable in x and trj (one autograd function, k-space in plan order)."""
    return SenseNormalFn.apply(x, trj, nufft, mps, dcf)


class _DotRe64(Function):
    """Re <a, b> = sum Re(conj(a) b) of complex64 vectors, accumulated in FP64 (a float64
    scalar). Backward returns complex64 gradients (g b, g a), so the FP64 scalar does not
    promote the vector gradients of the unrolled CG to complex128."""

    @staticmethod
    def forward(ctx, a, b):
        ctx.save_for_backward(a, b)
        return torch.real(torch.sum(a.conj() * b, dtype=torch.complex128))

    @staticmethod
    def backward(ctx, g):
        a, b = ctx.saved_tensors
        g32 = g.to(torch.float32)
        ga = g32 * b if ctx.needs_input_grad[0] else None
        gb = g32 * a if ctx.needs_input_grad[1] else None
        return ga, gb


def dot_re64(a, b):
    """Re <a, b> of complex64 tensors as an FP64 scalar (differentiable)."""
    return _DotRe64.apply(a, b)


def _vec(t: torch.Tensor) -> torch.Tensor:
    """A complex64 vector the kernels can read: conj/neg views materialised, contiguous."""
    return t.to(complex_dtype).resolve_conj().resolve_neg().contiguous()


class _CGStepFn(Function):
    """One CG iteration (algs.py:779-872 with P = identity) once nrm = (A^H D A) p is formed:
    ap = inv * nrm + lam * p, pAp = Re<p, ap>, alpha = rz / pAp, x1 = x + alpha p,
    r1 = r - alpha ap, rz1 = <r1, r1>, p1 = r1 + (rz1 / rz) p, with the inner products and
    step sizes in FP64 (alpha and beta cast to fp32 for the vector updates). Forward and
    backward are a handful of fused kernels (``mra_cg_step_fwd`` / ``mra_cg_step_bwd``)
    instead of a dozen torch ops, each with its own autograd node and scalar reductions."""

    @staticmethod
    def forward(ctx, x, r, p, rz, nrm, inv, lam):
        x, r, p, nrm = _vec(x), _vec(r), _vec(p), _vec(nrm)
        rz = rz.to(torch.float64).contiguous()
        dev = x.device
        L = _lib.lib()
        x1, r1, p1 = torch.empty_like(x), torch.empty_like(r), torch.empty_like(p)
        rz1 = torch.empty((), dtype=torch.float64, device=dev)
        sd = torch.empty(8, dtype=torch.float64, device=dev)
        part = torch.empty(L.mra_cg_partials(), dtype=torch.float64, device=dev)
        _lib.check(L.mra_cg_step_fwd(_lib.ptr(x), _lib.ptr(r), _lib.ptr(p), _lib.ptr(nrm), _lib.ptr(rz), x.numel(),
                                     float(inv), float(lam), _lib.ptr(x1), _lib.ptr(r1), _lib.ptr(p1),
                                     _lib.ptr(rz1), _lib.ptr(sd), _lib.ptr(part), _lib.stream_ptr(dev)))
        ctx.save_for_backward(p, nrm, r1, sd)
        ctx.inv, ctx.lam = float(inv), float(lam)
        return x1, r1, p1, rz1

    @staticmethod
    def backward(ctx, gx1, gr1, gp1, grz1):
        p, nrm, r1, sd = ctx.saved_tensors
        dev = p.device
        L = _lib.lib()
        gx1, gr1, gp1 = (None if g is None else _vec(g) for g in (gx1, gr1, gp1))
        grz1 = None if grz1 is None else grz1.to(torch.float64).contiguous()
        gr, gp, gnrm = torch.empty_like(p), torch.empty_like(p), torch.empty_like(p)
        grz = torch.empty((), dtype=torch.float64, device=dev)
        part = torch.empty(4 * L.mra_cg_partials(), dtype=torch.float64, device=dev)
        bs = torch.empty(8, dtype=torch.float64, device=dev)
        _lib.check(L.mra_cg_step_bwd(_lib.ptr(p), _lib.ptr(nrm), _lib.ptr(r1), _lib.ptr(sd), _lib.ptr(gx1),
                                     _lib.ptr(gr1), _lib.ptr(gp1), _lib.ptr(grz1), p.numel(), ctx.inv, ctx.lam,
                                     _lib.ptr(gr), _lib.ptr(gp), _lib.ptr(gnrm), _lib.ptr(grz), _lib.ptr(part),
                                     _lib.ptr(bs), _lib.stream_ptr(dev)))
        return gx1, gr, gp, grz, gnrm, None, None


def cg_step(x, r, p, rz, nrm, inv_eig: float, lamda: float):
    """Differentiable CG update (x, r, p, rz) -> (x1, r1, p1, rz1) given nrm = (A^H D A) p;
    see ``_CGStepFn``."""
    return _CGStepFn.apply(x, r, p, rz, nrm, inv_eig, lamda)


def _cg_iter(x, r, p, rz, trj, mps, nufft, dcf, lamda, inv_eig):
    """One CG step (algs.py:779-872 with P = identity): the normal operator as its autograd
    function, the vector work and its FP64 scalars as one fused differentiable step."""
    return cg_step(x, r, p, rz, sense_normal(p, trj, mps, nufft, dcf), inv_eig, lamda)


def unrolled_cg_sense(ksp, trj, mps, nufft: hip_nufft, num_iters: int = 10, lamda_l2: float = 0.0,
                      max_eigen: float = 1.0, dcf=None, checkpoint: bool = False):
    """``num_iters`` CG-SENSE iterations on (A^H D A / max_eigen + lamda I) x = A^H D y / sqrt(max_eigen),
    started at x0 = A^H D y / sqrt(max_eigen) (recons.py:77-156), differentiable in ksp, trj and mps.

    Without ``checkpoint`` autograd keeps each iteration's vectors and the normal
    operator's k-space (C x M complex64 per iteration: 17 MB at 16 coils x 131 k samples).
    With ``checkpoint`` each iteration keeps only its inputs and is recomputed during
    backward, so memory stays at a few images per iteration, for problems whose k-space
    would not fit; it gives the same bits and costs time (1.7x per loss+gradient evaluation
    at that size, measured).
    """
    inv = 1.0 / max_eigen
    ahb = sense_adjoint(ksp, trj, mps, nufft, dcf) * (max_eigen ** -0.5)
    x = ahb
    if num_iters == 0:
        return x
    r = ahb - (sense_normal(x, trj, mps, nufft, dcf) * inv + lamda_l2 * x)
    p = r
    rz = dot_re64(r, r)
    for _ in range(num_iters):
        if checkpoint:
            x, r, p, rz = ckpt.checkpoint(_cg_iter, x, r, p, rz, trj, mps, nufft, dcf, lamda_l2, inv,
                                          use_reentrant=False)
        else:
            x, r, p, rz = _cg_iter(x, r, p, rz, trj, mps, nufft, dcf, lamda_l2, inv)
    return x


__all__ = ["SenseForwardFn", "SenseAdjointFn", "SenseNormalFn", "nufft_forward", "nufft_adjoint", "sense_forward",
           "sense_adjoint", "sense_normal", "unrolled_cg_sense", "dot_re64", "cg_step"]
