"""Solvers on the NUFFT path: ``conjugate_gradient``, ``power_method_operator`` and
``density_compensation``.

Same signatures and iteration as the reference (algs.py:364-399, 779-872):
x0 = AHb, r = AHb - (AHA x0 + lamda_l2 x0), p = r; per iteration Ap, pAp,
alpha = <r,r>/pAp, x += alpha p, r -= alpha Ap, stop when ||r|| < tolerance,
beta = <r_new,r_new>/<r,r>, p = r + beta p.

On complex64 device tensors with no preconditioner the vector work runs in the
HIP CG kernels: inner products accumulate in FP64, alpha/beta stay on the
device, and a step issues no host synchronisation except the tolerance test
(skipped when ``tolerance`` is None or 0, which makes the loop capturable in a
HIP graph). With a preconditioner ``P`` or ``weights`` the reference's torch
arithmetic is used.
"""

from __future__ import annotations

from typing import Optional, Tuple

import torch
import torch.nn as nn

from . import _lib
from ._lib import check, lib, ptr, stream_ptr


class CGWorkspace:
    """Device buffers of the fused CG: r, p, FP64 partial sums and scalars."""

    def __init__(self, like: torch.Tensor):
        self.n = like.numel()
        self.r = torch.empty_like(like)
        self.p = torch.empty_like(like)
        self.part = torch.empty(lib().mra_cg_partials(), dtype=torch.float64, device=like.device)
        self.sc = torch.zeros(8, dtype=torch.float64, device=like.device)

    def init(self, b: torch.Tensor, ax: torch.Tensor, x: torch.Tensor, lam: float) -> None:
        check(lib().mra_cg_init(ptr(b), ptr(ax), ptr(x), ptr(self.r), ptr(self.p), self.n, float(lam),
                                ptr(self.part), ptr(self.sc), stream_ptr(x.device)))

    def step(self, x: torch.Tensor, ap: torch.Tensor, lam: float) -> None:
        check(lib().mra_cg_step(ptr(x), ptr(self.r), ptr(self.p), ptr(ap), self.n, float(lam), ptr(self.part),
                                ptr(self.sc), stream_ptr(x.device)))

    def rnorm(self) -> torch.Tensor:
        """||r|| as a device scalar (FP64)."""
        return self.sc[0].sqrt()


def _fused_ok(AHb: torch.Tensor, P, weights) -> bool:
    return P is None and weights is None and AHb.is_cuda and AHb.dtype == torch.complex64


def conjugate_gradient(AHA: nn.Module, AHb: torch.Tensor, P: Optional[nn.Module] = None,
                       num_iters: Optional[int] = 10, lamda_l2: Optional[float] = 0.0,
                       tolerance: Optional[float] = 1e-8, return_resids: Optional[bool] = False,
                       weights: Optional[torch.Tensor] = None, verbose=True):
    """Solve (AHA + lamda_l2 I) x = AHb starting at x0 = AHb."""
    if not _fused_ok(AHb, P, weights):
        return _cg_torch(AHA, AHb, P, num_iters, lamda_l2, tolerance, return_resids, weights)
    x = AHb.clone().contiguous()
    if num_iters == 0:
        return x
    b = AHb.contiguous()
    ws = CGWorkspace(x)
    ws.init(b, AHA(x).contiguous(), x, lamda_l2)
    resids = []
    xs = [x.clone()] if return_resids else None
    check_tol = tolerance is not None and tolerance > 0
    for _ in range(num_iters):
        ap = AHA(ws.p).contiguous()
        ws.step(x, ap, lamda_l2)
        if return_resids or check_tol:
            rn = float(ws.rnorm())
            if return_resids:
                resids.append(rn)
                xs.append(x.clone())
            if check_tol and rn < tolerance:
                break
    return (resids, xs) if return_resids else x


def _cg_torch(AHA, AHb, P, num_iters, lamda_l2, tolerance, return_resids, weights):
    """Reference arithmetic (algs.py:779-872), used with a preconditioner or weights."""
    if P is None:
        P = lambda x: x  # noqa: E731
    if weights is not None:
        wrap = lambda x: AHA(x) + lamda_l2 * weights * x  # noqa: E731
    else:
        wrap = lambda x: AHA(x) + lamda_l2 * x  # noqa: E731
    x0 = AHb.clone()
    if num_iters == 0:
        return x0
    r = AHb - wrap(x0)
    z = P(r)
    p = z.clone()
    resids, xs = [], [x0.clone()]
    for _ in range(num_iters):
        Ap = wrap(p)
        pAp = torch.real(torch.sum(p.conj() * Ap)).item()
        rz = torch.real(torch.sum(r.conj() * z))
        alpha = rz / pAp
        x0 = x0 + alpha * p
        r = r - alpha * Ap
        rnrm = torch.norm(r)
        if return_resids:
            resids.append(rnrm.item())
            xs.append(x0.clone())
        if tolerance is not None and rnrm < tolerance:
            break
        z = P(r)
        beta = torch.real(torch.sum(r.conj() * z)) / rz
        p = z + beta * p
    return (resids, xs) if return_resids else x0


def power_method_operator(A: callable, x0: torch.Tensor, num_iter: Optional[int] = 15,
                          verbose: Optional[bool] = True) -> Tuple[torch.Tensor, float]:
    """Largest eigenvalue of A by power iteration (algs.py:364-399)."""
    ll = torch.tensor(0.0)
    for _ in range(num_iter):
        z = A(x0)
        ll = torch.norm(z)
        x0 = z / ll
    if verbose:
        print(f"Max Eigenvalue = {ll}")
    return x0, ll.item()


__all__ = ["conjugate_gradient", "power_method_operator", "CGWorkspace"]


# Kaiser-Bessel gridding of the Pipe-Menon iteration: sigpy.mri.pipe_menon_dcf's
# defaults (width=4, beta=8), which density_compensation (algs.py:61-64) keeps.
PIPE_MENON_WIDTH = 4.0
PIPE_MENON_BETA = 8.0


def density_compensation(trj: torch.Tensor, im_size: tuple, num_iters: Optional[int] = 30,
                         method: str = "sigpy") -> torch.Tensor:
    """Pipe-Menon density compensation factors, shape ``trj.shape[:-1]`` (algs.py:20-106).

    ``trj`` (..., d) float32 device tensor in grid units (-N/2 .. N/2). The reference
    hands the trajectory to ``sigpy.mri.pipe_menon_dcf(trj, im_size, max_iter=num_iters)``:
    w = 1, then ``num_iters`` times w /= |G^H G w| with G the Kaiser-Bessel gridding
    (width 4, beta 8) of the samples onto an ``im_size`` grid -- no oversampling, no
    shift, grid indices wrapping mod N -- and finally dcf /= max(dcf) (algs.py:105).

    Here G is the HIP gridding (with w as its per-sample weight on unit samples) and
    G^H the HIP interpolation over one sample plan of the trajectory; the division
    and the normalisation are HIP kernels. The iteration keeps w and G^H G w in the
    plan's sorted order, so every pass over the samples is contiguous; the final
    normalisation stores through the sort permutation. Only ``method='sigpy'``
    exists in the reference (its other branches are commented out).
    """
    from . import ops
    from .plan import SamplePlan

    method = method.lower()
    if method != "sigpy":
        raise ValueError(f"density_compensation: unknown method {method!r} (only 'sigpy' is implemented)")
    _lib.require_device(trj)
    D = int(trj.shape[-1])
    im_size = tuple(int(n) for n in im_size)
    if len(im_size) != D:
        raise ValueError(f"im_size {im_size} does not match trajectory dimension {D}")
    if num_iters is None or num_iters < 0:
        raise ValueError("num_iters must be a non-negative integer")
    coords = trj.reshape(-1, D).to(torch.float32).contiguous()
    plan = SamplePlan(coords, im_size, PIPE_MENON_WIDTH, PIPE_MENON_BETA)
    M = plan.npts
    dev = trj.device
    st = stream_ptr(dev)
    L = lib()
    w = torch.ones(M, dtype=torch.float32, device=dev)  # plan (sorted) order
    if num_iters > 0:
        unit = torch.ones((1, M), dtype=torch.complex64, device=dev)
        grid = torch.empty((1, *im_size), dtype=torch.complex64, device=dev)
        ghg = torch.empty((1, M), dtype=torch.complex64, device=dev)
        scratch = torch.empty(ops.gridding_scratch_elems(plan, 1), dtype=torch.complex64, device=dev)
        geom = plan.geom(1, ksp_plan_order=True)
        for _ in range(int(num_iters)):
            ops.gridding_reduce(ops.gridding_partial(unit, plan, 1, dcf=w, scratch=scratch, ksp_plan_order=True),
                                plan, grid)
            ops.interp(grid, plan, out=ghg, geom=geom)
            check(L.mra_dcf_update(ptr(ghg), ptr(w), M, st))
    part = torch.empty(L.mra_dcf_partials(), dtype=torch.float32, device=dev)
    out = torch.empty(M, dtype=torch.float32, device=dev)
    check(L.mra_dcf_normalize(ptr(w), ptr(plan.perm), ptr(out), M, ptr(part), st))
    return out.reshape(trj.shape[:-1])
