"""Lower surface of the NUFFT: KB interpolation (``ungrid``) and gridding (``grid``).

Mirrors the reference package ``mr_recon.triton_interp`` call for call:
  ``ungrid``   ungrid.py:34-63      ``grid``                grid.py:32-67
  ``interpolate``  interp.py:77-87  ``interpolate_adjoint`` interp.py:157-177
with the same argument meaning, shape handling (prep_ungrid_shapes
ungrid.py:851-894, prep_grid_shapes grid.py:508-553) and errors. The compute is
the hand-written HIP kernels; there is no CPU path.

Differences from the reference, all additive: ``interpolate`` and
``interpolate_adjoint`` also return a gradient for ``locs`` in 1D to 3D (the
reference returns None, interp.py:66-74 and :145-154). 4D and 5D grids (the
reference's _ungrid4d / _ungrid5d, ungrid.py:498-845) run the direct kernels
``mra_interp_direct`` / ``mra_grid_direct``. Only the kernel the NUFFT uses is
implemented on the GPU (kaiser_bessel, norm '1', pad_mode 'circular'); the
other options raise NotImplementedError.
"""

from __future__ import annotations

import math

import torch
from torch.autograd import Function

from .. import ops
from ..plan import SamplePlan

__all__ = ["ungrid", "grid", "interpolate", "interpolate_adjoint", "InterpolateFn", "InterpolateAdjointFn"]

MAX_KERNEL_WIDTH_1D = 64  # ungrid.py:31 (block width limit; complex doubles the last dim)


def _width_tuple(width, ndim):
    if isinstance(width, (int, float)):
        return (float(width),) * ndim
    width = tuple(width)
    if len(width) != ndim:
        raise ValueError(
            "If width specified as tuple it must be same length as grid size but got "
            f"len(width) = {len(width)} and len(grid_size) = {ndim}")
    return width


def _lib_require(*ts):
    from .. import _lib

    _lib.require_device(*ts)


def _check_options(kernel, norm, pad_mode, width):
    if kernel not in ("kaiser_bessel", "spline"):
        raise ValueError(f"Unrecognized kernel type: {kernel}")
    if kernel != "kaiser_bessel":
        raise NotImplementedError("the HIP path implements kernel='kaiser_bessel' only")
    if norm not in ("1", "2"):
        raise ValueError(f"Unexpected norm: {norm}.")
    if norm != "1":
        raise NotImplementedError("the HIP path implements norm='1' (separable weights) only")
    if pad_mode != "circular":
        raise NotImplementedError("the HIP path implements pad_mode='circular' only")
    if len(set(width)) != 1:
        raise NotImplementedError("the HIP path needs the same kernel width in every dimension")
    bw = [2 ** math.ceil(math.log2(math.ceil(w + 1))) for w in width]
    bw[-1] *= 2
    if max(bw) > MAX_KERNEL_WIDTH_1D:
        raise ValueError(f"Necessary block width {bw} has entry which exceeds maximum width {MAX_KERNEL_WIDTH_1D}")
    if len(width) > 5:
        raise ValueError(f"interpolation supports 1 to 5 dimensions, got {len(width)}")
    if len(width) > 3 and math.floor(width[0]) + 1 > 7:
        raise NotImplementedError("4D / 5D interpolation supports kernel widths below 7")


def _beta(kernel_params):
    kp = {} if kernel_params is None else dict(kernel_params)
    return float(kp.get("beta", 1.0))


def _as_complex(v: torch.Tensor):
    if torch.is_complex(v):
        return v.to(torch.complex64), True
    return v.to(torch.float32).to(torch.complex64), False


def _plan(locs, grid_size, width, beta):
    return SamplePlan(locs.to(torch.float32), tuple(grid_size), width[0], beta)


def _direct(adjoint: bool, v: torch.Tensor, locs: torch.Tensor, grid_size, width, beta, nbatch: int):
    """4D / 5D: the direct kernels (no plan). Coordinates are wrapped to [0, G) with fp32
    ``remainder`` as prep_ungrid_shapes / prep_grid_shapes do (ungrid.py:864-865,
    grid.py:516-517)."""
    from .. import _lib

    D = len(grid_size)
    t = torch.remainder(locs.reshape(-1, D).to(torch.float32),
                        torch.tensor(grid_size, dtype=torch.float32, device=locs.device)).contiguous()
    M = t.shape[0]
    P = int(math.floor(width[0])) + 1
    betas = _lib.floats([beta] * D)
    st = _lib.stream_ptr(v.device)
    if adjoint:
        out = torch.zeros((nbatch, *grid_size), dtype=torch.complex64, device=v.device)
        _lib.check(_lib.lib().mra_grid_direct(D, P, _lib.ptr(v), _lib.ptr(t), M, nbatch, _lib.ints(grid_size), betas,
                                              float(width[0]), 1.0, _lib.ptr(out), st))
    else:
        out = torch.empty((nbatch, M), dtype=torch.complex64, device=v.device)
        _lib.check(_lib.lib().mra_interp_direct(D, P, _lib.ptr(v), _lib.ptr(t), M, nbatch, _lib.ints(grid_size),
                                                betas, float(width[0]), 1.0, _lib.ptr(out), st))
    return out


def ungrid(vals, locs, width, kernel: str = "kaiser_bessel", norm: str = "1", pad_mode: str = "circular",
           kernel_params: dict = None, plan: SamplePlan | None = None):
    """Interpolate on-grid ``vals`` (..., *G) to off-grid ``locs`` (..., D) in grid units."""
    ndim = locs.shape[-1]
    grid_size = tuple(vals.shape[-ndim:])
    width = _width_tuple(width, ndim)
    _check_options(kernel, norm, pad_mode, width)
    batch_shape = vals.shape[:-ndim]
    vc, was_complex = _as_complex(vals)
    vflat = vc.reshape(-1, *grid_size).contiguous()
    if ndim > 3:
        _lib_require(vals, locs)
        out = _direct(False, vflat, locs, grid_size, width, _beta(kernel_params), vflat.shape[0])
        out = out.reshape(*batch_shape, *locs.shape[:-1])
        return out if was_complex else out.real.contiguous()
    if plan is None:
        plan = _plan(locs, grid_size, width, _beta(kernel_params))
    out = ops.interp(vflat, plan)
    out = out.reshape(*batch_shape, *locs.shape[:-1])
    return out if was_complex else out.real.contiguous()


def grid(vals, locs, grid_size, width, kernel: str = "kaiser_bessel", norm: str = "1",
         pad_mode: str = "circular", kernel_params: dict = None, plan: SamplePlan | None = None):
    """Adjoint of ``ungrid``: accumulate off-grid ``vals`` (..., *locs_batch) onto a grid of ``grid_size``."""
    ndim = len(grid_size)
    if locs.shape[-1] != ndim:
        raise ValueError(
            f"Expected last dim of locs: to equal ndim but got locs.shape[-1] {locs.shape[-1]} != ndim {ndim}")
    width = _width_tuple(width, ndim)
    _check_options(kernel, norm, pad_mode, width)
    locs_batch = locs.shape[:-1]
    batch_shape = vals.shape[: vals.ndim - len(locs_batch)]
    vc, was_complex = _as_complex(vals)
    nbatch = int(math.prod(batch_shape))
    vflat = vc.reshape(nbatch, -1).contiguous()
    if ndim > 3:
        _lib_require(vals, locs)
        out = _direct(True, vflat, locs, tuple(grid_size), width, _beta(kernel_params), nbatch)
        out = out.reshape(*batch_shape, *grid_size)
        return out if was_complex else out.real.contiguous()
    if plan is None:
        plan = _plan(locs, grid_size, width, _beta(kernel_params))
    out = torch.empty((nbatch, *grid_size), dtype=torch.complex64, device=vals.device)
    ops.gridding(vflat, plan, out)
    out = out.reshape(*batch_shape, *grid_size)
    return out if was_complex else out.real.contiguous()


class InterpolateFn(Function):
    """ungrid with backward = grid for ``vals`` and the KB-derivative gather for ``locs``."""

    @staticmethod
    def forward(ctx, vals, locs, width, kernel, norm, pad_mode, kernel_params):
        ndim = locs.shape[-1]
        grid_size = tuple(vals.shape[-ndim:])
        w = _width_tuple(width, ndim)
        _check_options(kernel, norm, pad_mode, w)
        plan = _plan(locs.detach(), grid_size, w, _beta(kernel_params)) if ndim <= 3 else None
        out = ungrid(vals.detach(), locs.detach(), width, kernel, norm, pad_mode, kernel_params, plan=plan)
        ctx.locs = locs.detach()
        ctx.plan, ctx.grid_size, ctx.width = plan, grid_size, width
        ctx.opts = (kernel, norm, pad_mode, kernel_params)
        ctx.vals_complex = torch.is_complex(vals)
        ctx.save_for_backward(vals.detach() if ctx.needs_input_grad[1] else None)
        ctx.locs_shape = locs.shape
        return out

    @staticmethod
    def backward(ctx, grad_out):
        (vals,) = ctx.saved_tensors
        plan, ndim = ctx.plan, len(ctx.grid_size)
        g_vals = g_locs = None
        gc = grad_out.to(torch.complex64) if not torch.is_complex(grad_out) else grad_out
        batch_shape = grad_out.shape[: grad_out.ndim - (len(ctx.locs_shape) - 1)]
        nb = int(math.prod(batch_shape))
        if ctx.needs_input_grad[0]:
            if plan is None:  # 4D / 5D
                g = grid(gc.reshape(nb, -1), ctx.locs.reshape(-1, ndim), ctx.grid_size, ctx.width, *ctx.opts)
            else:
                g = torch.empty((nb, *ctx.grid_size), dtype=torch.complex64, device=grad_out.device)
                ops.gridding(gc.reshape(nb, -1), plan, g)
            g = g.reshape(*batch_shape, *ctx.grid_size)
            g_vals = g if ctx.vals_complex else g.real
        if ctx.needs_input_grad[1] and plan is None:
            raise NotImplementedError("the coordinate gradient is implemented for 1D to 3D grids")
        if ctx.needs_input_grad[1]:
            vc, _ = _as_complex(vals)
            gt = ops.interp_grad(vc.reshape(nb, *ctx.grid_size), gc.reshape(nb, -1).contiguous(), plan)
            g_locs = gt.reshape(ctx.locs_shape)
        return g_vals, g_locs, None, None, None, None, None


def interpolate(vals, locs, width, kernel="kaiser_bessel", norm: str = "1", pad_mode: str = "circular",
                kernel_params: dict = None):
    return InterpolateFn.apply(vals, locs, width, kernel, norm, pad_mode, kernel_params)


class InterpolateAdjointFn(Function):
    """grid with backward = ungrid for ``vals`` and the KB-derivative gather for ``locs``."""

    @staticmethod
    def forward(ctx, vals, locs, grid_size, width, kernel, norm, pad_mode, kernel_params):
        ndim = len(grid_size)
        w = _width_tuple(width, ndim)
        _check_options(kernel, norm, pad_mode, w)
        plan = _plan(locs.detach(), tuple(grid_size), w, _beta(kernel_params)) if ndim <= 3 else None
        out = grid(vals.detach(), locs.detach(), grid_size, width, kernel, norm, pad_mode, kernel_params, plan=plan)
        ctx.plan, ctx.grid_size = plan, tuple(grid_size)
        ctx.locs, ctx.width, ctx.opts = locs.detach(), width, (kernel, norm, pad_mode, kernel_params)
        ctx.vals_complex = torch.is_complex(vals)
        ctx.vals_shape = vals.shape
        ctx.locs_shape = locs.shape
        ctx.save_for_backward(vals.detach() if ctx.needs_input_grad[1] else None)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        (vals,) = ctx.saved_tensors
        plan = ctx.plan
        ndim = len(ctx.grid_size)
        gc = grad_out.to(torch.complex64) if not torch.is_complex(grad_out) else grad_out
        nb = int(math.prod(grad_out.shape[:-ndim]))
        gflat = gc.reshape(nb, *ctx.grid_size)
        g_vals = g_locs = None
        if ctx.needs_input_grad[0]:
            if plan is None:  # 4D / 5D
                g = ungrid(gflat, ctx.locs.reshape(-1, ndim), ctx.width, *ctx.opts).reshape(ctx.vals_shape)
            else:
                g = ops.interp(gflat, plan).reshape(ctx.vals_shape)
            g_vals = g if ctx.vals_complex else g.real
        if ctx.needs_input_grad[1] and plan is None:
            raise NotImplementedError("the coordinate gradient is implemented for 1D to 3D grids")
        if ctx.needs_input_grad[1]:
            vc, _ = _as_complex(vals)
            gt = ops.interp_grad(gflat, vc.reshape(nb, -1).contiguous(), plan)
            g_locs = gt.reshape(ctx.locs_shape)
        return g_vals, g_locs, None, None, None, None, None, None


def interpolate_adjoint(vals, locs, grid_size, width, kernel: str = "kaiser_bessel", norm: str = "1",
                        pad_mode: str = "circular", kernel_params: dict = None):
    return InterpolateAdjointFn.apply(vals, locs, grid_size, width, kernel, norm, pad_mode, kernel_params)
