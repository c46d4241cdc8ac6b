"""Thin tensor-level wrappers over the C-ABI kernels.

All functions take device tensors, launch on PyTorch's current HIP stream and
return without synchronising. Shapes are validated here so that every launch
matches what the kernel and its grid assume.
"""

from __future__ import annotations

import math

import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr
from .plan import SamplePlan


def _c64(t: torch.Tensor, name: str) -> torch.Tensor:
    if t.dtype != torch.complex64:
        raise TypeError(f"{name} must be complex64, got {t.dtype}")
    t = t.resolve_conj().resolve_neg()  # inputs: materialise lazy conj()/neg() views
    if not t.is_contiguous():
        t = t.contiguous()
    return t


def interp(grid: torch.Tensor, plan: SamplePlan, out_scale: float = 1.0, out: torch.Tensor | None = None,
           geom: torch.Tensor | None = None) -> torch.Tensor:
    """Gather: ``grid`` (B, *G) complex64 -> samples (B, M) complex64."""
    _lib.require_device(grid)
    grid = _c64(grid, "grid")
    D = plan.ndim
    if tuple(grid.shape[-D:]) != plan.grid_size:
        raise ValueError(f"grid shape {tuple(grid.shape[-D:])} does not match plan {plan.grid_size}")
    B = int(math.prod(grid.shape[:-D]))
    if out is None:
        out = torch.empty((B, plan.npts), dtype=torch.complex64, device=grid.device)
    elif out.numel() != B * plan.npts or not out.is_contiguous():
        raise ValueError("interp: out must be contiguous with B*M elements")
    g = geom if geom is not None else plan.geom(B, plan.npts, out_scale)
    cc = plan.coil_chunk(B)
    nch = -(-B // cc)
    with _lib.tile_scope(plan.ndim, plan.tile):
        check(lib().mra_interp(D, plan.P, ptr(grid), ptr(out), ptr(plan.wsort), ptr(plan.lbase), ptr(plan.perm),
                               ptr(g), ptr(plan.items), plan.nitems, cc, nch, stream_ptr(grid.device)))
    return out


def gridding(ksp: torch.Tensor, plan: SamplePlan, out: torch.Tensor, dcf: torch.Tensor | None = None,
             out_scale: float = 1.0, accumulate: bool = False, scratch: torch.Tensor | None = None,
             plan_order: bool = False) -> torch.Tensor:
    """Adjoint of ``interp``: samples (B, M) complex64 -> grid ``out`` (B, *G).

    Every point of ``out`` is written (``accumulate=True`` adds to its contents
    instead), so ``out`` needs no zeroing. ``dcf`` (M,) float32 multiplies each
    sample first. ``scratch`` (optional) holds the per-item partial subgrids,
    ``gridding_scratch_elems(plan, B)`` complex64 elements. ``plan_order``: ``ksp``
    and ``dcf`` are in sorted (plan) order.
    """
    B = _grid_batch(plan, out)
    scratch = gridding_partial(ksp, plan, B, dcf=dcf, out_scale=out_scale, scratch=scratch,
                               dcf_plan_order=plan_order, ksp_plan_order=plan_order)
    return gridding_reduce(scratch, plan, out, accumulate=accumulate)


def _grid_batch(plan: SamplePlan, out: torch.Tensor) -> int:
    D = plan.ndim
    if out.dtype != torch.complex64 or not out.is_contiguous():
        raise TypeError("gridding output must be a contiguous complex64 tensor")
    if tuple(out.shape[-D:]) != plan.grid_size:
        raise ValueError(f"output grid {tuple(out.shape[-D:])} does not match plan {plan.grid_size}")
    return int(math.prod(out.shape[:-D]))


def gridding_partial(ksp: torch.Tensor, plan: SamplePlan, nbatch: int, dcf: torch.Tensor | None = None,
                     out_scale: float = 1.0, scratch: torch.Tensor | None = None,
                     dcf_plan_order: bool = False, ksp_plan_order: bool = False) -> torch.Tensor:
    """Gridding phase 1: every work item's samples summed into its subgrids in
    ``scratch`` (returned; allocated when not given). ``dcf_plan_order``: ``dcf`` is
    given in sorted (plan) order, ``dcf[plan.perm]``, and is read contiguously;
    ``ksp_plan_order``: ``ksp`` and ``dcf`` both are."""
    _lib.require_device(ksp, dcf)
    ksp = _c64(ksp, "ksp")
    B = int(nbatch)
    if ksp.numel() != B * plan.npts:
        raise ValueError(f"gridding: ksp has {ksp.numel()} elements, expected {B}x{plan.npts}")
    if dcf is not None:
        if dcf.dtype != torch.float32 or dcf.numel() != plan.npts:
            raise ValueError("dcf must be float32 with one entry per sample")
        dcf = dcf.contiguous()
    g = plan.geom(B, plan.npts, out_scale, dcf_plan_order=dcf is not None and dcf_plan_order,
                  ksp_plan_order=ksp_plan_order)
    need = gridding_scratch_elems(plan, B)
    if scratch is None:
        scratch = torch.empty(need, dtype=torch.complex64, device=ksp.device)
    elif scratch.numel() < need or scratch.dtype != torch.complex64 or not scratch.is_contiguous():
        raise ValueError(f"gridding scratch must be contiguous complex64 with >= {need} elements")
    with _lib.tile_scope(plan.ndim, plan.tile):
        check(lib().mra_grid_partial(plan.ndim, plan.P, ptr(ksp), ptr(dcf), ptr(plan.wsort), ptr(plan.lbase),
                                     ptr(plan.perm), ptr(g), ptr(plan.items), plan.nitems, B, ptr(scratch),
                                     stream_ptr(ksp.device)))
    return scratch


def gridding_reduce(scratch: torch.Tensor, plan: SamplePlan, out: torch.Tensor,
                    accumulate: bool = False) -> torch.Tensor:
    """Gridding phase 2: each output tile sums the scratch subgrids covering it."""
    _lib.require_device(scratch, out)
    B = _grid_batch(plan, out)
    if scratch.numel() < gridding_scratch_elems(plan, B):
        raise ValueError("gridding scratch too small for this batch")
    g = plan.geom(B, plan.npts, 1.0)
    with _lib.tile_scope(plan.ndim, plan.tile):
        check(lib().mra_grid_reduce(plan.ndim, plan.P, ptr(out), ptr(scratch), ptr(g), 0, 1, ptr(plan.tile_items),
                                    plan.ntiles, B, int(accumulate), stream_ptr(out.device)))
    return out


def gridding_noalloc(ksp: torch.Tensor, plan: SamplePlan, out: torch.Tensor, dcf: torch.Tensor | None = None,
                     out_scale: float = 1.0, accumulate: bool = False) -> torch.Tensor:
    """``gridding`` without any device allocation: the phase-1 subgrids go to the library's
    static scratch (``mra_grid_noalloc``), so the call can run under HIP graph capture. Raises
    when the plan needs more scratch than that array holds. Calls that overlap on one GPU
    share the array and must not run concurrently."""
    import ctypes

    _lib.require_device(ksp, dcf, out)
    ksp = _c64(ksp, "ksp")
    B = _grid_batch(plan, out)
    if ksp.numel() != B * plan.npts:
        raise ValueError(f"gridding: ksp has {ksp.numel()} elements, expected {B}x{plan.npts}")
    if dcf is not None:
        if dcf.dtype != torch.float32 or dcf.numel() != plan.npts:
            raise ValueError("dcf must be float32 with one entry per sample")
        dcf = dcf.contiguous()
    gh = plan.geom_host(B, out_scale)
    with _lib.tile_scope(plan.ndim, plan.tile):
        check(lib().mra_grid_noalloc(plan.ndim, plan.P, ptr(ksp), ptr(dcf), ptr(plan.wsort), ptr(plan.lbase),
                                     ptr(plan.perm), ctypes.addressof(gh), ptr(plan.items), plan.nitems, B,
                                     ptr(plan.tile_items), plan.ntiles, ptr(out), int(accumulate),
                                     stream_ptr(out.device)))
    return out


def gridding_scratch_elems(plan: SamplePlan, nbatch: int) -> int:
    """complex64 elements of gridding scratch: one subgrid per (work item, batch entry)."""
    return max(1, plan.nitems * nbatch * plan.subgrid)


def interp_grad(grid: torch.Tensor, a: torch.Tensor, plan: SamplePlan, out_scale: float = 1.0,
                a_plan_order: bool = False) -> torch.Tensor:
    """d/dt of Re<a, interp(grid, t)> per sample: returns (M, D) float32 in grid units,
    in sample order.

    gt[m, d] = sum_b Re( conj(a[b, m]) * sum_p dW_p/dt_d * grid[b, g_p] )

    ``a_plan_order``: ``a`` is given in sorted (plan) order.
    """
    _lib.require_device(grid, a)
    if len(set(plan.betas)) > 1:
        raise NotImplementedError("interp_grad: plans with one beta per dimension are not supported")
    grid = _c64(grid, "grid")
    a = _c64(a, "a")
    D = plan.ndim
    B = int(math.prod(grid.shape[:-D]))
    if a.numel() != B * plan.npts:
        raise ValueError("interp_grad: a must hold B*M samples")
    gt = torch.empty((plan.npts, D), dtype=torch.float32, device=grid.device)
    g = plan.geom(B, plan.npts, out_scale, ksp_plan_order=a_plan_order)
    with _lib.tile_scope(plan.ndim, plan.tile):
        check(lib().mra_interp_grad(D, plan.P, ptr(grid), ptr(a), ptr(gt), ptr(plan.tsort), ptr(plan.perm), ptr(g),
                                    ptr(plan.items), plan.nitems, plan.S, stream_ptr(grid.device)))
    return gt


def pad_expand(x: torch.Tensor, im_size, grid_size, smaps: torch.Tensor | None = None,
               apod: torch.Tensor | None = None, scale: float = 1.0, center=None,
               out: torch.Tensor | None = None) -> torch.Tensor:
    """x (Bx, *N) [, smaps (C, *N)] -> grid (Bx, C, *G) in the unshifted-FFT layout.

    Pixel n lands at grid index (n - c) mod G with c = N//2 unless ``center`` is
    given. ``apod`` is the concatenation of the per-dim apodization tables.
    """
    _lib.require_device(x, smaps, apod)
    D = len(im_size)
    x = _c64(x, "x")
    if tuple(x.shape[-D:]) != tuple(im_size):
        raise ValueError(f"image shape {tuple(x.shape[-D:])} != {tuple(im_size)}")
    nx = int(math.prod(x.shape[:-D]))
    npix = int(math.prod(im_size))
    ns = 1
    if smaps is not None:
        smaps = _c64(smaps, "smaps")
        if tuple(smaps.shape[-D:]) != tuple(im_size):
            raise ValueError("maps must have the image shape")
        ns = int(math.prod(smaps.shape[:-D]))
    if apod is not None and (apod.dtype != torch.float32 or apod.numel() != sum(im_size)):
        raise ValueError("apod must be float32 with sum(im_size) entries")
    shape = (nx, ns, *grid_size) if smaps is not None else (nx, *grid_size)
    if out is None:
        out = torch.empty(shape, dtype=torch.complex64, device=x.device)
    elif out.numel() != math.prod(shape) or out.dtype != torch.complex64 or not out.is_contiguous():
        raise ValueError("pad_expand: bad output tensor")
    cen = None if center is None else _lib.ints(center)
    check(lib().mra_pad_expand(D, _lib.ints(im_size), _lib.ints(grid_size), cen, ptr(x), nx, npix, ptr(smaps), ns,
                               npix, ptr(apod), float(scale), ptr(out), stream_ptr(x.device)))
    return out


def crop_reduce(grid: torch.Tensor, im_size, grid_size, nx: int, smaps: torch.Tensor | None = None,
                apod: torch.Tensor | None = None, scale: float = 1.0, out: torch.Tensor | None = None,
                accumulate: bool = False, roll_out: bool = False, conj_maps: bool = True,
                center=None) -> torch.Tensor:
    """grid (nx, C, *G) -> image (nx, *N): crop, deapodize and (with maps) sum_c conj(S_c) * grid_c."""
    _lib.require_device(grid, smaps, apod, out)
    D = len(im_size)
    grid = _c64(grid, "grid")
    npix = int(math.prod(im_size))
    ns = 1
    if smaps is not None:
        smaps = _c64(smaps, "smaps")
        ns = int(math.prod(smaps.shape[:-D]))
    if grid.numel() != nx * ns * math.prod(grid_size):
        raise ValueError(f"crop_reduce: grid has {grid.numel()} elements, expected {nx}x{ns}x{tuple(grid_size)}")
    if apod is not None and (apod.dtype != torch.float32 or apod.numel() != sum(im_size)):
        raise ValueError("apod must be float32 with sum(im_size) entries")
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs an output tensor")
        out = torch.empty((nx, *im_size), dtype=torch.complex64, device=grid.device)
    elif out.numel() != nx * npix or out.dtype != torch.complex64 or not out.is_contiguous():
        raise ValueError("crop_reduce: bad output tensor")
    cen = None if center is None else _lib.ints(center)
    check(lib().mra_crop_reduce(D, _lib.ints(im_size), _lib.ints(grid_size), cen, ptr(grid), nx, ptr(smaps), ns,
                                npix, int(conj_maps), ptr(apod), float(scale), int(accumulate), int(roll_out),
                                ptr(out), npix, stream_ptr(grid.device)))
    return out


_fused_ok: dict = {}


def fused_supported(grid_size) -> bool:
    """Whether the fused pad+FFT / iFFT+crop transforms cover grids of this shape."""
    key = tuple(int(g) for g in grid_size)
    ok = _fused_ok.get(key)
    if ok is None:
        ok = bool(lib().mra_fft_fused_supported(len(key), _lib.ints(key)))
        _fused_ok[key] = ok
    return ok


def pad_fft(x: torch.Tensor, im_size, grid_size, smaps: torch.Tensor | None = None,
            apod: torch.Tensor | None = None, scale: float = 1.0, center=None,
            img_index: torch.Tensor | None = None, nq: int | None = None,
            out: torch.Tensor | None = None) -> torch.Tensor:
    """Spectrum of the zero-padded images: FFT(pad_expand(x, smaps, apod, scale)) as a
    grid tensor (nx[, C], *G), unnormalised, unshifted layout: new, or written into
    ``out`` (contiguous complex64 with as many elements). With ``img_index`` (batched
    problems), grid q pads image img_index[q] times map q (nq grids)."""
    from .gridfft import _fftn, _lds_fft

    D = len(im_size)
    G = tuple(int(g) for g in grid_size)
    if not fused_supported(G):
        if img_index is not None:
            raise ValueError("pad_fft: batched padding needs a grid with an LDS FFT plan")
        g = pad_expand(x, im_size, G, smaps=smaps, apod=apod, scale=scale, center=center)
        g = _fftn(g.reshape(-1, *G), D).reshape(g.shape)
        if out is None:
            return g
        if out.numel() != g.numel() or out.dtype != torch.complex64 or not out.is_contiguous():
            raise ValueError("pad_fft: bad output tensor")
        out.view(g.shape).copy_(g)
        return out
    _lib.require_device(x, smaps, apod, img_index)
    x = _c64(x, "x")
    if tuple(x.shape[-D:]) != tuple(im_size):
        raise ValueError(f"image shape {tuple(x.shape[-D:])} != {tuple(im_size)}")
    nx = int(math.prod(x.shape[:-D]))
    npix = int(math.prod(im_size))
    ns = 1
    if smaps is not None:
        smaps = _c64(smaps, "smaps")
        if tuple(smaps.shape[-D:]) != tuple(im_size):
            raise ValueError("maps must have the image shape")
        ns = int(math.prod(smaps.shape[:-D]))
    if apod is not None and (apod.dtype != torch.float32 or apod.numel() != sum(im_size)):
        raise ValueError("apod must be float32 with sum(im_size) entries")
    if img_index is not None:
        q = int(nq)
        if img_index.dtype != torch.int32 or img_index.numel() != q:
            raise ValueError("img_index must be int32 with one entry per grid")
        if smaps is not None and ns < q:
            raise ValueError("batched padding needs one map per grid")
        shape = (q, *G)
    else:
        q = nx * ns
        shape = (nx, ns, *G) if smaps is not None else (nx, *G)
    if out is None:
        out = torch.empty(shape, dtype=torch.complex64, device=x.device)
    else:
        _lib.require_device(out)
        if out.numel() != math.prod(shape) or out.dtype != torch.complex64 or not out.is_contiguous():
            raise ValueError("pad_fft: bad output tensor")
        out = out.view(shape)
    cen = None if center is None else _lib.ints(center)
    check(lib().mra_fft_pad_forward(D, _lib.ints(im_size), _lib.ints(G), cen, ptr(x), npix, ptr(smaps), ns, npix,
                                    ptr(img_index), q, ptr(apod), float(scale), ptr(out),
                                    _lds_fft.tables(x.device, G), stream_ptr(x.device)))
    return out


def ifft_crop(grid: torch.Tensor, im_size, grid_size, nx: int, smaps: torch.Tensor | None = None,
              apod: torch.Tensor | None = None, scale: float = 1.0, out: torch.Tensor | None = None,
              accumulate: bool = False, conj_maps: bool = True, center=None,
              coil_off: torch.Tensor | None = None) -> torch.Tensor:
    """crop_reduce(iFFT(grid)) (unnormalised inverse) into image (nx, *N). The grid is
    used as work space and overwritten. With ``coil_off`` (batched problems), image i
    reduces grids and maps coil_off[i] .. coil_off[i+1]-1."""
    from .gridfft import _ifftn, _lds_fft

    D = len(im_size)
    G = tuple(int(g) for g in grid_size)
    if not fused_supported(G):
        if coil_off is not None:
            raise ValueError("ifft_crop: batched reduction needs a grid with an LDS FFT plan")
        g = _ifftn(grid.reshape(-1, *G), D)
        return crop_reduce(g, im_size, G, nx, smaps=smaps, apod=apod, scale=scale, out=out,
                           accumulate=accumulate, conj_maps=conj_maps, center=center)
    _lib.require_device(grid, smaps, apod, out, coil_off)
    if grid.dtype != torch.complex64 or not grid.is_contiguous():
        raise TypeError("ifft_crop: grid must be contiguous complex64 (it is overwritten)")
    npix = int(math.prod(im_size))
    nq = grid.numel() // math.prod(G)
    ns = 1
    if smaps is not None:
        smaps = _c64(smaps, "smaps")
        ns = int(math.prod(smaps.shape[:-D]))
    if coil_off is None and nq != nx * ns:
        raise ValueError(f"ifft_crop: {nq} grids for {nx} images x {ns} maps")
    if coil_off is not None:
        if coil_off.dtype != torch.int32 or coil_off.numel() != nx + 1:
            raise ValueError("coil_off must be int32 with nx + 1 entries")
        if smaps is not None and ns < nq:
            raise ValueError("batched reduction needs one map per grid")
    if apod is not None and (apod.dtype != torch.float32 or apod.numel() != sum(im_size)):
        raise ValueError("apod must be float32 with sum(im_size) entries")
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs an output tensor")
        out = torch.empty((nx, *im_size), dtype=torch.complex64, device=grid.device)
    elif out.numel() != nx * npix or out.dtype != torch.complex64 or not out.is_contiguous():
        raise ValueError("ifft_crop: bad output tensor")
    cen = None if center is None else _lib.ints(center)
    check(lib().mra_fft_crop_adjoint(D, _lib.ints(im_size), _lib.ints(G), cen, ptr(grid), nq, nx, ptr(smaps), ns,
                                     npix, ptr(coil_off), int(conj_maps), ptr(apod), float(scale), int(accumulate),
                                     ptr(out), npix, _lds_fft.tables(grid.device, G), stream_ptr(grid.device)))
    return out


def toeplitz_normal(x: torch.Tensor, im_size, grid_size, smaps: torch.Tensor, kern_unshifted: torch.Tensor,
                    center=None, out: torch.Tensor | None = None, accumulate: bool = False) -> torch.Tensor:
    """sum_c conj(S_c) crop(iFFT(K * FFT(pad(x S_c)))) for one image x (*N) and maps (C, *N),
    K the PSF multiplier on grid ``grid_size`` (unshifted layout, normalisation included)."""
    from .gridfft import _fftn, _ifftn, _lds_fft

    D = len(im_size)
    G = tuple(int(g) for g in grid_size)
    x = _c64(x, "x").reshape(1, *im_size)
    smaps = _c64(smaps, "smaps")
    C = int(math.prod(smaps.shape[:-D]))
    if D < 2 or not fused_supported(G):
        g = pad_expand(x, im_size, G, smaps=smaps, center=center)
        g = _fftn(g.reshape(C, *G), D)
        cmul_bcast(g, kern_unshifted)
        g = _ifftn(g, D)
        return crop_reduce(g, im_size, G, 1, smaps=smaps, center=center, out=out, accumulate=accumulate)
    _lib.require_device(x, smaps, kern_unshifted, out)
    kern = _c64(kern_unshifted, "kern")
    if tuple(kern.shape[-D:]) != G:
        raise ValueError("toeplitz kernel shape != grid shape")
    npix = int(math.prod(im_size))
    if out is None:
        if accumulate:
            raise ValueError("accumulate needs an output tensor")
        out = torch.empty((1, *im_size), dtype=torch.complex64, device=x.device)
    elif out.numel() != npix or out.dtype != torch.complex64 or not out.is_contiguous():
        raise ValueError("toeplitz_normal: bad output tensor")
    work = torch.empty((C, *G), dtype=torch.complex64, device=x.device)
    cen = None if center is None else _lib.ints(center)
    check(lib().mra_toeplitz_fused(D, _lib.ints(im_size), _lib.ints(G), cen, ptr(x), ptr(smaps), C, npix, ptr(kern),
                                   ptr(work), int(accumulate), ptr(out), _lds_fft.tables(x.device, G),
                                   stream_ptr(x.device)))
    return out


def cmul_bcast(data: torch.Tensor, kern: torch.Tensor) -> torch.Tensor:
    """In place: data (B, *S) *= kern (*S), complex64."""
    _lib.require_device(data, kern)
    if not data.is_contiguous() or data.dtype != torch.complex64:
        raise ValueError("cmul_bcast: data must be contiguous complex64")
    kern = _c64(kern, "kern")
    n = kern.numel()
    if data.numel() % n:
        raise ValueError("cmul_bcast: data size is not a multiple of the kernel size")
    check(lib().mra_cmul_bcast(ptr(data), ptr(kern), data.numel() // n, n, stream_ptr(data.device)))
    return data
