"""ctypes binding of the HIP C-ABI library ``libmr_recon_amd.so``.

Every compute entry point takes raw device pointers plus the HIP stream that
PyTorch is currently using, so launches order correctly with torch work and can
be captured in a HIP graph. There is no CPU fallback: calling an op without the
library, or with tensors that are not on a ROCm device, raises.
"""

from __future__ import annotations

import ctypes as C
import threading
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libmr_recon_amd.so"

_lock = threading.Lock()
_lib = None

_vp = C.c_void_p
_i = C.c_int
_ll = C.c_longlong
_f = C.c_float
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_llp = C.POINTER(C.c_longlong)

# name -> (restype, argtypes); mirrors include/mr_recon_amd.h
SIGNATURES = {
    "mra_version": (_i, []),
    "mra_last_error": (C.c_char_p, []),
    "mra_geom_size": (_i, []),
    "mra_set_grid_kernel": (_i, [_i]),
    "mra_geom_set_dcf_plan_order": (_i, [_vp, _i]),
    "mra_geom_set_ksp_plan_order": (_i, [_vp, _i]),
    "mra_default_tile": (_i, [_i]),
    "mra_tile": (_i, [_i]),
    "mra_set_tile": (_i, [_i, _i]),
    "mra_grad_samples_per_block": (_i, []),
    "mra_subgrid_points": (_ll, [_i, _i]),
    "mra_max_coil_chunk": (_i, [_i, _i]),
    "mra_geom_fill": (_i, [_vp, _ll, _ll, _ll, _ll, _ll, _i, _i, _i, _ip, _i, _ip, _f, _f, _f, _i, _ll, _i]),
    "mra_plan_weights": (_i, [_i, _i, _vp, _i, _ip, _ip, _f, _f, _vp, _vp, _vp]),
    "mra_plan_weights_dims": (_i, [_i, _i, _vp, _i, _ip, _ip, _f, _fp, _vp, _vp, _vp]),
    "mra_scan_workspace": (_ll, [_ll]),
    "mra_scan_excl_i32": (_i, [_vp, _ll, _vp, _vp, _vp]),
    "mra_bin_workspace": (_ll, [_i, _ll, _ll, _i, _i]),
    "mra_bin_build": (_i, [_i, _vp, _ll, _ip, _fp, _fp, _i, _f, _i, _ip, _i, _vp, _vp, _vp, _vp, _vp]),
    "mra_bin_items": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp]),
    "mra_interp": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "mra_grid_partial": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp]),
    "mra_grid_reduce": (_i, [_i, _i, _vp, _vp, _vp, _i, _i, _vp, _i, _i, _i, _vp]),
    "mra_grid_noalloc": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp, _i, _vp, _i, _vp]),
    "mra_grid_noalloc_max_bytes": (_ll, []),
    "mra_interp_grad": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "mra_pad_expand": (_i, [_i, _ip, _ip, _ip, _vp, _ll, _ll, _vp, _ll, _ll, _vp, _f, _vp, _vp]),
    "mra_crop_reduce": (_i, [_i, _ip, _ip, _ip, _vp, _ll, _vp, _ll, _ll, _i, _vp, _f, _i, _i, _vp, _ll, _vp]),
    "mra_pad_expand_batched": (_i, [_i, _ip, _ip, _ip, _vp, _ll, _vp, _ll, _vp, _vp, _f, _vp, _vp]),
    "mra_crop_reduce_batched": (_i, [_i, _ip, _ip, _ip, _vp, _ll, _vp, _vp, _vp, _f, _vp, _vp]),
    "mra_cmul_bcast": (_i, [_vp, _vp, _ll, _ll, _vp]),
    "mra_fft_plan": (_i, [_i, _ip, _i, C.POINTER(_vp)]),
    "mra_fft_exec": (_i, [_vp, _vp, _i, _vp]),
    "mra_fft_destroy": (_i, [_vp]),
    "mra_fft_lds_supported": (_i, [_i, _ip]),
    "mra_fft_lds_twiddles": (_i, [_i, _fp]),
    "mra_fft_lds": (_i, [_i, _ip, _ll, _vp, C.POINTER(_vp), _i, _vp]),
    "mra_fft_fused_supported": (_i, [_i, _ip]),
    "mra_fft_pad_forward": (_i, [_i, _ip, _ip, _ip, _vp, _ll, _vp, _ll, _ll, _vp, _ll, _vp, _f, _vp, C.POINTER(_vp),
                                 _vp]),
    "mra_fft_crop_adjoint": (_i, [_i, _ip, _ip, _ip, _vp, _ll, _ll, _vp, _ll, _ll, _vp, _i, _vp, _f, _i, _vp, _ll,
                                  C.POINTER(_vp), _vp]),
    "mra_toeplitz_fused": (_i, [_i, _ip, _ip, _ip, _vp, _vp, _ll, _ll, _vp, _vp, _i, _vp, C.POINTER(_vp), _vp]),
    "mra_cg_partials": (_i, []),
    "mra_cg_seg_blocks": (_i, []),
    "mra_cdot_re": (_i, [_vp, _vp, _ll, _f, _vp, _vp, _i, _vp]),
    "mra_cg_init": (_i, [_vp, _vp, _vp, _vp, _vp, _ll, _f, _vp, _vp, _vp]),
    "mra_cg_step": (_i, [_vp, _vp, _vp, _vp, _ll, _f, _vp, _vp, _vp]),
    "mra_cg_step_fwd": (_i, [_vp, _vp, _vp, _vp, _vp, _ll, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mra_cg_step_bwd": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _ll, _f, _f, _vp, _vp, _vp, _vp, _vp, _vp,
                             _vp]),
    "mra_cg_init_seg": (_i, [_vp, _vp, _vp, _vp, _vp, _ll, _i, _f, _vp, _vp, _vp]),
    "mra_cg_step_seg": (_i, [_vp, _vp, _vp, _vp, _ll, _i, _f, _vp, _vp, _vp]),
    "mra_nn_gather": (_i, [_vp, _ll, _vp, _ll, _i, _f, _vp, _vp]),
    "mra_nn_scatter": (_i, [_vp, _vp, _ll, _i, _f, _vp, _ll, _vp]),
    "mra_nn_scatter_chunks": (_i, [_vp, _vp, _vp, _vp, _vp, _ll, _ll, _i, _f, _vp, _ll, _vp]),
    "mra_nn_gather_sum": (_i, [_vp, _ll, _vp, _vp, _i, _ll, _i, _f, _vp, _vp]),
    "mra_nn_scatter_sum_chunks": (_i, [_vp, _vp, _i, _vp, _vp, _vp, _vp, _ll, _ll, _i, _f, _vp, _ll, _vp]),
    "mra_interp_direct": (_i, [_i, _i, _vp, _vp, _ll, _i, _ip, _fp, _f, _f, _vp, _vp]),
    "mra_grid_direct": (_i, [_i, _i, _vp, _vp, _ll, _i, _ip, _fp, _f, _f, _vp, _vp]),
    "mra_lowrank_expand": (_i, [_vp, _vp, _i, _i, _ll, _vp, _vp]),
    "mra_lowrank_reduce": (_i, [_vp, _vp, _vp, _i, _i, _ll, _vp, _vp]),
    "mra_dcf_partials": (_i, []),
    "mra_dcf_update": (_i, [_vp, _vp, _ll, _vp]),
    "mra_dcf_normalize": (_i, [_vp, _vp, _vp, _ll, _vp, _vp]),
}


def lib():
    """Load (once) and return the ctypes library handle."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"HIP extension not built ({LIB_PATH} missing); run `python -m mr_recon_amd.csrc.build`"
                )
            h = C.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(h, name)
                fn.restype = res
                fn.argtypes = args
            _lib = h
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise RuntimeError(f"mr_recon_amd HIP call failed: {lib().mra_last_error().decode()}")


def require_device(*tensors: torch.Tensor) -> None:
    """The HIP path only runs on ROCm device tensors; there is no CPU fallback."""
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise RuntimeError(
                "mr_recon_amd kernels run on the GPU only; got a tensor on "
                f"{t.device}. Move inputs to a ROCm device (e.g. 'cuda:0')."
            )


def ptr(t: torch.Tensor | None):
    """Device address of ``t``'s data. Lazy conjugate / negative views (``x.conj()``)
    share the data of ``x``, so their address would hand the kernels the unconjugated
    values: they are refused."""
    if t is None:
        return None
    if t.is_conj() or t.is_neg():
        raise ValueError("a lazy conj()/neg() view cannot be passed to a kernel; call resolve_conj() first")
    return _vp(t.data_ptr())


def stream_ptr(device: torch.device | None = None):
    return _vp(torch.cuda.current_stream(device).cuda_stream)


class tile_scope:
    """Select the kernels' tile edge for ndim-dimensional calls of this thread (the C
    library keeps one per thread) and restore the previous one on exit."""

    __slots__ = ("ndim", "tile", "prev")

    def __init__(self, ndim: int, tile: int):
        self.ndim, self.tile = int(ndim), int(tile)

    def __enter__(self):
        self.prev = lib().mra_set_tile(self.ndim, self.tile)
        if self.prev < 0:
            raise ValueError(lib().mra_last_error().decode())
        return self

    def __exit__(self, *exc):
        lib().mra_set_tile(self.ndim, self.prev)
        return False


def ints(vals) -> C.Array:
    vals = [int(v) for v in vals]
    return (C.c_int * len(vals))(*vals)


def floats(vals) -> C.Array:
    vals = [float(v) for v in vals]
    return (C.c_float * len(vals))(*vals)
