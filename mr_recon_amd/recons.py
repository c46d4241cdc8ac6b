"""CG-SENSE reconstruction driver (reference recons.py:77-156)."""

from __future__ import annotations

import gc
import time
from typing import Optional

import torch

from .algs import conjugate_gradient, power_method_operator
from .dtypes import complex_dtype
from .linops import linop


def CG_SENSE_recon(A: linop, ksp: torch.Tensor, max_iter: Optional[int] = 15, lamda_l2: Optional[float] = 0.0,
                   max_eigen: Optional[float] = None, tolerance: Optional[float] = 1e-8,
                   weights: Optional[torch.Tensor] = None, ahb_init: Optional[torch.Tensor] = None,
                   clear_gpu_mem: Optional[bool] = True, verbose: Optional[bool] = True) -> torch.Tensor:
    """recon = (AHA/max_eigen + lamda_l2 I)^-1 AHb/sqrt(max_eigen), started at AHb."""
    device = ksp.device
    if max_eigen is None:
        x0 = torch.randn(A.ishape, dtype=complex_dtype, device=device)
        _, max_eigen = power_method_operator(A.normal, x0, verbose=verbose)
        max_eigen *= 1.01
    if ahb_init is None:
        start = time.perf_counter()
        AHb = A.adjoint(ksp.type(complex_dtype)) / (max_eigen ** 0.5)
        if verbose:
            print(f"AHb took {time.perf_counter() - start:.3f}(s)")
    else:
        AHb = ahb_init.type(complex_dtype) / (max_eigen ** 0.5)
    if max_iter == 0:
        return AHb
    if clear_gpu_mem and device.type == "cuda":
        gc.collect()
        torch.cuda.empty_cache()
    inv = 1.0 / max_eigen

    def AHA(x):
        return A.normal(x) * inv

    return conjugate_gradient(AHA=AHA, AHb=AHb, num_iters=max_iter, lamda_l2=lamda_l2, tolerance=tolerance,
                              weights=weights, verbose=verbose)
