"""Host-side utilities used on the NUFFT path.

``resize`` (reference utils.py:41-84), ``gen_grd`` (utils.py:166-200) and
``batch_iterator`` (utils.py:364-382) keep the reference's definitions; the
conversions mirror ``np_to_torch`` / ``torch_to_np`` (utils.py:300-362)
without the CuPy branch.
"""

from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .dtypes import real_dtype


def _expand_shapes(*shapes):
    shapes = [list(s) for s in shapes]
    nd = max(len(s) for s in shapes)
    return tuple([1] * (nd - len(s)) + s for s in shapes)


def resize(input: torch.Tensor, oshape, ishift=None, oshift=None) -> torch.Tensor:
    """Centred zero-pad or crop of every axis to ``oshape``.

    The input block starting at ishift = max(i//2 - o//2, 0) is copied to the
    output block starting at oshift = max(o//2 - i//2, 0).
    """
    ishape1, oshape1 = _expand_shapes(input.shape, oshape)
    if ishape1 == oshape1:
        return input.reshape(oshape)
    if ishift is None:
        ishift = [max(i // 2 - o // 2, 0) for i, o in zip(ishape1, oshape1)]
    if oshift is None:
        oshift = [max(o // 2 - i // 2, 0) for i, o in zip(ishape1, oshape1)]
    copy = [min(i - si, o - so) for i, si, o, so in zip(ishape1, ishift, oshape1, oshift)]
    isl = tuple(slice(si, si + c) for si, c in zip(ishift, copy))
    osl = tuple(slice(so, so + c) for so, c in zip(oshift, copy))
    out = torch.zeros(oshape1, dtype=input.dtype, device=input.device)
    out[osl] = input.reshape(ishape1)[isl]
    return out.reshape(oshape)


def gen_grd(im_size: tuple, fovs: Optional[tuple] = None, balanced: Optional[bool] = False) -> torch.Tensor:
    """Pixel-centre grid of shape (*im_size, d): r_i = (n_i - N_i//2) / N_i (times fov_i)."""
    if fovs is None:
        fovs = (1,) * len(im_size)
    if balanced:
        lins = [fovs[i] * torch.linspace(-0.5, 0.5, im_size[i]) for i in range(len(im_size))]
    else:
        lins = [fovs[i] * torch.arange(-(im_size[i] // 2), im_size[i] // 2 + (im_size[i] % 2)) / im_size[i]
                for i in range(len(im_size))]
    grds = torch.meshgrid(*lins, indexing="ij")
    return torch.stack(grds, dim=-1).type(real_dtype)


def batch_iterator(total: int, batch_size: Optional[int]):
    """(start, stop) pairs covering range(total) in steps of batch_size."""
    assert total > 0, f"batch_iterator called with {total} elements"
    if batch_size is None:
        return [(0, total)]
    edges = list(range(0, total, batch_size)) + [total]
    return list(zip(edges[:-1], edges[1:]))


def np_to_torch(*args):
    out = [torch.as_tensor(a) if isinstance(a, np.ndarray) else (a if isinstance(a, torch.Tensor) else None)
           for a in args]
    return out[0] if len(out) == 1 else out


def torch_to_np(*args):
    out = []
    for a in args:
        if isinstance(a, torch.Tensor):
            out.append(a.detach().cpu().numpy())
        elif isinstance(a, np.ndarray):
            out.append(a)
        else:
            out.append(None)
    return out[0] if len(out) == 1 else out
