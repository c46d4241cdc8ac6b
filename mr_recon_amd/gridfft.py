"""In-place FFTs of the oversampled grids (unnormalised, trailing dims).

One LDS-staged Stockham pass per axis for grid lengths with a compiled radix plan
(`mra_fft_lds`), rocFFT through hipFFT otherwise (`mra_fft_exec`). The fused
pad+FFT / iFFT+crop transforms (`ops.pad_fft`, `ops.ifft_crop`) use the same
twiddle tables.
"""

from __future__ import annotations

import ctypes
import math
import threading

import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr


class _FFTPlans:
    """hipFFT plans for in-place batched C2C transforms, per (device, batch, grid shape)."""

    def __init__(self):
        self._plans: dict = {}
        self._lock = threading.Lock()

    def get(self, dev_index: int, batch: int, shape: tuple):
        key = (dev_index, batch, shape)
        with self._lock:
            h = self._plans.get(key)
            if h is None:
                h = ctypes.c_void_p()
                with torch.cuda.device(dev_index):
                    check(lib().mra_fft_plan(len(shape), _lib.ints(shape), batch, ctypes.byref(h)))
                self._plans[key] = h
            return h


_fft_plans = _FFTPlans()


class _LdsFFT:
    """Twiddle tables (device) and shape support of the LDS-staged transforms."""

    def __init__(self):
        self._tw: dict = {}
        self._ok: dict = {}
        self._lock = threading.Lock()

    def supported(self, shape: tuple) -> bool:
        ok = self._ok.get(shape)
        if ok is None:
            ok = bool(lib().mra_fft_lds_supported(len(shape), _lib.ints(shape)))
            self._ok[shape] = ok
        return ok

    def tables(self, device: torch.device, shape: tuple):
        key = (device.index, shape)
        with self._lock:
            arr = self._tw.get(key)
            if arr is None:
                tabs = []
                for L in shape:
                    host = torch.empty(2 * L, dtype=torch.float32)
                    check(lib().mra_fft_lds_twiddles(L, ctypes.cast(host.data_ptr(), ctypes.POINTER(ctypes.c_float))))
                    tabs.append(host.to(device))
                arr = (tabs, (ctypes.c_void_p * len(shape))(*[t.data_ptr() for t in tabs]))
                self._tw[key] = arr
            return arr[1]


_lds_fft = _LdsFFT()


def _fft_inplace(x: torch.Tensor, d: int, inverse: bool) -> torch.Tensor:
    _lib.require_device(x)
    if x.dtype != torch.complex64 or not x.is_contiguous():
        raise TypeError("grid FFTs run in place on contiguous complex64 tensors")
    shape = tuple(int(s) for s in x.shape[-d:])
    batch = int(math.prod(x.shape[:-d]))
    if batch == 0:
        return x
    if _lds_fft.supported(shape):
        tw = _lds_fft.tables(x.device, shape)
        check(lib().mra_fft_lds(d, _lib.ints(shape), batch, ptr(x), tw, int(inverse), stream_ptr(x.device)))
        return x
    plan = _fft_plans.get(x.device.index, batch, shape)
    check(lib().mra_fft_exec(plan, ptr(x), int(inverse), stream_ptr(x.device)))
    return x


def _fftn(x: torch.Tensor, d: int) -> torch.Tensor:
    """Unnormalised forward FFT over the last d dims, IN PLACE on the caller's temporary
    grid ``x`` (returned): one LDS-staged pass per axis for lengths with a compiled radix
    plan, rocFFT (hipFFT) otherwise. An out-of-place transform that preserves its input
    costs a full copy of the grid first."""
    return _fft_inplace(x, d, False)


def _ifftn(x: torch.Tensor, d: int) -> torch.Tensor:
    """Unnormalised inverse FFT (no 1/n) over the last d dims, in place on ``x``."""
    return _fft_inplace(x, d, True)
