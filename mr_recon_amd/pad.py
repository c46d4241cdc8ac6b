"""Centred even zero-pad / crop of the trailing dims (reference pad.py:5-37)."""

import torch
import torch.nn as nn
import torch.nn.functional as F


class PadLast(nn.Module):
    def __init__(self, pad_im_size, im_size):
        super().__init__()
        assert len(pad_im_size) == len(im_size)
        self.im_dim = len(im_size)
        self.im_size = tuple(im_size)
        self.pad_im_size = tuple(pad_im_size)
        for psz in pad_im_size:
            assert not (psz % 2), "Pad sizes must be even"
        # (P - N)//2 on both sides of every dim; F.pad lists the last dim first
        lr = [(p - n) // 2 for p, n in zip(self.pad_im_size, self.im_size)]
        self.pad = [v for q in reversed(lr) for v in (q, q)]
        self.crop_slice = [slice(q, -q) if q > 0 else slice(None) for q in lr]

    @property
    def center(self):
        """Pixel that lands on index P//2 of the padded axis (the FFT centre)."""
        return tuple(p // 2 - (p - n) // 2 for p, n in zip(self.pad_im_size, self.im_size))

    def forward(self, x):
        assert tuple(x.shape[-self.im_dim:]) == self.im_size
        return F.pad(x, self.pad + [0, 0] * (x.ndim - self.im_dim))

    def adjoint(self, y):
        assert tuple(y.shape[-self.im_dim:]) == self.pad_im_size
        return y[(slice(None),) * (y.ndim - self.im_dim) + tuple(self.crop_slice)]
