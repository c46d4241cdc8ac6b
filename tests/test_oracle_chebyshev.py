"""Pins the oracle's Chebyshev NUFFT (fourier.py:1574-1774): on samples away from the upper
k-space edge (where the reference clamps the grid index, :1676-1680) it converges to the FP64
direct NDFT as the expansion order grows, with the reference's output scale grid_oversamp *
G^(-d/2) = grid_oversamp^(1 - d/2) N^(-d/2); forward and adjoint are adjoint to each other."""

import numpy as np
import pytest
import torch

from helpers import cdot
from mr_recon_amd import synth
from oracle import ndft
from oracle.nufft_ref import ChebyshevNufftRef


def _interior_k(im, os_, n, seed=0):
    rng = np.random.default_rng(seed)
    G = np.array([round(i * os_) for i in im])
    k = ((rng.random((n, len(im))) - 0.5) * np.array(im)).astype(np.float32)
    return k[np.all(k * os_ + G / 2 < G - 0.5, axis=1)]


@pytest.mark.parametrize("im,os_", [((32, 32), 1.0), ((32, 32), 1.25), ((16, 12, 8), 2.0)])
def test_chebyshev_oracle_converges_to_ndft(im, os_):
    k = _interior_k(im, os_, 1200)
    x = synth.random_image(im, seed=1)
    yex = ndft.ndft_forward(x.numpy(), k, im) * os_ ** (1 - len(im) / 2)
    errs = []
    for n in (3, 5, 7, 9):
        R = ChebyshevNufftRef(im, n, os_, dtype=torch.float64)
        t = R.rescale_trajectory(torch.from_numpy(k).to(torch.float64).clone())
        y = R.forward(x.to(torch.complex128)[None], t[None])[0].numpy()
        errs.append(np.linalg.norm(y - yex) / np.linalg.norm(yex))
    assert all(b < a / 10 for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-6


def test_chebyshev_oracle_adjoint_identity_and_edge_clamp():
    im, os_ = (16, 16), 1.0
    k = torch.from_numpy(((np.random.default_rng(2).random((400, 2)) - 0.5) * 16).astype(np.float32))
    k[0] = torch.tensor([7.9, -8.0])  # rounds past the last grid index: clamped to G - 1
    R = ChebyshevNufftRef(im, 5, os_, dtype=torch.float64)
    t = R.rescale_trajectory(k.to(torch.float64).clone())
    assert int(t[0, 0]) == 15 and int(t[0, 1]) == 0
    x = synth.random_image((2, *im), seed=3).to(torch.complex128)
    y = synth.random_image((2, 400), seed=4).to(torch.complex128)
    a = R.forward(x[None], t[None])[0]
    b = R.adjoint(y[None], t[None])[0]
    assert abs(cdot(y, a) - cdot(b, x)) < 1e-12 * abs(cdot(y, a))
