"""GPU parity of type3_nufft (image-domain gridding with per-dimension betas, inner
sigpy_nufft, phase / apodization factors on the HIP kernels) against the oracle."""

import pytest
import torch

from helpers import cdot, rel_l2
from oracle import type3_ref

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _case(D, n, T, seed):
    g = torch.Generator().manual_seed(seed)
    axes = torch.meshgrid(*([torch.arange(n) - n / 2] * D), indexing="ij")
    phis = torch.stack(axes, 0) / n + 0.05 * torch.randn((D, *(n,) * D), generator=g)
    alphas = (torch.rand((D, T), generator=g) - 0.5) * n
    x = torch.randn((2, *(n,) * D), dtype=torch.complex64, generator=g)
    return phis, alphas, x


@pytest.mark.parametrize("D,n,T,w", [(2, 16, 500, 4.0), (2, 12, 300, 6.0), (1, 40, 120, 4.0), (3, 8, 400, 4.0)])
def test_type3_vs_oracle(dev, D, n, T, w):
    from mr_recon_amd.linops import type3_nufft

    phis, alphas, x = _case(D, n, T, seed=D + int(w))
    A = type3_nufft(phis.to(dev), alphas.to(dev), 2.0, w)
    R = type3_ref.Type3Ref(phis, alphas, 2.0, w)
    assert A.grd_N_os == R.N_os
    y = A.forward(x.to(dev))
    assert y.shape == (2, T)
    assert rel_l2(y, R.forward(x)) < TOL
    yy = torch.randn((2, T), dtype=torch.complex64, generator=torch.Generator().manual_seed(2))
    xa = A.adjoint(yy.to(dev))
    assert rel_l2(xa, R.adjoint(yy)) < TOL
    assert abs(cdot(yy, y) - cdot(xa, x)) < 1e-5 * abs(cdot(yy, y))
    assert rel_l2(y, type3_ref.naive_forward(phis, alphas, x)) < 2e-3


def test_type3_toeplitz_and_naive(dev):
    from mr_recon_amd.linops import type3_nufft, type3_nufft_naive

    phis, alphas, x = _case(2, 12, 300, seed=3)
    A = type3_nufft(phis.to(dev), alphas.to(dev), 2.0, 4.0, use_toep=True)
    R = type3_ref.Type3Ref(phis, alphas, 2.0, 4.0, use_toep=True)
    assert rel_l2(A.normal(x.to(dev)), R.normal(x)) < TOL
    Nv = type3_nufft_naive(phis.to(dev), alphas.to(dev))
    assert rel_l2(Nv.forward(x.to(dev)), type3_ref.naive_forward(phis, alphas, x)) < 1e-4


@pytest.mark.parametrize("B,T", [(4, 200), (5, 150)])
def test_type3_four_and_five_bases(dev, B, T):
    """4 and 5 bases: the image-domain gridding and the inner sigpy_nufft run in 4D / 5D
    (direct KB kernels, torch FFTs) -- against the oracle and the exact sum."""
    from mr_recon_amd.linops import type3_nufft

    g = torch.Generator().manual_seed(B)
    im = (10, 10) if B == 4 else (6, 6)
    phis = torch.randn((B, *im), generator=g) * 0.3
    alphas = (torch.rand((B, T), generator=g) - 0.5) * (6 if B == 4 else 4)
    x = torch.randn((2, *im), dtype=torch.complex64, generator=g)
    A = type3_nufft(phis.to(dev), alphas.to(dev), 2.0, 4.0)
    R = type3_ref.Type3Ref(phis, alphas, 2.0, 4.0)
    assert A.grd_N_os == R.N_os
    y = A.forward(x.to(dev))
    assert rel_l2(y, R.forward(x)) < TOL
    yy = torch.randn((2, T), dtype=torch.complex64, generator=g)
    xa = A.adjoint(yy.to(dev))
    assert rel_l2(xa, R.adjoint(yy)) < TOL
    assert abs(cdot(yy, y) - cdot(xa, x)) < 1e-5 * abs(cdot(yy, y))
    assert rel_l2(y, type3_ref.naive_forward(phis, alphas, x)) < 3e-3
