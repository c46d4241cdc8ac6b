"""GPU parity of chebyshev_nufft and svd_nufft (fused pad+FFT over the L spatial functions, weighted
nearest-grid gather / chunked scatter, fused iFFT+crop with the conj(b) sum) against the oracle
restatement of fourier.py:1574-1774, including samples clamped at the upper k-space edge and
grid points hit by many samples (radial centres)."""

import numpy as np
import pytest
import torch

from helpers import cdot, rel_l2
from mr_recon_amd import synth
from oracle.nufft_ref import ChebyshevNufftRef

pytestmark = pytest.mark.gpu

CASES = [  # (im_size, n_cheby_per_dim, grid_oversamp, trajectory)
    ((32, 32), 5, 1.0, lambda: (torch.rand((900, 2)) - 0.5) * 32),
    ((40, 40), 4, 1.25, lambda: synth.radial_2d(60, 80, 40).reshape(-1, 2)),
    ((16, 12, 8), 3, 2.0, lambda: (torch.rand((700, 3)) - 0.5) * torch.tensor([16.0, 12.0, 8.0])),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_chebyshev_vs_oracle(dev, case):
    from mr_recon_amd.fourier import chebyshev_nufft

    im, n, os_, mk = CASES[case]
    torch.manual_seed(case)
    k = mk().to(torch.float32)
    k[0] = torch.tensor([im[0] / 2 - 0.01] + [-(i / 2) for i in im[1:]])  # clamped at the edge
    hn, rn = chebyshev_nufft(im, n, os_), ChebyshevNufftRef(im, n, os_)
    t = hn.rescale_trajectory(k.clone().to(dev))
    tr = rn.rescale_trajectory(k.clone())
    assert torch.equal(t.cpu(), tr)
    assert rel_l2(hn.h, rn.h) < 1e-5 and rel_l2(hn.b, rn.b) < 1e-5
    x = synth.random_image((1, 2, *im), seed=case)
    y = hn.forward(x.to(dev), t[None])
    y_ref = rn.forward(x, tr[None])
    assert y.shape == y_ref.shape == (1, 2, k.shape[0])
    assert rel_l2(y, y_ref) < 1e-5
    yy = synth.random_image(tuple(y.shape), seed=case + 7)
    z = hn.adjoint(yy.to(dev), t[None])
    z_ref = rn.adjoint(yy, tr[None])
    assert z.shape == z_ref.shape
    assert rel_l2(z, z_ref) < 1e-5
    assert abs(cdot(yy, y) - cdot(z, x)) < 1e-5 * abs(cdot(yy, y))


SVD_CASES = [  # (im_size, grid_oversamp, n_svd (complete eigenvalue groups), svd_mx_size)
    ((32, 32), 1.0, 6, None),
    ((40, 40), 1.25, 3, (20, 20)),
    ((12, 10, 8), 1.0, 4, (8, 8, 8)),
]


@pytest.mark.parametrize("case", range(len(SVD_CASES)))
def test_svd_nufft_vs_oracle(dev, case):
    from mr_recon_amd.fourier import svd_nufft
    from oracle.nufft_ref import SvdNufftRef

    im, os_, n, ss = SVD_CASES[case]
    torch.manual_seed(case)
    k = ((torch.rand((800, len(im))) - 0.5) * torch.tensor(im, dtype=torch.float32))
    if case == 1:
        k = synth.radial_2d(50, 80, 40).reshape(-1, 2)  # long runs at the centre
    hn, rn = svd_nufft(im, os_, n, svd_mx_size=ss), SvdNufftRef(im, os_, n, svd_mx_size=ss)
    t = hn.rescale_trajectory(k.clone().to(dev))
    tr = rn.rescale_trajectory(k.clone())
    assert torch.equal(t.cpu(), tr)
    x = synth.random_image((1, 2, *im), seed=case)
    y = hn.forward(x.to(dev), t[None])
    y_ref = rn.forward(x, tr[None])
    assert rel_l2(y, y_ref) < 1e-5
    yy = synth.random_image(tuple(y.shape), seed=case + 3)
    z = hn.adjoint(yy.to(dev), t[None])
    assert rel_l2(z, rn.adjoint(yy, tr[None])) < 1e-5
    assert abs(cdot(yy, y) - cdot(z, x)) < 1e-5 * abs(cdot(yy, y))
