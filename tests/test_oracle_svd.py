"""Pins the oracle's SVD NUFFT (fourier.py:1328-1572) and checks the host side of the product's:
  * the oracle converges to the FP64 direct NDFT as the rank grows (complete groups of equal
    eigenvalues: 3, 6, 10, 15 in 2D on a 30 x 30 model);
  * the product's spline interpolation equals scipy.ndimage.map_coordinates(order=3,
    mode='nearest') (the reference's cupyx call);
  * the product's Kronecker eigendecomposition spans the same model as the oracle's full
    eigendecomposition: sum_l h_l(k) b_l(r) agrees."""

import numpy as np
import pytest
import torch
from scipy.ndimage import map_coordinates

from mr_recon_amd import synth
from mr_recon_amd.fourier import svd_nufft
from oracle import ndft
from oracle.nufft_ref import SvdNufftRef


def test_svd_oracle_converges_to_ndft():
    im = (32, 32)
    k = (np.random.default_rng(0).random((1000, 2)) - 0.5) * 32
    x = synth.random_image(im, seed=1)
    yex = ndft.ndft_forward(x.numpy(), k, im)
    errs = []
    for n in (3, 6, 10, 15):
        R = SvdNufftRef(im, 1.0, n, svd_mx_size=(30, 30), dtype=torch.float64)
        t = R.rescale_trajectory(torch.from_numpy(k).clone())
        y = R.forward(x.to(torch.complex128)[None], t[None])[0].numpy()
        errs.append(np.linalg.norm(y - yex) / np.linalg.norm(yex))
    assert all(b < a / 2 for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-2


def test_spatial_interp_matches_map_coordinates():
    rng = np.random.default_rng(3)
    a = rng.standard_normal((2, 9, 7, 5)) + 1j * rng.standard_normal((2, 9, 7, 5))
    c = np.stack([rng.random(400) * 10 - 0.5, rng.random(400) * 8 - 0.5, rng.random(400) * 6 - 0.5], -1)
    got = svd_nufft.spatial_interp(torch.from_numpy(a), torch.from_numpy(c)).numpy()
    ref = np.stack([map_coordinates(a[i].real, c.T, order=3, mode="nearest")
                    + 1j * map_coordinates(a[i].imag, c.T, order=3, mode="nearest") for i in range(2)])
    assert np.abs(got - ref).max() < 1e-6 * np.abs(ref).max()
    with pytest.raises(NotImplementedError):
        svd_nufft.spatial_interp(torch.from_numpy(a), torch.from_numpy(c), order=1)


@pytest.mark.parametrize("im,n,ss", [((32, 32), 6, (30, 30)), ((12, 10, 8), 4, (8, 8, 8))])
def test_kronecker_model_matches_full_eigendecomposition(im, n, ss):
    k = torch.from_numpy((np.random.default_rng(1).random((600, len(im))) - 0.5) * np.array(im)).to(torch.float32)
    P, R = svd_nufft(im, 1.0, n, svd_mx_size=ss), SvdNufftRef(im, 1.0, n, svd_mx_size=ss)
    t, tr = P.rescale_trajectory(k.clone()), R.rescale_trajectory(k.clone())
    assert torch.equal(t, tr)
    A = torch.einsum("lm,lr->mr", P.h.to(torch.complex128), P.b.reshape(n, -1).to(torch.complex128))
    B = torch.einsum("lm,lr->mr", R.h.to(torch.complex128), R.b.reshape(n, -1).to(torch.complex128))
    assert float((A - B).norm() / B.norm()) < 1e-6
