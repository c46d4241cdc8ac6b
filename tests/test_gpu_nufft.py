"""GPU parity: hip_nufft (fused and split methods, Toeplitz) vs the reference restated on CPU."""

import numpy as np
import pytest
import torch

from helpers import cdot, rel_l2
from mr_recon_amd import synth
from oracle import ndft
from oracle.nufft_ref import TritonNufftRef

pytestmark = pytest.mark.gpu

TOL = 1e-5

CASES = [
    # (im_size, oversamp, width, trajectory)
    ((64, 64), 2.0, 6, lambda: synth.radial_2d(40, 128, 64)),
    ((48, 48), 1.25, 4, lambda: synth.spiral_2d(8, 600, 48)),
    ((40, 36), 1.3, 5, lambda: (torch.rand((2000, 2)) - 0.5) * torch.tensor([40.0, 36.0])),
    ((20, 20, 20), 1.25, 4, lambda: synth.kooshball_3d(60, 40, 20)),
    ((16, 18, 14), 2.0, 6, lambda: (torch.rand((1500, 3)) - 0.5) * torch.tensor([16.0, 18.0, 14.0])),
]


def _nuffts(im, os_, w):
    from mr_recon_amd.fourier import hip_nufft

    return hip_nufft(im, oversamp=os_, width=w), TritonNufftRef(im, oversamp=os_, width=w)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_forward_adjoint_vs_reference(dev, case):
    im, os_, w, mk = CASES[case]
    torch.manual_seed(case)
    trj = mk().reshape(-1, len(im))[None]
    img = synth.random_image((1, 2, *im), seed=case)
    hn, rn = _nuffts(im, os_, w)
    y = hn.forward(img.to(dev), trj.to(dev))
    y_ref = rn.forward(img, trj)
    assert y.shape == y_ref.shape
    assert rel_l2(y, y_ref) < TOL
    ksp = synth.random_image(y_ref.shape, seed=case + 50)
    x = hn.adjoint(ksp.to(dev), trj.to(dev))
    x_ref = rn.adjoint(ksp, trj)
    assert x.shape == x_ref.shape
    assert rel_l2(x, x_ref) < TOL
    # adjoint identity on the GPU operator itself
    assert abs(cdot(ksp, y) - cdot(x, img)) / abs(cdot(ksp, y)) < 1e-5


@pytest.mark.parametrize("case", [0, 3])
def test_split_methods_vs_reference(dev, case):
    im, os_, w, mk = CASES[case]
    trj = mk().reshape(-1, len(im))[None]
    img = synth.random_image((1, *im), seed=3)
    hn, rn = _nuffts(im, os_, w)
    g = hn.forward_FT_only(img.to(dev))
    g_ref = rn.forward_FT_only(img)
    assert rel_l2(g, g_ref) < TOL
    y = hn.forward_interp_only(g_ref.to(dev), trj.to(dev))
    assert rel_l2(y, rn.forward_interp_only(g_ref, trj)) < TOL
    ksp = synth.random_image(y.shape, seed=4)
    gg = hn.adjoint_grid_only(ksp.to(dev), trj.to(dev))
    gg_ref = rn.adjoint_grid_only(ksp, trj)
    assert rel_l2(gg, gg_ref) < TOL
    assert rel_l2(hn.adjoint_iFT_only(gg_ref.to(dev)), rn.adjoint_iFT_only(gg_ref)) < TOL


@pytest.mark.parametrize("os_,w,bound", [(2.0, 6, 3e-5), (1.25, 4, 1.2e-2), (1.25, 6, 1e-3), (2.0, 4, 1.5e-3)])
def test_ndft_accuracy(dev, os_, w, bound):
    """NUFFT vs FP64 direct NDFT; bounds follow the KB approximation error of each setting."""
    from mr_recon_amd.fourier import hip_nufft

    im = (48, 48)
    rng = np.random.default_rng(0)
    k = ((rng.random((3000, 2)) - 0.5) * 48).astype(np.float32)
    x = synth.random_image(im, seed=1)
    hn = hip_nufft(im, oversamp=os_, width=w)
    y = hn.forward(x[None].to(dev), torch.from_numpy(k)[None].to(dev))[0].cpu().numpy()
    y_ex = ndft.ndft_forward(x.numpy(), k, im)
    err = np.linalg.norm(y - y_ex) / np.linalg.norm(y_ex)
    assert err < bound
    z = hn.adjoint(torch.from_numpy(y_ex.astype(np.complex64))[None].to(dev), torch.from_numpy(k)[None].to(dev))
    z_ex = ndft.ndft_adjoint(y_ex, k, im)
    assert np.linalg.norm(z[0].cpu().numpy() - z_ex) / np.linalg.norm(z_ex) < bound


@pytest.mark.parametrize("case", [0, 1, 3])
def test_toeplitz_vs_reference(dev, case):
    im, os_, w, mk = CASES[case]
    trj = mk().reshape(-1, len(im))[None]
    hn, rn = _nuffts(im, os_, w)
    g = torch.Generator().manual_seed(2)
    wts = torch.rand((1, trj.shape[1]), generator=g) + 0.5
    K = hn.calc_teoplitz_kernels(trj.to(dev), wts.to(dev))
    K_ref = rn.calc_teoplitz_kernels(trj, wts)
    assert rel_l2(K, K_ref) < TOL
    img = synth.random_image((1, 2, *im), seed=5)
    out = hn.normal_toeplitz(img.to(dev), K_ref.to(dev))
    assert rel_l2(out, rn.normal_toeplitz(img, K_ref)) < TOL


def test_toeplitz_approximates_normal(dev):
    """Toeplitz normal op vs adjoint(forward) with the same weights (pattern of the
    reference's tests/test_nufft_scaling.py:87-100, tighter tolerance)."""
    from mr_recon_amd.fourier import hip_nufft

    im = (40, 40)
    trj = synth.radial_2d(30, 80, 40).reshape(1, -1, 2)
    hn = hip_nufft(im, oversamp=2.0, width=6)
    x = synth.random_image((1, *im), seed=9)
    K = hn.calc_teoplitz_kernels(trj.to(dev))
    a = hn.normal_toeplitz(x[:, None].to(dev), K)[:, 0]
    b = hn.adjoint(hn.forward(x.to(dev), trj.to(dev)), trj.to(dev))
    assert rel_l2(a, b) < 2e-3


@pytest.mark.parametrize("tile", [8, 32, "auto"])
def test_hip_nufft_tile_choice(dev, tile):
    """hip_nufft with a 2D tile edge of 8 or 32, or tile and work-item size both chosen by
    the autotuner: forward and adjoint against the oracle, and the chosen edge recorded."""
    from mr_recon_amd import autotune
    from mr_recon_amd.fourier import hip_nufft

    im = (64, 64)
    trj = synth.radial_2d(40, 128, 64)
    x = synth.random_image((1, 3, *im), seed=4)
    nf = hip_nufft(im, oversamp=2.0, width=6, tile=tile, samples_per_item="auto" if tile == "auto" else None)
    ref = TritonNufftRef(im, oversamp=2.0, width=6)
    y = nf.forward(x.to(dev), trj[None].to(dev))
    assert rel_l2(y, ref.forward(x, trj[None])) < TOL
    z = nf.adjoint(y, trj[None].to(dev))
    assert rel_l2(z, ref.adjoint(ref.forward(x, trj[None]), trj[None])) < TOL
    plan = nf.plan(trj.reshape(-1, 2).to(dev))
    assert plan.tile in autotune.TILES[2]
    if tile != "auto":
        assert plan.tile == tile
