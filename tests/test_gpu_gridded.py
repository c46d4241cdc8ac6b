"""GPU parity of gridded_nufft (nearest-grid gather / scatter-add on the HIP kernels)
against the oracle, and exactness on its own lattice against the FP64 NDFT."""

import numpy as np
import pytest
import torch

from helpers import cdot, rel_l2
from mr_recon_amd import synth
from oracle import ndft
from oracle.nufft_ref import GriddedNufftRef, calc_toep_kernel_helper

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("im,os_", [((24, 20), 1.0), ((24, 20), 2.0), ((10, 12, 8), 1.5)])
def test_gridded_vs_oracle(dev, im, os_):
    from mr_recon_amd.fourier import gridded_nufft

    g = torch.Generator().manual_seed(len(im))
    k = (torch.rand((2, 600, len(im)), generator=g) - 0.5) * torch.tensor(im, dtype=torch.float32)
    hn, rn = gridded_nufft(im, os_), GriddedNufftRef(im, os_)
    t = hn.rescale_trajectory(k.clone().to(dev))
    tr = rn.rescale_trajectory(k.clone())
    assert torch.equal(t.cpu(), tr)
    x = synth.random_image((2, 3, *im), seed=1)
    y = hn.forward(x.to(dev), t)
    assert y.shape == (2, 3, 600)
    assert rel_l2(y, rn.forward(x, tr)) < TOL
    yy = synth.random_image((2, 3, 600), seed=2)
    z = hn.adjoint(yy.to(dev), t)
    assert z.shape == (2, 3, *im)
    assert rel_l2(z, rn.adjoint(yy, tr)) < TOL
    assert abs(cdot(yy, y) - cdot(z, x)) < 1e-5 * abs(cdot(yy, y))
    # the split methods (centred intermediate grid) agree with the fused transforms
    ys = hn.forward_interp_only(hn.forward_FT_only(x.to(dev)), t)
    assert rel_l2(ys, y) < TOL
    zs = hn.adjoint_iFT_only(hn.adjoint_grid_only(yy.to(dev), t))
    assert rel_l2(zs, z) < TOL


def test_gridded_exact_on_lattice_and_toeplitz(dev):
    from mr_recon_amd.fourier import gridded_nufft

    im, os_ = (16, 12), 2.0
    rng = np.random.default_rng(0)
    k = (np.round((rng.random((500, 2)) - 0.5) * np.array(im) * os_) / os_).astype(np.float32)
    hn = gridded_nufft(im, os_)
    t = hn.rescale_trajectory(torch.from_numpy(k).to(dev))[None]
    x = synth.random_image(im, seed=4)
    y = hn.forward(x[None].to(dev), t)[0]
    assert rel_l2(y, torch.from_numpy(ndft.ndft_forward(x.numpy(), k, im))) < 1e-5
    w = torch.rand((1, 500), generator=torch.Generator().manual_seed(1)) + 0.5
    K = hn.calc_teoplitz_kernels(t, w.to(dev))
    ref = calc_toep_kernel_helper(GriddedNufftRef(hn.im_size_os, 1.0).adjoint, t.cpu(), w) * os_ ** 2
    assert rel_l2(K, ref) < TOL


@pytest.mark.parametrize("im,os_,nspokes", [((64, 64), 2.0, 400), ((16, 16, 16), 1.5, 400)])
def test_gridded_long_runs_vs_oracle(dev, im, os_, nspokes):
    """Radial spokes all cross the k-space centre, so the centre point collects far more than
    one scatter chunk (64 samples) -- the multi-chunk split of the fused adjoint runs -- and
    the adjoint, called twice on one trajectory, reuses its cached index tables."""
    from mr_recon_amd.fourier import gridded_nufft

    d = len(im)
    n = im[0]
    k = (synth.radial_2d(nspokes, 2 * n, n) if d == 2 else synth.kooshball_3d(nspokes, 2 * n, n))[None]
    hn, rn = gridded_nufft(im, os_), GriddedNufftRef(im, os_)
    t = hn.rescale_trajectory(k.clone().to(dev))
    tr = rn.rescale_trajectory(k.clone())
    assert torch.equal(t.cpu(), tr)
    idx = hn._flat_index(t[0], hn.im_size_os, unshifted=True)
    runs = torch.unique(idx, return_counts=True)[1]
    assert int(runs.max()) > 4 * hn.CHUNK
    x = synth.random_image((1, 2, *im), seed=3)
    y = hn.forward(x.to(dev), t)
    assert rel_l2(y, rn.forward(x, tr)) < TOL
    yy = synth.random_image(tuple(y.shape), seed=5)
    z_ref = rn.adjoint(yy, tr)
    for _ in range(2):
        z = hn.adjoint(yy.to(dev), t)
        assert rel_l2(z, z_ref) < TOL
