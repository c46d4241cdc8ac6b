"""GPU parity of subspace_linop (all coils, coefficients and segments in one HIP NUFFT
call) against the oracle's loop over them."""

import pytest
import torch

from helpers import cdot, rel_l2
from mr_recon_amd import synth
from oracle.nufft_ref import TritonNufftRef
from oracle.sense_ref import SubspaceRef

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _case(D, K, seed):
    g = torch.Generator().manual_seed(seed)
    im = (24, 24) if D == 2 else (12, 12, 12)
    trj = ((torch.rand((40, 3, 6, D), generator=g) - 0.5) * im[0]).to(torch.float32)   # nro npe ntr d
    mps = synth.coil_maps(4, im, seed=seed)
    phi = torch.randn((K, 6), dtype=torch.complex64, generator=g)
    alphas = torch.randn((K, *im), dtype=torch.complex64, generator=g)
    dcf = torch.rand(trj.shape[:-1], generator=g) + 0.5
    return im, trj, mps, phi, alphas, dcf


@pytest.mark.parametrize("D,K,segs", [(2, 3, False), (2, 2, True), (3, 2, False)])
def test_subspace_vs_oracle(dev, D, K, segs):
    from mr_recon_amd.fourier import hip_nufft
    from mr_recon_amd.linops import subspace_linop

    im, trj, mps, phi, alphas, dcf = _case(D, K, seed=D + K)
    g = torch.Generator().manual_seed(7)
    sp = torch.randn((2, *im), dtype=torch.complex64, generator=g) if segs else None
    tp = torch.randn((2, *trj.shape[:-1]), dtype=torch.complex64, generator=g) if segs else None
    A = subspace_linop(im, trj.to(dev), mps.to(dev), phi.to(dev), dcf=dcf.to(dev),
                       nufft=hip_nufft(im, oversamp=2.0, width=6),
                       spatial_funcs=None if sp is None else sp.to(dev),
                       temporal_funcs=None if tp is None else tp.to(dev))
    R = SubspaceRef(trj, mps, phi, dcf=dcf, nufft=TritonNufftRef(im, 2.0, 6), spatial_funcs=sp, temporal_funcs=tp)
    y = A.forward(alphas.to(dev))
    assert y.shape == (4, *trj.shape[:-1])
    assert rel_l2(y, R.forward(alphas)) < TOL
    ksp = torch.randn(y.shape, dtype=torch.complex64, generator=g)
    x = A.adjoint(ksp.to(dev))
    assert x.shape == alphas.shape
    assert rel_l2(x, R.adjoint(ksp)) < TOL
    if not segs:  # the adjoint carries the dcf: <y, A x> = <A^H (y / dcf), x>
        xa = A.adjoint((ksp / dcf).to(dev))
        assert abs(cdot(ksp, y) - cdot(xa, alphas)) < 1e-4 * abs(cdot(ksp, y))


def test_subspace_toeplitz(dev):
    from mr_recon_amd.fourier import hip_nufft
    from mr_recon_amd.linops import subspace_linop

    im, trj, mps, phi, alphas, dcf = _case(2, 3, seed=11)
    A = subspace_linop(im, trj.to(dev), mps.to(dev), phi.to(dev), dcf=dcf.to(dev),
                       nufft=hip_nufft(im, oversamp=2.0, width=6), use_toeplitz=True)
    R = SubspaceRef(trj, mps, phi, dcf=dcf, nufft=TritonNufftRef(im, 2.0, 6), use_toeplitz=True)
    assert rel_l2(A.toep_kerns, R.toep_kerns) < TOL
    assert rel_l2(A.normal(alphas.to(dev)), R.normal(alphas)) < TOL
